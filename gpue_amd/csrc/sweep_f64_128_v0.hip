// generated: sweep kernels for N=128, variant 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 128, 0, sweep_f64_128_v0)

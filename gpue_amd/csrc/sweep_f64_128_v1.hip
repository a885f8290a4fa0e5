// generated: sweep kernels for N=128, variant 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 128, 1, sweep_f64_128_v1)

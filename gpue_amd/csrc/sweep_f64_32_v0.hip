// generated: sweep kernels for N=32, variant 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 32, 0, sweep_f64_32_v0)

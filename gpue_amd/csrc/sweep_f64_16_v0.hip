// generated: sweep kernels for N=16, variant 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 16, 0, sweep_f64_16_v0)

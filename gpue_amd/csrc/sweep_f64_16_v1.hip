// generated: sweep kernels for N=16, variant 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 16, 1, sweep_f64_16_v1)

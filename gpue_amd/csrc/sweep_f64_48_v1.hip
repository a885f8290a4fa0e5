// generated: sweep kernels for N=48, variant 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 48, 1, sweep_f64_48_v1)

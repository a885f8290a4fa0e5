// generated: sweep kernels for N=64, variant 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 64, 1, sweep_f64_64_v1)

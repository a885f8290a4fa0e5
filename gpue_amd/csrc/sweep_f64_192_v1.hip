// generated: sweep kernels for N=192, variant 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 192, 1, sweep_f64_192_v1)

// generated: sweep kernels for N=192, variant 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 192, 0, sweep_f64_192_v0)

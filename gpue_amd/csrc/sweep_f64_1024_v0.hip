// generated: sweep kernels for N=1024, variant 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 1024, 0, sweep_f64_1024_v0)

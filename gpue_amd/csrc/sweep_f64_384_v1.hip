// generated: sweep kernels for N=384, variant 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 384, 1, sweep_f64_384_v1)

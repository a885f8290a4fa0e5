// generated: sweep kernels for N=384, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 384, sweep_f64_384)

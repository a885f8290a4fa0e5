// generated: sweep kernels for N=192, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 192, sweep_f64_192)

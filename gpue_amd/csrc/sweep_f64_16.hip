// generated: sweep kernels for N=16, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 16, sweep_f64_16)

// generated: sweep kernels for N=32, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 32, sweep_f64_32)

// generated: sweep kernels for N=64, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 64, sweep_f64_64)

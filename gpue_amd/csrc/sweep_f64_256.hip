// generated: sweep kernels for N=256, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 256, sweep_f64_256)

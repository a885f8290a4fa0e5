// generated: sweep kernels for N=128, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 128, sweep_f64_128)

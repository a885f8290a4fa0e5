// generated: sweep kernels for N=512, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 512, sweep_f64_512)

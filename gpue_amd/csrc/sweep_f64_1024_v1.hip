// written by gen_units.py: sweep kernels for N=1024, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 1024, 1, sweep_f64_1024_v1)

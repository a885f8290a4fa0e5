// written by gen_units.py: sweep kernels for N=1024, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 1024, 0, sweep_f64_1024_v0)

// written by gen_units.py: sweep kernels for N=4096, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 4096, 0, sweep_f64_4096_v0)

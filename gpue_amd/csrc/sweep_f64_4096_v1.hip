// written by gen_units.py: sweep kernels for N=4096, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 4096, 1, sweep_f64_4096_v1)

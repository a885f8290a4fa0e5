// written by gen_units.py: sweep kernels for N=192, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 192, 1, sweep_f64_192_v1)

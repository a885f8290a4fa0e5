// written by gen_units.py: sweep kernels for N=64, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 64, 0, sweep_f64_64_v0)

// written by gen_units.py: sweep kernels for N=96, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 96, 0, sweep_f64_96_v0)

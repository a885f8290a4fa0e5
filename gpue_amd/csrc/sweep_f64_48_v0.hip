// written by gen_units.py: sweep kernels for N=48, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 48, 0, sweep_f64_48_v0)

// written by gen_units.py: sweep kernels for N=48, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 48, 1, sweep_f64_48_v1)

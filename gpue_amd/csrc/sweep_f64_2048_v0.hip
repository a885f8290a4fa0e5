// written by gen_units.py: sweep kernels for N=2048, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 2048, 0, sweep_f64_2048_v0)

// written by gen_units.py: sweep kernels for N=32, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 32, 0, sweep_f64_32_v0)

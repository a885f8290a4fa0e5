// written by gen_units.py: sweep kernels for N=32, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 32, 1, sweep_f64_32_v1)

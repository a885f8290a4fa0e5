// written by gen_units.py: sweep kernels for N=16, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 16, 1, sweep_f64_16_v1)

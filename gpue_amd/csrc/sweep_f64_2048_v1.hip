// written by gen_units.py: sweep kernels for N=2048, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 2048, 1, sweep_f64_2048_v1)

// written by gen_units.py: sweep kernels for N=256, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 256, 0, sweep_f64_256_v0)

// written by gen_units.py: sweep kernels for N=128, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 128, 0, sweep_f64_128_v0)

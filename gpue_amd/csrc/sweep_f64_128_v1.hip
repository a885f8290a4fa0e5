// written by gen_units.py: sweep kernels for N=128, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 128, 1, sweep_f64_128_v1)

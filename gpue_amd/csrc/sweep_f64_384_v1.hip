// written by gen_units.py: sweep kernels for N=384, plan 1, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 384, 1, sweep_f64_384_v1)

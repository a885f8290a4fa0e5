// written by gen_units.py: sweep kernels for N=512, plan 0, double
#include "sweep_launch.h"
GPUE_SWEEP_TU(double, 512, 0, sweep_f64_512_v0)

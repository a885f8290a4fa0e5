// written by gen_units.py: sweep kernels for N=1024, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 1024, 0, sweep_f32_1024_v0)

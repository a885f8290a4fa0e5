// written by gen_units.py: sweep kernels for N=4096, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 4096, 0, sweep_f32_4096_v0)

// written by gen_units.py: sweep kernels for N=16, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 16, 0, sweep_f32_16_v0)

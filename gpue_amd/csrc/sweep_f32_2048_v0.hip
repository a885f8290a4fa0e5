// written by gen_units.py: sweep kernels for N=2048, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 2048, 0, sweep_f32_2048_v0)

// written by gen_units.py: sweep kernels for N=64, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 64, 0, sweep_f32_64_v0)

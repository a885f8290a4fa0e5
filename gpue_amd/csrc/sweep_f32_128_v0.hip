// written by gen_units.py: sweep kernels for N=128, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 128, 0, sweep_f32_128_v0)

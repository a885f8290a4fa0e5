// written by gen_units.py: sweep kernels for N=128, plan 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 128, 1, sweep_f32_128_v1)

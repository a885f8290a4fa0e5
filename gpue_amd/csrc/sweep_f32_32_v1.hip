// written by gen_units.py: sweep kernels for N=32, plan 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 32, 1, sweep_f32_32_v1)

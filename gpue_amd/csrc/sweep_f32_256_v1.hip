// written by gen_units.py: sweep kernels for N=256, plan 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 256, 1, sweep_f32_256_v1)

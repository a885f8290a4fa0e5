// written by gen_units.py: sweep kernels for N=48, plan 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 48, 1, sweep_f32_48_v1)

// written by gen_units.py: sweep kernels for N=192, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 192, 0, sweep_f32_192_v0)

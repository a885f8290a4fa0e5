// written by gen_units.py: sweep kernels for N=192, plan 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 192, 1, sweep_f32_192_v1)

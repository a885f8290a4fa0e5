// written by gen_units.py: sweep kernels for N=384, plan 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 384, 1, sweep_f32_384_v1)

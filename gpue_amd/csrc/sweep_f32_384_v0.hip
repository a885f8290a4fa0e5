// written by gen_units.py: sweep kernels for N=384, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 384, 0, sweep_f32_384_v0)

// written by gen_units.py: sweep kernels for N=512, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 512, 0, sweep_f32_512_v0)

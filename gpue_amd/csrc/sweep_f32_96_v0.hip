// written by gen_units.py: sweep kernels for N=96, plan 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 96, 0, sweep_f32_96_v0)

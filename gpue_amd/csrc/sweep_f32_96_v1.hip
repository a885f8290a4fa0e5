// written by gen_units.py: sweep kernels for N=96, plan 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 96, 1, sweep_f32_96_v1)

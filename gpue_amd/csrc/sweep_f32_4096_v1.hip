// written by gen_units.py: sweep kernels for N=4096, plan 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 4096, 1, sweep_f32_4096_v1)

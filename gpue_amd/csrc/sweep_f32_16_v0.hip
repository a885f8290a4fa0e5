// generated: sweep kernels for N=16, variant 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 16, 0, sweep_f32_16_v0)

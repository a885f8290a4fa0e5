// generated: sweep kernels for N=16, variant 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 16, 1, sweep_f32_16_v1)

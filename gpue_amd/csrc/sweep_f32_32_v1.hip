// generated: sweep kernels for N=32, variant 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 32, 1, sweep_f32_32_v1)

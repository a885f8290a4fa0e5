// generated: sweep kernels for N=128, variant 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 128, 0, sweep_f32_128_v0)

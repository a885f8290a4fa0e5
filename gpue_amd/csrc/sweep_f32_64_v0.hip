// generated: sweep kernels for N=64, variant 0, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 64, 0, sweep_f32_64_v0)

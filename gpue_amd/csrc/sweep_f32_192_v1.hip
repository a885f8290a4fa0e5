// generated: sweep kernels for N=192, variant 1, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 192, 1, sweep_f32_192_v1)

// generated: sweep kernels for N=16, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 16, sweep_f32_16)

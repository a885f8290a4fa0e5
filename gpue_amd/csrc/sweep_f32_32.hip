// generated: sweep kernels for N=32, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 32, sweep_f32_32)

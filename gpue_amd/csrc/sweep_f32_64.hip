// generated: sweep kernels for N=64, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 64, sweep_f32_64)

// generated: sweep kernels for N=192, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 192, sweep_f32_192)

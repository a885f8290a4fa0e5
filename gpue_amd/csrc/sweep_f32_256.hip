// generated: sweep kernels for N=256, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 256, sweep_f32_256)

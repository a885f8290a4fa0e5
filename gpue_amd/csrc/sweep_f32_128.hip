// generated: sweep kernels for N=128, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 128, sweep_f32_128)

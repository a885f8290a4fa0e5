// generated: sweep kernels for N=512, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 512, sweep_f32_512)

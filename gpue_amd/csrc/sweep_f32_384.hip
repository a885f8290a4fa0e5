// generated: sweep kernels for N=384, float
#include "sweep_launch.h"
GPUE_SWEEP_TU(float, 384, sweep_f32_384)

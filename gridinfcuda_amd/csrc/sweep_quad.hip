// Quad-per-node half-sweep kernels with 32-bit data costs (see sweep_quad.inc).
#define GINF_QUAD_D16 false
#define GINF_QUAD_KERNEL k_sweep_quad
#define GINF_QUAD_LAUNCH launch_sweep_quad
#include "sweep_quad.inc"

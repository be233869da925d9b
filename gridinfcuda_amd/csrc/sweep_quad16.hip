// Quad-per-node half-sweep kernels with 16-bit data costs (see sweep_quad.inc).
#define GINF_QUAD_D16 true
#define GINF_QUAD_KERNEL k_sweep_quad16
#define GINF_QUAD_LAUNCH launch_sweep_quad16
#include "sweep_quad.inc"

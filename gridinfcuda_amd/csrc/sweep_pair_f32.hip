// Fused pairs of half-sweeps, FP32 messages (see sweep_pair.inc).
#define GINF_PAIR_T float
#define GINF_PAIR_LAUNCH launch_sweep_pair_f32
#define GINF_PAIR_TILES pair_tiles_f32
#include "sweep_pair.inc"

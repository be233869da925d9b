// Fused pairs of half-sweeps, int32 messages (see sweep_pair.inc).
#define GINF_PAIR_T int
#define GINF_PAIR_LAUNCH launch_sweep_pair_i32
#define GINF_PAIR_TILES pair_tiles_i32
#include "sweep_pair.inc"

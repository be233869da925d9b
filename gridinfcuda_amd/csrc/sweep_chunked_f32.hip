// Chunked O(L) half-sweep kernels for float costs.
#define GINF_T float
#define GINF_CHUNKED_ENTRY launch_sweep_chunked_f32
#include "sweep_chunked.inc"

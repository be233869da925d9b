// Chunked O(L) half-sweep kernels for int costs.
#define GINF_T int
#define GINF_CHUNKED_ENTRY launch_sweep_chunked_i32
#include "sweep_chunked.inc"

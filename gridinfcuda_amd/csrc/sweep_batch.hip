// Batched grid-stride half-sweep, 32-bit data costs (see sweep_batch.inc).
#define GINF_BATCH_D16 false
#define GINF_BATCH_KERNEL k_sweep_batch
#define GINF_BATCH_LAUNCH launch_sweep_batch
#include "sweep_batch.inc"

// Batched grid-stride half-sweep, 16-bit data costs where they fit (see sweep_batch.inc).
#define GINF_BATCH_D16 true
#define GINF_BATCH_KERNEL k_sweep_batch16
#define GINF_BATCH_LAUNCH launch_sweep_batch16
#include "sweep_batch.inc"

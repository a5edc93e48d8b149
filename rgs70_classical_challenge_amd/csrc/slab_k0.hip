// k_tile specialisations with 0 lane bits.
#define RGS_C 0
#define RGS_LAUNCH_NAME launch_c0
#define RGS_MULTI_NAME launch_c0_multi
#include "slab_tiles.inc"

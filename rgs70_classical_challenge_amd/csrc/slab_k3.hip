// k_tile specialisations with 3 lane bits.
#define RGS_C 3
#define RGS_LAUNCH_NAME launch_c3
#define RGS_MULTI_NAME launch_c3_multi
#include "slab_tiles.inc"

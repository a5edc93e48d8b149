// k_tile specialisations with 2 lane bits.
#define RGS_C 2
#define RGS_LAUNCH_NAME launch_c2
#define RGS_MULTI_NAME launch_c2_multi
#include "slab_tiles.inc"

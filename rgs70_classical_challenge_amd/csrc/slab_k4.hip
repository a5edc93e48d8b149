// k_tile specialisations with 4 lane bits.
#define RGS_C 4
#define RGS_LAUNCH_NAME launch_c4
#define RGS_MULTI_NAME launch_c4_multi
#include "slab_tiles.inc"

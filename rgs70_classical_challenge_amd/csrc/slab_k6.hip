// k_tile specialisations with 6 lane bits.
#define RGS_C 6
#define RGS_LAUNCH_NAME launch_c6
#define RGS_MULTI_NAME launch_c6_multi
#include "slab_tiles.inc"

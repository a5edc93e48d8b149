// k_tile specialisations with 5 lane bits.
#define RGS_C 5
#define RGS_LAUNCH_NAME launch_c5
#define RGS_MULTI_NAME launch_c5_multi
#include "slab_tiles.inc"

// k_tile specialisations with 7 lane bits.
#define RGS_C 7
#define RGS_LAUNCH_NAME launch_c7
#define RGS_MULTI_NAME launch_c7_multi
#include "slab_tiles.inc"

// k_tile specialisations with 0 lane bits.
#define RGS_C 0
#define RGS_LAUNCH_NAME launch_c0
#include "slab_tiles.inc"

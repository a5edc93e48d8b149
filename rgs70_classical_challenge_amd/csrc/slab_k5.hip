// k_tile specialisations with 5 lane bits.
#define RGS_C 5
#define RGS_LAUNCH_NAME launch_c5
#include "slab_tiles.inc"

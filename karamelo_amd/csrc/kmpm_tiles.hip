// kmpm_tiles.hip -- one object file per tile variant: built with
// -DKMPM_TILE_DIM=<2|3> -DKMPM_TILE_S1=<2|4> -DKMPM_TILE_V=<variant>.
#include "kmpm_tile_launch.h"

template struct kmpm::TileLaunch<KMPM_TILE_DIM, KMPM_TILE_S1, KMPM_TILE_V>;

// Common.h -- constants of Include/Common.h:3-4.
#pragma once

#define PI 3.141592653589793238462643383
#define BLOCK_SLICES 32

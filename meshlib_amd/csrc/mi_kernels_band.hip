// mi_kernels_band.hip -- placeholder (filled below)
#include "mi_walk.hpp"

// sblas_tiles.hip -- segmented-tile kernels (COO, merge path, CSR5-style).
#include "internal.hpp"

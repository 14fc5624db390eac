// Stage, auxiliary and batched kernels for 7 nodes per direction (p = 6), 2D and 3D.
#include "dg_dispatch.hpp"

namespace dgx {
DGX_KERNELS_N(, 7)
} // namespace dgx

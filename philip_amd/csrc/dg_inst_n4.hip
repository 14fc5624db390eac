// Stage, auxiliary and batched kernels for 4 nodes per direction (p = 3), 2D and 3D.
#include "dg_dispatch.hpp"

namespace dgx {
DGX_KERNELS_N(, 4)
} // namespace dgx

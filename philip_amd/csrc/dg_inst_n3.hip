// Stage, auxiliary and batched kernels for 3 nodes per direction (p = 2), 2D and 3D.
#include "dg_dispatch.hpp"

namespace dgx {
DGX_KERNELS_N(, 3)
} // namespace dgx

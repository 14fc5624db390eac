// Stage, auxiliary and batched kernels for 2 nodes per direction (p = 1), 2D and 3D.
#include "dg_dispatch.hpp"

namespace dgx {
DGX_KERNELS_N(, 2)
} // namespace dgx

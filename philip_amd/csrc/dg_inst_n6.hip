// Stage, auxiliary and batched kernels for 6 nodes per direction (p = 5), 2D and 3D.
#include "dg_dispatch.hpp"

namespace dgx {
DGX_KERNELS_N(, 6)
} // namespace dgx

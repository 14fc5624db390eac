// Stage, auxiliary and batched kernels for 5 nodes per direction (p = 4), 2D and 3D.
#include "dg_dispatch.hpp"

namespace dgx {
DGX_KERNELS_N(, 5)
} // namespace dgx

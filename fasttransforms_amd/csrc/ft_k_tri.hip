// ft_k_tri.hip -- instantiates the pipelined engine for the triangle geometry
// (its own translation unit so the geometries compile in parallel).
#include "ft_pipeline.hpp"

namespace ftamd {
KernelFn pick_tri(int precision, int dir, int L, int W) { return pick_kind<FTAMD_TRI>(precision, dir, L, W); }
}  // namespace ftamd

// ft_k_sph.hip -- instantiates the pipelined engine for the sphere geometry
// (its own translation unit so the geometries compile in parallel).
#include "ft_pipeline.hpp"

namespace ftamd {
KernelFn pick_sph(int precision, int dir, int L, int W) { return pick_kind<FTAMD_SPH>(precision, dir, L, W); }
}  // namespace ftamd

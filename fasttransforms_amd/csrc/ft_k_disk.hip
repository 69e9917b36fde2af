// ft_k_disk.hip -- instantiates the pipelined engine for the disk geometry
// (its own translation unit so the geometries compile in parallel).
#include "ft_pipeline.hpp"

namespace ftamd {
KernelFn pick_disk(int precision, int dir, int L, int W) { return pick_kind<FTAMD_DISK>(precision, dir, L, W); }
}  // namespace ftamd

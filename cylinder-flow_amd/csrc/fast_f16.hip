// fast_f16.hip -- instantiations of the V-nodes-per-lane fused step for SF16 storage
// (split per storage type so the kernels compile in parallel).
#include "fast_kernel.hpp"

namespace lbm {
hipError_t fast_step_f16(int Q, int V, const StepBufs &b, bool force, bool mac, const FusedArgs &a, hipStream_t s) {
    return fast_step<SF16, 2, 4>(Q, V, b, force, mac, a, s);
}
}  // namespace lbm

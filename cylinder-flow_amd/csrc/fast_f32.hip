// fast_f32.hip -- instantiations of the V-nodes-per-lane fused step for SF32 storage
// (split per storage type so the kernels compile in parallel).
#include "fast_kernel.hpp"

namespace lbm {
hipError_t fast_step_f32(int Q, int V, const StepBufs &b, bool force, bool mac, const FusedArgs &a, hipStream_t s) {
    return fast_step<SF32, 1, 2>(Q, V, b, force, mac, a, s);
}
}  // namespace lbm

// Device helpers of the node-slice layout (see SliceView) shared by the multiply and the
// fused iteration.
#pragma once

#include "lfg_kernels.hpp"

namespace lfg {

// FP32 copy: the diagonal block (FP64, compensated) times the node's own pair, added
// after the off-diagonal blocks.
__device__ __forceinline__ void add_diag32(const SliceView& v, int node, dpair x, double& a0, double& a1) {
    const dpair du = reinterpret_cast<const dpair*>(v.diag32)[2 * node];
    const dpair dl = reinterpret_cast<const dpair*>(v.diag32)[2 * node + 1];
    a0 += du.x * x.x;
    a0 += du.y * x.y;
    a1 += dl.x * x.x;
    a1 += dl.y * x.y;
}

}  // namespace lfg

// XCD-aware block numbering for gather-heavy kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace vfvm {

// MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin (block b runs on
// XCD b % 8), and each XCD has its own L2.  Rows that are neighbours gather the same
// vector and field entries, so kernels that gather take their row range from
// xcd_block(): consecutive logical blocks run on the same XCD and share its L2.
constexpr unsigned NUM_XCD = 8;
__device__ inline unsigned xcd_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned x = b % NUM_XCD, q = nb / NUM_XCD, r = nb % NUM_XCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + b / NUM_XCD;
}

}  // namespace vfvm

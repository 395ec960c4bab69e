// Device helpers shared by the FP32 fast kernels (rhs_fast.hip, vjp_fast.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flouhip {
namespace fast {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float lg(float x) { return __logf(x); }
__device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }

constexpr int r4(int x) { return (x + 3) & ~3; }
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// Buffer resources: a uniform base in SGPRs, 32-bit lane byte offsets, and the
// hardware range check (loads past num_records return 0, stores are dropped).  The fast
// paths are only enabled for states of fewer than 2^30 nodes, so every per-variable byte
// offset fits in 32 bits.  Word 3 = 0x00020000: 32-bit data format (gfx9 family).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const float* base, uint32_t nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)nbytes, 0x00020000);
}
__device__ __forceinline__ float bload(rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, uint32_t off, float x) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, (int)off, 0, 0);
}

// The log-mean path of a wave's pair blocks: 2 when every line of the wave allows the
// first-order form, else 1 when every line allows the series, else 0 (each path is valid
// for the lines of the cheaper ones).  One path per wave: lines of one wave on different
// paths would run both pair blocks one after the other.
__device__ __forceinline__ int wave_path(bool first_order, bool series) {
    return __all(first_order) ? 2 : (__all(series) ? 1 : 0);
}

// slot-array position of block node j (PAD: rows of N nodes, one slot between rows)
template <int N, bool PAD>
__device__ __forceinline__ int pnode(int j) {
    return PAD ? j + j / N : j;
}

}  // namespace fast
}  // namespace flouhip

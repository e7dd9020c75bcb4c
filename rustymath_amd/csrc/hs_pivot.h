// Pivot search and broadcast helpers shared by the LU kernels (lu.hip, batched.hip): candidate keys with
// rustymath's tie rule, a DPP wave reduction, and lane-to-scalar reads.
#pragma once
#include "hs_common.h"

namespace hs {

// A candidate's key: larger magnitude first, then the smaller position, NaN never chosen over a number:
// `largest`'s first maximum (core/src/linalg/factor.rs:476). kv is the bit pattern of |x| plus one (the order of
// non-negative IEEE values is the order of their bit patterns), 0 for NaN; a row that is no candidate has kv = 0
// and position INT_MAX, so an all-NaN column picks its first row, as `largest` does.
template <typename T>
__device__ __forceinline__ unsigned long long mag_key(T x) {
    if constexpr (sizeof(T) == 8) {
        const double ax = fabs(x);
        return ax != ax ? 0ull : __builtin_bit_cast(unsigned long long, ax) + 1ull;
    } else {
        const float ax = fabsf(x);
        return ax != ax ? 0ull : (unsigned long long)__builtin_bit_cast(unsigned, ax) + 1ull;
    }
}
// The wave's maximum of a 32-bit value through DPP row shifts and broadcasts (it lands in lane 63): one max per step,
// with the row shift or broadcast as the DPP operand (0 is the identity of the maximum, for lanes with no source).
// An FP32 item's key (bits of |x| plus one) fits in 32 bits; a 64-bit key takes two such passes.
__device__ __forceinline__ unsigned dpp_max_u32(unsigned x) {
#define HS_DPP_MAX(ctrl, rm, bm) \
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, ctrl, rm, bm, false))
    HS_DPP_MAX(0x111, 0xf, 0xf);  // row_shr:1
    HS_DPP_MAX(0x112, 0xf, 0xf);  // row_shr:2
    HS_DPP_MAX(0x114, 0xf, 0xf);  // row_shr:4
    HS_DPP_MAX(0x118, 0xf, 0xf);  // row_shr:8
    HS_DPP_MAX(0x142, 0xa, 0xf);  // row_bcast:15
    HS_DPP_MAX(0x143, 0xc, 0xf);  // row_bcast:31
#undef HS_DPP_MAX
    return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}
__device__ __forceinline__ unsigned long long dpp_max_u64(unsigned long long x) {
    // The larger high word first, then the largest low word among the lanes that hold it; where one lane alone holds
    // the largest high word, its key is the maximum and the second pass is skipped.
    const unsigned hi = (unsigned)(x >> 32), bhi = dpp_max_u32(hi);
    const unsigned long long tied = __ballot(hi == bhi);
    if (tied != 0ull && (tied & (tied - 1)) == 0ull)
        return ((unsigned long long)bhi << 32) | (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, (int)__builtin_ctzll(tied));
    const unsigned blo = dpp_max_u32(hi == bhi ? (unsigned)x : 0u);
    return ((unsigned long long)bhi << 32) | blo;
}
// The wave's best key: the largest kv (32 or 64 bits), then among the lanes holding it the smallest position (one
// lane, save for ties). Returns wave-uniform values; `lane_of` gets the winning lane (-1 where no lane is a
// candidate: kp = INT_MAX everywhere).
template <typename K>
__device__ __forceinline__ void wave_best(K& kv, int& kp, int& lane_of) {
    K best;
    if constexpr (sizeof(K) == 4) best = dpp_max_u32(kv);
    else best = dpp_max_u64(kv);
    unsigned long long m = __ballot(kv == best && kp != 0x7fffffff);
    if (m == 0ull) { kv = 0ull; kp = 0x7fffffff; lane_of = -1; return; }
    int L = __builtin_ctzll(m);
    int bp = __builtin_amdgcn_readlane(kp, L);
    m &= m - 1;
    while (m) {  // ties of magnitude (rare): the smallest position
        const int L2 = __builtin_ctzll(m);
        const int p2 = __builtin_amdgcn_readlane(kp, L2);
        if (p2 < bp) { bp = p2; L = L2; }
        m &= m - 1;
    }
    kv = best;
    kp = bp;
    lane_of = L;
}

// Lane l's value of v (l uniform), as a scalar operand.
template <typename T>
__device__ __forceinline__ T lane_value(T v, int l) {
    if constexpr (sizeof(T) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    }
}

}  // namespace hs

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
__device__ __forceinline__ bool better(unsigned long long kv, int kp, unsigned long long kv2, int kp2) {
    return (kv2 > kv) | ((kv2 == kv) & (kp2 < kp));
}
// The wave's best key: the largest kv through DPP row shifts and broadcasts (the maximum lands in lane 63), then
// among the lanes holding it the smallest position (one lane, save for ties). Returns wave-uniform values; `lane_of`
// gets the winning lane (-1 where no lane is a candidate: kp = INT_MAX everywhere).
__device__ __forceinline__ unsigned long long dpp_max_u64(unsigned long long x) {
    unsigned lo = (unsigned)x, hi = (unsigned)(x >> 32);
#define HS_DPP_STEP(ctrl, rm, bm)                                                                      \
    do {                                                                                               \
        const unsigned lo2 = (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo, ctrl, rm, bm, false);   \
        const unsigned hi2 = (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi, ctrl, rm, bm, false);   \
        const bool gt = (hi2 > hi) | ((hi2 == hi) & (lo2 > lo));                                      \
        lo = gt ? lo2 : lo;                                                                            \
        hi = gt ? hi2 : hi;                                                                            \
    } while (0)
    HS_DPP_STEP(0x111, 0xf, 0xf);  // row_shr:1
    HS_DPP_STEP(0x112, 0xf, 0xf);  // row_shr:2
    HS_DPP_STEP(0x114, 0xf, 0xf);  // row_shr:4
    HS_DPP_STEP(0x118, 0xf, 0xf);  // row_shr:8: lane 15 of each row holds the row's maximum
    HS_DPP_STEP(0x142, 0xa, 0xf);  // row_bcast:15 into rows 1 and 3
    HS_DPP_STEP(0x143, 0xc, 0xf);  // row_bcast:31 into rows 2 and 3: lane 63 holds the wave's maximum
#undef HS_DPP_STEP
    const unsigned blo = (unsigned)__builtin_amdgcn_readlane((int)lo, 63);
    const unsigned bhi = (unsigned)__builtin_amdgcn_readlane((int)hi, 63);
    return ((unsigned long long)bhi << 32) | blo;
}
__device__ __forceinline__ void wave_best(unsigned long long& kv, int& kp, int& lane_of) {
    const unsigned long long best = dpp_max_u64(kv);
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

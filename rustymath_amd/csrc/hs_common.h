// Shared plumbing of the hip-sys library: status codes, the per-thread error text, the per-thread
// stream, the device-memory pool, and the MFMA operand traits that the kernels share.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <type_traits>
#include <mutex>

namespace hs {

// Status codes of every entry point: 0 = ok, -i = argument i is invalid (LAPACK style),
// > 0 = a HIP runtime failure (the hipError_t value), ERR_LIMIT = a device-side bounded wait ran out,
// ERR_UNSUPPORTED = valid arguments beyond what the library handles.
constexpr int OK = 0;
constexpr int ERR_LIMIT = 100000;
constexpr int ERR_UNSUPPORTED = 100001;  // valid arguments that this library does not handle (the message says which)

void set_error(const std::string& what);
const char* last_error();

// The calling thread's stream: one per host thread, so that calls from concurrent host threads
// (the reference runs cells on Rayon workers) neither serialise on nor race through one queue.
hipStream_t thread_stream();

// Device memory from a size-keyed pool that keeps buffers for reuse across calls, as `LuWork` keeps
// its buffers across factorizations (core/src/linalg/factor.rs:196-201).
void* pool_get(size_t bytes);
void pool_put(void* p);
void pool_release_all();

// Number of compute units of the current device (for grid sizing of persistent kernels).
int device_cus();
// The workgroup exchanges of the LU panels and single-column substitutions go through one XCD's L2 where launch
// indices congruent mod 8 share an XCD (probed once per device: at hs_init or the first call that may probe) and
// the caller has not asked for device scope; otherwise their words are published at device scope.
bool exchange_local(bool may_probe);

// Held by every launch sequence whose workgroups wait on one another and so must all be resident at once (the LU's
// panel launches, the single-column solves' runs): one such sequence at a time in the process, so that two of them on
// different streams never each hold part of the CUs the other needs.
std::mutex& resident_mutex();

int hip_status(hipError_t e, const char* where);

}  // namespace hs

#define HS_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) return hs::hip_status(e_, #call);               \
    } while (0)

// A loop over compile-time indices C..E-1: f(std::integral_constant<int, i>) for each i, so that
// register arrays indexed by i stay in registers.
namespace hs {
template <int C, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (C < E) {
        f(std::integral_constant<int, C>{});
        static_for<C + 1, E>(f);
    }
}
}  // namespace hs

// MFMA operand traits. Each 16x16x4 instruction takes one item of A (16x4) and one of B (4x16) per
// lane: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]. The 16x16 result gives lane l four items in
// column l % 16: rows 4 (l / 16) + e for FP32, but rows l / 16 + 4 e for FP64 (measured on gfx950).
namespace hs {
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4 acc;
    static __device__ __forceinline__ int row(int lane, int e) { return (lane >> 4) + 4 * e; }
    static __device__ __forceinline__ acc step(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef f4 acc;
    static __device__ __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
    static __device__ __forceinline__ acc step(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
}  // namespace hs

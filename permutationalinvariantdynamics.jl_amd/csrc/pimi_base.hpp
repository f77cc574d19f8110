// Types and address-space helpers shared by every device kernel (also
// compiled at run time, for the banded programs of a plan).
//
// Complex vectors are arrays of cx<R> (R = double or float), 8/16-byte pairs.
#pragma once

#if defined(__HIPCC_RTC__)
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace pimi {


template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// conj(a) * b
template <typename R>
__device__ __forceinline__ cx<R> cjmul(cx<R> a, cx<R> b) { return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }

// Address spaces: tables that no kernel writes while it runs are read through
// the constant address space (wave-uniform reads become scalar loads, also
// when the table pointer itself was loaded from a descriptor); vectors
// through the global one (global_load, not flat_load).
// (The attributes apply to the device pass only; the host pass sees plain
// pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define PIMI_CONST __attribute__((address_space(4)))
#define PIMI_GLOBAL __attribute__((address_space(1)))
#else
#define PIMI_CONST
#define PIMI_GLOBAL
#endif
template <typename T>
__device__ __forceinline__ const PIMI_CONST T *cptr(const T *p) { return (const PIMI_CONST T *)p; }
template <typename T>
__device__ __forceinline__ const PIMI_GLOBAL T *gptr(const T *p) { return (const PIMI_GLOBAL T *)p; }
template <typename T>
__device__ __forceinline__ PIMI_GLOBAL T *gptr_w(T *p) { return (PIMI_GLOBAL T *)p; }

constexpr int kApplyBlock = 256;
constexpr int kRedBlock = 256;
constexpr int kNumXcd = 8;

// An apply tile: rows a0.. of columns b0..b0+nb-1 of sector `sector`'s
// output block, one wave (the row count per tile is the kernel's).
struct ApplyTile {
    int32_t sector, b0, a0, nb;
};
constexpr int kApplyWave = 64;

}  // namespace pimi

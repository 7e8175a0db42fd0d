// common.h -- shared device helpers and host plumbing for libgdsjoin.
// gfx950 (CDNA4) only: wave64, ballot/mbcnt compaction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "gdsjoin.h"

#define GDSJ_WAVE 64

namespace gdsj {

// ---------------------------------------------------------------- errors --
void set_error(const char *fmt, ...);

#define GDSJ_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::gdsj::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,           \
                              hipGetErrorString(e_));                               \
            return e_ == hipErrorOutOfMemory ? GDSJ_EOOM : GDSJ_EDEVICE;            \
        }                                                                           \
    } while (0)

#define GDSJ_TRY(call)                                                              \
    do {                                                                            \
        int rc_ = (call);                                                           \
        if (rc_ != GDSJ_OK) return rc_;                                             \
    } while (0)

// ------------------------------------------------------- wave primitives --
// the wave's predicate mask (the compare result itself, no 0/1 materialisation)
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// number of set bits of m in lanes below this lane
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// add + number of set bits of m in lanes below this lane
__device__ __forceinline__ uint32_t mbcnt_add(uint64_t m, uint32_t add) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, add));
}

__device__ __forceinline__ uint32_t popc64(uint64_t m) { return (uint32_t)__popcll(m); }

// v_writelane_b32: lane `lane` (wave-uniform) of old replaced by the uniform
// value v; the LLVM intrinsic has no clang builtin, so it is bound by name.
extern "C" __device__ int gdsj_llvm_writelane_i32(int v, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ uint32_t writelane_u32(uint32_t v, uint32_t lane, uint32_t old) {
    return (uint32_t)gdsj_llvm_writelane_i32((int)v, (int)lane, (int)old);
}

// lanes 0 .. n-1 (all lanes for n >= 64)
__device__ __forceinline__ uint64_t prefix_mask(uint32_t n) {
    return n >= (uint32_t)GDSJ_WAVE ? ~0ull : ((1ull << n) - 1ull);
}

// element at byte offset bo (below 2^32) from a wave-uniform base: the
// access takes the base in scalar registers and a 32-bit offset per lane
// instead of a 64-bit address per lane
template <typename T>
__device__ __forceinline__ T load_at(const T *__restrict__ base, uint32_t bo) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + bo);
}
template <typename T>
__device__ __forceinline__ void store_at(T *__restrict__ base, uint32_t bo, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + bo) = v;
}

// raw buffer resource over [base, base + 4 GiB): a load takes it in scalar
// registers with a scalar and a 32-bit per-lane byte offset (third dword:
// 32-bit data format, the gfx9 family's raw-buffer setting)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)0xFFFFFFFF, 0x00020000);
}
__device__ __forceinline__ float buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ double buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

// value of lane i (wave-uniform i) as a scalar
__device__ __forceinline__ float lane_value(float v, uint32_t i) {
    return __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), i));
}
__device__ __forceinline__ double lane_value(double v, uint32_t i) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, i);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), i);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}


template <typename U>
__device__ __forceinline__ U uniform(U v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  This bijection on [0, nblocks) gives each XCD one contiguous range
// of logical blocks, so neighboring tiles (which share candidate cells) meet
// in the same L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
    constexpr uint32_t NX = 8;
    const uint32_t per = nblocks / NX, rem = nblocks % NX;
    const uint32_t x = b % NX, slot = b / NX;
    return x < rem ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}

// ------------------------------------------------- exact FP building blocks --
// Explicit fused multiply-add: the translation units are compiled with
// -ffp-contract=off so that only these calls fuse.
__device__ __forceinline__ float fma_rn(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_rn(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ------------------------------------------------------------ host helpers --
inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// device buffer that only grows (scratch reused across joins on one handle)
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return GDSJ_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = bytes + bytes / 8 + 256;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
            return GDSJ_EOOM;
        }
        cap = want;
        return GDSJ_OK;
    }
    template <typename U> U *as() const { return (U *)p; }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// pinned host buffer that only grows (device-to-host staging)
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return GDSJ_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
            return GDSJ_EOOM;
        }
        cap = bytes;
        return GDSJ_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

}  // namespace gdsj

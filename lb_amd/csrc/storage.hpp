// storage.hpp -- population storage formats in HBM.
//
// FP64 and FP32 store f_i.  F16DEV / BF16DEV store the deviation
// g_i = f_i - w_i*rho0 (well inside the 16-bit formats' range for the
// low-Mach flows this core targets) and do all arithmetic in FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

struct StoreF64 {
    using T = double;   // element in HBM
    using R = double;   // arithmetic type
    static constexpr bool dev = false;
    __device__ __forceinline__ static R load(const T* p, uint32_t i, R) { return p[i]; }
    __device__ __forceinline__ static void store(T* p, uint32_t i, R v, R) { p[i] = v; }
};

struct StoreF32 {
    using T = float;
    using R = float;
    static constexpr bool dev = false;
    __device__ __forceinline__ static R load(const T* p, uint32_t i, R) { return p[i]; }
    __device__ __forceinline__ static void store(T* p, uint32_t i, R v, R) { p[i] = v; }
};

struct StoreF16Dev {
    using T = _Float16;
    using R = float;
    static constexpr bool dev = true;
    __device__ __forceinline__ static R load(const T* p, uint32_t i, R rest) { return (float)p[i] + rest; }
    __device__ __forceinline__ static void store(T* p, uint32_t i, R v, R rest) { p[i] = (_Float16)(v - rest); }
};

// bfloat16 with round-to-nearest-even on store.
struct StoreBF16Dev {
    using T = uint16_t;
    using R = float;
    static constexpr bool dev = true;
    __device__ __forceinline__ static R load(const T* p, uint32_t i, R rest)
    {
        return __uint_as_float((uint32_t)p[i] << 16) + rest;
    }
    __device__ __forceinline__ static void store(T* p, uint32_t i, R v, R rest)
    {
        const uint32_t b = __float_as_uint(v - rest);
        p[i] = (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
    }
};

}  // namespace lbm

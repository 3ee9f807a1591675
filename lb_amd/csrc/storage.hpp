// storage.hpp -- population storage formats in HBM.
//
// FP64 and FP32 store f_i.  F16DEV / BF16DEV store the deviation
// g_i = f_i - w_i*rho0 (well inside the 16-bit formats' range for the
// low-Mach flows this core targets) and do all arithmetic in FP32.
//
// Every policy has store(p, i, v, rest) (round to nearest even) and
// put<NT>(p, i, v, rest, rnd).  The second form rounds stochastically in
// the *SR policies and to nearest in the others, and takes the streaming
// hint NT; the collide-stream kernel uses it for the populations it streams,
// all other writers use the first form.  load(p, i, rest) and
// get<NT>(p, i, rest) likewise.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

// Element i of a population plane.  The byte offset is formed in 32 bits (a
// plane holds at most 4 GiB, checked when the lattice is created), so the
// access is a 64-bit uniform base plus a 32-bit per-lane offset.
template <class T>
__device__ __forceinline__ const T& at(const T* p, uint32_t i)
{
    return *(const T*)((const char*)p + (uint32_t)(i * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ T& at(T* p, uint32_t i)
{
    return *(T*)((char*)p + (uint32_t)(i * (uint32_t)sizeof(T)));
}

// Plain or non-temporal (streaming) access to element i, chosen at compile
// time (a run-time choice between two accesses to one address lets the
// compiler merge them and drop the hint).  The collide-stream step streams
// every population once per step; where a wave's access covers whole cache
// lines (every access of the even step, and the directions without an x
// component on the odd step), the streaming hint keeps those lines from
// displacing the partial lines that neighbouring tiles share.
template <bool NT, class T>
__device__ __forceinline__ T ld(const T* p, uint32_t i)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(&at(p, i));
    else
        return at(p, i);
}
template <bool NT, class T>
__device__ __forceinline__ void st(T* p, uint32_t i, T v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, &at(p, i));
    else
        at(p, i) = v;
}

struct StoreF64 {
    using T = double;   // element in HBM
    using R = double;   // arithmetic type
    static constexpr bool dev = false;
    static constexpr bool sr = false;
    // the value a store then a load would return
    __device__ __forceinline__ static R round_trip(R v, R) { return v; }
    __device__ __forceinline__ static R load(const T* p, uint32_t i, R) { return at(p, i); }
    __device__ __forceinline__ static void store(T* p, uint32_t i, R v, R) { at(p, i) = v; }
    template <bool NT>
    __device__ __forceinline__ static R get(const T* p, uint32_t i, R)
    {
        return ld<NT>(p, i);
    }
    template <bool NT>
    __device__ __forceinline__ static void put(T* p, uint32_t i, R v, R, uint32_t)
    {
        st<NT>(p, i, v);
    }
};

struct StoreF32 {
    using T = float;
    using R = float;
    static constexpr bool dev = false;
    static constexpr bool sr = false;
    __device__ __forceinline__ static R round_trip(R v, R) { return v; }
    __device__ __forceinline__ static R load(const T* p, uint32_t i, R) { return at(p, i); }
    __device__ __forceinline__ static void store(T* p, uint32_t i, R v, R) { at(p, i) = v; }
    template <bool NT>
    __device__ __forceinline__ static R get(const T* p, uint32_t i, R)
    {
        return ld<NT>(p, i);
    }
    template <bool NT>
    __device__ __forceinline__ static void put(T* p, uint32_t i, R v, R, uint32_t)
    {
        st<NT>(p, i, v);
    }
};

struct StoreF16Dev {
    using T = _Float16;
    using R = float;
    static constexpr bool dev = true;
    static constexpr bool sr = false;
    __device__ __forceinline__ static R load(const T* p, uint32_t i, R rest) { return (float)at(p, i) + rest; }
    __device__ __forceinline__ static void store(T* p, uint32_t i, R v, R rest) { at(p, i) = (_Float16)(v - rest); }
    __device__ __forceinline__ static R round_trip(R v, R rest) { return (float)(_Float16)(v - rest) + rest; }
    template <bool NT>
    __device__ __forceinline__ static R get(const T* p, uint32_t i, R rest)
    {
        return (float)ld<NT>(p, i) + rest;
    }
    template <bool NT>
    __device__ __forceinline__ static void put(T* p, uint32_t i, R v, R rest, uint32_t)
    {
        st<NT>(p, i, (_Float16)(v - rest));
    }
};

// bfloat16 with round-to-nearest-even on store.
struct StoreBF16Dev {
    using T = uint16_t;
    using R = float;
    static constexpr bool dev = true;
    static constexpr bool sr = false;
    __device__ __forceinline__ static R load(const T* p, uint32_t i, R rest)
    {
        return __uint_as_float((uint32_t)at(p, i) << 16) + rest;
    }
    template <bool NT>
    __device__ __forceinline__ static R get(const T* p, uint32_t i, R rest)
    {
        return __uint_as_float((uint32_t)ld<NT>(p, i) << 16) + rest;
    }
    // round to nearest even in one instruction (v_cvt_pk_bf16_f32)
    __device__ __forceinline__ static uint16_t rne(R v, R rest)
    {
        const __bf16 h = (__bf16)(v - rest);
        return __builtin_bit_cast(uint16_t, h);
    }
    __device__ __forceinline__ static void store(T* p, uint32_t i, R v, R rest) { at(p, i) = rne(v, rest); }
    __device__ __forceinline__ static R round_trip(R v, R rest)
    {
        return __uint_as_float((uint32_t)rne(v, rest) << 16) + rest;
    }
    template <bool NT>
    __device__ __forceinline__ static void put(T* p, uint32_t i, R v, R rest, uint32_t)
    {
        st<NT>(p, i, rne(v, rest));
    }
};

typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Stochastic rounding (gfx950 v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32): the
// stored value is one of the two neighbours of g in the 16-bit format and is
// the upper one with probability (g - lower) / (upper - lower), so that
// E[stored] = g.  A rounding error that repeats from step to step on a slowly
// changing field then averages out instead of adding up.  For results in the
// normal range the conversion equals truncation of the FP32 bit pattern
// after adding the top 13 (FP16) or 16 (BF16) bits of rnd to the bits it
// drops.  Same layout in HBM as the round-to-nearest policies.
struct StoreF16DevSR : StoreF16Dev {
    static constexpr bool sr = true;
    template <bool NT>
    __device__ __forceinline__ static void put(T* p, uint32_t i, R v, R rest, uint32_t rnd)
    {
        // the conversion writes the low half and keeps the high half of its
        // first operand; passing the random word there saves a register copy
        f16x2_t h = __builtin_bit_cast(f16x2_t, rnd);
        h = __builtin_amdgcn_cvt_sr_f16_f32(h, v - rest, rnd, false);
        st<NT>(p, i, (_Float16)h[0]);
    }
};

struct StoreBF16DevSR : StoreBF16Dev {
    static constexpr bool sr = true;
    template <bool NT>
    __device__ __forceinline__ static void put(T* p, uint32_t i, R v, R rest, uint32_t rnd)
    {
        bf16x2_t h = __builtin_bit_cast(bf16x2_t, rnd);
        h = __builtin_amdgcn_cvt_sr_bf16_f32(h, v - rest, rnd, false);
        uint16_t b;
        __builtin_memcpy(&b, &h, 2);
        st<NT>(p, i, b);
    }
};

// Random bits of one cell and step for the stochastic-rounding stores: the
// "lowbias32" integer hash (C. Wellons, hash-prospector; xor-shift-multiply
// rounds) of the cell index xor the step seed; population q uses
// h + q * 0x9E3779B9 (a Weyl sequence, whose high bits the conversions read).
__device__ __forceinline__ uint32_t sr_hash(uint32_t i, uint32_t seed)
{
    uint32_t x = i ^ seed;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint32_t sr_bits(uint32_t h, int q) { return h + (uint32_t)q * 0x9E3779B9u; }

}  // namespace lbm

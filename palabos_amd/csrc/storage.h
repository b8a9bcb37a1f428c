// Population storage types of the engine: FP64, FP32, and 16-bit shifted
// populations (IEEE half or bfloat16) with FP32 arithmetic. Loads and stores
// carry the nontemporal hint: every population is touched once per step.
// store_raw_cached is an ordinary store (the staged odd step's z-moving
// outputs on lattices small enough for the last-level cache to keep them).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plbamd {

struct f16s { uint16_t b; };
struct bf16s { uint16_t b; };

template<typename S> struct Store;
template<> struct Store<double> {
    using T = double;
    using R = double;   // raw stored value
    static __device__ __forceinline__ R load_raw(const double* p) { return __builtin_nontemporal_load(p); }
    static __device__ __forceinline__ void store_raw(double* p, R v) { __builtin_nontemporal_store(v, p); }
    static __device__ __forceinline__ void store_raw_cached(double* p, R v) { *p = v; }
    static __device__ __forceinline__ T from_raw(R v) { return v; }
    static __device__ __forceinline__ R to_raw(T v) { return v; }
    static __device__ __forceinline__ double load(const double* p) { return __builtin_nontemporal_load(p); }
    static __device__ __forceinline__ void store(double* p, double v) { __builtin_nontemporal_store(v, p); }
    static __device__ __forceinline__ double round(double v) { return v; }
    static __device__ __forceinline__ uint16_t bits(double) { return 0; }
};
template<> struct Store<float> {
    using T = float;
    using R = float;
    static __device__ __forceinline__ R load_raw(const float* p) { return __builtin_nontemporal_load(p); }
    static __device__ __forceinline__ void store_raw(float* p, R v) { __builtin_nontemporal_store(v, p); }
    static __device__ __forceinline__ void store_raw_cached(float* p, R v) { *p = v; }
    static __device__ __forceinline__ T from_raw(R v) { return v; }
    static __device__ __forceinline__ R to_raw(T v) { return v; }
    static __device__ __forceinline__ float load(const float* p) { return __builtin_nontemporal_load(p); }
    static __device__ __forceinline__ void store(float* p, float v) { __builtin_nontemporal_store(v, p); }
    static __device__ __forceinline__ float round(float v) { return v; }
    static __device__ __forceinline__ uint16_t bits(float) { return 0; }
};
template<> struct Store<f16s> {
    using T = float;
    using R = uint16_t;
    static __device__ __forceinline__ R load_raw(const f16s* p) {
        return __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
    }
    static __device__ __forceinline__ void store_raw(f16s* p, R v) {
        __builtin_nontemporal_store(v, reinterpret_cast<uint16_t*>(p));
    }
    static __device__ __forceinline__ void store_raw_cached(f16s* p, R v) { *reinterpret_cast<uint16_t*>(p) = v; }
    static __device__ __forceinline__ T from_raw(R b) {
        _Float16 h;
        __builtin_memcpy(&h, &b, 2);
        return (float)h;
    }
    static __device__ __forceinline__ R to_raw(T v) { return bits(v); }
    static __device__ __forceinline__ float load(const f16s* p) {
        uint16_t b = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
        _Float16 h;
        __builtin_memcpy(&h, &b, 2);
        return (float)h;
    }
    static __device__ __forceinline__ void store(f16s* p, float v) {
        _Float16 h = (_Float16)v;  // round to nearest even
        uint16_t b;
        __builtin_memcpy(&b, &h, 2);
        __builtin_nontemporal_store(b, reinterpret_cast<uint16_t*>(p));
    }
    static __device__ __forceinline__ float round(float v) { return (float)(_Float16)v; }
    static __device__ __forceinline__ uint16_t bits(float v) {
        _Float16 h = (_Float16)v;
        uint16_t b;
        __builtin_memcpy(&b, &h, 2);
        return b;
    }
};
__device__ __forceinline__ uint16_t bf16_round(float v) {
    uint32_t x = __float_as_uint(v);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}
template<> struct Store<bf16s> {
    using T = float;
    using R = uint16_t;
    static __device__ __forceinline__ R load_raw(const bf16s* p) {
        return __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
    }
    static __device__ __forceinline__ void store_raw(bf16s* p, R v) {
        __builtin_nontemporal_store(v, reinterpret_cast<uint16_t*>(p));
    }
    static __device__ __forceinline__ void store_raw_cached(bf16s* p, R v) { *reinterpret_cast<uint16_t*>(p) = v; }
    static __device__ __forceinline__ T from_raw(R b) { return __uint_as_float((uint32_t)b << 16); }
    static __device__ __forceinline__ R to_raw(T v) { return bf16_round(v); }
    static __device__ __forceinline__ float load(const bf16s* p) {
        uint16_t b = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
        return __uint_as_float((uint32_t)b << 16);
    }
    static __device__ __forceinline__ void store(bf16s* p, float v) {
        __builtin_nontemporal_store(bf16_round(v), reinterpret_cast<uint16_t*>(p));
    }
    static __device__ __forceinline__ float round(float v) { return __uint_as_float((uint32_t)bf16_round(v) << 16); }
    static __device__ __forceinline__ uint16_t bits(float v) { return bf16_round(v); }
};

}  // namespace plbamd

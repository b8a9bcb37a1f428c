// Palabos checkpoint files on the device: the text of saveBinaryBlock /
// loadBinaryBlock (io/serializerIO_3D.cpp:37-66).
//
// A file is the base64 text of the data size (8-byte little-endian unsigned, or
// 4-byte with enforceUint; io/serializerIO.cpp Base64Writer::writeHeader),
// followed by the base64 text of the serialized cells (a separate base64 stream
// with its own '=' padding). The cells are the bounding box in C order (x, y,
// z; IndexOrdering::forward, multiBlock/multiBlockSerializer3D.cpp:71-105), each
// the raw T values of its 19 f-bar followed by its external scalars
// (core/cell.hh:103-111: the 3 force components for ForcedD3Q19Descriptor).
// The base64 alphabet and padding are RFC 4648's (io/base64.hh:36-38,104-114).
//
// Both conversions are byte streams at HBM speed: a thread takes 48 bytes and
// 64 characters in 128-bit accesses; the last partial group is one thread's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plbamd {

__host__ __device__ inline char b64_char(uint32_t v) {
    return (char)(v < 26 ? 'A' + v : (v < 52 ? 'a' + (v - 26) : (v < 62 ? '0' + (v - 52) : (v == 62 ? '+' : '/'))));
}

// Value of a base64 character, or 64 for a character outside the alphabet
// ('=' included).
__host__ __device__ inline uint32_t b64_value(uint32_t c) {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    return 64;
}

// Four characters of three bytes b0 b1 b2, packed little-endian in a word.
// The four 6-bit values sit one per byte and are mapped to the alphabet in
// one pass over the word: character = v + 'A', + 6 from 26 on ('a'), - 75
// from 52 on ('0'), - 15 from 62 on ('+'), + 3 from 63 on ('/'); every byte
// stays within 0..137 (no carry or borrow between bytes).
__host__ __device__ inline uint32_t b64_quad(uint32_t b0, uint32_t b1, uint32_t b2) {
    const uint32_t g = (b0 << 16) | (b1 << 8) | b2;
    const uint32_t v = (g >> 18) | (((g >> 12) & 63u) << 8) | (((g >> 6) & 63u) << 16) | ((g & 63u) << 24);
    constexpr uint32_t ones = 0x01010101u, high = 0x80808080u;
    const uint32_t ge26 = ((v + (128u - 26u) * ones) & high) >> 7;
    const uint32_t ge52 = ((v + (128u - 52u) * ones) & high) >> 7;
    const uint32_t ge62 = ((v + (128u - 62u) * ones) & high) >> 7;
    const uint32_t ge63 = ((v + (128u - 63u) * ones) & high) >> 7;
    return (v + 65u * ones + 6u * ge26 + 3u * ge63) - (75u * ge52 + 15u * ge62);
}

// Bytes per thread: 48 (three 128-bit loads) to 64 characters (four 128-bit
// stores); the last n % 48 bytes are one thread's.
constexpr uint64_t kB64Bytes = 48, kB64Chars = 64;

// in: n bytes (16-byte aligned), out: 4 * ceil(n / 3) characters (16-byte aligned).
static __global__ void __launch_bounds__(256) base64_encode(const uint8_t* __restrict__ in, uint64_t n, char* __restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t full = n / kB64Bytes;
    if (t < full) {
        const uint4* w = reinterpret_cast<const uint4*>(in + kB64Bytes * t);
        const uint4 a = w[0], b = w[1], c = w[2];
        const uint32_t words[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
        uint32_t q[16];
#pragma unroll
        for (int g = 0; g < 16; ++g) {
            uint32_t by[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int pos = 3 * g + k;
                by[k] = (words[pos / 4] >> (8 * (pos % 4))) & 255u;
            }
            q[g] = b64_quad(by[0], by[1], by[2]);
        }
        uint4* o = reinterpret_cast<uint4*>(out + kB64Chars * t);
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = make_uint4(q[4 * k], q[4 * k + 1], q[4 * k + 2], q[4 * k + 3]);
    } else if (t == full && n % kB64Bytes != 0) {
        // the last 1..47 bytes: whole groups, then a padded one
        const uint64_t base = kB64Bytes * full;
        const uint32_t rem = (uint32_t)(n - base);
        char* o = out + kB64Chars * full;
        uint32_t k = 0;
        for (; k + 3 <= rem; k += 3) {
            const uint32_t q = b64_quad(in[base + k], in[base + k + 1], in[base + k + 2]);
            for (int c = 0; c < 4; ++c) o[c] = (char)((q >> (8 * c)) & 255u);
            o += 4;
        }
        if (k < rem) {
            const uint32_t b0 = in[base + k], b1 = (k + 1 < rem) ? in[base + k + 1] : 0u;
            const uint32_t g = (b0 << 16) | (b1 << 8);
            o[0] = b64_char(g >> 18);
            o[1] = b64_char((g >> 12) & 63);
            o[2] = (k + 1 < rem) ? b64_char((g >> 6) & 63) : '=';
            o[3] = '=';
        }
    }
}

// in: 4 * ceil(n / 3) characters (16-byte aligned), out: n bytes (16-byte
// aligned). A character outside the alphabet where a value is expected sets
// *bad (padding is accepted only where the byte count ends).
static __global__ void __launch_bounds__(256) base64_decode(const char* __restrict__ in, uint64_t n, uint8_t* __restrict__ out,
                                                     int* bad) {
    // character -> value table (64: not in the alphabet)
    __shared__ uint8_t lut[256];
    for (uint32_t c = threadIdx.x; c < 256; c += blockDim.x) lut[c] = (uint8_t)b64_value(c);
    __syncthreads();
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t full = n / kB64Bytes;
    if (t < full) {
        const uint4* w = reinterpret_cast<const uint4*>(in + kB64Chars * t);
        uint32_t qs[16];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 v = w[k];
            qs[4 * k] = v.x; qs[4 * k + 1] = v.y; qs[4 * k + 2] = v.z; qs[4 * k + 3] = v.w;
        }
        uint32_t words[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t invalid = 0;
#pragma unroll
        for (int g = 0; g < 16; ++g) {
            uint32_t v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                v[c] = lut[(qs[g] >> (8 * c)) & 255u];
                invalid |= v[c] >> 6;
            }
            const uint32_t x = ((v[0] & 63) << 18) | ((v[1] & 63) << 12) | ((v[2] & 63) << 6) | (v[3] & 63);
            const uint32_t by[3] = {x >> 16, (x >> 8) & 255u, x & 255u};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int pos = 3 * g + k;
                words[pos / 4] |= by[k] << (8 * (pos % 4));
            }
        }
        uint4* o = reinterpret_cast<uint4*>(out + kB64Bytes * t);
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = make_uint4(words[4 * k], words[4 * k + 1], words[4 * k + 2], words[4 * k + 3]);
        if (invalid) atomicOr(bad, 1);
    } else if (t == full && n % kB64Bytes != 0) {
        const uint64_t base = kB64Bytes * full;
        const uint32_t rem = (uint32_t)(n - base);
        const char* s = in + kB64Chars * full;
        uint32_t invalid = 0;
        for (uint32_t k = 0; k < rem; k += 3, s += 4) {
            const uint32_t take = rem - k < 3 ? rem - k : 3;   // bytes of this group
            uint32_t v[4];
            for (int c = 0; c < 4; ++c) v[c] = b64_value((uint8_t)s[c]);
            // characters that carry bits: 2 for one byte, 3 for two, 4 for three
            for (uint32_t c = 0; c < take + 1; ++c) invalid |= v[c] >> 6;
            const uint32_t x = ((v[0] & 63) << 18) | ((v[1] & 63) << 12) | ((v[2] & 63) << 6) | (v[3] & 63);
            out[base + k] = (uint8_t)(x >> 16);
            if (take > 1) out[base + k + 1] = (uint8_t)((x >> 8) & 255u);
            if (take > 2) out[base + k + 2] = (uint8_t)(x & 255u);
        }
        if (invalid) atomicOr(bad, 1);
    }
}

// Cell records with the force externals: rec[c][0..18] = f-bar, rec[c][19..21] = force.
template<typename T>
__global__ void __launch_bounds__(256) append_force(const T* __restrict__ pops, T* __restrict__ rec, uint64_t cells,
                                                    T fx, T fy, T fz) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
#pragma unroll
    for (int i = 0; i < 19; ++i) rec[22 * c + i] = pops[19 * c + i];
    rec[22 * c + 19] = fx;
    rec[22 * c + 20] = fy;
    rec[22 * c + 21] = fz;
}

// The inverse; a cell whose force externals differ from (fx, fy, fz) sets *bad.
template<typename T>
__global__ void __launch_bounds__(256) split_force(const T* __restrict__ rec, T* __restrict__ pops, uint64_t cells,
                                                   T fx, T fy, T fz, int* bad) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
#pragma unroll
    for (int i = 0; i < 19; ++i) pops[19 * c + i] = rec[22 * c + i];
    if (!(rec[22 * c + 19] == fx && rec[22 * c + 20] == fy && rec[22 * c + 21] == fz)) atomicOr(bad, 1);
}

}  // namespace plbamd

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
// Both conversions are byte streams at HBM speed: a thread takes 12 bytes
// (three 32-bit words) and 16 characters (one 128-bit word); the last partial
// group is one thread's.
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
__host__ __device__ inline uint32_t b64_quad(uint32_t b0, uint32_t b1, uint32_t b2) {
    const uint32_t g = (b0 << 16) | (b1 << 8) | b2;
    return (uint32_t)(uint8_t)b64_char(g >> 18) | ((uint32_t)(uint8_t)b64_char((g >> 12) & 63) << 8) |
           ((uint32_t)(uint8_t)b64_char((g >> 6) & 63) << 16) | ((uint32_t)(uint8_t)b64_char(g & 63) << 24);
}

// in: n bytes (4-byte aligned), out: 4 * ceil(n / 3) characters (16-byte aligned).
__global__ void __launch_bounds__(256) base64_encode(const uint8_t* __restrict__ in, uint64_t n, char* __restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t full = n / 12;
    if (t < full) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(in + 12 * t);
        const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
        uint32_t b[12];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            b[k] = (w0 >> (8 * k)) & 255u;
            b[4 + k] = (w1 >> (8 * k)) & 255u;
            b[8 + k] = (w2 >> (8 * k)) & 255u;
        }
        uint4 q;
        q.x = b64_quad(b[0], b[1], b[2]);
        q.y = b64_quad(b[3], b[4], b[5]);
        q.z = b64_quad(b[6], b[7], b[8]);
        q.w = b64_quad(b[9], b[10], b[11]);
        *reinterpret_cast<uint4*>(out + 16 * t) = q;
    } else if (t == full && n % 12 != 0) {
        // the last 1..11 bytes: whole groups, then a padded one
        const uint64_t base = 12 * full;
        const uint32_t rem = (uint32_t)(n - base);
        char* o = out + 16 * full;
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

// in: 4 * ceil(n / 3) characters (16-byte aligned), out: n bytes (4-byte
// aligned). A character outside the alphabet where a value is expected sets
// *bad (padding is accepted only where the byte count ends).
__global__ void __launch_bounds__(256) base64_decode(const char* __restrict__ in, uint64_t n, uint8_t* __restrict__ out,
                                                     int* bad) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t full = n / 12;
    if (t < full) {
        const uint4 q = *reinterpret_cast<const uint4*>(in + 16 * t);
        const uint32_t qs[4] = {q.x, q.y, q.z, q.w};
        uint32_t bytes[12];
        uint32_t invalid = 0;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            uint32_t v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                v[c] = b64_value((qs[g] >> (8 * c)) & 255u);
                invalid |= v[c] >> 6;
            }
            const uint32_t x = ((v[0] & 63) << 18) | ((v[1] & 63) << 12) | ((v[2] & 63) << 6) | (v[3] & 63);
            bytes[3 * g] = x >> 16;
            bytes[3 * g + 1] = (x >> 8) & 255u;
            bytes[3 * g + 2] = x & 255u;
        }
        uint32_t* w = reinterpret_cast<uint32_t*>(out + 12 * t);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            w[k] = bytes[4 * k] | (bytes[4 * k + 1] << 8) | (bytes[4 * k + 2] << 16) | (bytes[4 * k + 3] << 24);
        if (invalid) atomicOr(bad, 1);
    } else if (t == full && n % 12 != 0) {
        const uint64_t base = 12 * full;
        const uint32_t rem = (uint32_t)(n - base);
        const char* s = in + 16 * full;
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

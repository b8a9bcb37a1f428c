// palabos_amd: plbamd_bandwidth_probe — the streaming rate this GPU sustains for
// the access patterns of the collide-and-stream kernels, with the same
// nontemporal loads and stores and no arithmetic. It is the practical ceiling
// the kernels' HBM fractions are read against:
//   PLBAMD_PROBE_COPY:  dst[i] = src[i], 16-byte elements, one per thread
//                       (bytes / 2 read, bytes / 2 written);
//   PLBAMD_PROBE_SLOTS: 19 slot arrays of bytes / 38 each, every thread's
//                       element of the 19 slots read and written back in
//                       place, the slots rotated by one: the data movement of
//                       an even step, 256-thread workgroups. 8-byte elements
//                       (one FP64 cell per thread); PLBAMD_PROBE_SLOTS_4B and
//                       PLBAMD_PROBE_SLOTS_16B use 4- and 16-byte elements
//                       (one FP32 cell, two FP64 cells per thread).
// Buffers are allocated and freed per call; the contents are not meaningful.
#include "engine.h"

namespace plbamd {

typedef double probe_v2d __attribute__((ext_vector_type(2)));

static __global__ void __launch_bounds__(256) probe_copy(const probe_v2d* __restrict__ src,
                                                         probe_v2d* __restrict__ dst, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

template<typename V>
static __global__ void __launch_bounds__(256) probe_slots(V* __restrict__ f, uint64_t slot, uint64_t cells) {
    const uint64_t c = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (c >= cells) return;
    V v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) v[i] = __builtin_nontemporal_load(f + (uint64_t)i * slot + c);
#pragma unroll
    for (int i = 0; i < 19; ++i) __builtin_nontemporal_store(v[(i + 1) % 19], f + (uint64_t)i * slot + c);
}

}  // namespace plbamd

using namespace plbamd;

extern "C" int plbamd_bandwidth_probe(int pattern, uint64_t bytes, int repeats, double* msBest, double* bytesMoved) {
    if (!msBest || !bytesMoved) return fail(PLBAMD_EINVAL, "bandwidth_probe: null output");
    if (repeats < 1 || repeats > 1000) return fail(PLBAMD_EINVAL, "bandwidth_probe: repeats in [1, 1000]");
    if (pattern < PLBAMD_PROBE_COPY || pattern > PLBAMD_PROBE_SLOTS_16B)
        return fail(PLBAMD_EINVAL, "bandwidth_probe: unknown pattern");
    const uint64_t elem = pattern == PLBAMD_PROBE_SLOTS_4B ? 4 : (pattern == PLBAMD_PROBE_SLOTS_16B ? 16 : 8);
    uint64_t n = 0, slot = 0;
    size_t alloc = 0;
    if (pattern == PLBAMD_PROBE_COPY) {
        n = bytes / 32;                       // 16-byte elements, read once and written once
        if (n == 0 || n > (uint64_t)INT32_MAX * 256u) return fail(PLBAMD_EINVAL, "bandwidth_probe: size out of range");
        alloc = (size_t)n * 16;
        *bytesMoved = 32.0 * (double)n;
    } else {
        n = bytes / (38 * elem);              // elements per slot: 19 read and 19 written per thread
        if (n == 0 || n > (uint64_t)INT32_MAX * 256u) return fail(PLBAMD_EINVAL, "bandwidth_probe: size out of range");
        slot = (n + 63) / 64 * 64;
        alloc = (size_t)slot * 19 * elem;
        *bytesMoved = 38.0 * (double)elem * (double)n;
    }
    void* a = nullptr;
    void* b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc(&a, alloc);
    if (err == hipSuccess && pattern == PLBAMD_PROBE_COPY) err = hipMalloc(&b, alloc);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) err = hipMemsetAsync(a, 0, alloc, s);
    if (err == hipSuccess && b) err = hipMemsetAsync(b, 0, alloc, s);
    const dim3 grid((unsigned)((n + 255) / 256));
    double best = 0;
    for (int r = 0; r <= repeats && err == hipSuccess; ++r) {   // r = 0: untimed first launch
        err = hipEventRecord(e0, s);
        if (err != hipSuccess) break;
        if (pattern == PLBAMD_PROBE_COPY)
            hipLaunchKernelGGL(probe_copy, grid, dim3(256), 0, s, static_cast<const probe_v2d*>(a),
                               static_cast<probe_v2d*>(b), n);
        else if (pattern == PLBAMD_PROBE_SLOTS)
            hipLaunchKernelGGL(probe_slots<double>, grid, dim3(256), 0, s, static_cast<double*>(a), slot, n);
        else if (pattern == PLBAMD_PROBE_SLOTS_4B)
            hipLaunchKernelGGL(probe_slots<uint32_t>, grid, dim3(256), 0, s, static_cast<uint32_t*>(a), slot, n);
        else
            hipLaunchKernelGGL(probe_slots<probe_v2d>, grid, dim3(256), 0, s, static_cast<probe_v2d*>(a), slot, n);
        err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(e1, s);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err == hipSuccess && r > 0 && (best == 0 || ms < best)) best = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("bandwidth_probe: ") + hipGetErrorString(err));
    *msBest = best;
    return PLBAMD_OK;
}

// palabos_amd: plbamd_reduce_moments_noalloc — the moment sums of
// plbamd_reduce_moments computed with no device allocation: the per-workgroup
// partial records go to a fixed file-scope device array instead of memory
// allocated per call.
//
// `scratch` below is a host-side stand-in for the device array: its device
// address on each GPU comes from hipGetSymbolAddress, taken once per device
// (plbamd_init, or the first call on the device) and kept in g_scratchPtr;
// only that address is ever handed to kernels.
#include "engine.h"

namespace plbamd {

// Partial records the array holds (8 doubles each). Any single-block lattice
// fits: each launch of the moments kernels has at most 65536 workgroups
// (gridFor), the bulk launch plus the plane launch at most 131072 records.
constexpr uint64_t kScratchRecords = 1u << 17;
constexpr uint64_t kScratchDoubles = (kScratchRecords + kMomentsLevel1 + 1) * 8;   // 8.02 MiB

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratchTable;            // guards g_scratchPtr
static double* g_scratchPtr[kMaxDevices] = {};
static std::mutex g_scratchUse;              // one call at a time uses the array

int scratchSetup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(PLBAMD_EINVAL, "scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchTable);
    if (g_scratchPtr[device]) return PLBAMD_OK;
    int current = -1;
    PLBAMD_HIP(hipGetDevice(&current));
    if (current != device) return fail(PLBAMD_ESTATE, "scratch: the device is not current");
    void* p = nullptr;
    PLBAMD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratchPtr[device] = static_cast<double*>(p);
    return PLBAMD_OK;
}

static double* scratchOf(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratchTable);
        if (g_scratchPtr[device]) return g_scratchPtr[device];
    }
    if (scratchSetup(device) != PLBAMD_OK) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratchTable);
    return g_scratchPtr[device];
}

}  // namespace plbamd

using namespace plbamd;

extern "C" int plbamd_reduce_moments_noalloc(int handle, const plbamd_box* box, double* out) {
    if (!out) return fail(PLBAMD_EINVAL, "reduce_moments_noalloc: null output");
    Lattice* L = find(handle);
    if (!L) return fail(PLBAMD_EINVAL, "reduce_moments_noalloc: unknown handle");
    if (!box || !boxValid(*L, *box)) return fail(PLBAMD_EINVAL, "reduce_moments_noalloc: box outside the lattice");
    double* base = scratchOf(L->device);
    if (!base)
        return fail(PLBAMD_ESTATE, "reduce_moments_noalloc: no scratch address for the lattice's device; call "
                                   "plbamd_init on it first");
    // every call on this device writes the same array: serialise them
    std::lock_guard<std::mutex> lock(g_scratchUse);
    const MomentsScratch s{base, kScratchRecords};
    switch (L->desc.storage) {
    case PLBAMD_F64: return moments<double>(*L, *box, nullptr, nullptr, out, nullptr, &s);
    case PLBAMD_F32: return moments<float>(*L, *box, nullptr, nullptr, out, nullptr, &s);
    case PLBAMD_F16: return moments<f16s>(*L, *box, nullptr, nullptr, out, nullptr, &s);
    default: return moments<bf16s>(*L, *box, nullptr, nullptr, out, nullptr, &s);
    }
}

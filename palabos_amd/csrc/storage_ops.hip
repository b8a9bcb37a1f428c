// palabos_amd: the operations of one storage type (PLBAMD_S) other than the
// step: envelope exchange, host transfers, checkpoint files, equilibrium
// initialisation and moments.
#include "engine.h"
#include "external.h"

#if !defined(PLBAMD_S)
#error "compile with -DPLBAMD_S=<storage type>"
#endif

namespace plbamd {

template<typename S>
int exchange(Lattice& L) {
    if (!L.multi || L.nSegs == 0) return PLBAMD_OK;
    hipLaunchKernelGGL((exchange_envelopes<S>), dim3(L.nTasks), dim3(256), 0, L.stream, L.arena, L.dGeom, L.dSegs,
                       L.dTasks);
    PLBAMD_HIP(hipGetLastError());
    return PLBAMD_OK;
}


// Populations of box b between the lattice and a device buffer of boxCells(b)
// * 19 elements of type elem, C order; launches only (no host copy, no wait).
template<typename S>
hipError_t transferDevice(Lattice& L, const plbamd_box& b, int elem, void* staging, bool up) {
    hipError_t err = hipSuccess;
    for (size_t bi = 0; bi < L.geom.size() && err == hipSuccess; ++bi) {
        BoxArgs a;
        if (!boxArgs(L, bi, b, a)) continue;
        const dim3 grid = gridFor((uint64_t)a.ilx * a.ily * a.ilz), block(256);
        if (up) {
            if (elem == PLBAMD_F64) hipLaunchKernelGGL((upload_box<S, double>), grid, block, 0, L.stream, L.arena, a, (const double*)staging);
            else if (elem == PLBAMD_F32) hipLaunchKernelGGL((upload_box<S, float>), grid, block, 0, L.stream, L.arena, a, (const float*)staging);
            else hipLaunchKernelGGL((upload_box_raw<S>), grid, block, 0, L.stream, L.arena, a, (const uint16_t*)staging);
        } else {
            if (elem == PLBAMD_F64) hipLaunchKernelGGL((download_box<S, double>), grid, block, 0, L.stream, L.arena, a, (double*)staging);
            else if (elem == PLBAMD_F32) hipLaunchKernelGGL((download_box<S, float>), grid, block, 0, L.stream, L.arena, a, (float*)staging);
            else hipLaunchKernelGGL((download_box_raw<S>), grid, block, 0, L.stream, L.arena, a, (uint16_t*)staging);
            err = hipGetLastError();
            if (err == hipSuccess && L.anyWrapped && L.nSpecial) {
                const dim3 g2 = gridFor(L.nSpecial);
                if (elem == PLBAMD_F64)
                    hipLaunchKernelGGL((complete_download<S, double>), g2, block, 0, L.stream, L.arena, a, L.dValues, (double*)staging, 0);
                else if (elem == PLBAMD_F32)
                    hipLaunchKernelGGL((complete_download<S, float>), g2, block, 0, L.stream, L.arena, a, L.dValues, (float*)staging, 0);
                else
                    hipLaunchKernelGGL((complete_download<S, float>), g2, block, 0, L.stream, L.arena, a, L.dValues, (float*)staging, 1);
            }
        }
        if (err == hipSuccess) err = hipGetLastError();
    }
    return err;
}

template<typename S>
int transfer(Lattice& L, const plbamd_box& b, int elem, void* data, size_t bytes, bool up) {
    const uint64_t n = boxCells(b);
    const size_t hostElem = (elem == PLBAMD_F64) ? 8 : (elem == PLBAMD_F32 ? 4 : 2);
    if (bytes != n * 19 * hostElem) return fail(PLBAMD_EINVAL, "transfer: byte count does not match the box");
    if ((elem == PLBAMD_F16 || elem == PLBAMD_BF16) && elem != L.desc.storage)
        return fail(PLBAMD_EINVAL, "transfer: 16-bit host data must match the storage type");
    if (up && L.multi && (L.time & 1))
        return fail(PLBAMD_ESTATE, "upload: a multi-block lattice accepts populations after an even number of steps");
    PLBAMD_TRY(rebuild(L));
    void* staging = nullptr;
    PLBAMD_HIP(hipMalloc(&staging, bytes));
    hipError_t err = hipSuccess;
    if (up) err = hipMemcpyAsync(staging, data, bytes, hipMemcpyHostToDevice, L.stream);
    if (err == hipSuccess) err = transferDevice<S>(L, b, elem, staging, up);
    if (err == hipSuccess && up && exchange<S>(L) != PLBAMD_OK) err = hipErrorUnknown;
    if (err == hipSuccess && !up) err = hipMemcpyAsync(data, staging, bytes, hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    hipFree(staging);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("transfer: ") + hipGetErrorString(err));
    return PLBAMD_OK;
}

// ---------------------------------------------------------------------------
// Palabos checkpoint files (saveBinaryBlock / loadBinaryBlock,
// io/serializerIO_3D.cpp:37-66): the cells are serialized and base64-coded on
// the device; the host writes or reads the text.

template<typename S, typename T>
int checkpointTyped(Lattice& L, const char* path, int elem, bool force, bool enforceUint, bool save) {
    const plbamd_box whole{0, L.desc.nx - 1, 0, L.desc.ny - 1, 0, L.desc.nz - 1};
    const uint64_t cells = boxCells(whole);
    const int recLen = force ? 22 : 19;
    const uint64_t dataBytes = cells * recLen * sizeof(T);
    const uint64_t nChars = (dataBytes + 2) / 3 * 4;
    if (enforceUint && dataBytes > 0xffffffffull)
        return fail(PLBAMD_EINVAL, "binary block: the data does not fit a 32-bit size header (enforce_uint)");
    if (dataBytes / kB64Bytes + 1 > (uint64_t)INT32_MAX * 256)
        return fail(PLBAMD_EINVAL, "binary block: lattice too large for one conversion launch");
    PLBAMD_TRY(rebuild(L));
    DeviceBuffers buf;
    T* pops = static_cast<T*>(buf.get(cells * 19 * sizeof(T)));
    T* rec = force ? static_cast<T*>(buf.get(dataBytes)) : pops;
    char* text = static_cast<char*>(buf.get((nChars + 63) / 64 * 64));
    int* bad = static_cast<int*>(buf.get(2 * sizeof(int)));
    if (!pops || !rec || !text || !bad) return fail(PLBAMD_ENOMEM, "binary block: cannot allocate device buffers");
    const T fx = (T)L.desc.force[0], fy = (T)L.desc.force[1], fz = (T)L.desc.force[2];
    if (save) {
        // cells in no block are zeros, as the reference's forward ordering writes them
        PLBAMD_HIP(hipMemsetAsync(pops, 0, cells * 19 * sizeof(T), L.stream));
        PLBAMD_HIP(transferDevice<S>(L, whole, elem, pops, false));
        if (force) hipLaunchKernelGGL((append_force<T>), itemsGrid(cells), dim3(256), 0, L.stream, pops, rec, cells, fx, fy, fz);
        hipLaunchKernelGGL(base64_encode, itemsGrid(dataBytes / kB64Bytes + 1), dim3(256), 0, L.stream,
                           reinterpret_cast<const uint8_t*>(rec), dataBytes, text);
        PLBAMD_HIP(hipGetLastError());
        std::vector<char> host(nChars);
        PLBAMD_HIP(hipMemcpyAsync(host.data(), text, nChars, hipMemcpyDeviceToHost, L.stream));
        PLBAMD_HIP(hipStreamSynchronize(L.stream));
        uint8_t header[8];
        const uint64_t size64 = dataBytes;
        std::memcpy(header, &size64, 8);   // little-endian host
        const std::string head = base64Host(header, enforceUint ? 4 : 8);
        FILE* fp = std::fopen(path, "wb");
        if (!fp) return fail(PLBAMD_EINVAL, std::string("save_binary_block: cannot open ") + path);
        const bool ok = std::fwrite(head.data(), 1, head.size(), fp) == head.size() &&
                        std::fwrite(host.data(), 1, host.size(), fp) == host.size();
        if (std::fclose(fp) != 0 || !ok) return fail(PLBAMD_EINVAL, std::string("save_binary_block: cannot write ") + path);
        return PLBAMD_OK;
    }
    if (L.multi && (L.time & 1))
        return fail(PLBAMD_ESTATE, "load_binary_block: a multi-block lattice accepts populations after an even number of steps");
    std::string file;
    {
        FILE* fp = std::fopen(path, "rb");
        if (!fp) return fail(PLBAMD_EINVAL, std::string("load_binary_block: cannot open ") + path);
        if (std::fseek(fp, 0, SEEK_END) == 0) {
            const long size = std::ftell(fp);
            if (size > 0) file.reserve((size_t)size);
            std::rewind(fp);
        }
        char chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) file.append(chunk, got);
        std::fclose(fp);
    }
    // the reference reads with operator>>, which skips white space (every
    // white-space character is at most ' '; every base64 character above it)
    if (std::any_of(file.begin(), file.end(), [](char c) { return (unsigned char)c <= ' '; }))
        file.erase(std::remove_if(file.begin(), file.end(), [](char c) { return std::isspace((unsigned char)c) != 0; }),
                   file.end());
    const size_t headChars = enforceUint ? 8 : 12;
    if (file.size() < headChars) return fail(PLBAMD_EINVAL, "load_binary_block: file shorter than its size header");
    uint8_t header[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!unbase64Host(file.data(), enforceUint ? 4 : 8, header))
        return fail(PLBAMD_EINVAL, "load_binary_block: the size header is not base64");
    uint64_t size64 = 0;
    std::memcpy(&size64, header, 8);
    if (size64 != dataBytes)
        return fail(PLBAMD_EINVAL, "load_binary_block: the file holds " + std::to_string(size64) +
                                       " bytes of cells, this lattice takes " + std::to_string(dataBytes));
    if (file.size() - headChars < nChars) return fail(PLBAMD_EINVAL, "load_binary_block: the file is truncated");
    PLBAMD_HIP(hipMemsetAsync(bad, 0, 2 * sizeof(int), L.stream));
    PLBAMD_HIP(hipMemcpyAsync(text, file.data() + headChars, nChars, hipMemcpyHostToDevice, L.stream));
    hipLaunchKernelGGL(base64_decode, itemsGrid(dataBytes / kB64Bytes + 1), dim3(256), 0, L.stream, text, dataBytes,
                       reinterpret_cast<uint8_t*>(rec), bad);
    if (force) hipLaunchKernelGGL((split_force<T>), itemsGrid(cells), dim3(256), 0, L.stream, rec, pops, cells, fx, fy, fz, bad + 1);
    PLBAMD_HIP(hipGetLastError());
    int flags[2] = {0, 0};
    PLBAMD_HIP(hipMemcpyAsync(flags, bad, sizeof(flags), hipMemcpyDeviceToHost, L.stream));
    PLBAMD_HIP(hipStreamSynchronize(L.stream));
    if (flags[0]) return fail(PLBAMD_EINVAL, "load_binary_block: the cell data is not base64");
    if (flags[1]) return fail(PLBAMD_EINVAL, "load_binary_block: the force externals differ from the lattice's force");
    PLBAMD_HIP(transferDevice<S>(L, whole, elem, pops, true));
    PLBAMD_TRY(exchange<S>(L));
    PLBAMD_HIP(hipStreamSynchronize(L.stream));
    return PLBAMD_OK;
}

template<typename S>
int checkpoint(Lattice& L, const char* path, int elem, bool force, bool enforceUint, bool save) {
    if (elem == PLBAMD_F64) return checkpointTyped<S, double>(L, path, elem, force, enforceUint, save);
    return checkpointTyped<S, float>(L, path, elem, force, enforceUint, save);
}

template<typename S>
int equilibrium(Lattice& L, const plbamd_box& b, const double* rho, const double* u) {
    const uint64_t n = boxCells(b);
    if (L.multi && (L.time & 1))
        return fail(PLBAMD_ESTATE, "initialize_at_equilibrium: a multi-block lattice accepts populations after an even number of steps");
    PLBAMD_TRY(rebuild(L));
    double* d = nullptr;
    PLBAMD_HIP(hipMalloc(&d, n * 4 * sizeof(double)));
    hipError_t err = hipMemcpyAsync(d, rho, n * sizeof(double), hipMemcpyHostToDevice, L.stream);
    if (err == hipSuccess) err = hipMemcpyAsync(d + n, u, 3 * n * sizeof(double), hipMemcpyHostToDevice, L.stream);
    for (size_t bi = 0; bi < L.geom.size() && err == hipSuccess; ++bi) {
        BoxArgs a;
        if (!boxArgs(L, bi, b, a)) continue;
        hipLaunchKernelGGL((equilibrium_box<S>), gridFor((uint64_t)a.ilx * a.ily * a.ilz), dim3(256), 0, L.stream,
                           L.arena, a, d, d + n);
        err = hipGetLastError();
    }
    if (err == hipSuccess && exchange<S>(L) != PLBAMD_OK) err = hipErrorUnknown;
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    hipFree(d);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("equilibrium: ") + hipGetErrorString(err));
    return PLBAMD_OK;
}

template<typename S>
int moments(Lattice& L, const plbamd_box& b, double* rho, double* u, double* sums, double* fields,
            const MomentsScratch* scratch) {
    const uint64_t n = boxCells(b);
    if (scratch) {
        if (rho || u || fields) return fail(PLBAMD_EINVAL, "moments: a fixed scratch takes the sums only");
        if (L.dirty)
            return fail(PLBAMD_ESTATE, "reduce_moments_noalloc: the lattice's tags, planes or porosity changed since "
                                       "its last step; run a step or plbamd_reduce_moments first");
    } else {
        PLBAMD_TRY(rebuild(L));
    }
    uint64_t nRecords = 0;
    const bool planes = L.nSpecial > 0 && L.dValues != nullptr;
    for (size_t bi = 0; bi < L.geom.size(); ++bi) {
        BoxArgs a;
        if (!boxArgs(L, bi, b, a)) continue;
        nRecords += gridFor((uint64_t)a.ilx * a.ily * a.ilz).x;
        if (planes) nRecords += gridFor(L.nSpecial).x;
    }
    const size_t fieldDoubles = (rho || u) ? n * 4 : 0;
    const size_t analysisDoubles = fields ? n * 7 : 0;
    // records reduced in two levels when there are many: kMomentsLevel1 workgroups, then one
    const uint32_t level1 = nRecords > 4096 ? kMomentsLevel1 : 0u;
    double* d = nullptr;
    double *dPart, *dLevel1;
    if (scratch) {
        if (nRecords > scratch->maxRecords)
            return fail(PLBAMD_EINVAL, "reduce_moments_noalloc: the box needs " + std::to_string(nRecords) +
                                           " partial records, the fixed scratch holds " +
                                           std::to_string(scratch->maxRecords));
        dPart = scratch->base;
        dLevel1 = dPart + 8 * scratch->maxRecords;
    } else {
        PLBAMD_HIP(hipMalloc(&d, (fieldDoubles + analysisDoubles + nRecords * 8 + 8 * level1 + 8) * sizeof(double)));
        dPart = d + fieldDoubles + analysisDoubles;
        dLevel1 = dPart + nRecords * 8;
    }
    double* dRho = rho ? d : nullptr;
    double* dU = u ? d + n : nullptr;
    double* dFields = fields ? d + fieldDoubles : nullptr;
    double* dSum = dLevel1 + 8 * (scratch ? kMomentsLevel1 : level1);
    MomentsArgs m{};
    m.tags = L.dTags;
    m.values = L.dValues;
    for (int c = 0; c < 3; ++c) m.force[c] = L.desc.force[c];
    m.forced = L.desc.force[0] != 0 || L.desc.force[1] != 0 || L.desc.force[2] != 0;
    m.fields = dFields;
    for (int k = 0; k < 3; ++k) m.solidRho[k] = L.solidRho[k];
    hipError_t err = hipSuccess;
    uint64_t rec = 0;
    for (size_t bi = 0; bi < L.geom.size(); ++bi) {
        BoxArgs a;
        if (!boxArgs(L, bi, b, a)) continue;
        const dim3 grid = gridFor((uint64_t)a.ilx * a.ily * a.ilz);
        if (a.parity)
            hipLaunchKernelGGL((moments_box<S, 1>), grid, dim3(256), 0, L.stream, L.arena, a, m, dRho, dU, dPart + rec * 8);
        else
            hipLaunchKernelGGL((moments_box<S, 0>), grid, dim3(256), 0, L.stream, L.arena, a, m, dRho, dU, dPart + rec * 8);
        rec += grid.x;
        err = hipGetLastError();
        if (err != hipSuccess) break;
        if (planes) {
            const dim3 pg = gridFor(L.nSpecial);
            hipLaunchKernelGGL((moments_planes<S>), pg, dim3(256), 0, L.stream, L.arena, a, m, dRho, dU, dPart + rec * 8);
            rec += pg.x;
            err = hipGetLastError();
            if (err != hipSuccess) break;
        }
    }
    if (err == hipSuccess) {
        if (level1) {
            hipLaunchKernelGGL(sum_records, dim3(level1), dim3(256), 0, L.stream, dPart, (uint32_t)nRecords, dLevel1);
            hipLaunchKernelGGL(sum_records, dim3(1), dim3(256), 0, L.stream, dLevel1, level1, dSum);
        } else {
            hipLaunchKernelGGL(sum_records, dim3(1), dim3(256), 0, L.stream, dPart, (uint32_t)nRecords, dSum);
        }
        err = hipGetLastError();
    }
    double host[8];
    if (err == hipSuccess && rho) err = hipMemcpyAsync(rho, dRho, n * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess && u) err = hipMemcpyAsync(u, dU, 3 * n * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess && fields)
        err = hipMemcpyAsync(fields, dFields, 7 * n * sizeof(double), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess) err = hipMemcpyAsync(host, dSum, sizeof(host), hipMemcpyDeviceToHost, L.stream);
    if (err == hipSuccess) err = hipStreamSynchronize(L.stream);
    if (d) hipFree(d);
    if (err != hipSuccess) return fail(PLBAMD_EHIP, std::string("moments: ") + hipGetErrorString(err));
    if (sums) {
        sums[0] = host[0];  // sum of rhoBar
        sums[1] = host[1];  // sum of |u|^2
        sums[2] = host[3];  // sum of u
        sums[3] = host[4];
        sums[4] = host[5];
    }
    return PLBAMD_OK;
}


template<typename S>
int launchProcessors(Lattice& L, int parity) {
    if (!L.dProc) return PLBAMD_OK;
    ProcArgs a{};
    a.arena = L.arena;
    a.blocks = L.dGeom;
    a.special = L.dSpecial;
    a.values = L.dValues;
    a.tags = L.dTags;
    a.walls = L.dWalls;
    a.cells = L.dProc;
    a.parity = parity;
    a.omega = L.desc.omega;
    for (int ph = 0; ph < 4; ++ph) {
        a.begin = L.procPhase[ph];
        a.end = L.procPhase[ph + 1];
        if (a.end == a.begin) continue;
        hipLaunchKernelGGL((boundary_processors<S>), dim3((a.end - a.begin + 255) / 256), dim3(256), 0, L.stream, a);
        PLBAMD_HIP(hipGetLastError());
    }
    return PLBAMD_OK;
}

// The external-moment field of a single-block lattice: op 0 fills it on box b
// from the populations (BoxRhoBarJfunctional3D), 1 copies its values on b to the
// host, 2 sets them from the host. Allocated on first use, [4][allocCells].
template<typename S>
int externalField(Lattice& L, const plbamd_box& b, double* rhoBar, double* j, int op) {
    using T = typename Store<S>::T;
    PLBAMD_TRY(rebuild(L));
    if (!L.dExtField) {
        PLBAMD_HIP(hipMalloc(&L.dExtField, 4 * L.allocCells * sizeof(T)));
        PLBAMD_HIP(hipMemsetAsync(L.dExtField, 0, 4 * L.allocCells * sizeof(T), L.stream));
    }
    T* field = static_cast<T*>(L.dExtField);
    BoxArgs a;
    if (!boxArgs(L, 0, b, a)) return fail(PLBAMD_EINVAL, "rhobar_j: box outside the lattice");
    const uint64_t n = boxCells(b);
    const dim3 grid = gridFor((uint64_t)a.ilx * a.ily * a.ilz);
    if (op == 0) {
        if (a.parity)
            hipLaunchKernelGGL((rhobar_j_box<S, 1>), grid, dim3(256), 0, L.stream, L.arena, a, L.dTags,
                               L.solidRho[0], field, L.allocCells);
        else
            hipLaunchKernelGGL((rhobar_j_box<S, 0>), grid, dim3(256), 0, L.stream, L.arena, a, L.dTags,
                               L.solidRho[0], field, L.allocCells);
        PLBAMD_HIP(hipGetLastError());
        PLBAMD_HIP(hipStreamSynchronize(L.stream));
        return PLBAMD_OK;
    }
    DeviceBuffers buf;
    double* d = static_cast<double*>(buf.get(4 * n * sizeof(double)));
    if (!d) return fail(PLBAMD_ENOMEM, "rhobar_j: staging allocation");
    if (op == 1) {
        hipLaunchKernelGGL((field_get<T>), grid, dim3(256), 0, L.stream, a, (const T*)field, L.allocCells, d, d + n);
        PLBAMD_HIP(hipGetLastError());
        PLBAMD_HIP(hipMemcpyAsync(rhoBar, d, n * sizeof(double), hipMemcpyDeviceToHost, L.stream));
        PLBAMD_HIP(hipMemcpyAsync(j, d + n, 3 * n * sizeof(double), hipMemcpyDeviceToHost, L.stream));
    } else {
        PLBAMD_HIP(hipMemcpyAsync(d, rhoBar, n * sizeof(double), hipMemcpyHostToDevice, L.stream));
        PLBAMD_HIP(hipMemcpyAsync(d + n, j, 3 * n * sizeof(double), hipMemcpyHostToDevice, L.stream));
        hipLaunchKernelGGL((field_set<T>), grid, dim3(256), 0, L.stream, a, field, L.allocCells,
                           (const double*)d, (const double*)(d + n));
        PLBAMD_HIP(hipGetLastError());
    }
    PLBAMD_HIP(hipStreamSynchronize(L.stream));
    return PLBAMD_OK;
}

// nsteps ExternalRhoJcollideAndStream3D steps with the field as it is.
template<typename S>
int externalSteps(Lattice& L, int nsteps) {
    using T = typename Store<S>::T;
    PLBAMD_TRY(rebuild(L));
    if (!L.dExtField) return fail(PLBAMD_ESTATE, "external_collide_and_stream: no rhoBar / j field (compute or set it first)");
    const BlockGeom& g = L.geom[0];
    const dim3 grid = itemsGrid((uint64_t)g.nx * g.ny * g.nz);
    const T* field = static_cast<const T*>(L.dExtField);
    const bool rbgk = L.desc.dynamics == PLBAMD_RBGK;
    for (int s = 0; s < nsteps; ++s) {
        const bool odd = (L.time & 1) != 0;
        if (rbgk) {
            if (odd) hipLaunchKernelGGL((external_step<S, 1, true>), grid, dim3(256), 0, L.stream, L.arena, g, L.dTags, field, L.desc.omega);
            else hipLaunchKernelGGL((external_step<S, 1, false>), grid, dim3(256), 0, L.stream, L.arena, g, L.dTags, field, L.desc.omega);
        } else {
            if (odd) hipLaunchKernelGGL((external_step<S, 0, true>), grid, dim3(256), 0, L.stream, L.arena, g, L.dTags, field, L.desc.omega);
            else hipLaunchKernelGGL((external_step<S, 0, false>), grid, dim3(256), 0, L.stream, L.arena, g, L.dTags, field, L.desc.omega);
        }
        PLBAMD_HIP(hipGetLastError());
        ++L.time;
    }
    L.statsValid = false;
    PLBAMD_HIP(hipStreamSynchronize(L.stream));
    return PLBAMD_OK;
}

template int externalField<PLBAMD_S>(Lattice&, const plbamd_box&, double*, double*, int);
template int externalSteps<PLBAMD_S>(Lattice&, int);
template int launchProcessors<PLBAMD_S>(Lattice&, int);
template int exchange<PLBAMD_S>(Lattice&);
template hipError_t transferDevice<PLBAMD_S>(Lattice&, const plbamd_box&, int, void*, bool);
template int transfer<PLBAMD_S>(Lattice&, const plbamd_box&, int, void*, size_t, bool);
template int checkpoint<PLBAMD_S>(Lattice&, const char*, int, bool, bool, bool);
template int equilibrium<PLBAMD_S>(Lattice&, const plbamd_box&, const double*, const double*);
template int moments<PLBAMD_S>(Lattice&, const plbamd_box&, double*, double*, double*, double*,
                                const MomentsScratch*);

}  // namespace plbamd

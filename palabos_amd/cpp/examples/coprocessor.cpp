// The co-processor plug-in path (coProcessors/coProcessor3D.h:59-186,
// coProcessorCommunication3D.hh:36-123): a domain is added, the host block's
// static data is sent, stepped on the device, and received back.
//
// usage: coprocessor nx ny nz omega steps in.bin out.bin [float|double]
// in.bin / out.bin hold T f-bar in the static cell layout [nx][ny][nz][19].
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "palabos_amd.hh"

using namespace plb;

template<typename T>
int run(plint nx, plint ny, plint nz, T omega, plint steps, const char* in, const char* out) {
    CoProcessor3D<T>& cp = global::defaultCoProcessor3D<T>();
    int handle = -1;
    if (!cp.addDomain(nx, ny, nz, omega, handle)) return 3;
    std::vector<char> data((size_t)(nx * ny * nz * 19) * sizeof(T));
    FILE* fi = std::fopen(in, "rb");
    if (!fi || std::fread(data.data(), 1, data.size(), fi) != data.size()) return 4;
    std::fclose(fi);
    Box3D all(0, nx - 1, 0, ny - 1, 0, nz - 1);
    if (!cp.send(handle, all, data)) return 5;
    for (plint i = 0; i < steps; ++i)
        if (!cp.collideAndStream(handle)) return 6;
    std::vector<char> result;
    if (!cp.receive(handle, all, result)) return 7;
    FILE* fo = std::fopen(out, "wb");
    std::fwrite(result.data(), 1, result.size(), fo);
    std::fclose(fo);
    // a sub-domain receive, as transferFromCoProcessors does per block
    std::vector<char> part;
    Box3D sub(1, nx - 2, 1, ny - 2, 1, nz - 2);
    if (!cp.receive(handle, sub, part) || part.size() != (size_t)sub.nCells() * 19 * sizeof(T)) return 8;
    std::printf("{\"ok\": true, \"handle\": %d}\n", handle);
    return 0;
}

int main(int argc, char* argv[]) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: coprocessor nx ny nz omega steps in.bin out.bin [float|double]\n");
        return 2;
    }
    plint nx = std::atol(argv[1]), ny = std::atol(argv[2]), nz = std::atol(argv[3]);
    double omega = std::atof(argv[4]);
    plint steps = std::atol(argv[5]);
    bool useFloat = argc > 8 && std::string(argv[8]) == "float";
    return useFloat ? run<float>(nx, ny, nz, (float)omega, steps, argv[6], argv[7])
                    : run<double>(nx, ny, nz, omega, steps, argv[6], argv[7]);
}

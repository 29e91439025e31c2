// Test driver of the .vdb reader (lightgraph_amd/host/vdb.cpp):
//   vdb_dump names FILE
//   vdb_dump grid FILE NAME OUT.raw [bx by bz dx dy dz]
//   vdb_dump volume FILE DENSITY.raw TEMPERATURE.raw
// prints one "key value ..." line per field and writes the dense float32
// arrays (z-y-x) raw; exit status 2 with the error on stderr when the reader
// throws.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "vdb.hpp"

using namespace lightgraph;

namespace {

void dump(const std::string &path, const std::vector<float> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    if (!v.empty()) std::fwrite(v.data(), sizeof(float), v.size(), f);
    std::fclose(f);
}

void print(const VdbFloatGrid &g) {
    std::printf("name %s\n", g.name.c_str());
    std::printf("voxel_size %.17g\n", g.voxel_size);
    std::printf("translation %.17g %.17g %.17g\n", g.translation[0], g.translation[1], g.translation[2]);
    std::printf("background %.9g\n", g.background);
    std::printf("empty %d\n", g.empty ? 1 : 0);
    std::printf("active_min %d %d %d\n", g.active_min[0], g.active_min[1], g.active_min[2]);
    std::printf("active_max %d %d %d\n", g.active_max[0], g.active_max[1], g.active_max[2]);
    std::printf("dense_min %d %d %d\n", g.dense_min[0], g.dense_min[1], g.dense_min[2]);
    std::printf("dims %d %d %d\n", g.dims[0], g.dims[1], g.dims[2]);
    std::printf("leaves %llu\n", (unsigned long long)g.leaves);
}

}  // namespace

int main(int argc, char **argv) {
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "names" && argc == 3) {
            for (const std::string &n : vdbGridNames(argv[2])) std::printf("grid %s\n", n.c_str());
        } else if (mode == "grid" && (argc == 5 || argc == 11)) {
            int32_t bmin[3], bdims[3];
            if (argc == 11)
                for (int a = 0; a < 3; ++a) {
                    bmin[a] = std::atoi(argv[5 + a]);
                    bdims[a] = std::atoi(argv[8 + a]);
                }
            const VdbFloatGrid g = readVdbFloatGrid(argv[2], argv[3], argc == 11 ? bmin : nullptr,
                                                    argc == 11 ? bdims : nullptr);
            print(g);
            dump(argv[4], g.dense);
        } else if (mode == "volume" && argc == 5) {
            const VdbVolume v = readVdbVolume(argv[2]);
            print(v.density);
            std::printf("has_temperature %d\n", v.has_temperature ? 1 : 0);
            std::printf("temperature_resampled %d\n", v.temperature_resampled ? 1 : 0);
            dump(argv[3], v.density.dense);
            dump(argv[4], v.temperature);
        } else {
            std::fprintf(stderr, "usage: vdb_dump names FILE | grid FILE NAME OUT [box] | volume FILE D T\n");
            return 1;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    return 0;
}

// Drives the host-only units of the library (STL parsing and the vertex weld,
// granule split, clipped volume fraction, RAW and TIFF volume decoders) over
// the files named on the command line, for the AddressSanitizer and
// UndefinedBehaviorSanitizer builds.  One line per input:
//   stl <path>                                -> nverts ntris ngranules vf xyz_sum
//   raw <folder> <w> <h> <bits> <signed> <be> -> width height depth bits fnv64
//   tiff <path>                               -> width height depth bits fnv64
// or "error <status> <message>" when the library refuses the input.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/s2hip.h"

static uint64_t fnv64(const uint8_t *p, uint64_t n) {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
    return h;
}

static void report_error(s2h_status st) { std::printf("error %d %s\n", (int)st, s2h_last_error()); }

static void run_stl(const char *path) {
    double *xyz = nullptr;
    uint32_t *tri = nullptr;
    uint64_t nv = 0, nt = 0;
    s2h_status st = s2h_stl_load(path, &xyz, &nv, &tri, &nt);
    if (st != S2H_OK) return report_error(st);
    std::vector<uint32_t> gid(nt ? nt : 1);
    uint32_t ng = 0;
    st = s2h_mesh_granules(tri, nt, nv, gid.data(), &ng);
    if (st != S2H_OK) {
        s2h_free(xyz);
        s2h_free(tri);
        return report_error(st);
    }
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1}, sum = 0.0;
    for (uint64_t v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a) {
            const double c = xyz[3 * v + a];
            if (v == 0 || c < lo[a]) lo[a] = c;
            if (v == 0 || c > hi[a]) hi[a] = c;
            sum += c;
        }
    double vf = 0.0;
    st = s2h_volume_fraction_in_bbox(xyz, nv, tri, nt, lo, hi, &vf);
    s2h_free(xyz);
    s2h_free(tri);
    if (st != S2H_OK) return report_error(st);
    std::printf("stl %" PRIu64 " %" PRIu64 " %u %.17g %.17g\n", nv, nt, ng, vf, sum);
}

static void run_volume(const s2h_volume_source &src) {
    s2h_volume_info info{};
    s2h_status st = s2h_volume_probe(&src, &info);
    if (st != S2H_OK) return report_error(st);
    const uint64_t n = info.width * info.height * info.depth * (info.bits / 8);
    std::vector<uint8_t> buf(n ? n : 1);
    st = s2h_volume_read(&src, buf.data(), n, &info);
    if (st != S2H_OK) return report_error(st);
    std::printf("vol %" PRIu64 " %" PRIu64 " %" PRIu64 " %u %" PRIu64 "\n", info.width, info.height, info.depth,
                info.bits, fnv64(buf.data(), n));
}

int main(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
        const std::string kind = argv[i];
        if (kind == "stl" && i + 1 < argc) {
            run_stl(argv[++i]);
        } else if (kind == "raw" && i + 6 < argc) {
            s2h_volume_source src{};
            src.path = argv[i + 1];
            src.kind = S2H_VOLUME_RAW;
            src.slice_start = src.slice_end = -1;
            src.width = std::strtoull(argv[i + 2], nullptr, 10);
            src.height = std::strtoull(argv[i + 3], nullptr, 10);
            src.bits = (uint32_t)std::strtoul(argv[i + 4], nullptr, 10);
            src.is_signed = std::atoi(argv[i + 5]);
            src.big_endian = std::atoi(argv[i + 6]);
            i += 6;
            run_volume(src);
        } else if (kind == "tiff" && i + 1 < argc) {
            s2h_volume_source src{};
            src.path = argv[++i];
            src.kind = S2H_VOLUME_TIFF;
            src.slice_start = src.slice_end = -1;
            run_volume(src);
        } else {
            std::fprintf(stderr, "bad argument %s\n", argv[i]);
            return 2;
        }
        std::fflush(stdout);
    }
    return 0;
}

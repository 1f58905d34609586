// Host-side mesh ingest for the S2 engine: STL load with the reference's weld
// and the granule (connected-component) split that defines per-particle parity.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/s2hip.h"

namespace s2h {
void set_error(const std::string &msg);
}

namespace {

struct Key {
    int64_t x, y, z;
    bool operator==(const Key &o) const { return x == o.x && y == o.y && z == o.z; }
};
struct KeyHash {
    size_t operator()(const Key &k) const {
        uint64_t h = (uint64_t)k.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)k.y + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.z + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
        return (size_t)(h ^ (h >> 31));
    }
};

// Rust `f as i64`: saturating, NaN -> 0.
int64_t f2i(double f) {
    if (std::isnan(f)) return 0;
    if (f >= 9.2233720368547758e18) return std::numeric_limits<int64_t>::max();
    if (f <= -9.2233720368547758e18) return std::numeric_limits<int64_t>::min();
    return (int64_t)f;
}

// src/io/stl.rs:55-80 quantize_key + dedup_vertex: key = round(v * 1e6) per axis
// (half away from zero), first occurrence keeps its index.
struct Welder {
    std::vector<double> xyz;
    std::unordered_map<Key, uint32_t, KeyHash> map;
    uint32_t add(double x, double y, double z) {
        Key k{f2i(std::round(x * 1000000.0)), f2i(std::round(y * 1000000.0)), f2i(std::round(z * 1000000.0))};
        auto it = map.find(k);
        if (it != map.end()) return it->second;
        uint32_t idx = (uint32_t)(xyz.size() / 3);
        xyz.push_back(x);
        xyz.push_back(y);
        xyz.push_back(z);
        map.emplace(k, idx);
        return idx;
    }
};

bool read_file(const char *path, std::vector<unsigned char> &out) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    unsigned char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
    bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

// src/io/stl.rs:201-212 looks_ascii_stl
bool looks_ascii(const std::vector<unsigned char> &b) {
    if (b.size() < 5 || std::memcmp(b.data(), "solid", 5) != 0) return false;
    std::string s(b.begin(), b.begin() + (long)std::min<size_t>(b.size(), 512));
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s.find("facet") != std::string::npos && s.find("vertex") != std::string::npos;
}

// src/io/stl.rs:157-189 parse_binary_reader: the face count at byte 80 and a
// file long enough for it (trailing bytes are ignored).
bool binary_count(const unsigned char *b, size_t n, uint32_t &count, std::string &err) {
    if (n < 84) {
        err = "Binary STL size mismatch";
        return false;
    }
    std::memcpy(&count, b + 80, 4);
    if (n < 84 + (uint64_t)count * 50) {
        err = "Binary STL size mismatch";
        return false;
    }
    return true;
}

// f32 little-endian vertices widened to f64, three corners per face.
template <class Add>
bool parse_binary(const std::vector<unsigned char> &b, Add &&add, std::string &err) {
    uint32_t count;
    if (!binary_count(b.data(), b.size(), count, err)) return false;
    for (uint32_t f = 0; f < count; ++f) {
        const unsigned char *r = b.data() + 84 + (size_t)f * 50;
        for (int k = 0; k < 3; ++k) {
            float v[3];
            std::memcpy(v, r + 12 + 12 * k, 12);
            add((double)v[0], (double)v[1], (double)v[2]);
        }
    }
    return true;
}

// src/io/stl.rs:82-153: line-by-line "vertex x y z" (exactly four tokens), a face
// per three vertices; no triangle -> binary parse of the same bytes.
template <class Add>
bool parse_ascii(const std::vector<unsigned char> &b, Add &&add) {
    size_t pos = 0, corners = 0;
    double pending[9];
    int np = 0;
    while (pos < b.size()) {
        size_t e = pos;
        while (e < b.size() && b[e] != '\n') ++e;
        std::string line(b.begin() + (long)pos, b.begin() + (long)e);
        pos = e + 1;
        char tok[4][128];
        int n = 0;
        const char *p = line.c_str();
        while (*p && n < 5) {
            while (*p && std::isspace((unsigned char)*p)) ++p;
            if (!*p) break;
            const char *s = p;
            while (*p && !std::isspace((unsigned char)*p)) ++p;
            if (n < 4) {
                size_t len = std::min<size_t>((size_t)(p - s), 127);
                std::memcpy(tok[n], s, len);
                tok[n][len] = 0;
            }
            ++n;
        }
        if (n != 4 || std::strcmp(tok[0], "vertex") != 0) continue;
        double v[3];
        bool ok = true;
        for (int k = 0; k < 3; ++k) {
            char *end = nullptr;
            v[k] = std::strtod(tok[k + 1], &end);
            if (end == tok[k + 1] || *end) ok = false;
        }
        if (!ok) continue;
        for (int k = 0; k < 3; ++k) pending[3 * np + k] = v[k];
        if (++np == 3) {
            for (int k = 0; k < 3; ++k) add(pending[3 * k], pending[3 * k + 1], pending[3 * k + 2]);
            corners += 3;
            np = 0;
        }
    }
    return corners > 0;
}

}  // namespace

namespace s2h {
namespace stlio {
bool read_file(const char *path, std::vector<unsigned char> &out) { return ::read_file(path, out); }
bool looks_ascii(const std::vector<unsigned char> &b) { return ::looks_ascii(b); }
bool ascii_corners(const std::vector<unsigned char> &b, std::vector<double> &xyz) {
    xyz.clear();
    return ::parse_ascii(b, [&](double x, double y, double z) {
        xyz.push_back(x);
        xyz.push_back(y);
        xyz.push_back(z);
    });
}
bool binary_count(const unsigned char *b, size_t n, uint32_t &count, std::string &err) {
    return ::binary_count(b, n, count, err);
}
}  // namespace stlio
}  // namespace s2h

extern "C" {

s2h_status s2h_stl_load(const char *path, double **xyz, uint64_t *nverts, uint32_t **tri, uint64_t *ntris) {
    if (!path || !xyz || !nverts || !tri || !ntris) {
        s2h::set_error("null argument");
        return S2H_ERR_ARG;
    }
    std::vector<unsigned char> bytes;
    if (!read_file(path, bytes)) {
        s2h::set_error(std::string("cannot read ") + path);
        return S2H_ERR_IO;
    }
    Welder w;
    std::vector<uint32_t> t;
    std::string err;
    bool ok = false;
    auto add = [&](double x, double y, double z) { t.push_back(w.add(x, y, z)); };
    if (looks_ascii(bytes)) ok = parse_ascii(bytes, add);
    if (!ok) {
        w = Welder();
        t.clear();
        if (!parse_binary(bytes, add, err)) {
            s2h::set_error(err + " in " + path);
            return S2H_ERR_IO;
        }
    }
    *nverts = w.xyz.size() / 3;
    *ntris = t.size() / 3;
    *xyz = (double *)std::malloc(std::max<size_t>(w.xyz.size(), 1) * sizeof(double));
    *tri = (uint32_t *)std::malloc(std::max<size_t>(t.size(), 1) * sizeof(uint32_t));
    if (!*xyz || !*tri) {
        std::free(*xyz);
        std::free(*tri);
        s2h::set_error("out of host memory");
        return S2H_ERR_IO;
    }
    std::memcpy(*xyz, w.xyz.data(), w.xyz.size() * sizeof(double));
    std::memcpy(*tri, t.data(), t.size() * sizeof(uint32_t));
    return S2H_OK;
}

void s2h_free(void *p) { std::free(p); }

// src/geometry/mesh_ops.rs:48-~120: components over shared vertex indices,
// numbered in the order their BFS starts (the smallest unvisited face index).
s2h_status s2h_mesh_granules(const uint32_t *tri, uint64_t ntris, uint64_t nverts, uint32_t *granule_of_tri,
                             uint32_t *ngranules) {
    if ((ntris && (!tri || !granule_of_tri)) || !ngranules) {
        s2h::set_error("null argument");
        return S2H_ERR_ARG;
    }
    *ngranules = 0;
    if (ntris == 0 || nverts == 0) return S2H_OK;
    for (uint64_t i = 0; i < 3 * ntris; ++i)
        if (tri[i] >= nverts) {
            s2h::set_error("vertex index out of range");
            return S2H_ERR_ARG;
        }
    // union-find over vertices, then components labelled by their smallest face
    std::vector<uint32_t> parent(nverts);
    for (uint64_t v = 0; v < nverts; ++v) parent[v] = (uint32_t)v;
    auto find = [&](uint32_t v) {
        while (parent[v] != v) {
            parent[v] = parent[parent[v]];
            v = parent[v];
        }
        return v;
    };
    for (uint64_t f = 0; f < ntris; ++f) {
        uint32_t a = find(tri[3 * f]);
        for (int k = 1; k < 3; ++k) {
            uint32_t b = find(tri[3 * f + k]);
            if (a != b) {
                if (a < b) parent[b] = a;
                else {
                    parent[a] = b;
                    a = b;
                }
            }
        }
    }
    std::vector<uint32_t> label(nverts, UINT32_MAX);
    uint32_t n = 0;
    for (uint64_t f = 0; f < ntris; ++f) {
        uint32_t r = find(tri[3 * f]);
        if (label[r] == UINT32_MAX) label[r] = n++;
        granule_of_tri[f] = label[r];
    }
    *ngranules = n;
    return S2H_OK;
}

}  // extern "C"

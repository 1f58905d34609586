// Geometric volume fraction of a mesh inside a measurement box (host).
//
// The measure report's "Volume Fraction:" and S2(0) of the continuous mesh
// Monte Carlo are the in-box volume of every granule over the box volume,
// volume_fraction_in_bbox (src/geometry/volume.rs:437-469).  A granule lying
// wholly inside the box counts its closed-mesh volume; one crossing the box is
// clipped by the six box planes, each cut capped with a triangulated polygon so
// the piece stays closed, and the piece's volume is counted.  The steps follow
// the reference one for one (f64, no contraction: this unit is compiled with
// -ffp-contract=off), because the printed fraction and the S2(0)-VF difference
// must carry the same bits:
//   granules           split_mesh_into_granules (src/geometry/mesh_ops.rs:48-122):
//                      BFS over shared vertices from the lowest unvisited face,
//                      faces in queue order, vertices renumbered on first use
//   closed volume      mesh_volume (src/geometry/volume.rs:6-15): |sum a.(b x c)/6|
//                      in face order, dot as (x x' + y y') + z z'
//   one plane          clip_mesh_by_plane_with_cap (src/geometry/volume.rs:352-400):
//                      Sutherland-Hodgman per triangle with eps 1e-9, fan
//                      triangulation, the triangle/plane segments, and the cap
//                      (triangulate_cap_from_segments, 211-350): points welded by
//                      the 1e-6 key, edges in ascending (i, j) order, rings walked
//                      from each unused edge, each ring fanned from its vertex
//                      mean in atan2 order, wound by its area sign
//   box                clip_mesh_by_bbox (403-421): planes x-min, x-max, y-min,
//                      y-max, z-min, z-max in turn, stopping once empty
//   per granule        particle_volume_in_bbox (424-430)
//   fraction           granule volumes summed in granule order, over
//                      max(box volume, 1e-12), clamped to [0, 1]
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/s2hip.h"

namespace s2h {
s2h_status fail(s2h_status code, const std::string &msg);
}

namespace {

struct P3 {
    double x, y, z;
};
inline P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 scale(P3 a, double f) { return {a.x * f, a.y * f, a.z * f}; }
inline double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct Tri {
    size_t a, b, c;
};
struct Mesh {
    std::vector<P3> v;
    std::vector<Tri> f;
    bool empty() const { return v.empty() || f.empty(); }
};

double mesh_volume(const Mesh &m) {
    double total = 0.0;
    for (const Tri &t : m.f) total += dot(m.v[t.a], cross(m.v[t.b], m.v[t.c])) / 6.0;
    return std::fabs(total);
}

void append(Mesh &dst, const Mesh &src) {
    const size_t off = dst.v.size();
    dst.v.insert(dst.v.end(), src.v.begin(), src.v.end());
    for (const Tri &t : src.f) dst.f.push_back({t.a + off, t.b + off, t.c + off});
}

// Connected components over shared vertex indices, the reference's order.
std::vector<Mesh> granules(const Mesh &m) {
    std::vector<Mesh> parts;
    if (m.f.empty() || m.v.empty()) return parts;
    const size_t nv = m.v.size(), nf = m.f.size();
    std::vector<size_t> start(nv + 1, 0);
    for (const Tri &t : m.f) {
        ++start[t.a + 1];
        ++start[t.b + 1];
        ++start[t.c + 1];
    }
    for (size_t i = 0; i < nv; ++i) start[i + 1] += start[i];
    std::vector<size_t> fill(start.begin(), start.end() - 1), adj(start[nv]);
    for (size_t fi = 0; fi < nf; ++fi)
        for (size_t vi : {m.f[fi].a, m.f[fi].b, m.f[fi].c}) adj[fill[vi]++] = fi;
    std::vector<char> seen(nf, 0);
    std::vector<size_t> owner(nv, SIZE_MAX), local(nv, 0), queue;
    for (size_t s = 0; s < nf; ++s) {
        if (seen[s]) continue;
        queue.clear();
        seen[s] = 1;
        queue.push_back(s);
        for (size_t h = 0; h < queue.size(); ++h) {
            const Tri &t = m.f[queue[h]];
            for (size_t vi : {t.a, t.b, t.c})
                for (size_t k = start[vi]; k < start[vi + 1]; ++k)
                    if (!seen[adj[k]]) {
                        seen[adj[k]] = 1;
                        queue.push_back(adj[k]);
                    }
        }
        const size_t id = parts.size();
        Mesh part;
        part.f.reserve(queue.size());
        auto remap = [&](size_t g) {
            if (owner[g] != id) {
                owner[g] = id;
                local[g] = part.v.size();
                part.v.push_back(m.v[g]);
            }
            return local[g];
        };
        for (size_t fi : queue) {
            const size_t a = remap(m.f[fi].a);
            const size_t b = remap(m.f[fi].b);
            const size_t c = remap(m.f[fi].c);
            part.f.push_back({a, b, c});
        }
        parts.push_back(std::move(part));
    }
    return parts;
}

inline double plane_dist(P3 p, P3 origin, P3 normal) { return dot(sub(p, origin), normal); }

P3 cut_point(P3 a, P3 b, double da, double db) {
    const double denom = da - db;
    if (std::fabs(denom) <= 1e-12) return a;
    const double t = std::min(std::max(da / denom, 0.0), 1.0);
    return add(a, scale(sub(b, a), t));
}

std::vector<P3> clip_polygon(const std::vector<P3> &poly, P3 origin, P3 normal, double eps) {
    std::vector<P3> out;
    if (poly.empty()) return out;
    for (size_t i = 0; i < poly.size(); ++i) {
        const P3 c = poly[i], n = poly[(i + 1) % poly.size()];
        const double dc = plane_dist(c, origin, normal), dn = plane_dist(n, origin, normal);
        const bool in_c = dc >= -eps, in_n = dn >= -eps;
        if (in_c && in_n) {
            out.push_back(n);
        } else if (in_c) {
            out.push_back(cut_point(c, n, dc, dn));
        } else if (in_n) {
            out.push_back(cut_point(c, n, dc, dn));
            out.push_back(n);
        }
    }
    if (out.size() <= 2) return out;
    std::vector<P3> kept;
    kept.reserve(out.size());
    for (const P3 &p : out) {
        if (!kept.empty()) {
            const P3 d = sub(p, kept.back());
            if (!(dot(d, d) > 1e-20)) continue;
        }
        kept.push_back(p);
    }
    if (kept.size() > 1) {
        const P3 d = sub(kept.front(), kept.back());
        if (dot(d, d) <= 1e-20) kept.pop_back();
    }
    return kept;
}

// Round half away from zero (f64::round) then the saturating `as i64`.
inline int64_t key_coord(double v) {
    const double r = std::round(v * 1000000.0);
    if (std::isnan(r)) return 0;
    if (r >= 9223372036854775807.0) return INT64_MAX;
    if (r <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)r;
}

struct KeyHash {
    size_t operator()(const std::tuple<int64_t, int64_t, int64_t> &k) const {
        const uint64_t a = (uint64_t)std::get<0>(k), b = (uint64_t)std::get<1>(k), c = (uint64_t)std::get<2>(k);
        return (size_t)(a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full ^
                        (c << 1));
    }
};

bool triangle_segment(const P3 (&tri)[3], P3 origin, P3 normal, double eps, P3 &s0, P3 &s1) {
    std::vector<P3> pts;
    static const int edges[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (const auto &e : edges) {
        const P3 a = tri[e[0]], b = tri[e[1]];
        const double da = plane_dist(a, origin, normal), db = plane_dist(b, origin, normal);
        if (std::fabs(da) <= eps) pts.push_back(a);
        if ((da > eps && db < -eps) || (da < -eps && db > eps)) pts.push_back(cut_point(a, b, da, db));
    }
    std::vector<P3> uniq;
    for (const P3 &p : pts) {
        bool dup = false;
        for (const P3 &q : uniq) {
            const P3 d = sub(p, q);
            if (dot(d, d) <= 1e-16) {
                dup = true;
                break;
            }
        }
        if (!dup) uniq.push_back(p);
    }
    if (uniq.size() < 2) return false;
    s0 = uniq[0];
    s1 = uniq[1];
    return true;
}

void plane_axes(P3 normal, P3 &u, P3 &v) {
    const P3 tangent = std::fabs(normal.x) < 0.5 ? P3{1.0, 0.0, 0.0} : P3{0.0, 1.0, 0.0};
    const P3 ur = cross(normal, tangent);
    const double un = std::sqrt(dot(ur, ur));
    u = un > 1e-12 ? scale(ur, 1.0 / un) : P3{0.0, 0.0, 1.0};
    const P3 vr = cross(normal, u);
    const double vn = std::sqrt(dot(vr, vr));
    v = vn > 1e-12 ? scale(vr, 1.0 / vn) : P3{1.0, 0.0, 0.0};
}

Mesh cap_from_segments(const std::vector<std::pair<P3, P3>> &segs, P3 normal) {
    Mesh cap;
    if (segs.empty()) return cap;
    std::vector<P3> pts;
    std::unordered_map<std::tuple<int64_t, int64_t, int64_t>, size_t, KeyHash> ids;
    std::unordered_map<size_t, std::vector<size_t>> nbr;
    std::set<std::pair<size_t, size_t>> edges;
    auto point_id = [&](P3 p) {
        const auto k = std::make_tuple(key_coord(p.x), key_coord(p.y), key_coord(p.z));
        auto it = ids.find(k);
        if (it != ids.end()) return it->second;
        const size_t id = pts.size();
        pts.push_back(p);
        ids.emplace(k, id);
        return id;
    };
    for (const auto &s : segs) {
        const size_t ia = point_id(s.first), ib = point_id(s.second);
        if (ia == ib) continue;
        if (edges.insert(ia < ib ? std::make_pair(ia, ib) : std::make_pair(ib, ia)).second) {
            nbr[ia].push_back(ib);
            nbr[ib].push_back(ia);
        }
    }
    std::set<std::pair<size_t, size_t>> used;
    P3 ax_u, ax_v;
    plane_axes(normal, ax_u, ax_v);
    for (const auto &e : edges) {
        const size_t a = e.first, b = e.second;
        if (used.count(e)) continue;
        std::vector<size_t> ring{a, b};
        used.insert(e);
        size_t prev = a, cur = b;
        while (cur != a) {
            auto it = nbr.find(cur);
            if (it == nbr.end()) break;
            bool found = false;
            size_t next = 0;
            for (size_t n : it->second) {
                if (n == prev) continue;
                if (!used.count(cur < n ? std::make_pair(cur, n) : std::make_pair(n, cur))) {
                    next = n;
                    found = true;
                    break;
                }
            }
            if (!found) break;
            used.insert(cur < next ? std::make_pair(cur, next) : std::make_pair(next, cur));
            ring.push_back(next);
            prev = cur;
            cur = next;
        }
        if (ring.size() < 4 || ring.back() != a) continue;
        ring.pop_back();
        if (ring.size() < 3) continue;
        P3 center{0.0, 0.0, 0.0};
        for (size_t i : ring) center = add(center, pts[i]);
        center = scale(center, 1.0 / (double)ring.size());
        std::vector<size_t> order(ring);
        std::vector<double> ang(pts.size(), 0.0);
        for (size_t i : ring) {
            const P3 d = sub(pts[i], center);
            ang[i] = std::atan2(dot(d, ax_v), dot(d, ax_u));
        }
        std::stable_sort(order.begin(), order.end(), [&](size_t i, size_t j) { return ang[i] < ang[j]; });
        const size_t ci = cap.v.size();
        cap.v.push_back(center);
        std::vector<size_t> ridx;
        ridx.reserve(order.size());
        for (size_t pid : order) {
            ridx.push_back(cap.v.size());
            cap.v.push_back(pts[pid]);
        }
        double area = 0.0;
        for (size_t i = 0; i < order.size(); ++i) {
            const P3 p = sub(pts[order[i]], center), q = sub(pts[order[(i + 1) % order.size()]], center);
            area += dot(cross(p, q), normal);
        }
        const bool ccw = area >= 0.0;
        for (size_t i = 0; i < ridx.size(); ++i) {
            const size_t p = ridx[i], q = ridx[(i + 1) % ridx.size()];
            if (ccw) cap.f.push_back({ci, q, p});
            else cap.f.push_back({ci, p, q});
        }
    }
    return cap;
}

Mesh clip_plane_capped(Mesh mesh, P3 origin, P3 normal) {
    const double eps = 1e-9;
    bool inside = false, outside = false;
    for (const Tri &t : mesh.f) {
        for (size_t vi : {t.a, t.b, t.c}) {
            if (plane_dist(mesh.v[vi], origin, normal) >= -eps) inside = true;
            else outside = true;
            if (inside && outside) break;
        }
        if (inside && outside) break;
    }
    if (!outside) return mesh;
    if (!inside) return Mesh{};
    Mesh body;
    std::vector<std::pair<P3, P3>> segs;
    for (const Tri &t : mesh.f) {
        const P3 tri[3] = {mesh.v[t.a], mesh.v[t.b], mesh.v[t.c]};
        const std::vector<P3> poly = clip_polygon({tri[0], tri[1], tri[2]}, origin, normal, eps);
        if (poly.size() >= 3) {
            const size_t base = body.v.size();
            body.v.insert(body.v.end(), poly.begin(), poly.end());
            for (size_t i = 1; i + 1 < poly.size(); ++i) body.f.push_back({base, base + i, base + i + 1});
        }
        P3 s0, s1;
        if (triangle_segment(tri, origin, normal, eps, s0, s1)) segs.emplace_back(s0, s1);
    }
    if (segs.empty()) return body;
    append(body, cap_from_segments(segs, normal));
    return body;
}

Mesh clip_box(const Mesh &mesh, const double mn[3], const double mx[3]) {
    const std::pair<P3, P3> planes[6] = {
        {{mn[0], 0.0, 0.0}, {1.0, 0.0, 0.0}},  {{mx[0], 0.0, 0.0}, {-1.0, 0.0, 0.0}},
        {{0.0, mn[1], 0.0}, {0.0, 1.0, 0.0}},  {{0.0, mx[1], 0.0}, {0.0, -1.0, 0.0}},
        {{0.0, 0.0, mn[2]}, {0.0, 0.0, 1.0}},  {{0.0, 0.0, mx[2]}, {0.0, 0.0, -1.0}},
    };
    Mesh out = mesh;
    for (const auto &pl : planes) {
        out = clip_plane_capped(std::move(out), pl.first, pl.second);
        if (out.empty()) break;
    }
    return out;
}

double volume_in_box(const Mesh &m, const double mn[3], const double mx[3]) {
    bool contained = true;
    for (const P3 &p : m.v)
        if (!(p.x >= mn[0] && p.x <= mx[0] && p.y >= mn[1] && p.y <= mx[1] && p.z >= mn[2] && p.z <= mx[2])) {
            contained = false;
            break;
        }
    if (contained) return mesh_volume(m);
    return mesh_volume(clip_box(m, mn, mx));
}

s2h_status to_mesh(const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris, Mesh &m) {
    if ((!xyz && nverts) || (!tri && ntris)) return s2h::fail(S2H_ERR_ARG, "null mesh array");
    m.v.resize(nverts);
    for (uint64_t i = 0; i < nverts; ++i) m.v[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    m.f.resize(ntris);
    for (uint64_t i = 0; i < ntris; ++i) {
        const uint32_t a = tri[3 * i], b = tri[3 * i + 1], c = tri[3 * i + 2];
        if (a >= nverts || b >= nverts || c >= nverts) return s2h::fail(S2H_ERR_ARG, "face index out of range");
        m.f[i] = {a, b, c};
    }
    return S2H_OK;
}

}  // namespace

extern "C" {

s2h_status s2h_mesh_volume_in_bbox(const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                                   const double bbox_min[3], const double bbox_max[3], double *volume) {
    if (!bbox_min || !bbox_max || !volume) return s2h::fail(S2H_ERR_ARG, "null argument");
    Mesh m;
    const s2h_status st = to_mesh(xyz, nverts, tri, ntris, m);
    if (st != S2H_OK) return st;
    *volume = volume_in_box(m, bbox_min, bbox_max);
    return S2H_OK;
}

s2h_status s2h_volume_fraction_in_bbox(const double *xyz, uint64_t nverts, const uint32_t *tri, uint64_t ntris,
                                       const double bbox_min[3], const double bbox_max[3], double *vf) {
    if (!bbox_min || !bbox_max || !vf) return s2h::fail(S2H_ERR_ARG, "null argument");
    Mesh m;
    const s2h_status st = to_mesh(xyz, nverts, tri, ntris, m);
    if (st != S2H_OK) return st;
    const double box = std::max(std::max(bbox_max[0] - bbox_min[0], 0.0) * std::max(bbox_max[1] - bbox_min[1], 0.0) *
                                    std::max(bbox_max[2] - bbox_min[2], 0.0),
                                1e-12);
    const std::vector<Mesh> parts = granules(m);
    double in_box = 0.0;
    if (parts.empty()) in_box = volume_in_box(m, bbox_min, bbox_max);
    for (const Mesh &p : parts) in_box += volume_in_box(p, bbox_min, bbox_max);
    *vf = std::min(std::max(in_box / box, 0.0), 1.0);
    return S2H_OK;
}

}  // extern "C"

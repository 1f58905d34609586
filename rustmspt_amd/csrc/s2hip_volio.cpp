// Volume file readers: RAW slice folders and TIFF stacks.
//
// RAW (load_raw_folder_typed, src/io/volume.rs:344-403): every regular file of
// the folder, sorted by file name, is one z slice of width*height samples of
// 8/16/32 bits; the inclusive slice range picks files (resolve_slice_range,
// src/io/volume.rs:258-285); a file whose length is not width*height*bytes is
// an error.  The samples stay in the spec's byte order; the device decodes them.
//
// TIFF (load_tiff_or_folder_typed_with_range, src/io/volume.rs:425-581, which
// decodes with the tiff crate 0.10.3): a path that is a file is a multi-page
// TIFF whose pages are the slices (the range picks pages); a folder holds
// .tif/.tiff files, sorted by name, each contributing all its pages (the range
// picks files).  Every page must match the first page's width and height and
// sample type.  Supported here as in the crate for one-channel images: 8/16/32
// bit unsigned or signed integer samples, classic or BigTIFF, strips or tiles,
// compression none, LZW, Deflate or PackBits, predictor none or horizontal.
#include "s2hip_volio.h"

#include <dirent.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

namespace s2h {
namespace volio {
namespace {

bool fail_with(std::string &err, const std::string &msg) {
    err = msg;
    return false;
}

bool lower_ext_is(const std::string &name, const char *const *exts, int n) {
    const size_t dot = name.rfind('.');
    if (dot == std::string::npos || dot == 0) return false;
    std::string ext = name.substr(dot + 1);
    for (auto &c : ext) c = (char)std::tolower((unsigned char)c);
    for (int i = 0; i < n; ++i)
        if (ext == exts[i]) return true;
    return false;
}

// collect_sorted_files (src/io/volume.rs:230-256): regular files (links
// followed), optionally filtered by a lower-cased extension, sorted by file name
// bytes.
bool sorted_files(const std::string &folder, const char *const *exts, int next, std::vector<std::string> &out,
                  std::string &err) {
    DIR *d = opendir(folder.c_str());
    if (!d) return fail_with(err, "cannot read folder " + folder);
    std::vector<std::string> names;
    while (dirent *e = readdir(d)) {
        std::string name = e->d_name;
        if (name == "." || name == "..") continue;
        struct stat st;
        const std::string full = folder + "/" + name;
        if (stat(full.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) continue;
        if (exts && !lower_ext_is(name, exts, next)) continue;
        names.push_back(name);
    }
    closedir(d);
    std::sort(names.begin(), names.end());  // char_traits<char> compares as unsigned bytes
    out.clear();
    for (auto &n : names) out.push_back(folder + "/" + n);
    return true;
}

bool slice_range(uint64_t total, int64_t start, int64_t end, uint64_t &s, uint64_t &e, std::string &err) {
    if (total == 0) return fail_with(err, "Input folder has no files");
    s = start < 0 ? 0 : (uint64_t)start;
    e = end < 0 ? total - 1 : (uint64_t)end;
    if (s >= total || e >= total || s > e)
        return fail_with(err, "Invalid slice range [" + std::to_string(start) + ", " + std::to_string(end) +
                                  "] for " + std::to_string(total) + " files");
    return true;
}

bool is_dir(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}
bool is_file(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

struct File {
    FILE *f = nullptr;
    uint64_t size = 0;
    explicit File(const std::string &path) {
        f = std::fopen(path.c_str(), "rb");
        if (f) {
            struct stat st;
            if (fstat(fileno(f), &st) == 0) size = (uint64_t)st.st_size;
        }
    }
    ~File() {
        if (f) std::fclose(f);
    }
    bool read_at(uint64_t off, void *dst, uint64_t n) {
        if (n == 0) return true;
        if (off > size || n > size - off) return false;
        if (fseeko(f, (off_t)off, SEEK_SET) != 0) return false;
        return std::fread(dst, 1, n, f) == n;
    }
};

void swap_samples(uint8_t *p, uint64_t count, uint32_t bytes) {
    if (bytes == 2)
        for (uint64_t i = 0; i < count; ++i) std::swap(p[2 * i], p[2 * i + 1]);
    else if (bytes == 4)
        for (uint64_t i = 0; i < count; ++i) {
            std::swap(p[4 * i], p[4 * i + 3]);
            std::swap(p[4 * i + 1], p[4 * i + 2]);
        }
}

// ------------------------------------------------------------------ TIFF

struct Page {
    uint64_t width = 0, height = 0;
    uint32_t bits = 0;
    int32_t is_signed = 0;
    uint32_t compression = 1, predictor = 1;
    bool tiled = false;
    uint64_t rows_per_strip = 0, tile_w = 0, tile_h = 0;
    std::vector<uint64_t> offsets, counts;
};

struct Tiff {
    File file;
    bool big_endian = false, bigtiff = false;
    std::vector<uint64_t> ifds;  // page IFD offsets in file order
    explicit Tiff(const std::string &path) : file(path) {}

    uint16_t u16(const uint8_t *p) const { return big_endian ? (uint16_t)(p[0] << 8 | p[1]) : (uint16_t)(p[1] << 8 | p[0]); }
    uint32_t u32(const uint8_t *p) const {
        return big_endian ? ((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3])
                          : ((uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]);
    }
    uint64_t u64(const uint8_t *p) const {
        const uint64_t a = u32(p), b = u32(p + 4);
        return big_endian ? (a << 32 | b) : (b << 32 | a);
    }

    bool open(const std::string &path, std::string &err) {
        if (!file.f) return fail_with(err, "cannot read " + path);
        uint8_t h[16];
        if (!file.read_at(0, h, 8)) return fail_with(err, "not a TIFF file: " + path);
        if (h[0] == 'I' && h[1] == 'I') big_endian = false;
        else if (h[0] == 'M' && h[1] == 'M') big_endian = true;
        else return fail_with(err, "not a TIFF file: " + path);
        const uint16_t version = u16(h + 2);
        uint64_t next;
        if (version == 42) {
            next = u32(h + 4);
        } else if (version == 43) {
            bigtiff = true;
            if (!file.read_at(0, h, 16) || u16(h + 4) != 8) return fail_with(err, "bad BigTIFF header: " + path);
            next = u64(h + 8);
        } else {
            return fail_with(err, "not a TIFF file: " + path);
        }
        while (next) {
            if (ifds.size() > (1u << 24) || std::find(ifds.begin(), ifds.end(), next) != ifds.end())
                return fail_with(err, "TIFF IFD chain loops in " + path);
            ifds.push_back(next);
            uint8_t c[8];
            const uint64_t csz = bigtiff ? 8 : 2, esz = bigtiff ? 20 : 12;
            if (!file.read_at(next, c, csz)) return fail_with(err, "truncated TIFF IFD in " + path);
            const uint64_t n = bigtiff ? u64(c) : u16(c);
            uint8_t l[8];
            if (!file.read_at(next + csz + n * esz, l, bigtiff ? 8 : 4)) return fail_with(err, "truncated TIFF IFD in " + path);
            next = bigtiff ? u64(l) : u32(l);
        }
        if (ifds.empty()) return fail_with(err, "TIFF has no pages: " + path);
        return true;
    }

    // Values of one IFD entry as unsigned integers.
    bool values(const uint8_t *e, std::vector<uint64_t> &out, std::string &err) {
        const uint16_t type = u16(e + 2);
        const uint64_t count = bigtiff ? u64(e + 4) : u32(e + 4);
        uint64_t size;
        switch (type) {
            case 1: case 2: case 6: case 7: size = 1; break;
            case 3: case 8: size = 2; break;
            case 4: case 9: case 13: size = 4; break;
            case 16: case 17: case 18: size = 8; break;
            default: return fail_with(err, "unsupported TIFF field type " + std::to_string(type));
        }
        if (count > (1ull << 32)) return fail_with(err, "TIFF field too long");
        const uint64_t bytes = size * count, inline_cap = bigtiff ? 8 : 4;
        std::vector<uint8_t> buf(std::max<uint64_t>(bytes, 1));
        const uint8_t *vp = e + (bigtiff ? 12 : 8);
        if (bytes <= inline_cap) {
            std::memcpy(buf.data(), vp, bytes);
        } else {
            const uint64_t off = bigtiff ? u64(vp) : u32(vp);
            if (!file.read_at(off, buf.data(), bytes)) return fail_with(err, "truncated TIFF field");
        }
        out.resize(count);
        for (uint64_t i = 0; i < count; ++i) {
            const uint8_t *p = buf.data() + i * size;
            switch (size) {
                case 1: out[i] = type == 6 ? (uint64_t)(int64_t)(int8_t)p[0] : p[0]; break;
                case 2: out[i] = type == 8 ? (uint64_t)(int64_t)(int16_t)u16(p) : u16(p); break;
                case 4: out[i] = type == 9 ? (uint64_t)(int64_t)(int32_t)u32(p) : u32(p); break;
                default: out[i] = u64(p); break;
            }
        }
        return true;
    }

    bool page(uint64_t index, Page &pg, const std::string &path, std::string &err) {
        const uint64_t at = ifds[index];
        uint8_t c[8];
        const uint64_t csz = bigtiff ? 8 : 2, esz = bigtiff ? 20 : 12;
        if (!file.read_at(at, c, csz)) return fail_with(err, "truncated TIFF IFD in " + path);
        const uint64_t n = bigtiff ? u64(c) : u16(c);
        std::vector<uint8_t> ents(n * esz);
        if (!file.read_at(at + csz, ents.data(), ents.size())) return fail_with(err, "truncated TIFF IFD in " + path);
        uint64_t spp = 1, sample_format = 1, planar = 1;
        std::vector<uint64_t> bits;
        bool have_w = false, have_h = false;
        std::vector<uint64_t> v;
        for (uint64_t i = 0; i < n; ++i) {
            const uint8_t *e = ents.data() + i * esz;
            const uint16_t tag = u16(e);
            switch (tag) {
                case 256: case 257: case 258: case 259: case 273: case 277: case 278: case 279:
                case 284: case 317: case 322: case 323: case 324: case 325: case 339:
                    if (!values(e, v, err)) return fail_with(err, err + " in " + path);
                    if (v.empty() && tag != 273 && tag != 279 && tag != 324 && tag != 325)
                        return fail_with(err, "empty TIFF field in " + path);
                    break;
                default:
                    continue;
            }
            switch (tag) {
                case 256: pg.width = v[0]; have_w = true; break;
                case 257: pg.height = v[0]; have_h = true; break;
                case 258: bits = v; break;
                case 259: pg.compression = (uint32_t)v[0]; break;
                case 273: case 324: pg.offsets = v; break;
                case 277: spp = v[0]; break;
                case 278: pg.rows_per_strip = v[0]; break;
                case 279: case 325: pg.counts = v; break;
                case 284: planar = v[0]; break;
                case 317: pg.predictor = (uint32_t)v[0]; break;
                case 322: pg.tile_w = v[0]; pg.tiled = true; break;
                case 323: pg.tile_h = v[0]; pg.tiled = true; break;
                case 339: sample_format = v[0]; break;
            }
        }
        (void)planar;
        if (!have_w || !have_h || pg.width == 0 || pg.height == 0)
            return fail_with(err, "TIFF page without dimensions in " + path);
        if (spp != 1) return fail_with(err, "Unsupported TIFF sample type; supported: U8/U16/U32/I8/I16/I32 (" +
                                                std::to_string(spp) + " samples per pixel in " + path + ")");
        if (bits.empty()) bits.push_back(1);
        if (bits[0] != 8 && bits[0] != 16 && bits[0] != 32)
            return fail_with(err, "Unsupported TIFF sample type; supported: U8/U16/U32/I8/I16/I32 (" +
                                      std::to_string(bits[0]) + "-bit samples in " + path + ")");
        if (sample_format != 1 && sample_format != 2)
            return fail_with(err, "Unsupported TIFF sample type; supported: U8/U16/U32/I8/I16/I32 (sample format " +
                                      std::to_string(sample_format) + " in " + path + ")");
        pg.bits = (uint32_t)bits[0];
        pg.is_signed = sample_format == 2;
        if (pg.compression != 1 && pg.compression != 5 && pg.compression != 8 && pg.compression != 32946 &&
            pg.compression != 32773)
            return fail_with(err, "unsupported TIFF compression " + std::to_string(pg.compression) + " in " + path);
        if (pg.predictor != 1 && pg.predictor != 2)
            return fail_with(err, "unsupported TIFF predictor " + std::to_string(pg.predictor) + " in " + path);
        if (pg.tiled) {
            if (pg.tile_w == 0 || pg.tile_h == 0) return fail_with(err, "bad TIFF tile size in " + path);
        } else if (pg.rows_per_strip == 0 || pg.rows_per_strip > pg.height) {
            pg.rows_per_strip = pg.height;
        }
        if (pg.offsets.size() != pg.counts.size() || pg.offsets.empty())
            return fail_with(err, "TIFF strip table mismatch in " + path);
        return true;
    }
};

// TIFF LZW (TIFF 6.0 section 13): MSB-first codes of 9..12 bits, Clear = 256,
// EndOfInformation = 257, the code width grows one code early.
bool lzw_decode(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t &written) {
    struct Entry {
        uint32_t prefix;
        uint16_t len;
        uint8_t first, last;
    };
    std::vector<Entry> table(4096);
    for (uint32_t i = 0; i < 256; ++i) table[i] = {0xFFFFFFFFu, 1, (uint8_t)i, (uint8_t)i};
    uint32_t next = 258, width = 9, prev = 0xFFFFFFFFu;
    uint64_t bitpos = 0, o = 0;
    std::vector<uint8_t> tmp(4096);
    auto emit = [&](uint32_t code) -> bool {
        const uint16_t len = table[code].len;
        uint32_t c = code;
        for (int k = len - 1; k >= 0; --k) {
            tmp[k] = table[c].last;
            c = table[c].prefix;
        }
        const uint64_t m = std::min<uint64_t>(len, cap - o);
        std::memcpy(out + o, tmp.data(), m);
        o += m;
        return true;
    };
    while (bitpos + width <= n * 8) {
        uint32_t code = 0;
        for (uint32_t b = 0; b < width; ++b, ++bitpos) code = code << 1 | ((in[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
        if (code == 257) break;
        if (code == 256) {
            next = 258;
            width = 9;
            prev = 0xFFFFFFFFu;
            continue;
        }
        if (prev == 0xFFFFFFFFu) {
            if (code > 255) return false;
            emit(code);
            prev = code;
            continue;
        }
        uint8_t first;
        if (code < next) first = table[code].first;
        else if (code == next) first = table[prev].first;
        else return false;
        if (next < 4096) {
            table[next] = {prev, (uint16_t)(table[prev].len + 1), table[prev].first, first};
            ++next;
        }
        emit(code);
        prev = code;
        if (next + 1 >= (1u << width) && width < 12) ++width;
        if (o >= cap) break;
    }
    written = o;
    return true;
}

bool packbits_decode(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t &written) {
    uint64_t i = 0, o = 0;
    while (i < n && o < cap) {
        const int8_t h = (int8_t)in[i++];
        if (h >= 0) {
            const uint64_t k = std::min<uint64_t>((uint64_t)h + 1, n - i);
            const uint64_t m = std::min(k, cap - o);
            std::memcpy(out + o, in + i, m);
            i += k;
            o += m;
        } else if (h != -128) {
            if (i >= n) break;
            const uint64_t k = std::min<uint64_t>((uint64_t)(1 - h), cap - o);
            std::memset(out + o, in[i++], k);
            o += k;
        }
    }
    written = o;
    return true;
}

bool inflate_decode(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t &written) {
    z_stream zs;
    std::memset(&zs, 0, sizeof zs);
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef *>(in);
    zs.avail_in = (uInt)n;
    zs.next_out = out;
    zs.avail_out = (uInt)cap;
    const int r = inflate(&zs, Z_FINISH);
    written = cap - zs.avail_out;
    inflateEnd(&zs);
    return r == Z_STREAM_END || r == Z_OK || r == Z_BUF_ERROR;
}

// Undo horizontal differencing on rows of `row_samples` samples stored in the
// file's byte order.
void undo_predictor(uint8_t *p, uint64_t rows, uint64_t row_samples, uint32_t bytes, bool big) {
    auto get = [&](const uint8_t *q) -> uint32_t {
        if (bytes == 1) return q[0];
        if (bytes == 2) return big ? (uint32_t)(q[0] << 8 | q[1]) : (uint32_t)(q[1] << 8 | q[0]);
        return big ? ((uint32_t)q[0] << 24 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 8 | q[3])
                   : ((uint32_t)q[3] << 24 | (uint32_t)q[2] << 16 | (uint32_t)q[1] << 8 | q[0]);
    };
    auto put = [&](uint8_t *q, uint32_t v) {
        for (uint32_t k = 0; k < bytes; ++k) q[big ? bytes - 1 - k : k] = (uint8_t)(v >> (8 * k));
    };
    for (uint64_t r = 0; r < rows; ++r) {
        uint8_t *row = p + r * row_samples * bytes;
        uint32_t acc = get(row);
        for (uint64_t x = 1; x < row_samples; ++x) {
            acc += get(row + x * bytes);
            put(row + x * bytes, acc);
        }
    }
}

// Decode one block (strip or tile) into dst (cap bytes expected).
bool decode_block(Tiff &t, const Page &pg, uint64_t b, std::vector<uint8_t> &raw, uint8_t *dst, uint64_t cap,
                  const std::string &path, std::string &err) {
    const uint64_t off = pg.offsets[b], cnt = pg.counts[b];
    if (pg.compression == 1) {
        const uint64_t m = std::min(cnt, cap);
        if (!t.file.read_at(off, dst, m)) return fail_with(err, "truncated TIFF strip in " + path);
        if (m < cap) std::memset(dst + m, 0, cap - m);
        return true;
    }
    raw.resize(std::max<uint64_t>(cnt, 1));
    if (!t.file.read_at(off, raw.data(), cnt)) return fail_with(err, "truncated TIFF strip in " + path);
    uint64_t got = 0;
    bool ok;
    if (pg.compression == 5) ok = lzw_decode(raw.data(), cnt, dst, cap, got);
    else if (pg.compression == 32773) ok = packbits_decode(raw.data(), cnt, dst, cap, got);
    else ok = inflate_decode(raw.data(), cnt, dst, cap, got);
    if (!ok) return fail_with(err, "corrupt compressed TIFF strip in " + path);
    if (got < cap) std::memset(dst + got, 0, cap - got);
    return true;
}

// Decode page pg into out (width*height samples) in byte order `big`.
bool decode_page(Tiff &t, const Page &pg, uint8_t *out, bool big, const std::string &path, std::string &err) {
    const uint32_t bytes = pg.bits / 8;
    const uint64_t row_bytes = pg.width * bytes;
    std::vector<uint8_t> raw, block;
    if (!pg.tiled) {
        const uint64_t nstrips = (pg.height + pg.rows_per_strip - 1) / pg.rows_per_strip;
        if (pg.offsets.size() < nstrips) return fail_with(err, "TIFF strip table too short in " + path);
        for (uint64_t s = 0; s < nstrips; ++s) {
            const uint64_t r0 = s * pg.rows_per_strip, rows = std::min(pg.rows_per_strip, pg.height - r0);
            uint8_t *dst = out + r0 * row_bytes;
            if (!decode_block(t, pg, s, raw, dst, rows * row_bytes, path, err)) return false;
            if (pg.predictor == 2) undo_predictor(dst, rows, pg.width, bytes, t.big_endian);
        }
    } else {
        const uint64_t across = (pg.width + pg.tile_w - 1) / pg.tile_w, down = (pg.height + pg.tile_h - 1) / pg.tile_h;
        if (pg.offsets.size() < across * down) return fail_with(err, "TIFF tile table too short in " + path);
        const uint64_t tile_row = pg.tile_w * bytes;
        block.resize(tile_row * pg.tile_h);
        for (uint64_t ty = 0; ty < down; ++ty)
            for (uint64_t tx = 0; tx < across; ++tx) {
                if (!decode_block(t, pg, ty * across + tx, raw, block.data(), block.size(), path, err)) return false;
                if (pg.predictor == 2) undo_predictor(block.data(), pg.tile_h, pg.tile_w, bytes, t.big_endian);
                const uint64_t x0 = tx * pg.tile_w, y0 = ty * pg.tile_h;
                const uint64_t w = std::min(pg.tile_w, pg.width - x0), h = std::min(pg.tile_h, pg.height - y0);
                for (uint64_t r = 0; r < h; ++r)
                    std::memcpy(out + (y0 + r) * row_bytes + x0 * bytes, block.data() + r * tile_row, w * bytes);
            }
    }
    if (t.big_endian != big) swap_samples(out, pg.width * pg.height, bytes);
    return true;
}

const char *const kTiffExts[] = {"tif", "tiff"};

struct TiffPlan {
    std::vector<std::string> files;     // in order
    std::vector<uint64_t> first, last;  // page range per file
};

bool tiff_plan(const s2h_volume_source &src, TiffPlan &plan, std::string &err) {
    const std::string path = src.path;
    uint64_t s, e;
    if (is_file(path)) {
        Tiff t(path);
        if (!t.open(path, err)) return false;
        if (!slice_range(t.ifds.size(), src.slice_start, src.slice_end, s, e, err)) return false;
        plan.files = {path};
        plan.first = {s};
        plan.last = {e};
        return true;
    }
    if (!is_dir(path)) return fail_with(err, "Input path is neither file nor folder: " + path);
    std::vector<std::string> files;
    if (!sorted_files(path, kTiffExts, 2, files, err)) return false;
    if (files.empty()) return fail_with(err, "No TIFF files found in " + path);
    if (!slice_range(files.size(), src.slice_start, src.slice_end, s, e, err)) return false;
    for (uint64_t i = s; i <= e; ++i) {
        Tiff t(files[i]);
        if (!t.open(files[i], err)) return false;
        plan.files.push_back(files[i]);
        plan.first.push_back(0);
        plan.last.push_back(t.ifds.size() - 1);
    }
    return true;
}

bool tiff_probe(const s2h_volume_source &src, s2h_volume_info &info, TiffPlan &plan, std::string &err) {
    if (!tiff_plan(src, plan, err)) return false;
    uint64_t depth = 0;
    bool first = true;
    for (size_t f = 0; f < plan.files.size(); ++f) {
        Tiff t(plan.files[f]);
        if (!t.open(plan.files[f], err)) return false;
        for (uint64_t p = plan.first[f]; p <= plan.last[f]; ++p) {
            Page pg;
            if (!t.page(p, pg, plan.files[f], err)) return false;
            if (first) {
                info.width = pg.width;
                info.height = pg.height;
                info.bits = pg.bits;
                info.is_signed = pg.is_signed;
                info.big_endian = t.big_endian;
                first = false;
            } else if (pg.width != info.width || pg.height != info.height) {
                return fail_with(err, "TIFF shape mismatch at " + plan.files[f]);
            } else if (pg.bits != info.bits || pg.is_signed != info.is_signed) {
                return fail_with(err, "TIFF numeric type mismatch at " + plan.files[f]);
            }
            ++depth;
        }
    }
    info.depth = depth;
    info.reserved = 0;
    return true;
}

bool raw_probe(const s2h_volume_source &src, s2h_volume_info &info, std::vector<std::string> &files, uint64_t &s,
               uint64_t &e, std::string &err) {
    if (src.width == 0 || src.height == 0) return fail_with(err, "RAW width/height must be > 0");
    if (src.bits != 8 && src.bits != 16 && src.bits != 32) return fail_with(err, "RAW bits must be 8/16/32");
    if (!sorted_files(src.path, nullptr, 0, files, err)) return false;
    if (!slice_range(files.size(), src.slice_start, src.slice_end, s, e, err)) return false;
    info.width = src.width;
    info.height = src.height;
    info.depth = e - s + 1;
    info.bits = src.bits;
    info.is_signed = src.is_signed != 0;
    info.big_endian = src.big_endian != 0;
    info.reserved = 0;
    return true;
}

}  // namespace

bool probe(const s2h_volume_source &src, s2h_volume_info &info, std::string &err) {
    if (!src.path) return fail_with(err, "null path");
    if (src.kind == S2H_VOLUME_RAW) {
        std::vector<std::string> files;
        uint64_t s, e;
        return raw_probe(src, info, files, s, e, err);
    }
    if (src.kind == S2H_VOLUME_TIFF) {
        TiffPlan plan;
        return tiff_probe(src, info, plan, err);
    }
    return fail_with(err, "unknown volume kind " + std::to_string(src.kind));
}

bool stream(const s2h_volume_source &src, const s2h_volume_info &info, const SliceSink &sink, std::string &err) {
    const uint32_t bytes = info.bits / 8;
    const uint64_t slice_bytes = info.width * info.height * bytes;
    if (src.kind == S2H_VOLUME_RAW) {
        std::vector<std::string> files;
        uint64_t s, e;
        s2h_volume_info chk;
        if (!raw_probe(src, chk, files, s, e, err)) return false;
        if (chk.width != info.width || chk.height != info.height || chk.depth != info.depth)
            return fail_with(err, "RAW folder changed while reading");
        for (uint64_t i = s; i <= e; ++i) {
            File f(files[i]);
            if (!f.f) return fail_with(err, "cannot read " + files[i]);
            if (f.size != slice_bytes)
                return fail_with(err, "RAW slice size mismatch for " + files[i] + ": expected " +
                                          std::to_string(slice_bytes) + ", got " + std::to_string(f.size));
            uint8_t *dst = sink.slot(slice_bytes, err);
            if (!dst) return false;
            if (!f.read_at(0, dst, slice_bytes)) return fail_with(err, "cannot read " + files[i]);
            if (!sink.commit(err)) return false;
        }
        return true;
    }
    TiffPlan plan;
    s2h_volume_info chk;
    if (!tiff_probe(src, chk, plan, err)) return false;
    if (chk.width != info.width || chk.height != info.height || chk.depth != info.depth || chk.bits != info.bits)
        return fail_with(err, "TIFF stack changed while reading");
    for (size_t f = 0; f < plan.files.size(); ++f) {
        Tiff t(plan.files[f]);
        if (!t.open(plan.files[f], err)) return false;
        for (uint64_t p = plan.first[f]; p <= plan.last[f]; ++p) {
            Page pg;
            if (!t.page(p, pg, plan.files[f], err)) return false;
            uint8_t *dst = sink.slot(slice_bytes, err);
            if (!dst) return false;
            if (!decode_page(t, pg, dst, info.big_endian != 0, plan.files[f], err)) return false;
            if (!sink.commit(err)) return false;
        }
    }
    return true;
}

}  // namespace volio

void set_error(const std::string &msg);
}  // namespace s2h

extern "C" {

s2h_status s2h_volume_probe(const s2h_volume_source *src, s2h_volume_info *info) {
    if (!src || !info) {
        s2h::set_error("null argument");
        return S2H_ERR_ARG;
    }
    std::string err;
    if (!s2h::volio::probe(*src, *info, err)) {
        s2h::set_error(err);
        return S2H_ERR_IO;
    }
    return S2H_OK;
}

s2h_status s2h_volume_read(const s2h_volume_source *src, void *samples, uint64_t nbytes, s2h_volume_info *info) {
    if (!src || !samples || !info) {
        s2h::set_error("null argument");
        return S2H_ERR_ARG;
    }
    std::string err;
    if (!s2h::volio::probe(*src, *info, err)) {
        s2h::set_error(err);
        return S2H_ERR_IO;
    }
    const uint64_t need = info->width * info->height * info->depth * (info->bits / 8);
    if (nbytes != need) {
        s2h::set_error("volume has " + std::to_string(need) + " sample bytes, buffer " + std::to_string(nbytes));
        return S2H_ERR_SIZE;
    }
    uint8_t *p = (uint8_t *)samples;
    s2h::volio::SliceSink sink;
    sink.slot = [&](uint64_t bytes, std::string &) { return p; };
    sink.commit = [&](std::string &) {
        p += info->width * info->height * (info->bits / 8);
        return true;
    };
    if (!s2h::volio::stream(*src, *info, sink, err)) {
        s2h::set_error(err);
        return S2H_ERR_IO;
    }
    return S2H_OK;
}

}  // extern "C"

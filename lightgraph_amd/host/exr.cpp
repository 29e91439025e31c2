// Scanline OpenEXR writer and reader (single part; NONE, ZIPS and ZIP
// compression; HALF and FLOAT channels), written from the published OpenEXR file
// layout: magic + version, attribute header, chunk offset table, chunks of
// (y, size, data) with channels in name order per scanline.  ZIP chunks hold
// zlib data of the byte-interleaved, delta-predicted scanlines.
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>

#include "lightgraph.hpp"

namespace lightgraph {
namespace {

enum : uint8_t { EXR_NONE = 0, EXR_ZIPS = 2, EXR_ZIP = 3 };
enum : int32_t { PIX_HALF = 1, PIX_FLOAT = 2 };

void put_u32(std::string &b, uint32_t v) { b.append(reinterpret_cast<const char *>(&v), 4); }
void put_i32(std::string &b, int32_t v) { b.append(reinterpret_cast<const char *>(&v), 4); }
void put_f32(std::string &b, float v) { b.append(reinterpret_cast<const char *>(&v), 4); }
void put_u64(std::string &b, uint64_t v) { b.append(reinterpret_cast<const char *>(&v), 8); }

void attr(std::string &h, const char *name, const char *type, const std::string &value) {
    h.append(name);
    h.push_back('\0');
    h.append(type);
    h.push_back('\0');
    put_i32(h, (int32_t)value.size());
    h.append(value);
}

// interleave the two byte halves, then store byte deltas (+128)
std::string zip_prepare(const std::string &raw) {
    const size_t n = raw.size();
    std::string t(n, '\0');
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) t[(i & 1) ? b++ : a++] = raw[i];
    for (size_t i = n; i-- > 1;) t[i] = (char)((uint8_t)t[i] - (uint8_t)t[i - 1] + 128u);
    return t;
}

std::string zip_restore(const std::string &t0) {
    std::string t = t0;
    const size_t n = t.size();
    for (size_t i = 1; i < n; ++i) t[i] = (char)((uint8_t)t[i - 1] + (uint8_t)t[i] - 128u);
    std::string raw(n, '\0');
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) raw[i] = t[(i & 1) ? b++ : a++];
    return raw;
}

float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31;
    uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff, bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: normalise
            int sh = 0;
            while (!(m & 0x400)) {
                m <<= 1;
                ++sh;
            }
            m &= 0x3ff;
            bits = s | ((uint32_t)(127 - 15 - sh + 1) << 23) | (m << 13);
        }
    } else if (e == 31) {
        bits = s | 0x7f800000u | (m << 13);
    } else {
        bits = s | ((e - 15 + 127) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

}  // namespace

void saveEXR(const std::string &fname, const Film &rgb, const Film &direct, const Film &scatter,
             const Film &emission, const SceneSettings &opts) {
    std::string filename = fname;
    if (filename.size() < 4 || filename.compare(filename.size() - 4, 4, ".exr") != 0) filename += ".exr";
    uint8_t comp;
    if (opts.compression == "none")
        comp = EXR_NONE;
    else if (opts.compression == "zip")
        comp = EXR_ZIP;
    else if (opts.compression == "zips")
        comp = EXR_ZIPS;
    else
        throw Error("expected none, zip or zips compression, got \"" + opts.compression + "\"", LG_ERR_UNSUPPORTED);
    const int W = (int)rgb.width(), H = (int)rgb.height();
    // channels in name order; each refers to a film and a component
    struct Ch {
        const char *name;
        const Film *film;
        int comp;
    };
    const Ch chans[13] = {{"A", &rgb, 3},       {"B", &rgb, 2},         {"G", &rgb, 1},
                          {"R", &rgb, 0},       {"direct.B", &direct, 2}, {"direct.G", &direct, 1},
                          {"direct.R", &direct, 0}, {"emission.B", &emission, 2}, {"emission.G", &emission, 1},
                          {"emission.R", &emission, 0}, {"scatter.B", &scatter, 2}, {"scatter.G", &scatter, 1},
                          {"scatter.R", &scatter, 0}};
    std::string head;
    put_u32(head, 20000630u);
    put_u32(head, 2u);
    std::string cl;
    for (const Ch &c : chans) {
        cl.append(c.name);
        cl.push_back('\0');
        put_i32(cl, PIX_FLOAT);
        cl.append(4, '\0');  // pLinear + reserved
        put_i32(cl, 1);
        put_i32(cl, 1);
    }
    cl.push_back('\0');
    attr(head, "channels", "chlist", cl);
    attr(head, "compression", "compression", std::string(1, (char)comp));
    std::string box;
    put_i32(box, 0);
    put_i32(box, 0);
    put_i32(box, W - 1);
    put_i32(box, H - 1);
    attr(head, "dataWindow", "box2i", box);
    attr(head, "displayWindow", "box2i", box);
    attr(head, "lineOrder", "lineOrder", std::string(1, '\0'));
    std::string f1;
    put_f32(f1, 1.0f);
    attr(head, "pixelAspectRatio", "float", f1);
    std::string v2;
    put_f32(v2, 0.0f);
    put_f32(v2, 0.0f);
    attr(head, "screenWindowCenter", "v2f", v2);
    attr(head, "screenWindowWidth", "float", f1);
    head.push_back('\0');
    const int lines = comp == EXR_ZIP ? 16 : 1;
    const int chunks = (H + lines - 1) / lines;
    std::vector<std::string> blocks((size_t)chunks);
    for (int c = 0; c < chunks; ++c) {
        const int y0 = c * lines, y1 = std::min(H, y0 + lines);
        std::string raw;
        raw.reserve((size_t)(y1 - y0) * 13 * W * 4);
        for (int y = y0; y < y1; ++y)
            for (const Ch &ch : chans)
                for (int x = 0; x < W; ++x) {
                    const Film::RGBA &p = ch.film->px[(size_t)y * W + x];
                    const float v = ch.comp == 0 ? p.r : ch.comp == 1 ? p.g : ch.comp == 2 ? p.b : p.a;
                    put_f32(raw, v);
                }
        std::string data = raw;
        if (comp != EXR_NONE) {
            const std::string t = zip_prepare(raw);
            uLongf zl = compressBound((uLong)t.size());
            std::string z(zl, '\0');
            if (compress2(reinterpret_cast<Bytef *>(&z[0]), &zl, reinterpret_cast<const Bytef *>(t.data()),
                          (uLong)t.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
                throw Error("zlib compression failed", LG_ERR_INVALID);
            z.resize(zl);
            if (z.size() < raw.size()) data = z;  // stored raw when compression does not pay
        }
        std::string blk;
        put_i32(blk, y0);
        put_i32(blk, (int32_t)data.size());
        blk += data;
        blocks[c] = std::move(blk);
    }
    std::ofstream out(filename, std::ios::binary);
    if (!out) throw Error("cannot open " + filename, LG_ERR_INVALID);
    uint64_t pos = head.size() + 8ull * chunks;
    std::string table;
    for (const std::string &b : blocks) {
        put_u64(table, pos);
        pos += b.size();
    }
    out.write(head.data(), (std::streamsize)head.size());
    out.write(table.data(), (std::streamsize)table.size());
    for (const std::string &b : blocks) out.write(b.data(), (std::streamsize)b.size());
    if (!out) throw Error("write failed: " + filename, LG_ERR_INVALID);
}

std::vector<float> readEXR_RGB(const std::string &fname, int *width, int *height) {
    std::ifstream in(fname, std::ios::binary);
    if (!in) throw Error("cannot open " + fname, LG_ERR_INVALID);
    const std::string f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    size_t p = 0;
    auto need = [&](size_t n) {
        if (p + n > f.size()) throw Error("truncated EXR file " + fname, LG_ERR_INVALID);
    };
    auto rd32 = [&]() {
        need(4);
        int32_t v;
        std::memcpy(&v, f.data() + p, 4);
        p += 4;
        return v;
    };
    auto cstr = [&]() {
        const size_t e = f.find('\0', p);
        if (e == std::string::npos) throw Error("bad EXR header in " + fname, LG_ERR_INVALID);
        std::string s = f.substr(p, e - p);
        p = e + 1;
        return s;
    };
    if ((uint32_t)rd32() != 20000630u) throw Error(fname + " is not an OpenEXR file", LG_ERR_INVALID);
    const uint32_t version = (uint32_t)rd32();
    if ((version & 0xff) != 2 || (version & 0x1e00)) throw Error(fname + ": only single-part scanline EXR files are read", LG_ERR_UNSUPPORTED);
    std::vector<std::pair<std::string, int32_t>> chans;
    int32_t comp = -1, x0 = 0, y0 = 0, x1 = -1, y1 = -1;
    for (;;) {
        const std::string name = cstr();
        if (name.empty()) break;
        const std::string type = cstr();
        const int32_t size = rd32();
        need((size_t)size);
        const size_t end = p + (size_t)size;
        if (name == "channels") {
            while (p < end && f[p] != '\0') {
                const std::string cn = cstr();
                const int32_t pt = rd32();
                p += 12;  // pLinear, reserved, sampling
                chans.push_back({cn, pt});
            }
        } else if (name == "compression") {
            comp = (uint8_t)f[p];
        } else if (name == "dataWindow") {
            x0 = rd32();
            y0 = rd32();
            x1 = rd32();
            y1 = rd32();
        }
        p = end;
    }
    if (comp != EXR_NONE && comp != EXR_ZIPS && comp != EXR_ZIP)
        throw Error(fname + ": compression " + std::to_string(comp) + " is not supported (NONE, ZIPS, ZIP)",
                    LG_ERR_UNSUPPORTED);
    const int W = x1 - x0 + 1, H = y1 - y0 + 1;
    if (W <= 0 || H <= 0 || chans.empty()) throw Error(fname + ": empty image", LG_ERR_INVALID);
    int idx[3] = {-1, -1, -1};
    size_t pix_bytes = 0;
    std::vector<size_t> ch_off;
    for (size_t c = 0; c < chans.size(); ++c) {
        if (chans[c].second != PIX_HALF && chans[c].second != PIX_FLOAT)
            throw Error(fname + ": channel " + chans[c].first + " is not HALF or FLOAT", LG_ERR_UNSUPPORTED);
        ch_off.push_back(pix_bytes);
        pix_bytes += chans[c].second == PIX_HALF ? 2 : 4;
        const std::string &n = chans[c].first;
        if (n == "R") idx[0] = (int)c;
        if (n == "G") idx[1] = (int)c;
        if (n == "B") idx[2] = (int)c;
    }
    for (int k = 0; k < 3; ++k)
        if (idx[k] < 0) throw Error(fname + ": no R, G, B channels", LG_ERR_INVALID);
    const int lines = comp == EXR_ZIP ? 16 : 1;
    const int chunks = (H + lines - 1) / lines;
    std::vector<uint64_t> offs((size_t)chunks);
    need(8ull * chunks);
    std::memcpy(offs.data(), f.data() + p, 8ull * chunks);
    std::vector<float> rgb((size_t)W * H * 3, 0.0f);
    for (int c = 0; c < chunks; ++c) {
        p = (size_t)offs[c];
        const int yb = rd32() - y0;
        const int32_t size = rd32();
        need((size_t)size);
        const int nl = std::min(lines, H - yb);
        if (yb < 0 || nl <= 0) throw Error(fname + ": bad chunk", LG_ERR_INVALID);
        const size_t raw_size = (size_t)nl * W * pix_bytes;
        std::string raw;
        if ((size_t)size == raw_size || comp == EXR_NONE) {
            raw = f.substr(p, raw_size);
        } else {
            std::string t(raw_size, '\0');
            uLongf dl = (uLongf)raw_size;
            if (uncompress(reinterpret_cast<Bytef *>(&t[0]), &dl, reinterpret_cast<const Bytef *>(f.data() + p),
                           (uLong)size) != Z_OK ||
                dl != raw_size)
                throw Error(fname + ": zlib data error", LG_ERR_INVALID);
            raw = zip_restore(t);
        }
        if (raw.size() != raw_size) throw Error(fname + ": short chunk", LG_ERR_INVALID);
        for (int l = 0; l < nl; ++l) {
            const size_t line = (size_t)l * W * pix_bytes;
            for (int k = 0; k < 3; ++k) {
                const int ci = idx[k];
                const size_t base = line + ch_off[ci] * W;
                for (int x = 0; x < W; ++x) {
                    float v;
                    if (chans[ci].second == PIX_FLOAT) {
                        std::memcpy(&v, raw.data() + base + 4 * (size_t)x, 4);
                    } else {
                        uint16_t h;
                        std::memcpy(&h, raw.data() + base + 2 * (size_t)x, 2);
                        v = half_to_float(h);
                    }
                    rgb[3 * ((size_t)(yb + l) * W + x) + k] = v;
                }
            }
        }
    }
    *width = W;
    *height = H;
    return rgb;
}

}  // namespace lightgraph

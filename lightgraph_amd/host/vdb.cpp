// OpenVDB FloatGrid reader (see vdb.hpp).  Layout followed, per grid:
//   uint32 compression flags (ZIP 1, ACTIVE_MASK 2, BLOSC 4)
//   metadata: uint32 count, then (string name, string type, uint32 size, bytes)
//   transform: string map type, then the map's doubles
//   int32 buffer count (1); root: float background, uint32 tiles, uint32 children,
//     tiles (int32 x3, float, uint8 active), children (int32 x3 origin, node)
//   32^3 / 16^3 internal node: child mask, value mask, mask-compressed values,
//     then its children in ascending offset order; leaf: value mask
//   then, at the grid's block offset, the leaves' buffers in the same order:
//     value mask again, mask-compressed values (512)
// Strings are uint32 length + bytes; masks are little-endian uint64 words;
// offsets are x-major (x slowest, z fastest).
#include "vdb.hpp"

#include <dlfcn.h>
#include <zlib.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>

namespace lightgraph {
namespace {

constexpr int64_t kMagic = 0x56444220;
constexpr uint32_t kMinVersion = 222;  // per-grid compression and mask-compressed node values
constexpr uint32_t kCompressZip = 0x1, kCompressMask = 0x2, kCompressBlosc = 0x4;
enum : int8_t {
    kNoMaskOrInactiveVals = 0,
    kNoMaskAndMinusBg = 1,
    kNoMaskAndOneInactiveVal = 2,
    kMaskAndNoInactiveVals = 3,
    kMaskAndOneInactiveVal = 4,
    kMaskAndTwoInactiveVals = 5,
    kNoMaskAndAllVals = 6
};
const char kSuffixSep = '\x1e';  // separates a duplicate grid name from its index

[[noreturn]] void bad(const std::string &msg) { throw std::runtime_error("vdb: " + msg); }

class Stream {
   public:
    explicit Stream(const std::string &path) : f_(path, std::ios::binary) {
        if (!f_) bad("cannot open " + path);
        f_.seekg(0, std::ios::end);
        size_ = (int64_t)f_.tellg();
        f_.seekg(0);
    }
    void read(void *p, size_t n) {
        if ((int64_t)n > size_ - pos()) bad("truncated stream");
        f_.read(static_cast<char *>(p), (std::streamsize)n);
        if (!f_) bad("read failed");
    }
    template <typename T>
    T get() {
        T v;
        read(&v, sizeof(T));
        return v;
    }
    std::string str() {
        const uint32_t n = get<uint32_t>();
        if ((int64_t)n > size_ - pos()) bad("bad string length");
        std::string s(n, '\0');
        if (n) read(&s[0], n);
        return s;
    }
    void skip(uint64_t n) { seek(pos() + (int64_t)n); }
    void seek(int64_t p) {
        if (p < 0 || p > size_) bad("offset outside the file");
        f_.seekg(p);
    }
    int64_t pos() { return (int64_t)f_.tellg(); }

   private:
    std::ifstream f_;
    int64_t size_ = 0;
};

struct Mask {
    std::vector<uint64_t> w;
    void load(Stream &s, size_t bits) {
        w.assign(bits / 64, 0);
        s.read(w.data(), bits / 8);
    }
    bool on(size_t i) const { return (w[i >> 6] >> (i & 63)) & 1u; }
    size_t count() const {
        size_t c = 0;
        for (uint64_t x : w) c += (size_t)__builtin_popcountll(x);
        return c;
    }
};

// c-blosc, loaded on first use (only files written with blosc need it)
using BloscDecompress = int (*)(const void *, void *, size_t);
BloscDecompress blosc_decompress_fn() {
    static BloscDecompress fn = [] {
        const char *names[] = {"libblosc.so.1", "libblosc.so", "/opt/conda/lib/libblosc.so.1"};
        for (const char *n : names)
            if (void *h = dlopen(n, RTLD_NOW | RTLD_LOCAL)) {
                if (auto init = reinterpret_cast<void (*)()>(dlsym(h, "blosc_init"))) init();
                if (auto f = reinterpret_cast<BloscDecompress>(dlsym(h, "blosc_decompress"))) return f;
            }
        return BloscDecompress(nullptr);
    }();
    return fn;
}

// io::readData: the bytes of `count` values, raw, zlib or blosc (an int64 size
// first; a negative size marks data stored uncompressed).
void readData(Stream &s, void *dst, size_t bytes, uint32_t compression) {
    if (!(compression & (kCompressZip | kCompressBlosc))) {
        s.read(dst, bytes);
        return;
    }
    const int64_t n = s.get<int64_t>();
    if (n <= 0) {
        if ((size_t)(-n) != bytes) bad("stored size does not match the node");
        s.read(dst, bytes);
        return;
    }
    std::vector<unsigned char> packed((size_t)n);
    s.read(packed.data(), (size_t)n);
    if (compression & kCompressBlosc) {
        const BloscDecompress f = blosc_decompress_fn();
        if (!f) bad("the file is blosc-compressed and libblosc was not found");
        if (f(packed.data(), dst, bytes) != (int)bytes) bad("blosc decompression failed");
    } else {
        uLongf out = (uLongf)bytes;
        if (uncompress(static_cast<Bytef *>(dst), &out, packed.data(), (uLong)n) != Z_OK || out != bytes)
            bad("zlib decompression failed");
    }
}

float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31, exp = (h >> 10) & 0x1f, man = h & 0x3ff;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {  // subnormal
            int e = -1;
            uint32_t m = man;
            do {
                ++e;
                m <<= 1;
            } while (!(m & 0x400));
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((m & 0x3ff) << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7f800000u | (man << 13);
    } else {
        bits = sign | ((exp - 15 + 127) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// io::readCompressedValues: n node values, of which only the active ones may be
// stored (inactive ones then rebuilt from up to two inactive values and a
// selection mask).
void readCompressedValues(Stream &s, float *dst, size_t n, const Mask &valueMask, uint32_t compression,
                          float background, bool half) {
    const int8_t meta = s.get<int8_t>();
    if (meta < kNoMaskOrInactiveVals || meta > kNoMaskAndAllVals) bad("bad value-compression flag");
    float inactive1 = background;
    float inactive0 = meta == kNoMaskOrInactiveVals ? background : -background;
    if (meta == kNoMaskAndOneInactiveVal || meta == kMaskAndOneInactiveVal || meta == kMaskAndTwoInactiveVals) {
        inactive0 = s.get<float>();
        if (meta == kMaskAndTwoInactiveVals) inactive1 = s.get<float>();
    }
    Mask sel;
    const bool has_sel = meta == kMaskAndNoInactiveVals || meta == kMaskAndOneInactiveVal || meta == kMaskAndTwoInactiveVals;
    if (has_sel) sel.load(s, n);
    size_t count = n;
    if ((compression & kCompressMask) && meta != kNoMaskAndAllVals) count = valueMask.count();
    std::vector<float> tmp(count + 1);  // + 1: a valid destination when no value is stored
    if (half) {
        std::vector<uint16_t> h(count + 1);
        readData(s, h.data(), count * 2, compression);
        for (size_t i = 0; i < count; ++i) tmp[i] = half_to_float(h[i]);
    } else {
        readData(s, tmp.data(), count * 4, compression);
    }
    if (count == n) {
        std::copy(tmp.begin(), tmp.begin() + n, dst);
        return;
    }
    for (size_t i = 0, k = 0; i < n; ++i) dst[i] = valueMask.on(i) ? tmp[k++] : (has_sel && sel.on(i) ? inactive1 : inactive0);
}

struct Leaf {
    std::array<int32_t, 3> o;
    Mask vmask;
    std::vector<float> v;
};
struct Int4 {
    std::array<int32_t, 3> o;
    Mask cmask, vmask;
    std::vector<float> v;  // 4096 tile values
    std::vector<size_t> leaves;
};
struct Int5 {
    std::array<int32_t, 3> o;
    Mask cmask, vmask;
    std::vector<float> v;  // 32768 tile values
    std::vector<size_t> kids;
};
struct RootTile {
    std::array<int32_t, 3> o;
    float value;
    bool active;
};
struct Tree {
    float background = 0.0f;
    std::vector<RootTile> tiles;
    std::vector<Int5> n5;
    std::vector<Int4> n4;
    std::vector<Leaf> leaves;
};

// origin of slot i of a node with 2^log2 slots per axis, each slot `span` voxels wide
std::array<int32_t, 3> slot_origin(const std::array<int32_t, 3> &o, size_t i, int log2, int span) {
    const size_t m = ((size_t)1 << log2) - 1;
    return {o[0] + (int32_t)((i >> (2 * log2)) & m) * span, o[1] + (int32_t)((i >> log2) & m) * span,
            o[2] + (int32_t)(i & m) * span};
}

void readTopology(Stream &s, Tree &t, uint32_t compression, bool half) {
    if (s.get<int32_t>() != 1) bad("expected one buffer per node");
    t.background = s.get<float>();
    const uint32_t ntiles = s.get<uint32_t>(), nkids = s.get<uint32_t>();
    for (uint32_t k = 0; k < ntiles; ++k) {
        RootTile rt;
        for (int a = 0; a < 3; ++a) rt.o[a] = s.get<int32_t>();
        rt.value = s.get<float>();
        rt.active = s.get<uint8_t>() != 0;
        t.tiles.push_back(rt);
    }
    for (uint32_t k = 0; k < nkids; ++k) {
        Int5 n;
        for (int a = 0; a < 3; ++a) n.o[a] = s.get<int32_t>();
        if ((n.o[0] | n.o[1] | n.o[2]) & 4095) bad("root child origin not on the 4096 lattice");
        n.cmask.load(s, 32768);
        n.vmask.load(s, 32768);
        n.v.resize(32768);
        readCompressedValues(s, n.v.data(), 32768, n.vmask, compression, t.background, half);
        for (size_t i = 0; i < 32768; ++i) {
            if (!n.cmask.on(i)) continue;
            Int4 c;
            c.o = slot_origin(n.o, i, 5, 128);
            c.cmask.load(s, 4096);
            c.vmask.load(s, 4096);
            c.v.resize(4096);
            readCompressedValues(s, c.v.data(), 4096, c.vmask, compression, t.background, half);
            for (size_t j = 0; j < 4096; ++j) {
                if (!c.cmask.on(j)) continue;
                Leaf l;
                l.o = slot_origin(c.o, j, 4, 8);
                l.vmask.load(s, 512);
                c.leaves.push_back(t.leaves.size());
                t.leaves.push_back(std::move(l));
            }
            n.kids.push_back(t.n4.size());
            t.n4.push_back(std::move(c));
        }
        t.n5.push_back(std::move(n));
    }
}

// leaf buffers, in the root's (x, y, z)-sorted child order and ascending offsets below
void readBuffers(Stream &s, Tree &t, uint32_t compression, bool half) {
    std::vector<size_t> order(t.n5.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return t.n5[a].o < t.n5[b].o; });
    for (size_t k : order)
        for (size_t c : t.n5[k].kids)
            for (size_t li : t.n4[c].leaves) {
                Leaf &l = t.leaves[li];
                l.vmask.load(s, 512);
                l.v.resize(512);
                readCompressedValues(s, l.v.data(), 512, l.vmask, compression, t.background, half);
            }
}

struct Descriptor {
    std::string name, type, parent;
    int64_t grid_pos = 0, block_pos = 0, end_pos = 0;
};

std::vector<Descriptor> readHeader(Stream &s, uint32_t *version) {
    if (s.get<int64_t>() != kMagic) bad("not a vdb file (bad magic number)");
    *version = s.get<uint32_t>();
    if (*version < kMinVersion) bad("file format version " + std::to_string(*version) + " predates 222");
    (void)s.get<uint32_t>();  // library major
    (void)s.get<uint32_t>();  // library minor
    const bool offsets = s.get<uint8_t>() != 0;
    s.skip(36);  // uuid (ASCII)
    // file-level metadata
    const uint32_t nmeta = s.get<uint32_t>();
    for (uint32_t k = 0; k < nmeta; ++k) {
        (void)s.str();
        (void)s.str();
        s.skip(s.get<uint32_t>());
    }
    if (!offsets) bad("streams without grid offsets are not supported");
    const int32_t ngrids = s.get<int32_t>();
    if (ngrids < 0 || ngrids > 1 << 20) bad("bad grid count");
    std::vector<Descriptor> out;
    for (int32_t k = 0; k < ngrids; ++k) {
        Descriptor d;
        d.name = s.str();
        const size_t sep = d.name.find(kSuffixSep);
        if (sep != std::string::npos) d.name.resize(sep);
        d.type = s.str();
        d.parent = s.str();
        d.grid_pos = s.get<int64_t>();
        d.block_pos = s.get<int64_t>();
        d.end_pos = s.get<int64_t>();
        out.push_back(d);
        s.seek(d.end_pos);
    }
    return out;
}

// index -> world of the supported linear maps: a uniform scale and a translation
void readTransform(Stream &s, VdbFloatGrid &g) {
    const std::string type = s.str();
    double t[3] = {0, 0, 0}, sc[3] = {1, 1, 1};
    auto vec = [&](double *v) {
        for (int a = 0; a < 3; ++a) v[a] = s.get<double>();
    };
    double skip[3];
    if (type == "TranslationMap") {
        vec(t);
    } else if (type == "ScaleMap" || type == "UniformScaleMap") {
        vec(sc);
        for (int k = 0; k < 4; ++k) vec(skip);  // voxel size, inverse scale, 1/s^2, 1/(2s)
    } else if (type == "ScaleTranslateMap" || type == "UniformScaleTranslateMap") {
        vec(t);
        vec(sc);
        for (int k = 0; k < 4; ++k) vec(skip);
    } else if (type == "AffineMap") {
        double m[16];
        for (double &x : m) x = s.get<double>();
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                if (r != c && std::fabs(m[4 * r + c]) > 1e-12 * std::fabs(m[0])) bad("rotated or sheared AffineMap");
        sc[0] = m[0];
        sc[1] = m[5];
        sc[2] = m[10];
        for (int a = 0; a < 3; ++a) t[a] = m[12 + a];
    } else {
        bad("unsupported transform " + type);
    }
    if (!(sc[0] > 0) || std::fabs(sc[1] - sc[0]) > 1e-9 * sc[0] || std::fabs(sc[2] - sc[0]) > 1e-9 * sc[0])
        bad("voxels must be uniform cubes with a positive size");
    g.voxel_size = sc[0];
    for (int a = 0; a < 3; ++a) g.translation[a] = t[a];
}

void grow_box(int32_t lo[3], int32_t hi[3], const std::array<int32_t, 3> &o, int32_t span) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], o[a]);
        hi[a] = std::max(hi[a], o[a] + span - 1);
    }
}

}  // namespace

std::vector<std::string> vdbGridNames(const std::string &path) {
    Stream s(path);
    uint32_t version = 0;
    std::vector<std::string> names;
    for (const Descriptor &d : readHeader(s, &version)) names.push_back(d.name);
    return names;
}

VdbFloatGrid readVdbFloatGrid(const std::string &path, const std::string &name, const int32_t *box_min,
                              const int32_t *box_dims) {
    Stream s(path);
    uint32_t version = 0;
    const std::vector<Descriptor> ds = readHeader(s, &version);
    const Descriptor *d = nullptr;
    for (const Descriptor &x : ds)
        if (x.name == name) {
            d = &x;
            break;
        }
    if (!d) bad("no grid named " + name + " in " + path);
    if (!d->parent.empty()) bad("grid " + name + " is an instance of " + d->parent + " (unsupported)");
    std::string type = d->type;
    const std::string half_suffix = "_HalfFloat";
    bool half = false;
    if (type.size() > half_suffix.size() && type.compare(type.size() - half_suffix.size(), half_suffix.size(), half_suffix) == 0) {
        half = true;
        type.resize(type.size() - half_suffix.size());
    }
    if (type != "Tree_float_5_4_3") bad("grid " + name + " is a " + d->type + ", not a FloatGrid");
    VdbFloatGrid g;
    g.name = name;
    s.seek(d->grid_pos);
    const uint32_t compression = s.get<uint32_t>();
    const uint32_t nmeta = s.get<uint32_t>();
    for (uint32_t k = 0; k < nmeta; ++k) {
        (void)s.str();
        (void)s.str();
        s.skip(s.get<uint32_t>());
    }
    readTransform(s, g);
    Tree t;
    readTopology(s, t, compression, half);
    s.seek(d->block_pos);
    readBuffers(s, t, compression, half);
    g.background = t.background;
    g.leaves = t.leaves.size();
    // active voxel bounding box
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (const RootTile &rt : t.tiles)
        if (rt.active) grow_box(lo, hi, rt.o, 4096);
    for (const Int5 &n : t.n5)
        for (size_t i = 0; i < 32768; ++i)
            if (!n.cmask.on(i) && n.vmask.on(i)) grow_box(lo, hi, slot_origin(n.o, i, 5, 128), 128);
    for (const Int4 &n : t.n4)
        for (size_t i = 0; i < 4096; ++i)
            if (!n.cmask.on(i) && n.vmask.on(i)) grow_box(lo, hi, slot_origin(n.o, i, 4, 8), 8);
    for (const Leaf &l : t.leaves)
        for (size_t i = 0; i < 512; ++i)
            if (l.vmask.on(i)) grow_box(lo, hi, slot_origin(l.o, i, 3, 1), 1);
    g.empty = lo[0] > hi[0];
    if (!g.empty)
        for (int a = 0; a < 3; ++a) {
            g.active_min[a] = lo[a];
            g.active_max[a] = hi[a];
        }
    // dense values over the box
    for (int a = 0; a < 3; ++a) {
        g.dense_min[a] = box_min ? box_min[a] : (g.empty ? 0 : lo[a]);
        g.dims[a] = box_dims ? box_dims[a] : (g.empty ? 1 : hi[a] - lo[a] + 1);
        if (g.dims[a] < 1) bad("empty dense box");
    }
    const size_t total = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    if (total > ((size_t)1 << 33)) bad("the active box holds more than 2^33 voxels");
    g.dense.assign(total, t.background);
    // the region [o, o + span) clipped to the box; vals: one value (leaf == false)
    // or a leaf's 512 values in x-major order
    auto fill = [&](const std::array<int32_t, 3> &o, int32_t span, const float *vals, bool leaf) {
        int32_t a0[3], a1[3];
        for (int a = 0; a < 3; ++a) {
            a0[a] = std::max(o[a], g.dense_min[a]);
            a1[a] = (int32_t)std::min<int64_t>((int64_t)o[a] + span, (int64_t)g.dense_min[a] + g.dims[a]);
            if (a0[a] >= a1[a]) return;
        }
        for (int32_t z = a0[2]; z < a1[2]; ++z)
            for (int32_t y = a0[1]; y < a1[1]; ++y) {
                float *row = &g.dense[((size_t)(z - g.dense_min[2]) * g.dims[1] + (size_t)(y - g.dense_min[1])) * g.dims[0]];
                for (int32_t x = a0[0]; x < a1[0]; ++x)
                    row[x - g.dense_min[0]] = leaf ? vals[((x - o[0]) << 6) + ((y - o[1]) << 3) + (z - o[2])] : vals[0];
            }
    };
    for (const RootTile &rt : t.tiles) fill(rt.o, 4096, &rt.value, false);
    for (const Int5 &n : t.n5)
        for (size_t i = 0; i < 32768; ++i)
            if (!n.cmask.on(i)) fill(slot_origin(n.o, i, 5, 128), 128, &n.v[i], false);
    for (const Int4 &n : t.n4)
        for (size_t i = 0; i < 4096; ++i)
            if (!n.cmask.on(i)) fill(slot_origin(n.o, i, 4, 8), 8, &n.v[i], false);
    for (const Leaf &l : t.leaves) fill(l.o, 8, l.v.data(), true);
    g.tiles = t.tiles.size();
    return g;
}

VdbVolume readVdbVolume(const std::string &path) {
    VdbVolume v;
    const std::vector<std::string> names = vdbGridNames(path);
    if (std::find(names.begin(), names.end(), "density") == names.end()) bad("no grid named density in " + path);
    v.density = readVdbFloatGrid(path, "density");
    if (std::find(names.begin(), names.end(), "temperature") == names.end()) return v;
    const VdbFloatGrid &d = v.density;
    VdbFloatGrid t = readVdbFloatGrid(path, "temperature", d.dense_min, d.dims);
    v.has_temperature = true;
    const bool same = t.voxel_size == d.voxel_size && t.translation[0] == d.translation[0] &&
                      t.translation[1] == d.translation[1] && t.translation[2] == d.translation[2];
    if (same) {
        v.temperature = std::move(t.dense);
        return v;
    }
    // another transform: BoxSampler at each density voxel centre, over the
    // temperature voxels those samples reach (stored values, active or not)
    auto to_temp = [&](int a, double idx) {
        return (d.translation[a] + d.voxel_size * idx - t.translation[a]) / t.voxel_size;
    };
    int32_t tmin[3], tdims[3];
    for (int a = 0; a < 3; ++a) {
        const double u0 = to_temp(a, d.dense_min[a]), u1 = to_temp(a, (double)d.dense_min[a] + d.dims[a] - 1);
        const double lo = std::floor(std::min(u0, u1)), hi = std::floor(std::max(u0, u1)) + 1;
        if (!(lo > -2e9 && hi < 2e9)) bad("temperature transform maps the density outside the index range");
        tmin[a] = (int32_t)lo;
        tdims[a] = (int32_t)(hi - lo) + 1;
    }
    t = readVdbFloatGrid(path, "temperature", tmin, tdims);
    v.temperature_resampled = true;
    v.temperature.assign(d.dense.size(), t.background);
    auto at = [&](int64_t x, int64_t y, int64_t z) -> double {
        x -= t.dense_min[0];
        y -= t.dense_min[1];
        z -= t.dense_min[2];
        if (x < 0 || y < 0 || z < 0 || x >= t.dims[0] || y >= t.dims[1] || z >= t.dims[2]) return t.background;
        return t.dense[((size_t)z * t.dims[1] + (size_t)y) * t.dims[0] + (size_t)x];
    };
    for (int32_t k = 0; k < d.dims[2]; ++k)
        for (int32_t j = 0; j < d.dims[1]; ++j)
            for (int32_t i = 0; i < d.dims[0]; ++i) {
                const int32_t idx[3] = {d.dense_min[0] + i, d.dense_min[1] + j, d.dense_min[2] + k};
                double u[3], f[3];
                int64_t c[3];
                for (int a = 0; a < 3; ++a) {
                    u[a] = to_temp(a, idx[a]);
                    c[a] = (int64_t)std::floor(u[a]);
                    f[a] = u[a] - (double)c[a];
                }
                double s = 0.0;
                for (int corner = 0; corner < 8; ++corner) {
                    const int dx = corner & 1, dy = (corner >> 1) & 1, dz = corner >> 2;
                    const double w = (dx ? f[0] : 1 - f[0]) * (dy ? f[1] : 1 - f[1]) * (dz ? f[2] : 1 - f[2]);
                    if (w != 0.0) s += w * at(c[0] + dx, c[1] + dy, c[2] + dz);
                }
                v.temperature[((size_t)k * d.dims[1] + j) * d.dims[0] + i] = (float)s;
            }
    return v;
}

}  // namespace lightgraph

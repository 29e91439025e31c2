"""Writer of OpenVDB files holding FloatGrids, for the tests of the .vdb reader
(lightgraph_amd/host/vdb.cpp), from OpenVDB's published file layout (format
version 224): header, file metadata, grid descriptors with stream offsets, and
per grid the compression flags, metadata, transform, Tree_float_5_4_3 topology
and leaf buffers.  Node values are written the way OpenVDB's
writeCompressedValues does: with mask compression the inactive values are
classified (MaskCompress) into one of seven layouts and only active values are
stored.  Values go out raw, zlib-compressed or blosc-compressed (c-blosc through
ctypes, when the library is found), as float32 or half."""
import ctypes
import ctypes.util
import os
import struct
import zlib

import numpy as np

ZIP, ACTIVE_MASK, BLOSC = 0x1, 0x2, 0x4
(NO_MASK_OR_INACTIVE_VALS, NO_MASK_AND_MINUS_BG, NO_MASK_AND_ONE_INACTIVE_VAL, MASK_AND_NO_INACTIVE_VALS,
 MASK_AND_ONE_INACTIVE_VAL, MASK_AND_TWO_INACTIVE_VALS, NO_MASK_AND_ALL_VALS) = range(7)
MAGIC = 0x56444220


def _blosc():
    for name in ("libblosc.so.1", "libblosc.so", "/opt/conda/lib/libblosc.so.1", ctypes.util.find_library("blosc")):
        if not name:
            continue
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            continue
        lib.blosc_init()
        return lib
    return None


BLOSC_LIB = _blosc()


def _s(x):
    b = x.encode() if isinstance(x, str) else x
    return struct.pack("<I", len(b)) + b


def _mask_bytes(bits):
    bits = np.asarray(bits, bool)
    return np.packbits(bits.reshape(-1, 8), axis=1, bitorder="little").tobytes()


def _data(raw, compression):
    """io::writeData: raw, or an int64 size and zlib / blosc bytes (negative size:
    stored uncompressed because compression did not shrink it)."""
    if not compression & (ZIP | BLOSC):
        return raw
    if compression & BLOSC:
        if BLOSC_LIB is None:
            raise RuntimeError("libblosc not found")
        dst = ctypes.create_string_buffer(len(raw) + 16)
        n = BLOSC_LIB.blosc_compress(9, 1, 4, len(raw), raw, dst, len(raw) + 16)
        if n > 0:
            return struct.pack("<q", n) + dst.raw[:n]
        return struct.pack("<q", -len(raw)) + raw
    z = zlib.compress(raw, 9)
    if len(z) < len(raw):
        return struct.pack("<q", len(z)) + z
    return struct.pack("<q", -len(raw)) + raw


def _eq(a, b):
    return np.float32(a).tobytes() == np.float32(b).tobytes()


def mask_compress(values, vmask, cmask, bg):
    """OpenVDB's MaskCompress: (metadata, inactive value 0, inactive value 1)."""
    inact = [bg, bg]
    n_unique = 0
    for i in np.nonzero(~vmask)[0]:
        if n_unique >= 3:
            break
        if cmask[i]:
            continue
        v = values[i]
        unique = not ((n_unique > 0 and _eq(v, inact[0])) or (n_unique > 1 and _eq(v, inact[1])))
        if unique:
            if n_unique < 2:
                inact[n_unique] = v
            n_unique += 1
    meta = NO_MASK_OR_INACTIVE_VALS
    if n_unique == 1:
        if not _eq(inact[0], bg):
            meta = NO_MASK_AND_MINUS_BG if _eq(inact[0], -bg) else NO_MASK_AND_ONE_INACTIVE_VAL
    elif n_unique == 2:
        if not _eq(inact[0], bg) and not _eq(inact[1], bg):
            meta = MASK_AND_TWO_INACTIVE_VALS
        elif _eq(inact[1], bg):
            meta = MASK_AND_NO_INACTIVE_VALS if _eq(inact[0], -bg) else MASK_AND_ONE_INACTIVE_VAL
        elif _eq(inact[0], bg):
            meta = MASK_AND_NO_INACTIVE_VALS if _eq(inact[1], -bg) else MASK_AND_ONE_INACTIVE_VAL
            inact = [inact[1], inact[0]]
    elif n_unique > 2:
        meta = NO_MASK_AND_ALL_VALS
    return meta, np.float32(inact[0]), np.float32(inact[1])


def compressed_values(values, vmask, cmask, bg, compression, half, stats=None):
    """io::writeCompressedValues of one node's values."""
    values = np.asarray(values, np.float32)
    n = len(values)
    out = b""
    store = values
    if not compression & ACTIVE_MASK:
        meta = NO_MASK_AND_ALL_VALS
        out += struct.pack("<b", meta)
    else:
        meta, i0, i1 = mask_compress(values, vmask, cmask, bg)
        out += struct.pack("<b", meta)
        if meta in (NO_MASK_AND_ONE_INACTIVE_VAL, MASK_AND_ONE_INACTIVE_VAL, MASK_AND_TWO_INACTIVE_VALS):
            cast = (lambda x: np.float32(np.float16(x))) if half else np.float32
            out += cast(i0).tobytes()
            if meta == MASK_AND_TWO_INACTIVE_VALS:
                out += cast(i1).tobytes()
        if meta != NO_MASK_AND_ALL_VALS:
            store = values[vmask]
            if meta in (MASK_AND_NO_INACTIVE_VALS, MASK_AND_ONE_INACTIVE_VAL, MASK_AND_TWO_INACTIVE_VALS):
                sel = np.zeros(n, bool)
                for i in range(n):
                    if not vmask[i] and _eq(values[i], i1):
                        sel[i] = True
                out += _mask_bytes(sel)
    if stats is not None:
        stats[meta] = stats.get(meta, 0) + 1
    raw = store.astype(np.float16).tobytes() if half else store.astype(np.float32).tobytes()
    return out + _data(raw, compression)


class Grid:
    """A FloatGrid: leaves {origin: (values[512] x-major, active[512])}, 16^3-node
    tiles {origin of the 8^3 slot: (value, active)}, 32^3-node tiles {origin of the
    128^3 slot: (value, active)}, root tiles [(origin, value, active)]."""

    def __init__(self, name, background=0.0, voxel_size=1.0, translation=(0.0, 0.0, 0.0), map_type=None):
        self.name = name
        self.background = np.float32(background)
        self.voxel_size = voxel_size
        self.translation = tuple(float(t) for t in translation)
        self.map_type = map_type or ("UniformScaleTranslateMap" if any(self.translation) else "UniformScaleMap")
        self.leaves, self.tiles8, self.tiles128, self.root_tiles = {}, {}, {}, []
        self.affine = None  # 4x4 row-vector matrix for AffineMap

    @classmethod
    def from_dense(cls, name, dense, active, index_min, tile_uniform=False, **kw):
        """Leaves (or, with tile_uniform, 8^3 tiles where a block is active and
        uniform) for every 8^3 block of [index_min, index_min + shape) holding an
        active voxel or a value other than the background; dense is z-y-x."""
        g = cls(name, **kw)
        nz, ny, nx = dense.shape
        x0, y0, z0 = index_min
        bx0, by0, bz0 = (x0 // 8) * 8, (y0 // 8) * 8, (z0 // 8) * 8
        for bx in range(bx0, x0 + nx, 8):
            for by in range(by0, y0 + ny, 8):
                for bz in range(bz0, z0 + nz, 8):
                    vals = np.full((8, 8, 8), g.background, np.float32)  # x, y, z
                    act = np.zeros((8, 8, 8), bool)
                    xs, ys, zs = (slice(max(b, o) - b, min(b + 8, o + n) - b) for b, o, n in
                                  ((bx, x0, nx), (by, y0, ny), (bz, z0, nz)))
                    src = (slice(zs.start + bz - z0, zs.stop + bz - z0), slice(ys.start + by - y0, ys.stop + by - y0),
                           slice(xs.start + bx - x0, xs.stop + bx - x0))
                    vals[xs, ys, zs] = dense[src].transpose(2, 1, 0)
                    act[xs, ys, zs] = active[src].transpose(2, 1, 0)
                    if not act.any() and np.all(vals == g.background):
                        continue
                    if tile_uniform and act.all() and np.all(vals == vals.flat[0]):
                        g.tiles8[(bx, by, bz)] = (float(vals.flat[0]), True)
                    else:
                        g.leaves[(bx, by, bz)] = (vals.reshape(-1), act.reshape(-1))
        return g

    def _transform(self):
        t, s = self.translation, (self.voxel_size,) * 3
        vec = lambda v: struct.pack("<3d", *v)
        inv = tuple(1.0 / x for x in s)
        tail = vec(s) + vec(inv) + vec(tuple(x * x for x in inv)) + vec(tuple(0.5 * x for x in inv))
        m = self.map_type
        if m in ("UniformScaleMap", "ScaleMap"):
            body = vec(s) + tail
        elif m in ("UniformScaleTranslateMap", "ScaleTranslateMap"):
            body = vec(t) + vec(s) + tail
        elif m == "TranslationMap":
            body = vec(t)
        elif m == "AffineMap":
            a = self.affine
            if a is None:
                a = np.diag([self.voxel_size] * 3 + [1.0])
                a[3, :3] = t
            body = np.asarray(a, "<f8").tobytes()
        else:
            body = b""
        return _s(m) + body

    def _nodes(self):
        """root children {origin5: {origin4: [leaf origins]}}"""
        n5 = {}
        for o in list(self.leaves) + list(self.tiles8):
            o5 = tuple((c >> 12) << 12 for c in o)
            o4 = tuple((c >> 7) << 7 for c in o)
            n5.setdefault(o5, {}).setdefault(o4, [])
            if o in self.leaves:
                n5[o5][o4].append(o)
        for o in self.tiles128:
            n5.setdefault(tuple((c >> 12) << 12 for c in o), {})
        return n5

    def write_body(self, compression, half, stats=None):
        """(bytes up to the buffers, buffer bytes)"""
        bg = self.background
        meta = [("class", "string", b"fog volume"), ("file_delayed_load", "__delayedload", b"\x00" * 24),
                ("file_bbox_min", "vec3i", struct.pack("<3i", 0, 0, 0))]
        head = struct.pack("<I", compression) + struct.pack("<I", len(meta))
        for k, t, v in meta:
            head += _s(k) + _s(t) + struct.pack("<I", len(v)) + v
        head += self._transform()
        topo = struct.pack("<i", 1) + bg.tobytes()
        n5 = self._nodes()
        topo += struct.pack("<II", len(self.root_tiles), len(n5))
        for o, v, a in self.root_tiles:
            topo += struct.pack("<3i", *o) + np.float32(v).tobytes() + struct.pack("<B", int(a))
        bufs = b""
        for o5 in sorted(n5):
            topo += struct.pack("<3i", *o5)
            c5, v5 = np.zeros(32768, bool), np.zeros(32768, bool)
            vals5 = np.full(32768, bg, np.float32)
            kids = {}
            for o4 in n5[o5]:
                i = (((o4[0] - o5[0]) >> 7) << 10) + (((o4[1] - o5[1]) >> 7) << 5) + ((o4[2] - o5[2]) >> 7)
                c5[i] = True
                vals5[i] = 0.0
                kids[i] = o4
            for o, (v, a) in self.tiles128.items():
                if tuple((c >> 12) << 12 for c in o) != o5:
                    continue
                i = (((o[0] - o5[0]) >> 7) << 10) + (((o[1] - o5[1]) >> 7) << 5) + ((o[2] - o5[2]) >> 7)
                vals5[i], v5[i] = v, a
            topo += _mask_bytes(c5) + _mask_bytes(v5) + compressed_values(vals5, v5, c5, bg, compression, half, stats)
            for i in sorted(kids):
                o4 = kids[i]
                c4, v4 = np.zeros(4096, bool), np.zeros(4096, bool)
                vals4 = np.full(4096, bg, np.float32)
                leaves = {}
                for o in n5[o5][o4]:
                    j = (((o[0] - o4[0]) >> 3) << 8) + (((o[1] - o4[1]) >> 3) << 4) + ((o[2] - o4[2]) >> 3)
                    c4[j] = True
                    vals4[j] = 0.0
                    leaves[j] = o
                for o, (v, a) in self.tiles8.items():
                    if tuple((c >> 7) << 7 for c in o) != o4:
                        continue
                    j = (((o[0] - o4[0]) >> 3) << 8) + (((o[1] - o4[1]) >> 3) << 4) + ((o[2] - o4[2]) >> 3)
                    vals4[j], v4[j] = v, a
                topo += _mask_bytes(c4) + _mask_bytes(v4) + compressed_values(vals4, v4, c4, bg, compression, half,
                                                                                stats)
                for j in sorted(leaves):
                    v, a = self.leaves[leaves[j]]
                    topo += _mask_bytes(a)
                    bufs += _mask_bytes(a) + compressed_values(v, a, np.zeros(512, bool), bg, compression, half, stats)
        return head + topo, bufs

    def dense_expected(self, box_min, dims, half=False):
        """What a reader returns over [box_min, box_min + dims): z-y-x values."""
        out = np.full(tuple(dims[::-1]), self.background, np.float32)

        def put(o, span, vals):
            lo = [max(o[a], box_min[a]) for a in range(3)]
            hi = [min(o[a] + span, box_min[a] + dims[a]) for a in range(3)]
            if any(lo[a] >= hi[a] for a in range(3)):
                return
            sl = tuple(slice(lo[a] - box_min[a], hi[a] - box_min[a]) for a in (2, 1, 0))
            if np.ndim(vals) == 0:
                out[sl] = vals
            else:
                v = vals.reshape(8, 8, 8)[lo[0] - o[0]:hi[0] - o[0], lo[1] - o[1]:hi[1] - o[1], lo[2] - o[2]:hi[2] - o[2]]
                out[sl] = v.transpose(2, 1, 0)

        for o, v, a in self.root_tiles:
            put(o, 4096, np.float32(v))
        for o, (v, a) in self.tiles128.items():
            put(o, 128, np.float32(v))
        for o, (v, a) in self.tiles8.items():
            put(o, 8, np.float32(v))
        for o, (v, a) in self.leaves.items():
            vv = np.asarray(v, np.float32)
            if half:
                vv = vv.astype(np.float16).astype(np.float32)
            put(o, 8, vv)
        return out

    def active_box(self):
        lo, hi = [1 << 31] * 3, [-(1 << 31)] * 3

        def grow(o, span):
            for a in range(3):
                lo[a] = min(lo[a], o[a])
                hi[a] = max(hi[a], o[a] + span - 1)

        for o, v, a in self.root_tiles:
            if a:
                grow(o, 4096)
        for o, (v, a) in self.tiles128.items():
            if a:
                grow(o, 128)
        for o, (v, a) in self.tiles8.items():
            if a:
                grow(o, 8)
        for o, (v, a) in self.leaves.items():
            idx = np.nonzero(np.asarray(a).reshape(8, 8, 8))
            if len(idx[0]):
                grow((o[0] + int(idx[0].min()), o[1] + int(idx[1].min()), o[2] + int(idx[2].min())), 1)
                grow((o[0] + int(idx[0].max()), o[1] + int(idx[1].max()), o[2] + int(idx[2].max())), 1)
        return lo, hi


def write(path, grids, compression=ZIP | ACTIVE_MASK, half=False, version=224, raw_grids=(), stats=None):
    """grids: Grid objects; raw_grids: (name, type, body bytes) written as opaque
    grids of another type (their descriptors are what a reader skips past)."""
    out = bytearray(struct.pack("<q", MAGIC) + struct.pack("<III", version, 11, 0) + b"\x01")
    out += b"0123abcd-0000-4000-8000-0123456789ab"
    out += struct.pack("<I", 1) + _s("creator") + _s("string") + struct.pack("<I", 4) + b"test"
    entries = [(n, t, b) for n, t, b in raw_grids] + [g for g in grids]
    out += struct.pack("<i", len(entries))
    seen = {}
    for e in entries:
        if isinstance(e, Grid):
            name, typ = e.name, "Tree_float_5_4_3" + ("_HalfFloat" if half else "")
            head, bufs = e.write_body(compression, half, stats)
        else:
            name, typ, head = e
            bufs = b""
        unique = name if name not in seen else f"{name}\x1e{seen[name]}"
        seen[name] = seen.get(name, 0) + 1
        out += _s(unique) + _s(typ) + _s("")
        pos = len(out)
        out += b"\0" * 24
        grid_pos = len(out)
        out += head
        block_pos = len(out)
        out += bufs
        end_pos = len(out)
        struct.pack_into("<qqq", out, pos, grid_pos, block_pos, end_pos)
    with open(path, "wb") as f:
        f.write(bytes(out))
    return path

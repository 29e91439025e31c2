"""Minimal scanline OpenEXR reader/writer for the tests (NONE and ZIP/ZIPS
compression, HALF/FLOAT channels), from the published file layout."""
import struct
import zlib

import numpy as np


def _restore(t):
    t = np.frombuffer(t, np.uint8).astype(np.int32)
    d = np.cumsum(np.concatenate([t[:1], t[1:] - 128])) % 256
    d = d.astype(np.uint8)
    n = len(d)
    out = np.empty(n, np.uint8)
    h = (n + 1) // 2
    out[0::2] = d[:h]
    out[1::2] = d[h:]
    return out.tobytes()


def _prepare(raw):
    r = np.frombuffer(raw, np.uint8)
    t = np.concatenate([r[0::2], r[1::2]]).astype(np.int32)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) % 256
    return d.astype(np.uint8).tobytes()


def read(path):
    """Returns (channels dict name -> float32 HxW array, header dict)."""
    b = open(path, "rb").read()
    assert struct.unpack_from("<I", b, 0)[0] == 20000630
    p = 8
    hdr, chans = {}, []
    while True:
        e = b.index(b"\0", p)
        name = b[p:e].decode()
        p = e + 1
        if not name:
            break
        e = b.index(b"\0", p)
        typ = b[p:e].decode()
        p = e + 1
        size = struct.unpack_from("<i", b, p)[0]
        p += 4
        v = b[p:p + size]
        p += size
        if name == "channels":
            q = 0
            while v[q] != 0:
                e = v.index(b"\0", q)
                cn = v[q:e].decode()
                pt = struct.unpack_from("<i", v, e + 1)[0]
                chans.append((cn, pt))
                q = e + 1 + 16
        elif name == "compression":
            hdr["compression"] = v[0]
        elif name == "dataWindow":
            hdr["dataWindow"] = struct.unpack("<4i", v)
        else:
            hdr[name] = (typ, v)
    x0, y0, x1, y1 = hdr["dataWindow"]
    W, H = x1 - x0 + 1, y1 - y0 + 1
    comp = hdr["compression"]
    lines = 16 if comp == 3 else 1
    nch = (H + lines - 1) // lines
    offs = struct.unpack_from(f"<{nch}Q", b, p)
    sizes = [2 if pt == 1 else 4 for _, pt in chans]
    out = {cn: np.zeros((H, W), np.float32) for cn, _ in chans}
    for c, off in enumerate(offs):
        y, sz = struct.unpack_from("<ii", b, off)
        data = b[off + 8:off + 8 + sz]
        nl = min(lines, H - (y - y0))
        raw_size = nl * W * sum(sizes)
        if comp != 0 and sz != raw_size:
            data = _restore(zlib.decompress(data))
        q = 0
        for l in range(nl):
            for (cn, pt), s in zip(chans, sizes):
                dt = np.float16 if pt == 1 else np.float32
                out[cn][y - y0 + l] = np.frombuffer(data, dt, W, q).astype(np.float32)
                q += W * s
    hdr["channels"] = [cn for cn, _ in chans]
    return out, hdr


def write_rgb(path, rgb, compression=0):
    """Writes an H x W x 3 float32 image as FLOAT R, G, B channels."""
    H, W, _ = rgb.shape
    chans = [("B", 2), ("G", 1), ("R", 0)]

    def attr(name, typ, val):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(val)) + val
    cl = b"".join(cn.encode() + b"\0" + struct.pack("<i", 2) + b"\0" * 4 + struct.pack("<ii", 1, 1)
                  for cn, _ in chans) + b"\0"
    box = struct.pack("<4i", 0, 0, W - 1, H - 1)
    h = struct.pack("<II", 20000630, 2)
    h += attr("channels", "chlist", cl) + attr("compression", "compression", bytes([compression]))
    h += attr("dataWindow", "box2i", box) + attr("displayWindow", "box2i", box)
    h += attr("lineOrder", "lineOrder", b"\0") + attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    h += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)) + attr("screenWindowWidth", "float",
                                                                             struct.pack("<f", 1.0))
    h += b"\0"
    lines = 16 if compression == 3 else 1
    blocks = []
    for y0 in range(0, H, lines):
        raw = b"".join(np.ascontiguousarray(rgb[y, :, k], np.float32).tobytes()
                       for y in range(y0, min(H, y0 + lines)) for _, k in chans)
        data = raw
        if compression:
            z = zlib.compress(_prepare(raw))
            if len(z) < len(raw):
                data = z
        blocks.append(struct.pack("<ii", y0, len(data)) + data)
    pos = len(h) + 8 * len(blocks)
    table = b""
    for bl in blocks:
        table += struct.pack("<Q", pos)
        pos += len(bl)
    open(path, "wb").write(h + table + b"".join(blocks))

"""The .vdb reader (lightgraph_amd/host/vdb.cpp), on the CPU and under the
address and undefined-behaviour sanitizers, over files written by tests/vdbio.py
from OpenVDB's published layout: every value-compression layout of
writeCompressedValues, raw / zlib / blosc / half-float streams, leaves, 16^3-
and 32^3-node tiles, root tiles, negative origins, the linear transforms, the
density + temperature pair the renderer reads (main.cc:488-502) with and
without a shared transform, and malformed files.

No OpenVDB library or .vdb file is available here: these files are written
from the published format, so parity with files written by OpenVDB itself is
unpinned."""
import os
import shutil
import struct
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
import vdbio  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("vdb") / "vdb_dump")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", *SAN, "-I", os.path.join(ROOT, "lightgraph_amd", "host"),
           os.path.join(ROOT, "tests", "host", "vdb_dump.cpp"), os.path.join(ROOT, "lightgraph_amd", "host", "vdb.cpp"),
           "-o", exe, "-lz", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, env=env, timeout=120)
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out, out
    return r


def _fields(stdout):
    f = {}
    for line in stdout.splitlines():
        k, _, v = line.partition(" ")
        f[k] = v.split()
    return f


def read_grid(exe, path, name, tmp_path, box=None):
    out = tmp_path / f"{name}.raw"
    r = _run(exe, "grid", path, name, out, *(list(box[0]) + list(box[1]) if box else []))
    assert r.returncode == 0, r.stderr
    f = _fields(r.stdout)
    dims = [int(x) for x in f["dims"]]
    dense = np.fromfile(out, np.float32).reshape(dims[::-1])
    return f, dense


def _sparse_cloud(rng, shape, half=False):
    """fog-volume-like block: values in (0, 1], active where above 0.35, inactive 0"""
    v = rng.random(shape, dtype=np.float32)
    if half:
        v = np.round(v * 64) / 64
    active = v > 0.35
    v = np.where(active, v, 0.0).astype(np.float32)
    return v, active


@pytest.mark.parametrize("compression,half", [
    (0, False), (vdbio.ACTIVE_MASK, False), (vdbio.ZIP, False), (vdbio.ZIP | vdbio.ACTIVE_MASK, False),
    (vdbio.BLOSC, False), (vdbio.BLOSC | vdbio.ACTIVE_MASK, False), (vdbio.ZIP | vdbio.ACTIVE_MASK, True)])
def test_round_trip_streams(dump, tmp_path, compression, half):
    if compression & vdbio.BLOSC and vdbio.BLOSC_LIB is None:
        pytest.skip("libblosc not found")
    rng = np.random.default_rng(7 + compression)
    shape = (21, 30, 40)  # z, y, x
    v, active = _sparse_cloud(rng, shape, half)
    imin = (-13, 5, -7)
    g = vdbio.Grid.from_dense("density", v, active, imin, voxel_size=0.5, translation=(1.0, -2.0, 3.5))
    path = vdbio.write(str(tmp_path / "c.vdb"), [g], compression=compression, half=half)
    f, dense = read_grid(dump, path, "density", tmp_path)
    lo, hi = g.active_box()
    assert [int(x) for x in f["active_min"]] == lo and [int(x) for x in f["active_max"]] == hi
    assert [int(x) for x in f["dense_min"]] == lo
    dims = [hi[a] - lo[a] + 1 for a in range(3)]
    assert [int(x) for x in f["dims"]] == dims
    assert float(f["voxel_size"][0]) == 0.5 and [float(x) for x in f["translation"]] == [1.0, -2.0, 3.5]
    assert int(f["leaves"][0]) == len(g.leaves)
    np.testing.assert_array_equal(dense, g.dense_expected(lo, dims, half))
    # the box is the active box of the source: the dense values are the source's
    src = v[lo[2] - imin[2]:hi[2] - imin[2] + 1, lo[1] - imin[1]:hi[1] - imin[1] + 1,
            lo[0] - imin[0]:hi[0] - imin[0] + 1]
    np.testing.assert_array_equal(dense, src)


def _leaf(values, active):
    return np.asarray(values, np.float32).reshape(-1), np.asarray(active, bool).reshape(-1)


def test_every_value_layout(dump, tmp_path):
    """one leaf per writeCompressedValues layout, background 0.25"""
    bg = np.float32(0.25)
    rng = np.random.default_rng(3)
    g = vdbio.Grid("density", background=bg)
    act = rng.random(512) < 0.5
    act[0] = True
    vals = rng.random(512).astype(np.float32) + 1.0
    c, d = np.float32(0.75), np.float32(-3.0)
    sel = rng.random(512) < 0.5
    layouts = {
        vdbio.NO_MASK_OR_INACTIVE_VALS: np.where(act, vals, bg),
        vdbio.NO_MASK_AND_MINUS_BG: np.where(act, vals, -bg),
        vdbio.NO_MASK_AND_ONE_INACTIVE_VAL: np.where(act, vals, c),
        vdbio.MASK_AND_NO_INACTIVE_VALS: np.where(act, vals, np.where(sel, bg, -bg)),
        vdbio.MASK_AND_ONE_INACTIVE_VAL: np.where(act, vals, np.where(sel, c, bg)),
        vdbio.MASK_AND_TWO_INACTIVE_VALS: np.where(act, vals, np.where(sel, c, d)),
        vdbio.NO_MASK_AND_ALL_VALS: np.where(act, vals, rng.random(512).astype(np.float32) - 5.0),
    }
    for k, lv in layouts.items():
        g.leaves[(8 * k, 0, 0)] = _leaf(lv, act)
    # the first inactive value seen is the background: the swapped orders too
    first_bg = np.where(act, vals, np.where(sel, c, bg))
    first_bg[np.nonzero(~act)[0][0]] = bg
    g.leaves[(0, 8, 0)] = _leaf(first_bg, act)
    first_bg2 = np.where(act, vals, np.where(sel, -bg, bg))
    first_bg2[np.nonzero(~act)[0][0]] = bg
    g.leaves[(8, 8, 0)] = _leaf(first_bg2, act)
    stats = {}
    path = vdbio.write(str(tmp_path / "layouts.vdb"), [g], compression=vdbio.ZIP | vdbio.ACTIVE_MASK, stats=stats)
    for k in layouts:
        assert stats.get(k, 0) >= 1, (k, stats)
    box = ((0, 0, 0), (56, 16, 8))
    f, dense = read_grid(dump, path, "density", tmp_path, box)
    assert [int(x) for x in f["dims"]] == [56, 16, 8]
    np.testing.assert_array_equal(dense, g.dense_expected(*box))
    assert float(f["background"][0]) == 0.25


def test_tiles_roots_and_negative_origins(dump, tmp_path):
    g = vdbio.Grid("density", voxel_size=2.0)
    rng = np.random.default_rng(11)
    # leaves on both sides of 0 (two 32^3-node children of the root: -4096 and 0)
    for o in [(-8, 0, 0), (0, 0, 0), (-16, -8, 8), (120, 8, -8)]:
        act = rng.random(512) < 0.6
        g.leaves[o] = _leaf(np.where(act, rng.random(512) + 0.5, 0.0), act)
    g.tiles8[(16, 16, 16)] = (0.625, True)       # 16^3-node tile, active
    g.tiles8[(-24, 16, 16)] = (0.125, False)     # inactive, value kept
    g.tiles128[(-128, -128, 128)] = (0.375, True)  # 32^3-node tile
    g.root_tiles.append(((8192, 0, 0), 3.0, False))  # inactive root tile, outside the box
    path = vdbio.write(str(tmp_path / "tiles.vdb"), [g])
    f, dense = read_grid(dump, path, "density", tmp_path)
    lo, hi = g.active_box()
    assert lo == [-128, -128, -8] and hi == [127, 23, 255]
    dims = [hi[a] - lo[a] + 1 for a in range(3)]
    np.testing.assert_array_equal(dense, g.dense_expected(lo, dims))
    assert dense[200 + 8, -100 + 128, -100 + 128] == np.float32(0.375)  # inside the 128^3 tile
    # an explicit box reaching the inactive root tile reads its value
    f2, d2 = read_grid(dump, path, "density", tmp_path, ((8190, 0, 0), (4, 1, 1)))
    np.testing.assert_array_equal(d2.ravel(), [0, 0, 3, 3])


@pytest.mark.parametrize("map_type", ["UniformScaleMap", "UniformScaleTranslateMap", "ScaleMap", "ScaleTranslateMap",
                                      "TranslationMap", "AffineMap"])
def test_transforms(dump, tmp_path, map_type):
    g = vdbio.Grid("density", voxel_size=1.0 if map_type == "TranslationMap" else 0.125,
                   translation=(0.0, 0.0, 0.0) if "Translat" not in map_type and map_type != "AffineMap"
                   else (4.0, -1.5, 0.25), map_type=map_type)
    g.leaves[(0, 0, 0)] = _leaf(np.full(512, 0.5), np.ones(512, bool))
    path = vdbio.write(str(tmp_path / "t.vdb"), [g])
    f, _ = read_grid(dump, path, "density", tmp_path)
    assert float(f["voxel_size"][0]) == g.voxel_size
    assert [float(x) for x in f["translation"]] == list(g.translation)


def _errors(dump, tmp_path, path, name="density"):
    r = _run(dump, "grid", path, name, tmp_path / "x.raw")
    assert r.returncode == 2, r.stdout
    return r.stderr


def test_malformed_and_unsupported(dump, tmp_path):
    g = vdbio.Grid("density")
    g.leaves[(0, 0, 0)] = _leaf(np.full(512, 0.5), np.ones(512, bool))
    good = vdbio.write(str(tmp_path / "g.vdb"), [g], raw_grids=[("vel", "Tree_vec3s_5_4_3", b"\0" * 64)])
    data = open(good, "rb").read()
    bad = tmp_path / "bad_magic.vdb"
    bad.write_bytes(b"\0" + data[1:])
    assert "bad magic" in _errors(dump, tmp_path, bad)
    old = tmp_path / "old.vdb"
    old.write_bytes(data[:8] + struct.pack("<I", 220) + data[12:])
    assert "predates 222" in _errors(dump, tmp_path, old)
    for cut in (20, len(data) // 2, len(data) - 3):
        t = tmp_path / f"trunc{cut}.vdb"
        t.write_bytes(data[:cut])
        assert "vdb:" in _errors(dump, tmp_path, t)
    assert "no grid named temperature" in _errors(dump, tmp_path, good, "temperature")
    assert "not a FloatGrid" in _errors(dump, tmp_path, good, "vel")
    assert "cannot open" in _errors(dump, tmp_path, tmp_path / "missing.vdb")
    sheared = vdbio.Grid("density", map_type="AffineMap")
    sheared.affine = np.eye(4)
    sheared.affine[0, 1] = 0.5
    sheared.leaves = g.leaves
    assert "rotated or sheared" in _errors(dump, tmp_path, vdbio.write(str(tmp_path / "sh.vdb"), [sheared]))
    aniso = vdbio.Grid("density", map_type="AffineMap")
    aniso.affine = np.diag([1.0, 2.0, 1.0, 1.0])
    aniso.leaves = g.leaves
    assert "uniform cubes" in _errors(dump, tmp_path, vdbio.write(str(tmp_path / "an.vdb"), [aniso]))
    unitary = vdbio.Grid("density", map_type="UnitaryMap")
    unitary.leaves = g.leaves
    assert "unsupported transform" in _errors(dump, tmp_path, vdbio.write(str(tmp_path / "u.vdb"), [unitary]))
    # a corrupt value-compression flag
    corrupt = tmp_path / "corrupt.vdb"
    flat = vdbio.write(str(tmp_path / "flat.vdb"), [g], compression=0)
    raw = bytearray(open(flat, "rb").read())
    raw[-(512 * 4 + 1)] = 9  # the leaf buffer's flag byte
    corrupt.write_bytes(bytes(raw))
    assert "bad value-compression flag" in _errors(dump, tmp_path, corrupt)


def test_names_duplicates_and_empty(dump, tmp_path):
    a = vdbio.Grid("density")
    a.leaves[(0, 0, 0)] = _leaf(np.full(512, 0.5), np.ones(512, bool))
    b = vdbio.Grid("density")
    b.leaves[(0, 0, 0)] = _leaf(np.full(512, 0.25), np.ones(512, bool))
    e = vdbio.Grid("empty")
    path = vdbio.write(str(tmp_path / "n.vdb"), [a, b, e], raw_grids=[("vel", "Tree_vec3s_5_4_3", b"\1" * 40)])
    r = _run(dump, "names", path)
    assert r.returncode == 0 and r.stdout.split() == ["grid", "vel", "grid", "density", "grid", "density", "grid",
                                                       "empty"]
    f, dense = read_grid(dump, path, "density", tmp_path)
    assert np.all(dense == 0.5)  # the first grid of that name
    f, dense = read_grid(dump, path, "empty", tmp_path)
    assert f["empty"] == ["1"] and dense.shape == (1, 1, 1) and dense[0, 0, 0] == 0


def _volume(dump, path, tmp_path):
    d, t = tmp_path / "d.raw", tmp_path / "t.raw"
    r = _run(dump, "volume", path, d, t)
    assert r.returncode == 0, r.stderr
    f = _fields(r.stdout)
    dims = [int(x) for x in f["dims"]][::-1]
    return f, np.fromfile(d, np.float32).reshape(dims), np.fromfile(t, np.float32)


def test_density_and_temperature(dump, tmp_path):
    rng = np.random.default_rng(5)
    v, act = _sparse_cloud(rng, (12, 14, 18))
    dens = vdbio.Grid.from_dense("density", v, act, (3, -4, 9), voxel_size=0.5, translation=(1.0, 2.0, 3.0))
    tv = (rng.random((20, 20, 24)) * 1000).astype(np.float32)
    temp = vdbio.Grid.from_dense("temperature", tv, tv > 300, (0, -8, 4), voxel_size=0.5, translation=(1.0, 2.0, 3.0))
    path = vdbio.write(str(tmp_path / "v.vdb"), [temp, dens])
    f, d, t = _volume(dump, path, tmp_path)
    lo, hi = dens.active_box()
    dims = [hi[a] - lo[a] + 1 for a in range(3)]
    assert f["has_temperature"] == ["1"] and f["temperature_resampled"] == ["0"]
    np.testing.assert_array_equal(d, dens.dense_expected(lo, dims))
    np.testing.assert_array_equal(t.reshape(d.shape), temp.dense_expected(lo, dims))
    # a temperature grid on another lattice: BoxSampler at the density's voxel centres
    temp2 = vdbio.Grid.from_dense("temperature", tv, tv > 300, (0, -8, 4), voxel_size=0.75,
                                  translation=(0.4, 1.1, 2.9))
    path2 = vdbio.write(str(tmp_path / "v2.vdb"), [dens, temp2])
    f2, d2, t2 = _volume(dump, path2, tmp_path)
    assert f2["temperature_resampled"] == ["1"]
    # every stored temperature value (inactive ones included), background around them
    tlo, tdims = (-16, -24, -16), (64, 64, 64)
    td = temp2.dense_expected(tlo, tdims).astype(np.float64)
    exp = np.empty(d2.shape, np.float64)
    for k in range(dims[2]):
        for j in range(dims[1]):
            for i in range(dims[0]):
                idx = (lo[0] + i, lo[1] + j, lo[2] + k)
                u = [(dens.translation[a] + 0.5 * idx[a] - temp2.translation[a]) / 0.75 for a in range(3)]
                c = [int(np.floor(x)) for x in u]
                fr = [u[a] - c[a] for a in range(3)]
                s = 0.0
                for dz in (0, 1):
                    for dy in (0, 1):
                        for dx in (0, 1):
                            w = ((fr[0] if dx else 1 - fr[0]) * (fr[1] if dy else 1 - fr[1]) *
                                 (fr[2] if dz else 1 - fr[2]))
                            x, y, z = c[0] + dx - tlo[0], c[1] + dy - tlo[1], c[2] + dz - tlo[2]
                            assert 0 <= x < tdims[0] and 0 <= y < tdims[1] and 0 <= z < tdims[2]
                            s += w * td[z, y, x]
                exp[k, j, i] = s
    np.testing.assert_allclose(t2.reshape(d2.shape), exp, rtol=1e-6, atol=1e-4)
    # no temperature grid; no density grid
    path3 = vdbio.write(str(tmp_path / "v3.vdb"), [dens])
    f3, _, t3 = _volume(dump, path3, tmp_path)
    assert f3["has_temperature"] == ["0"] and t3.size == 0
    path4 = vdbio.write(str(tmp_path / "v4.vdb"), [temp])
    r = _run(dump, "volume", path4, tmp_path / "a.raw", tmp_path / "b.raw")
    assert r.returncode == 2 and "no grid named density" in r.stderr


def test_corrupted_files_fail_cleanly(tmp_path):
    """Bytes flipped, zeroed or cut anywhere in a valid file: the reader either
    reads a grid or throws its error (exit status 0 or 2), with no crash, no
    undefined behaviour and no hang.  Run without the address sanitizer and
    under a 4 GiB address-space limit, so that a corrupted size asking for a
    huge allocation ends in std::bad_alloc (caught: status 2)."""
    import resource
    exe = str(tmp_path / "vdb_dump_ub")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=undefined", "-fno-sanitize-recover=undefined",
           "-I", os.path.join(ROOT, "lightgraph_amd", "host"), os.path.join(ROOT, "tests", "host", "vdb_dump.cpp"),
           os.path.join(ROOT, "lightgraph_amd", "host", "vdb.cpp"), "-o", exe, "-lz", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    rng = np.random.default_rng(9)
    v, act = _sparse_cloud(rng, (10, 12, 14))
    g = vdbio.Grid.from_dense("density", v, act, (-5, 3, -9), voxel_size=0.5, translation=(1, 2, 3))
    g.tiles8[(64, 64, 64)] = (0.5, True)
    t = vdbio.Grid.from_dense("temperature", v * 100, act, (-5, 3, -9), voxel_size=0.5, translation=(1, 2, 3))
    good = open(vdbio.write(str(tmp_path / "good.vdb"), [g, t]), "rb").read()

    def limit():
        resource.setrlimit(resource.RLIMIT_AS, (4 << 30, 4 << 30))

    outcomes = {0: 0, 2: 0}
    for k in range(150):
        b = bytearray(good)
        kind = k % 3
        if kind == 0:  # flip a few bytes
            for _ in range(int(rng.integers(1, 4))):
                b[int(rng.integers(0, len(b)))] ^= int(rng.integers(1, 256))
        elif kind == 1:  # zero a run
            a, n = int(rng.integers(0, len(b))), int(rng.integers(1, 16))
            b[a:a + n] = bytes(len(b[a:a + n]))
        else:  # cut
            b = b[:int(rng.integers(0, len(b)))]
        f = tmp_path / f"bad{k}.vdb"
        f.write_bytes(bytes(b))
        for args in (["grid", f, "density", tmp_path / "o.raw"], ["volume", f, tmp_path / "d.raw", tmp_path / "t.raw"]):
            r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, errors="replace", timeout=60,
                               preexec_fn=limit,
                               env=dict(os.environ, UBSAN_OPTIONS="print_stacktrace=1"))
            assert r.returncode in (0, 2), (k, args[0], r.returncode, r.stderr[-2000:])
            assert "runtime error" not in r.stderr, (k, r.stderr[-2000:])
            outcomes[r.returncode] += 1
    assert outcomes[2] > 0

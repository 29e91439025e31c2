"""The command-line front end on the device: a procedural scene through
LightGraph iterations, splat and render equals, film for film and bit for bit,
the same stages called through the Python mirror of the C ABI; EXR output with
and without compression holds the same pixels; env lights load from EXR."""
import os
import subprocess

import numpy as np
import pytest

import exrio

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(__file__), "..", "lightgraph_amd", "lightgraph")
AOVS = [("", ("R", "G", "B", "A")), ("direct.", ("R", "G", "B")), ("scatter.", ("R", "G", "B")),
        ("emission.", ("R", "G", "B"))]


def _cli(*args):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _api_films(lga, ctx, size, seed, n, iters, radius, w, h, spp, lights, sh, p, cam_pos, focal, grids=None):
    if grids is None:
        spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, size, seed=seed, octaves=4, radius=0.375 * size)
        grid, _ = lga.Grid.generate(ctx, spec)
        temp = None
    else:
        grid, temp = grids
    s = lga.Solver(ctx, n * iters)
    s.solve_batch([(grid, n, i, temp) for i in range(iters)], sh, lights, p, radius=radius)
    res = s.batch_results()
    pts = np.concatenate([r["points"] for r in res])
    dif = np.concatenate([r["diffuse"] for r in res])
    sg = lga.ScatterGrid(ctx, grid, 4)
    rad, _ = lga.splat_radius(pts, 3 * iters)
    pb, rb = ctx.upload(pts), ctx.upload(dif)
    sg.splat(pb, rb, len(pts), rad)
    cam = lga.camera_setup(cam_translate=cam_pos, focal=focal, width=w, height=h)
    films = lga.Films(ctx, w, h)
    lga.render(ctx, grid, sg, temp, cam, sh, lights, p, spp, films, seed=1)
    out = films.read()
    films.free()
    for b in (pb, rb):
        b.free()
    sg.destroy()
    s.destroy()
    grid.destroy()
    if temp is not None:
        temp.destroy()
    return out


def test_cli_equals_api(lga, ctx, tmp_path):
    common = ["-scene", "sphere", "-sceneSize", "64", "-sceneSeed", "3", "-points", "400", "-iter", "3",
              "-radius", "6", "-resWidth", "64", "-resHeight", "48", "-samples", "2", "-camPos", "31.5,31.5,150",
              "-focal", "60", "-light", "dir:0:1,1,1:1:-45,30,0", "-light", "point:0:1,0.8,0.6:2000:20,70,30",
              "-scattering", "1", "-absorption", "0.1", "-shadowStep", "2"]
    a, b = str(tmp_path / "none.exr"), str(tmp_path / "zip.exr")
    out = _cli(*common, "-compression", "none", "-exrOut", a)
    for stage in ("LightGraphs Calculation:", "Rasterize LightGraphs to Grid:", "Render:"):
        assert stage in out
    _cli(*common, "-exrOut", b)  # default zip
    ca, ha = exrio.read(a)
    cb, hb = exrio.read(b)
    assert ha["compression"] == 0 and hb["compression"] == 3
    assert len(ha["channels"]) == 13
    for k in ca:
        np.testing.assert_array_equal(ca[k], cb[k])
    sh = lga.default_shader(absorption=0.1, scattering=1.0)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=2.0)
    lights = [lga.parse_light("dir:0:1,1,1:1:-45,30,0"), lga.parse_light("point:0:1,0.8,0.6:2000:20,70,30")]
    want = _api_films(lga, ctx, 64, 3, 400, 3, 6.0, 64, 48, 2, lights, sh, p, (31.5, 31.5, 150.0), 60.0)
    assert want[0][..., 3].max() > 0.3
    for f, (prefix, comps) in enumerate(AOVS):
        for c, comp in enumerate(comps):
            np.testing.assert_array_equal(ca[prefix + comp], want[f][..., c], err_msg=prefix + comp)


def test_cli_env_light_from_exr(lga, ctx, tmp_path):
    h, w = 16, 32
    y = np.linspace(0, 1, h, dtype=np.float32)[:, None, None]
    sky = np.broadcast_to(np.concatenate([0.3 + 0.5 * y, 0.4 + 0.4 * y, 0.9 - 0.3 * y], -1), (h, w, 3)).copy()
    sky[3, 20] = 400.0
    env_path = str(tmp_path / "sky.exr")
    exrio.write_rgb(env_path, sky, compression=3)
    back, _ = exrio.read(env_path)
    np.testing.assert_array_equal(back["R"], sky[..., 0])
    out = str(tmp_path / "env.exr")
    _cli("-scene", "sphere", "-sceneSize", "48", "-points", "300", "-iter", "2", "-radius", "5", "-resWidth", "40",
         "-resHeight", "30", "-camPos", "23.5,23.5,120", "-focal", "60", "-light", f"env:0:1,1,1:0.5:1,{env_path}",
         "-compression", "none", "-exrOut", out)
    c, _ = exrio.read(out)
    assert np.all(np.isfinite(c["direct.R"])) and c["direct.R"].max() > 0


def test_cli_geo_dump(lga, ctx, tmp_path):
    """-geoDump writes, per iteration, the reference's five comma-separated
    files (LightGraph.h:604-696); their values are the API's solve outputs and
    the edges' transmittance exp(-sigma D) normalised by max(1, column sum)."""
    n, iters = 300, 2
    args = ["-scene", "sphere", "-sceneSize", "48", "-sceneSeed", "5", "-points", str(n), "-iter", str(iters),
            "-radius", "6", "-resWidth", "16", "-resHeight", "12", "-camPos", "23.5,23.5,120",
            "-light", "dir:0:1,1,1:1:-45,30,0", "-scattering", "1,0.9,0.8", "-absorption", "0.1",
            "-exrOut", str(tmp_path / "o.exr"), "-geoDump"]
    r = subprocess.run([os.path.abspath(CLI), *args], capture_output=True, text=True, timeout=180, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    d = tmp_path / "geoDump"
    names = sorted(os.listdir(d))
    assert names == sorted(f"sphere_{k}_{i:04d}.txt" for k in ("pc", "edge", "marching", "light", "diffuse")
                           for i in range(iters))

    def load(kind, i, dtype=float):
        return np.loadtxt(d / f"sphere_{kind}_{i:04d}.txt", delimiter=",", dtype=dtype, ndmin=2)

    spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, 48, seed=5, octaves=4, radius=0.375 * 48)
    grid, _ = lga.Grid.generate(ctx, spec)
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params()
    lights = [lga.parse_light("dir:0:1,1,1:1:-45,30,0")]
    s = lga.Solver(ctx, n, dense=True)
    ext = np.array(sh.extinction)
    for i in range(iters):
        s.solve(grid, sh, lights, p, n=n, radius=6.0, solve=i)
        res = s.results()
        for kind, key in (("pc", "points"), ("light", "L"), ("diffuse", "diffuse")):
            # the files hold 6 significant digits (the stream default of boost::format)
            np.testing.assert_allclose(load(kind, i), res[key], rtol=1e-5, atol=1e-30, err_msg=kind)
        e = load("edge", i, int)
        np.testing.assert_array_equal(e, res["edges"])
        D = res["D"][:n, :n].astype(np.float64)
        T = np.exp(D[..., None] * ext[None, None, :])  # +INF -> 0
        col = np.maximum(T.sum(axis=0), 1.0)
        want = T[e[:, 0], e[:, 1]] / col[e[:, 1]]
        got = load("marching", i)
        assert got.shape == (len(e), 3) and np.all(got >= 0) and np.all(got <= 1)
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-9)
    s.destroy()
    grid.destroy()


def test_cli_dart_sampler(lga, ctx, tmp_path):
    """-sampler dart: each iteration's points come from the reference's dart
    thrower (lg_dart_throw; bit-exact against the oracle in test_gpu_dart.py)
    and its graph radius is 1.7 x the thrower's final minimum distance
    (LightGraph.h:169-170); the dumped points, edges, light and radiance equal
    the API's solve on those points."""
    iters = 2
    args = ["-scene", "sphere", "-sceneSize", "48", "-sceneSeed", "5", "-sampler", "dart", "-scatterCount", "100000",
            "-iter", str(iters), "-resWidth", "16", "-resHeight", "12", "-camPos", "23.5,23.5,120",
            "-light", "dir:0:1,1,1:1:-45,30,0", "-scattering", "1,0.9,0.8", "-absorption", "0.1",
            "-exrOut", str(tmp_path / "o.exr"), "-geoDump"]
    r = subprocess.run([os.path.abspath(CLI), *args], capture_output=True, text=True, timeout=180, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    d = tmp_path / "geoDump"

    def load(kind, i, dtype=float):
        return np.loadtxt(d / f"sphere_{kind}_{i:04d}.txt", delimiter=",", dtype=dtype, ndmin=2)

    spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, 48, seed=5, octaves=4, radius=0.375 * 48)
    grid, _ = lga.Grid.generate(ctx, spec)
    pts, md = lga.dart_throw(ctx, grid, range(iters), scatter_count=100000, target=750)
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params()
    lights = [lga.parse_light("dir:0:1,1,1:1:-45,30,0")]
    for i in range(iters):
        n = len(pts[i])
        assert 100 < n <= 751
        np.testing.assert_allclose(load("pc", i), pts[i], rtol=1e-5, atol=1e-30)
        s = lga.Solver(ctx, n)
        s.write_points(pts[i])
        radius = float(np.float32(md[i]) * np.float32(1.7))
        s.solve(grid, sh, lights, p, n=n, radius=radius, solve=i, external_points=True)
        res = s.results(want_D=False)
        np.testing.assert_array_equal(load("edge", i, int), res["edges"])
        np.testing.assert_allclose(load("light", i), res["L"], rtol=1e-5, atol=1e-30)
        np.testing.assert_allclose(load("diffuse", i), res["diffuse"], rtol=1e-5, atol=1e-30)
        s.destroy()
    grid.destroy()
    bad = subprocess.run([os.path.abspath(CLI), *args[:6], "-sampler", "poisson"], capture_output=True, text=True,
                         timeout=60, cwd=tmp_path)
    assert bad.returncode != 0 and "not density or dart" in bad.stdout


def test_cli_vdb_input_equals_api(lga, ctx, tmp_path):
    """-vdbIn: a density + temperature .vdb (tests/vdbio.py, off-origin index box,
    voxel 0.5, translated) renders, film for film and bit for bit, what the same
    dense arrays give through the Python mirror of the C ABI; the same volume
    stored with other encodings (raw, blosc, 16^3-node tiles) renders the same."""
    import vdbio
    rng = np.random.default_rng(12)
    z, y, x = np.mgrid[0:40, 0:36, 0:44].astype(np.float64)
    r = np.sqrt(((x - 21.5) / 20) ** 2 + ((y - 17.5) / 16) ** 2 + ((z - 19.5) / 18) ** 2)
    dens = np.where(r < 1, (1 - r) * (0.6 + 0.4 * rng.random(r.shape)), 0.0).astype(np.float32)
    dens[8:16, 8:16, 16:24] = 0.5  # index box [0, 8) x [0, 8) x [-16, -8): one 8^3 block
    temp = np.where(r < 0.8, 1000.0 * (0.8 - r), 0.0).astype(np.float32)
    imin, vs, tr = (-16, -8, -24), 0.5, (1.0, 2.0, 3.0)
    files = {}
    for tag, comp, tiles in (("zip", vdbio.ZIP | vdbio.ACTIVE_MASK, False), ("raw", 0, False),
                             ("blosc", vdbio.BLOSC | vdbio.ACTIVE_MASK, True)):
        if comp & vdbio.BLOSC and vdbio.BLOSC_LIB is None:
            continue
        g = vdbio.Grid.from_dense("density", dens, dens > 0, imin, tile_uniform=tiles, voxel_size=vs, translation=tr)
        t = vdbio.Grid.from_dense("temperature", temp, temp > 0, imin, voxel_size=vs, translation=tr)
        if tiles:
            assert len(g.tiles8) == 1
        files[tag] = vdbio.write(str(tmp_path / f"{tag}.vdb"), [g, t], compression=comp)
    ramp = "0.0,0.33,0.66,1.0/0,0,0:1.0,0.076,0.0:1.0,0.322,0.0:1.0,0.45,0.15/0.35/1.5"
    common = ["-points", "400", "-iter", "2", "-radius", "3", "-resWidth", "48", "-resHeight", "40", "-samples", "2",
              "-camPos", "1.5,3,60", "-focal", "50", "-light", "dir:0:1,1,1:1:-45,30,0", "-tempColorRamp", ramp,
              "-emissionScale", "0.002", "-shadowStep", "2", "-compression", "none"]
    films = {}
    for tag, path in files.items():
        out = _cli("-vdbIn", path, *common, "-exrOut", str(tmp_path / f"{tag}.exr"))
        assert f"Reading in: {path}" in out
        films[tag], _ = exrio.read(str(tmp_path / f"{tag}.exr"))
    for tag in films:
        for k in films["zip"]:
            np.testing.assert_array_equal(films[tag][k], films["zip"][k], err_msg=f"{tag} {k}")
    # the grids the reader gives: the active box of the density
    nz_z, nz_y, nz_x = np.nonzero(dens)
    lo = (nz_x.min(), nz_y.min(), nz_z.min())
    hi = (nz_x.max(), nz_y.max(), nz_z.max())
    box = (slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
    origin = tuple(tr[a] + vs * (imin[a] + lo[a]) for a in range(3))
    grid = lga.Grid.from_dense(ctx, np.ascontiguousarray(dens[box]), vs, origin)
    tgrid = lga.Grid.from_dense(ctx, np.ascontiguousarray(temp[box]), vs, origin)
    sh = lga.default_shader(ramp=ramp, emission_scale=0.002)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=2.0)
    lights = [lga.parse_light("dir:0:1,1,1:1:-45,30,0")]
    want = _api_films(lga, ctx, 0, 0, 400, 2, 3.0, 48, 40, 2, lights, sh, p, (1.5, 3.0, 60.0), 50.0,
                      grids=(grid, tgrid))
    assert want[0][..., 3].max() > 0.3 and want[3].max() > 0  # opaque somewhere; emission from the temperature
    for f, (prefix, comps) in enumerate(AOVS):
        for c, comp in enumerate(comps):
            np.testing.assert_array_equal(films["zip"][prefix + comp], want[f][..., c], err_msg=prefix + comp)

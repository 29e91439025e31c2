"""Device parity of the splat and the per-pixel march (kernel d) against the
oracle on seeded inputs.  Bar: every pixel of every AOV within 1e-4 relative
(floor 1e-6 of the image maximum); topology and splat radius exact."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(lga, ctx):
    spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, 64, seed=1, octaves=4, radius=24.0)
    grid, _ = lga.Grid.generate(ctx, spec)
    dense = grid.read_dense()
    yield grid, dense
    grid.destroy()


def _cam(lga, w, h, **kw):
    return lga.camera_setup(cam_translate=(31.5, 31.5, 150.0), focal=60.0, width=w, height=h, **kw)


def _films_parity(got, want, tol=1e-4):
    for k in range(4):
        assert rel_err(got[k], want[k]) < tol, f"AOV {k}: {rel_err(got[k], want[k])}"


def test_render_empty_volume(lga, ctx):
    g = lga.Grid.from_dense(ctx, np.zeros((16, 16, 16), np.float32))
    films = lga.Films(ctx, 20, 12)
    cam = lga.camera_setup(cam_translate=(8, 8, 50), width=20, height=12)
    lga.render(ctx, g, None, None, cam, lga.default_shader(), [lga.parse_light("dir:0:1,1,1:1:0,0,0")],
               lga.march_params(), 2, films)
    assert np.all(films.read() == 0.0)
    films.free()
    g.destroy()


def test_render_c1_parity(lga, ctx, oracle, scene):
    """The C1 image settings (256 x 256, 1 spp, primary 1, shadow 3, one directional light)."""
    grid, dense = scene
    sh = lga.default_shader(absorption=0.1, scattering=1.0)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=3.0)
    lights = [lga.parse_light("dir:0:1,1,1:1:-45,30,0")]
    cam = _cam(lga, 256, 256)
    films = lga.Films(ctx, 256, 256)
    lga.render(ctx, grid, None, None, cam, sh, lights, p, 1, films, seed=1)
    got = films.read()
    want = oracle.render(dense, cam, sh, lights, p, 1, seed=1, threads=16)
    assert got[0][..., 3].max() > 0.5
    _films_parity(got, want)
    films.free()


def test_render_all_terms_parity(lga, ctx, oracle, scene):
    """Scatter grid from a splat, temperature emission, point + directional
    lights, 3 AA samples, jittered shadow steps."""
    grid, dense = scene
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8), emission_scale=2.0,
                            ramp="0.0,0.33,0.66,1.0/0,0,0:1.0,0.076,0.0:1.0,0.322,0.0:1.0,0.45,0.15/0.35/1.5")
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=2.0)
    lights = [lga.parse_light("dir:0:1,0.89,0.79:0.7:-48.8,-311.86,0"),
              lga.point_light(0.0, (1.0, 0.8, 0.6), 5e3, (10.0, 90.0, 40.0))]
    temp = np.clip(dense * 2.0, 0, 1.5).astype(np.float32)
    tgrid = lga.Grid.from_dense(ctx, temp)
    # scatter grid from a solve's points
    solver = lga.Solver(ctx, 2048)
    solver.solve(grid, sh, lights, lga.march_params(shadow_step=0.5), n=2048, radius=8.0, seed=5, solve=0,
                 temperature=tgrid)
    r = solver.results(want_D=False)
    sg = lga.ScatterGrid(ctx, grid, 4)
    radius, vs = lga.splat_radius(r["points"], 3 * 1)
    pb, rb = ctx.upload(r["points"]), ctx.upload(r["diffuse"])
    sg.splat(pb, rb, len(r["points"]), radius)
    srgb, sact = sg.read()
    n = (sg.sg.nx, sg.sg.ny, sg.sg.nz)
    i0 = tuple(sg.sg.i0)
    assert srgb.max() > 0
    cam = _cam(lga, 96, 80, cam_rotate=(4.0, -3.0, 0.0))
    films = lga.Films(ctx, 96, 80)
    lga.render(ctx, grid, sg, tgrid, cam, sh, lights, p, 3, films, seed=7)
    got = films.read()
    want = oracle.render(dense, cam, sh, lights, p, 3, seed=7, temp_dense=temp, scatter=(srgb, n, i0, 4.0), threads=16)
    assert got[3][..., :3].max() > 0 and got[2][..., :3].max() > 0 and got[1][..., :3].max() > 0
    _films_parity(got, want)
    for b in (pb, rb):
        b.free()
    films.free()
    sg.destroy()
    solver.destroy()
    tgrid.destroy()


def test_scatter_topology_and_splat_parity(lga, ctx, oracle, scene):
    grid, dense = scene
    sg = lga.ScatterGrid(ctx, grid, 4)
    n = (sg.sg.nx, sg.sg.ny, sg.sg.nz)
    i0 = tuple(sg.sg.i0)
    _, act = sg.read()
    np.testing.assert_array_equal(act, oracle.scatter_topology(dense, 4, n, i0, 4.0))
    rng = np.random.default_rng(2)
    pts = oracle.sample_points(dense, 3, 0, 3000)
    rad = rng.uniform(0, 1, size=(3000, 3)).astype(np.float32)
    radius, vs = lga.splat_radius(pts, 6)
    assert vs == pytest.approx(oracle.point_voxel_size(pts), rel=0, abs=0)
    pb, rb = ctx.upload(pts), ctx.upload(rad)
    sg.splat(pb, rb, len(pts), radius)
    got, _ = sg.read()
    want = oracle.splat(pts, rad, radius, n, i0, 4.0, act)
    assert rel_err(got, want) < 1e-6
    pb.free()
    rb.free()
    sg.destroy()


def test_envmap_cdf_device_equals_oracle(lga, ctx, oracle):
    """The FP64 importance-sampling CDF built on the device (Light.h:161-176,
    Sampling.h:76-167) for a 1024 x 512 HDR map, sky gradient plus a 5000x sun
    disk, equals the oracle's (same sequential sums)."""
    h, w = 512, 1024
    y = np.linspace(0, 1, h, dtype=np.float32)[:, None, None]
    sky = np.broadcast_to(np.concatenate([0.25 + 0.5 * y, 0.35 + 0.45 * y, 0.8 - 0.2 * y], -1), (h, w, 3)).copy()
    yy, xx = np.mgrid[0:h, 0:w]
    sky[(xx - 0.7 * w) ** 2 + (yy - 0.22 * h) ** 2 <= (0.012 * w) ** 2 + 1] *= 5000.0
    env = lga.EnvMap(ctx, sky)
    cond, marg, rowint, total = env.read_cdf()
    env.destroy()
    want_cond, want_marg = oracle.env_cdf(sky)
    np.testing.assert_allclose(cond, want_cond, rtol=0, atol=1e-14)
    np.testing.assert_allclose(marg, want_marg, rtol=0, atol=1e-14)
    # the sun's rows hold most of the marginal mass
    sun_rows = slice(int(0.22 * h) - 13, int(0.22 * h) + 14)
    assert marg[sun_rows.stop] - marg[sun_rows.start] > 0.5


def test_ramp_knots_on_device(lga, ctx, oracle):
    """Emission only (no lights): the vertex light of points in a constant
    temperature grid is ramp(T) x emissionScale, at the knot temperatures of
    scenes/explo.sh:1 and between them."""
    ramp = "0.0,0.33,0.66,1.0/0,0,0:1.0,0.076,0.0:1.0,0.322,0.0:1.0,0.45,0.15/0.35/1.5"
    sh = lga.default_shader(absorption=0.1, scattering=1.0, emission_scale=3.0, ramp=ramp)
    p = lga.march_params()
    dens = lga.Grid.from_dense(ctx, np.full((16, 16, 16), 0.5, np.float32))
    pts = np.array([[8.0, 8.0, 8.0], [5.5, 9.25, 3.0]], np.float32)
    pb, ob = ctx.upload(pts), ctx.alloc(len(pts) * 12)
    for T in (0.0, 0.35, 0.35 + 0.33 * 1.15, 0.35 + 0.495 * 1.15, 0.35 + 0.66 * 1.15, 1.5, 3.0):
        tg = lga.Grid.from_dense(ctx, np.full((16, 16, 16), T, np.float32))
        lga.vertex_light(ctx, dens, tg, sh, [], p, 4564446, 0, pb, len(pts), ob)
        got = ob.read(np.float32, (len(pts), 3))
        want = oracle.ramp_color(sh, float(np.float32(T))) * 3.0
        for row in got:
            np.testing.assert_allclose(row, want, rtol=2e-7, atol=1e-7)
        tg.destroy()
    pb.free()
    ob.free()
    dens.destroy()


@pytest.mark.parametrize("voxel", [(37, 20, 9), (8, 8, 8)])
def test_scatter_topology_single_voxel_on_device(lga, ctx, voxel):
    """One active voxel v: the coarse voxel floor((v + 2) / 4) of the factor-4
    grid and its 26 neighbours, nothing else."""
    a = np.zeros((64, 64, 64), np.float32)
    a[voxel[2], voxel[1], voxel[0]] = 0.7
    g = lga.Grid.from_dense(ctx, a)
    sg = lga.ScatterGrid(ctx, g, 4)
    _, act = sg.read()
    i0 = tuple(sg.sg.i0)
    want = np.zeros_like(act)
    c = [(voxel[k] + 2) // 4 - i0[k] for k in range(3)]
    want[max(c[2] - 1, 0):c[2] + 2, max(c[1] - 1, 0):c[1] + 2, max(c[0] - 1, 0):c[0] + 2] = 1
    np.testing.assert_array_equal(act, want)
    sg.destroy()
    g.destroy()

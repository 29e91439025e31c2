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

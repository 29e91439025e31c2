"""Device marches over a sparse volume against the oracle: three small blobs in
a 160^3 grid, far apart, so that rays cross wide empty regions where the
brick walk jumps over empty boxes (the spans, step origins and jitter draws
must stay those of the cell-by-cell walk) and the vertex light's batches jump
over the empty box around their last sample (no sample in an existing brick
may be skipped, the one-brick specks included).  Bars as in the dense tests:
edge optical depths within float32 rounding of the oracle's FP64 value, vertex
light within 1e-5, every render AOV within 1e-4 relative."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

CENTRES = ((28.0, 30.0, 34.0), (120.0, 46.0, 118.0), (70.0, 128.0, 60.0))


def _sparse_volume(n=160, radius=13.0, seed=11):
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float32)
    vol = np.zeros((n, n, n), np.float32)
    for c in CENTRES:
        r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
        vol = np.maximum(vol, np.clip(1.0 - r / radius, 0.0, 1.0))
    noise = rng.uniform(0.5, 1.0, size=vol.shape).astype(np.float32)
    vol = np.where(vol > 0, vol * noise, 0.0).astype(np.float32)
    # a few isolated single-voxel specks make one-brick islands in the gaps
    for _ in range(6):
        p = rng.integers(8, n - 8, size=3)
        vol[p[2], p[1], p[0]] = 0.75
    return vol


@pytest.fixture(scope="module")
def sparse(lga, ctx):
    dense = _sparse_volume()
    grid = lga.Grid.from_dense(ctx, dense)
    yield grid, dense
    grid.destroy()


def test_sparse_edge_march_parity(lga, ctx, oracle, sparse):
    grid, dense = sparse
    pts = oracle.sample_points(dense, 77, 0, 600)
    edges = oracle.build_edges(pts, 140.0, 0)
    sh = lga.default_shader(absorption=0.1, scattering=1.0)
    p = lga.march_params(cutoff=0.005, shadow_step=0.5)
    pb, ebuf = ctx.upload(pts), ctx.upload(edges)
    wb = ctx.alloc(len(edges) * 4)
    lga.march_edges(ctx, grid, sh, p, pb, ebuf, len(edges), w_buf=wb)
    w = wb.read(np.float32, (len(edges),))
    T, tau = oracle.march_edges(dense, sh, p, pts, edges)
    assert np.array_equal(np.isinf(w), np.isinf(tau))
    fin = np.isfinite(tau)
    assert fin.sum() > 0.5 * len(tau)
    assert rel_err(w[fin], tau[fin], floor=1e-30) < 2e-7
    for b in (pb, ebuf, wb):
        b.free()


@pytest.mark.parametrize("step", [0.5, 2.0])
def test_sparse_vertex_light_parity(lga, ctx, oracle, sparse, step):
    """Points of all three blobs lit by point lights placed beyond the other
    blobs (each shadow ray leaves its blob, crosses a gap and enters another
    blob) and by a directional light: the batch jumps over empty boxes and the
    samples in existing bricks after them equal the oracle's span walk."""
    grid, dense = sparse
    pts = oracle.sample_points(dense, 91, 0, 900)
    sh = lga.default_shader(absorption=0.1, scattering=1.0)
    p = lga.march_params(cutoff=0.005, shadow_step=step)
    c = np.array(CENTRES)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 3e4, tuple(c[0] + 2.0 * (c[1] - c[0]))),
              lga.point_light(0.0, (0.8, 0.9, 1.0), 3e4, tuple(c[1] + 2.0 * (c[2] - c[1]))),
              lga.point_light(0.0, (0.9, 1.0, 0.9), 3e4, tuple(c[2] + 2.0 * (c[0] - c[2]))),
              lga.parse_light("dir:0:1,1,1:1:-45,30,0")]
    pb = ctx.upload(pts)
    lb = ctx.alloc(len(pts) * 12)
    lga.vertex_light(ctx, grid, None, sh, lights, p, 4564446, 0, pb, len(pts), lb)
    got = lb.read(np.float32, (len(pts), 3))
    want = oracle.vertex_light(dense, sh, p, lights, pts)
    assert np.all(want > 0)
    assert rel_err(got, want) < 1e-5
    pb.free()
    lb.free()


@pytest.mark.parametrize("rotate", [(0.0, 0.0, 0.0), (6.0, -8.0, 3.0)])
def test_sparse_render_parity(lga, ctx, oracle, sparse, rotate):
    grid, dense = sparse
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=2.0)
    lights = [lga.parse_light("dir:0:1,0.89,0.79:0.7:-48.8,-31.86,0"),
              lga.point_light(0.0, (1.0, 0.8, 0.6), 5e3, (80.0, 80.0, 80.0))]
    cam = lga.camera_setup(cam_translate=(80.0, 80.0, 260.0), cam_rotate=rotate, focal=50.0, width=160, height=128)
    films = lga.Films(ctx, 160, 128)
    lga.render(ctx, grid, None, None, cam, sh, lights, p, 2, films, seed=3)
    got = films.read()
    want = oracle.render(dense, cam, sh, lights, p, 2, seed=3, threads=16)
    assert got[0][..., 3].max() > 0.2
    for k in range(4):
        assert rel_err(got[k], want[k]) < 1e-4, f"AOV {k}: {rel_err(got[k], want[k])}"
    films.free()

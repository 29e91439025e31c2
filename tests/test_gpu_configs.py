"""Parity at the configured sizes.

* Production solve: 512^3 cloud of 20 seeded ellipsoid blobs x 6-octave fBm
  (~30% active voxels), N = 8192 points, edges within 12 voxels (cap 12), a
  point and a directional light, edge march at 0.5 voxel: points and edges
  bit-exact against the oracle, vertex light, optical depths and diffused
  radiance within 1e-5 relative of the oracle's FP64 max-product solve
  (LightGraph.h:158-196).
* The dense 8192 x 8192 Floyd-Warshall closure of that graph (the whole matrix,
  not its components) against Dijkstra in FP64 (1e-5 relative), with integer
  weights bit-exact, predecessors by the canonical rule.
* The 1920 x 1080 x 4 render of that volume with the scatter grid of one
  solve: the whole image against the oracle's render, every AOV of every pixel
  within 1e-4 relative (Renderer.h:506-693).
* The batch: 64 volumes of side 128..384 with N in [512, 6144]: the batch
  equals the plain per-volume solves bit for bit in the ascending and the
  nested-dissection pivot orders, and one graph equals the oracle's solve.
* Serving: a 768^3, N = 16000, 1920 x 1080 x 4 job with directional, point and
  environment lights gives the same films alone and mixed with other jobs."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import dijkstra

from conftest import rel_err

pytestmark = pytest.mark.gpu

N_C2 = 8192
RADIUS = 12.0


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c2(lga, ctx):
    _say("configs: generating the 512^3 blob volume")
    spec = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, 512, seed=2, octaves=6, blobs=20, axis_min=30.0, axis_max=120.0,
                          active_fraction=0.30)
    grid, _ = lga.Grid.generate(ctx, spec)
    dense = grid.read_dense()
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 4.0e5, (256.0, 700.0, 256.0)),
              lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")]
    yield dict(grid=grid, dense=dense, sh=sh, p=p, lights=lights)
    grid.destroy()


@pytest.fixture(scope="module")
def c2_solve(lga, ctx, c2):
    _say("configs: C2 solve on the device (N=8192, dense D)")
    s = lga.Solver(ctx, N_C2, dense=True)
    s.solve(c2["grid"], c2["sh"], c2["lights"], c2["p"], n=N_C2, radius=RADIUS, seed=854123, solve=0)
    r = s.results()
    s.destroy()
    return r


def test_c2_solve_against_oracle(lga, ctx, oracle, c2, c2_solve):
    r = c2_solve
    dense, sh = c2["dense"], c2["sh"]
    np.testing.assert_array_equal(r["points"], oracle.sample_points(dense, 854123, 0, N_C2))
    _say("configs: oracle solve at N=8192")
    ref = oracle.solve_points(dense, sh, c2["p"], c2["lights"], r["points"], radius=RADIUS, cap=12, want_P=True)
    np.testing.assert_array_equal(r["edges"], ref["edges"])
    assert len(r["edges"]) > N_C2 // 2 and r["n_components"] > 100
    assert rel_err(r["L"], ref["L"]) < 1e-5
    assert rel_err(r["diffuse"], ref["diffuse"]) < 1e-5
    P0 = ref["P"][..., 0]
    del ref["P"]
    sig0 = -sh.extinction[0]
    D = r["D"]
    reach = P0 > 0.0
    # reachable in the reference <=> finite optical depth on the device, except
    # where exp(-sigma tau) underflows FP64 in the reference's transmittance
    assert np.all(np.isfinite(D[reach]))
    lost = np.isfinite(D) & ~reach
    assert np.all(D[lost] * sig0 > 700.0)
    fin = P0 > 1e-290
    tau_ref = -np.log(P0[fin]) / sig0
    assert rel_err(D[fin], tau_ref, floor=1e-30) < 1e-5


def _canonical_pred(D, edges, w):
    """pred[i][j] = smallest neighbour k of j minimising D[i][k] + w(k, j) in
    float32, -1 where j is unreachable from i, i on the diagonal (column by
    column, vectorised over i)."""
    n = len(D)
    nbr = [[] for _ in range(n)]
    for (a, b), x in zip(edges, w):
        if np.isfinite(x):
            nbr[a].append((b, np.float32(x)))
            nbr[b].append((a, np.float32(x)))
    out = np.full((n, n), -1, np.int32)
    for j in range(n):
        if not nbr[j]:
            continue
        lst = sorted(nbr[j])
        ks = np.array([k for k, _ in lst], np.int64)
        wv = np.array([x for _, x in lst], np.float32)
        vals = D[:, ks] + wv[None, :]
        best = np.argmin(vals, axis=1)
        reach = np.isfinite(D[:, j])
        out[reach, j] = ks[best[reach]]
    np.fill_diagonal(out, np.arange(n, dtype=np.int32))
    return out


def _dense_matrix(lga, n, e, w):
    ld = lga.fw_ld(n)
    W = np.full((ld, ld), np.inf, np.float32)
    np.fill_diagonal(W, 0.0)
    W[e[:, 0], e[:, 1]] = w
    W[e[:, 1], e[:, 0]] = w
    return W, ld


def test_c2_dense_fw_against_dijkstra(lga, ctx, c2_solve):
    """The dense 8192^2 closure (every tile of the whole matrix, pivots in
    ascending order) of the production graph."""
    r = c2_solve
    n, e, w = N_C2, r["edges"], r["weights"]
    fin = np.isfinite(w)
    for integer in (False, True):
        wt = np.maximum(np.round(w * 4.0), 1.0).astype(np.float32) if integer else w
        wt = np.where(fin, wt, np.inf).astype(np.float32)
        W, ld = _dense_matrix(lga, n, e, wt)
        buf = ctx.upload(W)
        del W
        pred = ctx.alloc(ld * ld * 4) if integer else None
        eb, wb = ctx.upload(e), ctx.upload(wt)
        _say(f"configs: dense 8192^2 closure (integer weights: {integer})")
        lga.fw_solve(ctx, n, buf, ld, pred, eb, wb, len(e))
        got = buf.read(np.float32, (ld, ld))[:n, :n]
        A = csr_matrix((wt[fin].astype(np.float64), (e[fin, 0], e[fin, 1])), shape=(n, n))
        ref = dijkstra(A, directed=False)
        assert np.array_equal(np.isinf(got), np.isinf(ref))
        if integer:
            np.testing.assert_array_equal(got.astype(np.float64), ref)
            _say("configs: canonical predecessors")
            np.testing.assert_array_equal(pred.read(np.int32, (ld, ld))[:n, :n], _canonical_pred(got, e, wt))
            pred.free()
        else:
            f = np.isfinite(ref)
            assert rel_err(got[f], ref[f], floor=1e-30) < 1e-5
            # the per-component closure of the solve agrees with the whole-matrix one
            assert np.array_equal(np.isinf(r["D"]), np.isinf(got))
            assert rel_err(r["D"][f], got[f], floor=1e-30) < 1e-5
        for b in (buf, eb, wb):
            b.free()


def test_c2_render_against_oracle(lga, ctx, oracle, c2, c2_solve):
    """1920 x 1080, 4 AA samples, primary step 1, shadow step 3, scatter grid
    (factor 4) splatted from the solve's radiance with 3 points per kernel: the
    whole image, every AOV of every pixel within 1e-4 relative (floor 1e-6 of
    the image maximum) of the oracle's FP64 render."""
    r = c2_solve
    grid, dense, sh, lights = c2["grid"], c2["dense"], c2["sh"], c2["lights"]
    sg = lga.ScatterGrid(ctx, grid, 4)
    radius, _ = lga.splat_radius(r["points"], 3)
    pb, rb = ctx.upload(r["points"]), ctx.upload(r["diffuse"])
    sg.splat(pb, rb, len(r["points"]), radius)
    srgb, _ = sg.read()
    n = (sg.sg.nx, sg.sg.ny, sg.sg.nz)
    i0 = tuple(sg.sg.i0)
    w, h, spp = 1920, 1080, 4
    cam = lga.camera_setup(cam_translate=(256.0, 256.0, 1400.0), focal=50.0, width=w, height=h)
    rp = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=3.0)
    films = lga.Films(ctx, w, h)
    _say("configs: 1920x1080x4 render on the device")
    lga.render(ctx, grid, sg, None, cam, sh, lights, rp, spp, films, seed=1)
    got = films.read()
    films.free()
    assert got[0][..., 3].mean() > 0.3 and got[2][..., :3].max() > 0
    want = np.zeros_like(got)
    for j0 in range(0, h, 120):
        j1 = min(h, j0 + 120)
        _say(f"configs: oracle rows {j0}..{j1}")
        part = oracle.render(dense, cam, sh, lights, rp, spp, seed=1, scatter=(srgb, n, i0, 4.0), threads=16,
                             rows=(j0, j1))
        for k in range(4):
            want[k][j0:j1] = part[k]
    assert want[0][..., 3].max() > 0.3
    for k in range(4):
        e = rel_err(got[k], want[k])
        _say(f"configs: AOV {k} max rel err {e:.3g}")
        assert e < 1e-4, f"AOV {k}: {e}"
    for b in (pb, rb):
        b.free()
    sg.destroy()


def test_c3_batch_equals_plain_and_oracle(lga, ctx, oracle):
    """64 volumes, side U{128..384} step 8, N U{512..6144}, seeds 3000+k."""
    rng = np.random.default_rng(3000)
    count = 64
    sides = rng.integers(0, (384 - 128) // 8 + 1, count) * 8 + 128
    ns = rng.integers(512, 6145, count)
    _say("configs: generating 64 volumes")
    grids = []
    for k in range(count):
        spec = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, int(sides[k]), seed=3000 + k, octaves=6, blobs=20,
                              axis_min=30.0, axis_max=120.0, active_fraction=0.30)
        grids.append(lga.Grid.generate(ctx, spec)[0])
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 4.0e5, (256.0, 700.0, 256.0)),
              lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")]
    items = [(g, int(n), 5000 + k) for k, (g, n) in enumerate(zip(grids, ns))]
    plain = lga.Solver(ctx, int(ns.max()))
    batch = lga.Solver(ctx, int(ns.sum()))
    for order in ("reference", "nd"):
        _say(f"configs: batch of 64 vs plain, {order} order")
        batch.solve_batch(items, sh, lights, p, radius=RADIUS, order=order)
        got = batch.batch_results()
        for k, (g, n, sv) in enumerate(items):
            plain.solve(g, sh, lights, p, n=n, radius=RADIUS, solve=sv, order=order)
            want = plain.results(want_D=False)
            for key in ("points", "edges", "weights", "L", "diffuse"):
                np.testing.assert_array_equal(got[k][key], want[key], err_msg=f"volume {k} {key} ({order})")
        if order == "reference":
            ref_batch = got
    # one graph against the oracle: the median-size graph
    k = int(np.argsort(ns)[count // 2])
    _say(f"configs: oracle solve of volume {k} (side {sides[k]}, N={ns[k]})")
    dense = grids[k].read_dense()
    r = ref_batch[k]
    np.testing.assert_array_equal(r["points"], oracle.sample_points(dense, 854123, 5000 + k, int(ns[k])))
    ref = oracle.solve_points(dense, sh, p, lights, r["points"], radius=RADIUS, cap=12, solve=5000 + k)
    np.testing.assert_array_equal(r["edges"], ref["edges"])
    assert rel_err(r["L"], ref["L"]) < 1e-5
    assert rel_err(r["diffuse"], ref["diffuse"]) < 1e-5
    plain.destroy()
    batch.destroy()
    for g in grids:
        g.destroy()


def _sky(h, w):
    y = np.linspace(0, 1, h, dtype=np.float32)[:, None, None]
    sky = np.broadcast_to(np.concatenate([0.2 + 0.6 * y, 0.3 + 0.5 * y, 0.9 - 0.2 * y], -1), (h, w, 3)).copy()
    sky[20:23, 150:153] = 5000.0
    return sky


def test_c5_large_job_alone_equals_mixed(lga, ctx):
    """A job at the largest serving size (768^3 volume, N = 16000, 1920 x 1080
    at 4 spp, directional + point + environment light) alone, then arriving with
    three other jobs that are batched and overlapped with it: the same films."""
    env = lga.EnvMap(ctx, _sky(128, 256))
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    mp = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=3.0)
    lights = [lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0"),
              lga.point_light(0, (1, 0.9, 0.8), 4e5, (384, 1100, 384)),
              lga.env_light(env, 0.0, (1, 1, 1), 1.0, samples=1)]

    def mk(side, seed, n, w, h, spp, base):
        v = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, side, seed=seed, octaves=6, blobs=20,
                           axis_min=side * 30 / 512, axis_max=side * 120 / 512, active_fraction=0.30)
        cam = lga.camera_setup(cam_translate=(side / 2, side / 2, 2.8 * side), focal=50.0, width=w, height=h)
        return lga.job(v, n, cam, sh, lights, mp, radius=12.0, samples=spp, solve_base=base, order="nd")

    big = mk(768, 7008, 16000, 1920, 1080, 4, 0)
    others = [mk(256, 7001, 4000, 512, 512, 1, 1), mk(384, 7002, 8000, 1280, 720, 2, 2),
              mk(128, 7003, 1000, 512, 512, 3, 3)]
    srv = lga.Server(96 << 30, workers=3, max_batch=8)
    _say("configs: 768^3 N=16000 1080p x4 job alone")
    alone = srv.wait(srv.submit(big))
    assert alone["status"] == 0 and alone["batch"] == 1
    _say("configs: the same job mixed with three others")
    ids = [srv.submit(j) for j in [others[0], big, others[1], others[2]]]
    mixed = [srv.wait(i) for i in ids]
    assert all(m["status"] == 0 for m in mixed)
    np.testing.assert_array_equal(mixed[1]["films"], alone["films"])
    assert alone["films"][0][..., 3].max() > 0.3
    st = srv.stats()
    assert st["failed"] == 0 and st["peak_reserved_bytes"] <= st["budget_bytes"]
    srv.close()
    env.destroy()

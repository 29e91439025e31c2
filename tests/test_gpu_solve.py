"""Device parity of the solve stages (kernels a, b, c, diffusion) against the
CPU oracle on seeded inputs.  Bars: bit-exact for integer/index work (sampler,
edges, integer-weighted Floyd-Warshall and predecessors); optical depth within
float32 rounding; Floyd-Warshall within 1e-5 relative on float weights; light
and diffusion within 1e-5 relative."""
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fw_golden.npz")


@pytest.fixture(scope="module")
def c1(lga, ctx):
    """C1 volume: 64^3 sphere r=24 x fBm4, seed 1 (generated on the device)."""
    spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, 64, seed=1, octaves=4, radius=24.0)
    grid, _ = lga.Grid.generate(ctx, spec)
    dense = grid.read_dense()
    yield grid, dense, spec
    grid.destroy()


def _c1_shader(lga):
    return lga.default_shader(absorption=0.1, scattering=1.0)


def test_generator_device_equals_host(lga, ctx, c1):
    grid, dense, spec = c1
    np.testing.assert_array_equal(dense, lga.generate_host(spec, threads=16))
    spec2 = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, 96, seed=7, octaves=6, blobs=12, axis_min=6, axis_max=24,
                           active_fraction=0.3)
    g2, _ = lga.Grid.generate(ctx, spec2)
    np.testing.assert_array_equal(g2.read_dense(), lga.generate_host(spec2, threads=16))
    g2.destroy()


def test_grid_roundtrip_and_topology(lga, ctx, oracle):
    rng = np.random.default_rng(4)
    a = np.zeros((29, 37, 45), np.float32)
    a[3:20, 5:30, 10:40] = rng.uniform(0, 1, size=(17, 25, 30)).astype(np.float32)
    a[a < 0.3] = 0.0
    for dtype in (lga.LG_F32,):
        g = lga.Grid.from_dense(ctx, a, dtype=dtype)
        np.testing.assert_array_equal(g.read_dense(), a)
        info = g.info()
        nz = np.argwhere(a != 0)
        bz, by, bx = (nz // 8).min(0), (nz // 8).max(0), None
        assert list(info.bbox_min) == [8 * (nz[:, 2] // 8).min(), 8 * (nz[:, 1] // 8).min(), 8 * (nz[:, 0] // 8).min()]
        assert list(info.bbox_max) == [8 * ((nz[:, 2] // 8).max() + 1), 8 * ((nz[:, 1] // 8).max() + 1),
                                       8 * ((nz[:, 0] // 8).max() + 1)]
        assert info.active_voxels == int(np.count_nonzero(a))
        g.destroy()


@pytest.mark.parametrize("seed,solve,n", [(854123, 0, 1024), (854123, 7, 333), (42, 3, 4096)])
def test_sampler_bit_exact(lga, ctx, oracle, c1, seed, solve, n):
    grid, dense, _ = c1
    buf = ctx.alloc(n * 12)
    lga.sample_points(ctx, grid, seed, solve, n, buf)
    got = buf.read(np.float32, (n, 3))
    buf.free()
    np.testing.assert_array_equal(got, oracle.sample_points(dense, seed, solve, n))


# radius 5-12 on the C1 sphere gives >= 4 candidates per vertex (rows decided in
# parallel); radius 2.5-3 gives fewer (the single-wave sweep)
@pytest.mark.parametrize("radius,cap", [(8.0, 12), (8.0, 0), (5.0, 3), (12.0, 12), (3.0, 3), (2.5, 1)])
def test_edges_bit_exact(lga, ctx, oracle, c1, radius, cap):
    grid, dense, _ = c1
    n = 1024
    pts = oracle.sample_points(dense, 854123, 1, n)
    pb = ctx.upload(pts)
    cap_e = n * n // 2
    eb = ctx.alloc(cap_e * 8)
    ne = lga.build_edges(ctx, pb, n, radius, cap, eb, cap_e)
    got = eb.read(np.int32, (ne, 2))
    np.testing.assert_array_equal(got, oracle.build_edges(pts, radius, cap))
    pb.free()
    eb.free()


def test_edges_overflow_reports(lga, ctx, oracle):
    rng = np.random.default_rng(3)
    pts = rng.uniform(0.0, 0.1, size=(20, 3)).astype(np.float32)
    pb = ctx.upload(pts)
    eb = ctx.alloc(10 * 8)
    with pytest.raises(lga.LGError):
        lga.build_edges(ctx, pb, 20, 10.0, 12, eb, 10)
    pb.free()
    eb.free()


def test_edge_march_parity(lga, ctx, oracle, c1):
    grid, dense, _ = c1
    n = 1024
    pts = oracle.sample_points(dense, 854123, 2, n)
    edges = oracle.build_edges(pts, 8.0, 12)
    sh = _c1_shader(lga)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    pb, ebuf = ctx.upload(pts), ctx.upload(edges)
    wb = ctx.alloc(len(edges) * 4)
    lga.march_edges(ctx, grid, sh, p, pb, ebuf, len(edges), w_buf=wb)
    w = wb.read(np.float32, (len(edges),))
    T, tau = oracle.march_edges(dense, sh, p, pts, edges)
    assert np.array_equal(np.isinf(w), np.isinf(tau))
    fin = np.isfinite(tau)
    assert fin.sum() > 0.5 * len(tau)
    # float32 rounding of the FP64 optical depth
    assert rel_err(w[fin], tau[fin], floor=1e-30) < 2e-7
    for b in (pb, ebuf, wb):
        b.free()


# an arbitrary step and a power of two
@pytest.mark.parametrize("step", [0.7, 0.5])
def test_edge_march_dense_dropouts(lga, ctx, oracle, c1, step):
    grid, dense, _ = c1
    pts = oracle.sample_points(dense, 5, 0, 512)
    edges = oracle.build_edges(pts, 20.0, 0)
    sh = lga.default_shader(absorption=0.5, scattering=2.0, scatter_density_scale=1.0)
    p = lga.march_params(cutoff=0.01, shadow_step=step)
    pb, ebuf = ctx.upload(pts), ctx.upload(edges)
    wb = ctx.alloc(len(edges) * 4)
    lga.march_edges(ctx, grid, sh, p, pb, ebuf, len(edges), w_buf=wb)
    w = wb.read(np.float32, (len(edges),))
    T, tau = oracle.march_edges(dense, sh, p, pts, edges)
    assert 0 < np.isinf(tau).sum() < len(tau)
    assert np.array_equal(np.isinf(w), np.isinf(tau))
    fin = np.isfinite(tau)
    assert rel_err(w[fin], tau[fin], floor=1e-30) < 2e-7
    for b in (pb, ebuf, wb):
        b.free()


def _lights(lga):
    return [lga.parse_light("dir:0:1,1,1:1:-45,30,0"), lga.point_light(0, (1.0, 0.8, 0.6), 2000.0, (20, 70, 30))]


@pytest.mark.parametrize("step", [0.5, 3.0, 0.3])
def test_vertex_light_parity(lga, ctx, oracle, c1, step):
    """Steps 0.5 and 3 take the wave-per-ray kernel (exact sample positions),
    0.3 the one-lane-per-vertex sequential march."""
    grid, dense, _ = c1
    n = 1024
    pts = oracle.sample_points(dense, 854123, 0, n)
    sh = _c1_shader(lga)
    p = lga.march_params(cutoff=0.005, shadow_step=step)
    lights = _lights(lga)
    pb = ctx.upload(pts)
    lb = ctx.alloc(n * 12)
    lga.vertex_light(ctx, grid, None, sh, lights, p, 4564446, 0, pb, n, lb)
    got = lb.read(np.float32, (n, 3))
    want = oracle.vertex_light(dense, sh, p, lights, pts)
    assert rel_err(got, want) < 1e-5
    pb.free()
    lb.free()


def _sky(h=32, w=64):
    """Sky gradient with one bright sun texel block."""
    y = np.linspace(0.0, 1.0, h, dtype=np.float32)[:, None, None]
    rgb = np.broadcast_to(np.concatenate([0.2 + 0.6 * y, 0.3 + 0.5 * y, 0.9 - 0.2 * y], axis=-1), (h, w, 3)).copy()
    rgb[6:8, 40:42] = 500.0
    return rgb


@pytest.mark.parametrize("step", [0.5, 0.3])
def test_vertex_light_env_and_multisample(lga, ctx, oracle, c1, step):
    """Env light (2 draws per sample, 10x samples), a multi-sample point light
    and a directional light in one list: the per-ray kernel's stream offsets and
    the per-vertex sum order against the oracle."""
    grid, dense, _ = c1
    n = 512
    pts = oracle.sample_points(dense, 854123, 2, n)
    sh = _c1_shader(lga)
    p = lga.march_params(cutoff=0.005, shadow_step=step)
    rgb = _sky()
    env = lga.EnvMap(ctx, rgb)
    lights = [lga.point_light(0, (1.0, 0.8, 0.6), 2000.0, (20, 70, 30)),
              lga.env_light(env, y_rot=30.0, color=(1.0, 0.9, 0.8), intensity=0.5, samples=2),
              lga.parse_light("dir:0:1,1,1:1:-45,30,0")]
    lights[0].samples = 2
    pb = ctx.upload(pts)
    lb = ctx.alloc(n * 12)
    lga.vertex_light(ctx, grid, None, sh, lights, p, 4564446, 7, pb, n, lb)
    got = lb.read(np.float32, (n, 3))
    want = oracle.vertex_light(dense, sh, p, lights, pts, solve=7, envs=[None, rgb, None])
    assert rel_err(got, want) < 1e-5
    pb.free()
    lb.free()
    env.destroy()


def _fw_device(lga, ctx, W):
    n = len(W)
    ld = lga.fw_ld(n)
    Dm = np.full((ld, ld), np.inf, np.float32)
    np.fill_diagonal(Dm, 0.0)
    Dm[:n, :n] = W
    db = ctx.upload(Dm)
    lga.fw_solve(ctx, n, db, ld)
    out = db.read(np.float32, (ld, ld))
    db.free()
    return out[:n, :n]


def test_fw_golden_fixtures(lga, ctx, oracle):
    gold = np.load(GOLDEN)
    for name in ("int16", "int64", "int200", "sparse130", "flt64"):
        W, want = gold[name + "_W"], gold[name + "_D"]
        got = _fw_device(lga, ctx, W.astype(np.float32))
        if name.startswith("int") or name.startswith("sparse"):
            np.testing.assert_array_equal(got.astype(np.float64), want)
        else:
            assert np.array_equal(np.isinf(got), np.isinf(want))
            assert rel_err(got[np.isfinite(want)], want[np.isfinite(want)]) < 1e-5


def _random_graph(rng, n, deg, integer):
    W = np.full((n, n), np.inf, np.float32)
    np.fill_diagonal(W, 0)
    m = int(n * deg / 2)
    i = rng.integers(0, n, m)
    j = rng.integers(0, n, m)
    w = rng.integers(1, 1000, m).astype(np.float32) if integer else rng.uniform(0.01, 2.0, m).astype(np.float32)
    for a, b, x in zip(i, j, w):
        if a != b:
            W[a, b] = W[b, a] = min(W[a, b], x)
    return W


@pytest.mark.parametrize("n", [1, 2, 100, 128, 129, 257, 600, 1024])
def test_fw_integer_bit_exact(lga, ctx, oracle, n):
    rng = np.random.default_rng(n)
    W = _random_graph(rng, n, 3.0, True)
    got = _fw_device(lga, ctx, W)
    want = oracle.floyd_minplus(W.astype(np.float64))
    np.testing.assert_array_equal(got.astype(np.float64), want)
    np.testing.assert_array_equal(got, got.T)


@pytest.mark.parametrize("n,deg", [(300, 2.0), (1000, 6.0), (1500, 1.0)])
def test_fw_float_within_1e5(lga, ctx, oracle, n, deg):
    rng = np.random.default_rng(7 + n)
    W = _random_graph(rng, n, deg, False)
    got = _fw_device(lga, ctx, W)
    want = oracle.floyd_minplus(W.astype(np.float64))
    assert np.array_equal(np.isinf(got), np.isinf(want))
    fin = np.isfinite(want)
    assert rel_err(got[fin], want[fin], floor=1e-30) < 1e-5


def test_fw_predecessors_bit_exact(lga, ctx, oracle):
    rng = np.random.default_rng(11)
    n = 400
    W = _random_graph(rng, n, 3.0, True)
    iu = np.argwhere(np.triu(np.isfinite(W), 1))
    edges = iu.astype(np.int32)
    w = W[iu[:, 0], iu[:, 1]].astype(np.float32)
    ld = lga.fw_ld(n)
    Dm = np.full((ld, ld), np.inf, np.float32)
    np.fill_diagonal(Dm, 0.0)
    Dm[:n, :n] = W
    db, pbuf = ctx.upload(Dm), ctx.alloc(ld * ld * 4)
    eb, wb = ctx.upload(edges), ctx.upload(w)
    lga.fw_solve(ctx, n, db, ld, pbuf, eb, wb, len(edges))
    pred = pbuf.read(np.int32, (ld, ld))[:n, :n]
    Dd = oracle.floyd_minplus(W.astype(np.float64))
    np.testing.assert_array_equal(pred, oracle.predecessors(Dd, edges, w.astype(np.float64)))
    for b in (db, pbuf, eb, wb):
        b.free()


def test_diffuse_parity(lga, ctx, oracle):
    rng = np.random.default_rng(13)
    n = 700
    W = _random_graph(rng, n, 4.0, False) * np.float32(0.3)
    np.fill_diagonal(W, 0)
    Dd = oracle.floyd_minplus(W.astype(np.float64))
    Df = Dd.astype(np.float32)
    L = rng.uniform(0, 2, size=(n, 3)).astype(np.float32)
    sigma = np.array([1.1, 1.0, 0.9])
    ld = lga.fw_ld(n)
    Dm = np.full((ld, ld), np.inf, np.float32)
    Dm[:n, :n] = Df
    db, lb, ob = ctx.upload(Dm), ctx.upload(L), ctx.alloc(n * 12)
    lga.diffuse(ctx, n, db, ld, sigma, lb, ob)
    got = ob.read(np.float32, (n, 3))
    P = np.where(np.isfinite(Df)[..., None], np.exp(-Df.astype(np.float64)[..., None] * sigma), 0.0)
    want = oracle.diffuse(P, L.astype(np.float64))
    assert rel_err(got, want) < 1e-5
    for b in (db, lb, ob):
        b.free()


def test_full_solve_c1_against_oracle(lga, ctx, oracle, c1):
    """One LightGraph solve at the C1 configuration (N=1024, r=8, cap 12, edge
    step 0.5) through lg_solve, against the oracle's literal max-product solve."""
    grid, dense, _ = c1
    sh = _c1_shader(lga)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.parse_light("dir:0:1,1,1:1:-45,30,0")]
    s = lga.Solver(ctx, 1024, dense=True, with_pred=True)
    s.solve(grid, sh, lights, p, n=1024, radius=8.0, seed=854123, solve=5)
    r = s.results()
    ref = oracle.solve_points(dense, sh, p, lights, r["points"], radius=8.0, cap=12, want_P=True)
    np.testing.assert_array_equal(r["points"], oracle.sample_points(dense, 854123, 5, 1024))
    np.testing.assert_array_equal(r["edges"], ref["edges"])
    assert rel_err(r["L"], ref["L"]) < 1e-5
    tau_ref = np.where(ref["P"][..., 0] > 0, -np.log(np.maximum(ref["P"][..., 0], 1e-300)) / 1.1, np.inf)
    assert np.array_equal(np.isinf(r["D"]), np.isinf(tau_ref))
    fin = np.isfinite(tau_ref)
    assert rel_err(r["D"][fin], tau_ref[fin], floor=1e-30) < 1e-5
    assert rel_err(r["diffuse"], ref["diffuse"]) < 1e-5
    # determinism: the same solve again gives the same bits
    s.solve(grid, sh, lights, p, n=1024, radius=8.0, seed=854123, solve=5)
    r2 = s.results()
    for k in ("points", "edges", "D", "L", "diffuse", "pred"):
        np.testing.assert_array_equal(r[k], r2[k])
    s.destroy()


def _pred_rule(D, edges, w):
    """Canonical predecessor in float32 arithmetic: smallest neighbour k of j
    minimising D[i][k] + w(k, j); -1 if unreachable; i on the diagonal."""
    n = len(D)
    nbrs = [[] for _ in range(n)]
    for (a, b), x in zip(edges, w):
        if np.isfinite(x):
            nbrs[a].append((b, np.float32(x)))
            nbrs[b].append((a, np.float32(x)))
    out = np.full((n, n), -1, np.int32)
    for j in range(n):
        ks = np.array(sorted(k for k, _ in nbrs[j]), dtype=np.int64)
        ws = dict(nbrs[j])
        wv = np.array([ws[k] for k in ks], dtype=np.float32)
        for i in range(n):
            if i == j:
                out[i, j] = i
            elif np.isfinite(D[i, j]) and len(ks):
                vals = (D[i, ks] + wv).astype(np.float32)
                out[i, j] = ks[int(np.argmin(vals))]
    return out


def test_full_solve_sparse_components(lga, ctx, oracle):
    """A sparse graph (many small components, closed in LDS) against the oracle's
    dense max-product solve; predecessors against the canonical rule."""
    spec = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, 128, seed=9, octaves=6, blobs=20, axis_min=10, axis_max=40,
                          active_fraction=0.3)
    grid, _ = lga.Grid.generate(ctx, spec)
    dense = grid.read_dense()
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 4.0e4, (64.0, 200.0, 64.0)),
              lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")]
    n = 1500
    s = lga.Solver(ctx, n, dense=True, with_pred=True)
    s.solve(grid, sh, lights, p, n=n, radius=5.0, seed=77, solve=3)
    r = s.results()
    assert r["n_components"] > 50
    ref = oracle.solve_points(dense, sh, p, lights, r["points"], radius=5.0, cap=12, want_P=True)
    np.testing.assert_array_equal(r["edges"], ref["edges"])
    assert rel_err(r["L"], ref["L"]) < 1e-5
    sig0 = -sh.extinction[0]
    tau_ref = np.where(ref["P"][..., 0] > 0, -np.log(np.maximum(ref["P"][..., 0], 1e-300)) / sig0, np.inf)
    assert np.array_equal(np.isinf(r["D"]), np.isinf(tau_ref))
    fin = np.isfinite(tau_ref)
    assert rel_err(r["D"][fin], tau_ref[fin], floor=1e-30) < 1e-5
    assert rel_err(r["diffuse"], ref["diffuse"]) < 1e-5
    np.testing.assert_array_equal(r["pred"], _pred_rule(r["D"], r["edges"], r["weights"]))
    # without the dense outputs the radiance is the same
    s.solve(grid, sh, lights, p, n=n, radius=5.0, seed=77, solve=3, dense=False)
    np.testing.assert_array_equal(s.results(want_D=False)["diffuse"], r["diffuse"])
    s.destroy()
    grid.destroy()


def test_full_solve_giant_component_pred(lga, ctx, oracle, c1):
    """C1 graph: one component far above the LDS limit goes through the blocked
    kernels; its predecessors follow the canonical rule."""
    grid, dense, _ = c1
    sh = _c1_shader(lga)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    s = lga.Solver(ctx, 600, dense=True, with_pred=True)
    s.solve(grid, sh, [lga.parse_light("dir:0:1,1,1:1:-45,30,0")], p, n=600, radius=8.0, seed=3, solve=1)
    r = s.results()
    assert r["n_components"] < 10
    np.testing.assert_array_equal(r["pred"], _pred_rule(r["D"], r["edges"], r["weights"]))
    s.destroy()


@pytest.mark.parametrize("order", ["bfs", "nd"])
def test_full_solve_bfs_order(lga, ctx, oracle, c1, order):
    """LG_SOLVE_BFS_ORDER / LG_SOLVE_ND_ORDER on the C1 graph (one component far
    above 128 vertices): the same points, edges, weights and vertex light as the
    ascending order; the closure in the other pivot order agrees with the FP64
    oracle and with the ascending order within FP32 rounding (1e-5 relative, same
    +INF pattern); its predecessors follow the canonical rule; it computes no
    more remainder tiles."""
    grid, dense, _ = c1
    sh = _c1_shader(lga)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.parse_light("dir:0:1,1,1:1:-45,30,0")]
    n = 1024
    s = lga.Solver(ctx, n, dense=True, with_pred=True)
    s.solve(grid, sh, lights, p, n=n, radius=8.0, seed=854123, solve=5, stats=True)
    ref_order = s.results()
    tiles_ref = s.stats()["fw_updates"]
    s.solve(grid, sh, lights, p, n=n, radius=8.0, seed=854123, solve=5, stats=True, order=order)
    r = s.results()
    st = s.stats()
    assert st["big_components"] >= 1 and st["fw_updates"] <= tiles_ref
    for k in ("points", "edges", "weights", "L"):
        np.testing.assert_array_equal(r[k], ref_order[k], err_msg=k)
    assert np.array_equal(np.isinf(r["D"]), np.isinf(ref_order["D"]))
    fin = np.isfinite(r["D"])
    assert rel_err(r["D"][fin], ref_order["D"][fin], floor=1e-30) < 1e-5
    assert rel_err(r["diffuse"], ref_order["diffuse"]) < 1e-5
    ref = oracle.solve_points(dense, sh, p, lights, r["points"], radius=8.0, cap=12, want_P=True)
    tau_ref = np.where(ref["P"][..., 0] > 0, -np.log(np.maximum(ref["P"][..., 0], 1e-300)) / 1.1, np.inf)
    assert np.array_equal(np.isinf(r["D"]), np.isinf(tau_ref))
    assert rel_err(r["D"][fin], tau_ref[fin], floor=1e-30) < 1e-5
    assert rel_err(r["diffuse"], ref["diffuse"]) < 1e-5
    np.testing.assert_array_equal(r["pred"], _pred_rule(r["D"], r["edges"], r["weights"]))
    with pytest.raises(ValueError):
        s.solve(grid, sh, lights, p, n=n, radius=8.0, order="morton")
    s.destroy()

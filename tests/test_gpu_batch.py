"""Batched form: uneven graphs in one launch sequence give, graph by graph, the
same results as the plain per-graph path (bit-identical: every graph is
computed by the same per-component arithmetic whatever else is in the batch),
and the batched blocked Floyd-Warshall matches the oracle."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

pytestmark = pytest.mark.gpu


def _random_graph(rng, n, deg, integer):
    W = np.full((n, n), np.inf, np.float32)
    np.fill_diagonal(W, 0.0)
    m = int(deg * n / 2)
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    x = rng.integers(1, 20, m).astype(np.float32) if integer else rng.uniform(0.01, 3.0, m).astype(np.float32)
    for i, j, v in zip(a, b, x):
        if i != j:
            W[i, j] = W[j, i] = min(W[i, j], v)
    return W


def _padded(lga, W):
    n = W.shape[0]
    ld = lga.fw_ld(n)
    Dm = np.full((ld, ld), np.inf, np.float32)
    np.fill_diagonal(Dm, 0.0)
    Dm[:n, :n] = W
    return Dm, ld


SIZES = [1, 7, 128, 129, 300, 700, 1030, 2]


def test_fw_batched_equals_single_and_oracle(lga, ctx, oracle):
    rng = np.random.default_rng(21)
    Ws = [_random_graph(rng, n, 2.5, integer=(k % 2 == 0)) for k, n in enumerate(SIZES)]
    single, batch = [], []
    for W in Ws:
        Dm, ld = _padded(lga, W)
        b1, b2 = ctx.upload(Dm), ctx.upload(Dm)
        lga.fw_solve(ctx, W.shape[0], b1, ld)
        single.append(b1)
        batch.append(b2)
    lga.fw_solve_batched(ctx, SIZES, batch)
    for k, W in enumerate(Ws):
        n = W.shape[0]
        ld = lga.fw_ld(n)
        got = batch[k].read(np.float32, (ld, ld))
        ref = single[k].read(np.float32, (ld, ld))
        np.testing.assert_array_equal(got, ref)
        want = oracle.floyd_minplus(W.astype(np.float64))
        g = got[:n, :n].astype(np.float64)
        if k % 2 == 0:
            np.testing.assert_array_equal(g, want)
        else:
            assert np.array_equal(np.isinf(g), np.isinf(want))
            fin = np.isfinite(want)
            np.testing.assert_allclose(g[fin], want[fin], rtol=1e-5, atol=0)
        single[k].free()
        batch[k].free()


def test_fw_split_and_graph_schedules_are_identical(lga, ctx):
    """Quadrant-split tiles and the hipGraph-captured round schedule give the
    bytes of the plain schedule; the captured graph is replayed three times on
    fresh inputs of the same shape, with eager work and a synchronisation
    between replays, and every replay equals the eager result."""
    rng = np.random.default_rng(5)
    sizes = [300, 129, 700, 1030, 256, 7, 513]
    Ws = [_random_graph(rng, n, 2.5, integer=(k % 2 == 1)) for k, n in enumerate(sizes)]
    inputs = [_padded(lga, W)[0] for W in Ws]

    def run(split_below, use_graph):
        lga.set_fw_tuning(ctx, split_below, use_graph)
        bufs = [ctx.upload(Dm) for Dm in inputs]
        lga.fw_solve_batched(ctx, sizes, bufs)
        out = [b.read(np.float32, Dm.shape) for b, Dm in zip(bufs, inputs)]
        for b in bufs:
            b.free()
        return out

    saved = lga.fw_tuning(ctx)
    try:
        plain = run(0, 0)
        split = run(1 << 30, 0)
        for a, b in zip(plain, split):
            np.testing.assert_array_equal(a, b)
        for rep in range(3):  # capture, then two replays of the same schedule
            graph = run(1 << 30, 1)
            for a, b in zip(plain, graph):
                np.testing.assert_array_equal(a, b)
            eager = run(1 << 30, 0)  # eager work between replays
            for a, b in zip(plain, eager):
                np.testing.assert_array_equal(a, b)
            ctx.sync()
        for a, b in zip(plain, run(0, 1)):  # whole tiles, captured
            np.testing.assert_array_equal(a, b)
    finally:
        lga.set_fw_tuning(ctx, saved["split_below"], saved["use_graph"])


def test_fw_split_panels_in_large_launches(lga, ctx):
    """600 copies of one matrix in one batched closure: each split panel launch
    holds 2400 workgroups, more than the GPU keeps resident, so late workgroups
    start after others of the same tile have finished writing.  Every copy equals
    the whole-tile closure of the same input, split or not, and the one-matrix
    closure of it."""
    rng = np.random.default_rng(11)
    W = _random_graph(rng, 300, 3.0, integer=False)
    Dm = _padded(lga, W)[0]
    count = 600
    saved = lga.fw_tuning(ctx)
    try:
        results = []
        for split_below in (0, 1 << 30):
            lga.set_fw_tuning(ctx, split_below, 0)
            bufs = [ctx.upload(Dm) for _ in range(count)]
            lga.fw_solve_batched(ctx, [300] * count, bufs)
            out = [b.read(np.float32, Dm.shape) for b in bufs]
            for b in bufs:
                b.free()
            results.append(out)
        want = results[0][0]
        single = ctx.upload(Dm)
        lga.fw_solve(ctx, 300, single, Dm.shape[0])
        np.testing.assert_array_equal(single.read(np.float32, Dm.shape), want)
        single.free()
        for out in results:
            for k, got in enumerate(out):
                np.testing.assert_array_equal(got, want, err_msg=f"copy {k}")
    finally:
        lga.set_fw_tuning(ctx, saved["split_below"], saved["use_graph"])


def test_fw_autotune_measures_both_choices(lga, ctx):
    t = lga.fw_autotune(ctx)
    assert t["tuned"] == 1 and len(t["tiles"]) == 12 and t["tiles"][0] == 1 and t["tiles"][-1] == 4095
    assert all(x > 0 for x in t["whole_us"] + t["split_us"])
    assert t["stream_ms"] > 0 and t["graph_ms"] > 0
    assert 1 <= t["split_below"] and t["use_graph"] in (0, 1)
    # the choice follows the table: split below the first size where it stops winning
    wins = [s < w for s, w in zip(t["split_us"], t["whole_us"])]
    want = next((n for n, w in zip(t["tiles"], wins) if not w), 1 << 30)
    assert t["split_below"] == want
    # whole-tile launches grow with the tile count once the GPU is full
    assert t["whole_us"][-1] > t["whole_us"][0]
    assert lga.fw_tuning(ctx)["split_below"] == t["split_below"]


def test_fw_batched_empty_and_bad_ld(lga, ctx):
    lga.fw_solve_batched(ctx, [], [])
    b = ctx.alloc(4)
    with pytest.raises(lga.LGError):
        k = 1
        import ctypes as C
        n_arr = (C.c_int32 * k)(5)
        ld_arr = (C.c_int32 * k)(5)  # must be lg_fw_ld(5) = 128
        p_arr = (C.c_void_p * k)(b.ptr)
        lga._check(lga.lib().lg_fw_solve_batched(ctx.handle, k, n_arr, p_arr, ld_arr), "fw_solve_batched")
    b.free()


@pytest.fixture(scope="module")
def volumes(lga, ctx):
    """Six blob+fBm volumes of uneven side (64..160) with seeds 3000+k."""
    sides = [64, 96, 160, 72, 128, 88]
    grids = []
    for k, side in enumerate(sides):
        spec = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, side, seed=3000 + k, octaves=6, blobs=20,
                              axis_min=side * 30 // 512 + 4, axis_max=side * 120 // 512 + 8, active_fraction=0.3)
        g, _ = lga.Grid.generate(ctx, spec)
        grids.append(g)
    yield grids
    for g in grids:
        g.destroy()


def test_solve_batch_equals_plain(lga, ctx, volumes):
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 4.0e4, (64.0, 200.0, 64.0)),
              lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")]
    ns = [512, 1400, 3000, 700, 2100, 900]
    radius = 6.0
    items = [(g, n, 40 + k) for k, (g, n) in enumerate(zip(volumes, ns))]
    bs = lga.Solver(ctx, sum(ns))
    bs.solve_batch(items, sh, lights, p, radius=radius, seed=854123)
    got = bs.batch_results()
    plain = lga.Solver(ctx, max(ns))
    largest = 0
    for k, (g, n, solve) in enumerate(items):
        plain.solve(g, sh, lights, p, n=n, radius=radius, seed=854123, solve=solve)
        ref = plain.results(want_D=False)
        for key in ("points", "L", "diffuse", "edges", "weights"):
            np.testing.assert_array_equal(got[k][key], ref[key], err_msg=f"item {k} ({key})")
        assert np.all(np.isfinite(ref["diffuse"])) and ref["diffuse"].max() > 0
        e = ref["edges"][np.isfinite(ref["weights"])]
        A = csr_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
        _, lab = connected_components(A, directed=False)
        largest = max(largest, np.bincount(lab).max())
    assert largest > 128, "the batch should hold a component for the blocked Floyd-Warshall"
    bs.destroy()
    plain.destroy()


@pytest.mark.parametrize("order", ["bfs", "nd"])
def test_solve_batch_bfs_order(lga, ctx, volumes, order):
    """Breadth-first and nested-dissection pivot orders (LG_SOLVE_BFS_ORDER,
    LG_SOLVE_ND_ORDER): the batch equals the plain solves in the same order bit
    for bit, agrees with the ascending order within FP32 rounding, and skips
    remainder tiles of the large components."""
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 4.0e4, (64.0, 200.0, 64.0)),
              lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")]
    ns = [512, 1400, 3000, 700, 2100, 900]
    items = [(g, n, 40 + k) for k, (g, n) in enumerate(zip(volumes, ns))]
    bs = lga.Solver(ctx, sum(ns))
    bs.solve_batch(items, sh, lights, p, radius=6.0, seed=854123, order=order, stats=True)
    got = bs.batch_results()
    tiles_bfs = bs.stats()["fw_updates"]
    bs.solve_batch(items, sh, lights, p, radius=6.0, seed=854123, stats=True)
    asc = bs.batch_results()
    tiles_asc = bs.stats()["fw_updates"]
    assert 0 < tiles_bfs < tiles_asc
    plain = lga.Solver(ctx, max(ns))
    for k, (g, n, solve) in enumerate(items):
        plain.solve(g, sh, lights, p, n=n, radius=6.0, seed=854123, solve=solve, order=order)
        ref = plain.results(want_D=False)
        for key in ("points", "L", "diffuse", "edges", "weights"):
            np.testing.assert_array_equal(got[k][key], ref[key], err_msg=f"item {k} ({key})")
        d, a = got[k]["diffuse"].astype(np.float64), asc[k]["diffuse"].astype(np.float64)
        assert np.max(np.abs(d - a) / np.maximum(np.abs(a), 1e-6 * np.abs(a).max())) < 1e-5, f"item {k}"
    bs.destroy()
    plain.destroy()


def test_solve_batch_large_call_equals_plain(lga, ctx, volumes):
    """A call of 81,920 vertices (40 solves of 2048 points over the six volumes)
    takes the large-call forms of two stages -- the candidate search with 4
    lanes per vertex, and the shadow rays traced light by light in Morton order
    of their vertices, four rays per wave, where a plain solve gives each ray a
    wave -- and each solve still equals its plain solve bit for bit; with the
    work counters on, the batch's shadow rays and samples are the plain
    solves' sums."""
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 4.0e4, (64.0, 200.0, 64.0)),
              lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")]
    n = 2048
    items = [(volumes[k % len(volumes)], n, 200 + k) for k in range(40)]
    assert len(items) * n >= 65536
    bs = lga.Solver(ctx, n * len(items))
    plain = lga.Solver(ctx, n)
    for stats in (False, True):
        bs.solve_batch(items, sh, lights, p, radius=6.0, seed=854123, stats=stats)
        got = bs.batch_results()
        want = {"shadow_rays": 0, "shadow_samples": 0}
        for k, (g, nk, solve) in enumerate(items):
            plain.solve(g, sh, lights, p, n=nk, radius=6.0, seed=854123, solve=solve, stats=stats)
            ref = plain.results(want_D=False)
            for key in ("points", "L", "diffuse", "edges", "weights"):
                np.testing.assert_array_equal(got[k][key], ref[key], err_msg=f"item {k} ({key}, stats {stats})")
            if stats:
                st = plain.stats()
                for key in want:
                    want[key] += st[key]
        if stats:
            st = bs.stats()
            assert want["shadow_rays"] == len(items) * n * 2
            for key in want:
                assert st[key] == want[key], key
    bs.destroy()
    plain.destroy()


def test_solve_batch_rejects_mixed_dtypes(lga, ctx, volumes):
    d = volumes[0].read_dense()
    g16 = lga.Grid.from_dense(ctx, d, dtype=lga.LG_F16)
    sh = lga.default_shader()
    s = lga.Solver(ctx, 400)
    with pytest.raises(lga.LGError):
        s.solve_batch([(volumes[0], 200, 0), (g16, 200, 1)], sh, [lga.parse_light("dir:0:1,1,1:1:0,0,0")],
                      lga.march_params(), radius=4.0)
    s.destroy()
    g16.destroy()


def test_solve_stats_counts(lga, ctx, volumes):
    """Device work counters: rays = vertices x light samples, LDS closure work =
    sum of m^3 over the components of 2..128 vertices, tiles only for larger ones."""
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 4.0e4, (64.0, 200.0, 64.0)),
              lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")]
    n = 1400
    s = lga.Solver(ctx, n)
    s.solve(volumes[1], sh, lights, p, n=n, radius=6.0, seed=854123, solve=3, stats=True)
    st = s.stats()
    r = s.results(want_D=False)
    fin = np.isfinite(r["weights"])
    e = r["edges"][fin]
    A = csr_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
    nc, lab = connected_components(A, directed=False)
    sizes = np.bincount(lab)
    assert st["points"] == n and st["edges"] == len(r["edges"]) and st["components"] == nc
    assert st["shadow_rays"] == 2 * n and st["shadow_samples"] > st["shadow_rays"]
    assert st["edge_samples"] > 0
    small = sizes[(sizes >= 2) & (sizes <= 128)].astype(np.int64)
    assert st["small_updates"] == int((small ** 3).sum())
    assert st["small_updates_32"] == int((small[small <= 32] ** 3).sum())
    assert st["small_updates_64"] == int((small[(small > 32) & (small <= 64)] ** 3).sum())
    assert st["big_components"] == int((sizes > 128).sum())
    assert (st["fw_updates"] + st["fw_panel_updates"] > 0) == (st["big_components"] > 0)
    # whole 32-pivot chunks of 128 x 128 tile parts
    for k in ("fw_updates", "fw_panel_updates"):
        assert st[k] % (32 * 32 * 64) == 0
    storage = (np.array(volumes[1].read_dense().shape) // 8 + 2).prod()
    assert 0 < st["light_bricks"] <= storage and 0 < st["edge_bricks"] <= storage
    # without the flag the counters are not collected
    s.solve(volumes[1], sh, lights, p, n=n, radius=6.0, seed=854123, solve=3)
    assert s.stats()["shadow_samples"] == 0
    s.destroy()


def test_light_bricks_closed_form(lga, ctx):
    """One vertex at (20.5, 20.3, 20.7) of a uniform 64^3 grid lit by a point
    light straight above it: the shadow ray's samples run up the line x = 20.5,
    z = 20.7 from y = 20.8 to 63.8 in steps of 0.5, so their taps read the
    bricks (2, 2..8, 2) -- the last one the zero border above the grid --
    7 distinct bricks; with no edges the edge march reads none."""
    g = lga.Grid.from_dense(ctx, np.full((64, 64, 64), 0.01, np.float32))
    sh = lga.default_shader(absorption=0.1, scattering=1.0)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1, 1, 1), 1.0, (20.5, 200.0, 20.7))]
    s = lga.Solver(ctx, 4)
    s.write_points(np.array([[20.5, 20.3, 20.7]], np.float32))
    s.solve(g, sh, lights, p, n=1, radius=4.0, external_points=True, stats=True)
    st = s.stats()
    assert st["shadow_rays"] == 1 and st["shadow_samples"] == 87
    assert st["light_bricks"] == 7 and st["edge_bricks"] == 0
    s.destroy()
    g.destroy()

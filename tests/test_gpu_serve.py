"""Serving form (lg_server): continuous batching with HBM-budget admission.
Each job's films are the same alone and in a mix (bit for bit: the solve is
per-graph arithmetic, the render per job), equal the same stages called
directly, and the reserved memory never exceeds the budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GB = 1 << 30


def _setup(lga):
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=1.0)
    lights = [lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0"), lga.point_light(0, (1, 0.9, 0.8), 2e4, (40, 150, 40))]
    return sh, p, lights


def _job(lga, kind, size, seed, n, w, h, spp, iterations=1, solve_base=0, order="reference"):
    sh, p, lights = _setup(lga)
    if kind == "sphere":
        spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, size, seed=seed, octaves=4, radius=0.375 * size)
    else:
        spec = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, size, seed=seed, octaves=6, blobs=12, axis_min=size / 16,
                              axis_max=size / 4, active_fraction=0.3)
    cam = lga.camera_setup(cam_translate=(size / 2, size / 2, 2.6 * size), focal=60.0, width=w, height=h)
    return lga.job(spec, n, cam, sh, lights, p, iterations=iterations, radius=6.0, samples=spp,
                   solve_base=solve_base, order=order)


JOBS = [("sphere", 64, 1, 600, 48, 40, 1, 2), ("blobs", 96, 7, 1500, 64, 48, 2, 1), ("blobs", 80, 8, 900, 40, 32, 1, 1),
        ("sphere", 64, 2, 700, 32, 32, 3, 2)]


@pytest.mark.parametrize("order", ["reference", "bfs", "nd"])
def test_alone_equals_mixed_and_direct(lga, ctx, order):
    jobs = [_job(lga, *j, order=order) for j in JOBS]
    srv = lga.Server(8 * GB, workers=2, max_batch=4)
    alone = []
    for j in jobs:
        r = srv.wait(srv.submit(j))
        assert r["status"] == 0 and r["batch"] == 1
        alone.append(r["films"])
    ids = [srv.submit(j) for j in jobs]  # arrive together: batched and overlapped
    mixed = [srv.wait(i) for i in ids]
    for a, m in zip(alone, mixed):
        assert m["status"] == 0
        np.testing.assert_array_equal(a, m["films"])
    assert max(m["batch"] for m in mixed) > 1
    st = srv.stats()
    assert st["completed"] == 2 * len(jobs) and st["failed"] == 0
    assert st["peak_reserved_bytes"] <= st["budget_bytes"]
    srv.close()
    # the first job through the stages directly
    kind, size, seed, n, w, h, spp, iters = JOBS[0]
    sh, p, lights = _setup(lga)
    spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, size, seed=seed, octaves=4, radius=0.375 * size)
    grid, _ = lga.Grid.generate(ctx, spec)
    s = lga.Solver(ctx, n * iters)
    s.solve_batch([(grid, n, i) for i in range(iters)], sh, lights, p, radius=6.0, order=order)
    res = s.batch_results()
    pts = np.concatenate([r["points"] for r in res])
    dif = np.concatenate([r["diffuse"] for r in res])
    sg = lga.ScatterGrid(ctx, grid, 4)
    rad, _ = lga.splat_radius(pts, 3 * iters)
    pb, rb = ctx.upload(pts), ctx.upload(dif)
    sg.splat(pb, rb, len(pts), rad)
    cam = lga.camera_setup(cam_translate=(size / 2, size / 2, 2.6 * size), focal=60.0, width=w, height=h)
    films = lga.Films(ctx, w, h)
    lga.render(ctx, grid, sg, None, cam, sh, lights, p, spp, films, seed=1)
    np.testing.assert_array_equal(films.read(), alone[0])
    assert alone[0][0][..., 3].max() > 0.2
    films.free()
    for b in (pb, rb):
        b.free()
    sg.destroy()
    s.destroy()
    grid.destroy()


def test_budget_admission(lga):
    jobs = [_job(lga, "blobs", 96, 20 + k, 1200, 48, 40, 1) for k in range(4)]
    probe = lga.Server(8 * GB, workers=1, max_batch=1)
    need = probe.wait(probe.submit(jobs[0]))["reserved_bytes"]
    probe.close()
    # room for one job at a time: the others wait for memory, all complete
    srv = lga.Server(int(need * 1.5), workers=2, max_batch=4)
    ids = [srv.submit(j) for j in jobs]
    rs = [srv.wait(i) for i in ids]
    assert all(r["status"] == 0 for r in rs)
    st = srv.stats()
    assert st["peak_reserved_bytes"] <= st["budget_bytes"] and st["completed"] == 4
    # runs were serialised by the budget: each started after the previous finished
    starts = sorted((r["t_start"], r["t_done"]) for r in rs)
    for (s0, d0), (s1, d1) in zip(starts, starts[1:]):
        assert s1 >= d0 - 1e-6
    # a job larger than the whole budget is refused at submit
    big = _job(lga, "blobs", 384, 5, 4000, 256, 256, 1)
    with pytest.raises(lga.LGError) as e:
        srv.submit(big)
    assert e.value.code == -4
    srv.close()


def test_concurrent_scene_generation(lga):
    """Two contexts on two threads generate different blob scenes at the same
    time (as serving workers do); each volume equals the one generated alone.
    The blob table lives per context: a table shared by the process let one
    thread's volume be generated with the other's blobs."""
    import threading

    specs = [lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, 128, seed=900 + k, octaves=4, blobs=20 - 8 * k, axis_min=6.0,
                            axis_max=30.0, active_fraction=0.3) for k in range(2)]
    want = []
    with lga.Context(0) as c:
        for spec in specs:
            g, _ = lga.Grid.generate(c, spec)
            want.append(g.read_dense())
            g.destroy()
    assert not np.array_equal(want[0], want[1])
    errors = []

    def work(k):
        try:
            with lga.Context(0) as c:
                for _ in range(8):
                    g, _ = lga.Grid.generate(c, specs[k])
                    got = g.read_dense()
                    g.destroy()
                    if not np.array_equal(got, want[k]):
                        errors.append(k)
        except Exception as e:  # reported below, not lost in the thread
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert errors == []

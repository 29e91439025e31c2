"""LightGraph solves/s on one MI355X.

Workload (the plain method at production size): a 512^3 fp32 cloud made of 20
seeded ellipsoid blobs (semi-axes 30-120 voxels) times 6-octave fBm, thresholded
to ~30% active voxels in 8^3 bricks; sigma_s = (1.0, 0.9, 0.8), sigma_a = 0.1; a
point light and a directional light; N = 8192 density-sampled points per solve,
edges within 12 voxels (12 per vertex at most), edge march at 0.5 voxel.  A step
is one LightGraph solve (sample -> edges -> edge march -> all-pairs -> vertex
light -> diffusion) with the volume resident in HBM.

Beside the headline, the line carries:
  * "latency": p50/p99 of one solve issued alone (host clock, solve + sync);
  * "batched": the batching workload -- 64 seeded blob+fBm volumes of side
    128..384 (step 8, seeds 3000+k), N drawn from [512, 6144] -- solved one graph
    at a time (plain) and as one batch, solves/s of each, their ratio, and the
    largest difference between a graph's batched and plain radiance (0 when
    the batch reproduces every graph exactly).

Prints one JSON line.  --gpus above 1 is refused: the benchmark runs on one GPU.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Kernel-level timers (one kernel, or one kernel per size class) that can carry
# the roofline, and stage timers reported beside them.
KERNELS = ("k_light_rays", "k_march_edges", "k_fw_minplus", "k_fw_panel", "k_fw_pivot", "k_comp_small_128",
           "k_comp_small_64", "k_comp_small_32", "k_sample_points", "k_big_fill", "k_diffuse")
STAGES = ("k_vertex_light", "k_graph", "k_cand", "k_greedy", "k_components", "k_union_find", "k_comp_sort")
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
# VALU issue ceiling for 32-bit operations: 256 CUs x 4 SIMDs x 2.4 GHz, a wave64
# FP32/INT32 instruction every 2 cycles per SIMD (32 lanes a cycle: the 157.3
# TFLOP/s FP32 peak is 256 x 4 x 32 lanes x 2 x 2.4 GHz) = 1228.8e9 wave
# instructions/s = 78.6e12 lane instructions/s; FP64 runs at half that.  The
# dense min-plus kernel measures 8.93e11 wave instructions/s under SQ_INSTS_VALU,
# above what a 4-cycle issue would allow.  A min-plus update costs 1.5 32-bit
# instructions (two adds and one v_min3_f32 per pivot pair), so Floyd-Warshall
# tops out at 52.4e12 updates/s = 104.9 T ops/s (add + min).
VALU_LANE_INSTR_PER_S = 256 * 4 * 2.4e9 * 32
FW_OPS_CEILING_TFLOPS = VALU_LANE_INSTR_PER_S / 1.5 * 2 / 1e12
# Marches: t_roof = max(B_comp / 8 TB/s, S * 32 / 157.3 TFLOP/s) with
# S trilinear samples at 32 FP32 ops each and B_comp the compulsory bytes: the
# distinct 8^3 bricks the samples' taps read (2 KiB each in FP32) plus inputs and
# outputs.  kernel -> (sample counter, brick counter, input + output bytes)
MARCH = {
    # points in (12 B each), one FP64 RGB sum out per ray (24 B)
    "k_light_rays": ("shadow_samples", "light_bricks", lambda w: 12.0 * w.get("points", 0) + 24.0 * w.get("shadow_rays", 0)),
    # points in, edges in (8 B), optical depth out (4 B)
    "k_march_edges": ("edge_samples", "edge_bricks", lambda w: 12.0 * w.get("points", 0) + 12.0 * w.get("edges", 0)),
}
# Floyd-Warshall and the LDS closures: 2 ops (add + min) per executed update
FW = {
    "k_fw_minplus": lambda w: w.get("fw_updates", 0),
    "k_fw_panel": lambda w: w.get("fw_panel_updates", 0),
    "k_fw_pivot": lambda w: w.get("fw_pivot_updates", 0),
    "k_comp_small_128": lambda w: w.get("small_updates", 0) - w.get("small_updates_64", 0) - w.get("small_updates_32", 0),
    "k_comp_small_64": lambda w: w.get("small_updates_64", 0),
    "k_comp_small_32": lambda w: w.get("small_updates_32", 0),
}
BRICK_BYTES = 512 * 4
OPS_PER_SAMPLE = 32.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
RENDER_KERNELS = ("k_render_count", "k_render_fill", "k_render_shadow", "k_render_combine")

WORKLOAD = {"workload": "c2_plain_solve", "volume": "512^3 blobs20+fBm6, ~30% active, fp32 bricks",
            "points": 8192, "edge_radius": 12.0, "max_connect": 12, "edge_step": 0.5,
            "lights": "point + directional", "sigma_s": [1.0, 0.9, 0.8], "sigma_a": 0.1}


def setup(lga, ctx, n_points):
    spec = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, 512, seed=2, octaves=6, blobs=20, axis_min=30.0,
                          axis_max=120.0, active_fraction=0.30)
    grid, _ = lga.Grid.generate(ctx, spec)
    shader = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    params = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5)
    lights = [lga.point_light(0.0, (1.0, 0.9, 0.8), 4.0e5, (256.0, 700.0, 256.0)),
              lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")]
    return spec, grid, shader, params, lights


def batch_workload(lga, ctx, count, seed=3000):
    """count volumes: side U{128..384} step 8, N U{512..6144}, volume seeds seed+k."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = rng.integers(0, (384 - 128) // 8 + 1, count) * 8 + 128
    ns = rng.integers(512, 6145, count)
    grids = []
    for k in range(count):
        spec = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, int(sides[k]), seed=seed + k, octaves=6, blobs=20,
                              axis_min=30.0, axis_max=120.0, active_fraction=0.30)
        g, _ = lga.Grid.generate(ctx, spec)
        grids.append(g)
    return grids, [int(x) for x in sides], [int(x) for x in ns]


ORDERS = ("reference", "bfs", "nd")


def bench_batched(lga, ctx, shader, params, lights, count, reps):
    """C3: the 64 volumes solved one by one (plain) and as one batch, each in the
    three pivot orders: ascending (the reference's, the drop-in path),
    breadth-first (LG_SOLVE_BFS_ORDER) and nested dissection (LG_SOLVE_ND_ORDER).
    Every speedup is against the plain ascending-order pass; the roofline is that
    of the nested-dissection batch."""
    import numpy as np
    radius = 12.0
    grids, sides, ns = batch_workload(lga, ctx, count)
    items = [(g, n, 5000 + k) for k, (g, n) in enumerate(zip(grids, ns))]
    plain = lga.Solver(ctx, max(ns))
    batch = lga.Solver(ctx, sum(ns))

    def run_plain(order, keep=False):
        outs = []
        for g, n, sv in items:
            plain.solve(g, shader, lights, params, n=n, radius=radius, seed=854123, solve=sv, order=order)
            if keep:
                outs.append(plain.results(want_D=False)["diffuse"])
        return outs

    def run_batch(order, stats=False):
        batch.solve_batch(items, shader, lights, params, radius=radius, seed=854123, order=order, stats=stats)

    def rel_diff(got, ref):
        worst = 0.0
        for a, b in zip(got, ref):
            if len(a):
                a64, b64 = a.astype(np.float64), b.astype(np.float64)
                worst = max(worst, float(np.max(np.abs(a64 - b64) / np.maximum(np.abs(b64), 1e-6 * np.abs(b64).max()))))
        return worst

    plain_out = {order: run_plain(order, keep=True) for order in ORDERS}
    diffs = {}
    for order in ORDERS:
        run_batch(order)
        got = [r["diffuse"] for r in batch.batch_results()]
        diffs[order] = {"max_abs_diff_vs_plain_same_order": max(float(np.max(np.abs(a - b))) if len(a) else 0.0
                                                                for a, b in zip(got, plain_out[order])),
                        "max_rel_diff_vs_plain_reference_order": rel_diff(got, plain_out["reference"])}
    ms = {}
    for name, fn in (("plain", run_plain), ("batched", run_batch)):
        for order in ORDERS:
            ctx.sync()
            ctx.mark(2)
            if name == "batched" and order == "nd":
                ctx.kernel_time_reset()
                ctx.kernel_timing(True)
            for _ in range(reps):
                fn(order)
            ctx.mark(3)
            ctx.sync()
            ctx.kernel_timing(False)
            ms[(name, order)] = ctx.elapsed_ms(2, 3) / reps
    btimes = kernel_times(ctx, KERNELS + STAGES)
    tiles = {}
    for order in ORDERS:
        run_batch(order, stats=True)
        st = batch.stats()
        tiles[order] = int(st["fw_updates"])
    bwork = {k: v * reps for k, v in st.items()}
    broof = kernel_roofline(btimes, bwork, "c3_batch_nd_order")
    plain.destroy()
    batch.destroy()
    for g in grids:
        g.destroy()
    t_plain = ms[("plain", "reference")]
    per_s = {f"{a}_{b}": round(count * 1000.0 / t, 2) for (a, b), t in ms.items()}
    return {"workload": "c3_batch", "volumes": count, "side_range": [min(sides), max(sides)],
            "points_total": int(sum(ns)), "n_range": [min(ns), max(ns)], "edge_radius": radius, "reps": reps,
            "ms": {f"{a}_{b}": round(t, 3) for (a, b), t in ms.items()},
            "solves_per_s": per_s,
            "plain_ms": round(t_plain, 3), "batched_ms": round(ms[("batched", "reference")], 3),
            "batched_nd_ms": round(ms[("batched", "nd")], 3),
            "plain_solves_per_s": per_s["plain_reference"],
            "batched_solves_per_s": per_s["batched_reference"],
            "batched_nd_solves_per_s": per_s["batched_nd"],
            "speedup": round(t_plain / ms[("batched", "reference")], 3),
            "speedup_bfs": round(t_plain / ms[("batched", "bfs")], 3),
            "speedup_nd": round(t_plain / ms[("batched", "nd")], 3),
            "speedup_note": "plain = one lg_solve per volume in ascending pivot order (the drop-in path); "
                            "speedup / speedup_bfs / speedup_nd = the batch in ascending / breadth-first / "
                            "nested-dissection order, each over plain ascending order",
            "fw_remainder_updates_per_pass": tiles,
            "diffs": diffs,
            "max_abs_diff_batched_vs_plain": diffs["reference"]["max_abs_diff_vs_plain_same_order"],
            "roofline": broof,
            "kernels": {k: {"launches": v["launches"], "avg_ms": v["avg_ms"]} for k, v in btimes.items()}}


def bench_render(lga, ctx, grid, shader, lights, solver, n, reps):
    """The C2 image: 1920 x 1080, 4 AA samples, the scatter grid splatted from one
    solve's radiance (3 points per kernel, factor-4 coarse grid), primary step 1,
    shadow step 3 (SceneSettings defaults)."""
    solver.solve(grid, shader, lights, lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=0.5),
                 n=n, radius=12.0, seed=854123, solve=777)
    r = solver.results(want_D=False)
    sg = lga.ScatterGrid(ctx, grid, 4)
    radius, _ = lga.splat_radius(r["points"], 3)
    pb, rb = ctx.upload(r["points"]), ctx.upload(r["diffuse"])
    ctx.sync()
    ctx.mark(5)
    sg.splat(pb, rb, len(r["points"]), radius)
    ctx.mark(6)
    ctx.sync()
    splat_ms = ctx.elapsed_ms(5, 6)
    w, h, spp = 1920, 1080, 4
    cam = lga.camera_setup(cam_translate=(256.0, 256.0, 1400.0), focal=50.0, width=w, height=h)
    films = lga.Films(ctx, w, h)
    rp = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=3.0)
    lga.render(ctx, grid, sg, None, cam, shader, lights, rp, spp, films, seed=1)  # warm-up
    ctx.sync()
    ctx.mark(5)
    for _ in range(reps):
        lga.render(ctx, grid, sg, None, cam, shader, lights, rp, spp, films, seed=1)
    ctx.mark(6)
    ctx.sync()
    ms = ctx.elapsed_ms(5, 6) / reps
    img = films.read()
    alpha_mean = float(img[0][..., 3].mean())
    # one more render with kernel timing and the device work counters on
    ctx.kernel_time_reset()
    ctx.kernel_timing(True)
    lga.render(ctx, grid, sg, None, cam, shader, lights, rp, spp, films, seed=1)
    ctx.sync()
    ctx.kernel_timing(False)
    st = lga.render_stats(ctx)
    phases = {}
    for nm in RENDER_KERNELS:
        launches, tot = ctx.kernel_time(nm)
        if launches:
            phases[nm] = {"launches": launches, "ms": round(tot, 3)}
    # and one counting the distinct bricks each shadow launch reads (the marking
    # slows it: not timed)
    ctx.brick_counting(True)
    ctx.kernel_timing(True)
    lga.render(ctx, grid, sg, None, cam, shader, lights, rp, spp, films, seed=1)
    ctx.sync()
    ctx.kernel_timing(False)
    ctx.brick_counting(False)
    bricks = lga.render_stats(ctx)["shadow_bricks"]
    ctx.kernel_time_reset()
    films.free()
    sg.destroy()
    pb.free()
    rb.free()
    out = {"workload": "c2_render", "width": w, "height": h, "spp": spp, "ms": round(ms, 3),
           "camera_rays_per_s": round(w * h * spp / (ms * 1e-3), 1), "splat_ms": round(splat_ms, 4),
           "primary_step": 1.0, "shadow_step": 3.0, "reps": reps, "mean_alpha": round(alpha_mean, 4),
           "phases": phases, "work": {k: v for k, v in st.items() if k not in ("shadow_lane_slots", "shadow_bricks")},
           "shadow_lane_efficiency": round(st["shadow_samples"] / max(st["shadow_lane_slots"], 1), 4)}
    sh = phases.get("k_render_shadow")
    if sh:
        # the shadow phase (kernel d's dominant kernel): shadow-ray trilinear
        # samples and the bricks they read; inputs 32 B per shading point, light
        # sums out 24 B per (light, point)
        io = 32.0 * st["shading_points"] + 24.0 * len(lights) * st["shading_points"]
        roof = {"kernel": "k_render_shadow", "traffic": None, "launches": sh["launches"],
                "avg_launch_ms": round(sh["ms"] / sh["launches"], 4)}
        roof.update(march_roofline(st["shadow_samples"], bricks, io, sh["ms"] * 1e-3))
        roof["shadow_samples_per_s"] = round(st["shadow_samples"] / (sh["ms"] * 1e-3), 1)
        # what the counters show holds it (profiles/k_render_c2_1080p_kernel_trace.md):
        # neither term of t_roof but the latency of dependent loads
        roof["limiter"] = "latency: dependent brick-code and tap loads, 5 waves/SIMD, 0.58 lane efficiency"
        out["roofline"] = roof
    return out


def bench_iter(lga, ctx, grid, shader, params, lights, n, radius, iters_list, reps=2):
    """The reference's own shape: `-iter` solves of one volume (main.cc:534-538,
    scenes/*.sh use 25..500), as LightGraph::operator()(0, iter) issues them --
    one lg_solve_batch of `iter` solves of the production volume -- in the
    ascending (drop-in) and nested-dissection pivot orders."""
    out = {"workload": "c2_iter_batch", "points_per_solve": n, "runs": {}}
    big = lga.Solver(ctx, n * max(iters_list))
    for iters in iters_list:
        items = [(grid, n, k) for k in range(iters)]
        row = {}
        for order in ("reference", "nd"):
            big.solve_batch(items, shader, lights, params, radius=radius, order=order)  # warm-up
            ctx.sync()
            ctx.mark(8)
            for _ in range(reps):
                big.solve_batch(items, shader, lights, params, radius=radius, order=order)
            ctx.mark(9)
            ctx.sync()
            ms = ctx.elapsed_ms(8, 9) / reps
            row[order] = {"ms": round(ms, 3), "solves_per_s": round(iters * 1000.0 / ms, 1)}
        out["runs"][str(iters)] = row
    big.destroy()
    return out


def bench_fw_dense(lga, ctx, solver, n):
    """Kernel (c) at the configured size: blocked Floyd-Warshall on the dense
    n x n optical-depth matrix of one C2 graph (edges of the last solve), the
    whole matrix rather than its components.  Algorithmic ops: 2 n^3."""
    import numpy as np
    r = solver.results(want_D=False)
    ld = lga.fw_ld(n)
    W = np.full((ld, ld), np.inf, np.float32)
    np.fill_diagonal(W, 0.0)
    e, w = r["edges"], r["weights"]
    W[e[:, 0], e[:, 1]] = w
    W[e[:, 1], e[:, 0]] = w
    buf = ctx.upload(W)
    del W
    ctx.kernel_time_reset()
    ctx.kernel_timing(True)
    ctx.sync()
    ctx.mark(5)
    lga.fw_solve(ctx, n, buf, ld)
    ctx.mark(6)
    ctx.sync()
    ctx.kernel_timing(False)
    ms = ctx.elapsed_ms(5, 6)
    t = kernel_times(ctx, ("k_fw_minplus", "k_fw_panel", "k_fw_pivot"))
    buf.free()
    alg = 2.0 * float(n) ** 3
    return {"workload": "c2_fw_dense", "n": n, "ms": round(ms, 3),
            "algorithmic_tflops": round(alg / (ms * 1e-3) / 1e12, 2),
            "frac_of_fp32_peak": round(alg / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
            "executed_tflops": round(alg / 2 / (ms * 1e-3) / 1e12, 2),
            "issue_ceiling_tflops": round(FW_OPS_CEILING_TFLOPS, 2),
            "executed_frac_of_issue_ceiling": round(alg / 2 / (ms * 1e-3) / 1e12 / FW_OPS_CEILING_TFLOPS, 4),
            "note": "2 n^3 ops as the reference's triple loop; the symmetric blocked kernel computes half the "
                    "tiles (executed = n^3 ops); the issue ceiling is 1.5 VALU instructions per update",
            "kernels": {k: {"launches": v["launches"], "avg_ms": v["avg_ms"]} for k, v in t.items()}}


def bench_serving(lga, ctx, rates, jobs_per_rate, workers, max_batch, seed=5):
    """Serving workload (scaled-down C5): Poisson arrivals at each rate of render
    jobs drawn from 12 seeded blob volumes of side 128..768 (step 64), N in
    [1000, 16000], images 512^2 / 1280x720 / 1920x1080 at 1-4 spp, one of three
    light sets (directional, point, environment); one LightGraph solve per job.
    Latency = arrival -> films on the host.  The served runs close the graphs in
    nested-dissection pivot order; the same job list is also run, in the
    ascending (drop-in) pivot order, through a one-worker, one-job-per-batch
    server ("plain": volumes, solver and buffers stay resident) and job by job
    the way one command-line run does it ("dropin": everything built and freed
    per job), back to back."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = [int(x) for x in rng.integers(0, (768 - 128) // 64 + 1, 12) * 64 + 128]
    pool = [lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, sides[k], seed=7000 + k, octaves=6, blobs=20,
                           axis_min=sides[k] * 30 / 512, axis_max=sides[k] * 120 / 512, active_fraction=0.30)
            for k in range(12)]
    h, w = 128, 256
    y = np.linspace(0, 1, h, dtype=np.float32)[:, None, None]
    sky = np.broadcast_to(np.concatenate([0.2 + 0.6 * y, 0.3 + 0.5 * y, 0.9 - 0.2 * y], -1), (h, w, 3)).copy()
    sky[20:23, 150:153] = 5000.0
    env = lga.EnvMap(ctx, sky)
    sh = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8))
    mp = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=3.0)
    light_sets = [[lga.parse_light("dir:0:1,0.89,0.79:0.7:-41.8,-180,0")],
                  [lga.point_light(0, (1, 0.9, 0.8), 4e5, (256, 900, 256))],
                  [lga.env_light(env, 0.0, (1, 1, 1), 1.0, samples=1)]]
    images = [(512, 512), (1280, 720), (1920, 1080)]

    def make_jobs(n, order):
        out = []
        for k in range(n):
            v = pool[int(rng.integers(12))]
            iw, ih = images[int(rng.integers(3))]
            cam = lga.camera_setup(cam_translate=(v.n / 2, v.n / 2, 2.8 * v.n), focal=50.0, width=iw, height=ih)
            out.append(lga.job(v, int(rng.integers(1000, 16001)), cam, sh, light_sets[int(rng.integers(3))], mp,
                               radius=12.0, samples=int(rng.integers(1, 5)), solve_base=k, order=order))
        return out

    def run(jobs, arrivals, nworkers, nbatch):
        srv = lga.Server(160 << 30, workers=nworkers, max_batch=nbatch)
        t0 = time.perf_counter()
        ids = []
        for k, (j, a) in enumerate(zip(jobs, arrivals)):
            dt = a - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(j))
            if k % 50 == 49:
                print(f"serving: {k + 1} of {len(jobs)} jobs submitted", file=sys.stderr, flush=True)
        rs = [srv.wait(i) for i in ids]
        st = srv.stats()
        srv.close()
        lat = sorted(r["t_done"] - r["t_submit"] for r in rs)
        span = max(r["t_done"] for r in rs) - min(r["t_submit"] for r in rs)
        pct = lambda q: round(1e3 * lat[min(len(lat) - 1, int(q * len(lat)))], 2)
        return {"jobs": len(rs), "failed": sum(r["status"] != 0 for r in rs), "solves_per_s": round(len(rs) / span, 3),
                "p50_ms": pct(0.5), "p99_ms": pct(0.99), "mean_batch": round(float(np.mean([r["batch"] for r in rs])), 2),
                "peak_reserved_gb": round(st["peak_reserved_bytes"] / 2 ** 30, 2)}

    out = {"workload": "c5_serving", "volumes": {"count": 12, "sides": sides}, "n_range": [1000, 16000],
           "images": ["512x512", "1280x720", "1920x1080"], "spp": [1, 4], "workers": workers, "max_batch": max_batch,
           "rates": {}}
    for lam in rates:
        jobs = make_jobs(jobs_per_rate, "nd")
        arrivals = np.cumsum(rng.exponential(1.0 / lam, jobs_per_rate))
        out["rates"][str(lam)] = run(jobs, arrivals, workers, max_batch)
    def dropin(jobs):
        """Each job the way one command-line run does it, back to back: build the
        volume, a solver, the scatter grid and the films, solve, splat, render,
        read the films back, free everything."""
        t0 = time.perf_counter()
        for j in jobs:
            grid, temp = lga.Grid.generate(ctx, j.scene)
            lights = [j.lights[k] for k in range(j.n_lights)]
            n = j.points
            s = lga.Solver(ctx, n)
            s.solve(grid, j.shader, lights, j.march, n=n, radius=j.radius, seed=j.seed, solve=j.solve_base,
                    max_connect=j.max_connect, temperature=temp, env_seed=j.env_seed)
            r = s.results(want_D=False)
            radius, _ = lga.splat_radius(r["points"], j.pts_per_kernel)
            sg = lga.ScatterGrid(ctx, grid, j.scatter_res_factor)
            pb, rb = ctx.upload(r["points"]), ctx.upload(r["diffuse"])
            sg.splat(pb, rb, len(r["points"]), radius)
            films = lga.Films(ctx, j.camera.width, j.camera.height)
            lga.render(ctx, grid, sg, temp, j.camera, j.shader, lights, j.march, j.samples, films, seed=j.render_seed,
                       env_seed=j.env_seed)
            films.read()
            for x in (films, pb, rb):
                x.free()
            sg.destroy()
            s.destroy()
            grid.destroy()
            if temp is not None:
                temp.destroy()
        ctx.sync()
        return {"jobs": len(jobs), "solves_per_s": round(len(jobs) / (time.perf_counter() - t0), 3)}

    state = rng.bit_generator.state
    jobs = make_jobs(jobs_per_rate, "nd")
    out["saturated"] = run(jobs, np.zeros(len(jobs)), workers, max_batch)
    rng.bit_generator.state = state  # the same jobs in the drop-in pivot order
    plain_jobs = make_jobs(jobs_per_rate, "reference")
    out["plain"] = run(plain_jobs, np.zeros(len(jobs)), 1, 1)
    out["dropin"] = dropin(plain_jobs)
    out["served_vs_plain"] = round(out["saturated"]["solves_per_s"] / out["plain"]["solves_per_s"], 3)
    out["served_vs_dropin"] = round(out["saturated"]["solves_per_s"] / out["dropin"]["solves_per_s"], 3)
    env.destroy()
    return out


def bench_precision(lga, ctx, reps=3):
    """Precision workload (C4): 1024^3 plume with temperature, bricks stored as
    FP32 / FP16 / BF16; solve (N = 12288, r = 12, env map 1024 x 512 with a 5000x
    sun + directional light) and a 960 x 540 render per brick type.  Errors
    against the FP64 oracle for the same workload are measured by
    tests/test_gpu_precision.py."""
    import numpy as np
    h, w = 512, 1024
    y = np.linspace(0, 1, h, dtype=np.float32)[:, None, None]
    sky = np.broadcast_to(np.concatenate([0.25 + 0.5 * y, 0.35 + 0.45 * y, 0.8 - 0.2 * y], -1), (h, w, 3)).copy()
    yy, xx = np.mgrid[0:h, 0:w]
    sky[(xx - 0.7 * w) ** 2 + (yy - 0.22 * h) ** 2 <= (0.012 * w) ** 2 + 1] *= 5000.0
    env = lga.EnvMap(ctx, sky)
    ramp = "0.0,0.33,0.66,1.0/0,0,0:1.0,0.076,0.0:1.0,0.322,0.0:1.0,0.45,0.15/0.35/1.5"
    sh = lga.default_shader(absorption=0.1, scattering=1.0, emission_scale=2.0, ramp=ramp)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=1.0)
    lights = [lga.env_light(env, 20.0, (1, 1, 1), 0.5, samples=2), lga.parse_light("dir:0:1,0.9,0.8:0.8:-40,30,0")]
    out = {"workload": "c4_precision", "volume": "plume 1024^3 + temperature", "points": 12288, "radius": 12.0,
           "image": "960x540x1", "types": {}}
    temp = None
    for name, dt in (("fp32", lga.LG_F32), ("fp16", lga.LG_F16), ("bf16", lga.LG_BF16)):
        spec = lga.scene_spec(lga.LG_SCENE_PLUME, 1024, seed=4, octaves=6, with_temperature=1, dtype=dt)
        grid, t = lga.Grid.generate(ctx, spec)
        if temp is None:
            temp = t
        else:
            t.destroy()
        s = lga.Solver(ctx, 12288)
        s.solve(grid, sh, lights, p, n=12288, radius=12.0, solve=0, temperature=temp)
        ctx.sync()
        ctx.mark(5)
        for k in range(reps):
            s.solve(grid, sh, lights, p, n=12288, radius=12.0, solve=1 + k, temperature=temp)
        ctx.mark(6)
        ctx.sync()
        solve_ms = ctx.elapsed_ms(5, 6) / reps
        s.destroy()
        cam = lga.camera_setup(cam_translate=(512.0, 460.0, 2450.0), focal=50.0, width=960, height=540)
        films = lga.Films(ctx, 960, 540)
        lga.render(ctx, grid, None, temp, cam, sh, lights, p, 1, films, seed=1)
        ctx.sync()
        ctx.mark(5)
        lga.render(ctx, grid, None, temp, cam, sh, lights, p, 1, films, seed=1)
        ctx.mark(6)
        ctx.sync()
        render_ms = ctx.elapsed_ms(5, 6)
        films.free()
        out["types"][name] = {"grid_gb": round(grid.info().device_bytes / 2 ** 30, 3), "solve_ms": round(solve_ms, 3),
                              "render_ms": round(render_ms, 3)}
        grid.destroy()
    temp.destroy()
    env.destroy()
    return out


def kernel_times(ctx, names):
    out = {}
    for nm in names:
        launches, total = ctx.kernel_time(nm)
        if launches:
            out[nm] = {"launches": launches, "total_ms": total, "avg_ms": round(total / launches, 5)}
    return out


def march_roofline(samples, bricks, io_bytes, t_s):
    """t_roof / t for one march launch (or a sum of launches): the compulsory-byte
    term against HBM and the trilinear-sample term against the FP32 peak; the
    bound is the larger term."""
    b_comp = bricks * BRICK_BYTES + io_bytes
    t_hbm = b_comp / (PEAK_HBM_GBS * 1e9)
    t_valu = samples * OPS_PER_SAMPLE / (PEAK_FP32_TFLOPS * 1e12)
    if t_hbm >= t_valu:
        achieved = b_comp / t_s / 1e9
        out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
               "frac": round(achieved / PEAK_HBM_GBS, 4)}
    else:
        achieved = samples * OPS_PER_SAMPLE / t_s / 1e12
        out = {"bound": "valu", "achieved": round(achieved, 4), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
               "frac": round(achieved / PEAK_FP32_TFLOPS, 5)}
    out.update({"compulsory_bytes": int(b_comp), "bricks": int(bricks), "samples": int(samples),
                "t_roof_hbm_us": round(t_hbm * 1e6, 3), "t_roof_valu_us": round(t_valu * 1e6, 3),
                "issued_gather_gbs": round(samples * 8 * 4 / t_s / 1e9, 1)})
    return out


def kernel_roofline(times, work, workload, name=None):
    """Roofline object for one kernel (by default the one with the largest total
    time among KERNELS) over its launches in `times`: the marches by t_roof =
    max(B_comp / 8 TB/s, S * 32 / 157.3 TFLOP/s) (compulsory bytes or trilinear
    samples, whichever bounds), the Floyd-Warshall kernels by executed min-plus
    updates (2 ops each) against the FP32 peak and the 32-bit VALU issue
    ceiling.  Work comes from the device counters of the same solves."""
    if name is None:
        cands = [(t["total_ms"], nm) for nm, t in times.items() if nm in KERNELS]
        if not cands:
            return None
        _, name = max(cands)
    if name not in times:
        return None
    t = times[name]
    t_s = t["total_ms"] * 1e-3
    roof = {"kernel": name, "bound": None, "achieved": None, "peak": None, "unit": None, "frac": None,
            "traffic": None, "avg_launch_ms": t["avg_ms"], "launches": t["launches"]}
    if name in MARCH:
        skey, bkey, io = MARCH[name]
        roof.update(march_roofline(work.get(skey, 0), work.get(bkey, 0), io(work), t_s))
        roof["per_launch"] = {"samples": round(work.get(skey, 0) / t["launches"], 1),
                              "bricks": round(work.get(bkey, 0) / t["launches"], 1),
                              "compulsory_bytes": round(roof["compulsory_bytes"] / t["launches"], 1)}
    elif name in FW:
        ops = 2.0 * FW[name](work)
        achieved = ops / t_s / 1e12
        roof.update({"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 5), "work_unit": "executed min-plus updates x 2 ops",
                     "issue_ceiling_tflops": round(FW_OPS_CEILING_TFLOPS, 2),
                     "frac_of_issue_ceiling": round(achieved / FW_OPS_CEILING_TFLOPS, 4),
                     "ops_per_launch": round(ops / t["launches"], 1)})
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f).get(workload, {}).get(name)
        if tr:
            roof["traffic"] = tr["hbm_bytes_per_launch"]
            roof["traffic_gbs"] = round(tr["hbm_bytes_per_launch"] / (t["avg_ms"] * 1e-3) / 1e9, 1)
            roof["traffic_source"] = tr.get("source")
    return roof


def bench_dart(lga, ctx, grid, iters, scatter, cpu):
    """The reference's sampler (dart throwing, LightGraph::scatterPoints) on the
    C2 volume: `iters` iterations of `scatter` throws each in one lg_dart_throw
    call, timed with the context's markers; with `cpu`, one iteration of the
    oracle's sequential restatement on the host for comparison."""
    import numpy as np
    lga.dart_throw(ctx, grid, [0], scatter_count=scatter, target=750)  # warm-up
    ctx.sync()
    ctx.mark(6)
    pts, md = lga.dart_throw(ctx, grid, range(iters), scatter_count=scatter, target=750)
    ctx.mark(7)
    ctx.sync()
    ms = ctx.elapsed_ms(6, 7)
    out = {"workload": "c2_dart_throw", "iterations": iters, "throws_per_iteration": scatter, "target": 750,
           "ms": round(ms, 3), "iterations_per_s": round(iters * 1000.0 / ms, 2),
           "points_per_iteration": [int(np.min([len(p) for p in pts])), int(np.max([len(p) for p in pts]))]}
    if cpu:
        from oracle import lgoracle
        dense = grid.read_dense()
        t0 = time.perf_counter()
        want, _ = lgoracle.dart_throw(dense, 0, scatter_count=scatter, target=750,
                                      min_dist=float(_dart_start(lga, ctx, grid)))
        dt = time.perf_counter() - t0
        out["cpu_oracle_s_per_iteration"] = round(dt, 3)
        out["cpu_equal_to_gpu_iteration_0"] = bool(np.array_equal(want, pts[0]))
    return out


def _dart_start(lga, ctx, grid):
    """MIN_DIST_MULT x mean side of the active-voxel box (main.cc:512-518)."""
    import numpy as np
    lo, hi = lga.active_bbox(ctx, grid)
    desc = grid.info().desc
    size = [np.float32(desc.voxel_size * (hi[a] - lo[a])) for a in range(3)]
    return np.float32(0.13) * (np.float32(np.float32(size[0] + size[1]) + size[2]) / np.float32(3.0))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(lga, ctx, grid, shader, params, lights, n_points, radius, rounds=3):
    """The oracle's serial solve (max-product Floyd-Warshall on Vec3 transmittance,
    FP64, built -O3 like the reference) on the same volume and points, one solve
    per host thread, as the reference parallelises across solves
    (main.cc:534-538).  Threads: the host CPUs this process may run on, capped
    at 16, the host share the GPU box gives one GPU.  Median of `rounds` rounds
    after one warm-up round."""
    from oracle import lgoracle
    dense = grid.read_dense()
    visible = os.cpu_count() or 1
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else visible
    cores = max(1, min(16, allowed))
    solver = lga.Solver(ctx, n_points)
    pts = []
    for s in range(cores):
        solver.solve(grid, shader, lights, params, n=n_points, radius=radius, seed=854123, solve=1000 + s)
        pts.append(solver.results(want_D=False)["points"])
    solver.destroy()

    def one(k):
        lgoracle.solve_points(dense, shader, params, lights, pts[k], radius=radius, cap=12)

    walls = []
    for r in range(rounds + 1):
        t0 = time.perf_counter()
        th = [threading.Thread(target=one, args=(k,)) for k in range(cores)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if r > 0:
            walls.append(time.perf_counter() - t0)
    walls.sort()
    dt = walls[len(walls) // 2]
    return {"value": round(cores / dt, 4), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"{cores} concurrent full oracle solves (N={n_points}, same volume, points and lights) per "
                      f"round, median of {rounds} rounds after a warm-up: {dt:.2f} s wall",
            "cpu_model": _cpu_model(), "host_cpus_visible": visible, "host_cpus_allowed": allowed,
            "build": "g++ -O3 -ffp-contract=off, FP64"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=8192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-volumes", type=int, default=64)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--latency-solves", type=int, default=50)
    ap.add_argument("--render-reps", type=int, default=3, help="0: skip the 1080p render timing")
    ap.add_argument("--no-precision", action="store_true", help="skip the FP32/FP16/BF16 brick timing")
    ap.add_argument("--serve-jobs", type=int, default=40, help="jobs per arrival rate; 0: skip serving")
    ap.add_argument("--serve-rates", default="2,10,50", help="Poisson arrival rates, jobs/s")
    ap.add_argument("--serve-workers", type=int, default=3)
    ap.add_argument("--serve-batch", type=int, default=8)
    ap.add_argument("--no-fw-dense", action="store_true", help="skip the dense n x n Floyd-Warshall timing")
    ap.add_argument("--iters", default="25,100,500",
                    help="solves per lg_solve_batch of the production volume (the reference's -iter); empty: skip")
    ap.add_argument("--dart-iters", type=int, default=64,
                    help="iterations of the reference's dart thrower (1e6 throws each) to time; 0: skip")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device three times with k_copy "
                         "(the known-size kernel that traffic counters are scaled against)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        return 2

    import lightgraph_amd as lga
    ctx = lga.Context(0)
    # the Floyd-Warshall tile choices are measured once per device, on the first
    # batched closure or this query; taking them here keeps the measurement out
    # of every timed region, whatever the warm-up count
    lga.fw_tuning(ctx)
    spec, grid, shader, params, lights = setup(lga, ctx, args.points)
    n = args.points
    solver = lga.Solver(ctx, n)
    radius = 12.0
    for k in range(args.warmup):
        solver.solve(grid, shader, lights, params, n=n, radius=radius, seed=854123, solve=k)
    ctx.sync()
    # the timed solves once untimed with every kernel and stage timed: the
    # per-kernel breakdown, and the dominant kernel for the roofline
    ctx.kernel_time_reset()
    ctx.kernel_timing(True)
    for k in range(args.steps):
        solver.solve(grid, shader, lights, params, n=n, radius=radius, seed=854123, solve=args.warmup + k)
    ctx.sync()
    ctx.kernel_timing(False)
    times = kernel_times(ctx, KERNELS + STAGES)
    dominant = max(((t["total_ms"], nm) for nm, t in times.items() if nm in KERNELS), default=(0, None))[1]
    # timed region: HIP events around the dominant kernel's launches only
    ctx.kernel_time_reset()
    ctx.kernel_timing(True, only=dominant)
    ctx.sync()
    ctx.mark(0)
    t0 = time.perf_counter()
    for k in range(args.steps):
        solver.solve(grid, shader, lights, params, n=n, radius=radius, seed=854123, solve=args.warmup + k)
    ctx.mark(1)
    ctx.sync()
    wall = time.perf_counter() - t0
    gpu_ms = ctx.elapsed_ms(0, 1)
    ctx.kernel_timing(False)
    ms_per_step = gpu_ms / args.steps
    timed = kernel_times(ctx, [dominant]) if dominant else {}
    # the same solves again, untimed, with the device work counters on
    work = {}
    for k in range(args.steps):
        solver.solve(grid, shader, lights, params, n=n, radius=radius, seed=854123, solve=args.warmup + k,
                     stats=True)
        for key, v in solver.stats().items():
            work[key] = work.get(key, 0) + v
    roof = kernel_roofline(timed, work, WORKLOAD["workload"])
    if roof:
        # how dominant the dominant kernel is: its share of the kernel time of
        # the pass with every kernel timed (the two streams' kernels summed)
        total = sum(t["total_ms"] for nm, t in times.items() if nm in KERNELS)
        share = times[dominant]["total_ms"] / total if total > 0 else None
        roof["share_of_roofline_kernels_time"] = round(share, 4) if share is not None else None
        roof["note"] = ("largest of the solve's march and closure kernels by total time; the plain solve is a chain of "
                        "~46 small launches, see kernel_rooflines for every kernel's fraction")
    # every roofline kernel of the solve, from the pass with every kernel timed
    # (its per-kernel events add a little to each launch)
    kroofs = {}
    for nm in KERNELS:
        r = kernel_roofline(times, work, WORKLOAD["workload"], name=nm)
        if r and r.get("frac") is not None:
            kroofs[nm] = {k: r.get(k) for k in ("bound", "frac", "achieved", "unit", "avg_launch_ms", "launches",
                                                "traffic", "traffic_gbs") if k in r}
    b = solver.buffers()
    lat = []
    for k in range(args.latency_solves):
        t1 = time.perf_counter()
        solver.solve(grid, shader, lights, params, n=n, radius=radius, seed=854123, solve=10000 + k)
        ctx.sync()
        lat.append((time.perf_counter() - t1) * 1e3)
    lat.sort()

    def pct(q):
        return round(lat[min(len(lat) - 1, int(q * len(lat)))], 4) if lat else None
    line = {"metric": "LightGraph solves/s", "value": round(1000.0 / ms_per_step, 3), "unit": "solves/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "wall_ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 (fp64 ray/DDA positions)", "data": "synthetic",
            "config": dict(WORKLOAD, edges_last_solve=b.n_edges), "roofline": roof,
            "kernels": {k: {"launches": v["launches"], "avg_ms": v["avg_ms"]} for k, v in times.items()},
            "kernel_rooflines": kroofs,
            "work_per_solve": {k: round(v / args.steps, 1) for k, v in work.items()},
            "latency": {"p50_ms": pct(0.50), "p99_ms": pct(0.99), "solves": len(lat), "mode": "one solve alone"}}
    if not args.no_fw_dense:
        line["fw_dense"] = bench_fw_dense(lga, ctx, solver, n)
    iters = [int(x) for x in args.iters.split(",") if x]
    if iters:
        line["iter_batch"] = bench_iter(lga, ctx, grid, shader, params, lights, n, radius, iters)
    if args.render_reps > 0:
        line["render"] = bench_render(lga, ctx, grid, shader, lights, solver, n, args.render_reps)
    solver.destroy()
    if args.dart_iters > 0:
        line["dart_throw"] = bench_dart(lga, ctx, grid, args.dart_iters, 1000000, not args.no_cpu_baseline)
    if args.batch_volumes > 0:
        line["batched"] = bench_batched(lga, ctx, shader, params, lights, args.batch_volumes, args.batch_reps)
    t = lga.fw_tuning(ctx)
    line["fw_tuning"] = {"split_below": t["split_below"], "use_graph": t["use_graph"], "tuned": t["tuned"],
                         "tiles": t["tiles"], "whole_us": [round(x, 2) for x in t["whole_us"]],
                         "split_us": [round(x, 2) for x in t["split_us"]], "stream_ms": round(t["stream_ms"], 4),
                         "graph_ms": round(t["graph_ms"], 4), "capture_ms": round(t["capture_ms"], 4)}
    if not args.no_precision:
        line["precision"] = bench_precision(lga, ctx)
    if args.serve_jobs > 0:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        line["serving"] = bench_serving(lga, ctx, rates, args.serve_jobs, args.serve_workers, args.serve_batch)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(lga, ctx, grid, shader, params, lights, n, radius)
    else:
        line["cpu_baseline"] = None
    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        a, b2 = ctx.alloc(nbytes), ctx.alloc(nbytes)
        lga.lib().lg_memset_d(ctx.handle, a.ptr, 1, nbytes)
        for _ in range(3):
            lga._check(lga.lib().lg_copy_d(ctx.handle, b2.ptr, a.ptr, nbytes), "copy")
        ctx.sync()
        a.free()
        b2.free()
        line["copy_calibration_bytes"] = nbytes
    grid.destroy()
    ctx.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

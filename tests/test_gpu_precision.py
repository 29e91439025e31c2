"""Precision form: density bricks stored as FP32, FP16 and BF16 (FP32/FP64
accumulation in every kernel), error of each against the FP64 oracle on the
FP32 volume, per quantity: edge optical depth, all-pairs optical depth,
vertex transmittance-weighted light, diffused radiance and pixels.

The workload is a smoke plume with a temperature grid (emission ramp of
scenes/explo.sh:1), an environment light from a synthetic HDR map (sky gradient
and a 5000x sun disk, importance-sampled through the FP64 CDF) and a directional
light.  All three runs share the points sampled on the FP32 volume, so the graph
is the same and only the brick precision differs.

Default size: the full configuration (1024^3, N = 12288, 1024 x 512 map,
160 x 90 image).  LG_PRECISION_SMALL=1 runs a reduced one (128^3, N = 2048,
128 x 64 map, 80 x 60 image); LG_PRECISION_JSON=<path> also writes the table
there."""
import json
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import dijkstra

pytestmark = pytest.mark.gpu

FULL = os.environ.get("LG_PRECISION_SMALL") != "1"
SIZE, NPTS, ENV_W, IMG = (1024, 12288, 1024, (160, 90)) if FULL else (128, 2048, 128, (80, 60))
RAMP = "0.0,0.33,0.66,1.0/0,0,0:1.0,0.076,0.0:1.0,0.322,0.0:1.0,0.45,0.15/0.35/1.5"

# error bars per brick type (quantisation: FP16 2^-11, BF16 2^-8 per density).
# FP32: the maximum relative error (stat 0).  Reduced types: optical depths on the
# maximum relative error; light, radiance and pixels on the largest error relative
# to the reference maximum (stat 3), because a transmittance exp(-sigma tau) carries
# a relative error sigma * tau * eps that grows without bound in deep shadow, and a
# quantised density that crosses the cutoff changes which samples a ray takes.
BARS = {"fp32": {"tau": (0, 1e-6), "D": (0, 1e-5), "L": (0, 1e-5), "diffuse": (0, 1e-4), "pixels": (0, 1e-4)},
        "fp16": {"tau": (0, 1e-3), "D": (0, 1e-3), "L": (3, 2e-3), "diffuse": (3, 2e-3), "pixels": (3, 5e-2)},
        "bf16": {"tau": (0, 8e-3), "D": (0, 8e-3), "L": (3, 2e-2), "diffuse": (3, 2e-2), "pixels": (3, 1e-1)}}


def _say(msg):
    if FULL:
        print(msg, flush=True)


def _rel(got, want, floor_frac=1e-6):
    """(max, 99.9th percentile, mean) relative error and the largest error
    relative to the reference's maximum."""
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    top = float(np.abs(want).max())
    floor = max(top * floor_frac, 1e-30)
    d = np.abs(got - want).ravel()
    r = d / np.maximum(np.abs(want).ravel(), floor)
    return float(r.max()), float(np.quantile(r, 0.999)), float(r.mean()), float(d.max() / max(top, 1e-30))


def _sky(h, w):
    y = np.linspace(0.0, 1.0, h, dtype=np.float32)[:, None, None]
    rgb = np.broadcast_to(np.concatenate([0.25 + 0.5 * y, 0.35 + 0.45 * y, 0.8 - 0.2 * y], -1), (h, w, 3)).copy()
    yy, xx = np.mgrid[0:h, 0:w]
    sun = (xx - 0.7 * w) ** 2 + (yy - 0.22 * h) ** 2 <= (0.012 * w) ** 2 + 1
    rgb[sun] *= 5000.0
    return rgb


@pytest.fixture(scope="module")
def setup(lga, ctx, oracle):
    spec = lga.scene_spec(lga.LG_SCENE_PLUME, SIZE, seed=4, octaves=6, with_temperature=1)
    grid32, temp = lga.Grid.generate(ctx, spec)
    dense = grid32.read_dense()
    tdense = temp.read_dense()
    sky = _sky(ENV_W // 2, ENV_W)
    env = lga.EnvMap(ctx, sky)
    sh = lga.default_shader(absorption=0.1, scattering=1.0, emission_scale=2.0, ramp=RAMP)
    p = lga.march_params(cutoff=0.005, primary_step=1.0, shadow_step=1.0)
    lights = [lga.env_light(env, 20.0, (1.0, 1.0, 1.0), 0.5, samples=2),
              lga.parse_light("dir:0:1,0.9,0.8:0.8:-40,30,0")]
    envs = [sky, None]
    radius = 6.0 * SIZE / 128.0 if not FULL else 12.0
    s = lga.Solver(ctx, NPTS)
    s.solve(grid32, sh, lights, p, n=NPTS, radius=radius, seed=854123, solve=0)
    pts = s.results(want_D=False)["points"]
    s.destroy()
    yield dict(spec=spec, grid32=grid32, temp=temp, dense=dense, tdense=tdense, env=env, sky=sky, sh=sh, p=p,
               lights=lights, envs=envs, radius=radius, pts=pts)
    env.destroy()
    temp.destroy()
    grid32.destroy()


def test_precision_errors_vs_fp64_oracle(lga, ctx, oracle, setup):
    S = setup
    n = len(S["pts"])
    table = {}
    ref = None
    for name, dtype in (("fp32", lga.LG_F32), ("fp16", lga.LG_F16), ("bf16", lga.LG_BF16)):
        spec = lga.scene_spec(lga.LG_SCENE_PLUME, SIZE, seed=4, octaves=6, with_temperature=1, dtype=dtype)
        _say(f"precision: {name} solve")
        grid, temp = lga.Grid.generate(ctx, spec)
        temp.destroy()
        s = lga.Solver(ctx, n, dense=True)
        s.write_points(S["pts"])
        s.solve(grid, S["sh"], S["lights"], S["p"], n=n, radius=S["radius"], seed=854123, solve=0,
                temperature=S["temp"], external_points=True)
        r = s.results()
        s.destroy()
        if ref is None:
            # FP64 oracle on the FP32 volume, same points and edges
            e = r["edges"]
            _say(f"precision: oracle edges ({len(e)})")
            T_o, tau_o = oracle.march_edges(S["dense"], S["sh"], S["p"], S["pts"], e)
            tau_o = np.where(T_o[:, 0] > 0, tau_o, np.inf)
            _say("precision: oracle vertex light")
            L_o = oracle.vertex_light(S["dense"], S["sh"], S["p"], S["lights"], S["pts"], solve=0,
                                      temp_dense=S["tdense"], envs=S["envs"])
            _say("precision: oracle all-pairs")
            fin = np.isfinite(tau_o)
            A = csr_matrix((tau_o[fin], (e[fin, 0], e[fin, 1])), shape=(n, n))
            D_o = dijkstra(A, directed=False)
            sig = -np.array(S["sh"].extinction)
            P_o = np.exp(-D_o[..., None] * sig[None, None, :])
            dif_o = oracle.diffuse(P_o, L_o)
            del P_o
            ref = dict(edges=e, tau=tau_o, L=L_o, D=D_o, diffuse=dif_o)
        assert np.array_equal(r["edges"], ref["edges"])
        both = np.isfinite(ref["tau"]) & np.isfinite(r["weights"])
        row = {"tau": _rel(r["weights"][both], ref["tau"][both]),
               "edges_dropped_differently": int(np.sum(np.isfinite(ref["tau"]) != np.isfinite(r["weights"])))}
        Dfin = np.isfinite(ref["D"]) & np.isfinite(r["D"])
        row["D"] = _rel(r["D"][Dfin], ref["D"][Dfin])
        row["L"] = _rel(r["L"], ref["L"])
        row["diffuse"] = _rel(r["diffuse"], ref["diffuse"])
        # pixels: single scattering + emission (no scatter grid), env + directional light
        w, h = IMG
        cam = lga.camera_setup(cam_translate=(SIZE / 2, 0.45 * SIZE, 2.4 * SIZE), focal=50.0, width=w, height=h)
        films = lga.Films(ctx, w, h)
        lga.render(ctx, grid, None, S["temp"], cam, S["sh"], S["lights"], S["p"], 1, films, seed=1)
        got = films.read()
        films.free()
        if "pix" not in ref:
            _say("precision: oracle render")
            ref["pix"] = oracle.render(S["dense"], cam, S["sh"], S["lights"], S["p"], 1, seed=1,
                                       temp_dense=S["tdense"], envs=S["envs"], threads=16)
        errs = [_rel(got[k][..., :3], ref["pix"][k][..., :3]) for k in (0, 1, 3)]
        row["pixels"] = tuple(max(e[i] for e in errs) for i in range(4))
        row["alpha"] = _rel(got[0][..., 3], ref["pix"][0][..., 3])
        info = grid.info()
        row["grid_bytes"] = int(info.device_bytes)
        grid.destroy()
        table[name] = row
    out = {"config": {"volume": f"plume {SIZE}^3", "points": n, "env_map": f"{ENV_W}x{ENV_W // 2}",
                      "image": f"{IMG[0]}x{IMG[1]}", "radius": S["radius"]},
           "errors": table,
           "note": "(max, 99.9th percentile, mean) relative error (floor 1e-6 of the reference maximum), "
                   "largest error / reference maximum"}
    print(json.dumps(out))
    path = os.environ.get("LG_PRECISION_JSON")
    if path:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
    assert ref["pix"][0][..., 3].max() > 0.3 and ref["pix"][3][..., :3].max() > 0
    for name, bars in BARS.items():
        for q, (stat, bar) in bars.items():
            assert table[name][q][stat] < bar, f"{name} {q}: {table[name][q]} (bar {bar} on stat {stat})"

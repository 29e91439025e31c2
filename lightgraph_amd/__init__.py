"""Python access to the MI355X LightGraph library (liblightgraph.so).

The compute path is the HIP/C++ library behind the C ABI of
include/lightgraph.h; this module only marshals numpy arrays and parameter
blocks for tests, the benchmark and the smoke check.  Nothing here falls back to
a CPU implementation: if the library or a GPU is missing, calls raise.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import (LG_F32, LG_F16, LG_BF16, LG_LIGHT_DIRECTIONAL, LG_LIGHT_POINT, LG_LIGHT_ENV,
                   LG_SCENE_SPHERE_FBM, LG_SCENE_BLOBS_FBM, LG_SCENE_PLUME, lg_shader, lg_light,
                   lg_march_params, lg_camera, lg_scene_spec, lg_grid_desc, lg_films, lg_scatter_grid,
                   lg_solve_params)

_LIB = None


class LGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[lightgraph {code}] {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _abi.load()
    return _LIB


def _check(rc, what=""):
    if rc != 0:
        msg = lib().lg_last_error()
        raise LGError(rc, f"{what}: {msg.decode() if msg else ''}")


def _hcheck(rc, what=""):
    if rc != 0:
        msg = lib().lg_host_last_error()
        raise LGError(rc, f"{what}: {msg.decode() if msg else ''}")


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def _vec3(v):
    if np.isscalar(v):
        return [float(v)] * 3
    v = list(v)
    return [float(x) for x in v] if len(v) == 3 else [float(v[0])] * 3


# ------------------------------------------------------------ parameters ----
def default_shader(absorption=None, scattering=None, volume_color=None, **fields):
    """Shader::Shader defaults (Shader.h:123-138) plus overrides."""
    s = lg_shader()
    _hcheck(lib().lg_shader_default(C.byref(s)), "lg_shader_default")
    if absorption is not None or scattering is not None or volume_color is not None:
        a = _vec3(absorption if absorption is not None else 0.1)
        sc = _vec3(scattering if scattering is not None else 1.5)
        vc = _vec3(volume_color if volume_color is not None else 1.0)
        _hcheck(lib().lg_shader_set_coefficients(C.byref(s), _d3(a), _d3(sc), _d3(vc)), "coefficients")
    ramp = fields.pop("ramp", None)
    if ramp is not None:
        _hcheck(lib().lg_shader_parse_ramp(C.byref(s), ramp.encode()), "ramp")
    for k, v in fields.items():
        setattr(s, k, v)
    return s


def march_params(cutoff=0.005, primary_step=1.0, shadow_step=3.0):
    p = lg_march_params()
    p.cutoff = cutoff
    p.step[0] = primary_step
    p.step[1] = shadow_step
    return p


def directional_light(y_rot=0.0, color=(1, 1, 1), intensity=1.0, rot=(0, 0, 0)):
    l = lg_light()
    _hcheck(lib().lg_light_directional(y_rot, _d3(_vec3(color)), intensity, _d3(_vec3(rot)), C.byref(l)), "dir")
    return l


def point_light(y_rot=0.0, color=(1, 1, 1), intensity=1.0, pos=(0, 0, 0)):
    l = lg_light()
    _hcheck(lib().lg_light_point(y_rot, _d3(_vec3(color)), intensity, _d3(_vec3(pos)), C.byref(l)), "point")
    return l


def env_light(envmap, y_rot=0.0, color=(1, 1, 1), intensity=1.0, samples=3):
    l = lg_light()
    _hcheck(lib().lg_light_env(y_rot, _d3(_vec3(color)), intensity, samples, envmap.handle, C.byref(l)), "env")
    return l


def parse_light(text):
    l = lg_light()
    _hcheck(lib().lg_light_parse(text.encode(), C.byref(l)), f"light '{text}'")
    return l


def light_array(lights):
    arr = (lg_light * max(1, len(lights)))()
    for k, l in enumerate(lights):
        arr[k] = l
    return arr


def scene_spec(kind, n, seed, octaves=4, radius=0.0, blobs=0, axis_min=30.0, axis_max=120.0,
               active_fraction=0.0, dtype=LG_F32, with_temperature=0):
    s = lg_scene_spec()
    s.kind, s.n, s.seed, s.octaves = kind, n, seed, octaves
    s.radius, s.blobs, s.axis_min, s.axis_max = radius, blobs, axis_min, axis_max
    s.active_fraction, s.dtype, s.with_temperature = active_fraction, dtype, with_temperature
    return s


def generate_host(spec, threads=0):
    """Host twin of the procedural generator: dense float32 volume [z, y, x]."""
    out = np.zeros((spec.n, spec.n, spec.n), dtype=np.float32)
    _check(lib().lg_scene_generate_host(C.byref(spec), out.ctypes.data, threads), "generate_host")
    return out


def fw_ld(n):
    return lib().lg_fw_ld(n)


# --------------------------------------------------------------- device -----
class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().lg_malloc(ctx.handle, self.nbytes, C.byref(p)), "lg_malloc")
        self.ptr = p.value

    def free(self):
        if self.ptr:
            _check(lib().lg_free(self.ctx.handle, self.ptr), "lg_free")
            self.ptr = None

    def write(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes, "write exceeds buffer"
        _check(lib().lg_memcpy_h2d(self.ctx.handle, self.ptr, arr.ctypes.data, arr.nbytes), "h2d")

    def read(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes, "read exceeds buffer"
        _check(lib().lg_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out


class Context:
    """One HIP device + stream (lg_ctx)."""

    def __init__(self, device=0, hbm_budget=0):
        h = C.c_void_p()
        _check(lib().lg_ctx_create(device, hbm_budget, C.byref(h)), "lg_ctx_create")
        self.handle = h.value

    def close(self):
        if self.handle:
            lib().lg_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _check(lib().lg_sync(self.handle), "lg_sync")

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, max(arr.nbytes, 4))
        b.write(arr)
        return b

    def bytes_in_use(self):
        v = C.c_uint64()
        _check(lib().lg_ctx_bytes_in_use(self.handle, C.byref(v)), "bytes_in_use")
        return v.value

    def kernel_timing(self, enable=True, only=None):
        """Per-kernel HIP-event timing; only: time the launches of that kernel alone."""
        _check(lib().lg_ctx_set_kernel_timing_only(self.handle, only.encode() if only else None),
               "kernel_timing_only")
        _check(lib().lg_ctx_set_kernel_timing(self.handle, 1 if enable else 0), "kernel_timing")

    def brick_counting(self, enable=True):
        """Renders under kernel timing also count the distinct bricks of each
        shadow launch (render_stats()["shadow_bricks"]); slows those renders."""
        _check(lib().lg_ctx_set_brick_counting(self.handle, 1 if enable else 0), "brick_counting")

    def kernel_time(self, name):
        n = C.c_uint64()
        ms = C.c_double()
        _check(lib().lg_kernel_time(self.handle, name.encode(), C.byref(n), C.byref(ms)), "kernel_time")
        return n.value, ms.value

    def kernel_time_reset(self):
        _check(lib().lg_kernel_time_reset(self.handle), "kernel_time_reset")

    def mark(self, slot):
        _check(lib().lg_timer_mark(self.handle, slot), "timer_mark")

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        _check(lib().lg_timer_elapsed(self.handle, a, b, C.byref(ms)), "timer_elapsed")
        return ms.value


class Grid:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle

    @staticmethod
    def from_dense(ctx, dense_zyx, voxel_size=1.0, origin=(0.0, 0.0, 0.0), dtype=LG_F32):
        a = np.ascontiguousarray(dense_zyx, dtype=np.float32)
        d = lg_grid_desc()
        d.nz, d.ny, d.nx = a.shape
        d.dtype = dtype
        d.voxel_size = voxel_size
        for c in range(3):
            d.origin[c] = origin[c]
        h = C.c_void_p()
        _check(lib().lg_grid_create(ctx.handle, C.byref(d), a.ctypes.data, C.byref(h)), "lg_grid_create")
        return Grid(ctx, h.value)

    @staticmethod
    def generate(ctx, spec):
        d = C.c_void_p()
        t = C.c_void_p()
        _check(lib().lg_grid_generate(ctx.handle, C.byref(spec), C.byref(d), C.byref(t)), "lg_grid_generate")
        return Grid(ctx, d.value), (Grid(ctx, t.value) if t.value else None)

    def info(self):
        i = _abi.lg_grid_info()
        _check(lib().lg_grid_info_get(self.handle, C.byref(i)), "grid_info")
        return i

    def read_dense(self):
        i = self.info()
        out = np.empty((i.desc.nz, i.desc.ny, i.desc.nx), dtype=np.float32)
        _check(lib().lg_grid_read_dense(self.ctx.handle, self.handle, out.ctypes.data), "read_dense")
        return out

    def destroy(self):
        if self.handle:
            lib().lg_grid_destroy(self.ctx.handle, self.handle)
            self.handle = None


class EnvMap:
    def __init__(self, ctx, rgb_hw3):
        self.ctx = ctx
        a = np.ascontiguousarray(rgb_hw3, dtype=np.float32)
        self.height, self.width = a.shape[:2]
        self.rgb = a
        h = C.c_void_p()
        _check(lib().lg_envmap_create(ctx.handle, self.width, self.height, a.ctypes.data, C.byref(h)), "envmap")
        self.handle = h.value

    def read_cdf(self):
        cond = np.empty((self.height, self.width + 1), dtype=np.float64)
        marg = np.empty(self.height + 1, dtype=np.float64)
        rowint = np.empty(self.height, dtype=np.float64)
        total = C.c_double()
        _check(lib().lg_envmap_read_cdf(self.ctx.handle, self.handle, cond.ctypes.data_as(_abi.DP),
                                        marg.ctypes.data_as(_abi.DP), rowint.ctypes.data_as(_abi.DP),
                                        C.byref(total)), "read_cdf")
        return cond, marg, rowint, total.value

    def destroy(self):
        if self.handle:
            lib().lg_envmap_destroy(self.ctx.handle, self.handle)
            self.handle = None


# ---------------------------------------------------------- solve stages ----
def sample_points(ctx, grid, seed, solve, n, out_buf):
    _check(lib().lg_sample_points(ctx.handle, grid.handle, seed, solve, n, out_buf.ptr), "sample_points")


def active_bbox(ctx, grid):
    """Inclusive voxel bbox of the voxels != 0 (evalActiveVoxelBoundingBox)."""
    lo, hi = (C.c_int32 * 3)(), (C.c_int32 * 3)()
    _check(lib().lg_grid_active_bbox(ctx.handle, grid.handle, lo, hi), "grid_active_bbox")
    return tuple(lo), tuple(hi)


def dart_throw(ctx, grid, iters, scatter_count=1000000, target=750, min_dist=None, bbox=None, out_cap=None):
    """LightGraph::scatterPoints (the reference's dart thrower) for each
    iteration index in `iters`.  bbox: (min, max) world corners, by default the
    active-voxel box of the grid (main.cc:512); min_dist: starting minimum
    distance per iteration (a float or a sequence), by default
    MIN_DIST_MULT (0.13) * mean bbox side in float arithmetic (main.cc:514-518).
    Returns (list of
    (count, 3) float32 point arrays, array of final minimum distances)."""
    iters = [int(i) for i in iters]
    k = len(iters)
    if bbox is None:
        lo, hi = active_bbox(ctx, grid)
        desc = grid.info().desc
        o, v = list(desc.origin), desc.voxel_size
        bbox = ([o[a] + v * lo[a] for a in range(3)], [o[a] + v * hi[a] for a in range(3)])
    bmin, bmax = (C.c_double * 3)(*bbox[0]), (C.c_double * 3)(*bbox[1])
    if min_dist is None:
        size = np.array([np.float32(bbox[1][a] - bbox[0][a]) for a in range(3)], np.float32)
        avg = np.float32(np.float32(size[0] + size[1]) + size[2]) / np.float32(3.0)
        min_dist = np.float32(np.float32(0.13) * avg)
    md = np.broadcast_to(np.asarray(min_dist, np.float32), (k,)).astype(np.float32)
    cap = int(out_cap or target + 1)
    out = DeviceBuffer(ctx, k * cap * 3 * 4)
    counts = (C.c_int32 * k)()
    mdc = (C.c_float * k)(*md.tolist())
    _check(lib().lg_dart_throw(ctx.handle, grid.handle, bmin, bmax, k, (C.c_uint32 * k)(*iters), scatter_count,
                               target, mdc, out.ptr, cap, counts), "dart_throw")
    allp = out.read(np.float32, (k, cap, 3))
    out.free()
    pts = [allp[i, :min(counts[i], cap)].copy() for i in range(k)]
    return pts, np.array(mdc[:], np.float32)


def build_edges(ctx, pts_buf, n, radius, max_connect, edges_buf, edge_cap):
    ne = C.c_int32()
    _check(lib().lg_build_edges(ctx.handle, pts_buf.ptr, n, radius, max_connect, edges_buf.ptr, edge_cap,
                                C.byref(ne)), "build_edges")
    return ne.value


def march_edges(ctx, grid, shader, params, pts_buf, edges_buf, n_edges, w_buf=None, D_buf=None, ld=0):
    _check(lib().lg_march_edges(ctx.handle, grid.handle, C.byref(shader), C.byref(params), pts_buf.ptr,
                                edges_buf.ptr if edges_buf else None, n_edges, w_buf.ptr if w_buf else None,
                                D_buf.ptr if D_buf else None, ld), "march_edges")


def vertex_light(ctx, grid, temp_grid, shader, lights, params, env_seed, solve, pts_buf, n, out_buf):
    arr = light_array(lights)
    _check(lib().lg_vertex_light(ctx.handle, grid.handle, temp_grid.handle if temp_grid else None,
                                 C.byref(shader), arr, len(lights), C.byref(params), env_seed, solve,
                                 pts_buf.ptr, n, out_buf.ptr), "vertex_light")


def fw_init(ctx, n, D_buf, ld):
    _check(lib().lg_fw_init(ctx.handle, n, D_buf.ptr, ld), "fw_init")


def fw_solve(ctx, n, D_buf, ld, pred_buf=None, edges_buf=None, w_buf=None, n_edges=0):
    _check(lib().lg_fw_solve(ctx.handle, n, D_buf.ptr, ld, pred_buf.ptr if pred_buf else None,
                             edges_buf.ptr if edges_buf else None, w_buf.ptr if w_buf else None, n_edges),
           "fw_solve")


def fw_solve_batched(ctx, ns, D_bufs):
    k = len(ns)
    n_arr = (C.c_int32 * k)(*ns)
    ld_arr = (C.c_int32 * k)(*[fw_ld(n) for n in ns])
    p_arr = (C.c_void_p * k)(*[b.ptr for b in D_bufs])
    _check(lib().lg_fw_solve_batched(ctx.handle, k, n_arr, p_arr, ld_arr), "fw_solve_batched")


def _tuning_dict(t):
    k = t.points
    return {"split_below": t.split_below, "use_graph": t.use_graph, "tuned": t.tuned,
            "tiles": list(t.tiles[:k]), "whole_us": list(t.whole_us[:k]), "split_us": list(t.split_us[:k]),
            "stream_ms": t.stream_ms, "graph_ms": t.graph_ms, "capture_ms": t.capture_ms}


def fw_autotune(ctx):
    """Measure the batched Floyd-Warshall choices on this device now (lg_fw_autotune)."""
    t = _abi.lg_fw_tuning()
    _check(lib().lg_fw_autotune(ctx.handle, C.byref(t)), "fw_autotune")
    return _tuning_dict(t)


def fw_tuning(ctx):
    t = _abi.lg_fw_tuning()
    _check(lib().lg_fw_get_tuning(ctx.handle, C.byref(t)), "fw_get_tuning")
    return _tuning_dict(t)


def set_fw_tuning(ctx, split_below, use_graph):
    t = _abi.lg_fw_tuning()
    _check(lib().lg_fw_get_tuning(ctx.handle, C.byref(t)), "fw_get_tuning")
    t.split_below, t.use_graph = int(split_below), int(use_graph)
    _check(lib().lg_fw_set_tuning(ctx.handle, C.byref(t)), "fw_set_tuning")


def diffuse(ctx, n, D_buf, ld, sigma_t, L_buf, out_buf):
    sig = (C.c_double * 3)(*sigma_t)
    _check(lib().lg_diffuse(ctx.handle, n, D_buf.ptr, ld, sig, L_buf.ptr, out_buf.ptr), "diffuse")


LG_SOLVE_DENSE, LG_SOLVE_STATS, LG_SOLVE_BFS_ORDER, LG_SOLVE_ND_ORDER = 1, 2, 4, 8
ORDERS = {"reference": 0, "bfs": LG_SOLVE_BFS_ORDER, "nd": LG_SOLVE_ND_ORDER}


def _order_flag(order):
    if order not in ORDERS:
        raise ValueError(f"order must be one of {sorted(ORDERS)}, not {order!r}")
    return ORDERS[order]


class Solver:
    """Per-solve workspace (lg_solver): LightGraph::operator() for one index."""

    def __init__(self, ctx, n_max, dense=False, with_pred=False):
        self.ctx = ctx
        h = C.c_void_p()
        flags = (1 if dense else 0) | (2 if with_pred else 0)
        _check(lib().lg_solver_create(ctx.handle, n_max, 0, flags, C.byref(h)), "solver_create")
        self.handle = h.value
        self.dense = dense or with_pred
        self.last_dense = False

    def solve(self, density, shader, lights, params, n, radius, seed=854123, solve=0, max_connect=12,
              temperature=None, env_seed=4564446, external_points=False, timing=False, dense=None, stats=False,
              order="reference"):
        """order: "reference" closes each component in ascending vertex order (the
        reference's pivot order); "bfs" and "nd" close the components of more
        than 128 vertices in breadth-first (LG_SOLVE_BFS_ORDER) or
        nested-dissection (LG_SOLVE_ND_ORDER) order."""
        sp = lg_solve_params()
        sp.n, sp.max_connect, sp.radius, sp.seed, sp.env_seed = n, max_connect, radius, seed, env_seed
        sp.solve, sp.external_points, sp.timing = solve, 1 if external_points else 0, 1 if timing else 0
        dense = self.dense if dense is None else dense
        sp.flags = (LG_SOLVE_DENSE if dense else 0) | (LG_SOLVE_STATS if stats else 0) | _order_flag(order)
        self.last_dense = dense
        arr = light_array(lights)
        _check(lib().lg_solve(self.ctx.handle, self.handle, density.handle,
                              temperature.handle if temperature else None, C.byref(shader), arr, len(lights),
                              C.byref(params), C.byref(sp)), "lg_solve")

    def solve_batch(self, items, shader, lights, params, radius, seed=854123, max_connect=12, env_seed=4564446,
                    timing=False, stats=False, order="reference"):
        """items: sequence of (density Grid, n, solve index[, temperature Grid]).
        Item k owns vertices [voff_k, voff_k + n_k) of the buffers.  order as in
        solve."""
        arr = (_abi.lg_batch_item * len(items))()
        self.batch_n = []
        for k, it in enumerate(items):
            g, n, solve = it[0], it[1], it[2]
            t = it[3] if len(it) > 3 else None
            arr[k].density, arr[k].temperature = g.handle, (t.handle if t else None)
            arr[k].n, arr[k].solve = n, solve
            self.batch_n.append(n)
        sp = lg_solve_params()
        sp.max_connect, sp.radius, sp.seed, sp.env_seed = max_connect, radius, seed, env_seed
        sp.timing = 1 if timing else 0
        sp.flags = (LG_SOLVE_STATS if stats else 0) | _order_flag(order)
        self.last_dense = False
        la = light_array(lights)
        _check(lib().lg_solve_batch(self.ctx.handle, self.handle, arr, len(items), C.byref(shader), la, len(lights),
                                    C.byref(params), C.byref(sp)), "lg_solve_batch")

    def batch_results(self):
        """Per-item dicts (points, edges with item-local ids, weights, L, diffuse)."""
        r = self.results(want_D=False)
        out, voff = [], 0
        e = r["edges"]
        for n in self.batch_n:
            sel = (e[:, 0] >= voff) & (e[:, 0] < voff + n)
            out.append({"points": r["points"][voff:voff + n], "L": r["L"][voff:voff + n],
                        "diffuse": r["diffuse"][voff:voff + n], "edges": e[sel] - voff,
                        "weights": r["weights"][sel]})
            voff += n
        return out

    def buffers(self):
        b = _abi.lg_solver_buffers()
        _check(lib().lg_solver_get(self.handle, C.byref(b)), "solver_get")
        return b

    def stats(self):
        """Work counters of the last call made with stats=True (lg_solver_stats)."""
        st = _abi.lg_solve_stats()
        _check(lib().lg_solver_stats(self.ctx.handle, self.handle, C.byref(st)), "solver_stats")
        return {k: int(getattr(st, k)) for k, _ in st._fields_}

    def timing(self):
        t = _abi.lg_solve_timing()
        _check(lib().lg_solver_timing(self.handle, C.byref(t)), "solver_timing")
        return {k: getattr(t, k) for k, _ in t._fields_}

    def _read(self, ptr, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        _check(lib().lg_memcpy_d2h(self.ctx.handle, out.ctypes.data, ptr, out.nbytes), "d2h")
        return out

    def write_points(self, pts):
        b = self.buffers()
        a = np.ascontiguousarray(pts, dtype=np.float32)
        _check(lib().lg_memcpy_h2d(self.ctx.handle, b.d_points, a.ctypes.data, a.nbytes), "h2d")

    def results(self, want_D=True):
        b = self.buffers()
        r = {"points": self._read(b.d_points, np.float32, (b.n, 3)),
             "edges": self._read(b.d_edges, np.int32, (b.n_edges, 2)) if b.n_edges else np.zeros((0, 2), np.int32),
             "weights": self._read(b.d_weights, np.float32, (b.n_edges,)) if b.n_edges else np.zeros(0, np.float32),
             "L": self._read(b.d_L, np.float32, (b.n, 3)),
             "diffuse": self._read(b.d_diffuse, np.float32, (b.n, 3))}
        r["n_components"] = b.n_components
        if want_D and self.last_dense and b.d_D:
            r["D"] = self._read(b.d_D, np.float32, (b.n, b.ld))[:, : b.n]
        if self.last_dense and b.d_pred:
            r["pred"] = self._read(b.d_pred, np.int32, (b.n, b.ld))[:, : b.n]
        return r

    def destroy(self):
        if self.handle:
            lib().lg_solver_destroy(self.ctx.handle, self.handle)
            self.handle = None


def camera_setup(cam_rotate=(0, 0, 0), cam_translate=(0, 0, 0), cam_y_rotate=0.0, focal=50.0, aperture=36.0,
                 znear=1e-3, zfar=3.4028234663852886e38, width=960, height=540):
    cam = lg_camera()
    _check(lib().lg_camera_setup(_d3(cam_rotate), _d3(cam_translate), cam_y_rotate, focal, aperture, znear, zfar,
                                 width, height, C.byref(cam)), "camera_setup")
    return cam


# ---------------------------------------------------------- splat / render ---
def splat_radius(points, kernel_total):
    """(radius, point-grid voxel size) of the splat kernel (ScatterGrid.h:71-73)."""
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    r = C.c_double()
    v = C.c_double()
    _check(lib().lg_splat_radius(p.ctypes.data, len(p), int(kernel_total), C.byref(r), C.byref(v)), "splat_radius")
    return r.value, v.value


class ScatterGrid:
    """Coarse Vec3f multiple-scattering grid (ScatterGrid.h)."""

    def __init__(self, ctx, density, res_factor=4):
        self.ctx = ctx
        self.sg = lg_scatter_grid()
        _check(lib().lg_scatter_grid_create(ctx.handle, density.handle, res_factor, C.byref(self.sg)),
               "scatter_grid_create")

    @property
    def shape(self):
        return (self.sg.nz, self.sg.ny, self.sg.nx)

    def splat(self, pts_buf, rad_buf, count, radius):
        _check(lib().lg_splat(self.ctx.handle, pts_buf.ptr if pts_buf else None, rad_buf.ptr if rad_buf else None,
                              int(count), float(radius), C.byref(self.sg)), "splat")

    def read(self):
        nz, ny, nx = self.shape
        rgb = np.empty((nz, ny, nx, 3), np.float32)
        act = np.empty((nz, ny, nx), np.uint8)
        _check(lib().lg_memcpy_d2h(self.ctx.handle, rgb.ctypes.data, self.sg.d_rgb, rgb.nbytes), "d2h")
        _check(lib().lg_memcpy_d2h(self.ctx.handle, act.ctypes.data, self.sg.d_active, act.nbytes), "d2h")
        return rgb, act

    def write_rgb(self, rgb):
        a = np.ascontiguousarray(rgb, dtype=np.float32)
        _check(lib().lg_memcpy_h2d(self.ctx.handle, self.sg.d_rgb, a.ctypes.data, a.nbytes), "h2d")

    def destroy(self):
        if self.sg.d_rgb:
            lib().lg_scatter_grid_destroy(self.ctx.handle, C.byref(self.sg))


class Films:
    """The four RGBA float32 AOV films (Renderer.h:68-121) on the device."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, width, height
        self.bufs = [ctx.alloc(width * height * 16) for _ in range(4)]
        self.s = lg_films(*[b.ptr for b in self.bufs])

    def read(self):
        return np.stack([b.read(np.float32, (self.height, self.width, 4)) for b in self.bufs])

    def free(self):
        for b in self.bufs:
            b.free()


def render(ctx, density, scatter, temperature, camera, shader, lights, params, samples, films, seed=1,
           env_seed=4564446):
    arr = light_array(lights)
    _check(lib().lg_render(ctx.handle, density.handle, C.byref(scatter.sg) if scatter else None,
                           temperature.handle if temperature else None, C.byref(camera), C.byref(shader), arr,
                           len(lights), C.byref(params), samples, seed, env_seed, C.byref(films.s)), "render")


def render_stats(ctx):
    """Work counters of the last render on ctx (counted while kernel timing is on)."""
    st = _abi.lg_render_stats()
    _check(lib().lg_render_get_stats(ctx.handle, C.byref(st)), "render_get_stats")
    return {k: int(getattr(st, k)) for k, _ in st._fields_}


def render_workspace(width, height, samples, n_lights, max_points_per_ray=0):
    """Device scratch bytes a render of this size keeps on its context."""
    b = C.c_uint64()
    _check(lib().lg_render_workspace(width, height, samples, n_lights, max_points_per_ray, C.byref(b)),
           "render_workspace")
    return int(b.value)


# ------------------------------------------------------------- serving -------
def job(scene, points, camera, shader, lights, march, iterations=1, radius=12.0, max_connect=12, samples=1,
        seed=854123, env_seed=4564446, render_seed=1, solve_base=0, pts_per_kernel=3, scatter_res_factor=4,
        order="reference"):
    """An lg_job: volume -> `iterations` solves -> splat -> render.  order: the
    pivot order of the solves ("reference" or "bfs", as in Solver.solve)."""
    j = _abi.lg_job()
    j.solve_flags = _order_flag(order)
    j.scene, j.points, j.iterations, j.radius, j.max_connect = scene, points, iterations, radius, max_connect
    j.samples, j.seed, j.env_seed, j.render_seed, j.solve_base = samples, seed, env_seed, render_seed, solve_base
    j.pts_per_kernel, j.scatter_res_factor = pts_per_kernel, scatter_res_factor
    j.shader, j.march, j.camera = shader, march, camera
    if len(lights) > len(j.lights):
        raise LGError(-1, f"at most {len(j.lights)} lights")
    j.n_lights = len(lights)
    for k, l in enumerate(lights):
        j.lights[k] = l
    return j


class Server:
    """lg_server: continuous-batching request loop with HBM-budget admission."""

    def __init__(self, budget_bytes, workers=2, max_batch=8, device=0):
        h = C.c_void_p()
        _check(lib().lg_server_create(device, int(budget_bytes), workers, max_batch, C.byref(h)), "server_create")
        self.handle = h.value
        self._films = {}

    def submit(self, j):
        w, h = j.camera.width, j.camera.height
        films = np.zeros((4, h, w, 4), np.float32)
        ptrs = (C.c_void_p * 4)(*[films[k].ctypes.data for k in range(4)])
        jid = C.c_uint64()
        _check(lib().lg_server_submit(self.handle, C.byref(j), ptrs, C.byref(jid)), "server_submit")
        self._films[jid.value] = films
        return jid.value

    def wait(self, jid):
        r = _abi.lg_job_result()
        _check(lib().lg_server_wait(self.handle, jid, C.byref(r)), "server_wait")
        res = {k: getattr(r, k) for k, _ in r._fields_}
        res["films"] = self._films.pop(jid)
        return res

    def stats(self):
        st = _abi.lg_server_stats()
        _check(lib().lg_server_get_stats(self.handle, C.byref(st)), "server_stats")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.handle:
            lib().lg_server_destroy(self.handle)
            self.handle = None

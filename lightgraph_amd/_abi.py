"""ctypes declarations of include/lightgraph.h (the library's C ABI)."""
import ctypes as C
import os

LG_OK = 0
LG_ERR_INVALID = -1
LG_ERR_NOMEM = -2
LG_ERR_HIP = -3
LG_ERR_BUDGET = -4
LG_ERR_UNSUPPORTED = -5
LG_ERR_OVERFLOW = -6

LG_F32, LG_F16, LG_BF16 = 0, 1, 2
LG_LIGHT_DIRECTIONAL, LG_LIGHT_POINT, LG_LIGHT_ENV = 0, 1, 2
LG_SCENE_SPHERE_FBM, LG_SCENE_BLOBS_FBM, LG_SCENE_PLUME = 0, 1, 2
LG_MAX_RAMP = 8
LG_MAX_LIGHTS = 8

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblightgraph.so")


class lg_grid_desc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("dtype", C.c_int32),
                ("voxel_size", C.c_double), ("origin", C.c_double * 3)]


class lg_grid_info(C.Structure):
    _fields_ = [("desc", lg_grid_desc), ("bricks", C.c_int32 * 3), ("active_bricks", C.c_int32),
                ("bbox_min", C.c_int32 * 3), ("bbox_max", C.c_int32 * 3),
                ("active_voxels", C.c_uint64), ("device_bytes", C.c_uint64)]


class lg_shader(C.Structure):
    _fields_ = [("density_scale", C.c_double), ("density_max", C.c_double),
                ("shadow_density_scale", C.c_double), ("scatter_density_scale", C.c_double),
                ("scatter_density_min", C.c_double), ("albedo", C.c_double * 3),
                ("extinction", C.c_double * 3), ("scattering", C.c_double * 3),
                ("emission_scale", C.c_double), ("scatter_scale", C.c_double),
                ("scatter_density_mask_power", C.c_double), ("ramp_count", C.c_int32),
                ("ramp_pad_", C.c_int32), ("ramp_in", C.c_double), ("ramp_out", C.c_double),
                ("ramp_pos", C.c_double * LG_MAX_RAMP), ("ramp_color", (C.c_double * 3) * LG_MAX_RAMP)]


class lg_light(C.Structure):
    _fields_ = [("type", C.c_int32), ("samples", C.c_int32), ("intensity", C.c_float), ("pad_", C.c_float),
                ("color", C.c_double * 3), ("dir", C.c_double * 3), ("pos", C.c_double * 3),
                ("xform", C.c_double * 16), ("env", C.c_void_p)]


class lg_march_params(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("step", C.c_double * 2)]


class lg_camera(C.Structure):
    _fields_ = [("screen_to_world", C.c_double * 16), ("scale_width", C.c_double),
                ("scale_height", C.c_double), ("znear", C.c_double), ("zfar", C.c_double),
                ("width", C.c_int32), ("height", C.c_int32)]


class lg_films(C.Structure):
    _fields_ = [("d_rgb", C.c_void_p), ("d_direct", C.c_void_p), ("d_scatter", C.c_void_p),
                ("d_emission", C.c_void_p)]


class lg_scatter_grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("i0", C.c_int32 * 3),
                ("voxel_size", C.c_double), ("d_rgb", C.c_void_p), ("d_active", C.c_void_p)]


class lg_scene_spec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n", C.c_int32), ("octaves", C.c_int32), ("blobs", C.c_int32),
                ("seed", C.c_uint64), ("radius", C.c_double), ("axis_min", C.c_double),
                ("axis_max", C.c_double), ("active_fraction", C.c_double), ("dtype", C.c_int32),
                ("with_temperature", C.c_int32)]


class lg_solve_timing(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("sample_ms", "graph_ms", "march_ms", "light_ms", "fw_ms",
                                         "diffuse_ms", "total_ms")]


class lg_solver_buffers(C.Structure):
    _fields_ = [("n", C.c_int32), ("ld", C.c_int32), ("n_edges", C.c_int32), ("n_components", C.c_int32),
                ("d_points", C.c_void_p), ("d_edges", C.c_void_p), ("d_weights", C.c_void_p),
                ("d_D", C.c_void_p), ("d_L", C.c_void_p), ("d_diffuse", C.c_void_p),
                ("d_pred", C.c_void_p)]


class lg_solve_params(C.Structure):
    _fields_ = [("n", C.c_int32), ("max_connect", C.c_int32), ("radius", C.c_double),
                ("seed", C.c_uint64), ("env_seed", C.c_uint64), ("solve", C.c_uint32),
                ("external_points", C.c_int32), ("timing", C.c_int32), ("flags", C.c_int32)]


class lg_solve_stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("points", "edges", "edge_samples", "shadow_rays", "shadow_samples",
                                          "fw_updates", "fw_panel_updates", "fw_pivot_updates", "small_updates", "components",
                                          "big_components", "light_bricks", "edge_bricks", "small_updates_64",
                                          "small_updates_32")]


class lg_fw_tuning(C.Structure):
    _fields_ = [("split_below", C.c_int32), ("use_graph", C.c_int32), ("tuned", C.c_int32), ("points", C.c_int32),
                ("tiles", C.c_int32 * 16), ("whole_us", C.c_double * 16), ("split_us", C.c_double * 16),
                ("stream_ms", C.c_double), ("graph_ms", C.c_double), ("capture_ms", C.c_double)]


class lg_render_stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("camera_rays", "primary_samples", "shading_points", "shadow_rays",
                                          "shadow_samples", "shadow_lane_slots", "chunks", "shadow_bricks")]


LG_MAX_LIGHTS_ = 8


class lg_job(C.Structure):
    _fields_ = [("scene", lg_scene_spec), ("points", C.c_int32), ("iterations", C.c_int32), ("radius", C.c_double),
                ("max_connect", C.c_int32), ("samples", C.c_int32), ("seed", C.c_uint64), ("env_seed", C.c_uint64),
                ("render_seed", C.c_uint64), ("solve_base", C.c_uint32), ("pts_per_kernel", C.c_int32),
                ("scatter_res_factor", C.c_int32), ("n_lights", C.c_int32), ("shader", lg_shader),
                ("march", lg_march_params), ("camera", lg_camera), ("lights", lg_light * LG_MAX_LIGHTS_),
                ("solve_flags", C.c_int32)]


class lg_job_result(C.Structure):
    _fields_ = [("status", C.c_int32), ("batch", C.c_int32), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double), ("reserved_bytes", C.c_uint64)]


class lg_server_stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("queued", "completed", "failed", "batches", "batched_jobs", "solves",
                                          "reserved_bytes", "peak_reserved_bytes", "budget_bytes")] + \
               [("clock_s", C.c_double)]


class lg_batch_item(C.Structure):
    _fields_ = [("density", C.c_void_p), ("temperature", C.c_void_p), ("n", C.c_int32), ("solve", C.c_uint32)]


P = C.c_void_p
I32 = C.c_int32
U32 = C.c_uint32
U64 = C.c_uint64
I64 = C.c_int64
D = C.c_double
SZ = C.c_size_t
PP = C.POINTER(C.c_void_p)
DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); every entry is declared in include/lightgraph.h
SIGNATURES = {
    "lg_last_error": (C.c_char_p, []),
    "lg_abi_version": (I32, []),
    "lg_ctx_create": (I32, [I32, U64, PP]),
    "lg_ctx_destroy": (I32, [P]),
    "lg_sync": (I32, [P]),
    "lg_ctx_stream": (I32, [P, PP]),
    "lg_ctx_bytes_in_use": (I32, [P, C.POINTER(U64)]),
    "lg_ctx_set_kernel_timing": (I32, [P, I32]),
    "lg_ctx_set_kernel_timing_only": (I32, [P, C.c_char_p]),
    "lg_ctx_set_brick_counting": (I32, [P, I32]),
    "lg_kernel_time": (I32, [P, C.c_char_p, C.POINTER(U64), DP]),
    "lg_kernel_time_reset": (I32, [P]),
    "lg_malloc": (I32, [P, SZ, PP]),
    "lg_free": (I32, [P, P]),
    "lg_memcpy_h2d": (I32, [P, P, P, SZ]),
    "lg_memcpy_d2h": (I32, [P, P, P, SZ]),
    "lg_copy_d": (I32, [P, P, P, SZ]),
    "lg_memset_d": (I32, [P, P, I32, SZ]),
    "lg_timer_mark": (I32, [P, I32]),
    "lg_timer_elapsed": (I32, [P, I32, I32, C.POINTER(C.c_float)]),
    "lg_grid_create": (I32, [P, C.POINTER(lg_grid_desc), P, PP]),
    "lg_grid_generate": (I32, [P, C.POINTER(lg_scene_spec), PP, PP]),
    "lg_scene_generate_host": (I32, [C.POINTER(lg_scene_spec), P, I32]),
    "lg_grid_info_get": (I32, [P, C.POINTER(lg_grid_info)]),
    "lg_grid_read_dense": (I32, [P, P, P]),
    "lg_grid_destroy": (I32, [P, P]),
    "lg_sample_points": (I32, [P, P, U64, U32, I32, P]),
    "lg_grid_active_bbox": (I32, [P, P, C.POINTER(I32), C.POINTER(I32)]),
    "lg_dart_throw": (I32, [P, P, DP, DP, I32, C.POINTER(U32), U32, U32, C.POINTER(C.c_float), P, I32,
                            C.POINTER(I32)]),
    "lg_build_edges": (I32, [P, P, I32, D, I32, P, I32, C.POINTER(I32)]),
    "lg_march_edges": (I32, [P, P, C.POINTER(lg_shader), C.POINTER(lg_march_params), P, P, I32, P, P, I32]),
    "lg_vertex_light": (I32, [P, P, P, C.POINTER(lg_shader), C.POINTER(lg_light), I32,
                              C.POINTER(lg_march_params), U64, U32, P, I32, P]),
    "lg_fw_init": (I32, [P, I32, P, I32]),
    "lg_fw_ld": (I32, [I32]),
    "lg_fw_solve": (I32, [P, I32, P, I32, P, P, P, I32]),
    "lg_diffuse": (I32, [P, I32, P, I32, DP, P, P]),
    "lg_diffuse_noalloc": (I32, [P, I32, P, I32, DP, P, P]),
    "lg_solver_create": (I32, [P, I32, I32, I32, PP]),
    "lg_solver_destroy": (I32, [P, P]),
    "lg_solver_get": (I32, [P, C.POINTER(lg_solver_buffers)]),
    "lg_solver_timing": (I32, [P, C.POINTER(lg_solve_timing)]),
    "lg_solve": (I32, [P, P, P, P, C.POINTER(lg_shader), C.POINTER(lg_light), I32,
                       C.POINTER(lg_march_params), C.POINTER(lg_solve_params)]),
    "lg_fw_solve_batched": (I32, [P, I32, C.POINTER(I32), C.POINTER(C.c_void_p), C.POINTER(I32)]),
    "lg_solver_stats": (I32, [P, P, C.POINTER(lg_solve_stats)]),
    "lg_server_create": (I32, [I32, U64, I32, I32, PP]),
    "lg_server_submit": (I32, [P, C.POINTER(lg_job), C.POINTER(C.c_void_p), C.POINTER(U64)]),
    "lg_server_wait": (I32, [P, U64, C.POINTER(lg_job_result)]),
    "lg_server_poll": (I32, [P, U64, C.POINTER(lg_job_result), C.POINTER(I32)]),
    "lg_server_get_stats": (I32, [P, C.POINTER(lg_server_stats)]),
    "lg_server_destroy": (I32, [P]),
    "lg_solve_batch": (I32, [P, P, C.POINTER(lg_batch_item), I32, C.POINTER(lg_shader), C.POINTER(lg_light), I32,
                             C.POINTER(lg_march_params), C.POINTER(lg_solve_params)]),
    "lg_splat_radius": (I32, [P, I64, I64, DP, DP]),
    "lg_scatter_grid_create": (I32, [P, P, I32, C.POINTER(lg_scatter_grid)]),
    "lg_scatter_grid_destroy": (I32, [P, C.POINTER(lg_scatter_grid)]),
    "lg_splat": (I32, [P, P, P, I64, D, C.POINTER(lg_scatter_grid)]),
    "lg_camera_setup": (I32, [DP, DP, D, D, D, D, D, I32, I32, C.POINTER(lg_camera)]),
    "lg_render": (I32, [P, P, C.POINTER(lg_scatter_grid), P, C.POINTER(lg_camera), C.POINTER(lg_shader),
                        C.POINTER(lg_light), I32, C.POINTER(lg_march_params), I32, U64, U64,
                        C.POINTER(lg_films)]),
    "lg_fw_autotune": (I32, [P, C.POINTER(lg_fw_tuning)]),
    "lg_fw_get_tuning": (I32, [P, C.POINTER(lg_fw_tuning)]),
    "lg_fw_set_tuning": (I32, [P, C.POINTER(lg_fw_tuning)]),
    "lg_render_workspace": (I32, [I32, I32, I32, I32, I64, C.POINTER(C.c_uint64)]),
    "lg_render_get_stats": (I32, [P, C.POINTER(lg_render_stats)]),
    "lg_envmap_create": (I32, [P, I32, I32, P, PP]),
    "lg_envmap_destroy": (I32, [P, P]),
    "lg_envmap_read_cdf": (I32, [P, P, DP, DP, DP, DP]),
    "lg_host_last_error": (C.c_char_p, []),
    "lg_shader_default": (I32, [C.POINTER(lg_shader)]),
    "lg_shader_set_coefficients": (I32, [C.POINTER(lg_shader), DP, DP, DP]),
    "lg_shader_parse_ramp": (I32, [C.POINTER(lg_shader), C.c_char_p]),
    "lg_march_params_default": (I32, [C.POINTER(lg_march_params)]),
    "lg_light_directional": (I32, [D, DP, D, DP, C.POINTER(lg_light)]),
    "lg_light_point": (I32, [D, DP, D, DP, C.POINTER(lg_light)]),
    "lg_light_env": (I32, [D, DP, D, I32, P, C.POINTER(lg_light)]),
    "lg_light_parse": (I32, [C.c_char_p, C.POINTER(lg_light)]),
}


def load(path=LIB_PATH):
    """Load liblightgraph.so and attach the prototypes.  Raises OSError if the
    library has not been built (lightgraph_amd/Makefile)."""
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `make -C lightgraph_amd` "
                      "(or __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib

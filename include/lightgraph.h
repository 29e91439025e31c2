/*
 * lightgraph.h - C ABI of the MI355X (gfx950) LightGraph multiple-scattering path.
 *
 * The reference (asiroisvigneux/LightGraph) is a single C++ program; its hot path
 * is the LightGraph estimator (src/LightGraph.h), the scatter-grid splat
 * (src/ScatterGrid.h) and the per-pixel march (src/Renderer.h).  This header is the
 * device boundary that the C++ host classes (lightgraph_amd/host) call.  Every
 * entry point names the reference code it replaces.
 *
 * Conventions
 *   - Every function returns int32_t: LG_OK (0) or a negative LG_ERR_* code.  The
 *     message of the last failure on the calling thread is lg_last_error().
 *   - No C++ exception crosses this boundary.
 *   - Pointers named d_* are device pointers (allocated with lg_malloc or any other
 *     HIP allocator on the context's device).  Pointers named h_* are host pointers.
 *   - A context owns one HIP stream; all work of a context is ordered on that stream
 *     and is asynchronous unless the entry point says it synchronises.  A context is
 *     not thread-safe: use one per host thread.
 *   - World space follows the reference: index space of the density grid maps to
 *     world space by world = origin + voxel_size * index (uniform voxels,
 *     Renderer.h:287-290 rejects non-uniform ones).
 *   - Distances between graph vertices are optical depths tau >= 0 (float32);
 *     +INFINITY means "no edge / unreachable".  Transmittance per channel is
 *     exp(-sigma_t[c] * tau), with sigma_t = -extinction (Shader.h:133).
 */
#ifndef LIGHTGRAPH_H_
#define LIGHTGRAPH_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LG_ABI_VERSION 1

enum {
    LG_OK = 0,
    LG_ERR_INVALID = -1,     /* bad argument                                  */
    LG_ERR_NOMEM = -2,       /* device allocation failed                      */
    LG_ERR_HIP = -3,         /* HIP runtime error                             */
    LG_ERR_BUDGET = -4,      /* request exceeds the context's HBM budget      */
    LG_ERR_UNSUPPORTED = -5, /* valid request this library does not support   */
    LG_ERR_OVERFLOW = -6     /* an output capacity given by the caller is too small */
};

enum { LG_F32 = 0, LG_F16 = 1, LG_BF16 = 2 };

enum { LG_LIGHT_DIRECTIONAL = 0, LG_LIGHT_POINT = 1, LG_LIGHT_ENV = 2 };

#define LG_MAX_RAMP 8
#define LG_MAX_LIGHTS 8

typedef struct lg_ctx lg_ctx;       /* opaque: device, stream, HBM budget   */
typedef struct lg_grid lg_grid;     /* opaque: dense-brick voxel grid       */
typedef struct lg_envmap lg_envmap; /* opaque: env map + FP64 sampling CDF  */
typedef struct lg_solver lg_solver; /* opaque: per-solve device workspace   */

/* Voxel grid description.  Replaces openvdb::FloatGrid + its linear transform
 * (main.cc:491-502).  Index (x,y,z), 0 <= x < nx ..., addresses voxel centres. */
typedef struct lg_grid_desc {
    int32_t nx, ny, nz;
    int32_t dtype;      /* LG_F32 | LG_F16 | LG_BF16: storage type of the bricks */
    double voxel_size;  /* uniform voxel edge in world units                     */
    double origin[3];   /* world position of index (0,0,0)                       */
} lg_grid_desc;

typedef struct lg_grid_info {
    lg_grid_desc desc;
    int32_t bricks[3];        /* ceil(n/8) per axis                               */
    int32_t active_bricks;    /* bricks holding at least one non-zero voxel       */
    int32_t bbox_min[3];      /* active bbox in voxel units at brick resolution,  */
    int32_t bbox_max[3];      /* exclusive max: [8*bmin, 8*(bmax+1))              */
    uint64_t active_voxels;   /* voxels != 0                                      */
    uint64_t device_bytes;    /* HBM held by the grid                             */
} lg_grid_info;

/* Shader parameter block.  Replaces class Shader (Shader.h:100-274) and its
 * ColorRamp (Shader.h:23-93); defaults are those of Shader::Shader (Shader.h:123-138). */
typedef struct lg_shader {
    double density_scale;          /* Shader.h:124 */
    double density_max;            /* Shader.h:125 */
    double shadow_density_scale;   /* Shader.h:126 */
    double scatter_density_scale;  /* Shader.h:127 */
    double scatter_density_min;    /* Shader.h:128 */
    double albedo[3];              /* volumeColor*scattering/(scattering+absorption), Shader.h:132 */
    double extinction[3];          /* -(scattering+absorption), Shader.h:133 (negative) */
    double scattering[3];          /* Shader.h:130, used as 1/scattering in Renderer.h:516 */
    double emission_scale;         /* Shader.h:134 */
    double scatter_scale;          /* Shader.h:136 */
    double scatter_density_mask_power; /* Shader.h:137 */
    int32_t ramp_count;            /* 0: no ramp, emission colour is 0 (Shader.h:66-68) */
    int32_t ramp_pad_;
    double ramp_in, ramp_out;      /* ColorRamp mIn/mOut (Shader.h:26-27) */
    double ramp_pos[LG_MAX_RAMP];
    double ramp_color[LG_MAX_RAMP][3];
} lg_shader;

/* One light.  Replaces class Light and its subclasses (Light.h:26-235). */
typedef struct lg_light {
    int32_t type;        /* LG_LIGHT_*                                             */
    int32_t samples;     /* Light::mSamples (1 for directional and point lights)   */
    float intensity;     /* Light::mIntensity is a float in the reference (Light.h:31) */
    float pad_;
    double color[3];     /* Light::mColor                                          */
    double dir[3];       /* directional: DistantLight::mDir (unit, Light.h:111)    */
    double pos[3];       /* point: PointLight::mPos (Light.h:72)                   */
    double xform[16];    /* env: Light::mXform, row-major, row-vector convention   */
    const lg_envmap *env;/* env: map + CDF created by lg_envmap_create            */
} lg_light;

/* March settings.  Replaces the SceneSettings fields the marches read
 * (SceneSettings.h:31-32). */
typedef struct lg_march_params {
    double cutoff;   /* SceneSettings::cutoff, default 0.005                      */
    double step[2];  /* SceneSettings::step: primary, shadow (index units)        */
} lg_march_params;

/* Pinhole camera.  Replaces PerspectiveCamera (Renderer.h:129-248); the host
 * fills it with lg_camera_setup or its own equivalent. */
typedef struct lg_camera {
    double screen_to_world[16]; /* AffineMap matrix, row-major, world = [x y z 1]*M */
    double scale_width;         /* 0.5*aperture/focal (Renderer.h:162)              */
    double scale_height;        /* scale_width*height/width (Renderer.h:163)        */
    double znear, zfar;         /* ray time span (Renderer.h:234)                   */
    int32_t width, height;
} lg_camera;

/* Device pointers to the four AOV films, each width*height RGBA float32 in the
 * layout of Film::RGBA (Renderer.h:72-101).  Any of them may be NULL. */
typedef struct lg_films {
    float *d_rgb, *d_direct, *d_scatter, *d_emission;
} lg_films;

/* Scatter grid (multiple-scattering radiance, Vec3f) on a coarse lattice:
 * world = index * voxel_size (no origin, like createLinearTransform in
 * ScatterGrid.h:310).  Dense storage covers coarse indices [i0, i0 + n) per axis,
 * x fastest, 3 floats per voxel; voxels outside it or inactive read 0. */
typedef struct lg_scatter_grid {
    int32_t nx, ny, nz;
    int32_t i0[3];         /* coarse index of storage element (0,0,0) */
    double voxel_size;
    float *d_rgb;          /* nx*ny*nz*3 floats                    */
    uint8_t *d_active;     /* nx*ny*nz flags (1 = active voxel)    */
} lg_scatter_grid;

/* Procedural volumes used for the synthetic workloads. */
enum { LG_SCENE_SPHERE_FBM = 0, LG_SCENE_BLOBS_FBM = 1, LG_SCENE_PLUME = 2 };
typedef struct lg_scene_spec {
    int32_t kind;        /* LG_SCENE_*                                     */
    int32_t n;           /* cube side in voxels                            */
    int32_t octaves;     /* fBm octaves                                    */
    int32_t blobs;       /* BLOBS: number of ellipsoids                    */
    uint64_t seed;
    double radius;       /* SPHERE: sphere radius in voxels                */
    double axis_min, axis_max;   /* BLOBS: semi-axis range in voxels       */
    double active_fraction;      /* BLOBS: target fraction of active voxels (0 = no tuning) */
    int32_t dtype;       /* storage type of the density bricks             */
    int32_t with_temperature;    /* PLUME: also produce a temperature grid */
} lg_scene_spec;

/* Per-stage timing of the last lg_solve on a solver (milliseconds, HIP events). */
typedef struct lg_solve_timing {
    float sample_ms, graph_ms, march_ms, light_ms, fw_ms, diffuse_ms, total_ms;
} lg_solve_timing;

/* Device buffers owned by a solver (valid until the next lg_solve on it). */
typedef struct lg_solver_buffers {
    int32_t n, ld, n_edges, n_components;
    float *d_points;      /* n*3, world space                                      */
    int32_t *d_edges;     /* n_edges*2, (i,j) with i<j, in the reference's scan order */
    float *d_weights;     /* n_edges edge optical depths (+INF = dropped)           */
    float *d_D;           /* n x ld all-pairs optical depth (LG_SOLVE_DENSE), or NULL */
    float *d_L;           /* n*3 direct light per vertex                             */
    float *d_diffuse;     /* n*3 diffused radiance per vertex                        */
    int32_t *d_pred;      /* n x ld canonical predecessors, or NULL                  */
} lg_solver_buffers;

/* ---------------------------------------------------------------- context ---- */
const char *lg_last_error(void);
int32_t lg_abi_version(void);
/* hbm_budget_bytes = 0 means "no limit beyond the device". */
int32_t lg_ctx_create(int32_t device, uint64_t hbm_budget_bytes, lg_ctx **out);
int32_t lg_ctx_destroy(lg_ctx *ctx);
int32_t lg_sync(lg_ctx *ctx);
int32_t lg_ctx_stream(lg_ctx *ctx, void **hip_stream_out);
int32_t lg_ctx_bytes_in_use(lg_ctx *ctx, uint64_t *bytes);
/* Per-kernel HIP-event timing of the library's own launches (for benchmarks). */
int32_t lg_ctx_set_kernel_timing(lg_ctx *ctx, int32_t enable);
/* While timing is on, time only the launches of the named kernel (NULL or "":
 * every kernel), so that a timed region carries two events per launch of that
 * kernel and none elsewhere.  No reference counterpart. */
int32_t lg_ctx_set_kernel_timing_only(lg_ctx *ctx, const char *kernel);
int32_t lg_kernel_time(lg_ctx *ctx, const char *kernel, uint64_t *launches, double *total_ms);
/* Renders made while kernel timing is on also count the distinct bricks each
 * shadow launch reads (lg_render_stats.shadow_bricks); the marking slows the
 * shadow phase, so a render counted this way is not the one to time. */
int32_t lg_ctx_set_brick_counting(lg_ctx *ctx, int32_t enable);
int32_t lg_kernel_time_reset(lg_ctx *ctx);

/* Device memory, charged against the context budget. */
int32_t lg_malloc(lg_ctx *ctx, size_t bytes, void **d_out);
int32_t lg_free(lg_ctx *ctx, void *d_ptr);
int32_t lg_memcpy_h2d(lg_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int32_t lg_memcpy_d2h(lg_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int32_t lg_memset_d(lg_ctx *ctx, void *d_dst, int32_t value, size_t bytes);
/* Stream-ordered device-to-device copy by a plain streaming kernel (k_copy);
 * bytes a multiple of 16, pointers 16-byte aligned.  No reference counterpart:
 * it is the known-size kernel that HBM traffic counters are calibrated on. */
int32_t lg_copy_d(lg_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);

/* Stream-ordered timer: lg_timer_mark records an event, lg_timer_elapsed syncs. */
int32_t lg_timer_mark(lg_ctx *ctx, int32_t slot);
int32_t lg_timer_elapsed(lg_ctx *ctx, int32_t slot_a, int32_t slot_b, float *ms);

/* ------------------------------------------------------------------ grids ---- */
/* Upload a dense host volume (x fastest, nx*ny*nz float32) into a dense-brick
 * device grid.  Replaces io::File::readGrid + GridSampler set-up (main.cc:489-502).
 * Voxels != 0 are active; bricks holding an active voxel exist (leaf nodes). */
int32_t lg_grid_create(lg_ctx *ctx, const lg_grid_desc *desc, const float *h_dense, lg_grid **out);
/* Generate a procedural grid on the device (no counterpart in the reference: it
 * only reads VDB files).  temperature_out may be NULL. */
int32_t lg_grid_generate(lg_ctx *ctx, const lg_scene_spec *spec, lg_grid **density_out,
                         lg_grid **temperature_out);
/* Host twin of lg_grid_generate's density (same arithmetic; for CPU-only use). */
int32_t lg_scene_generate_host(const lg_scene_spec *spec, float *h_dense_out, int32_t threads);
int32_t lg_grid_info_get(const lg_grid *grid, lg_grid_info *info);
/* Download the grid as a dense float32 volume (x fastest, nx*ny*nz). */
int32_t lg_grid_read_dense(lg_ctx *ctx, const lg_grid *grid, float *h_out);
int32_t lg_grid_destroy(lg_ctx *ctx, lg_grid *grid);

/* ------------------------------------------------ reference-mode sampler ---- */
/* Active-voxel bounding box at voxel resolution, inclusive (voxels != 0):
 * grid->evalActiveVoxelBoundingBox() of main.cc:512.  An empty grid gives
 * vmin = {0,0,0}, vmax = {-1,-1,-1}. */
int32_t lg_grid_active_bbox(lg_ctx *ctx, const lg_grid *grid, int32_t vmin[3], int32_t vmax[3]);

/* The reference's adaptive dart throwing, LightGraph::scatterPoints
 * (LightGraph.h:206-274), for `count` iterations at once (one workgroup each,
 * as the tbb::parallel_for over iterations of main.cc:534-538).  Iteration k
 * draws from std::mt19937(854123 + h_iters[k]) through
 * uniform_real_distribution<float>(0,1) (RandGen.h:16-44); candidates fall in
 * [bbox_min, bbox_max] (world; the reference's mBbox); a candidate is rejected
 * when the density there is below 0.001 or an accepted point is closer than the
 * adaptive minimum distance; the squared distance is rescaled every
 * scatter_count/100 throws by accepted/expected; throwing stops once more than
 * `target` points are accepted.  h_min_dist[k]: the starting minimum distance
 * in, sqrt of the squared distance at the last acceptance out (minDistVec[k]).
 * d_points_out: count x out_cap x 3 floats (world, acceptance order);
 * h_counts[k]: points accepted (only the first out_cap are written).  The
 * per-iteration state is reset on each call (the reference carries pcSize over
 * within one TBB range).  scatter_count >= 100, target <= 2047. */
int32_t lg_dart_throw(lg_ctx *ctx, const lg_grid *density, const double bbox_min[3], const double bbox_max[3],
                      int32_t count, const uint32_t *h_iters, uint32_t scatter_count, uint32_t target,
                      float *h_min_dist, float *d_points_out, int32_t out_cap, int32_t *h_counts);

/* ---------------------------------------------------------- solve stages ---- */
/* Kernel (a): density-weighted point sampling over the dense bricks.  Replaces
 * the dart thrower LightGraph::scatterPoints (LightGraph.h:206-274) with the
 * density-weighted sampler: voxel v is drawn with probability
 * floor(rho_v*2^24)/sum, then jittered uniformly in its cell.  Stream: Philox4x32-10
 * keyed by seed, counter (point, solve, block, 1).  d_pts_out: n*3 floats (world). */
int32_t lg_sample_points(lg_ctx *ctx, const lg_grid *grid, uint64_t seed, uint32_t solve,
                         int32_t n, float *d_pts_out);

/* Graph build.  Replaces LightGraph::buildGraph (LightGraph.h:287-318): for i
 * ascending, j>i ascending, add (i,j) when |p_i-p_j|^2 < radius^2 (FP64), skipping
 * j whose count reached max_connect and ending row i when its count reaches it.
 * max_connect <= 0 means uncapped.  d_edges_out: edge_cap*2 int32 in row-major
 * order; *n_edges_out (host) receives the count (the call synchronises). */
int32_t lg_build_edges(lg_ctx *ctx, const float *d_pts, int32_t n, double radius,
                       int32_t max_connect, int32_t *d_edges_out, int32_t edge_cap,
                       int32_t *n_edges_out);

/* Kernel (b): edge optical depth.  Replaces LightGraph::raymarchConnection
 * (LightGraph.h:347-418).  For edge e=(i,j): tau_e = sum over samples of
 * max(rho', scatterDensityMin)*h, tau_e = +INF when the ray misses the grid bbox or
 * the reference's dropout (|T|^2 < cutoff) fires.  Writes d_w[e] (may be NULL) and,
 * when d_D is not NULL, D[i*ld+j] = D[j*ld+i] = tau_e. */
int32_t lg_march_edges(lg_ctx *ctx, const lg_grid *grid, const lg_shader *shader,
                       const lg_march_params *params, const float *d_pts,
                       const int32_t *d_edges, int32_t n_edges, float *d_w, float *d_D,
                       int32_t ld);

/* Vertex lighting.  Replaces LightGraph::computeLighting (LightGraph.h:427-505):
 * d_L_out[i] = sum_lights albedo*mean_s(I_s*T_s) + ramp(temp(p_i))*emissionScale.
 * temperature may be NULL (all zero).  Env-light samples use Philox counter
 * (vertex, solve, block, 4) with env_seed. */
int32_t lg_vertex_light(lg_ctx *ctx, const lg_grid *density, const lg_grid *temperature,
                        const lg_shader *shader, const lg_light *lights, int32_t n_lights,
                        const lg_march_params *params, uint64_t env_seed, uint32_t solve,
                        const float *d_pts, int32_t n, float *d_L_out);

/* Fill D (ld rows of ld floats, ld >= n) with +INF off the diagonal and 0 on it. */
int32_t lg_fw_init(lg_ctx *ctx, int32_t n, float *d_D, int32_t ld);
/* Required leading dimension of D for lg_fw_solve with n vertices. */
int32_t lg_fw_ld(int32_t n);

/* Kernel (c): all-pairs min-plus closure of the symmetric matrix D (ld rows of
 * ld floats, ld = lg_fw_ld(n), zero diagonal, +INF = no edge; rows and columns
 * n..ld-1 are padding, +INF off the diagonal as lg_fw_init leaves them, and are
 * never taken as pivots) in place.  Replaces
 * LightGraph::floydAlgo (LightGraph.h:512-542): the reference's per-channel
 * (max, x) Floyd-Warshall on transmittance equals (min, +) on tau because
 * T_c = exp(-sigma_t[c]*tau) with a global sigma_t.  Pivots are taken in
 * ascending order in blocks of lg_fw_ld's tile.  If d_pred is not NULL it receives
 * canonical predecessors: pred[i][j] = smallest k adjacent to j minimising
 * D[i][k] + w(k,j) (D[i][i]=0), -1 if unreachable, i on the diagonal; this needs
 * the edge list and its weights. */
int32_t lg_fw_solve(lg_ctx *ctx, int32_t n, float *d_D, int32_t ld, int32_t *d_pred,
                    const int32_t *d_edges, const float *d_w, int32_t n_edges);

/* Radiance diffusion.  Replaces LightGraph::diffuseLight (LightGraph.h:554-587):
 * P_c = exp(-sigma_t[c]*D) (so P_ii = 1, LightGraph.h:530-534), columns normalised
 * by max(sum_i P_ij, 1), out[i] = sum_j L[j] (x) P_ij. */
int32_t lg_diffuse(lg_ctx *ctx, int32_t n, const float *d_D, int32_t ld,
                   const double sigma_t[3], const float *d_L, float *d_out);

/* lg_diffuse without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as lg_diffuse (bit for bit); the column sums and normalised light
 * live in a fixed per-GPU scratch array inside the library, sized for
 * n <= LG_DIFFUSE_NOALLOC_MAX_N (2.25 MiB, held by each GPU the library runs on
 * until the process ends); larger n returns LG_ERR_INVALID.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other
 * contexts, other threads) race and must be serialised by the caller.  Kernel
 * timing does not cover this entry point. */
#define LG_DIFFUSE_NOALLOC_MAX_N 65536
int32_t lg_diffuse_noalloc(lg_ctx *ctx, int32_t n, const float *d_D, int32_t ld,
                           const double sigma_t[3], const float *d_L, float *d_out);

/* ------------------------------------------------------------ full solve ---- */
/* Per-solve workspace for graphs (or batches, lg_solve_batch) of up to n_max
 * vertices in total, n_max <= 2^24.  flags: LG_SOLVER_DENSE_D allocates the dense
 * n x ld optical-depth matrix, LG_SOLVER_PRED also the predecessor matrix; both
 * need n_max <= 65536.  edge_cap is ignored (the edge
 * buffers grow on demand); it is kept for callers that size their own copies. */
int32_t lg_solver_create(lg_ctx *ctx, int32_t n_max, int32_t edge_cap, int32_t flags,
                         lg_solver **out);
int32_t lg_solver_destroy(lg_ctx *ctx, lg_solver *solver);
int32_t lg_solver_get(const lg_solver *solver, lg_solver_buffers *out);
int32_t lg_solver_timing(const lg_solver *solver, lg_solve_timing *out);

typedef struct lg_solve_params {
    int32_t n;             /* points in the graph                                  */
    int32_t max_connect;   /* GRAPH_MAX_CONNECT (main.cc:62) = 12; <= 0 uncapped    */
    double radius;         /* edge radius in world units                           */
    uint64_t seed;         /* point-sampling seed (reference engines: 854123+i)     */
    uint64_t env_seed;     /* env-light sampling seed (Light.h:157: 4564446)        */
    uint32_t solve;        /* solve index (the reference's LightGraph iteration i)  */
    int32_t external_points; /* 1: use the points already in the solver's d_points */
    int32_t timing;        /* 1: record per-stage HIP events                       */
    int32_t flags;         /* LG_SOLVE_DENSE: also write the dense n x ld outputs;
                            * LG_SOLVE_STATS: count the work (lg_solver_stats);
                            * LG_SOLVE_BFS_ORDER, LG_SOLVE_ND_ORDER: see below        */
} lg_solve_params;

/* LG_SOLVE_BFS_ORDER (no reference counterpart; off in the drop-in path): close
 * each component of more than 128 vertices in breadth-first vertex order (level
 * from the component's smallest vertex, then vertex index) instead of ascending
 * vertex order.  The shortest paths are the same; the blocked closure skips the
 * many remainder tiles whose panel operands are still +INF in that order, and the
 * FP32 sums are rounded in the new pivot order (distances and diffused radiance
 * agree with the ascending order to FP32 rounding, bit-exact for integer
 * weights).  Points, edges and the outputs keep the vertex indices of the call;
 * at most 2^21 vertices per call. */
/* LG_SOLVE_ND_ORDER (no reference counterpart; off in the drop-in path; not
 * with LG_SOLVE_BFS_ORDER): nested-dissection vertex order for the same
 * components.  Each component's vertex list is split recursively at the
 * midpoint of its points' longest bounding-box side; the left vertices with a
 * neighbour on the right (the separator) are ordered after both halves, down to
 * parts of 256 vertices, each part in ascending vertex order.  While the closure
 * pivots through one half the other stays unreachable, so far more remainder
 * tiles are skipped than in breadth-first order; the same FP32 rounding note
 * applies. */
enum { LG_SOLVE_DENSE = 1, LG_SOLVE_STATS = 2, LG_SOLVE_BFS_ORDER = 4, LG_SOLVE_ND_ORDER = 8 };

/* Work done by the last lg_solve / lg_solve_batch call, counted on the device
 * when it ran with LG_SOLVE_STATS (otherwise the counters read 0).  The counts are
 * the algorithmic units of the kernels' rooflines. */
typedef struct lg_solve_stats {
    uint64_t points;          /* vertices of the call                                   */
    uint64_t edges;           /* accepted edges (kernel b rays)                         */
    uint64_t edge_samples;    /* trilinear samples of the edge march                    */
    uint64_t shadow_rays;     /* vertex-light shadow rays                               */
    uint64_t shadow_samples;  /* their trilinear samples                                */
    /* Floyd-Warshall (kernel c): executed (i, j, k) min-plus updates, one add and
     * one min each; the pivots of a round are counted in the 32-pivot chunks that
     * run, each tile part (quadrant, half, quarter) counts what it computed */
    uint64_t fw_updates;         /* remainder tiles                                     */
    uint64_t fw_panel_updates;   /* row-panel tiles                                     */
    uint64_t fw_pivot_updates;   /* pivot-tile closures (pivots x 128 x 128)            */
    uint64_t small_updates;   /* sum of m^3 over the components closed in LDS (m <= 128) */
    uint64_t components;      /* connected components                                   */
    uint64_t big_components;  /* components above 128 vertices                          */
    /* compulsory bytes of the marches: distinct 8^3 bricks whose voxels the
     * samples' trilinear taps read, summed over the grids of the call */
    uint64_t light_bricks;    /* vertex-light shadow rays                               */
    uint64_t edge_bricks;     /* edge march                                             */
    /* the part of small_updates closed by the 33..64- and 2..32-vertex kernels
     * (the rest by the 65..128-vertex one) */
    uint64_t small_updates_64;
    uint64_t small_updates_32;
} lg_solve_stats;
int32_t lg_solver_stats(lg_ctx *ctx, const lg_solver *solver, lg_solve_stats *out);
enum { LG_SOLVER_DENSE_D = 1, LG_SOLVER_PRED = 2 };

/* One LightGraph solve (LightGraph::operator() body, LightGraph.h:163-189):
 * sample -> edges -> edge march -> Floyd-Warshall -> vertex light -> diffuse.
 * The all-pairs closure runs per connected component of the graph (optical
 * depth across components is +INF, the reference's zero transmittance). */
int32_t lg_solve(lg_ctx *ctx, lg_solver *solver, const lg_grid *density,
                 const lg_grid *temperature, const lg_shader *shader, const lg_light *lights,
                 int32_t n_lights, const lg_march_params *params, const lg_solve_params *sp);

/* ----------------------------------------------------------- batched form ---- */
/* Many independent min-plus closures of uneven n (lg_fw_solve for each matrix)
 * in one launch sequence: each 128-pivot round closes the pivot tiles of every
 * matrix still that large in one launch, then all their panels, then all their
 * remainder tiles.  d_D[g] is ld[g] x ld[g] with ld[g] = lg_fw_ld(n[g]).  The host
 * arrays are read during the call; the device matrices are updated in place.
 * Replaces the per-solve floydAlgo calls of a tbb::parallel_for over solves
 * (LightGraph.h:512-542 under main.cc:534-538). */
int32_t lg_fw_solve_batched(lg_ctx *ctx, int32_t count, const int32_t *h_n,
                            float *const *h_D, const int32_t *h_ld);

/* Tile and schedule choices of the batched Floyd-Warshall, picked by a
 * measurement on the device (no reference counterpart: the reference runs one
 * scalar loop per solve).  split_below: panel and remainder launches of fewer
 * 128 x 128 tiles run split: a panel tile as two 128 x 64 column-half
 * workgroups (four 128 x 32 column quarters when the launch has fewer than 128
 * panel tiles) and a remainder tile as four 64 x 64 quadrant workgroups;
 * split_below = 0 keeps one workgroup per tile everywhere (results are
 * identical either way).  use_graph: the rounds of a batch are captured once into a
 * hipGraph and replayed while the schedule's shape repeats.  The first batched
 * closure on a device runs lg_fw_autotune unless the environment sets
 * LG_FW_AUTOTUNE=0 (then the defaults hold); LG_FW_SPLIT_BELOW and LG_FW_GRAPH
 * override the two choices. */
typedef struct lg_fw_tuning {
    int32_t split_below;
    int32_t use_graph;
    int32_t tuned;          /* 1: measured on this device                           */
    int32_t points;         /* entries of the tables below                           */
    int32_t tiles[16];      /* tile counts of one remainder launch                   */
    double whole_us[16];    /* its time with one workgroup per tile                  */
    double split_us[16];    /* and with four                                         */
    double stream_ms;       /* a synthetic 24-matrix schedule launched on the stream */
    double graph_ms;        /* the same schedule replayed as a captured hipGraph     */
    double capture_ms;      /* capturing and instantiating that graph                */
} lg_fw_tuning;
int32_t lg_fw_autotune(lg_ctx *ctx, lg_fw_tuning *out);
int32_t lg_fw_get_tuning(lg_ctx *ctx, lg_fw_tuning *out);
int32_t lg_fw_set_tuning(lg_ctx *ctx, const lg_fw_tuning *tuning);

/* One graph of a batched solve. */
typedef struct lg_batch_item {
    const lg_grid *density;      /* density grid of this solve (all items: one dtype) */
    const lg_grid *temperature;  /* float32 temperature grid or NULL                  */
    int32_t n;                   /* points in this graph                              */
    uint32_t solve;              /* solve index: selects the RNG stream               */
} lg_batch_item;

/* count independent LightGraph solves (LightGraph::operator() for every index of
 * the tbb::parallel_for at main.cc:534-538, each on its own volume) as one
 * pipeline: one sampling launch, one vertex-light launch, one graph build, one
 * fused edge-march launch, one small-component launch and the batched
 * Floyd-Warshall for the components above 128 vertices.  Item k owns vertices
 * [voff_k, voff_k + n_k) of the solver's buffers, voff_k = n_0 + ... + n_{k-1};
 * edges carry global vertex ids.  shader, lights, params and sp (except n, solve
 * and flags, which must be 0) are shared.  Each item's outputs equal those of
 * lg_solve on that item alone. */
int32_t lg_solve_batch(lg_ctx *ctx, lg_solver *solver, const lg_batch_item *items, int32_t count,
                       const lg_shader *shader, const lg_light *lights, int32_t n_lights,
                       const lg_march_params *params, const lg_solve_params *sp);

/* ------------------------------------------------------------ serving ---- */
/* A render job: volume -> `iterations` LightGraph solves of `points` points ->
 * splat -> render (main.cc:505-128 as one request).  Iteration i uses solve
 * index solve_base + i. */
typedef struct lg_job {
    lg_scene_spec scene;          /* procedural volume (kept resident between jobs) */
    int32_t points, iterations;
    double radius;                /* edge radius, world units                       */
    int32_t max_connect;
    int32_t samples;              /* AA samples per pixel                           */
    uint64_t seed, env_seed, render_seed;
    uint32_t solve_base;
    int32_t pts_per_kernel;       /* SceneSettings::ptsPerKernel                    */
    int32_t scatter_res_factor;   /* SceneSettings::scatterGridResFactor            */
    int32_t n_lights;
    lg_shader shader;
    lg_march_params march;
    lg_camera camera;
    lg_light lights[LG_MAX_LIGHTS];
    int32_t solve_flags;          /* 0, LG_SOLVE_BFS_ORDER or LG_SOLVE_ND_ORDER; jobs share a solve
                                   * launch only with equal flags                  */
} lg_job;

typedef struct lg_job_result {
    int32_t status;               /* LG_OK or the failing call's code               */
    int32_t batch;                /* jobs that shared this job's solve launch       */
    double t_submit, t_admit, t_start, t_done;   /* seconds on the server clock     */
    uint64_t reserved_bytes;      /* device memory the job was admitted with        */
} lg_job_result;

typedef struct lg_server_stats {
    uint64_t queued, completed, failed, batches, batched_jobs, solves;
    uint64_t reserved_bytes, peak_reserved_bytes, budget_bytes;
    double clock_s;
} lg_server_stats;

typedef struct lg_server lg_server;

/* Request loop with continuous batching (no reference counterpart: the reference
 * renders one scene per process).  `workers` threads, each with its own HIP
 * stream, take the oldest admissible job plus queued jobs that can share its
 * solve launch (same brick type, shader, lights, march and graph parameters), up
 * to max_batch, while the estimated device memory fits hbm_budget; resident
 * volumes are evicted least-recently-used when the budget needs room. */
int32_t lg_server_create(int32_t device, uint64_t hbm_budget, int32_t workers, int32_t max_batch,
                         lg_server **out);
/* Queues a job; the four films (width*height RGBA float32 each, NULL to skip) are
 * written when it completes.  LG_ERR_BUDGET: the job alone exceeds the budget. */
int32_t lg_server_submit(lg_server *srv, const lg_job *job, float *const h_films[4], uint64_t *id_out);
int32_t lg_server_wait(lg_server *srv, uint64_t id, lg_job_result *out);
int32_t lg_server_poll(lg_server *srv, uint64_t id, lg_job_result *out, int32_t *done);
int32_t lg_server_get_stats(lg_server *srv, lg_server_stats *out);
/* Finishes every submitted job, then stops the workers and frees the device memory. */
int32_t lg_server_destroy(lg_server *srv);

/* ----------------------------------------------------------- splat/render ---- */
/* Splat radius of ScatterGrid::operator() (ScatterGrid.h:71-73) with the voxel size
 * of LightGraph::buildVdbPointGrid (LightGraph.h:713-714) restated on the host. */
int32_t lg_splat_radius(const float *h_pts, int64_t count, int64_t pts_per_kernel_total,
                        double *radius_out, double *point_voxel_out);
/* Coarse topology (ScatterGrid::buildScatterGrid, ScatterGrid.h:105-144): a coarse
 * voxel is active if any density voxel in its footprint is active, then the set is
 * dilated by one voxel (26-neighbourhood).  Allocates sg->d_rgb / d_active. */
int32_t lg_scatter_grid_create(lg_ctx *ctx, const lg_grid *density, int32_t res_factor,
                               lg_scatter_grid *sg);
int32_t lg_scatter_grid_destroy(lg_ctx *ctx, lg_scatter_grid *sg);
/* Gaussian splat (ScatterGrid::operator(), ScatterGrid.h:61-98) of count points
 * (world, float3) with radiance (float3) into every active coarse voxel. */
int32_t lg_splat(lg_ctx *ctx, const float *d_pts, const float *d_rad, int64_t count,
                 double radius, lg_scatter_grid *sg);

/* Camera set-up of main.cc:96-102 + PerspectiveCamera ctor (Renderer.h:148-171). */
int32_t lg_camera_setup(const double cam_rotate[3], const double cam_translate[3],
                        double cam_y_rotate, double focal, double aperture, double znear,
                        double zfar, int32_t width, int32_t height, lg_camera *out);

/* Kernel (d): per-pixel march.  Replaces VolumeRender::operator()
 * (Renderer.h:506-693).  samples = AA samples per pixel, in [1, 2^20), and
 * width * height * samples < 2^32.  Streams (reference: one RandGen(1) per TBB
 * range, Renderer.h:518): the primary ray of (pixel, sample s) draws camera
 * jitter and steps from Philox counter (pixel, s, block, 2) keyed by seed; at its
 * q-th shading point, light k draws shadow steps from (pixel, q, block,
 * 5 | k << 8 | s << 12) keyed by seed and env-map samples from (pixel, q, block,
 * 3 | k << 8 | s << 12) keyed by env_seed.  Device scratch: lg_render_workspace. */
int32_t lg_render(lg_ctx *ctx, const lg_grid *density, const lg_scatter_grid *scatter,
                  const lg_grid *temperature, const lg_camera *camera,
                  const lg_shader *shader, const lg_light *lights, int32_t n_lights,
                  const lg_march_params *params, int32_t samples, uint64_t seed,
                  uint64_t env_seed, const lg_films *films);
/* Device scratch an lg_render of this size keeps on its context (for admission
 * control), given the most shading points one camera ray can take (<= 0:
 * unknown; a ray through n voxels at primary step h takes at most n / (0.75 h)
 * + 1).  A render of more than one chunk of pixels keeps two chunk workspaces
 * (consecutive chunks run on the context's two streams).  It is exceeded only
 * when one block of 64 pixels holds more than 2 GiB of shading points.  No
 * reference counterpart. */
int32_t lg_render_workspace(int32_t width, int32_t height, int32_t samples, int32_t n_lights,
                            int64_t max_points_per_ray, uint64_t *bytes);

/* Work of the last lg_render on a context, counted on the device while kernel
 * timing is on (lg_ctx_set_kernel_timing), otherwise 0: the algorithmic units of
 * kernel (d)'s roofline.  No reference counterpart. */
typedef struct lg_render_stats {
    uint64_t camera_rays;      /* (pixel, AA sample) primary rays                     */
    uint64_t primary_samples;  /* trilinear samples of the primary march              */
    uint64_t shading_points;   /* primary samples that passed the cutoff              */
    uint64_t shadow_rays;      /* shadow rays (shading points x lights x samples)     */
    uint64_t shadow_samples;   /* their trilinear samples                             */
    uint64_t shadow_lane_slots; /* 64 x the most samples of a lane, summed over the
                                 * shadow waves: shadow_samples / this = lane efficiency */
    uint64_t chunks;           /* shading-point chunks the render ran in              */
    uint64_t shadow_bricks;    /* distinct 8^3 bricks the shadow taps read, summed over
                                * the shadow launches (with lg_ctx_set_brick_counting) */
} lg_render_stats;
int32_t lg_render_get_stats(lg_ctx *ctx, lg_render_stats *out);

/* Environment map: width x height RGB float32 (row 0 = top), luminance*sin(theta)
 * weighted 2D distribution (Light.h:165-176, Sampling.h:128-167) with its CDF
 * prefix sums computed on the device in FP64. */
int32_t lg_envmap_create(lg_ctx *ctx, int32_t width, int32_t height, const float *h_rgb,
                         lg_envmap **out);
int32_t lg_envmap_destroy(lg_ctx *ctx, lg_envmap *env);
/* Download the FP64 CDFs: conditional (height rows x (width+1)), marginal (height+1),
 * per-row integrals (height) and the total. */
int32_t lg_envmap_read_cdf(lg_ctx *ctx, const lg_envmap *env, double *h_cond,
                           double *h_marg, double *h_rowint, double *h_total);

/* ------------------------------------------------- host-side parameters ---- */
/* Message of the last failure of the functions in this section. */
const char *lg_host_last_error(void);
/* Shader::Shader defaults (Shader.h:123-138). */
int32_t lg_shader_default(lg_shader *s);
/* Shader::setAbsorption/setScattering/setVolumeColor (Shader.h:165-211). */
int32_t lg_shader_set_coefficients(lg_shader *s, const double absorption[3], const double scattering[3],
                                   const double volume_color[3]);
/* parseColorRamp + Shader::setEmissionColorRamp (main.cc:280-319). */
int32_t lg_shader_parse_ramp(lg_shader *s, const char *text);
/* SceneSettings cutoff/step defaults (SceneSettings.h:51-52). */
int32_t lg_march_params_default(lg_march_params *p);
/* parseLight "dir" + DistantLight (main.cc:255-269, Light.h:106-111). */
int32_t lg_light_directional(double y_rot, const double color[3], double intensity, const double rot[3],
                             lg_light *out);
/* PointLight (Light.h:56-91); the reference has no CLI form for it. */
int32_t lg_light_point(double y_rot, const double color[3], double intensity, const double pos[3],
                       lg_light *out);
/* EnvLight (Light.h:132-153) over a map from lg_envmap_create. */
int32_t lg_light_env(double y_rot, const double color[3], double intensity, int32_t samples,
                     const lg_envmap *env, lg_light *out);
/* parseLight (main.cc:234-274) for "dir:..." and "point:yRot:r,g,b:intensity:x,y,z". */
int32_t lg_light_parse(const char *text, lg_light *out);

#ifdef __cplusplus
}
#endif

#endif /* LIGHTGRAPH_H_ */

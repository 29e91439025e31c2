// Splat (ScatterGrid.h), camera (Renderer.h:129-248), environment map CDFs
// (Light.h:132-176, Sampling.h:76-167) and the per-pixel march (kernel d,
// Renderer.h:506-693).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "lg_internal.h"
#include "lg_lights.h"

namespace lg {

// ------------------------------------------------------------- camera ------
// OpenVDB Mat4 (row-major, row-vector convention): M = M * R(axis, angle).
static void post_rotate(double M[16], int axis, double angle) {
    const double c = cos(angle), s = sin(angle);
    double R[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (axis == 0) {
        R[5] = c;
        R[6] = s;
        R[9] = -s;
        R[10] = c;
    } else if (axis == 1) {
        R[0] = c;
        R[2] = -s;
        R[8] = s;
        R[10] = c;
    } else {
        R[0] = c;
        R[1] = s;
        R[4] = -s;
        R[5] = c;
    }
    double out[16];
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            double acc = 0.0;
            for (int m = 0; m < 4; ++m) acc += M[4 * r + m] * R[4 * m + k];
            out[4 * r + k] = acc;
        }
    memcpy(M, out, sizeof(out));
}

// ------------------------------------------------------------ env map ------
// Conditional CDFs, one lane per row, sequential FP64 prefix (Sampling.h:81-93
// with double accumulation).
__global__ void k_env_cond(const float *__restrict__ func, int w, int h, double *__restrict__ cond,
                           double *__restrict__ rowint) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= h) return;
    double *cdf = cond + (long long)v * (w + 1);
    const float *f = func + (long long)v * w;
    cdf[0] = 0.0;
    for (int i = 1; i <= w; ++i) cdf[i] = cdf[i - 1] + (double)f[i - 1] / (double)w;
    const double I = cdf[w];
    if (I == 0.0) {
        for (int i = 1; i <= w; ++i) cdf[i] = (double)i / (double)w;
    } else {
        for (int i = 1; i <= w; ++i) cdf[i] = cdf[i] / I;
    }
    rowint[v] = I;
}

__global__ void k_env_marg(const double *__restrict__ rowint, int h, double *__restrict__ marg,
                           double *__restrict__ total) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    marg[0] = 0.0;
    for (int i = 1; i <= h; ++i) marg[i] = marg[i - 1] + rowint[i - 1] / (double)h;
    const double I = marg[h];
    if (I == 0.0) {
        for (int i = 1; i <= h; ++i) marg[i] = (double)i / (double)h;
    } else {
        for (int i = 1; i <= h; ++i) marg[i] = marg[i] / I;
    }
    *total = I;
}

}  // namespace lg

using namespace lg;

extern "C" {

int32_t lg_camera_setup(const double rot[3], const double tr[3], double yrot, double focal, double aperture,
                        double znear, double zfar, int32_t width, int32_t height, lg_camera *out) {
    LG_CHECK_ARG(rot && tr && out, "lg_camera_setup: NULL argument");
    LG_CHECK_ARG(width > 0 && height > 0, "lg_camera_setup: image size must be positive");
    LG_CHECK_ARG(znear > 0 && zfar > znear, "lg_camera_setup: need 0 < znear < zfar");
    memset(out, 0, sizeof(*out));
    // main.cc:98-102: rotation = camRotate + (0, camYRotate, 0),
    // translation = camTranslate * yRotXform(camYRotate)
    double Y[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    post_rotate(Y, 1, (yrot / 180.0) * M_PI);
    double t[3];
    for (int c = 0; c < 3; ++c) t[c] = tr[0] * Y[c] + tr[1] * Y[4 + c] + tr[2] * Y[8 + c] + Y[12 + c];
    const double r[3] = {rot[0], rot[1] + yrot, rot[2]};
    double M[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    // PerspectiveCamera ctor (Renderer.h:166-169)
    post_rotate(M, 0, r[0] * M_PI / 180.0);
    post_rotate(M, 1, r[1] * M_PI / 180.0);
    post_rotate(M, 2, r[2] * M_PI / 180.0);
    M[12] += t[0];
    M[13] += t[1];
    M[14] += t[2];
    memcpy(out->screen_to_world, M, sizeof(M));
    out->scale_width = 0.5 * aperture / focal;
    out->scale_height = out->scale_width * double(height) / double(width);
    out->znear = znear;
    out->zfar = zfar;
    out->width = width;
    out->height = height;
    return LG_OK;
}

int32_t lg_envmap_create(lg_ctx *ctx, int32_t w, int32_t h, const float *h_rgb, lg_envmap **out) {
    LG_CHECK_ARG(ctx && h_rgb && out && w >= 2 && h >= 2, "lg_envmap_create: bad argument");
    *out = nullptr;
    // luminance * sin(theta) weights in float (Light.h:165-176)
    std::vector<float> Y((size_t)w * h);
    for (int i = 0; i < h; ++i) {
        const float sin_theta = (float)std::sin(M_PI * (double)((float)i / (float)(h - 1)));
        for (int j = 0; j < w; ++j) {
            const float *p = h_rgb + 3 * ((size_t)i * w + j);
            float y = (float)(0.2126 * (double)p[0] + 0.7152 * (double)p[1] + 0.0722 * (double)p[2]);
            y *= sin_theta;
            Y[(size_t)i * w + j] = y;
        }
    }
    lg_envmap *e = new lg_envmap();
    e->width = w;
    e->height = h;
    float *func = nullptr;
    double *total = nullptr;
    int32_t rc = dalloc(ctx, (size_t)w * h * 3, &e->rgb);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)h * (w + 1), &e->cond);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)h + 1, &e->marg);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)h, &e->rowint);
    if (rc == LG_OK) rc = dalloc(ctx, (size_t)w * h, &func);
    if (rc == LG_OK) rc = dalloc(ctx, 1, &total);
    if (rc == LG_OK) {
        hipError_t err = hipMemcpyAsync(e->rgb, h_rgb, (size_t)w * h * 3 * 4, hipMemcpyHostToDevice, ctx->stream);
        if (err == hipSuccess) err = hipMemcpyAsync(func, Y.data(), Y.size() * 4, hipMemcpyHostToDevice, ctx->stream);
        if (err == hipSuccess) {
            k_env_cond<<<(h + 63) / 64, 64, 0, ctx->stream>>>(func, w, h, e->cond, e->rowint);
            k_env_marg<<<1, 64, 0, ctx->stream>>>(e->rowint, h, e->marg, total);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipMemcpyAsync(&e->total, total, 8, hipMemcpyDeviceToHost, ctx->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
        if (err != hipSuccess) {
            set_error("lg_envmap_create: %s", hipGetErrorString(err));
            rc = LG_ERR_HIP;
        }
    }
    dfree(ctx, func);
    dfree(ctx, total);
    if (rc != LG_OK) {
        lg_envmap_destroy(ctx, e);
        return rc;
    }
    *out = e;
    return LG_OK;
}

int32_t lg_envmap_destroy(lg_ctx *ctx, lg_envmap *e) {
    if (!e) return LG_OK;
    dfree(ctx, e->rgb);
    dfree(ctx, e->cond);
    dfree(ctx, e->marg);
    dfree(ctx, e->rowint);
    delete e;
    return LG_OK;
}

int32_t lg_envmap_read_cdf(lg_ctx *ctx, const lg_envmap *e, double *cond, double *marg, double *rowint,
                           double *total) {
    LG_CHECK_ARG(ctx && e, "lg_envmap_read_cdf: NULL argument");
    if (cond) LG_HIP(hipMemcpyAsync(cond, e->cond, (size_t)e->height * (e->width + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (marg) LG_HIP(hipMemcpyAsync(marg, e->marg, ((size_t)e->height + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (rowint) LG_HIP(hipMemcpyAsync(rowint, e->rowint, (size_t)e->height * 8, hipMemcpyDeviceToHost, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    if (total) *total = e->total;
    return LG_OK;
}

int32_t lg_splat_radius(const float *h_pts, int64_t count, int64_t kernel_total, double *radius_out,
                        double *voxel_out) {
    (void)h_pts;
    (void)count;
    (void)kernel_total;
    (void)radius_out;
    (void)voxel_out;
    set_error("lg_splat_radius: not built yet");
    return LG_ERR_UNSUPPORTED;
}

int32_t lg_scatter_grid_create(lg_ctx *, const lg_grid *, int32_t, lg_scatter_grid *) {
    set_error("lg_scatter_grid_create: not built yet");
    return LG_ERR_UNSUPPORTED;
}

int32_t lg_scatter_grid_destroy(lg_ctx *, lg_scatter_grid *) { return LG_OK; }

int32_t lg_splat(lg_ctx *, const float *, const float *, int64_t, double, lg_scatter_grid *) {
    set_error("lg_splat: not built yet");
    return LG_ERR_UNSUPPORTED;
}

int32_t lg_render(lg_ctx *, const lg_grid *, const lg_scatter_grid *, const lg_grid *, const lg_camera *,
                  const lg_shader *, const lg_light *, int32_t, const lg_march_params *, int32_t, uint64_t, uint64_t,
                  const lg_films *) {
    set_error("lg_render: not built yet");
    return LG_ERR_UNSUPPORTED;
}

int32_t lg_fw_solve_batched(lg_ctx *ctx, int32_t count, const int32_t *h_n, float *const *h_D, const int32_t *h_ld) {
    LG_CHECK_ARG(ctx && count >= 0 && (count == 0 || (h_n && h_D && h_ld)), "lg_fw_solve_batched: bad argument");
    for (int g = 0; g < count; ++g) {
        LG_CHECK_ARG(h_ld[g] == fw_ld(h_n[g]), "lg_fw_solve_batched: ld[%d] must be lg_fw_ld(n)", g);
        LG_TRY(fw_solve_impl(ctx, h_n[g], h_D[g], h_ld[g]));
    }
    return LG_OK;
}

}  // extern "C"

// Device-side lights: Light::illuminate of the directional, point and environment
// lights (Light.h:51-231) with the environment map's 2D distribution sampled from
// FP64 CDFs (Sampling.h:60-167 restated with double prefix sums).
#pragma once

#include <cmath>

#include "lg_internal.h"

namespace lgl {

struct DevLight {
    int type, samples;
    float intensity;
    int pad;
    double inv_samples;  // 1/samples when samples is a power of two (exact), else 0
    double color[3];
    double dir[3];
    double pos[3];
    double xf[12];  // rows 0..3, cols 0..2 of the 4x4 transform (row-vector convention)
    const float *env_rgb;
    const double *cond;
    const double *marg;
    int ew, eh;
};

struct DevLights {
    int n;
    int pad;
    DevLight l[LG_MAX_LIGHTS];
};

inline int32_t make_dev_lights(const lg_light *lights, int n, DevLights *out) {
    LG_CHECK_ARG(n >= 0 && n <= LG_MAX_LIGHTS, "lights: %d lights exceed the limit of %d", n, LG_MAX_LIGHTS);
    LG_CHECK_ARG(n == 0 || lights, "lights: NULL light array");
    out->n = n;
    for (int k = 0; k < n; ++k) {
        const lg_light &s = lights[k];
        DevLight &d = out->l[k];
        LG_CHECK_ARG(s.type >= LG_LIGHT_DIRECTIONAL && s.type <= LG_LIGHT_ENV, "lights: bad type %d", s.type);
        LG_CHECK_ARG(s.samples >= 1, "lights: samples must be >= 1");
        d.type = s.type;
        d.samples = s.samples;
        d.inv_samples = (s.samples & (s.samples - 1)) == 0 ? 1.0 / (double)s.samples : 0.0;
        d.intensity = s.intensity;
        for (int c = 0; c < 3; ++c) {
            d.color[c] = s.color[c];
            d.dir[c] = s.dir[c];
            d.pos[c] = s.pos[c];
        }
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 3; ++c) d.xf[3 * r + c] = s.xform[4 * r + c];
        d.env_rgb = nullptr;
        d.cond = d.marg = nullptr;
        d.ew = d.eh = 0;
        if (s.type == LG_LIGHT_ENV) {
            LG_CHECK_ARG(s.env, "lights: env light without a map");
            d.env_rgb = s.env->rgb;
            d.cond = s.env->cond;
            d.marg = s.env->marg;
            d.ew = s.env->width;
            d.eh = s.env->height;
        }
    }
    return LG_OK;
}

// Distribution1D::SampleContinuous with FindInterval (Sampling.h:60-72,102-119);
// cdf has count+1 FP64 entries.
__device__ __forceinline__ double sample_1d(const double *cdf, int count, double u, int *off) {
    const int size = count + 1;
    int first = 0, len = size;
    while (len > 0) {
        const int half = len >> 1, middle = first + half;
        if (cdf[middle] <= u) {
            first = middle + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    int offset = first - 1;
    offset = offset < 0 ? 0 : (offset > size - 2 ? size - 2 : offset);
    if (off) *off = offset;
    double du = u - cdf[offset];
    const double span = cdf[offset + 1] - cdf[offset];
    if (span > 0) du /= span;
    return ((double)offset + du) / (double)count;
}

// Correctly rounded float sin/cos (evaluated in FP64 and rounded once), so host
// and device agree on every bit.
__device__ __forceinline__ float sin_f(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cos_f(float x) { return (float)cos((double)x); }

// kRcp: a point light's direction and falloff from reciprocals (see
// lgd::isect_world_ray) instead of four divisions.
template <bool kRcp = false>
__device__ __forceinline__ void illuminate(const DevLight &lt, const double p[3], lgd::Stream &rng, double ldir[3],
                                           double inten[3], double &distance) {
    if (lt.type == LG_LIGHT_DIRECTIONAL) {  // DistantLight::illuminate, Light.h:118-126
        for (int c = 0; c < 3; ++c) {
            ldir[c] = lt.dir[c];
            inten[c] = lt.color[c] * (double)lt.intensity;
        }
        distance = lgd::kDblMax;
    } else if (lt.type == LG_LIGHT_POINT) {  // PointLight::illuminate, Light.h:79-89
        double v[3];
        for (int c = 0; c < 3; ++c) v[c] = p[c] - lt.pos[c];
        const float r2 = (float)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        distance = (double)sqrtf(r2);
        const double fall = 1.0 + 4.0 * M_PI * (double)r2;
        if (kRcp) {
            const double rd = lgd::rcp_newton(distance), rf = lgd::rcp_newton(fall);
            for (int c = 0; c < 3; ++c) {
                ldir[c] = v[c] * rd;
                inten[c] = (lt.color[c] * (double)lt.intensity) * rf;
            }
        } else {
            for (int c = 0; c < 3; ++c) {
                ldir[c] = v[c] / distance;
                inten[c] = (lt.color[c] * (double)lt.intensity) / fall;
            }
        }
    } else {  // EnvLight::illuminate, Light.h:161-231
        const float u0 = rng.next();
        const float u1 = rng.next();
        int v = 0;
        const float sp1 = (float)sample_1d(lt.marg, lt.eh, (double)u1, &v);
        const float sp0 = (float)sample_1d(lt.cond + (long long)v * (lt.ew + 1), lt.ew, (double)u0, nullptr);
        int col_lo = (int)floorf(sp0 * (float)lt.ew), row_lo = (int)floorf(sp1 * (float)lt.eh);
        int col_hi = (int)ceilf(sp0 * (float)lt.ew), row_hi = (int)ceilf(sp1 * (float)lt.eh);
        const float col_ratio = sp0 * (float)lt.ew - (float)col_lo;
        const float row_ratio = sp1 * (float)lt.eh - (float)row_lo;
        // the reference can read one past the last texel here; clamp instead
        col_lo = col_lo < 0 ? 0 : (col_lo > lt.ew - 1 ? lt.ew - 1 : col_lo);
        col_hi = col_hi < 0 ? 0 : (col_hi > lt.ew - 1 ? lt.ew - 1 : col_hi);
        row_lo = row_lo < 0 ? 0 : (row_lo > lt.eh - 1 ? lt.eh - 1 : row_lo);
        row_hi = row_hi < 0 ? 0 : (row_hi > lt.eh - 1 ? lt.eh - 1 : row_hi);
        const float theta = (float)((double)sp0 * M_PI * 2.0);
        const float phi = (float)((double)sp1 * M_PI);
        const double sin_phi = (double)sin_f(phi);
        double sd[3] = {(double)cos_f(theta) * sin_phi, (double)cos_f(phi), (double)sin_f(theta) * sin_phi};
        double t[3];
        for (int c = 0; c < 3; ++c) t[c] = sd[0] * lt.xf[c] + sd[1] * lt.xf[3 + c] + sd[2] * lt.xf[6 + c] + lt.xf[9 + c];
        const double len = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int c = 0; c < 3; ++c) sd[c] = t[c] / len;
        const float wll = (1.0f - row_ratio) * (1.0f - col_ratio), wlh = (1.0f - row_ratio) * col_ratio;
        const float whl = row_ratio * (1.0f - col_ratio), whh = row_ratio * col_ratio;
        const float *LL = lt.env_rgb + 3 * ((long long)row_lo * lt.ew + col_lo);
        const float *LH = lt.env_rgb + 3 * ((long long)row_lo * lt.ew + col_hi);
        const float *HL = lt.env_rgb + 3 * ((long long)row_hi * lt.ew + col_lo);
        const float *HH = lt.env_rgb + 3 * ((long long)row_hi * lt.ew + col_hi);
        for (int c = 0; c < 3; ++c) {
            const double img = (double)wll * (double)LL[c] + (double)wlh * (double)LH[c] + (double)whl * (double)HL[c] +
                               (double)whh * (double)HH[c];
            ldir[c] = -sd[c];
            inten[c] = img * lt.color[c] * (double)lt.intensity;
        }
        distance = lgd::kDblMax;
    }
}

}  // namespace lgl

// Host-side parameter construction mirroring the reference's Shader, ColorRamp,
// light and CLI parsing code (Shader.h, Light.h, main.cc:214-319).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>
#include <string>
#include <unordered_set>
#include <vector>

#include "lightgraph.h"

namespace lg {
void set_error(const char *fmt, ...);  // the library's thread-local error message (lg_last_error)
}

namespace {

thread_local std::string g_host_err;

int32_t fail(const std::string &msg) {
    g_host_err = msg;
    return LG_ERR_INVALID;
}

// OpenVDB Mat4 row-major, row-vector convention: M = M * R(axis, angle)
// (math::rotation, Mat4::postRotate).
void post_rotate(double M[16], int axis, double angle) {
    const double c = std::cos(angle), s = std::sin(angle);
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
    std::memcpy(M, out, sizeof(out));
}

void identity(double M[16]) {
    for (int i = 0; i < 16; ++i) M[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

void matmul(const double A[16], const double B[16], double out[16]) {
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            double acc = 0.0;
            for (int m = 0; m < 4; ++m) acc += A[4 * r + m] * B[4 * m + k];
            out[4 * r + k] = acc;
        }
}

// yRotXform (main.cc:68-70)
void y_rot_xform(double deg, double M[16]) {
    identity(M);
    post_rotate(M, 1, (deg / 180.0) * M_PI);
}

// Mat4::transform(Vec3): v * M with translation
void transform(const double M[16], const double v[3], double out[3]) {
    for (int c = 0; c < 3; ++c) out[c] = v[0] * M[c] + v[1] * M[4 + c] + v[2] * M[8 + c] + M[12 + c];
}

// strToVec (main.cc:214-229)
bool str_to_vec(const std::string &s, double out[3]) {
    std::vector<double> r;
    std::stringstream ss(s);
    try {
        while (ss.good()) {
            std::string sub;
            std::getline(ss, sub, ',');
            r.push_back(std::stod(sub));
        }
    } catch (...) {
        return false;
    }
    if (r.size() == 1) {
        out[0] = out[1] = out[2] = r[0];
        return true;
    }
    if (r.size() < 3) return false;
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
    return true;
}

std::vector<std::string> split(const std::string &s, char sep) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    while (ss.good()) {
        std::string sub;
        std::getline(ss, sub, sep);
        out.push_back(sub);
    }
    return out;
}

}  // namespace

extern "C" {

const char *lg_host_last_error(void) { return g_host_err.c_str(); }

// Shader::Shader defaults (Shader.h:123-138).
int32_t lg_shader_default(lg_shader *s) {
    if (!s) return fail("lg_shader_default: NULL");
    std::memset(s, 0, sizeof(*s));
    s->density_scale = 1.0;
    s->density_max = 100.0;
    s->shadow_density_scale = 1.0;
    s->scatter_density_scale = 0.35;
    s->scatter_density_min = 0.025;
    const double absorption[3] = {0.1, 0.1, 0.1}, scattering[3] = {1.5, 1.5, 1.5}, color[3] = {1, 1, 1};
    for (int c = 0; c < 3; ++c) {
        s->scattering[c] = scattering[c];
        s->albedo[c] = (color[c] * scattering[c]) / (scattering[c] + absorption[c]);
        s->extinction[c] = -(scattering[c] + absorption[c]);
    }
    s->emission_scale = 0.0;
    s->scatter_scale = 4.0;
    s->scatter_density_mask_power = 0.25;
    s->ramp_count = 0;
    s->ramp_in = 0.0;
    s->ramp_out = 1.0;
    return LG_OK;
}

// Shader::setAbsorption/setScattering/setVolumeColor with recomputeAlbedo and
// recomputeExtinction (Shader.h:165-211).
int32_t lg_shader_set_coefficients(lg_shader *s, const double absorption[3], const double scattering[3],
                                   const double volume_color[3]) {
    if (!s || !absorption || !scattering || !volume_color) return fail("lg_shader_set_coefficients: NULL");
    for (int c = 0; c < 3; ++c) {
        s->scattering[c] = scattering[c];
        s->albedo[c] = (volume_color[c] * scattering[c]) / (scattering[c] + absorption[c]);
        s->extinction[c] = -(scattering[c] + absorption[c]);
    }
    return LG_OK;
}

// parseColorRamp (main.cc:280-319): "pos,pos,.../r,g,b:r,g,b:.../in/out".
int32_t lg_shader_parse_ramp(lg_shader *s, const char *text) {
    if (!s || !text) return fail("lg_shader_parse_ramp: NULL");
    const auto parts = split(text, '/');
    if (parts.size() != 4) return fail("This is not a valid color ramp argument");
    std::vector<double> pos;
    std::vector<std::string> cols;
    try {
        for (auto &p : split(parts[0], ',')) pos.push_back(std::stod(p));
        cols = split(parts[1], ':');
        s->ramp_in = std::stod(parts[2]);
        s->ramp_out = std::stod(parts[3]);
    } catch (...) {
        return fail("This is not a valid color ramp argument");
    }
    if (pos.size() != cols.size()) return fail("Those two vectors should be the same size");
    if ((int)pos.size() > LG_MAX_RAMP) return fail("color ramp has more than LG_MAX_RAMP knots");
    s->ramp_count = (int32_t)pos.size();
    for (size_t k = 0; k < pos.size(); ++k) {
        s->ramp_pos[k] = pos[k];
        if (!str_to_vec(cols[k], s->ramp_color[k])) return fail("bad color ramp colour");
    }
    return LG_OK;
}

int32_t lg_march_params_default(lg_march_params *p) {
    if (!p) return fail("lg_march_params_default: NULL");
    p->cutoff = 0.005;  // SceneSettings.h:51-52
    p->step[0] = 1.0;
    p->step[1] = 3.0;
    return LG_OK;
}

// DistantLight from parseLight "dir" (main.cc:255-269, Light.h:106-111).
int32_t lg_light_directional(double y_rot, const double color[3], double intensity, const double rot[3],
                             lg_light *out) {
    if (!color || !rot || !out) return fail("lg_light_directional: NULL");
    std::memset(out, 0, sizeof(*out));
    double orient[16], yr[16], M[16];
    identity(orient);
    post_rotate(orient, 0, rot[0] * M_PI / 180.0);
    post_rotate(orient, 1, rot[1] * M_PI / 180.0);
    post_rotate(orient, 2, rot[2] * M_PI / 180.0);
    y_rot_xform(y_rot, yr);
    matmul(orient, yr, M);
    const double d0[3] = {0.0, 0.0, -1.0};
    double d[3];
    transform(M, d0, d);
    const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    out->type = LG_LIGHT_DIRECTIONAL;
    out->samples = 1;
    out->intensity = (float)intensity;
    for (int c = 0; c < 3; ++c) {
        out->color[c] = color[c];
        out->dir[c] = d[c] / len;
    }
    std::memcpy(out->xform, M, sizeof(M));
    return LG_OK;
}

// PointLight (Light.h:56-91); the reference has no CLI form, this one follows
// the "dir" layout: yRot, colour, intensity, position.
int32_t lg_light_point(double y_rot, const double color[3], double intensity, const double pos[3], lg_light *out) {
    if (!color || !pos || !out) return fail("lg_light_point: NULL");
    std::memset(out, 0, sizeof(*out));
    double M[16];
    y_rot_xform(y_rot, M);
    out->type = LG_LIGHT_POINT;
    out->samples = 1;
    out->intensity = (float)intensity;
    for (int c = 0; c < 3; ++c) out->color[c] = color[c];
    transform(M, pos, out->pos);
    std::memcpy(out->xform, M, sizeof(M));
    return LG_OK;
}

// EnvLight (Light.h:132-153, main.cc:244-254) over a map made by lg_envmap_create.
int32_t lg_light_env(double y_rot, const double color[3], double intensity, int32_t samples, const lg_envmap *env,
                     lg_light *out) {
    if (!color || !out || !env) return fail("lg_light_env: NULL");
    if (samples < 1) return fail("lg_light_env: samples must be >= 1");
    std::memset(out, 0, sizeof(*out));
    double M[16];
    y_rot_xform(y_rot, M);
    out->type = LG_LIGHT_ENV;
    out->samples = samples;
    out->intensity = (float)intensity;
    for (int c = 0; c < 3; ++c) out->color[c] = color[c];
    std::memcpy(out->xform, M, sizeof(M));
    out->env = env;
    return LG_OK;
}

// parseLight (main.cc:234-274) for "dir:yRot:r,g,b:intensity:rx,ry,rz" and
// "point:yRot:r,g,b:intensity:x,y,z".  Env lights need a map: see lg_light_env.
int32_t lg_light_parse(const char *text, lg_light *out) {
    if (!text || !out) return fail("lg_light_parse: NULL");
    const auto f = split(text, ':');
    if (f.empty()) return fail("empty light");
    try {
        if (f[0] == "dir" && f.size() >= 5) {
            double col[3], rot[3];
            if (!str_to_vec(f[2], col) || !str_to_vec(f[4], rot)) return fail("bad light vector");
            return lg_light_directional(std::stod(f[1]), col, std::stod(f[3]), rot, out);
        }
        if (f[0] == "point" && f.size() >= 5) {
            double col[3], pos[3];
            if (!str_to_vec(f[2], col) || !str_to_vec(f[4], pos)) return fail("bad light vector");
            return lg_light_point(std::stod(f[1]), col, std::stod(f[3]), pos, out);
        }
        if (f[0] == "env") return fail("env lights need an environment map (lg_light_env)");
    } catch (...) {
        return fail("bad light number");
    }
    return fail(f[0] + " is not a valid light type!");
}

}  // extern "C"

// ---------------------------------------------------- camera and splat ----
// Pure host computations of the render and splat set-up; errors go to
// lg_last_error like the device entry points'.
namespace {
int32_t arg_error(const char *msg) {
    lg::set_error("%s", msg);
    return LG_ERR_INVALID;
}
}  // namespace

extern "C" {

int32_t lg_camera_setup(const double rot[3], const double tr[3], double yrot, double focal, double aperture,
                        double znear, double zfar, int32_t width, int32_t height, lg_camera *out) {
    if (!rot || !tr || !out) return arg_error("lg_camera_setup: NULL argument");
    if (width <= 0 || height <= 0) return arg_error("lg_camera_setup: image size must be positive");
    if (!(znear > 0 && zfar > znear)) return arg_error("lg_camera_setup: need 0 < znear < zfar");
    std::memset(out, 0, sizeof(*out));
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
    std::memcpy(out->screen_to_world, M, sizeof(M));
    out->scale_width = 0.5 * aperture / focal;
    out->scale_height = out->scale_width * double(height) / double(width);
    out->znear = znear;
    out->zfar = zfar;
    out->width = width;
    out->height = height;
    return LG_OK;
}

// computeVoxelSize(points, 8) (used by LightGraph::buildVdbPointGrid,
// LightGraph.h:713-714) restated: voxel volume = bbox volume / (count / 8),
// refined by counting the occupied voxels of a cell-centred mask until the
// estimate stops shrinking by more than 10%, truncated to 5 decimals.  Then the
// splat radius of ScatterGrid.h:71-73.  Parity unpinned against OpenVDB.
int32_t lg_splat_radius(const float *h_pts, int64_t count, int64_t kernel_total, double *radius_out,
                        double *voxel_out) {
    if (!h_pts || !radius_out || count < 0 || kernel_total < 0) return arg_error("lg_splat_radius: bad argument");
    float vs = 0.1f;
    if (count > 1) {
        int64_t target = count / 8;
        if (target == 0) target = 1;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t i = 0; i < count; ++i)
            for (int c = 0; c < 3; ++c) {
                lo[c] = std::min(lo[c], (double)h_pts[3 * i + c]);
                hi[c] = std::max(hi[c], (double)h_pts[3 * i + c]);
            }
        double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        bool coincident = ext[0] <= 1e-7 && ext[1] <= 1e-7 && ext[2] <= 1e-7;
        if (!coincident) {
            double volume = ext[0] * ext[1] * ext[2];
            if (volume <= 1e-7) {  // collinear or coplanar points
                std::sort(ext, ext + 3);
                volume = (ext[1] <= 1e-7) ? ext[2] * ext[2] * ext[2] : ext[2] * ext[2] * ext[1];
            }
            auto from_volume = [&](double vol, float &v) {
                double vv = vol / (double)target;
                bool ok = true;
                if (vv < 3e-15) {
                    vv = 3e-15;
                    ok = false;
                }
                v = (float)std::pow(vv, 1.0 / 3.0);
                return ok;
            };
            double prev_volume = volume;
            if (from_volume(volume, vs)) {
                int64_t prev_count = 0, vcount = 1;
                std::unordered_set<uint64_t> cells;
                while (vcount > prev_count) {
                    cells.clear();
                    const double inv = 1.0 / (double)vs;
                    for (int64_t i = 0; i < count; ++i) {
                        uint64_t key = 0;
                        for (int c = 0; c < 3; ++c) {
                            const int64_t q = (int64_t)std::floor((double)h_pts[3 * i + c] * inv + 0.5);
                            key = key * 2097152ull + (uint64_t)((q + 1048576) & 2097151);
                        }
                        cells.insert(key);
                    }
                    prev_count = vcount;
                    vcount = (int64_t)cells.size();
                    volume = (double)(vs * vs * vs * (float)vcount);
                    if (volume >= prev_volume) break;
                    prev_volume = volume;
                    const float prev_vs = vs;
                    if (!from_volume(volume, vs)) {
                        vs = prev_vs;
                        break;
                    }
                    if (vs / prev_vs > 0.9f) break;
                }
            }
            for (int d = 5; d < 11; ++d) {  // truncate for readability (5 decimals)
                const double t = std::pow(10.0, d);
                const float tv = (float)(std::floor((double)vs * t + 0.5) / t);
                if (tv != 0.0f) {
                    vs = tv;
                    break;
                }
            }
        }
    }
    const double v = (double)vs;
    if (voxel_out) *voxel_out = v;
    *radius_out = std::pow((3.0 * (double)kernel_total * v * v * v) / (32.0 * M_PI), 1.0 / 3.0);
    return LG_OK;
}

}  // extern "C"

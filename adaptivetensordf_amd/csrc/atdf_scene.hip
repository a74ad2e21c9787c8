// atdf_scene.hip -- the data side of the path: detector rays from the camera
// model and the per-image poses, the X-ray log conversion of raw
// projections, bilinear reads of the resident projections, and random ray
// batches assembled on the device.
//
// Rays follow SceneBase::GetRays (data/SceneBase.cpp:369-449): pixel
// coordinates from uv in f32 (UVToPixel, modules/CameraPose.h:14-30), the
// pinhole unprojection ((px - c) / f, 1) in f64
// (CameraModelModuleImpl::Unproject, modules/CameraPose.cu:469-481), the
// rotation of a unit quaternion q = (x, y, z, w) in f64
// (t = 2 q x v, v' = (v + w t) + q x t), and the cast to f32 at the end.
//   pinhole:       origin = translation[image], direction = R (u / |u|)
//   orthographic:  direction = R (0, 0, 1),
//                  origin = (R (u_x, u_y, 0) + volume_translation) - sqrt(2) direction
// Every expression is evaluated in the written order without FMA
// contraction, so the rays are bit-identical to a C restatement.
//
// All kernels are one lane per ray / pixel and bound by HBM traffic
// (rays: 16 B in, 24 B out; projection reads: four 4-byte gathers from an
// image that stays in L2 for a batch grouped by image).
#include <hip/hip_runtime.h>

#include <cmath>

#include "atdf_device.h"
#include "atdf_internal.h"

namespace atdf {

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// q = (x, y, z, w), unit length.
__device__ __forceinline__ void rotate_point(const double* q, const double* v, double* out)
{
    double t[3];
    cross3(q, v, t);
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = t[k] + t[k];
    double c[3];
    cross3(q, t, c);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = (v[k] + q[3] * t[k]) + c[k];
}

struct RigDev {
    const double* intrinsics;
    const double* rotation;
    const double* translation;
    int64_t num_cameras, num_images;
    float vol_translate[3];
    int w, h, align_corners, orthographic;
};

struct ProjDev {
    const float* projection;
    const float* mask;
    int64_t num_images;
    int w, h, align_corners;
};

__device__ void ray_for_uv(const RigDev& r, float u, float v, int32_t image, int32_t camera, float* o_out,
                           float* d_out)
{
    if (image < 0 || image >= r.num_images || camera < 0 || camera >= r.num_cameras) {
#pragma unroll
        for (int k = 0; k < 3; ++k) o_out[k] = d_out[k] = __builtin_nanf("");
        return;
    }
    float px, py;
    if (r.align_corners) {
        px = u * (float)(r.w - 1);
        py = v * (float)(r.h - 1);
    } else {
        px = u * (float)r.w - 0.5f;
        py = v * (float)r.h - 0.5f;
    }
    const double* K = r.intrinsics + (int64_t)camera * 5;
    double q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = r.rotation[(int64_t)image * 4 + k];
    double un[3] = {((double)px - K[2]) / K[0], ((double)py - K[3]) / K[1], 1.0};
    double o[3], d[3];
    if (!r.orthographic) {
        const double nrm = sqrt((un[0] * un[0] + un[1] * un[1]) + un[2] * un[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k) un[k] = un[k] / nrm;
        rotate_point(q, un, d);
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = r.translation[(int64_t)image * 3 + k];
    } else {
        un[2] = 0.0;
        double p[3];
        rotate_point(q, un, p);
        const double z[3] = {0.0, 0.0, 1.0};
        rotate_point(q, z, d);
        const double s2 = sqrt(2.0);
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = (p[k] + (double)r.vol_translate[k]) - d[k] * s2;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o_out[k] = (float)o[k];
        d_out[k] = (float)d[k];
    }
}

// grid_sample source coordinate with border padding: g = uv * 2 - 1 mapped to
// pixels and clamped to [0, size - 1].
__device__ __forceinline__ float source_coord(float uv, int size, int align_corners)
{
    const float g = uv * 2.0f - 1.0f;
    float x;
    if (align_corners) {
        const float sf = (float)(size - 1) / 2.0f;
        x = (g + 1.0f) * sf;
    } else {
        const float sf = (float)size / 2.0f;
        x = (g + 1.0f) * sf - 0.5f;
    }
    return fminf(fmaxf(x, 0.0f), (float)(size - 1));
}

__device__ float bilinear(const float* __restrict__ img, int w, int h, float x, float y)
{
    const float x0 = floorf(x), y0 = floorf(y);
    const float a = x - x0, e = 1.0f - a;
    const float b = y - y0, s = 1.0f - b;
    const float nw = s * e, ne = s * a, sw = b * e, se = b * a;
    const int ix = (int)x0, iy = (int)y0;
    const bool east = ix + 1 < w, south = iy + 1 < h;
    const float v_nw = img[(int64_t)iy * w + ix];
    const float v_ne = east ? img[(int64_t)iy * w + ix + 1] : 0.0f;
    const float v_sw = south ? img[(int64_t)(iy + 1) * w + ix] : 0.0f;
    const float v_se = (east && south) ? img[(int64_t)(iy + 1) * w + ix + 1] : 0.0f;
    return ((v_nw * nw + v_ne * ne) + v_sw * sw) + v_se * se;
}

__device__ void sample_at(const ProjDev& p, float u, float v, int32_t image, float* target, float* mask)
{
    if (image < 0 || image >= p.num_images) {
        *target = __builtin_nanf("");
        *mask = __builtin_nanf("");
        return;
    }
    const float x = source_coord(u, p.w, p.align_corners);
    const float y = source_coord(v, p.h, p.align_corners);
    const int64_t off = (int64_t)image * p.w * p.h;
    *target = bilinear(p.projection + off, p.w, p.h, x, y);
    *mask = p.mask ? bilinear(p.mask + off, p.w, p.h, x, y) : 1.0f;
}

// Draw k of ray i of a batch: a 64-bit hash of (seed, i, k).
__device__ __forceinline__ uint64_t batch_draw(uint64_t seed, int64_t i, int k)
{
    const uint64_t key = ((uint64_t)i << 2) ^ (uint64_t)k ^ 0x3c6ef372fe94f82bull;
    return splitmix64(seed ^ splitmix64(key));
}

// Uniform integer in [0, n) from the high 32 bits of h (multiply-shift).
__device__ __forceinline__ int64_t draw_below(uint64_t h, int64_t n)
{
    return (int64_t)(((h >> 32) * (uint64_t)n) >> 32);
}

__device__ __forceinline__ float draw_unit(uint64_t h) { return (float)(h >> 40) * (1.0f / 16777216.0f); }

__global__ __launch_bounds__(256) void ray_gen_kernel(RigDev r, const float* __restrict__ uv,
                                                      const int32_t* __restrict__ image_id,
                                                      const int32_t* __restrict__ camera_id, int64_t n,
                                                      float* __restrict__ origin, float* __restrict__ direction)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float o[3], d[3];
    ray_for_uv(r, uv[i * 2], uv[i * 2 + 1], image_id[i], camera_id[i], o, d);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        origin[i * 3 + k] = o[k];
        direction[i * 3 + k] = d[k];
    }
}

struct LogArgs {
    const float* raw;
    const float* mask;
    int64_t n;
    float xray_max;
    float factor;
    int mode;  // 0: -log10, 1: -ln, 2: raw / xray_max (input already in log space)
    float* out;
};

__global__ __launch_bounds__(256) void log_convert_kernel(LogArgs a)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    float v;
    if (a.mode == 2) {
        v = a.raw[i] / a.xray_max;
    } else {
        const float r = fmaxf(a.raw[i], 1e-5f);
        const float x = r / a.xray_max;
        v = -(a.mode == 0 ? log10f(x) : logf(x));
        v = v * a.factor;
    }
    if (a.mask) v = v * a.mask[i];
    a.out[i] = v;
}

__global__ __launch_bounds__(256) void sample_projection_kernel(ProjDev p, const float* __restrict__ uv,
                                                                const int32_t* __restrict__ image_id, int64_t n,
                                                                float* __restrict__ target,
                                                                float* __restrict__ target_mask)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t, m;
    sample_at(p, uv[i * 2], uv[i * 2 + 1], image_id[i], &t, &m);
    target[i] = t;
    if (target_mask) target_mask[i] = m;
}

struct BatchArgs {
    RigDev rig;
    ProjDev proj;
    int has_rig, has_proj;
    const int32_t* indices;
    const int32_t* image_camera;
    int64_t num_indices;
    int w, h, align_corners, interpolate;
    uint64_t seed;
    int64_t n;
    float* uv;
    int32_t* image_id;
    int32_t* camera_id;
    float* target;
    float* target_mask;
    float* origin;
    float* direction;
};

__global__ __launch_bounds__(256) void random_batch_kernel(BatchArgs a)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int64_t k = draw_below(batch_draw(a.seed, i, 0), a.num_indices);
    const int32_t image = a.indices[k];
    const int32_t camera = a.image_camera ? a.image_camera[k] : 0;
    const uint64_t hx = batch_draw(a.seed, i, 1), hy = batch_draw(a.seed, i, 2);
    float u, v;
    if (a.interpolate) {
        u = draw_unit(hx);
        v = draw_unit(hy);
    } else {
        const float x = (float)draw_below(hx, a.w), y = (float)draw_below(hy, a.h);
        if (a.align_corners) {
            u = x / (float)(a.w - 1);
            v = y / (float)(a.h - 1);
        } else {
            u = (x + 0.5f) / (float)a.w;
            v = (y + 0.5f) / (float)a.h;
        }
    }
    a.uv[i * 2] = u;
    a.uv[i * 2 + 1] = v;
    a.image_id[i] = image;
    a.camera_id[i] = camera;
    if (a.has_proj && a.target) {
        float t, m;
        sample_at(a.proj, u, v, image, &t, &m);
        a.target[i] = t;
        if (a.target_mask) a.target_mask[i] = m;
    }
    if (a.has_rig && a.origin && a.direction) {
        float o[3], d[3];
        ray_for_uv(a.rig, u, v, image, camera, o, d);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a.origin[i * 3 + c] = o[c];
            a.direction[i * 3 + c] = d[c];
        }
    }
}

struct VisArgs {
    RigDev rig;
    float roi_min[3], roi_max[3];
    const float* points;
    int64_t n;
    float* mask;
};

// One lane per point; the images are visited in order until one sees the
// point (the reference launches a block per image and sets the flag
// concurrently; the result is the same "any").
__global__ __launch_bounds__(256) void point_visible_kernel(VisArgs a)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float pf[3] = {a.points[i * 3], a.points[i * 3 + 1], a.points[i * 3 + 2]};
    bool inside = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (a.roi_min[k] > pf[k] || a.roi_max[k] < pf[k]) inside = false;
    float vis = 0.f;
    if (inside) {
        const RigDev& r = a.rig;
        const double p[3] = {(double)pf[0], (double)pf[1], (double)pf[2]};
        const double fx = r.intrinsics[0], fy = r.intrinsics[1], cx = r.intrinsics[2], cy = r.intrinsics[3];
        const double sk = 0.0;
        for (int64_t img = 0; img < r.num_images; ++img) {
            const double* q = r.rotation + img * 4;
            const double* t = r.translation + img * 3;
            const double qi[4] = {-q[0], -q[1], -q[2], q[3]};
            const double mt[3] = {-t[0], -t[1], -t[2]};
            double a1[3], a2[3], v[3];
            rotate_point(qi, p, a1);
            rotate_point(qi, mt, a2);
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] = a1[k] + a2[k];
            double ix, iy;
            if (!r.orthographic) {
                const double invz = 1.0 / v[2];
                const double x = v[0] * invz, y = v[1] * invz;
                ix = (fx * x + sk * y) + cx;
                iy = fy * y + cy;
            } else {
                ix = (fx * v[0] + sk * v[1]) + cx;
                iy = fy * v[1] + cy;
            }
            const double rx = round(ix), ry = round(iy);
            if (rx >= 0.0 && rx < (double)r.w && ry >= 0.0 && ry < (double)r.h) {
                vis = 1.f;
                break;
            }
        }
    }
    a.mask[i] = vis;
}

atdf_status make_rig(const atdf_camera_rig* rig, RigDev* out, const char* who)
{
    ATDF_REQUIRE(rig->camera_model == 0 || rig->camera_model == 1,
                 "%s: camera_model must be 0 (pinhole) or 1 (orthographic), got %d", who, rig->camera_model);
    ATDF_REQUIRE(rig->w >= 2 && rig->h >= 2, "%s: bad detector size %d x %d", who, rig->w, rig->h);
    ATDF_REQUIRE(rig->num_cameras >= 1 && rig->num_images >= 1, "%s: empty rig", who);
    ATDF_REQUIRE(rig->intrinsics && rig->rotation && (rig->camera_model == 1 || rig->translation),
                 "%s: null rig buffer", who);
    out->intrinsics = rig->intrinsics;
    out->rotation = rig->rotation;
    out->translation = rig->translation;
    out->num_cameras = rig->num_cameras;
    out->num_images = rig->num_images;
    for (int k = 0; k < 3; ++k) out->vol_translate[k] = rig->volume_translation[k];
    out->w = rig->w;
    out->h = rig->h;
    out->align_corners = rig->align_corners != 0;
    out->orthographic = rig->camera_model == 1;
    return ATDF_OK;
}

atdf_status make_proj(const atdf_projection_set* s, ProjDev* out, const char* who)
{
    ATDF_REQUIRE(s->projection, "%s: null projection buffer", who);
    ATDF_REQUIRE(s->w >= 2 && s->h >= 2 && s->num_images >= 1, "%s: bad projection set %d x %d x %lld", who, s->w,
                 s->h, (long long)s->num_images);
    out->projection = s->projection;
    out->mask = s->mask;
    out->num_images = s->num_images;
    out->w = s->w;
    out->h = s->h;
    out->align_corners = s->align_corners != 0;
    return ATDF_OK;
}

}  // namespace atdf

using namespace atdf;

extern "C" {

atdf_status atdf_generate_rays(const atdf_camera_rig* rig, const float* uv, const int32_t* image_id,
                               const int32_t* camera_id, int64_t num_rays, float* ray_origin, float* ray_direction,
                               atdf_stream_t stream)
{
    ATDF_REQUIRE(rig && num_rays >= 0, "atdf_generate_rays: bad arguments");
    RigDev r;
    atdf_status st = make_rig(rig, &r, "atdf_generate_rays");
    if (st != ATDF_OK) return st;
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(uv && image_id && camera_id && ray_origin && ray_direction, "atdf_generate_rays: null buffer");
    hipLaunchKernelGGL(ray_gen_kernel, dim3((unsigned)div_up(num_rays, 256)), dim3(256), 0, to_hip(stream), r, uv,
                       image_id, camera_id, num_rays, ray_origin, ray_direction);
    return check_launch("atdf_generate_rays");
}

atdf_status atdf_log_convert(const float* raw, const float* mask, int64_t n, double xray_min, double xray_max,
                             double projection_factor, int32_t log10_conversion, int32_t log_space_input,
                             float* out, atdf_stream_t stream)
{
    ATDF_REQUIRE(n >= 0 && xray_max > 0, "atdf_log_convert: bad arguments (n %lld, xray_max %g)", (long long)n,
                 xray_max);
    ATDF_REQUIRE(log_space_input || (xray_min > 0 && xray_min != xray_max),
                 "atdf_log_convert: xray_min must be positive and differ from xray_max (got %g, %g)", xray_min,
                 xray_max);
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(raw && out, "atdf_log_convert: null buffer");
    LogArgs a{};
    a.raw = raw;
    a.mask = mask;
    a.n = n;
    a.xray_max = (float)xray_max;
    a.mode = log_space_input ? 2 : (log10_conversion ? 0 : 1);
    if (!log_space_input) {
        const double den = log10_conversion ? -std::log10(xray_min / xray_max) : -std::log(xray_min / xray_max);
        a.factor = (float)(projection_factor / den);
    }
    a.out = out;
    hipLaunchKernelGGL(log_convert_kernel, dim3((unsigned)div_up(n, 256)), dim3(256), 0, to_hip(stream), a);
    return check_launch("atdf_log_convert");
}

atdf_status atdf_sample_projection(const atdf_projection_set* set, const float* uv, const int32_t* image_id,
                                   int64_t n, float* target, float* target_mask, atdf_stream_t stream)
{
    ATDF_REQUIRE(set && n >= 0, "atdf_sample_projection: bad arguments");
    ProjDev p;
    atdf_status st = make_proj(set, &p, "atdf_sample_projection");
    if (st != ATDF_OK) return st;
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(uv && image_id && target, "atdf_sample_projection: null buffer");
    hipLaunchKernelGGL(sample_projection_kernel, dim3((unsigned)div_up(n, 256)), dim3(256), 0, to_hip(stream), p, uv,
                       image_id, n, target, target_mask);
    return check_launch("atdf_sample_projection");
}

atdf_status atdf_random_ray_batch(const atdf_camera_rig* rig, const atdf_projection_set* set,
                                  const int32_t* image_indices, int64_t num_indices, const int32_t* image_camera,
                                  int32_t w, int32_t h, int32_t align_corners, int32_t interpolate, uint64_t seed,
                                  int64_t num_rays, float* uv, int32_t* image_id, int32_t* camera_id, float* target,
                                  float* target_mask, float* ray_origin, float* ray_direction,
                                  atdf_stream_t stream)
{
    ATDF_REQUIRE(num_rays >= 0 && num_indices >= 1 && num_indices < (int64_t)1 << 31,
                 "atdf_random_ray_batch: bad sizes (rays %lld, images %lld)", (long long)num_rays,
                 (long long)num_indices);
    ATDF_REQUIRE(w >= 2 && h >= 2, "atdf_random_ray_batch: bad detector size %d x %d", w, h);
    BatchArgs a{};
    if (rig) {
        atdf_status st = make_rig(rig, &a.rig, "atdf_random_ray_batch");
        if (st != ATDF_OK) return st;
        ATDF_REQUIRE(rig->w == w && rig->h == h, "atdf_random_ray_batch: rig detector %d x %d differs from %d x %d",
                     rig->w, rig->h, w, h);
        a.has_rig = 1;
    }
    if (set) {
        atdf_status st = make_proj(set, &a.proj, "atdf_random_ray_batch");
        if (st != ATDF_OK) return st;
        ATDF_REQUIRE(set->w == w && set->h == h,
                     "atdf_random_ray_batch: projections %d x %d differ from the detector %d x %d", set->w, set->h, w,
                     h);
        a.has_proj = 1;
    }
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(image_indices && uv && image_id && camera_id, "atdf_random_ray_batch: null buffer");
    ATDF_REQUIRE(!target || set, "atdf_random_ray_batch: targets need a projection set");
    ATDF_REQUIRE(!(ray_origin || ray_direction) || (rig && ray_origin && ray_direction),
                 "atdf_random_ray_batch: rays need a rig and both ray buffers");
    a.indices = image_indices;
    a.image_camera = image_camera;
    a.num_indices = num_indices;
    a.w = w;
    a.h = h;
    a.align_corners = align_corners != 0;
    a.interpolate = interpolate != 0;
    a.seed = seed;
    a.n = num_rays;
    a.uv = uv;
    a.image_id = image_id;
    a.camera_id = camera_id;
    a.target = target;
    a.target_mask = target_mask;
    a.origin = ray_origin;
    a.direction = ray_direction;
    hipLaunchKernelGGL(random_batch_kernel, dim3((unsigned)div_up(num_rays, 256)), dim3(256), 0, to_hip(stream), a);
    return check_launch("atdf_random_ray_batch");
}

atdf_status atdf_point_in_any_image(const atdf_camera_rig* rig, const float* roi_min, const float* roi_max,
                                    const float* points, int64_t n, float* mask, atdf_stream_t stream)
{
    ATDF_REQUIRE(rig && roi_min && roi_max && n >= 0, "atdf_point_in_any_image: bad arguments");
    VisArgs a{};
    atdf_status st = make_rig(rig, &a.rig, "atdf_point_in_any_image");
    if (st != ATDF_OK) return st;
    ATDF_REQUIRE(rig->translation, "atdf_point_in_any_image: the rig needs translations");
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(points && mask, "atdf_point_in_any_image: null buffer");
    for (int k = 0; k < 3; ++k) {
        a.roi_min[k] = roi_min[k];
        a.roi_max[k] = roi_max[k];
    }
    a.points = points;
    a.n = n;
    a.mask = mask;
    hipLaunchKernelGGL(point_visible_kernel, dim3((unsigned)div_up(n, 256)), dim3(256), 0, to_hip(stream), a);
    return check_launch("atdf_point_in_any_image");
}

}  // extern "C"

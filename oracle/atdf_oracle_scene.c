/*
 * atdf_oracle_scene.c -- CPU oracle of the scene data path: detector rays,
 * X-ray log conversion, bilinear projection reads and random ray batches.
 *
 * TEST INFRASTRUCTURE ONLY (see atdf_oracle.h).  Follows
 *   SceneBase::GetRays                 data/SceneBase.cpp:369-449
 *   UVToPixel / PixelToUV              modules/CameraPose.h:14-48
 *   CameraModelModuleImpl::Unproject   modules/CameraPose.cu:469-481
 *   RotatePointGPUk                    modules/CameraPose.cu:27-40; the
 *       rotation itself is Saiga's RotatePoint(q, p) (Saiga is not vendored
 *       in the reference), the unit-quaternion product q p q^-1 written as
 *       Eigen's Quaternion::_transformVector: t = 2 q.vec x v,
 *       v' = (v + w t) + q.vec x t
 *   SceneBase::LoadImagesCT            data/SceneBase.cpp:939-983
 *   UnifiedImage::SampleProjection     data/SceneBase.cpp:76-97; the read is
 *       libtorch's grid_sample (bilinear, padding border) on the CPU:
 *       x = (g + 1) * ((size - 1) / 2) (align_corners) or
 *       (g + 1) * (size / 2) - 0.5, clamped to [0, size - 1]; weights
 *       (1-a)(1-b), a(1-b), (1-a)b, ab; sum in that order; neighbours past
 *       the last pixel read 0
 *   RandomRaybasedSampleDataset::get_batch  data/Dataloader.cpp:9-48 (the
 *       draws are this library's counter-based hash, not libtorch's RNG)
 * Evaluated in the written order with no FMA contraction (bit-exact with the
 * kernels).  Transcendentals (log10f / logf) are the C library's, so the log
 * conversion matches the kernels within a few ulp, not bit for bit.
 */
#include <math.h>
#include <stdint.h>

#include "atdf_oracle.h"

static void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

void atdfo_rotate_point(const double* q, const double* v, double* out)
{
    double t[3], c[3];
    cross3(q, v, t);
    for (int k = 0; k < 3; ++k) t[k] = t[k] + t[k];
    cross3(q, t, c);
    for (int k = 0; k < 3; ++k) out[k] = (v[k] + q[3] * t[k]) + c[k];
}

static void one_ray(int camera_model, int w, int h, int align_corners, const double* K, const double* q,
                    const double* t, const float* vol_translate, float u, float v, float* o_out, float* d_out)
{
    float px, py;
    if (align_corners) {
        px = u * (float)(w - 1);
        py = v * (float)(h - 1);
    } else {
        px = u * (float)w - 0.5f;
        py = v * (float)h - 0.5f;
    }
    double un[3] = {((double)px - K[2]) / K[0], ((double)py - K[3]) / K[1], 1.0};
    double o[3], d[3];
    if (camera_model == 0) {
        double nrm = sqrt((un[0] * un[0] + un[1] * un[1]) + un[2] * un[2]);
        for (int k = 0; k < 3; ++k) un[k] = un[k] / nrm;
        atdfo_rotate_point(q, un, d);
        for (int k = 0; k < 3; ++k) o[k] = t[k];
    } else {
        double p[3];
        const double z[3] = {0.0, 0.0, 1.0};
        un[2] = 0.0;
        atdfo_rotate_point(q, un, p);
        atdfo_rotate_point(q, z, d);
        double s2 = sqrt(2.0);
        for (int k = 0; k < 3; ++k) o[k] = (p[k] + (double)vol_translate[k]) - d[k] * s2;
    }
    for (int k = 0; k < 3; ++k) {
        o_out[k] = (float)o[k];
        d_out[k] = (float)d[k];
    }
}

void atdfo_generate_rays(int camera_model, int w, int h, int align_corners, const double* intrinsics,
                         const double* rotation, const double* translation, const float* vol_translate,
                         const float* uv, const int32_t* image_id, const int32_t* camera_id, int64_t n,
                         float* origin, float* direction)
{
    for (int64_t i = 0; i < n; ++i)
        one_ray(camera_model, w, h, align_corners, intrinsics + (int64_t)camera_id[i] * 5,
                rotation + (int64_t)image_id[i] * 4, translation ? translation + (int64_t)image_id[i] * 3 : 0,
                vol_translate, uv[i * 2], uv[i * 2 + 1], origin + i * 3, direction + i * 3);
}

void atdfo_log_convert(const float* raw, const float* mask, int64_t n, double xray_min, double xray_max,
                       double projection_factor, int log10_conversion, int log_space_input, float* out)
{
    float xm = (float)xray_max;
    float factor = 0.f;
    if (!log_space_input) {
        double den = log10_conversion ? -log10(xray_min / xray_max) : -log(xray_min / xray_max);
        factor = (float)(projection_factor / den);
    }
    for (int64_t i = 0; i < n; ++i) {
        float v;
        if (log_space_input) {
            v = raw[i] / xm;
        } else {
            float r = raw[i] < 1e-5f ? 1e-5f : raw[i];
            float x = r / xm;
            v = -(log10_conversion ? log10f(x) : logf(x));
            v = v * factor;
        }
        if (mask) v = v * mask[i];
        out[i] = v;
    }
}

static float source_coord(float uv, int size, int align_corners)
{
    float g = uv * 2.0f - 1.0f;
    float x;
    if (align_corners) {
        float sf = (float)(size - 1) / 2.0f;
        x = (g + 1.0f) * sf;
    } else {
        float sf = (float)size / 2.0f;
        x = (g + 1.0f) * sf - 0.5f;
    }
    return fminf(fmaxf(x, 0.0f), (float)(size - 1));
}

static float bilinear(const float* img, int w, int h, float x, float y)
{
    float x0 = floorf(x), y0 = floorf(y);
    float a = x - x0, e = 1.0f - a;
    float b = y - y0, s = 1.0f - b;
    float nw = s * e, ne = s * a, sw = b * e, se = b * a;
    int ix = (int)x0, iy = (int)y0;
    int east = ix + 1 < w, south = iy + 1 < h;
    float v_nw = img[(int64_t)iy * w + ix];
    float v_ne = east ? img[(int64_t)iy * w + ix + 1] : 0.0f;
    float v_sw = south ? img[(int64_t)(iy + 1) * w + ix] : 0.0f;
    float v_se = (east && south) ? img[(int64_t)(iy + 1) * w + ix + 1] : 0.0f;
    return ((v_nw * nw + v_ne * ne) + v_sw * sw) + v_se * se;
}

void atdfo_sample_projection(const float* projection, const float* mask, int w, int h, int align_corners,
                             const float* uv, const int32_t* image_id, int64_t n, float* target, float* target_mask)
{
    for (int64_t i = 0; i < n; ++i) {
        float x = source_coord(uv[i * 2], w, align_corners);
        float y = source_coord(uv[i * 2 + 1], h, align_corners);
        int64_t off = (int64_t)image_id[i] * w * h;
        target[i] = bilinear(projection + off, w, h, x, y);
        if (target_mask) target_mask[i] = mask ? bilinear(mask + off, w, h, x, y) : 1.0f;
    }
}

static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static uint64_t batch_draw(uint64_t seed, int64_t i, int k)
{
    uint64_t key = ((uint64_t)i << 2) ^ (uint64_t)k ^ 0x3c6ef372fe94f82bull;
    return splitmix64(seed ^ splitmix64(key));
}

static int64_t draw_below(uint64_t h, int64_t n) { return (int64_t)(((h >> 32) * (uint64_t)n) >> 32); }

void atdfo_random_pixels(const int32_t* indices, int64_t num_indices, const int32_t* image_camera, int w, int h,
                         int align_corners, int interpolate, uint64_t seed, int64_t n, float* uv, int32_t* image_id,
                         int32_t* camera_id)
{
    for (int64_t i = 0; i < n; ++i) {
        int64_t k = draw_below(batch_draw(seed, i, 0), num_indices);
        uint64_t hx = batch_draw(seed, i, 1), hy = batch_draw(seed, i, 2);
        float u, v;
        if (interpolate) {
            u = (float)(hx >> 40) * (1.0f / 16777216.0f);
            v = (float)(hy >> 40) * (1.0f / 16777216.0f);
        } else {
            float x = (float)draw_below(hx, w), y = (float)draw_below(hy, h);
            if (align_corners) {
                u = x / (float)(w - 1);
                v = y / (float)(h - 1);
            } else {
                u = (x + 0.5f) / (float)w;
                v = (y + 0.5f) / (float)h;
            }
        }
        uv[i * 2] = u;
        uv[i * 2 + 1] = v;
        image_id[i] = indices[k];
        camera_id[i] = image_camera ? image_camera[k] : 0;
    }
}

/* SceneBase::PointInAnyImage (data/SceneBase.cu:13-95): inside the closed
 * ROI box and projected (inverse pose R^T p + R^T (-t), camera 0, skew 0,
 * pinhole through x * (1 / z)) into some detector after rounding half away
 * from zero.  The pinhole projection is Saiga's IntrinsicsPinhole::project
 * (not vendored), restated as the perspective division it publishes. */
void atdfo_point_in_any_image(int camera_model, int w, int h, const double* intrinsics, const double* rotation,
                              const double* translation, int64_t num_images, const float* roi_min,
                              const float* roi_max, const float* points, int64_t n, float* mask)
{
    double fx = intrinsics[0], fy = intrinsics[1], cx = intrinsics[2], cy = intrinsics[3], sk = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const float* pf = points + i * 3;
        int inside = 1;
        for (int k = 0; k < 3; ++k)
            if (roi_min[k] > pf[k] || roi_max[k] < pf[k]) inside = 0;
        float vis = 0.f;
        if (inside) {
            double p[3] = {pf[0], pf[1], pf[2]};
            for (int64_t img = 0; img < num_images && vis == 0.f; ++img) {
                const double* q = rotation + img * 4;
                const double* t = translation + img * 3;
                double qi[4] = {-q[0], -q[1], -q[2], q[3]};
                double mt[3] = {-t[0], -t[1], -t[2]};
                double a1[3], a2[3], v[3], ix, iy;
                atdfo_rotate_point(qi, p, a1);
                atdfo_rotate_point(qi, mt, a2);
                for (int k = 0; k < 3; ++k) v[k] = a1[k] + a2[k];
                if (camera_model == 0) {
                    double invz = 1.0 / v[2];
                    double x = v[0] * invz, y = v[1] * invz;
                    ix = (fx * x + sk * y) + cx;
                    iy = fy * y + cy;
                } else {
                    ix = (fx * v[0] + sk * v[1]) + cx;
                    iy = fy * v[1] + cy;
                }
                double rx = round(ix), ry = round(iy);
                if (rx >= 0.0 && rx < (double)w && ry >= 0.0 && ry < (double)h) vis = 1.f;
            }
        }
        mask[i] = vis;
    }
}

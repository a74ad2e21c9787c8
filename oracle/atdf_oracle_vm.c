/*
 * atdf_oracle_vm.c -- CPU oracle of the 2-D / 1-D indirect gathers (the
 * tensorial plane and line features of in-ROI nodes), of GradientCorrection
 * and of the per-sample ROI test.
 *
 * TEST INFRASTRUCTURE ONLY (see atdf_oracle.h).  Follows
 *   IndirectGridSample2D  modules/IndirectGridSample2D.cu:22-162 (forward),
 *                         175-393 (backward: atomics into clipped corners)
 *   IndirectGridSample1D  modules/IndirectGridSample1D.cu:20-140
 *   GradientCorrection    modules/GradientCorrection.cu:22-124
 *   ComputeBoundWeight    structure/HyperTree.cu:962-993
 * Forward values are evaluated in the reference's order with no FMA
 * contraction (bit-exact with the kernels); gradients accumulate in FP64.
 * grad_uv is the analytic derivative of the forward (the reference leaves it
 * at zero in 2-D and sums a different expression in 1-D).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "atdf_oracle.h"

static int clip(int v, int lim)
{
    int a = v > 0 ? v : 0;
    return a < lim - 1 ? a : lim - 1;
}

void atdfo_grid_sample2d_fwd(const float* grid, int64_t N, int C, int H, int W, const int64_t* index,
                             const float* uv, int64_t S, float* out)
{
    (void)N;
    const int64_t plane = (int64_t)H * W;
    for (int64_t s = 0; s < S; ++s) {
        float ix = ((uv[s * 2] + 1) / 2) * (float)(W - 1);
        float iy = ((uv[s * 2 + 1] + 1) / 2) * (float)(H - 1);
        int x0 = (int)floorf(ix), y0 = (int)floorf(iy);
        int x1 = x0 + 1, y1 = y0 + 1;
        float fx1 = (float)x1 - ix, fx0 = ix - (float)x0;
        float fy1 = (float)y1 - iy, fy0 = iy - (float)y0;
        float w[4] = {fx1 * fy1, fx0 * fy1, fx1 * fy0, fx0 * fy0}; /* tnw tne tsw tse */
        int cx[2] = {clip(x0, W), clip(x1, W)}, cy[2] = {clip(y0, H), clip(y1, H)};
        const float* g = grid + index[s] * C * plane;
        for (int c = 0; c < C; ++c) {
            float acc = 0;
            for (int k = 0; k < 4; ++k) acc += g[c * plane + (int64_t)cy[k >> 1] * W + cx[k & 1]] * w[k];
            out[s * C + c] = acc;
        }
    }
}

void atdfo_grid_sample2d_bwd(const float* grid, int64_t N, int C, int H, int W, const int64_t* index,
                             const float* uv, int64_t S, const float* grad_out, double* grad_grid, double* grad_uv)
{
    (void)N;
    const int64_t plane = (int64_t)H * W;
    for (int64_t s = 0; s < S; ++s) {
        double ix = ((uv[s * 2] + 1.0) / 2) * (W - 1);
        double iy = ((uv[s * 2 + 1] + 1.0) / 2) * (H - 1);
        int x0 = (int)floor(ix), y0 = (int)floor(iy);
        double fx[2] = {x0 + 1 - ix, ix - x0}, fy[2] = {y0 + 1 - iy, iy - y0};
        int cx[2] = {clip(x0, W), clip(x0 + 1, W)}, cy[2] = {clip(y0, H), clip(y0 + 1, H)};
        double gx = 0, gy = 0;
        for (int c = 0; c < C; ++c) {
            double go = grad_out[s * C + c];
            for (int k = 0; k < 4; ++k) {
                int bx = k & 1, by = k >> 1;
                int64_t off = (index[s] * C + c) * plane + (int64_t)cy[by] * W + cx[bx];
                grad_grid[off] += fx[bx] * fy[by] * go;
                double v = grid[off];
                gx += v * (bx ? 1.0 : -1.0) * fy[by] * go;
                gy += v * fx[bx] * (by ? 1.0 : -1.0) * go;
            }
        }
        if (grad_uv) {
            grad_uv[s * 2] += gx * (W - 1) / 2.0;
            grad_uv[s * 2 + 1] += gy * (H - 1) / 2.0;
        }
    }
}

void atdfo_grid_sample1d_fwd(const float* grid, int64_t N, int C, int W, const int64_t* index, const float* uv,
                             int64_t S, float* out)
{
    (void)N;
    for (int64_t s = 0; s < S; ++s) {
        float ix = ((uv[s] + 1) / 2) * (float)(W - 1);
        int x0 = (int)floorf(ix);
        int x1 = x0 + 1;
        float w0 = (float)x1 - ix, w1 = ix - (float)x0;
        int c0 = clip(x0, W), c1 = clip(x1, W);
        const float* g = grid + index[s] * C * W;
        for (int c = 0; c < C; ++c) {
            float acc = 0;
            acc += g[c * W + c0] * w0;
            acc += g[c * W + c1] * w1;
            out[s * C + c] = acc;
        }
    }
}

void atdfo_grid_sample1d_bwd(const float* grid, int64_t N, int C, int W, const int64_t* index, const float* uv,
                             int64_t S, const float* grad_out, double* grad_grid, double* grad_uv)
{
    (void)N;
    for (int64_t s = 0; s < S; ++s) {
        double ix = ((uv[s] + 1.0) / 2) * (W - 1);
        int x0 = (int)floor(ix);
        double w0 = x0 + 1 - ix, w1 = ix - x0;
        int c0 = clip(x0, W), c1 = clip(x0 + 1, W);
        double gx = 0;
        for (int c = 0; c < C; ++c) {
            double go = grad_out[s * C + c];
            int64_t base = (index[s] * C + c) * W;
            grad_grid[base + c0] += w0 * go;
            grad_grid[base + c1] += w1 * go;
            gx += (grid[base + c1] - grid[base + c0]) * go;
        }
        if (grad_uv) grad_uv[s] += gx * (W - 1) / 2.0;
    }
}

void atdfo_gradient_correction(const float* ray_direction, const int64_t* ray_index, const float* grad_out,
                               int64_t S, float* grad_pos)
{
    for (int64_t s = 0; s < S; ++s) {
        const float* d = ray_direction + ray_index[s] * 3;
        const float* g = grad_out + s * 3;
        float dot = g[0] * d[0] + g[1] * d[1];
        dot = dot + g[2] * d[2];
        for (int k = 0; k < 3; ++k) grad_pos[s * 3 + k] = g[k] - dot * d[k];
    }
}

void atdfo_compute_bound_weight(const int64_t* node_id, int64_t S, const int64_t* in_roi_node, int64_t n_in,
                                int64_t* out)
{
    for (int64_t s = 0; s < S; ++s) {
        out[s] = -1;
        for (int64_t i = 0; i < n_in; ++i)
            if (node_id[s] == in_roi_node[i]) out[s] = s;
    }
}

/* ---- TVLoss::forward (geometry/geometry.h:31-81) -----------------------
 * x[b, c, s...] at b*sb + c*sc + sum s_d*ss[d]; loss (FP64, sequential sum)
 * = scale * sum over elements and dimensions of w_b * coef_d * |x[s] -
 * x[s + e_d]|, coef_d = C / (B * C * (n_d - 1) * prod_{k != d} n_k); grad
 * (nullable, added) gathered per element as sum_d (float)(scale * coef_d) *
 * (sgn(x[s] - x[s + e_d]) - sgn(x[s - e_d] - x[s])) * w_b. */
static float sgnf_(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); }

double atdfo_tv_loss(const float* x, int64_t B, int64_t C, int D, const int* n, int64_t sb, int64_t sc,
                     const int64_t* ss, const float* weight, double scale, float* grad)
{
    double coef[3];
    float gcoef[3];
    int64_t per = 1;
    for (int d = 0; d < D; ++d) per *= n[d];
    for (int d = 0; d < D; ++d) {
        double count = (double)C * (double)(n[d] - 1);
        for (int k = 0; k < D; ++k)
            if (k != d) count *= (double)n[k];
        coef[d] = (double)C / ((double)B * count);
        gcoef[d] = (float)(scale * coef[d]);
    }
    double acc = 0.0;
    for (int64_t b = 0; b < B; ++b)
        for (int64_t c = 0; c < C; ++c)
            for (int64_t p = 0; p < per; ++p) {
                int s[3] = {0, 0, 0};
                int64_t r = p;
                for (int d = D - 1; d >= 0; --d) {
                    s[d] = (int)(r % n[d]);
                    r /= n[d];
                }
                int64_t addr = b * sb + c * sc;
                for (int d = 0; d < D; ++d) addr += (int64_t)s[d] * ss[d];
                float v = x[addr];
                float wb = weight ? weight[b] : 1.f;
                float g = 0.f;
                for (int d = 0; d < D; ++d) {
                    float up = 0.f, dn = 0.f;
                    if (s[d] + 1 < n[d]) {
                        float diff = v - x[addr + ss[d]];
                        acc += (double)fabsf(diff) * coef[d] * (double)wb;
                        up = sgnf_(diff);
                    }
                    if (s[d] > 0) dn = sgnf_(x[addr - ss[d]] - v);
                    g = g + gcoef[d] * (up - dn);
                }
                if (grad) grad[addr] = grad[addr] + g * wb;
            }
    return acc * scale;
}

/* ------------------------------------------------------------------------ */
/* Ring means of the Fourier regulariser (geometry_tensor_qo.cpp:784-808),   */
/* pointwise: every one of the m^2 meshgrid pairs is interpolated with the   */
/* float corner weights of grid_sample (bilinear, border, align_corners)     */
/* and the values summed in double.  img [Y, X, Z]; out [X / 2, Y], row 0    */
/* +inf (left uninitialised by the reference).                               */
/* ------------------------------------------------------------------------ */
static float ring_axis_coord(int k, float step, int i, int centre, int use_sin, int X, int size)
{
    /* torch on the GPU: th = k * step; th / 180 as th * (1 / 180); * (float)pi;
     * cos / sin; * i; + centre; / (X - 1) as * (1 / (X - 1)); * 2; - 1; then
     * grid_sample's ((g + 1) / 2) * (size - 1) and the border clamp */
    float th = (float)k * step;
    float a = (th * (1.0f / 180.0f)) * 3.14159265358979323846f;
    float cs = use_sin ? sinf(a) : cosf(a);
    float p = (float)i * cs + (float)centre;
    float gn = (p * (1.0f / (float)(X - 1))) * 2.0f - 1.0f;
    float u = ((gn + 1.0f) / 2.0f) * (float)(size - 1);
    if (u < 0.0f) u = 0.0f;
    if (u > (float)(size - 1)) u = (float)(size - 1);
    return u;
}

void atdfo_fourier_ring_means(const float* img, int64_t Y, int X, int Z, float* out)
{
    const int rings = X / 2;
    for (int64_t y = 0; y < Y; ++y) out[y] = INFINITY;
    for (int i = 1; i < rings; ++i) {
        const int n = (int)(2.0 * 3.14159265358979323846 * (double)i);
        const float step = (float)(360.0 / (double)n);
        const int m = (int)ceil(360.0 / (double)step);
        float* rr = (float*)malloc(sizeof(float) * m);
        float* cc = (float*)malloc(sizeof(float) * m);
        for (int k = 0; k < m; ++k) {
            rr[k] = ring_axis_coord(k, step, i, X / 2, 0, X, X);
            cc[k] = ring_axis_coord(k, step, i, Z / 2, 1, X, Z);
        }
        for (int64_t y = 0; y < Y; ++y) {
            const float* im = img + y * (int64_t)X * Z;
            double acc = 0.0;
            for (int a = 0; a < m; ++a) {
                const float iy = rr[a];
                const float iy_n = floorf(iy), iy_s = iy_n + 1.0f;
                for (int b = 0; b < m; ++b) {
                    const float ix = cc[b];
                    const float ix_w = floorf(ix), ix_e = ix_w + 1.0f;
                    const float nw = (ix_e - ix) * (iy_s - iy);
                    const float ne = (ix - ix_w) * (iy_s - iy);
                    const float sw = (ix_e - ix) * (iy - iy_n);
                    const float se = (ix - ix_w) * (iy - iy_n);
                    const int r0 = (int)iy_n, c0 = (int)ix_w;
                    float v = 0.0f;
                    if (r0 < X && c0 < Z) v += im[(int64_t)r0 * Z + c0] * nw;
                    if (r0 < X && c0 + 1 < Z) v += im[(int64_t)r0 * Z + c0 + 1] * ne;
                    if (r0 + 1 < X && c0 < Z) v += im[(int64_t)(r0 + 1) * Z + c0] * sw;
                    if (r0 + 1 < X && c0 + 1 < Z) v += im[(int64_t)(r0 + 1) * Z + c0 + 1] * se;
                    acc += v;
                }
            }
            out[(int64_t)i * Y + y] = (float)(acc / ((double)m * (double)m));
        }
        free(rr);
        free(cc);
    }
}

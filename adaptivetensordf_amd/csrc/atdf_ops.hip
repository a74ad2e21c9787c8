// atdf_ops.hip -- the per-sample operators of the reference API (local
// coordinates, indirect 3-D gather and its backward, line integral, per-node
// FP64 loss), the MSE loss, Adam, and grid layout conversion.
#include <hip/hip_runtime.h>

#include "atdf_device.h"
#include "atdf_internal.h"

namespace atdf {

__global__ __launch_bounds__(256) void local_samples_kernel(const float* __restrict__ pmin, const float* __restrict__ pmax,
                                                            const float* __restrict__ g,
                                                            const int64_t* __restrict__ node, int64_t S,
                                                            float* __restrict__ local)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    int64_t nd = node[s];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float mn = pmin[nd * 3 + k];
        float mx = pmax[nd * 3 + k];
        float size = mx - mn;
        local[s * 3 + k] = ((g[s * 3 + k] - mn) / size) * 2 - 1;
    }
}

// One thread per sample, grid [N, C, Z, Y, X] (reference layout).
__global__ __launch_bounds__(256) void grid_sample3d_fwd_kernel(const float* __restrict__ grid, int C, int Z, int Y,
                                                                int X, const int64_t* __restrict__ index,
                                                                const float* __restrict__ uv, int64_t S,
                                                                float* __restrict__ out)
{
    // one lane per (sample, channel): the eight corner loads of a channel
    // plane are independent of the other channels', so spreading the
    // channels over lanes keeps C times more loads in flight than a lane
    // walking all channels (the reference layout puts them a plane apart)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= S * C) return;
    const int64_t s = i / C;
    const int c = (int)(i - s * C);
    Tri t;
    tri_setup(uv[s * 3 + 0], uv[s * 3 + 1], uv[s * 3 + 2], Z, Y, X, t);
    const int64_t plane = (int64_t)Z * Y * X;
    const float* gc = grid + (index[s] * C + c) * plane;
    const int64_t off[8] = {((int64_t)t.z0 * Y + t.y0) * X + t.x0, ((int64_t)t.z0 * Y + t.y0) * X + t.x1,
                            ((int64_t)t.z0 * Y + t.y1) * X + t.x0, ((int64_t)t.z0 * Y + t.y1) * X + t.x1,
                            ((int64_t)t.z1 * Y + t.y0) * X + t.x0, ((int64_t)t.z1 * Y + t.y0) * X + t.x1,
                            ((int64_t)t.z1 * Y + t.y1) * X + t.x0, ((int64_t)t.z1 * Y + t.y1) * X + t.x1};
    float acc = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += gc[off[k]] * t.w[k];
    out[s * C + c] = acc;
}

__global__ __launch_bounds__(256) void grid_sample3d_bwd_kernel(const float* __restrict__ grid, int C, int Z, int Y,
                                                                int X, const int64_t* __restrict__ index,
                                                                const float* __restrict__ uv, int64_t S,
                                                                const float* __restrict__ grad_out,
                                                                float* __restrict__ grad_grid,
                                                                float* __restrict__ grad_uv)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    const float u = uv[s * 3 + 0], v = uv[s * 3 + 1], w = uv[s * 3 + 2];
    Tri t;
    tri_setup(u, v, w, Z, Y, X, t);
    const int64_t plane = (int64_t)Z * Y * X;
    const int64_t base = index[s] * C * plane;
    const int64_t off[8] = {((int64_t)t.z0 * Y + t.y0) * X + t.x0, ((int64_t)t.z0 * Y + t.y0) * X + t.x1,
                            ((int64_t)t.z0 * Y + t.y1) * X + t.x0, ((int64_t)t.z0 * Y + t.y1) * X + t.x1,
                            ((int64_t)t.z1 * Y + t.y0) * X + t.x0, ((int64_t)t.z1 * Y + t.y0) * X + t.x1,
                            ((int64_t)t.z1 * Y + t.y1) * X + t.x0, ((int64_t)t.z1 * Y + t.y1) * X + t.x1};
    float dw[8][3];
    if (grad_uv) tri_dweights(u, v, w, Z, Y, X, dw);
    float gu[3] = {0.f, 0.f, 0.f};
    for (int c = 0; c < C; ++c) {
        float go = grad_out[s * C + c];
        float* gg = grad_grid + base + c * plane;
#pragma unroll
        for (int k = 0; k < 8; ++k) unsafeAtomicAdd(gg + off[k], t.w[k] * go);
        if (grad_uv) {
            const float* gc = grid + base + c * plane;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                float val = gc[off[k]] * go;
                gu[0] += val * dw[k][0];
                gu[1] += val * dw[k][1];
                gu[2] += val * dw[k][2];
            }
        }
    }
    if (grad_uv) {
        grad_uv[s * 3 + 0] += gu[0] * ((float)(X - 1) / 2);
        grad_uv[s * 3 + 1] += gu[1] * ((float)(Y - 1) / 2);
        grad_uv[s * 3 + 2] += gu[2] * ((float)(Z - 1) / 2);
    }
}

// 2-D indirect gather, grid [N, C, H, W] (modules/IndirectGridSample2D.cu:22-162),
// corner setup in bil_setup (atdf_device.h).
__global__ __launch_bounds__(256) void grid_sample2d_fwd_kernel(const float* __restrict__ grid, int C, int H, int W,
                                                                const int64_t* __restrict__ index,
                                                                const float* __restrict__ uv, int64_t S,
                                                                float* __restrict__ out)
{
    // one lane per (sample, channel), as the 3-D forward
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= S * C) return;
    const int64_t s = i / C;
    const int c = (int)(i - s * C);
    Bil b;
    bil_setup(uv[s * 2], uv[s * 2 + 1], H, W, b);
    const int64_t plane = (int64_t)H * W;
    const float* gc = grid + (index[s] * C + c) * plane;
    const int64_t off[4] = {(int64_t)b.y0 * W + b.x0, (int64_t)b.y0 * W + b.x1, (int64_t)b.y1 * W + b.x0,
                            (int64_t)b.y1 * W + b.x1};
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc += gc[off[k]] * b.w[k];
    out[s * C + c] = acc;
}

// Backward: atomics into the clipped corners (IndirectGridSample2D.cu:175-393);
// grad_uv (nullable) += the analytic derivative of the forward.
__global__ __launch_bounds__(256) void grid_sample2d_bwd_kernel(const float* __restrict__ grid, int C, int H, int W,
                                                                const int64_t* __restrict__ index,
                                                                const float* __restrict__ uv, int64_t S,
                                                                const float* __restrict__ grad_out,
                                                                float* __restrict__ grad_grid,
                                                                float* __restrict__ grad_uv)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    const float u = uv[s * 2], v = uv[s * 2 + 1];
    Bil b;
    bil_setup(u, v, H, W, b);
    const int64_t plane = (int64_t)H * W;
    const int64_t base = index[s] * C * plane;
    const int64_t off[4] = {(int64_t)b.y0 * W + b.x0, (int64_t)b.y0 * W + b.x1, (int64_t)b.y1 * W + b.x0,
                            (int64_t)b.y1 * W + b.x1};
    const float ix = ((u + 1) / 2) * (float)(W - 1);
    const float iy = ((v + 1) / 2) * (float)(H - 1);
    const float fx[2] = {(float)((int)floorf(ix) + 1) - ix, ix - (float)(int)floorf(ix)};
    const float fy[2] = {(float)((int)floorf(iy) + 1) - iy, iy - (float)(int)floorf(iy)};
    float gx = 0.f, gy = 0.f;
    for (int c = 0; c < C; ++c) {
        const float go = grad_out[s * C + c];
        float* gg = grad_grid + base + c * plane;
#pragma unroll
        for (int k = 0; k < 4; ++k) unsafeAtomicAdd(gg + off[k], b.w[k] * go);
        if (grad_uv) {
            const float* gc = grid + base + c * plane;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int bx = k & 1, by = k >> 1;
                const float val = gc[off[k]] * go;
                gx += val * (bx ? fy[by] : -fy[by]);
                gy += val * (by ? fx[bx] : -fx[bx]);
            }
        }
    }
    if (grad_uv) {
        grad_uv[s * 2] += gx * ((float)(W - 1) / 2);
        grad_uv[s * 2 + 1] += gy * ((float)(H - 1) / 2);
    }
}

// 1-D indirect gather, grid [N, C, W, 1] (modules/IndirectGridSample1D.cu:20-140).
__global__ __launch_bounds__(256) void grid_sample1d_fwd_kernel(const float* __restrict__ grid, int C, int W,
                                                                const int64_t* __restrict__ index,
                                                                const float* __restrict__ uv, int64_t S,
                                                                float* __restrict__ out)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    Lin l;
    lin_setup(uv[s], W, l);
    const float w0 = l.w0, w1 = l.w1;
    const int c0 = l.x0, c1 = l.x1;
    const float* g = grid + index[s] * C * W;
    for (int c = 0; c < C; ++c) {
        float acc = 0.f;
        acc += g[c * W + c0] * w0;
        acc += g[c * W + c1] * w1;
        out[s * C + c] = acc;
    }
}

__global__ __launch_bounds__(256) void grid_sample1d_bwd_kernel(const float* __restrict__ grid, int C, int W,
                                                                const int64_t* __restrict__ index,
                                                                const float* __restrict__ uv, int64_t S,
                                                                const float* __restrict__ grad_out,
                                                                float* __restrict__ grad_grid,
                                                                float* __restrict__ grad_uv)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    Lin l;
    lin_setup(uv[s], W, l);
    const float w0 = l.w0, w1 = l.w1;
    const int c0 = l.x0, c1 = l.x1;
    const int64_t base = index[s] * C * W;
    float gx = 0.f;
    for (int c = 0; c < C; ++c) {
        const float go = grad_out[s * C + c];
        unsafeAtomicAdd(grad_grid + base + c * W + c0, w0 * go);
        unsafeAtomicAdd(grad_grid + base + c * W + c1, w1 * go);
        if (grad_uv) gx += (grid[base + c * W + c1] - grid[base + c * W + c0]) * go;
    }
    if (grad_uv) grad_uv[s] += gx * ((float)(W - 1) / 2);
}

// GradientCorrection backward (modules/GradientCorrection.cu:22-124):
// g - (g . d) d with d the direction of the sample's ray.
__global__ __launch_bounds__(256) void gradient_correction_kernel(const float* __restrict__ dir,
                                                                  const int64_t* __restrict__ ray, int64_t S,
                                                                  const float* __restrict__ grad_out,
                                                                  float* __restrict__ grad_pos)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int64_t r = ray[s];
    const float d0 = dir[r * 3], d1 = dir[r * 3 + 1], d2 = dir[r * 3 + 2];
    const float g0 = grad_out[s * 3], g1 = grad_out[s * 3 + 1], g2 = grad_out[s * 3 + 2];
    float dot = g0 * d0 + g1 * d1;
    dot = dot + g2 * d2;
    grad_pos[s * 3] = g0 - dot * d0;
    grad_pos[s * 3 + 1] = g1 - dot * d1;
    grad_pos[s * 3 + 2] = g2 - dot * d2;
}

// ComputeBoundWeight (structure/HyperTree.cu:962-993) with an O(1) node flag
// instead of a scan over the in-ROI list: out[s] = s if in_roi_flag[node[s]]
// else -1.
__global__ __launch_bounds__(256) void bound_weight_kernel(const int64_t* __restrict__ node,
                                                           const int32_t* __restrict__ flag, int64_t S,
                                                           int64_t* __restrict__ out)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    out[s] = flag[node[s]] ? s : -1;
}

__global__ __launch_bounds__(256) void integrate_kernel(const float* __restrict__ v, const float* __restrict__ w,
                                                        const int64_t* __restrict__ ray, int64_t S, int C, int64_t R,
                                                        float* __restrict__ out)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    int64_t r = ray[s];
    float ws = w[s];
    for (int c = 0; c < C; ++c) unsafeAtomicAdd(out + c * R + r, v[s * C + c] * ws);
}

__global__ __launch_bounds__(256) void loss_per_node_kernel(const float* __restrict__ prl, int C, int64_t R,
                                                            const int64_t* __restrict__ ray,
                                                            const float* __restrict__ w,
                                                            const int64_t* __restrict__ node, int64_t S,
                                                            double* __restrict__ loss_sum,
                                                            double* __restrict__ weight_sum)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= S) return;
    int64_t r = ray[s];
    double l = 0;
    for (int c = 0; c < C; ++c) l += (double)prl[c * R + r];
    l /= C;
    double ws = (double)w[s];
    unsafeAtomicAdd(loss_sum + node[s], l * ws);
    unsafeAtomicAdd(weight_sum + node[s], ws);
}

// ---------------------------------------------------------------------------
// MSE: per-block FP64 partial sums, then one block reduces them in a fixed
// order (deterministic).
// ---------------------------------------------------------------------------
constexpr int kRedThreads = 256;

__device__ __forceinline__ double block_sum_d(double v, double* smem)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) smem[wave] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kRedThreads / 64; ++w) t += smem[w];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kRedThreads) void mse_partial_kernel(const float* __restrict__ P,
                                                                  const float* __restrict__ T, int64_t R,
                                                                  float scale, float* __restrict__ dLdP,
                                                                  double* __restrict__ partial)
{
    __shared__ double smem[kRedThreads / 64];
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < R; i += (int64_t)gridDim.x * blockDim.x) {
        double diff = (double)P[i] - (double)T[i];
        acc += diff * diff;
        if (dLdP) dLdP[i] = scale * (2.0f * (P[i] - T[i]) / (float)R);
    }
    double t = block_sum_d(acc, smem);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(kRedThreads) void mse_final_kernel(const double* __restrict__ partial, int nb, int64_t R,
                                                                float scale, double* __restrict__ loss)
{
    __shared__ double smem[kRedThreads / 64];
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partial[i];
    double t = block_sum_d(acc, smem);
    if (threadIdx.x == 0) *loss = (double)scale * t / (double)R;
}

// ---------------------------------------------------------------------------
// Adam (torch::optim::Adam defaults: no weight decay, no amsgrad)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                                                   float b1, float b2, float eps, float bc1, float bc2_sqrt,
                                                   int zero_grad)
{
    int64_t i4 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * 4;
    if (i4 + 3 < n) {
        float4 gg = *reinterpret_cast<float4*>(g + i4);
        float4 mm = *reinterpret_cast<float4*>(m + i4);
        float4 vv = *reinterpret_cast<float4*>(v + i4);
        float4 pp = *reinterpret_cast<float4*>(p + i4);
        float* gp = &gg.x;
        float* mp = &mm.x;
        float* vp = &vv.x;
        float* pq = &pp.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) pq[k] = adam_update(pq[k], gp[k], mp[k], vp[k], lr, b1, b2, eps, bc1, bc2_sqrt);
        *reinterpret_cast<float4*>(m + i4) = mm;
        *reinterpret_cast<float4*>(v + i4) = vv;
        *reinterpret_cast<float4*>(p + i4) = pp;
        if (zero_grad) *reinterpret_cast<float4*>(g + i4) = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        for (int64_t i = i4; i < n; ++i) {
            float mi = m[i], vi = v[i];
            p[i] = adam_update(p[i], g[i], mi, vi, lr, b1, b2, eps, bc1, bc2_sqrt);
            m[i] = mi;
            v[i] = vi;
            if (zero_grad) g[i] = 0.f;
        }
    }
}

__global__ __launch_bounds__(256) void adam_bf16_kernel(float* __restrict__ p, uint16_t* __restrict__ pb,
                                                        float* __restrict__ g, float* __restrict__ m,
                                                        float* __restrict__ v, int64_t n, float lr, float b1,
                                                        float b2, float eps, float bc1, float bc2_sqrt,
                                                        int zero_grad)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float mi = m[i], vi = v[i];
    const float pi = adam_update(p[i], g[i], mi, vi, lr, b1, b2, eps, bc1, bc2_sqrt);
    m[i] = mi;
    v[i] = vi;
    p[i] = pi;
    pb[i] = f32_to_bf16_rne(pi);
    if (zero_grad) g[i] = 0.f;
}

__global__ __launch_bounds__(256) void layout_kernel(const float* __restrict__ src, float* __restrict__ dst, int64_t N,
                                                     int C, int G, int direction)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t cells = (int64_t)G * G * G;
    if (i >= N * C * cells) return;
    // i indexes the channels-last destination / source: [n][cell][c]
    int64_t c = i % C;
    int64_t cell = (i / C) % cells;
    int64_t n = i / (C * cells);
    int64_t ncz = (n * C + c) * cells + cell;
    if (direction == 0)
        dst[i] = src[ncz];
    else
        dst[ncz] = src[i];
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float* __restrict__ src, uint16_t* __restrict__ dst,
                                                      int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = f32_to_bf16_rne(src[i]);
}

}  // namespace atdf

using namespace atdf;

static unsigned blocks_for(int64_t n, int threads = 256) { return (unsigned)atdf::div_up(n, threads); }

extern "C" {

atdf_status atdf_local_samples(const float* pmin, const float* pmax, const float* g, const int64_t* node, int64_t S,
                               float* local, atdf_stream_t stream)
{
    ATDF_REQUIRE(S >= 0, "atdf_local_samples: negative size");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(pmin && pmax && g && node && local, "atdf_local_samples: null buffer");
    hipLaunchKernelGGL(local_samples_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), pmin, pmax, g, node, S,
                       local);
    return check_launch("atdf_local_samples");
}

atdf_status atdf_grid_sample3d_fwd(const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                                   const int64_t* index, const float* uv, int64_t S, float* out, atdf_stream_t stream)
{
    ATDF_REQUIRE(N > 0 && C > 0 && Z > 0 && Y > 0 && X > 0 && S >= 0, "atdf_grid_sample3d_fwd: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(grid && index && uv && out, "atdf_grid_sample3d_fwd: null buffer");
    hipLaunchKernelGGL(grid_sample3d_fwd_kernel, dim3(blocks_for(S * C)), dim3(256), 0, to_hip(stream), grid, C, Z, Y,
                       X, index, uv, S, out);
    return check_launch("atdf_grid_sample3d_fwd");
}

atdf_status atdf_grid_sample3d_bwd(const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                                   const int64_t* index, const float* uv, int64_t S, const float* grad_out,
                                   float* grad_grid, float* grad_uv, atdf_stream_t stream)
{
    ATDF_REQUIRE(N > 0 && C > 0 && Z > 0 && Y > 0 && X > 0 && S >= 0, "atdf_grid_sample3d_bwd: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(grid && index && uv && grad_out && grad_grid, "atdf_grid_sample3d_bwd: null buffer");
    hipLaunchKernelGGL(grid_sample3d_bwd_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), grid, C, Z, Y, X,
                       index, uv, S, grad_out, grad_grid, grad_uv);
    return check_launch("atdf_grid_sample3d_bwd");
}

atdf_status atdf_grid_sample2d_fwd(const float* grid, int64_t N, int32_t C, int32_t H, int32_t W,
                                   const int64_t* index, const float* uv, int64_t S, float* out, atdf_stream_t stream)
{
    ATDF_REQUIRE(N > 0 && C > 0 && H > 0 && W > 0 && S >= 0, "atdf_grid_sample2d_fwd: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(grid && index && uv && out, "atdf_grid_sample2d_fwd: null buffer");
    hipLaunchKernelGGL(grid_sample2d_fwd_kernel, dim3(blocks_for(S * C)), dim3(256), 0, to_hip(stream), grid, C, H, W,
                       index, uv, S, out);
    return check_launch("atdf_grid_sample2d_fwd");
}

atdf_status atdf_grid_sample2d_bwd(const float* grid, int64_t N, int32_t C, int32_t H, int32_t W,
                                   const int64_t* index, const float* uv, int64_t S, const float* grad_out,
                                   float* grad_grid, float* grad_uv, atdf_stream_t stream)
{
    ATDF_REQUIRE(N > 0 && C > 0 && H > 0 && W > 0 && S >= 0, "atdf_grid_sample2d_bwd: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(grid && index && uv && grad_out && grad_grid, "atdf_grid_sample2d_bwd: null buffer");
    hipLaunchKernelGGL(grid_sample2d_bwd_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), grid, C, H, W,
                       index, uv, S, grad_out, grad_grid, grad_uv);
    return check_launch("atdf_grid_sample2d_bwd");
}

atdf_status atdf_grid_sample1d_fwd(const float* grid, int64_t N, int32_t C, int32_t W, const int64_t* index,
                                   const float* uv, int64_t S, float* out, atdf_stream_t stream)
{
    ATDF_REQUIRE(N > 0 && C > 0 && W > 0 && S >= 0, "atdf_grid_sample1d_fwd: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(grid && index && uv && out, "atdf_grid_sample1d_fwd: null buffer");
    hipLaunchKernelGGL(grid_sample1d_fwd_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), grid, C, W, index,
                       uv, S, out);
    return check_launch("atdf_grid_sample1d_fwd");
}

atdf_status atdf_grid_sample1d_bwd(const float* grid, int64_t N, int32_t C, int32_t W, const int64_t* index,
                                   const float* uv, int64_t S, const float* grad_out, float* grad_grid,
                                   float* grad_uv, atdf_stream_t stream)
{
    ATDF_REQUIRE(N > 0 && C > 0 && W > 0 && S >= 0, "atdf_grid_sample1d_bwd: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(grid && index && uv && grad_out && grad_grid, "atdf_grid_sample1d_bwd: null buffer");
    hipLaunchKernelGGL(grid_sample1d_bwd_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), grid, C, W, index,
                       uv, S, grad_out, grad_grid, grad_uv);
    return check_launch("atdf_grid_sample1d_bwd");
}

atdf_status atdf_gradient_correction_bwd(const float* ray_direction, const int64_t* ray_index, int64_t S,
                                         const float* grad_out, float* grad_position, atdf_stream_t stream)
{
    ATDF_REQUIRE(S >= 0, "atdf_gradient_correction_bwd: negative sample count");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(ray_direction && ray_index && grad_out && grad_position, "atdf_gradient_correction_bwd: null buffer");
    hipLaunchKernelGGL(gradient_correction_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), ray_direction,
                       ray_index, S, grad_out, grad_position);
    return check_launch("atdf_gradient_correction_bwd");
}

atdf_status atdf_compute_bound_weight(const int64_t* node_id, int64_t S, const int32_t* in_roi_flag,
                                      int64_t* out, atdf_stream_t stream)
{
    ATDF_REQUIRE(S >= 0, "atdf_compute_bound_weight: negative sample count");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(node_id && in_roi_flag && out, "atdf_compute_bound_weight: null buffer");
    hipLaunchKernelGGL(bound_weight_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), node_id, in_roi_flag,
                       S, out);
    return check_launch("atdf_compute_bound_weight");
}

atdf_status atdf_integrate_xray(const float* values, const float* weight, const int64_t* ray_index, int64_t S,
                                int32_t C, int64_t R, float* out, atdf_stream_t stream)
{
    ATDF_REQUIRE(S >= 0 && C > 0 && R >= 0, "atdf_integrate_xray: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(values && weight && ray_index && out, "atdf_integrate_xray: null buffer");
    hipLaunchKernelGGL(integrate_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), values, weight, ray_index,
                       S, C, R, out);
    return check_launch("atdf_integrate_xray");
}

atdf_status atdf_accumulate_loss_per_node(const float* per_ray_loss, int32_t C, int64_t R, const int64_t* ray_index,
                                          const float* weight, const int64_t* node_id, int64_t S, int64_t num_nodes,
                                          double* loss_sum, double* weight_sum, atdf_stream_t stream)
{
    ATDF_REQUIRE(S >= 0 && C > 0 && num_nodes > 0, "atdf_accumulate_loss_per_node: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(per_ray_loss && ray_index && weight && node_id && loss_sum && weight_sum,
                 "atdf_accumulate_loss_per_node: null buffer");
    hipLaunchKernelGGL(loss_per_node_kernel, dim3(blocks_for(S)), dim3(256), 0, to_hip(stream), per_ray_loss, C, R,
                       ray_index, weight, node_id, S, loss_sum, weight_sum);
    return check_launch("atdf_accumulate_loss_per_node");
}

static int mse_blocks(int64_t R)
{
    int64_t b = atdf::div_up(R > 0 ? R : 1, kRedThreads * 4);
    return (int)(b < 1024 ? b : 1024);
}

size_t atdf_mse_workspace_bytes(int64_t num_rays) { return (size_t)mse_blocks(num_rays) * sizeof(double); }

atdf_status atdf_mse_loss(const float* P, const float* T, int64_t R, float scale, float* dLdP, double* loss,
                          void* workspace, size_t workspace_bytes, atdf_stream_t stream)
{
    ATDF_REQUIRE(R > 0 && P && T && loss && workspace, "atdf_mse_loss: bad arguments");
    ATDF_REQUIRE(workspace_bytes >= atdf_mse_workspace_bytes(R), "atdf_mse_loss: workspace too small");
    int nb = mse_blocks(R);
    hipStream_t s = to_hip(stream);
    double* partial = reinterpret_cast<double*>(workspace);
    hipLaunchKernelGGL(mse_partial_kernel, dim3(nb), dim3(kRedThreads), 0, s, P, T, R, scale, dLdP, partial);
    hipLaunchKernelGGL(mse_final_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, nb, R, scale, loss);
    return check_launch("atdf_mse_loss");
}

atdf_status atdf_adam_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n, float lr,
                           float beta1, float beta2, float eps, int64_t step, int32_t zero_grad, atdf_stream_t stream)
{
    ATDF_REQUIRE(n >= 0 && step >= 1, "atdf_adam_step: bad arguments");
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(param && grad && exp_avg && exp_avg_sq, "atdf_adam_step: null buffer");
    ATDF_REQUIRE(((uintptr_t)param | (uintptr_t)grad | (uintptr_t)exp_avg | (uintptr_t)exp_avg_sq) % 16 == 0,
                 "atdf_adam_step: buffers must be 16-byte aligned");
    float bc1 = 1.0f - (float)pow((double)beta1, (double)step);
    float bc2 = 1.0f - (float)pow((double)beta2, (double)step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(atdf::div_up(n, 4))), dim3(256), 0, to_hip(stream), param, grad,
                       exp_avg, exp_avg_sq, n, lr, beta1, beta2, eps, bc1, sqrtf(bc2), zero_grad);
    return check_launch("atdf_adam_step");
}

atdf_status atdf_adam_step_bf16(float* master, uint16_t* param_bf16, float* grad, float* exp_avg, float* exp_avg_sq,
                                int64_t n, float lr, float beta1, float beta2, float eps, int64_t step,
                                int32_t zero_grad, atdf_stream_t stream)
{
    ATDF_REQUIRE(n >= 0 && step >= 1, "atdf_adam_step_bf16: bad arguments");
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(master && param_bf16 && grad && exp_avg && exp_avg_sq, "atdf_adam_step_bf16: null buffer");
    float bc1 = 1.0f - (float)pow((double)beta1, (double)step);
    float bc2 = 1.0f - (float)pow((double)beta2, (double)step);
    hipLaunchKernelGGL(adam_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, to_hip(stream), master, param_bf16, grad,
                       exp_avg, exp_avg_sq, n, lr, beta1, beta2, eps, bc1, sqrtf(bc2), zero_grad);
    return check_launch("atdf_adam_step_bf16");
}

atdf_status atdf_grid_layout(const float* src, float* dst, int64_t N, int32_t C, int32_t G, int32_t direction,
                             atdf_stream_t stream)
{
    ATDF_REQUIRE(N >= 0 && C > 0 && G > 0 && (direction == 0 || direction == 1), "atdf_grid_layout: bad arguments");
    int64_t total = N * C * (int64_t)G * G * G;
    if (total == 0) return ATDF_OK;
    ATDF_REQUIRE(src && dst && src != dst, "atdf_grid_layout: null or aliased buffers");
    hipLaunchKernelGGL(layout_kernel, dim3(blocks_for(total)), dim3(256), 0, to_hip(stream), src, dst, N, C, G,
                       direction);
    return check_launch("atdf_grid_layout");
}

atdf_status atdf_f32_to_bf16(const float* src, uint16_t* dst, int64_t n, atdf_stream_t stream)
{
    ATDF_REQUIRE(n >= 0, "atdf_f32_to_bf16: bad size");
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(src && dst, "atdf_f32_to_bf16: null buffer");
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, to_hip(stream), src, dst, n);
    return check_launch("atdf_f32_to_bf16");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Slot-major adjoint of the indirect gathers (the backward of
// IndirectGridSample3D / 2D / 1D, modules/IndirectGridSample2D.cu:441-477
// pattern).  The per-sample kernels above add every corner with a global
// float atomic; the samples of one grid slot fall on the same few hundred
// cells, and a tensorial field's line takes every sample of the batch, so
// those atomics serialise.  Here the samples are counting-sorted by slot (LDS
// histograms and LDS ranks, one global atomic per slot and workgroup), cut
// into chunks of at most kGsChunk samples of one slot, and a workgroup per
// chunk accumulates w * grad_out into an LDS copy of the slot's gradient (LDS
// float atomics), then adds it to grad_grid once: a plain read-add-write when
// the chunk is the slot's only one, global atomics otherwise.  grad_uv is per
// sample and added in the same pass, with the per-sample kernels' arithmetic.
// ---------------------------------------------------------------------------
namespace atdf {

constexpr int kGsSortItems = 8192;  // samples per sort workgroup
constexpr int kGsHistMax = 8192;    // slots an LDS histogram holds
constexpr int kGsChunk = 4096;      // samples per accumulation chunk
constexpr int kGsBlock = 512;

struct GsSort {
    const int64_t* index;
    int64_t S, N;
    int32_t* cnt;     // [N]
    int64_t* off;     // [N + 1]
    int32_t* cursor;  // [N]
    int32_t* sorted;  // [S]
    int32_t* ccnt;    // [N]
    int64_t* coff;    // [N + 1]
    int4* desc;       // [chunks]: slot, begin, end, single chunk
};

__global__ __launch_bounds__(256) void gs_zero_kernel(int32_t* __restrict__ a, int32_t* __restrict__ b, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = 0;
        b[i] = 0;
    }
}

__global__ __launch_bounds__(256) void gs_hist_kernel(GsSort a)
{
    __shared__ int h[kGsHistMax];
    const bool lds = a.N <= kGsHistMax;
    if (lds)
        for (int i = threadIdx.x; i < a.N; i += 256) h[i] = 0;
    __syncthreads();
    const int64_t b0 = blockIdx.x * (int64_t)kGsSortItems;
    const int64_t b1 = b0 + kGsSortItems < a.S ? b0 + kGsSortItems : a.S;
    for (int64_t s = b0 + threadIdx.x; s < b1; s += 256) {
        const int64_t k = a.index[s];
        if (lds)
            atomicAdd(&h[k], 1);
        else
            atomicAdd(&a.cnt[k], 1);
    }
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i < a.N; i += 256)
            if (h[i]) atomicAdd(&a.cnt[i], h[i]);
}

__global__ __launch_bounds__(256) void gs_scatter_kernel(GsSort a)
{
    __shared__ int h[kGsHistMax];
    const bool lds = a.N <= kGsHistMax;
    const int64_t b0 = blockIdx.x * (int64_t)kGsSortItems;
    const int64_t b1 = b0 + kGsSortItems < a.S ? b0 + kGsSortItems : a.S;
    if (!lds) {
        for (int64_t s = b0 + threadIdx.x; s < b1; s += 256) {
            const int64_t k = a.index[s];
            a.sorted[a.off[k] + atomicAdd(&a.cursor[k], 1)] = (int32_t)s;
        }
        return;
    }
    for (int i = threadIdx.x; i < a.N; i += 256) h[i] = 0;
    __syncthreads();
    for (int64_t s = b0 + threadIdx.x; s < b1; s += 256) atomicAdd(&h[a.index[s]], 1);
    __syncthreads();
    // reserve this workgroup's range in every slot it touches
    for (int i = threadIdx.x; i < a.N; i += 256)
        if (h[i]) h[i] = atomicAdd(&a.cursor[i], h[i]);
    __syncthreads();
    for (int64_t s = b0 + threadIdx.x; s < b1; s += 256) {
        const int64_t k = a.index[s];
        a.sorted[a.off[k] + atomicAdd(&h[k], 1)] = (int32_t)s;
    }
}

__global__ __launch_bounds__(256) void gs_chunk_count_kernel(GsSort a)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < a.N) a.ccnt[i] = (a.cnt[i] + kGsChunk - 1) / kGsChunk;
}

__global__ __launch_bounds__(256) void gs_chunk_fill_kernel(GsSort a)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const int n = a.ccnt[i];
    const int64_t o = a.off[i], e = a.off[i + 1];
    for (int k = 0; k < n; ++k) {
        const int64_t b = o + (int64_t)k * kGsChunk;
        const int64_t f = b + kGsChunk < e ? b + kGsChunk : e;
        a.desc[a.coff[i] + k] = make_int4((int)i, (int)b, (int)f, n == 1 ? 1 : 0);
    }
}

struct GsAcc {
    const float* grid;
    const int64_t* index;
    const float* uv;
    const float* grad_out;
    float* grad_grid;
    float* grad_uv;
    const int32_t* sorted;
    const int4* desc;
    const int64_t* total_chunks;
    int C, Z, Y, X;  // D = 2: Y = H, X = W; D = 1: X = W
    int slot_size;   // C * cells
};

template <int D>
__global__ __launch_bounds__(kGsBlock) void gs_accumulate_kernel(GsAcc a)
{
    extern __shared__ __attribute__((aligned(16))) float lg[];
    const int C = a.C, Z = a.Z, Y = a.Y, X = a.X;
    const int cells = a.slot_size / C;
    const int64_t total = *a.total_chunks;
    for (int64_t c = blockIdx.x; c < total; c += gridDim.x) {
        const int4 ch = a.desc[c];
        for (int i = threadIdx.x; i < a.slot_size; i += kGsBlock) lg[i] = 0.f;
        __syncthreads();
        const int64_t base = (int64_t)ch.x * a.slot_size;
        const float* gslot = a.grid + base;
        for (int p = ch.y + threadIdx.x; p < ch.z; p += kGsBlock) {
            const int64_t s = a.sorted[p];
            const float* go_s = a.grad_out + s * C;
            if constexpr (D == 3) {
                const float u = a.uv[s * 3 + 0], v = a.uv[s * 3 + 1], w = a.uv[s * 3 + 2];
                Tri t;
                tri_setup(u, v, w, Z, Y, X, t);
                const int off[8] = {(t.z0 * Y + t.y0) * X + t.x0, (t.z0 * Y + t.y0) * X + t.x1,
                                    (t.z0 * Y + t.y1) * X + t.x0, (t.z0 * Y + t.y1) * X + t.x1,
                                    (t.z1 * Y + t.y0) * X + t.x0, (t.z1 * Y + t.y0) * X + t.x1,
                                    (t.z1 * Y + t.y1) * X + t.x0, (t.z1 * Y + t.y1) * X + t.x1};
                float dw[8][3];
                if (a.grad_uv) tri_dweights(u, v, w, Z, Y, X, dw);
                float gu[3] = {0.f, 0.f, 0.f};
                for (int cc = 0; cc < C; ++cc) {
                    const float go = go_s[cc];
                    float* l = lg + cc * cells;
#pragma unroll
                    for (int k = 0; k < 8; ++k) atomicAdd(l + off[k], t.w[k] * go);
                    if (a.grad_uv) {
                        const float* gc = gslot + (int64_t)cc * cells;
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const float val = gc[off[k]] * go;
                            gu[0] += val * dw[k][0];
                            gu[1] += val * dw[k][1];
                            gu[2] += val * dw[k][2];
                        }
                    }
                }
                if (a.grad_uv) {
                    a.grad_uv[s * 3 + 0] += gu[0] * ((float)(X - 1) / 2);
                    a.grad_uv[s * 3 + 1] += gu[1] * ((float)(Y - 1) / 2);
                    a.grad_uv[s * 3 + 2] += gu[2] * ((float)(Z - 1) / 2);
                }
            } else if constexpr (D == 2) {
                const float u = a.uv[s * 2], v = a.uv[s * 2 + 1];
                Bil b;
                bil_setup(u, v, Y, X, b);
                const int off[4] = {b.y0 * X + b.x0, b.y0 * X + b.x1, b.y1 * X + b.x0, b.y1 * X + b.x1};
                const float ix = ((u + 1) / 2) * (float)(X - 1);
                const float iy = ((v + 1) / 2) * (float)(Y - 1);
                const float fx[2] = {(float)((int)floorf(ix) + 1) - ix, ix - (float)(int)floorf(ix)};
                const float fy[2] = {(float)((int)floorf(iy) + 1) - iy, iy - (float)(int)floorf(iy)};
                float gx = 0.f, gy = 0.f;
                for (int cc = 0; cc < C; ++cc) {
                    const float go = go_s[cc];
                    float* l = lg + cc * cells;
#pragma unroll
                    for (int k = 0; k < 4; ++k) atomicAdd(l + off[k], b.w[k] * go);
                    if (a.grad_uv) {
                        const float* gc = gslot + (int64_t)cc * cells;
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const int bx = k & 1, by = k >> 1;
                            const float val = gc[off[k]] * go;
                            gx += val * (bx ? fy[by] : -fy[by]);
                            gy += val * (by ? fx[bx] : -fx[bx]);
                        }
                    }
                }
                if (a.grad_uv) {
                    a.grad_uv[s * 2] += gx * ((float)(X - 1) / 2);
                    a.grad_uv[s * 2 + 1] += gy * ((float)(Y - 1) / 2);
                }
            } else {
                const float ix = ((a.uv[s] + 1) / 2) * (float)(X - 1);
                const int x0 = (int)floorf(ix), x1 = x0 + 1;
                const float w0 = (float)x1 - ix, w1 = ix - (float)x0;
                const int c0 = clampi(x0, X), c1 = clampi(x1, X);
                float gx = 0.f;
                for (int cc = 0; cc < C; ++cc) {
                    const float go = go_s[cc];
                    atomicAdd(lg + cc * X + c0, w0 * go);
                    atomicAdd(lg + cc * X + c1, w1 * go);
                    if (a.grad_uv) gx += (gslot[cc * X + c1] - gslot[cc * X + c0]) * go;
                }
                if (a.grad_uv) a.grad_uv[s] += gx * ((float)(X - 1) / 2);
            }
        }
        __syncthreads();
        float* gg = a.grad_grid + base;
        if (ch.w) {
            for (int i = threadIdx.x; i < a.slot_size; i += kGsBlock) gg[i] += lg[i];
        } else {
            for (int i = threadIdx.x; i < a.slot_size; i += kGsBlock)
                if (lg[i] != 0.f) unsafeAtomicAdd(gg + i, lg[i]);
        }
        __syncthreads();
    }
}

struct GsLayout {
    size_t cnt, off, cursor, ccnt, coff, sorted, desc, scan, total;
};

size_t gs_align(size_t x) { return (x + 255) & ~(size_t)255; }

GsLayout gs_layout(int64_t N, int64_t S)
{
    GsLayout L{};
    size_t p = 0;
    auto take = [&](size_t bytes) {
        const size_t at = p;
        p += gs_align(bytes);
        return at;
    };
    L.cnt = take(sizeof(int32_t) * N);
    L.off = take(sizeof(int64_t) * (N + 1));
    L.cursor = take(sizeof(int32_t) * N);
    L.ccnt = take(sizeof(int32_t) * N);
    L.coff = take(sizeof(int64_t) * (N + 1));
    L.sorted = take(sizeof(int32_t) * S);
    L.desc = take(sizeof(int4) * (size_t)(S / kGsChunk + N + 1));
    L.scan = take(atdf_scan_workspace_bytes(N));
    L.total = p;
    return L;
}

int gs_max_lds()
{
    static int v = 0;
    if (v == 0) {
        int dev = 0, x = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || x <= 0) {
            (void)hipGetLastError();
            x = 64 * 1024;
        }
        v = x;
    }
    return v;
}

int gs_blocks()
{
    static int n = 0;
    if (n == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
            (void)hipGetLastError();
            cus = 256;
        }
        n = 2 * cus;
    }
    return n;
}

}  // namespace atdf

extern "C" {

int64_t atdf_grid_sample_bwd_max_slot_floats(void) { return gs_max_lds() / (int)sizeof(float); }

size_t atdf_grid_sample_bwd_workspace_bytes(int64_t num_slots, int64_t num_samples)
{
    if (num_slots <= 0 || num_samples <= 0) return 0;
    return gs_layout(num_slots, num_samples).total;
}

atdf_status atdf_grid_sample_bwd_slots(int32_t D, const float* grid, int64_t N, int32_t C, const int32_t* size,
                                       const int64_t* index, const float* uv, int64_t S, const float* grad_out,
                                       float* grad_grid, float* grad_uv, void* workspace, size_t workspace_bytes,
                                       atdf_stream_t stream)
{
    ATDF_REQUIRE(D >= 1 && D <= 3 && size, "atdf_grid_sample_bwd_slots: D must be 1, 2 or 3");
    int Z = 1, Y = 1, X = 1;
    if (D == 3) {
        Z = size[0];
        Y = size[1];
        X = size[2];
    } else if (D == 2) {
        Y = size[0];
        X = size[1];
    } else {
        X = size[0];
    }
    ATDF_REQUIRE(N > 0 && C > 0 && Z > 0 && Y > 0 && X > 0 && S >= 0, "atdf_grid_sample_bwd_slots: bad shape");
    if (S == 0) return ATDF_OK;
    ATDF_REQUIRE(grid && index && uv && grad_out && grad_grid, "atdf_grid_sample_bwd_slots: null buffer");
    ATDF_REQUIRE(S < (int64_t)INT32_MAX && N < (int64_t)INT32_MAX, "atdf_grid_sample_bwd_slots: %lld samples / %lld "
                 "slots, at most 2^31 - 1", (long long)S, (long long)N);
    const int64_t slot = (int64_t)C * Z * Y * X;
    ATDF_REQUIRE(slot * (int64_t)sizeof(float) <= gs_max_lds(), "atdf_grid_sample_bwd_slots: a slot's gradient "
                 "(%lld floats) exceeds the %d bytes of LDS a workgroup can hold; use the per-sample backward",
                 (long long)slot, gs_max_lds());
    const GsLayout L = gs_layout(N, S);
    ATDF_REQUIRE(workspace && workspace_bytes >= L.total, "atdf_grid_sample_bwd_slots: workspace of %zu bytes, %zu "
                 "needed", workspace_bytes, L.total);
    unsigned char* w = static_cast<unsigned char*>(workspace);
    GsSort a{};
    a.index = index;
    a.S = S;
    a.N = N;
    a.cnt = reinterpret_cast<int32_t*>(w + L.cnt);
    a.off = reinterpret_cast<int64_t*>(w + L.off);
    a.cursor = reinterpret_cast<int32_t*>(w + L.cursor);
    a.ccnt = reinterpret_cast<int32_t*>(w + L.ccnt);
    a.coff = reinterpret_cast<int64_t*>(w + L.coff);
    a.sorted = reinterpret_cast<int32_t*>(w + L.sorted);
    a.desc = reinterpret_cast<int4*>(w + L.desc);
    hipStream_t s = to_hip(stream);
    const unsigned nb = blocks_for(N), sb = (unsigned)div_up(S, kGsSortItems);
    hipLaunchKernelGGL(gs_zero_kernel, dim3(nb), dim3(256), 0, s, a.cnt, a.cursor, N);
    hipLaunchKernelGGL(gs_hist_kernel, dim3(sb), dim3(256), 0, s, a);
    atdf_status st = atdf_exclusive_scan_i32(a.cnt, N, a.off, w + L.scan, atdf_scan_workspace_bytes(N), stream);
    if (st) return st;
    hipLaunchKernelGGL(gs_scatter_kernel, dim3(sb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(gs_chunk_count_kernel, dim3(nb), dim3(256), 0, s, a);
    st = atdf_exclusive_scan_i32(a.ccnt, N, a.coff, w + L.scan, atdf_scan_workspace_bytes(N), stream);
    if (st) return st;
    hipLaunchKernelGGL(gs_chunk_fill_kernel, dim3(nb), dim3(256), 0, s, a);
    GsAcc g{};
    g.grid = grid;
    g.index = index;
    g.uv = uv;
    g.grad_out = grad_out;
    g.grad_grid = grad_grid;
    g.grad_uv = grad_uv;
    g.sorted = a.sorted;
    g.desc = a.desc;
    g.total_chunks = a.coff + N;
    g.C = C;
    g.Z = Z;
    g.Y = Y;
    g.X = X;
    g.slot_size = (int)slot;
    const size_t lds = sizeof(float) * (size_t)slot;
    const dim3 grid_dim((unsigned)gs_blocks()), block(kGsBlock);
#define ATDF_GS(DD) hipLaunchKernelGGL(gs_accumulate_kernel<DD>, grid_dim, block, lds, s, g)
    if (D == 3)
        ATDF_GS(3);
    else if (D == 2)
        ATDF_GS(2);
    else
        ATDF_GS(1);
#undef ATDF_GS
    return check_launch("atdf_grid_sample_bwd_slots");
}

}  // extern "C"

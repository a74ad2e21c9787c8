// atdf_regularize.hip -- total-variation regulariser of feature grids, loss
// and gradient in one pass.
//
// TVLoss::forward (geometry/geometry.h:31-81) on x [B, C, n_0 .. n_{D-1}]:
//   loss = C * sum_d mean_b( w_b * mean_{c, s}( |x[s] - x[s + e_d]| ) )
// so every forward difference of dimension d carries the coefficient
// coef_d = C / (B * count_d), count_d = C * (n_d - 1) * prod_{k != d} n_k.
// The gradient of element s is gathered, not scattered: for each dimension
// sgn(x[s] - x[s + e_d]) where s has a successor minus sgn(x[s - e_d] - x[s])
// where it has a predecessor (sgn(0) = 0, the subgradient libtorch's abs
// uses), so no two lanes write one element and no atomics are needed.  The
// loss is reduced in FP64: lane sums, a fixed-order tree per workgroup, and
// a second one-workgroup pass over the workgroup partials.
//
// One lane per element, grid-stride; HBM-bound (4 B read, 4 B read-modify-
// written per element; the neighbour reads hit L1/L2).
#include <hip/hip_runtime.h>

#include "atdf_device.h"
#include "atdf_internal.h"

namespace atdf {

constexpr int kTVBlock = 256;
constexpr int kTVMaxBlocks = 2048;

struct TVArgs {
    const float* x;
    float* grad;
    const float* weight;
    int64_t B, C, total;
    int D;
    int n[3];
    int64_t sb, sc, ss[3];
    double coef[3];   // C / (B * count_d)
    float gcoef[3];   // (float)(scale * coef_d)
    double* partial;  // [gridDim.x]
};

__device__ __forceinline__ float sgnf(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); }

__global__ __launch_bounds__(kTVBlock) void tv_kernel(TVArgs a)
{
    __shared__ double red[kTVBlock];
    double acc = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)kTVBlock + threadIdx.x; e < a.total;
         e += (int64_t)gridDim.x * kTVBlock) {
        int64_t r = e;
        int s[3] = {0, 0, 0};
        for (int d = a.D - 1; d >= 0; --d) {
            s[d] = (int)(r % a.n[d]);
            r /= a.n[d];
        }
        const int64_t c = r % a.C, b = r / a.C;
        int64_t addr = b * a.sb + c * a.sc;
        for (int d = 0; d < a.D; ++d) addr += (int64_t)s[d] * a.ss[d];
        const float v = a.x[addr];
        const float wb = a.weight ? a.weight[b] : 1.f;
        float g = 0.f;
        for (int d = 0; d < a.D; ++d) {
            float up = 0.f, dn = 0.f;
            if (s[d] + 1 < a.n[d]) {
                const float diff = v - a.x[addr + a.ss[d]];
                acc += (double)fabsf(diff) * a.coef[d] * (double)wb;
                up = sgnf(diff);
            }
            if (s[d] > 0) dn = sgnf(a.x[addr - a.ss[d]] - v);
            g = g + a.gcoef[d] * (up - dn);
        }
        if (a.grad) a.grad[addr] = a.grad[addr] + g * wb;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kTVBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kTVBlock) void tv_finish_kernel(const double* __restrict__ partial, int blocks,
                                                             double scale, double* __restrict__ loss)
{
    __shared__ double red[kTVBlock];
    double acc = 0.0;
    for (int i = threadIdx.x; i < blocks; i += kTVBlock) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kTVBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] * scale;
}

int tv_blocks(int64_t total) { return (int)std::min<int64_t>(div_up(total, kTVBlock), kTVMaxBlocks); }

// ---------------------------------------------------------------------------
// Ring means of the Fourier regulariser (geometry_tensor_qo.cpp:784-808) on
// img [Y, X, Z] (the shifted FFT magnitudes, one X x Z slice per y): ring i
// (1 <= i < X / 2) has n = (int)(2 pi i) angles of step s = (float)(360 / n),
// th_k = k s for k < m = ceil(360 / s) (torch::arange(0, 360, s), so m is n
// or n + 1), rows r_a = i cos(th_a pi / 180) + X / 2 and columns
// c_b = i sin(th_b pi / 180) + Z / 2, and its value is the mean over all m^2
// pairs (a, b) -- torch::meshgrid of the two coordinate lists, not the m
// points of the circle -- of the bilinear value at (r_a, c_b): grid_sample
// with align_corners and border padding, both coordinates normalised by
// X - 1 and mapped back onto X - 1 rows and Z - 1 columns, then clamped.
// Bilinear weights factor into a row weight and a column weight, so the
// mean is sum_{row, col} A_i[row] B_i[col] img[y, row, col] / m^2 with
// A_i[row] = sum_a (weight of row in r_a's interpolation) and B_i likewise:
// ring_weights_kernel builds A and B (FP64, each row summed over the angles
// in order) and the index range where each is non-zero, ring_means_kernel
// forms the bilinear form over that support in FP64 with a fixed-order
// reduction.  Deterministic.
// ---------------------------------------------------------------------------
constexpr int kRingMaxAngles = 8192;
constexpr int kRingBlock = 256;

struct RingGeom {
    int X, Z, rings;
    int64_t Y;
};

__host__ __device__ inline void ring_angles(int i, float& step, int& m)
{
    const int n = (int)(2.0 * 3.14159265358979323846 * (double)i);
    step = (float)(360.0 / (double)n);
    m = (int)ceil(360.0 / (double)step);
}

// Coordinate along one axis after the reference's float operations:
// th / 180 (times the float reciprocal), * pi, cos or sin, * i, + centre,
// / (X - 1) (reciprocal), * 2 - 1, then grid_sample's ((g + 1) / 2) * (size - 1)
// and the border clamp.
__device__ __forceinline__ float ring_coord(int k, float step, int i, int centre, bool use_sin, int X, int size)
{
    const float th = (float)k * step;
    const float a = (th * (1.0f / 180.0f)) * 3.14159265358979323846f;
    const float cs = use_sin ? sinf(a) : cosf(a);
    const float p = (float)i * cs + (float)centre;
    const float gn = (p * (1.0f / (float)(X - 1))) * 2.0f - 1.0f;
    float u = ((gn + 1.0f) / 2.0f) * (float)(size - 1);
    u = fminf((float)(size - 1), fmaxf(u, 0.0f));
    return u;
}

// grid: rings x 2 workgroups (axis 0 = rows over X, axis 1 = columns over Z);
// bounds[4 i + 2 axis + {0, 1}] = first / last index with a non-zero weight
// (the columns are scaled by (Z - 1) / (X - 1), so the support is not the
// ring's box around the centre when Z != X)
__global__ __launch_bounds__(kRingBlock) void ring_weights_kernel(RingGeom gm, double* __restrict__ A,
                                                                  double* __restrict__ B, int* __restrict__ bounds)
{
    __shared__ float coord[kRingMaxAngles];
    __shared__ int s_lo, s_hi;
    const int i = blockIdx.x >> 1, axis = blockIdx.x & 1;
    const int size = axis ? gm.Z : gm.X;
    double* W = (axis ? B + (int64_t)i * gm.Z : A + (int64_t)i * gm.X);
    if (i == 0) {
        for (int r = threadIdx.x; r < size; r += kRingBlock) W[r] = 0.0;
        if (threadIdx.x == 0) {
            bounds[4 * i + 2 * axis] = 0;
            bounds[4 * i + 2 * axis + 1] = -1;
        }
        return;
    }
    if (threadIdx.x == 0) {
        s_lo = size;
        s_hi = -1;
    }
    float step;
    int m;
    ring_angles(i, step, m);
    const int centre = axis ? gm.Z / 2 : gm.X / 2;
    for (int k = threadIdx.x; k < m; k += kRingBlock) coord[k] = ring_coord(k, step, i, centre, axis == 1, gm.X, size);
    __syncthreads();
    for (int r = threadIdx.x; r < size; r += kRingBlock) {
        double w = 0.0;
        for (int k = 0; k < m; ++k) {
            const float u = coord[k];
            const float f0 = floorf(u);
            const int r0 = (int)f0;
            if (r == r0) w += (double)((f0 + 1.0f) - u);
            else if (r == r0 + 1) w += (double)(u - f0);
        }
        W[r] = w;
        if (w != 0.0) {
            atomicMin(&s_lo, r);
            atomicMax(&s_hi, r);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bounds[4 * i + 2 * axis] = s_lo;
        bounds[4 * i + 2 * axis + 1] = s_hi;
    }
}

// grid: rings * Y workgroups, one per (ring, slice)
__global__ __launch_bounds__(kRingBlock) void ring_means_kernel(RingGeom gm, const float* __restrict__ img,
                                                                const double* __restrict__ A,
                                                                const double* __restrict__ B,
                                                                const int* __restrict__ bounds, float* __restrict__ out)
{
    __shared__ double red[kRingBlock];
    const int64_t blk = blockIdx.x;
    const int i = (int)(blk / gm.Y);
    const int64_t y = blk % gm.Y;
    if (i == 0) {
        // the reference leaves this row uninitialised (torch::empty); +inf
        // keeps its entries out of the masked means
        if (threadIdx.x == 0) out[y] = INFINITY;
        return;
    }
    const int r_lo = bounds[4 * i], r_hi = bounds[4 * i + 1];
    const int c_lo = bounds[4 * i + 2], c_hi = bounds[4 * i + 3];
    const int nc = c_hi - c_lo + 1;
    const int64_t ne = (r_hi < r_lo || nc <= 0) ? 0 : (int64_t)(r_hi - r_lo + 1) * nc;
    const double* Ai = A + (int64_t)i * gm.X;
    const double* Bi = B + (int64_t)i * gm.Z;
    const float* im = img + y * (int64_t)gm.X * gm.Z;
    double acc = 0.0;
    for (int64_t e = threadIdx.x; e < ne; e += kRingBlock) {
        const int r = r_lo + (int)(e / nc), c = c_lo + (int)(e % nc);
        acc += Ai[r] * Bi[c] * (double)im[(int64_t)r * gm.Z + c];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kRingBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float step;
        int m;
        ring_angles(i, step, m);
        out[(int64_t)i * gm.Y + y] = (float)(red[0] / ((double)m * (double)m));
    }
}

}  // namespace atdf

using namespace atdf;

extern "C" {

size_t atdf_tv_workspace_bytes(int64_t num_elements)
{
    return num_elements <= 0 ? 0 : (size_t)tv_blocks(num_elements) * sizeof(double);
}

atdf_status atdf_tv_loss(const float* x, const atdf_tv_shape* shape, const float* weight, double scale, double* loss,
                         float* grad, void* workspace, size_t workspace_bytes, atdf_stream_t stream)
{
    ATDF_REQUIRE(shape, "atdf_tv_loss: null shape");
    const atdf_tv_shape& sh = *shape;
    ATDF_REQUIRE(sh.D >= 1 && sh.D <= 3, "atdf_tv_loss: D must be 1, 2 or 3, got %d", sh.D);
    ATDF_REQUIRE(sh.B >= 1 && sh.C >= 1, "atdf_tv_loss: empty batch or channels");
    int64_t total = sh.B * sh.C;
    for (int d = 0; d < sh.D; ++d) {
        ATDF_REQUIRE(sh.n[d] >= 2, "atdf_tv_loss: spatial size %d of dimension %d (the mean over an empty "
                     "difference is undefined)", sh.n[d], d);
        total *= sh.n[d];
    }
    ATDF_REQUIRE(x && loss, "atdf_tv_loss: null buffer");
    const int blocks = tv_blocks(total);
    ATDF_REQUIRE(workspace && workspace_bytes >= (size_t)blocks * sizeof(double),
                 "atdf_tv_loss: workspace of %zu bytes, %zu needed", workspace_bytes, (size_t)blocks * sizeof(double));
    TVArgs a{};
    a.x = x;
    a.grad = grad;
    a.weight = weight;
    a.B = sh.B;
    a.C = sh.C;
    a.total = total;
    a.D = sh.D;
    a.sb = sh.stride_b;
    a.sc = sh.stride_c;
    for (int d = 0; d < 3; ++d) {
        a.n[d] = d < sh.D ? sh.n[d] : 1;
        a.ss[d] = d < sh.D ? sh.stride_s[d] : 0;
    }
    for (int d = 0; d < sh.D; ++d) {
        double count = (double)sh.C * (double)(sh.n[d] - 1);
        for (int k = 0; k < sh.D; ++k)
            if (k != d) count *= (double)sh.n[k];
        a.coef[d] = (double)sh.C / ((double)sh.B * count);
        a.gcoef[d] = (float)(scale * a.coef[d]);
    }
    a.partial = static_cast<double*>(workspace);
    hipStream_t s = to_hip(stream);
    hipLaunchKernelGGL(tv_kernel, dim3((unsigned)blocks), dim3(kTVBlock), 0, s, a);
    hipLaunchKernelGGL(tv_finish_kernel, dim3(1), dim3(kTVBlock), 0, s, a.partial, blocks, scale, loss);
    return check_launch("atdf_tv_loss");
}

size_t atdf_fourier_ring_workspace_bytes(int32_t X, int32_t Z)
{
    if (X < 2 || Z < 2) return 0;
    return sizeof(double) * (size_t)(X / 2) * (size_t)(X + Z) + sizeof(int) * 4 * (size_t)(X / 2);
}

atdf_status atdf_fourier_ring_means(const float* img, int64_t Y, int32_t X, int32_t Z, float* out, void* workspace,
                                    size_t workspace_bytes, atdf_stream_t stream)
{
    ATDF_REQUIRE(Y >= 1 && X >= 2 && Z >= 2, "atdf_fourier_ring_means: image [%lld, %d, %d] too small", (long long)Y,
                 X, Z);
    ATDF_REQUIRE(img && out, "atdf_fourier_ring_means: null buffer");
    const int rings = X / 2;
    float step;
    int m = 0;
    if (rings > 1) ring_angles(rings - 1, step, m);
    ATDF_REQUIRE(m <= kRingMaxAngles, "atdf_fourier_ring_means: %d angles on the outer ring, at most %d (X <= %d)", m,
                 kRingMaxAngles, 2 * (int)(kRingMaxAngles / 6.2832));
    ATDF_REQUIRE((int64_t)rings * Y < (int64_t)INT32_MAX, "atdf_fourier_ring_means: too many slices");
    const size_t need = atdf_fourier_ring_workspace_bytes(X, Z);
    ATDF_REQUIRE(workspace && workspace_bytes >= need, "atdf_fourier_ring_means: workspace of %zu bytes, %zu needed",
                 workspace_bytes, need);
    RingGeom gm{X, Z, rings, Y};
    double* A = static_cast<double*>(workspace);
    double* B = A + (size_t)rings * X;
    int* bounds = reinterpret_cast<int*>(B + (size_t)rings * Z);
    hipStream_t s = to_hip(stream);
    hipLaunchKernelGGL(ring_weights_kernel, dim3((unsigned)(2 * rings)), dim3(kRingBlock), 0, s, gm, A, B, bounds);
    hipLaunchKernelGGL(ring_means_kernel, dim3((unsigned)((int64_t)rings * Y)), dim3(kRingBlock), 0, s, gm, img, A, B,
                       bounds, out);
    return check_launch("atdf_fourier_ring_means");
}

}  // extern "C"

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

}  // extern "C"

// atdf_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// Everything here that feeds traversal, sample placement, local coordinates
// or the trilinear weights is evaluated in the written order with no FMA
// contraction (the library is compiled with -ffp-contract=off), so those
// values are bit-identical to a C restatement compiled the same way.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace atdf {

// Packed node used by the traversal kernels (48 B, three 16-byte loads).
struct alignas(16) PNode {
    float mn[3];
    int32_t first_child;  // -1 for a leaf
    float mx[3];
    int32_t parent;       // -1 for the root
    int32_t flags;        // bit 0: active, bit 1: has an active descendant
    int32_t first_sib;    // first child of the parent (0 for the root)
    int32_t pad[2];
};
static_assert(sizeof(PNode) == 48, "PNode must stay 48 bytes");

constexpr int kFlagActive = 1;
constexpr int kFlagActiveBelow = 2;
// the quad-tree representation: four children, bit 0 selects the upper half
// in x, bit 1 in z (every node spans the same y range)
constexpr int kFlagQuad = 4;

__device__ __forceinline__ int child_count(const PNode& nd) { return (nd.flags & kFlagQuad) ? 4 : 8; }

// Child-index bits of the near half along the ray, so child k ^ mask for
// k = 0, 1, ... visits the children front to back.
__device__ __forceinline__ int near_child_mask(const float* d, bool quad)
{
    return quad ? ((d[0] < 0 ? 1 : 0) | (d[2] < 0 ? 2 : 0)) : ((d[0] < 0 ? 1 : 0) | (d[1] < 0 ? 2 : 0) | (d[2] < 0 ? 4 : 0));
}

// Slab test of structure/HyperTree.h:338-380: float divisions, running
// near/far values kept in double, the parallel-axis test asymmetric
// (o < min or o >= max misses).
__device__ __forceinline__ bool intersect_box_ray(const float* bmin, const float* bmax, const float* o,
                                                  const float* d, float& tmin, float& tmax)
{
    double t_near = -236436436;
    double t_far = 43637575;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (d[i] == 0) {
            if ((o[i] < bmin[i]) || (o[i] >= bmax[i])) {
                tmin = 0;
                tmax = 0;
                return false;
            }
        } else {
            float T_1 = (bmin[i] - o[i]) / d[i];
            float T_2 = (bmax[i] - o[i]) / d[i];
            if (T_1 > T_2) {
                float tmp = T_1;
                T_1 = T_2;
                T_2 = tmp;
            }
            if (T_1 > t_near) t_near = T_1;
            if (T_2 < t_far) t_far = T_2;
            if ((t_near > t_far) || (t_far < 0)) {
                tmin = 0;
                tmax = 0;
                return false;
            }
        }
    }
    tmin = (float)t_near;
    tmax = (float)t_far;
    return true;
}

__device__ __forceinline__ int iceil_f(float v)
{
    int iv = (int)v;
    return (v > (float)iv) ? iv + 1 : iv;
}

// Samples of one hit node: n = ceil((tmax - tmin) / |max - min| * K)
// (structure/HyperTree.cu:350-356).
__device__ __forceinline__ int node_sample_count(const float* bmin, const float* bmax, float tmin, float tmax, int K)
{
    float dx = bmax[0] - bmin[0];
    float dy = bmax[1] - bmin[1];
    float dz = bmax[2] - bmin[2];
    float diag = sqrtf(dx * dx + dy * dy + dz * dz);
    float dis = tmax - tmin;
    float rel = dis / diag;
    return iceil_f(rel * (float)K);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Stratified jitter, 24-bit uniform in [0,1), keyed by (seed, ray, index in ray).
__device__ __forceinline__ float hash_jitter(uint64_t seed, int64_t ray, int64_t local_index)
{
    uint64_t key = ((uint64_t)ray << 24) ^ (uint64_t)local_index;
    uint64_t h = splitmix64(seed ^ splitmix64(key));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// Uniform in [0,1) for point `point` of node `node` (random node samples).
__device__ __forceinline__ float point_uniform(uint64_t seed, int64_t node, int64_t point, int axis)
{
    uint64_t key = ((uint64_t)node << 34) ^ ((uint64_t)point << 2) ^ (uint64_t)axis;
    uint64_t h = splitmix64(seed ^ splitmix64(key ^ 0x5bd1e995ull));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// t of sample j in [tmin, tmax] split into n strata (structure/HyperTree.cu:466-474).
__device__ __forceinline__ float sample_t(float tmin, float step, int j, float a)
{
    float t1 = tmin + (float)j * step;
    float t2 = tmin + (float)(j + 1) * step;
    return t1 * (1 - a) + t2 * a;
}

// Midpoint-rule weight of sample j (structure/HyperTree.cu:495-512).
__device__ __forceinline__ float sample_weight(float tmin, float tmax, int n, int j, float t_prev, float t,
                                               float t_next)
{
    float lo = (j == 0) ? tmin : (t_prev + t) * 0.5f;
    float hi = (j == n - 1) ? tmax : (t_next + t) * 0.5f;
    return hi - lo;
}

__device__ __forceinline__ int clampi(int v, int lim)
{
    int a = v > 0 ? v : 0;
    return a < (lim - 1) ? a : (lim - 1);
}

// Trilinear setup with align_corners=True and clamped corners
// (modules/IndirectGridSample2D.cu:40-91 extended to 3-D).  Weights are in
// torch grid_sample order tnw tne tsw tse bnw bne bsw bse (t/b = z0/z1,
// n/s = y0/y1, w/e = x0/x1), computed from the unclamped corner indices.
struct Tri {
    int x0, x1, y0, y1, z0, z1;
    float w[8];
};

// Lattice coordinates (ix, iy, iz) of a local point in [-1, 1]^3.
__device__ __forceinline__ float tri_index(float u, int n)
{
    return ((u + 1) / 2) * (float)(n - 1);
}

// Corners and weights from lattice coordinates (the second half of tri_setup).
__device__ __forceinline__ void tri_from_index(float ix, float iy, float iz, int Z, int Y, int X, Tri& t)
{
    int ix0 = (int)floorf(ix), iy0 = (int)floorf(iy), iz0 = (int)floorf(iz);
    int ix1 = ix0 + 1, iy1 = iy0 + 1, iz1 = iz0 + 1;
    float fx1 = (float)ix1 - ix, fx0 = ix - (float)ix0;
    float fy1 = (float)iy1 - iy, fy0 = iy - (float)iy0;
    float fz1 = (float)iz1 - iz, fz0 = iz - (float)iz0;
    t.w[0] = fx1 * fy1 * fz1;
    t.w[1] = fx0 * fy1 * fz1;
    t.w[2] = fx1 * fy0 * fz1;
    t.w[3] = fx0 * fy0 * fz1;
    t.w[4] = fx1 * fy1 * fz0;
    t.w[5] = fx0 * fy1 * fz0;
    t.w[6] = fx1 * fy0 * fz0;
    t.w[7] = fx0 * fy0 * fz0;
    t.x0 = clampi(ix0, X);
    t.x1 = clampi(ix1, X);
    t.y0 = clampi(iy0, Y);
    t.y1 = clampi(iy1, Y);
    t.z0 = clampi(iz0, Z);
    t.z1 = clampi(iz1, Z);
}

__device__ __forceinline__ void tri_setup(float u, float v, float s, int Z, int Y, int X, Tri& t)
{
    tri_from_index(tri_index(u, X), tri_index(v, Y), tri_index(s, Z), Z, Y, X, t);
}

// Bilinear read of IndirectGridSample2D, grid [N, C, H, W]
// (modules/IndirectGridSample2D.cu:22-162):
// corners tnw tne tsw tse from floor(ix), floor(iy), weights from the
// unclipped corners, indices clipped (border padding), summed in that order.
struct Bil {
    int x0, x1, y0, y1;
    float w[4];
};

__device__ __forceinline__ void bil_setup(float u, float v, int H, int W, Bil& b)
{
    const float ix = ((u + 1) / 2) * (float)(W - 1);
    const float iy = ((v + 1) / 2) * (float)(H - 1);
    const int x0 = (int)floorf(ix), y0 = (int)floorf(iy);
    const int x1 = x0 + 1, y1 = y0 + 1;
    const float fx1 = (float)x1 - ix, fx0 = ix - (float)x0;
    const float fy1 = (float)y1 - iy, fy0 = iy - (float)y0;
    b.w[0] = fx1 * fy1;
    b.w[1] = fx0 * fy1;
    b.w[2] = fx1 * fy0;
    b.w[3] = fx0 * fy0;
    b.x0 = clampi(x0, W);
    b.x1 = clampi(x1, W);
    b.y0 = clampi(y0, H);
    b.y1 = clampi(y1, H);
}

// 1-D read of IndirectGridSample1D (modules/IndirectGridSample1D.cu:20-140):
// weights from the unclipped corners, indices clipped.
struct Lin {
    int x0, x1;
    float w0, w1;
};

__device__ __forceinline__ void lin_setup(float u, int W, Lin& l)
{
    const float ix = ((u + 1) / 2) * (float)(W - 1);
    const int x0 = (int)floorf(ix), x1 = x0 + 1;
    l.w0 = (float)x1 - ix;
    l.w1 = ix - (float)x0;
    l.x0 = clampi(x0, W);
    l.x1 = clampi(x1, W);
}

// Partial derivatives of the 8 weights w.r.t. (ix, iy, iz).
__device__ __forceinline__ void tri_dweights(float u, float v, float s, int Z, int Y, int X, float dw[8][3])
{
    float ix = ((u + 1) / 2) * (float)(X - 1);
    float iy = ((v + 1) / 2) * (float)(Y - 1);
    float iz = ((s + 1) / 2) * (float)(Z - 1);
    int ix0 = (int)floorf(ix), iy0 = (int)floorf(iy), iz0 = (int)floorf(iz);
    float fx[2] = {(float)(ix0 + 1) - ix, ix - (float)ix0};
    float fy[2] = {(float)(iy0 + 1) - iy, iy - (float)iy0};
    float fz[2] = {(float)(iz0 + 1) - iz, iz - (float)iz0};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
        float sx = bx ? 1.f : -1.f, sy = by ? 1.f : -1.f, sz = bz ? 1.f : -1.f;
        dw[c][0] = sx * fy[by] * fz[bz];
        dw[c][1] = fx[bx] * sy * fz[bz];
        dw[c][2] = fx[bx] * fy[by] * sz;
    }
}

// softplus(y, beta) with torch's threshold 20 (geometry/geometry.cpp:210-213).
__device__ __forceinline__ float softplus_f(float y, float beta)
{
    float yb = y * beta;
    if (yb > 20.0f) return y;
    return log1pf(expf(yb)) / beta;
}

__device__ __forceinline__ float softplus_grad_f(float y, float beta)
{
    float yb = y * beta;
    if (yb > 20.0f) return 1.0f;
    float z = expf(yb);
    return z / (z + 1.0f);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h)
{
    return __uint_as_float(((uint32_t)h) << 16);
}

// Active node holding p, or -1: from the root, step into the first child (in
// child order) whose closed box holds p (ActiveNodeIdForGlobalPosition,
// structure/HyperTree.cu:1339-1399).
// nchild = 0 descends into every child of a node (eight, or four in the
// quad-tree representation); nchild = 4 descends only into the first four
// children of every node, as ActiveNodeIdForGlobalPosition with use_quad
// (structure/HyperTree.cu:1340-1400) does on any tree.
__device__ __forceinline__ int32_t active_node_for_position(const PNode* __restrict__ nodes, const float* p,
                                                            int nchild = 0)
{
    int32_t node = 0;
    while (true) {
        const PNode nd = nodes[node];
        if (nd.flags & kFlagActive) return node;
        if (nd.first_child < 0) return -1;
        int32_t next = -1;
        const int nc = nchild > 0 ? nchild : child_count(nd);
        for (int c = 0; c < nc; ++c) {
            const PNode ch = nodes[nd.first_child + c];
            if (ch.mn[0] <= p[0] && ch.mx[0] >= p[0] && ch.mn[1] <= p[1] && ch.mx[1] >= p[1] && ch.mn[2] <= p[2] &&
                ch.mx[2] >= p[2]) {
                next = nd.first_child + c;
                break;
            }
        }
        if (next < 0) return -1;
        node = next;
    }
}

// One torch::optim::Adam update of a parameter (bias corrections bc1 and
// sqrt(bc2) precomputed); m and v are updated in place, returns the new value.
__device__ __forceinline__ float adam_update(float p, float g, float& m, float& v, float lr, float b1, float b2,
                                             float eps, float bc1, float bc2_sqrt)
{
    m = b1 * m + (1.f - b1) * g;
    v = b2 * v + (1.f - b2) * g * g;
    const float denom = sqrtf(v) / bc2_sqrt + eps;
    return p - (lr / bc1) * (m / denom);
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan
        return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    }
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

}  // namespace atdf

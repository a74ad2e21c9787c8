// atdf_project.hip -- fused per-ray projection and its adjoint.
//
// Work unit: one wavefront owns `rays_per_wave` consecutive rays and walks
// each ray's samples in tiles of 16.  Samples are recomputed from the ray's
// node intervals (never stored): lane (g, j) = (lane / 16, lane % 16) handles
// sample j of the tile and channels [CPL*g, CPL*g + CPL) of its gather, so
// the four lanes of a column read one contiguous channels-last corner record.
//
// Decoder C -> 64 -> 64 -> 1 on v_mfma_f32_16x16x4f32 with activations kept
// in the MFMA result layout: samples are the N (column) index, hidden units
// the M (row) index.  The result layout of one layer is directly the B
// operand of the next: lane (g, j) holds rows 4g + r of column j, and the
// K-step kk of a 16-row block uses rows {4g + kk}, so no data moves between
// layers.  Weights are the A operands, read from LDS.
//
// The adjoint recomputes the forward per tile and accumulates
//   dW2 += dH2 . H1^T and dW1 += dH1 . X^T
// on MFMA with the tile's samples as the K index (operands transposed
// through a per-wave LDS buffer), dX = W1^T . dH1 on MFMA, and scatters
// w_corner * dX into an LDS copy of the node's channels-last grid gradient
// (node-major order, below), which is added to HBM once per chunk.
// Per-wave parameter-gradient partials go to a workspace and are summed in a
// fixed order by a second kernel (deterministic decoder gradients).
//
// The BF16 decoder variant runs the same contractions on
// v_mfma_f32_16x16x16_bf16 / v_mfma_f32_16x16x32_bf16 with FP32 accumulation
// (a K = 64 contraction is two 16x16x32 steps; a K = 16 one is one 16x16x16).
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>

#include "atdf_device.h"
#include "atdf_internal.h"

namespace atdf {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ s4 pack_bf4(float a, float b, float c, float d)
{
    bf4 v = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
    return __builtin_bit_cast(s4, v);
}

__device__ __forceinline__ bf8 pack_bf8(const v4f& lo, const v4f& hi)
{
    bf8 v = {(__bf16)lo[0], (__bf16)lo[1], (__bf16)lo[2], (__bf16)lo[3],
             (__bf16)hi[0], (__bf16)hi[1], (__bf16)hi[2], (__bf16)hi[3]};
    return v;
}

// BF16 inputs, FP32 accumulation: K = 16 (4 per lane) and K = 32 (8 per lane).
__device__ __forceinline__ v4f mfma_bf16_k16(s4 a, s4 b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4f mfma_bf16_k32(bf8 a, bf8 b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Compiler + wavefront ordering point for LDS data exchanged between lanes
// of one wave (LDS executes a wave's operations in issue order).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

constexpr int kH = 64;        // hidden width
constexpr int kTile = 16;     // samples per tile
constexpr int kWaves = 4;     // waves per block
constexpr int kW2S = kH + 4;  // padded LDS row stride of W2
constexpr int kTS = kTile + 4;  // padded row stride of the transpose buffers

struct ProjArgs {
    const PNode* nodes;
    const float* O;
    const float* D;
    int64_t R;
    int jitter_mode;
    uint64_t seed;
    const int64_t* iv_off;
    const int32_t* iv_node;
    const float* iv_tmin;
    const float* iv_tmax;
    const int32_t* iv_count;
    const int32_t* iv_first;
    const void* grid;
    const int32_t* node_to_grid;
    const float* mlp;
    int G;
    float beta;
    int rpw;
    float* bins;  // [R, 32] mass histogram (binned forward only)
    // tensorial (VM) field of the in-ROI nodes (atdf_vm); planes null = none
    struct {
        const float* planes;         // [num_planes, 3, Gp, Gp, C]
        const float* lines;          // [3, Gl, C]
        const int32_t* node_to_plane;
        int64_t NG;                  // explicit grid slots; plane slot p sorts as NG + p
        int Gp, Gl;
        float rx0, rx1;              // ROI x range of the line x coordinate
    } vm;
};

template <int C>
struct MlpSmem {
    static constexpr int W1S = C + 4;
    static constexpr int kW1 = 0;
    static constexpr int kW2 = kW1 + kH * W1S;
    static constexpr int kB1 = kW2 + kH * kW2S;
    static constexpr int kB2 = kB1 + kH;
    static constexpr int kW3 = kB2 + kH;
    static constexpr int kB3 = kW3 + kH;
    static constexpr int kTotal = kB3 + 4;
};

// MlpSmem plus transposed copies of W2 (W2T[i][o] = W2[o][i]) and W1 (W1T[c][h]
// = W1[h][c], 16 rows, rows c >= C zero) for float4 operand reads in the
// FP32 decoder backward.
template <int C>
struct MlpSmemT {
    static constexpr int kW2T = MlpSmem<C>::kTotal;
    static constexpr int kW1T = kW2T + kH * kW2S;
    static constexpr int kTotal = kW1T + 16 * kW2S;
};

template <int C>
__device__ __forceinline__ void load_mlp_to_lds(const float* __restrict__ mlp, float* s)
{
    using L = MlpSmem<C>;
    const float* W1 = mlp;
    const float* b1 = W1 + kH * C;
    const float* W2 = b1 + kH;
    const float* b2 = W2 + kH * kH;
    const float* W3 = b2 + kH;
    const float* b3 = W3 + kH;
    for (int i = threadIdx.x; i < kH * C; i += blockDim.x) s[L::kW1 + (i / C) * L::W1S + (i % C)] = W1[i];
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) s[L::kW2 + (i / kH) * kW2S + (i % kH)] = W2[i];
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        s[L::kB1 + i] = b1[i];
        s[L::kB2 + i] = b2[i];
        s[L::kW3 + i] = W3[i];
    }
    if (threadIdx.x == 0) s[L::kB3] = b3[0];
}

template <int C>
__device__ __forceinline__ void load_mlp_to_lds_t(const float* __restrict__ mlp, float* s)
{
    using T = MlpSmemT<C>;
    load_mlp_to_lds<C>(mlp, s);
    const float* W1 = mlp;
    const float* W2 = W1 + kH * C + kH;
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) s[T::kW2T + (i % kH) * kW2S + i / kH] = W2[i];
    for (int i = threadIdx.x; i < 16 * kH; i += blockDim.x) {
        const int c = i / kH, h = i % kH;
        s[T::kW1T + c * kW2S + h] = c < C ? W1[h * C + c] : 0.f;
    }
}

// Per-lane sample state of one tile.
struct SampleCtx {
    bool valid;
    float t;       // ray parameter
    float w;       // integration weight
    Tri tri;
    float li[3];   // lattice coordinates of the sample in its node grid
    int64_t gbase; // element offset of the node grid, -1 if none
    // tensorial field only: local coordinates, global x and z, plane slot
    float loc[3];
    float gx, gz;
    int pslot;
};

// Recompute sample jj of a node interval (ray r, sample index s_local in the ray).
template <bool kVM = false>
__device__ __forceinline__ void make_sample_iv(const ProjArgs& a, int64_t r, const float* o, const float* d, int node,
                                               float tmin, float tmax, int n, int jj, int s_local, int C,
                                               SampleCtx& sc)
{
    const float step = (tmax - tmin) / (float)n;
    float aj = 0.5f, ap = 0.5f, an = 0.5f;
    if (a.jitter_mode == 2) {
        aj = hash_jitter(a.seed, r, s_local);
        if (jj > 0) ap = hash_jitter(a.seed, r, s_local - 1);
        if (jj < n - 1) an = hash_jitter(a.seed, r, s_local + 1);
    }
    const float t = sample_t(tmin, step, jj, aj);
    const float tp = (jj > 0) ? sample_t(tmin, step, jj - 1, ap) : 0.f;
    const float tn = (jj < n - 1) ? sample_t(tmin, step, jj + 1, an) : 0.f;
    sc.w = sample_weight(tmin, tmax, n, jj, tp, t, tn);
    sc.t = t;
    const PNode nd = a.nodes[node];
    float loc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float pos = o[k] + d[k] * t;
        float size = nd.mx[k] - nd.mn[k];
        loc[k] = ((pos - nd.mn[k]) / size) * 2 - 1;
        if constexpr (kVM) {
            sc.loc[k] = loc[k];
            if (k == 0) sc.gx = pos;
            if (k == 2) sc.gz = pos;
        }
    }
    tri_setup(loc[0], loc[1], loc[2], a.G, a.G, a.G, sc.tri);
    const int slot = a.node_to_grid[node];
    sc.gbase = slot >= 0 ? (int64_t)slot * a.G * a.G * a.G * C : -1;
    if constexpr (kVM) sc.pslot = slot < 0 ? a.vm.node_to_plane[node] : -1;
}

// Index (0..16) of the lane group's candidate interval holding sample t:
// the number of candidates whose exclusive sample end is <= t.  `ends` is
// nondecreasing over the 16 lanes of the group (lanes past the list hold
// INT_MAX); a branch-free binary search with five cross-lane reads.
__device__ __forceinline__ int group_search(int ends, int t, int grp)
{
    // four halving steps count the first 15 candidates, the last read the 16th
    int pos = 0;
#pragma unroll
    for (int s = 8; s > 0; s >>= 1) {
        const int e = __shfl(ends, grp | (pos + s - 1), 64);
        if (e <= t) pos += s;
    }
    const int e15 = __shfl(ends, grp | 15, 64);
    return e15 <= t ? 16 : pos;
}

// Inclusive prefix sum over each 16-lane row (DPP row shifts; lanes shifted
// in from outside the row read 0).
__device__ __forceinline__ int scan16_inclusive(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true);  // row_shr:8
    return v;
}

// Whether another lane of the same 16-lane row holds `key` (DPP row
// rotations 1..15, no LDS traffic).
__device__ __forceinline__ bool dup16(int key)
{
    bool dup = false;
#define ATDF_ROR(n) dup |= (__builtin_amdgcn_update_dpp(0, key, 0x120 + (n), 0xF, 0xF, false) == key)
    ATDF_ROR(1); ATDF_ROR(2); ATDF_ROR(3); ATDF_ROR(4); ATDF_ROR(5); ATDF_ROR(6); ATDF_ROR(7); ATDF_ROR(8);
    ATDF_ROR(9); ATDF_ROR(10); ATDF_ROR(11); ATDF_ROR(12); ATDF_ROR(13); ATDF_ROR(14); ATDF_ROR(15);
#undef ATDF_ROR
    return dup;
}

// Recompute sample `s_local` of ray r (lane-private interval cursor `cur`).
__device__ __forceinline__ void make_sample(const ProjArgs& a, int64_t r, const float* o, const float* d, int ns,
                                            int s_local, int64_t h1, int64_t& cur, int C, SampleCtx& sc)
{
    sc.valid = s_local < ns;
    sc.w = 0.f;
    sc.gbase = -1;
    if (!sc.valid) return;
    while (cur + 1 < h1 && a.iv_first[cur] + a.iv_count[cur] <= s_local) ++cur;
    make_sample_iv(a, r, o, d, a.iv_node[cur], a.iv_tmin[cur], a.iv_tmax[cur], a.iv_count[cur],
                   s_local - a.iv_first[cur], s_local, C, sc);
}

__device__ __forceinline__ int64_t corner_cell(const Tri& t, int k, int G)
{
    int x = (k & 1) ? t.x1 : t.x0;
    int y = (k & 2) ? t.y1 : t.y0;
    int z = (k & 4) ? t.z1 : t.z0;
    return ((int64_t)z * G + y) * G + x;
}

// Gather CPL channels starting at c0 (canonical corner order per channel,
// accumulated as a fused multiply-add chain).
template <int C, bool kBF16>
__device__ __forceinline__ void gather_channels(const ProjArgs& a, const SampleCtx& sc, int c0, float* x)
{
    constexpr int CPL = C / 4;
#pragma unroll
    for (int q = 0; q < CPL; ++q) x[q] = 0.f;
    if (!sc.valid || sc.gbase < 0) return;
    float v[8][CPL];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        int64_t off = sc.gbase + corner_cell(sc.tri, k, a.G) * C + c0;
        if constexpr (!kBF16) {
            const float* g = reinterpret_cast<const float*>(a.grid) + off;
            if constexpr (CPL == 4) {
                float4 f = *reinterpret_cast<const float4*>(g);
                v[k][0] = f.x;
                v[k][1] = f.y;
                v[k][2] = f.z;
                v[k][3] = f.w;
            } else {
                float2 f = *reinterpret_cast<const float2*>(g);
                v[k][0] = f.x;
                v[k][1] = f.y;
            }
        } else {
            const uint16_t* g = reinterpret_cast<const uint16_t*>(a.grid) + off;
            if constexpr (CPL == 4) {
                uint2 u = *reinterpret_cast<const uint2*>(g);
                v[k][0] = __uint_as_float(u.x << 16);
                v[k][1] = __uint_as_float(u.x & 0xffff0000u);
                v[k][2] = __uint_as_float(u.y << 16);
                v[k][3] = __uint_as_float(u.y & 0xffff0000u);
            } else {
                uint32_t u = *reinterpret_cast<const uint32_t*>(g);
                v[k][0] = __uint_as_float(u << 16);
                v[k][1] = __uint_as_float(u & 0xffff0000u);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc = fmaf(v[k][q], sc.tri.w[k], acc);
        x[q] = acc;
    }
}

// Tensorial (VM) features of an in-ROI sample, tensor_op "add"
// (geometry/geometry_tensor_qo.cpp:279-327): planes yz (u = ly, v = lz),
// xz (lx, lz), xy (lx, ly) and lines x (the ROI-normalised global x),
// y (ly), z (the raw global z); F = Pyz Lx + Pxz Ly + Pxy Lz with the
// products rounded and added in that order, each read accumulated
// acc += value * weight from 0 in corner order (as the 2-D / 1-D gathers).
// Channels c0 .. c0 + NCH - 1; P and L keep the reads for the adjoint.
struct VMSetup {
    Bil b[3];
    Lin l[3];
};

__device__ __forceinline__ void vm_setup(const ProjArgs& a, const SampleCtx& sc, VMSetup& st)
{
    bil_setup(sc.loc[1], sc.loc[2], a.vm.Gp, a.vm.Gp, st.b[0]);
    bil_setup(sc.loc[0], sc.loc[2], a.vm.Gp, a.vm.Gp, st.b[1]);
    bil_setup(sc.loc[0], sc.loc[1], a.vm.Gp, a.vm.Gp, st.b[2]);
    const float xr = (2.0f * (sc.gx - a.vm.rx0)) / (a.vm.rx1 - a.vm.rx0) - 1.0f;
    lin_setup(xr, a.vm.Gl, st.l[0]);
    lin_setup(sc.loc[1], a.vm.Gl, st.l[1]);
    lin_setup(sc.gz, a.vm.Gl, st.l[2]);
}

template <int C, int NCH>
__device__ __forceinline__ void load_channels(const float* p, float* v)
{
    if constexpr (NCH == 4) {
        const float4 f = *reinterpret_cast<const float4*>(p);
        v[0] = f.x;
        v[1] = f.y;
        v[2] = f.z;
        v[3] = f.w;
    } else {
        const float2 f = *reinterpret_cast<const float2*>(p);
        v[0] = f.x;
        v[1] = f.y;
    }
}

template <int C, int NCH>
__device__ __forceinline__ void vm_reads(const ProjArgs& a, int pslot, const VMSetup& st, int c0, float (*P)[NCH],
                                         float (*L)[NCH])
{
    const int Gp = a.vm.Gp;
    const float* pl = a.vm.planes + (int64_t)pslot * 3 * Gp * Gp * C + c0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const Bil& b = st.b[p];
        const float* pp = pl + (int64_t)p * Gp * Gp * C;
        const int off[4] = {(b.y0 * Gp + b.x0) * C, (b.y0 * Gp + b.x1) * C, (b.y1 * Gp + b.x0) * C,
                            (b.y1 * Gp + b.x1) * C};
        float acc[NCH];
#pragma unroll
        for (int q = 0; q < NCH; ++q) acc[q] = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float v[NCH];
            load_channels<C, NCH>(pp + off[k], v);
#pragma unroll
            for (int q = 0; q < NCH; ++q) acc[q] += v[q] * b.w[k];
        }
        const Lin& l = st.l[p];
        const float* ln = a.vm.lines + (int64_t)p * a.vm.Gl * C + c0;
        float v0[NCH], v1[NCH];
        load_channels<C, NCH>(ln + l.x0 * C, v0);
        load_channels<C, NCH>(ln + l.x1 * C, v1);
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            P[p][q] = acc[q];
            float la = 0.f;
            la += v0[q] * l.w0;
            la += v1[q] * l.w1;
            L[p][q] = la;
        }
    }
}

// vm_setup + vm_reads for plane p and line p alone (the same arithmetic):
// a sample's tensorial reads one pair at a time, in fewer registers.
template <int C, int NCH>
__device__ __forceinline__ void vm_pair_read(const ProjArgs& a, int pslot, const SampleCtx& sc, int p, int c0,
                                             float* P, float* L)
{
    const int Gp = a.vm.Gp;
    Bil b;
    Lin l;
    if (p == 0) {
        bil_setup(sc.loc[1], sc.loc[2], Gp, Gp, b);
        const float xr = (2.0f * (sc.gx - a.vm.rx0)) / (a.vm.rx1 - a.vm.rx0) - 1.0f;
        lin_setup(xr, a.vm.Gl, l);
    } else if (p == 1) {
        bil_setup(sc.loc[0], sc.loc[2], Gp, Gp, b);
        lin_setup(sc.loc[1], a.vm.Gl, l);
    } else {
        bil_setup(sc.loc[0], sc.loc[1], Gp, Gp, b);
        lin_setup(sc.gz, a.vm.Gl, l);
    }
    const float* pp = a.vm.planes + ((int64_t)pslot * 3 + p) * Gp * Gp * C + c0;
    const int off[4] = {(b.y0 * Gp + b.x0) * C, (b.y0 * Gp + b.x1) * C, (b.y1 * Gp + b.x0) * C,
                        (b.y1 * Gp + b.x1) * C};
    float acc[NCH];
#pragma unroll
    for (int q = 0; q < NCH; ++q) acc[q] = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float v[NCH];
        load_channels<C, NCH>(pp + off[k], v);
#pragma unroll
        for (int q = 0; q < NCH; ++q) acc[q] += v[q] * b.w[k];
    }
    const float* ln = a.vm.lines + (int64_t)p * a.vm.Gl * C + c0;
    float v0[NCH], v1[NCH];
    load_channels<C, NCH>(ln + l.x0 * C, v0);
    load_channels<C, NCH>(ln + l.x1 * C, v1);
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
        P[q] = acc[q];
        float la = 0.f;
        la += v0[q] * l.w0;
        la += v1[q] * l.w1;
        L[q] = la;
    }
}

template <int NCH>
__device__ __forceinline__ void vm_combine(const float (*P)[NCH], const float (*L)[NCH], float* x)
{
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
        float f = P[0][q] * L[0][q];
        f = f + P[1][q] * L[1][q];
        f = f + P[2][q] * L[2][q];
        x[q] = f;
    }
}

// Forward decoder for one tile.  x: this lane's CPL features of sample j.
// Returns y (pre-softplus) of sample j on every lane of column j; h1/h2 are
// the post-ReLU activations in result layout.
template <int C>
__device__ __forceinline__ float decoder_fwd(const float* s, const float* x, int g, int j, v4f* h1, v4f* h2)
{
    using L = MlpSmem<C>;
    constexpr int CPL = C / 4;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < CPL; ++kk) acc = mfma4(s[L::kW1 + (16 * rb + j) * L::W1S + CPL * g + kk], x[kk], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = acc[r] + s[L::kB1 + 16 * rb + 4 * g + r];
            h1[rb][r] = v > 0.f ? v : 0.f;
        }
    }
#pragma unroll
    for (int rb2 = 0; rb2 < 4; ++rb2) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const float* wrow = s + L::kW2 + (16 * rb2 + j) * kW2S + 16 * rb + 4 * g;
            float4 wv = *reinterpret_cast<const float4*>(wrow);
            acc = mfma4(wv.x, h1[rb][0], acc);
            acc = mfma4(wv.y, h1[rb][1], acc);
            acc = mfma4(wv.z, h1[rb][2], acc);
            acc = mfma4(wv.w, h1[rb][3], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = acc[r] + s[L::kB2 + 16 * rb2 + 4 * g + r];
            h2[rb2][r] = v > 0.f ? v : 0.f;
        }
    }
    float part = 0.f;
#pragma unroll
    for (int rb2 = 0; rb2 < 4; ++rb2)
#pragma unroll
        for (int r = 0; r < 4; ++r) part += s[L::kW3 + 16 * rb2 + 4 * g + r] * h2[rb2][r];
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    return part + s[L::kB3];
}

// ---------------------------------------------------------------------------
// BF16 decoder.  Weights are stored in LDS as bf16, pre-permuted so each
// lane's MFMA operand is one contiguous 8- or 16-byte read.  For a K = 64
// contraction (hidden -> hidden) K-step s and lane group g use the rows
//   {32 s + 4 g + r} U {32 s + 16 + 4 g + r},  r = 0..3,
// i.e. the result-layout registers of row blocks 2s and 2s+1, so activations
// chain between layers without moving between lanes.
// ---------------------------------------------------------------------------
constexpr int kBW = 72;  // padded row stride (bf16 elements) of the K = 64 tables
constexpr int kTB = 24;  // padded row stride (bf16 elements) of the transpose buffers

template <int C>
struct BSmem {
    // bf16 element offsets
    static constexpr int kW1b = 0;                 // [64][16]  W1 rows, lane-group order
    static constexpr int kW2b = kW1b + kH * 16;    // [64][72]  W2 rows, permuted K
    static constexpr int kW2Tb = kW2b + kH * kBW;  // [64][72]  W2^T rows, permuted K
    static constexpr int kW1Tb = kW2Tb + kH * kBW; // [16][72]  W1^T rows (channel), permuted K
    static constexpr int kBf16Total = kW1Tb + 16 * kBW;
    // f32 element offsets (after the bf16 region)
    static constexpr int kB1 = 0, kB2 = kH, kW3 = 2 * kH, kB3 = 3 * kH, kF32Total = 3 * kH + 4;
};

__device__ __forceinline__ int permuted_k(int p)  // p in [0, 64): (s, g, q) -> hidden index
{
    int s = p >> 5, g = (p >> 3) & 3, q = p & 7;
    return 32 * s + 16 * (q >> 2) + 4 * g + (q & 3);
}

template <int C>
__device__ __forceinline__ void load_mlp_bf16_to_lds(const float* __restrict__ mlp, __bf16* sb, float* sf)
{
    using L = BSmem<C>;
    const float* W1 = mlp;
    const float* b1 = W1 + kH * C;
    const float* W2 = b1 + kH;
    const float* b2 = W2 + kH * kH;
    const float* W3 = b2 + kH;
    const float* b3 = W3 + kH;
    for (int i = threadIdx.x; i < kH * 16; i += blockDim.x) {
        int row = i >> 4, p = i & 15, g = p >> 2, q = p & 3;
        int c = (C == 16) ? 4 * g + q : (q < 2 ? 2 * g + q : -1);
        sb[L::kW1b + i] = (__bf16)(c >= 0 ? W1[row * C + c] : 0.f);
    }
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
        int row = i >> 6, p = i & 63, k = permuted_k(p);
        sb[L::kW2b + row * kBW + p] = (__bf16)W2[row * kH + k];
        sb[L::kW2Tb + row * kBW + p] = (__bf16)W2[k * kH + row];
    }
    for (int i = threadIdx.x; i < 16 * kH; i += blockDim.x) {
        int c = i >> 6, p = i & 63, k = permuted_k(p);
        sb[L::kW1Tb + c * kBW + p] = (__bf16)(c < C ? W1[k * C + c] : 0.f);
    }
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        sf[L::kB1 + i] = b1[i];
        sf[L::kB2 + i] = b2[i];
        sf[L::kW3 + i] = W3[i];
    }
    if (threadIdx.x == 0) sf[L::kB3] = b3[0];
}

template <int C>
__device__ __forceinline__ float decoder_fwd_bf16(const __bf16* sb, const float* sf, const float* x, int g, int j,
                                                  v4f* h1, v4f* h2)
{
    using L = BSmem<C>;
    const s4 xb = (C == 16) ? pack_bf4(x[0], x[1], x[2], x[3]) : pack_bf4(x[0], x[1], 0.f, 0.f);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_bf16_k16(*reinterpret_cast<const s4*>(sb + L::kW1b + (16 * rb + j) * 16 + 4 * g), xb, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = acc[r] + sf[L::kB1 + 16 * rb + 4 * g + r];
            h1[rb][r] = v > 0.f ? v : 0.f;
        }
    }
#pragma unroll
    for (int rb2 = 0; rb2 < 4; ++rb2) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            bf8 av = *reinterpret_cast<const bf8*>(sb + L::kW2b + (16 * rb2 + j) * kBW + 32 * st + 8 * g);
            acc = mfma_bf16_k32(av, pack_bf8(h1[2 * st], h1[2 * st + 1]), acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = acc[r] + sf[L::kB2 + 16 * rb2 + 4 * g + r];
            h2[rb2][r] = v > 0.f ? v : 0.f;
        }
    }
    float part = 0.f;
#pragma unroll
    for (int rb2 = 0; rb2 < 4; ++rb2)
#pragma unroll
        for (int r = 0; r < 4; ++r) part += sf[L::kW3 + 16 * rb2 + 4 * g + r] * h2[rb2][r];
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    return part + sf[L::kB3];
}

template <int C, bool kDecBF16>
struct DecoderSmem {
    static constexpr int kF32 = kDecBF16 ? BSmem<C>::kF32Total : MlpSmem<C>::kTotal;
    static constexpr int kBF16 = kDecBF16 ? BSmem<C>::kBf16Total : 8;
};

// Bin of the mass histogram: b = (int)((t - t0) / (t1 - t0) * 32), clamped.
__device__ __forceinline__ int sample_bin(float t, float t0, float t1)
{
    const float span = t1 - t0;
    int b = 0;
    if (span > 0.f) {
        const float rel = (t - t0) / span;
        b = (int)(rel * 32.0f);
    }
    return b < 0 ? 0 : (b > 31 ? 31 : b);
}

// Four waves per SIMD: the weight reads below go through an opaque copy of
// the LDS pointers, so the compiler reloads them per tile instead of keeping
// every weight in registers (about 95 VGPRs instead of 180-250).
template <int C, bool kBF16, bool kDecBF16, bool kBins, bool kVM = false>
__global__ __launch_bounds__(64 * kWaves, 4) void project_fwd_kernel(ProjArgs a, float* __restrict__ P)
{
    using SM = DecoderSmem<C, kDecBF16>;
    constexpr int CPL = C / 4;
    __shared__ __attribute__((aligned(16))) float s[SM::kF32];
    __shared__ __attribute__((aligned(16))) __bf16 sb[SM::kBF16];
    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds<C>(a.mlp, s);
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int64_t wave = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
    const int64_t r0 = wave * a.rpw;
    const int64_t r1 = (r0 + a.rpw < a.R) ? r0 + a.rpw : a.R;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t h0 = a.iv_off[r], h1e = a.iv_off[r + 1];
        if (h0 == h1e) {
            if (lane == 0) P[r] = 0.f;
            if constexpr (kBins) {
                if (j == 0) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) a.bins[r * 32 + 8 * g + q] = 0.f;
                }
            }
            continue;
        }
        const int ns = a.iv_first[h1e - 1] + a.iv_count[h1e - 1];
        float o[3], d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            o[k] = a.O[r * 3 + k];
            d[k] = a.D[r * 3 + k];
        }
        int64_t cur = h0;
        float acc = 0.f;
        // binned forward: lane group g owns bins 8g .. 8g+7
        float bacc[8];
        float bt0 = 0.f, bt1 = 0.f;
        if constexpr (kBins) {
#pragma unroll
            for (int q = 0; q < 8; ++q) bacc[q] = 0.f;
            const PNode root = a.nodes[0];
            intersect_box_ray(root.mn, root.mx, o, d, bt0, bt1);
        }
        for (int base = 0; base < ns; base += kTile) {
            // the tile's 16 samples lie in the 16 intervals from `cur` on
            // (every interval holds at least one sample): load them once,
            // then each lane finds its own by a cross-lane binary search
            SampleCtx sc;
            const int64_t cand = cur + j;
            int c_end = INT_MAX, c_cnt = 0, c_node = 0;
            float c_lo = 0.f, c_hi = 0.f;
            if (cand < h1e) {
                c_cnt = a.iv_count[cand];
                c_end = a.iv_first[cand] + c_cnt;
                c_node = a.iv_node[cand];
                c_lo = a.iv_tmin[cand];
                c_hi = a.iv_tmax[cand];
            }
            const int t = base + j;
            const int grp = lane & 48;
            // a window of one-sample intervals maps sample j to candidate j
            const bool ones = __all(cand >= h1e || c_cnt == 1);
            int pos, posn, s_cnt, s_end, s_node;
            float s_lo, s_hi;
            if (ones) {
                pos = j;
                posn = (int)min((int64_t)kTile, h1e - cur);
                s_cnt = c_cnt;
                s_end = c_end;
                s_node = c_node;
                s_lo = c_lo;
                s_hi = c_hi;
            } else {
                pos = group_search(c_end, t, grp);
                posn = group_search(c_end, base + kTile, grp);
                const int src = grp | (pos < kTile ? pos : kTile - 1);
                s_cnt = __shfl(c_cnt, src, 64);
                s_end = __shfl(c_end, src, 64);
                s_node = __shfl(c_node, src, 64);
                s_lo = __shfl(c_lo, src, 64);
                s_hi = __shfl(c_hi, src, 64);
            }
            sc.valid = t < ns;
            sc.w = 0.f;
            sc.gbase = -1;
            if (sc.valid) {
                if (pos < kTile) {
                    make_sample_iv<kVM>(a, r, o, d, s_node, s_lo, s_hi, s_cnt, t - (s_end - s_cnt), t, C, sc);
                } else {
                    // only with zero-sample intervals in the list: walk on
                    int64_t q = cur + kTile - 1;
                    while (q + 1 < h1e && a.iv_first[q] + a.iv_count[q] <= t) ++q;
                    make_sample_iv<kVM>(a, r, o, d, a.iv_node[q], a.iv_tmin[q], a.iv_tmax[q], a.iv_count[q],
                                        t - a.iv_first[q], t, C, sc);
                }
            }
            cur += posn;
            float x[CPL];
            gather_channels<C, kBF16>(a, sc, CPL * g, x);
            if constexpr (kVM) {
                if (sc.valid && sc.pslot >= 0) {
                    VMSetup st;
                    vm_setup(a, sc, st);
                    float Pv[3][CPL], Lv[3][CPL];
                    vm_reads<C, CPL>(a, sc.pslot, st, CPL * g, Pv, Lv);
                    vm_combine<CPL>(Pv, Lv, x);
                }
            }
            v4f h1[4], h2[4];
            float y;
            int opq = 0;
            asm volatile("" : "+v"(opq));  // no hoisting of the weight reads
            const float* so = s + opq;
            const __bf16* sbo = sb + opq;
            if constexpr (kDecBF16)
                y = decoder_fwd_bf16<C>(sbo, so, x, g, j, h1, h2);
            else
                y = decoder_fwd<C>(so, x, g, j, h1, h2);
            if (sc.valid) {
                const float v = softplus_f(y, a.beta) * sc.w;
                acc += v;
                if constexpr (kBins) {
                    const int b = sample_bin(sc.t, bt0, bt1);
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (b == 8 * g + q) bacc[q] += v;
                }
            }
        }
        acc += __shfl_xor(acc, 1, 64);
        acc += __shfl_xor(acc, 2, 64);
        acc += __shfl_xor(acc, 4, 64);
        acc += __shfl_xor(acc, 8, 64);
        if (lane == 0) P[r] = acc;
        if constexpr (kBins) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                float v = bacc[q];
                v += __shfl_xor(v, 1, 64);
                v += __shfl_xor(v, 2, 64);
                v += __shfl_xor(v, 4, 64);
                v += __shfl_xor(v, 8, 64);
                bacc[q] = v;
            }
            if (j == 0) {
#pragma unroll
                for (int q = 0; q < 8; ++q) a.bins[r * 32 + 8 * g + q] = bacc[q];
            }
        }
    }
}

template <int C>
constexpr int mlp_numel()
{
    return kH * C + kH + kH * kH + kH + kH + 1;
}

// Decoder-parameter gradient accumulators of one wave.
struct DecGrads {
    static constexpr bool kRS = false;
    v4f W2[4][4];
    v4f W1[4];
    float W3[4][4], B2[4][4], B1[4][4];  // per-lane sums over its samples
    float B3;
};

// The same with the 64-entry gradients of b1, b2 and W3 summed over the
// tile's 16 sample lanes at the end of every tile (row_reduce_scatter16), so
// that each lane keeps one entry of each (row 16 (j / 4) + 4 g + j % 4):
// 45 fewer registers for the price of about 135 lane-permute VALU
// operations per tile.
struct DecGradsRS {
    static constexpr bool kRS = true;
    v4f W2[4][4];
    v4f W1[4];
    float W3, B2, B1;
    float B3;
};

__device__ __forceinline__ void zero_grads(DecGrads& q)
{
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        q.W1[p] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            q.W2[p][r] = (v4f){0.f, 0.f, 0.f, 0.f};
            q.W3[p][r] = q.B2[p][r] = q.B1[p][r] = 0.f;
        }
    }
    q.B3 = 0.f;
}

__device__ __forceinline__ void zero_grads(DecGradsRS& q)
{
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        q.W1[p] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < 4; ++r) q.W2[p][r] = (v4f){0.f, 0.f, 0.f, 0.f};
    }
    q.W3 = q.B2 = q.B1 = q.B3 = 0.f;
}

template <int kCtrl>
__device__ __forceinline__ float dpp_f(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), kCtrl, 0xF, 0xF,
                                                                 false));
}

// Sum of v[i] over the 16 lanes of each 16-lane row; lane j of the row
// returns the sum for i = j.  A butterfly in which each step halves the
// entries a lane keeps: the partner differs in bit 3 (row_ror:8), bit 2
// (row_half_mirror: both lanes still hold the same entries), bit 1 and bit 0
// (quad permutations); the lane whose bit is set keeps the upper half.
__device__ __forceinline__ float row_reduce_scatter16(const float* v, int j)
{
    float u[8], w[4], x[2];
    const bool b3 = j & 8, b2 = j & 4, b1 = j & 2, b0 = j & 1;
#pragma unroll
    for (int k = 0; k < 8; ++k) u[k] = (b3 ? v[k + 8] : v[k]) + dpp_f<0x128>(b3 ? v[k] : v[k + 8]);
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = (b2 ? u[k + 4] : u[k]) + dpp_f<0x141>(b2 ? u[k] : u[k + 4]);
#pragma unroll
    for (int k = 0; k < 2; ++k) x[k] = (b1 ? w[k + 2] : w[k]) + dpp_f<0x4E>(b1 ? w[k] : w[k + 2]);
    return (b0 ? x[1] : x[0]) + dpp_f<0xB1>(b0 ? x[0] : x[1]);
}

// Backward of the decoder for one tile given dL/dy (dy) of this lane's
// sample; accumulates the parameter gradients and returns dX (channels
// 4g .. 4g+3 of sample j, result layout).  bufA/bufB (f32) or bufAb/bufBb
// (bf16) are this wave's transpose buffers.
// kT (FP32 decoder): the weights carry the MlpSmemT transposed copies.
template <int C, bool kDecBF16, class Q, bool kT = false>
__device__ __forceinline__ v4f tile_backward(const float* s, const __bf16* sb, float* bufA, float* bufB, __bf16* bufAb,
                                             __bf16* bufBb, const float* x, const v4f* h1, const v4f* h2, float dy,
                                             int g, int j, Q& q)
{
    using LT = MlpSmemT<C>;
    using L = MlpSmem<C>;
    using LB = BSmem<C>;
    constexpr int CPL = C / 4;
    const float* sW3 = kDecBF16 ? s + LB::kW3 : s + L::kW3;
    if (g == 0) q.B3 += dy;
    v4f dh2[4];
    float tw3[16];
#pragma unroll
    for (int rb2 = 0; rb2 < 4; ++rb2)
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
            const int row = 16 * rb2 + 4 * g + r4;
            float v = (h2[rb2][r4] > 0.f) ? dy * sW3[row] : 0.f;
            dh2[rb2][r4] = v;
            if constexpr (Q::kRS) {
                tw3[4 * rb2 + r4] = dy * h2[rb2][r4];
            } else {
                q.W3[rb2][r4] += dy * h2[rb2][r4];
                q.B2[rb2][r4] += v;
            }
        }
    if constexpr (Q::kRS) {
        q.W3 += row_reduce_scatter16(tw3, j);
        float tb2[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) tb2[i] = dh2[i >> 2][i & 3];
        q.B2 += row_reduce_scatter16(tb2, j);
    }
    v4f dh1[4];
    v4f dx = {0.f, 0.f, 0.f, 0.f};
    if constexpr (!kDecBF16) {
        // dW2 += dH2 . H1^T (K = the tile's 16 samples)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                bufA[(16 * rb + 4 * g + r4) * kTS + j] = h1[rb][r4];
                bufB[(16 * rb + 4 * g + r4) * kTS + j] = dh2[rb][r4];
            }
        wave_lds_sync();
#pragma unroll
        for (int rb2 = 0; rb2 < 4; ++rb2) {
            float4 av = *reinterpret_cast<const float4*>(bufB + (16 * rb2 + j) * kTS + 4 * g);
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                float4 bv = *reinterpret_cast<const float4*>(bufA + (16 * rb + j) * kTS + 4 * g);
                v4f acc = q.W2[rb2][rb];
                acc = mfma4(av.x, bv.x, acc);
                acc = mfma4(av.y, bv.y, acc);
                acc = mfma4(av.z, bv.z, acc);
                acc = mfma4(av.w, bv.w, acc);
                q.W2[rb2][rb] = acc;
            }
        }
        wave_lds_sync();
        // dH1 = W2^T . dH2, masked by ReLU
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int rb2 = 0; rb2 < 4; ++rb2) {
                if constexpr (kT) {
                    const float4 w = *reinterpret_cast<const float4*>(s + LT::kW2T + (16 * rb + j) * kW2S + 16 * rb2 +
                                                                      4 * g);
                    acc = mfma4(w.x, dh2[rb2][0], acc);
                    acc = mfma4(w.y, dh2[rb2][1], acc);
                    acc = mfma4(w.z, dh2[rb2][2], acc);
                    acc = mfma4(w.w, dh2[rb2][3], acc);
                } else {
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk)
                        acc = mfma4(s[L::kW2 + (16 * rb2 + 4 * g + kk) * kW2S + 16 * rb + j], dh2[rb2][kk], acc);
                }
            }
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                float v = (h1[rb][r4] > 0.f) ? acc[r4] : 0.f;
                dh1[rb][r4] = v;
                if constexpr (!Q::kRS) q.B1[rb][r4] += v;
            }
        }
        // dX = W1^T . dH1 (rows = channels, padded to 16)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            if constexpr (kT) {
                const float4 w = *reinterpret_cast<const float4*>(s + LT::kW1T + j * kW2S + 16 * rb + 4 * g);
                dx = mfma4(w.x, dh1[rb][0], dx);
                dx = mfma4(w.y, dh1[rb][1], dx);
                dx = mfma4(w.z, dh1[rb][2], dx);
                dx = mfma4(w.w, dh1[rb][3], dx);
            } else {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    float wv = (j < C) ? s[L::kW1 + (16 * rb + 4 * g + kk) * L::W1S + j] : 0.f;
                    dx = mfma4(wv, dh1[rb][kk], dx);
                }
            }
        }
        // dW1 += dH1 . X^T
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) bufA[(16 * rb + 4 * g + r4) * kTS + j] = dh1[rb][r4];
#pragma unroll
        for (int qq = 0; qq < CPL; ++qq) bufB[(CPL * g + qq) * kTS + j] = x[qq];
        wave_lds_sync();
        float4 bv = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < C) bv = *reinterpret_cast<const float4*>(bufB + j * kTS + 4 * g);
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            float4 av = *reinterpret_cast<const float4*>(bufA + (16 * rb + j) * kTS + 4 * g);
            v4f acc = q.W1[rb];
            acc = mfma4(av.x, bv.x, acc);
            acc = mfma4(av.y, bv.y, acc);
            acc = mfma4(av.z, bv.z, acc);
            acc = mfma4(av.w, bv.w, acc);
            q.W1[rb] = acc;
        }
        wave_lds_sync();
    } else {
        // dW2 += dH2 . H1^T on bf16 MFMA (K = the tile's 16 samples)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                bufAb[(16 * rb + 4 * g + r4) * kTB + j] = (__bf16)h1[rb][r4];
                bufBb[(16 * rb + 4 * g + r4) * kTB + j] = (__bf16)dh2[rb][r4];
            }
        wave_lds_sync();
#pragma unroll
        for (int rb2 = 0; rb2 < 4; ++rb2) {
            s4 av = *reinterpret_cast<const s4*>(bufBb + (16 * rb2 + j) * kTB + 4 * g);
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                s4 bv = *reinterpret_cast<const s4*>(bufAb + (16 * rb + j) * kTB + 4 * g);
                q.W2[rb2][rb] = mfma_bf16_k16(av, bv, q.W2[rb2][rb]);
            }
        }
        wave_lds_sync();
        // dH1 = W2^T . dH2 (K = 64 in two bf16 steps), masked by ReLU
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int st = 0; st < 2; ++st) {
                bf8 av = *reinterpret_cast<const bf8*>(sb + LB::kW2Tb + (16 * rb + j) * kBW + 32 * st + 8 * g);
                acc = mfma_bf16_k32(av, pack_bf8(dh2[2 * st], dh2[2 * st + 1]), acc);
            }
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                float v = (h1[rb][r4] > 0.f) ? acc[r4] : 0.f;
                dh1[rb][r4] = v;
                if constexpr (!Q::kRS) q.B1[rb][r4] += v;
            }
        }
        // dX = W1^T . dH1 (rows = channels, padded to 16)
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            bf8 av = *reinterpret_cast<const bf8*>(sb + LB::kW1Tb + j * kBW + 32 * st + 8 * g);
            dx = mfma_bf16_k32(av, pack_bf8(dh1[2 * st], dh1[2 * st + 1]), dx);
        }
        // dW1 += dH1 . X^T
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) bufAb[(16 * rb + 4 * g + r4) * kTB + j] = (__bf16)dh1[rb][r4];
#pragma unroll
        for (int qq = 0; qq < CPL; ++qq) bufBb[(CPL * g + qq) * kTB + j] = (__bf16)x[qq];
        wave_lds_sync();
        s4 bv = {0, 0, 0, 0};
        if (j < C) bv = *reinterpret_cast<const s4*>(bufBb + j * kTB + 4 * g);
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            s4 av = *reinterpret_cast<const s4*>(bufAb + (16 * rb + j) * kTB + 4 * g);
            q.W1[rb] = mfma_bf16_k16(av, bv, q.W1[rb]);
        }
        wave_lds_sync();
    }
    if constexpr (Q::kRS) {
        float tb1[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) tb1[i] = dh1[i >> 2][i & 3];
        q.B1 += row_reduce_scatter16(tb1, j);
    }
    return dx;
}

// Writes this wave's decoder-gradient partials to row `out` (every entry).
template <int C>
__device__ __forceinline__ void store_partials(DecGrads& q, float* out, int g, int j, int lane)
{
    constexpr int oW1 = 0, oB1 = kH * C, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4)
            if (j < C) out[oW1 + (16 * rb + 4 * g + r4) * C + j] = q.W1[rb][r4];
#pragma unroll
    for (int rb2 = 0; rb2 < 4; ++rb2)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) out[oW2 + (16 * rb2 + 4 * g + r4) * kH + 16 * rb + j] = q.W2[rb2][rb][r4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v1 = q.B1[p][r], v2 = q.B2[p][r], v3 = q.W3[p][r];
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                v1 += __shfl_xor(v1, off, 64);
                v2 += __shfl_xor(v2, off, 64);
                v3 += __shfl_xor(v3, off, 64);
            }
            if (j == 0) {
                const int row = 16 * p + 4 * g + r;
                out[oB1 + row] = v1;
                out[oB2 + row] = v2;
                out[oW3 + row] = v3;
            }
        }
    float b3 = q.B3;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) b3 += __shfl_xor(b3, off, 64);
    if (lane == 0) out[oB3] = b3;
}

// kAdd: add to the row (written by an earlier kernel on the same stream)
template <int C, bool kAdd = false>
__device__ __forceinline__ void store_partials(DecGradsRS& q, float* out, int g, int j, int lane)
{
    constexpr int oW1 = 0, oB1 = kH * C, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
    auto put = [&](int i, float v) { out[i] = kAdd ? out[i] + v : v; };
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4)
            if (j < C) put(oW1 + (16 * rb + 4 * g + r4) * C + j, q.W1[rb][r4]);
#pragma unroll
    for (int rb2 = 0; rb2 < 4; ++rb2)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) put(oW2 + (16 * rb2 + 4 * g + r4) * kH + 16 * rb + j, q.W2[rb2][rb][r4]);
    const int row = 16 * (j >> 2) + 4 * g + (j & 3);
    put(oB1 + row, q.B1);
    put(oB2 + row, q.B2);
    put(oW3 + row, q.W3);
    float b3 = q.B3;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) b3 += __shfl_xor(b3, off, 64);
    if (lane == 0) put(oB3, b3);
}

// ---------------------------------------------------------------------------
// Node-major adjoint.  Intervals are counting-sorted by grid slot and cut
// into chunks of one node.  A persistent workgroup (one per CU: the node's
// gradient lives in LDS) takes chunks from a global counter, splits the
// chunk's intervals over its 4 waves, recomputes every sample (interval ->
// t, weight, position, gather, decoder), runs the decoder backward on MFMA,
// and scatters w_corner * dX into the LDS copy of the node's grid gradient
// under a workgroup lock (see the scatter below).  At the end of the chunk
// the LDS gradient is added to the global gradient once: a plain
// read-add-write when the chunk covers the whole node, a global atomic add
// when the node was split over several chunks.
//
// The counting sort writes every interval's fields, its ray's origin /
// direction and dL/dP into one 48-byte record at its sorted position, so the
// node kernel reads its candidate intervals as contiguous records instead of
// through three levels of indirection (sorted id -> interval -> ray).
//
// A chunk holds about kChunkSamples samples: its interval count is
// kChunkSamples / (mean samples per interval of the list), so stratified
// lists (~6 samples per interval) and learned one-sample lists both give
// chunks of similar work, and a node is split only when it carries more
// samples than that.
// ---------------------------------------------------------------------------
#ifndef ATDF_NODE_CHUNK_SAMPLES
#define ATDF_NODE_CHUNK_SAMPLES 8192
#endif
constexpr int64_t kChunkSamples = ATDF_NODE_CHUNK_SAMPLES;
constexpr int kMinChunk = 64;     // intervals
constexpr int kMaxChunk = 16384;  // intervals

__device__ __forceinline__ int64_t chunk_intervals(int64_t H, int64_t S)
{
    int64_t c = S > 0 ? (kChunkSamples * H + S - 1) / S : kMaxChunk;
    return c < kMinChunk ? kMinChunk : (c > kMaxChunk ? kMaxChunk : c);
}

// Adam on the node grids, applied by the adjoint itself (atdf_project_bwd_adam).
struct GridAdam {
    float* p;      // f32 grids (BF16 field: the master copy)
    uint16_t* pb;  // BF16 grid copy or null
    float* m;
    float* v;
    float lr, b1, b2, eps, bc1, bc2s;
    int enabled;
};

// Waves per workgroup of the adjoint (they share one LDS node gradient).
#ifndef ATDF_BWD_WAVES_BF16
#define ATDF_BWD_WAVES_BF16 4
#endif
template <bool kDecBF16>
__host__ __device__ constexpr int bwd_waves()
{
    return kDecBF16 ? ATDF_BWD_WAVES_BF16 : kWaves;
}
constexpr int kMaxBwdWaves = ATDF_BWD_WAVES_BF16 > kWaves ? ATDF_BWD_WAVES_BF16 : kWaves;

struct NodeBwdArgs {
    ProjArgs a;
    GridAdam ad;
    const float* dLdP;
    const int32_t* sorted;    // interval ids, grouped by grid slot
    const int4* chunks;       // (slot, begin, end, split)
    const int64_t* chunk_off; // [num_slots + 1]; chunk_off[num_slots] = chunk count
    int64_t num_slots;
    int* counter;             // zeroed before the launch
    const int32_t* iv_ray;
    const float4* rec;        // [H][3] sorted interval records (write_record)
    float* grad_grid;
    float* ws;                // decoder-gradient partial rows
    int lds_f32_off;          // float offset of the weight region in dynamic LDS
    int lds_bf16_off;         // byte offset of the bf16 weight region
    int lds_tb_off;           // byte offset of the transpose buffers
    // tensorial field: plane gradients, per-workgroup line partial rows
    // ([blocks][3 Gl C]) and the float offset of the line partials in LDS
    float* grad_planes;
    float* line_part;
    int lds_line_off;
    int lds_grad_floats;      // LDS node-gradient region, zeroed at start
    // two-pass adjoint (the staging region of the workspace)
    const int* mode;          // 1: the two-pass kernels run, 0: the fused kernel; null: fused
    const int64_t* soff;      // [H + 1] first staged sample of each sorted interval
    const int4* parts;        // decoder-pass work items (slot, begin, end, chunk)
    const int64_t* part_off;  // [max_chunks + 1]; part_off[max_chunks] = item count
    int64_t max_chunks;
    float4* stage;            // per sample: lattice coordinates, then dX (C/4 float4)
    float* ws2;               // decoder-pass gradient partial rows
    int* counter2;            // decoder-pass item counter
    int* counter3;            // scatter-pass chunk counter
    int lds_pl_off;           // byte offset of the decoder pass's per-wave sample tables
    // tensorial field in the two-pass form: slots [0, exp_slots) are explicit
    // grids; the in-ROI samples (sorted after every explicit one, from
    // soff[slot_off[exp_slots]]) are staged in records of their own
    // (vm_record_f4) after the explicit region and its padding
    const int64_t* slot_off;  // [num_slots + 1] first sorted interval of each slot
    int64_t exp_slots;
    int* counter4;            // tensorial plane-pass chunk counter
    int* counter5;            // tensorial decoder-pass item counter
    float* line_rows;         // line partial rows of the tensorial line pass
};

template <int C, bool kDecBF16>
__host__ __device__ constexpr int weights_f32_floats()
{
    return kDecBF16 ? BSmem<C>::kF32Total : MlpSmem<C>::kTotal;
}

// Sorted interval record k (three float4): {tmin, tmax, count, first sample
// in the ray}, {origin, direction.x}, {direction.y, direction.z, interval id,
// ray}.
// Written by the counting sort's scatter pass, which reads the intervals in
// list order (coalesced) and their rays (mostly the same few lines).
struct RecordSrc {
    const float* tmin;
    const float* tmax;
    const int32_t* count;
    const int32_t* first;
    const int32_t* ray;
    const float* O;
    const float* D;
    float4* rec;
    int32_t* scnt;  // sample count per sorted position (two-pass adjoint), or null
};

__device__ __forceinline__ void write_record(const RecordSrc& rs, int64_t i, int64_t k)
{
    const int64_t r = rs.ray[i];
    if (rs.scnt) rs.scnt[k] = rs.count[i];
    rs.rec[k * 3] = make_float4(rs.tmin[i], rs.tmax[i], __int_as_float(rs.count[i]), __int_as_float(rs.first[i]));
    rs.rec[k * 3 + 1] = make_float4(rs.O[r * 3], rs.O[r * 3 + 1], rs.O[r * 3 + 2], rs.D[r * 3]);
    rs.rec[k * 3 + 2] = make_float4(rs.D[r * 3 + 1], rs.D[r * 3 + 2], __int_as_float((int)i), __int_as_float((int)r));
}

// make_sample_iv for a node whose box and grid offset are already known
// (with_tri false: an in-ROI sample of a tensorial field, no grid corners).
template <bool kVM = false>
__device__ __forceinline__ void make_sample_node(const ProjArgs& a, const PNode& nd, int64_t gbase, int64_t r,
                                                 const float* o, const float* d, float tmin, float tmax, int n,
                                                 int jj, int s_local, SampleCtx& sc, bool with_tri = true)
{
    const float step = (tmax - tmin) / (float)n;
    float aj = 0.5f, ap = 0.5f, an = 0.5f;
    if (a.jitter_mode == 2) {
        aj = hash_jitter(a.seed, r, s_local);
        if (jj > 0) ap = hash_jitter(a.seed, r, s_local - 1);
        if (jj < n - 1) an = hash_jitter(a.seed, r, s_local + 1);
    }
    const float t = sample_t(tmin, step, jj, aj);
    const float tp = (jj > 0) ? sample_t(tmin, step, jj - 1, ap) : 0.f;
    const float tn = (jj < n - 1) ? sample_t(tmin, step, jj + 1, an) : 0.f;
    sc.w = sample_weight(tmin, tmax, n, jj, tp, t, tn);
    sc.t = t;
    float loc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float pos = o[k] + d[k] * t;
        float size = nd.mx[k] - nd.mn[k];
        loc[k] = ((pos - nd.mn[k]) / size) * 2 - 1;
        if constexpr (kVM) {
            sc.loc[k] = loc[k];
            if (k == 0) sc.gx = pos;
            if (k == 2) sc.gz = pos;
        }
    }
    sc.gbase = gbase;
    if (!with_tri) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) sc.li[k] = tri_index(loc[k], a.G);
    tri_from_index(sc.li[0], sc.li[1], sc.li[2], a.G, a.G, a.G, sc.tri);
}

// Software pipeline of the adjoint's tile loop.  The node gradient fills
// most of the LDS, so a CU runs one workgroup (one wave per SIMD) and the
// loop hides its own memory latency: the stages below split a tile into
// (1) mapping its samples to intervals and loading their records,
// (2) placing the samples and issuing the corner gathers, (3) the decoder,
// its adjoint and the scatter, and the loop runs them on three consecutive
// tiles.  Each tile's arithmetic is the same as in an unpipelined loop and
// comes in the same order.

// Wave-uniform interval cursor over the wave's sorted intervals [.., ie):
// `cur` is the interval holding the next tile's first sample, `cum` the
// number of the wave's samples before it, and `cnt` the sample count of
// candidate interval cur + j (0 past ie), loaded one tile ahead.
struct TileCursor {
    int cur, cum, cnt;
};

// Stage 1 result for this lane's sample: its interval's records.
struct TileMap {
    bool valid;
    int jj, n_s;
    float4 A, B, D;
};

// Stage 1 of tile ti: maps the tile's 16 samples onto the intervals from
// the cursor on (a window of one-sample intervals, as learned lists have,
// maps sample j to candidate j directly; otherwise the counts are
// prefix-summed across the lane group and each lane binary-searches),
// loads the records of this lane's interval, advances the cursor and loads
// the next window's counts.  Every lane of the wave runs it.
__device__ __forceinline__ void map_tile(const float4* __restrict__ rec, int ie, int ti, int tot, int j, int grp,
                                         TileCursor& tc, TileMap& m)
{
    const int base = ti * kTile;
    const int t = base + j;
    const int cand = tc.cur + j;
    const int c_cnt = tc.cnt;
    const bool ones = __all(cand >= ie || c_cnt == 1);
    int c_end = c_cnt;
    int pos, posn, s_cnt, s_end, last_end = 0;
    if (ones) {
        pos = j;
        posn = min(kTile, ie - tc.cur);
        s_cnt = 1;
        s_end = tc.cum + j + 1;
    } else {
        c_end = scan16_inclusive(c_end);
        c_end = cand < ie ? c_end + tc.cum : INT_MAX;
        pos = group_search(c_end, t, grp);
        posn = group_search(c_end, base + kTile, grp);
        const int src = grp | (pos < kTile ? pos : kTile - 1);
        s_cnt = __shfl(c_cnt, src, 64);
        s_end = __shfl(c_end, src, 64);
        last_end = __shfl(c_end, grp | (kTile - 1), 64);
    }
    m.valid = t < tot;
    if (m.valid) {
        int k = tc.cur + pos, n_s = s_cnt, start = s_end - s_cnt;
        if (pos >= kTile) {
            // only with zero-sample intervals in the list: walk on
            k = tc.cur + kTile;
            start = last_end;
            n_s = __float_as_int(rec[3 * (int64_t)k].z);
            while (start + n_s <= t) {
                start += n_s;
                ++k;
                n_s = __float_as_int(rec[3 * (int64_t)k].z);
            }
        }
        m.n_s = n_s;
        m.jj = t - start;
        m.A = rec[3 * (int64_t)k];
        m.B = rec[3 * (int64_t)k + 1];
        m.D = rec[3 * (int64_t)k + 2];
    }
    if (ones)
        tc.cum += posn;
    else if (posn > 0)
        tc.cum = __shfl(c_end, grp | (posn - 1), 64);
    tc.cur += posn;
    const int nc = tc.cur + j;
    tc.cnt = nc < ie ? __float_as_int(rec[3 * (int64_t)nc].z) : 0;
}

// Number of the wave's 64 lanes whose nondecreasing `ends` are <= t (0..64):
// a binary search with seven cross-lane reads.
__device__ __forceinline__ int wave_search(int ends, int t)
{
    int pos = 0;
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) {
        const int e = __shfl(ends, pos + st - 1, 64);
        if (e <= t) pos += st;
    }
    const int e63 = __shfl(ends, 63, 64);
    return e63 <= t ? 64 : pos;
}

// Inclusive prefix sum over the wave's 64 lanes.
__device__ __forceinline__ int wave_scan_inclusive(int v, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// map_tile for 64 samples at once (lane = sample base + lane): a window of 64
// candidate intervals from the cursor (every interval holds at least one
// sample, so 64 samples span at most 64 intervals), their counts scanned
// across the wave; the cursor advances past the intervals the 64 samples
// finish.  tc.cnt is not used.
__device__ __forceinline__ void map_tile64(const float4* __restrict__ rec, int ie, int base, int tot, int lane,
                                           TileCursor& tc, TileMap& m)
{
    const int t = base + lane;
    const int cand = tc.cur + lane;
    const int c_cnt = cand < ie ? __float_as_int(rec[3 * (int64_t)cand].z) : 0;
    const bool ones = __all(cand >= ie || c_cnt == 1);
    int c_end = c_cnt;
    int pos, posn, s_cnt, s_end, last_end = 0;
    if (ones) {
        pos = lane;
        posn = min(64, ie - tc.cur);
        s_cnt = 1;
        s_end = tc.cum + lane + 1;
    } else {
        c_end = wave_scan_inclusive(c_end, lane);
        c_end = cand < ie ? c_end + tc.cum : INT_MAX;
        pos = wave_search(c_end, t);
        posn = wave_search(c_end, base + 64);
        const int src = pos < 64 ? pos : 63;
        s_cnt = __shfl(c_cnt, src, 64);
        s_end = __shfl(c_end, src, 64);
        last_end = __shfl(c_end, 63, 64);
    }
    m.valid = t < tot;
    if (m.valid) {
        int k = tc.cur + pos, n_s = s_cnt, start = s_end - s_cnt;
        if (pos >= 64) {
            // only with zero-sample intervals in the list: walk on
            k = tc.cur + 64;
            start = last_end;
            n_s = __float_as_int(rec[3 * (int64_t)k].z);
            while (start + n_s <= t) {
                start += n_s;
                ++k;
                n_s = __float_as_int(rec[3 * (int64_t)k].z);
            }
        }
        m.n_s = n_s;
        m.jj = t - start;
        m.A = rec[3 * (int64_t)k];
        m.B = rec[3 * (int64_t)k + 1];
        m.D = rec[3 * (int64_t)k + 2];
    }
    if (ones)
        tc.cum += posn;
    else if (posn > 0)
        tc.cum = __shfl(c_end, posn - 1, 64);
    tc.cur += posn;
}

// Raw corner records of one lane's CPL channels (f32, or bf16 pairs), as
// loaded; decoded by corners_combine.
template <int C, bool kBF16>
struct Corners {
    static constexpr int CPL = C / 4;
    static constexpr int W = kBF16 ? CPL / 2 : CPL;  // dwords per corner
    uint32_t u[8][W];
};

// Stage 2: places this lane's sample, issues the gathers of its 8 corners
// (channels 4g.., C = 16; 2g.., C = 8) and returns dL/dP of its ray (0 for
// a lane without a sample).
template <int C, bool kBF16, bool kVM = false>
__device__ __forceinline__ float sample_stage(const ProjArgs& a, const float* __restrict__ dLdP, const PNode& nd,
                                              int64_t gbase, const TileMap& m, int g, SampleCtx& sc,
                                              Corners<C, kBF16>& cr, bool vm_chunk = false)
{
    constexpr int CPL = C / 4;
    constexpr int W = Corners<C, kBF16>::W;
    sc.valid = m.valid;
    sc.w = 0.f;
    sc.gbase = -1;
    if (!m.valid) return 0.f;
    const float o[3] = {m.B.x, m.B.y, m.B.z};
    const float d[3] = {m.B.w, m.D.x, m.D.y};
    const int r = __float_as_int(m.D.w);
    make_sample_node<kVM>(a, nd, gbase, r, o, d, m.A.x, m.A.y, m.n_s, m.jj, __float_as_int(m.A.w) + m.jj, sc);
    const float dp = dLdP[r];
    if (kVM && vm_chunk) return dp;  // tensorial reads happen in stage 3
    const unsigned char* grid = reinterpret_cast<const unsigned char*>(a.grid);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int64_t off = sc.gbase + corner_cell(sc.tri, k, a.G) * C + CPL * g;
        const unsigned char* p = grid + off * (kBF16 ? 2 : 4);
        if constexpr (W == 4) {
            const uint4 v = *reinterpret_cast<const uint4*>(p);
            cr.u[k][0] = v.x;
            cr.u[k][1] = v.y;
            cr.u[k][2] = v.z;
            cr.u[k][3] = v.w;
        } else if constexpr (W == 2) {
            const uint2 v = *reinterpret_cast<const uint2*>(p);
            cr.u[k][0] = v.x;
            cr.u[k][1] = v.y;
        } else {
            cr.u[k][0] = *reinterpret_cast<const uint32_t*>(p);
        }
    }
    return dp;
}

// The gathers of sample_stage for a placed sample: its 8 corner records at
// channels CPL g .. (C = 16: 4g.., C = 8: 2g..).
template <int C, bool kBF16>
__device__ __forceinline__ void gather_corners(const ProjArgs& a, const SampleCtx& sc, int g, Corners<C, kBF16>& cr)
{
    constexpr int CPL = C / 4;
    constexpr int W = Corners<C, kBF16>::W;
    const unsigned char* grid = reinterpret_cast<const unsigned char*>(a.grid);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int64_t off = sc.gbase + corner_cell(sc.tri, k, a.G) * C + CPL * g;
        const unsigned char* p = grid + off * (kBF16 ? 2 : 4);
        if constexpr (W == 4) {
            const uint4 v = *reinterpret_cast<const uint4*>(p);
            cr.u[k][0] = v.x;
            cr.u[k][1] = v.y;
            cr.u[k][2] = v.z;
            cr.u[k][3] = v.w;
        } else if constexpr (W == 2) {
            const uint2 v = *reinterpret_cast<const uint2*>(p);
            cr.u[k][0] = v.x;
            cr.u[k][1] = v.y;
        } else {
            cr.u[k][0] = *reinterpret_cast<const uint32_t*>(p);
        }
    }
}

// Trilinear combination of the loaded corners: canonical corner order per
// channel, accumulated as a fused multiply-add chain (as gather_channels).
template <int C, bool kBF16>
__device__ __forceinline__ void corners_combine(const SampleCtx& sc, const Corners<C, kBF16>& cr, float* x)
{
    constexpr int CPL = C / 4;
#pragma unroll
    for (int q = 0; q < CPL; ++q) x[q] = 0.f;
    if (!sc.valid) return;
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float v;
            if constexpr (kBF16)
                v = (q & 1) ? __uint_as_float(cr.u[k][q >> 1] & 0xffff0000u) : __uint_as_float(cr.u[k][q >> 1] << 16);
            else
                v = __uint_as_float(cr.u[k][q]);
            acc = fmaf(v, sc.tri.w[k], acc);
        }
        x[q] = acc;
    }
}

// Adds w_corner * (d0, d1, d2, d3) into the LDS node gradient at channels
// 4g .. 4g+3 of the sample's 8 corners.  `plain` lanes (a sample off the
// grid border whose floor cell no other lane of its 16-lane row shares)
// read-add-write float4; `atomic` lanes (a shared cell, or a clamped corner
// at the border) add afterwards with LDS atomics.  The caller keeps other
// waves out (workgroup lock).
template <int C>
__device__ __forceinline__ void scatter_corners(float* lgrad, const Tri& tri, int G, int g, float d0, float d1, float d2,
                                                float d3, bool plain, bool atomic)
{
    if (plain) {
        // corner by corner: a corner of one lane can be another corner of a
        // lane whose floor cell is adjacent, and LDS runs a wave's operations
        // in order, so each read sees the previous corners' writes
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float4* gp = reinterpret_cast<float4*>(lgrad + corner_cell(tri, k, G) * C + 4 * g);
            const float wk = tri.w[k];
            float4 v = *gp;
            v.x = fmaf(wk, d0, v.x);
            v.y = fmaf(wk, d1, v.y);
            v.z = fmaf(wk, d2, v.z);
            v.w = fmaf(wk, d3, v.w);
            *gp = v;
        }
    }
    wave_lds_sync();
    if (atomic) {
        const float dv[4] = {d0, d1, d2, d3};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float* gp = lgrad + corner_cell(tri, k, G) * C + 4 * g;
            const float wk = tri.w[k];
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) atomicAdd(gp + r4, wk * dv[r4]);
        }
    }
}

// Workgroup lock in LDS guarding the node gradient.  Lane 0 polls with a
// compare-and-swap until it gets the lock.  The wait ends: a holder runs a
// fixed sequence of LDS read-add-writes (it waits on no other wave and on no
// global memory) and releases, and the waves of a workgroup take the lock a
// finite number of times (once per sample tile of the node's chunk), so a
// waiting wave gets it at the latest when the other waves have finished
// their scatters.
__device__ __forceinline__ void lds_lock(int* lock, int lane)
{
    if (lane == 0) {
        while (true) {
            int expected = 0;
            if (__hip_atomic_compare_exchange_strong(lock, &expected, 1, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                break;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void lds_unlock(int* lock, int lane)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(lock, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// End of a chunk: adds the LDS node gradient (n4 float4) to HBM once and
// clears it for the next chunk (float4 accesses, kU in flight per thread).
// adam_here: the chunk holds all of the node's samples, so its gradient is
// complete and Adam is applied to the node's grid (parameters at gbase)
// instead of going through the gradient buffer.  split: other chunks of the
// node add to the same gradient, through device-scope atomics.
// Channel-pair swizzle of the scatter passes' LDS gradients (cell-major, C
// floats per cell): channels c, c + 1 of cell x sit at c ^ pair_swz(x).  The
// low bits of x already move a cell across bank rows (64 / C cells fill the
// 64 banks of one row), so the swizzle takes the bits above them: lanes of
// one float2 access at random cells then reach every bank pair with 32 or
// with 64 banks, where 2 (x mod C/2) reached half of them.
__device__ __forceinline__ int pair_swz(int cell, int C)
{
    const int sh = C >= 16 ? 2 : (C == 8 ? 3 : 4);
    return 2 * ((cell >> sh) & (C / 2 - 1));
}

template <int kSwzC>
__device__ __forceinline__ float4 take_lds4(float* lgrad, int i)
{
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (kSwzC == 0) {
        float4* lg4 = reinterpret_cast<float4*>(lgrad);
        const float4 v = lg4[i];
        lg4[i] = z4;
        return v;
    } else {
        // float4 i = channels 4q .. 4q+3 of cell i / (C/4), stored as two
        // swizzled channel pairs (the scatter pass's layout)
        constexpr int Q = kSwzC / 4;
        const int cell = i / Q, q = i % Q;
        const int swz = pair_swz(cell, kSwzC);
        float2* a = reinterpret_cast<float2*>(lgrad + cell * kSwzC + ((4 * q) ^ swz));
        float2* b = reinterpret_cast<float2*>(lgrad + cell * kSwzC + ((4 * q + 2) ^ swz));
        const float2 va = *a, vb = *b;
        *a = make_float2(0.f, 0.f);
        *b = make_float2(0.f, 0.f);
        return make_float4(va.x, va.y, vb.x, vb.y);
    }
}

template <int kU, int kSwzC = 0>
__device__ __forceinline__ void flush_node_grad(const GridAdam& ad, bool adam_here, bool split, float* lgrad,
                                                float* gdst, int64_t gbase, int n4)
{
    float4* gg4 = reinterpret_cast<float4*>(gdst);
    if (adam_here) {
        float4* p4 = reinterpret_cast<float4*>(ad.p + gbase);
        float4* m4 = reinterpret_cast<float4*>(ad.m + gbase);
        float4* v4 = reinterpret_cast<float4*>(ad.v + gbase);
        constexpr int kA = 4;
        for (int i0 = threadIdx.x; i0 < n4; i0 += kA * blockDim.x) {
            float4 gv[kA], pv[kA], mv[kA], vv[kA];
#pragma unroll
            for (int u = 0; u < kA; ++u) {
                const int i = i0 + u * blockDim.x;
                if (i < n4) {
                    gv[u] = take_lds4<kSwzC>(lgrad, i);
                    pv[u] = p4[i];
                    mv[u] = m4[i];
                    vv[u] = v4[i];
                }
            }
#pragma unroll
            for (int u = 0; u < kA; ++u) {
                const int i = i0 + u * blockDim.x;
                if (i >= n4) continue;
                float* gp = &gv[u].x;
                float* pp = &pv[u].x;
                float* mp = &mv[u].x;
                float* vp = &vv[u].x;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    pp[k] = adam_update(pp[k], gp[k], mp[k], vp[k], ad.lr, ad.b1, ad.b2, ad.eps, ad.bc1, ad.bc2s);
                p4[i] = pv[u];
                m4[i] = mv[u];
                v4[i] = vv[u];
                if (ad.pb) {
                    ushort4 h;
                    h.x = f32_to_bf16_rne(pp[0]);
                    h.y = f32_to_bf16_rne(pp[1]);
                    h.z = f32_to_bf16_rne(pp[2]);
                    h.w = f32_to_bf16_rne(pp[3]);
                    reinterpret_cast<ushort4*>(ad.pb + gbase)[i] = h;
                }
            }
        }
        return;
    }
    for (int i0 = threadIdx.x; i0 < n4; i0 += kU * blockDim.x) {
        float4 v[kU], o[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < n4) {
                v[u] = take_lds4<kSwzC>(lgrad, i);
                if (!split) o[u] = gg4[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i >= n4) continue;
            if (split) {
                float* dst = reinterpret_cast<float*>(gg4 + i);
                if (v[u].x != 0.f) unsafeAtomicAdd(dst + 0, v[u].x);
                if (v[u].y != 0.f) unsafeAtomicAdd(dst + 1, v[u].y);
                if (v[u].z != 0.f) unsafeAtomicAdd(dst + 2, v[u].z);
                if (v[u].w != 0.f) unsafeAtomicAdd(dst + 3, v[u].w);
            } else {
                gg4[i] = make_float4(o[u].x + v[u].x, o[u].y + v[u].y, o[u].z + v[u].z, o[u].w + v[u].w);
            }
        }
    }
}

template <int C, bool kBF16, bool kDecBF16, bool kVM = false>
__global__ __launch_bounds__(64 * bwd_waves<kDecBF16>()) void project_bwd_node_kernel(NodeBwdArgs b)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dsm[];
    __shared__ int s_chunk;
    __shared__ int s_lock;
    if (b.mode && *b.mode) return;  // the two-pass kernels run instead
    constexpr int kNW = bwd_waves<kDecBF16>();
    const ProjArgs& a = b.a;
    const float4* __restrict__ rec = b.rec;
    constexpr int CPL = C / 4;
    constexpr int NP = mlp_numel<C>();
    const int G = a.G;
    const int gsz = G * G * G * C;
    float* lgrad = reinterpret_cast<float*>(dsm);
    float* s = reinterpret_cast<float*>(dsm) + b.lds_f32_off;
    __bf16* sb = reinterpret_cast<__bf16*>(dsm + b.lds_bf16_off);
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int wl = threadIdx.x >> 6;
    float* bufA = reinterpret_cast<float*>(dsm + b.lds_tb_off) + wl * 2 * kH * kTS;
    float* bufB = bufA + kH * kTS;
    __bf16* bufAb = reinterpret_cast<__bf16*>(dsm + b.lds_tb_off) + wl * 2 * kH * kTB;
    __bf16* bufBb = bufAb + kH * kTB;

    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds<C>(a.mlp, s);
    for (int i = threadIdx.x; i < b.lds_grad_floats; i += blockDim.x) lgrad[i] = 0.f;
    if (threadIdx.x == 0) s_lock = 0;
    // tensorial chunks (slots >= NG) come after every explicit one in the
    // chunk table, and a workgroup takes chunks in increasing order, so once
    // it reaches them the line partials at lds_line_off (zero until then:
    // explicit chunks clear their gradient after each flush) stay in LDS
    // until the workgroup ends
    float* lline = lgrad + b.lds_line_off;
    const int Gp = kVM ? a.vm.Gp : 0, Gl = kVM ? a.vm.Gl : 0;

    DecGrads q;
    zero_grads(q);
    const int64_t total_chunks = b.chunk_off[b.num_slots];
    __syncthreads();

    while (true) {
        if (threadIdx.x == 0) s_chunk = atomicAdd(b.counter, 1);
        __syncthreads();
        const int64_t c = s_chunk;
        __syncthreads();
        if (c >= total_chunks) break;
        const int4 ch = b.chunks[c];
        const bool vmc = kVM && ch.x >= a.vm.NG;
        const int pslot = vmc ? (int)(ch.x - a.vm.NG) : -1;
        const int64_t gbase = vmc ? 0 : (int64_t)ch.x * gsz;
        const PNode nd = a.nodes[a.iv_node[b.sorted[ch.y]]];  // one node per chunk
        const int n_iv = ch.z - ch.y;
        const int per = (n_iv + kNW - 1) / kNW;
        const int ib = ch.y + wl * per;
        const int ie = min(ib + per, ch.z);
        // samples of this wave's intervals
        int tot = 0;
        for (int k = ib + lane; k < ie; k += 64) tot += __float_as_int(rec[3 * (int64_t)k].z);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
        const int ntiles = (tot + kTile - 1) / kTile;

        // Software-pipelined tile loop: while tile ti runs the decoder and
        // the scatter, the corner gathers of tile ti+1, the interval records
        // of tile ti+2 and the interval counts of tile ti+3 are in flight.
        const int grp = lane & 48;
        TileCursor tc;
        tc.cur = ib;
        tc.cum = 0;
        tc.cnt = ib + j < ie ? __float_as_int(rec[3 * (int64_t)(ib + j)].z) : 0;
        TileMap m1;
        m1.valid = false;
        if (ntiles > 0) map_tile(rec, ie, 0, tot, j, grp, tc, m1);
        SampleCtx sc0;
        Corners<C, kBF16> cr0;
        float dp0 = sample_stage<C, kBF16, kVM>(a, b.dLdP, nd, gbase, m1, g, sc0, cr0, vmc);
        m1.valid = false;
        if (ntiles > 1) map_tile(rec, ie, 1, tot, j, grp, tc, m1);
        for (int ti = 0; ti < ntiles; ++ti) {
            SampleCtx sc1;
            Corners<C, kBF16> cr1;
            const float dp1 = sample_stage<C, kBF16, kVM>(a, b.dLdP, nd, gbase, m1, g, sc1, cr1, vmc);
            m1.valid = false;
            if (ti + 2 < ntiles) map_tile(rec, ie, ti + 2, tot, j, grp, tc, m1);

            const SampleCtx& sc = sc0;
            float x[CPL];
            VMSetup vst;
            float Pv[3][CPL], Lv[3][CPL];
            if (kVM && vmc) {
#pragma unroll
                for (int qq = 0; qq < CPL; ++qq) x[qq] = 0.f;
                if (sc.valid) {
                    vm_setup(a, sc, vst);
                    vm_reads<C, CPL>(a, pslot, vst, CPL * g, Pv, Lv);
                    vm_combine<CPL>(Pv, Lv, x);
                }
            } else {
                corners_combine<C, kBF16>(sc, cr0, x);
            }
            v4f h1[4], h2[4];
            float y;
            if constexpr (kDecBF16)
                y = decoder_fwd_bf16<C>(sb, s, x, g, j, h1, h2);
            else
                y = decoder_fwd<C>(s, x, g, j, h1, h2);
            const float dy = sc.valid ? dp0 * sc.w * softplus_grad_f(y, a.beta) : 0.f;
            v4f dx = tile_backward<C, kDecBF16>(s, sb, bufA, bufB, bufAb, bufBb, x, h1, h2, dy, g, j, q);

            // Scatter w_corner * dX into the LDS node gradient.  One wave at a
            // time holds the lock, so waves never interleave their updates,
            // and the others keep computing meanwhile.  Inside a wave one
            // instruction can only collide between samples of the same floor
            // cell (LDS runs a wave's operations in order), so samples with a
            // unique cell in the tile use a plain float4 read-add-write and the
            // rest (shared cell, or a clamped corner at the grid border) use
            // LDS atomics.
            if (kVM && vmc) {
                // dL/dplane = w dX L, dL/dline = w dX P into the LDS plane
                // gradient and the line partials, all by LDS atomics (no
                // lock: no plain read-add-write in this chunk)
                if (sc.valid && dy != 0.f && 4 * g < C) {
                    float P4[3][4], L4[3][4];
                    if constexpr (C == 16) {
#pragma unroll
                        for (int pp = 0; pp < 3; ++pp)
#pragma unroll
                            for (int qq = 0; qq < 4; ++qq) {
                                P4[pp][qq] = Pv[pp][qq];
                                L4[pp][qq] = Lv[pp][qq];
                            }
                    } else {
                        vm_reads<C, 4>(a, pslot, vst, 4 * g, P4, L4);
                    }
#pragma unroll
                    for (int pp = 0; pp < 3; ++pp) {
                        const Bil& bb = vst.b[pp];
                        float* gpl = lgrad + pp * Gp * Gp * C + 4 * g;
                        const int off[4] = {(bb.y0 * Gp + bb.x0) * C, (bb.y0 * Gp + bb.x1) * C,
                                            (bb.y1 * Gp + bb.x0) * C, (bb.y1 * Gp + bb.x1) * C};
                        const Lin& ll = vst.l[pp];
                        float* gl = lline + pp * Gl * C + 4 * g;
#pragma unroll
                        for (int qq = 0; qq < 4; ++qq) {
                            const float dP = dx[qq] * L4[pp][qq];
                            const float dL = dx[qq] * P4[pp][qq];
#pragma unroll
                            for (int k = 0; k < 4; ++k) atomicAdd(gpl + off[k] + qq, bb.w[k] * dP);
                            atomicAdd(gl + ll.x0 * C + qq, ll.w0 * dL);
                            atomicAdd(gl + ll.x1 * C + qq, ll.w1 * dL);
                        }
                    }
                }
                sc0 = sc1;
                cr0 = cr1;
                dp0 = dp1;
                continue;
            }
            const bool live = sc.valid && dy != 0.f;
            const bool border = sc.tri.x0 == sc.tri.x1 || sc.tri.y0 == sc.tri.y1 || sc.tri.z0 == sc.tri.z1;
            const int key = live ? (border ? -1 : (sc.tri.z0 * G + sc.tri.y0) * G + sc.tri.x0) : -2 - j;
            const bool dup = dup16(key);
            const bool plain = live && !border && !dup && 4 * g < C;
            const bool atomic = live && (border || dup) && 4 * g < C;
            if (__any(plain || atomic)) {
                lds_lock(&s_lock, lane);
                scatter_corners<C>(lgrad, sc.tri, G, g, dx[0], dx[1], dx[2], dx[3], plain, atomic);
                lds_unlock(&s_lock, lane);
            }
            sc0 = sc1;
            cr0 = cr1;
            dp0 = dp1;
        }
        __syncthreads();
        // flush the node gradient once, clear LDS for the next chunk (G^3 C,
        // or a tensorial chunk's three planes 3 Gp^2 C: a multiple of 4
        // floats for C in {8, 16})
        float* gdst = vmc ? b.grad_planes + (int64_t)pslot * 3 * Gp * Gp * C : b.grad_grid + gbase;
        const int n4 = (vmc ? 3 * Gp * Gp * C : gsz) >> 2;
        flush_node_grad<kDecBF16 ? 16 : 4>(b.ad, b.ad.enabled && !ch.w && !vmc, ch.w != 0, lgrad, gdst, gbase, n4);
        __syncthreads();
    }
    if constexpr (kVM) {
        // this workgroup's line partials (all zero if it took no tensorial
        // chunk), summed over workgroups in a fixed order afterwards
        float* row = b.line_part + (int64_t)blockIdx.x * 3 * Gl * C;
        for (int i = threadIdx.x; i < 3 * Gl * C; i += blockDim.x) row[i] = lline[i];
    }
    store_partials<C>(q, b.ws + ((int64_t)blockIdx.x * kNW + wl) * NP, g, j, lane);
}

// ---------------------------------------------------------------------------
// Two-pass adjoint.  The fused kernel above keeps a node's whole gradient
// (G^3 C f32, 85 KB at G = 11, C = 16) in LDS next to the decoder, so a CU
// runs one workgroup and each SIMD one wave, and the decoder's dependent
// LDS and MFMA steps run without another wave to overlap them.  With a
// staging buffer in the workspace (atdf_project_bwd_staging_bytes) the same
// work runs as two kernels over the same counting sort and chunk table:
//
// (1) project_bwd_dec_kernel: the decoder adjoint.  No node gradient in LDS
//     (weights and per-wave transpose buffers only) and at most 256
//     registers per wave, so a CU runs eight waves, two per SIMD.  Waves
//     take work items of about kPartSamples samples (a run of one chunk's
//     sorted intervals) from a global counter and write, per sample, its
//     lattice coordinates and dX = dL/dx (the dL/dy-weighted feature
//     gradient, f32) at the sample's position in the sorted order.
//     Decoder-gradient partials per wave, reduced in a fixed order
//     afterwards.
// (2) project_bwd_scatter_kernel: per chunk, C/2 waves (two channels each)
//     add w_corner * dX of the staged samples into the LDS node gradient,
//     then flush it (and apply Adam) as the fused kernel does.
//
// Staged bytes per sample: 16 (lattice coordinates) + 4 C (dX).
// ---------------------------------------------------------------------------
#ifndef ATDF_PART_SAMPLES
#define ATDF_PART_SAMPLES 256
#endif
constexpr int64_t kPartSamples = ATDF_PART_SAMPLES;
constexpr int kMinPartIv = 16;    // intervals per decoder-pass work item, at least
// Staged records past the list's last sample: the scatter pass reads up to
// 64 + 2 records beyond a chunk's end without a branch.
constexpr int64_t kStagePad = 128;
constexpr int kDecWaves = 8;  // one workgroup per CU, two waves per SIMD

// Work items of a chunk of n_iv intervals holding `samples` samples.
__host__ __device__ inline int parts_of(int n_iv, int64_t samples)
{
    if (n_iv <= 0) return 0;
    int64_t p = (samples + kPartSamples - 1) / kPartSamples;
    const int64_t pmax = (n_iv + kMinPartIv - 1) / kMinPartIv;
    p = p < pmax ? p : pmax;
    p = p > 1 ? p : 1;
    const int per = (int)((n_iv + p - 1) / p);
    return (n_iv + per - 1) / per;
}

// Staged records of an in-ROI sample of a tensorial field (float4 each):
// a plane record {local x, y, z, 0} + dL/dplane_p / w = dX * L_p for the
// planes yz, xz, xy (C floats each), in node-sorted order, and a line record
// {ROI-normalised x, local y, global z, 0} + dL/dline_p / w = dX * P_p for
// the lines x, y, z, in the line pass's interleaved order: with n in-ROI
// samples and span = ceil(n / 64), sample v sits at (v mod span) 64 +
// v / span, so that the 64 lanes of a line-pass step (lane l walking
// samples [l span, (l + 1) span)) read 64 consecutive records.
template <int C>
__host__ __device__ constexpr int vm_record_f4()
{
    return 1 + 3 * C / 4;
}

struct VMStage {
    float4* plane;  // [n + kStagePad] plane records
    float4* line;   // [64 span + 2 kStagePad] line records
    int64_t s_vm0;  // first in-ROI sample in the sorted order
    int64_t n, span;
};

// The in-ROI regions follow the explicit records (and their padding); the
// line region follows the plane region (and its padding).
template <int C>
__device__ __forceinline__ VMStage vm_stage(const NodeBwdArgs& b)
{
    VMStage v;
    v.s_vm0 = b.soff[b.slot_off[b.exp_slots]];
    v.n = b.soff[b.slot_off[b.num_slots]] - v.s_vm0;
    v.span = (v.n + 63) / 64;
    v.plane = b.stage + (v.s_vm0 + kStagePad) * (1 + C / 4);
    v.line = v.plane + (v.n + kStagePad) * vm_record_f4<C>();
    return v;
}

// (32-bit division: a list holds fewer than 2^31 samples)
__device__ __forceinline__ int64_t vm_line_pos(int64_t v, int64_t span)
{
    const unsigned q = (unsigned)v / (unsigned)span, r = (unsigned)v - q * (unsigned)span;
    return (int64_t)r * 64 + q;
}

template <int C, bool kBF16, bool kDecBF16, bool kVM = false>
__global__ __launch_bounds__(64 * kDecWaves, 2) void project_bwd_dec_kernel(NodeBwdArgs b)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dsm[];
    if (*b.mode == 0) return;
    const ProjArgs& a = b.a;
    const float4* __restrict__ rec = b.rec;
    constexpr int CPL = C / 4;
    constexpr int NP = mlp_numel<C>();
    constexpr int RF = 1 + C / 4;
    constexpr int RFV = vm_record_f4<C>();
    const int G = a.G;
    const int gsz = G * G * G * C;
    float* s = reinterpret_cast<float*>(dsm) + b.lds_f32_off;
    __bf16* sb = reinterpret_cast<__bf16*>(dsm + b.lds_bf16_off);
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int wl = threadIdx.x >> 6;
    float* bufA = reinterpret_cast<float*>(dsm + b.lds_tb_off) + wl * 2 * kH * kTS;
    float* bufB = bufA + kH * kTS;
    __bf16* bufAb = reinterpret_cast<__bf16*>(dsm + b.lds_tb_off) + wl * 2 * kH * kTB;
    __bf16* bufBb = bufAb + kH * kTB;
    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds_t<C>(a.mlp, s);
    VMStage vs{};
    if constexpr (kVM) vs = vm_stage<C>(b);
    __syncthreads();

    DecGradsRS q;
    zero_grads(q);
    // the explicit-grid items come first in the part table, the in-ROI
    // items of a tensorial field after them; each instance takes its own
    const int64_t split_item = b.part_off[b.chunk_off[b.exp_slots]];
    const int64_t i0 = kVM ? split_item : 0;
    const int64_t i1 = kVM ? b.part_off[b.max_chunks] : split_item;
    int* const counter = kVM ? b.counter5 : b.counter2;
    while (true) {
        int item = 0;
        if (lane == 0) item = (int)i0 + atomicAdd(counter, 1);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= i1) break;
        const int4 it = b.parts[item];
        const PNode nd = a.nodes[a.iv_node[b.sorted[it.y]]];
        constexpr bool vmc = kVM;
        const int pslot = vmc ? it.x - a.vm.NG : -1;
        const int64_t gbase = vmc ? 0 : (int64_t)it.x * gsz;
        const int ib = it.y, ie = it.z;
        const int64_t s_base = b.soff[ib];
        const int tot = (int)(b.soff[ie] - s_base);
        const int64_t v_base = s_base - vs.s_vm0;  // in-ROI index of the item's first sample

        // 64 samples per step.  (A) lane l maps and places sample l of the
        // step once (the interval search, jitter hashes, divisions and
        // trilinear set-up are done by one lane per sample instead of the four
        // lane rows of a tile), writes its lattice coordinates to the stage
        // and its packed corners, weights and dL/dP * w to the wave's sample
        // table in LDS; (B) the step's four tiles of 16 read the table, gather,
        // run the decoder and its adjoint and stage dX.  An in-ROI sample of a
        // tensorial field keeps its local and global coordinates in the table
        // instead, gathers its planes and lines, and stages dX * L_p and
        // dX * P_p (the plane and line gradients before the corner weights).
        float4* pl = reinterpret_cast<float4*>(dsm + b.lds_pl_off) + wl * 64 * 3;
        TileCursor tc;
        tc.cur = ib;
        tc.cum = 0;
        tc.cnt = 0;
        const int nsteps = (tot + 63) / 64;
        for (int st = 0; st < nsteps; ++st) {
            const int base = st * 64;
            {
                TileMap m;
                m.valid = false;
                map_tile64(rec, ie, base, tot, lane, tc, m);
                int packed = -1;
                float dpw = 0.f;
                SampleCtx sc;
                sc.valid = m.valid;
                if (kVM && vmc) {
                    sc.loc[0] = sc.loc[1] = sc.loc[2] = sc.gx = sc.gz = 0.f;
                    if (m.valid) {
                        const float o[3] = {m.B.x, m.B.y, m.B.z};
                        const float d[3] = {m.B.w, m.D.x, m.D.y};
                        const int r = __float_as_int(m.D.w);
                        make_sample_node<true>(a, nd, 0, r, o, d, m.A.x, m.A.y, m.n_s, m.jj,
                                               __float_as_int(m.A.w) + m.jj, sc, false);
                        dpw = b.dLdP[r] * sc.w;
                        // (the line x coordinate as vm_setup forms it)
                        const float xr = (2.0f * (sc.gx - a.vm.rx0)) / (a.vm.rx1 - a.vm.rx0) - 1.0f;
                        const int64_t v = v_base + base + lane;
                        vs.plane[v * RFV] = make_float4(sc.loc[0], sc.loc[1], sc.loc[2], 0.f);
                        vs.line[vm_line_pos(v, vs.span) * RFV] = make_float4(xr, sc.loc[1], sc.gz, 0.f);
                    }
                    pl[lane * 3] = make_float4(sc.loc[0], sc.loc[1], sc.loc[2], sc.gx);
                    pl[lane * 3 + 1] = make_float4(sc.gz, dpw, m.valid ? 1.f : 0.f, 0.f);
                } else {
                    if (m.valid) {
                        const float o[3] = {m.B.x, m.B.y, m.B.z};
                        const float d[3] = {m.B.w, m.D.x, m.D.y};
                        const int r = __float_as_int(m.D.w);
                        make_sample_node(a, nd, gbase, r, o, d, m.A.x, m.A.y, m.n_s, m.jj,
                                         __float_as_int(m.A.w) + m.jj, sc);
                        dpw = b.dLdP[r] * sc.w;
                        packed = sc.tri.x0 | (sc.tri.x1 << 5) | (sc.tri.y0 << 10) | (sc.tri.y1 << 15) |
                                 (sc.tri.z0 << 20) | (sc.tri.z1 << 25);
                        b.stage[(s_base + base + lane) * RF] = make_float4(sc.li[0], sc.li[1], sc.li[2], 0.f);
                    }
                    pl[lane * 3] = make_float4(__int_as_float(packed), sc.tri.w[0], sc.tri.w[1], sc.tri.w[2]);
                    pl[lane * 3 + 1] = make_float4(sc.tri.w[3], sc.tri.w[4], sc.tri.w[5], sc.tri.w[6]);
                    pl[lane * 3 + 2] = make_float4(sc.tri.w[7], dpw, 0.f, 0.f);
                }
            }
            wave_lds_sync();
            const int ntile = min(4, (tot - base + kTile - 1) / kTile);
            if (kVM && vmc) {
                for (int q4 = 0; q4 < ntile; ++q4) {
                    const int e = q4 * kTile + j;
                    SampleCtx sc;
                    float dpw;
                    {
                        const float4 p0 = pl[e * 3], p1 = pl[e * 3 + 1];
                        sc.valid = p1.z != 0.f;
                        sc.loc[0] = p0.x;
                        sc.loc[1] = p0.y;
                        sc.loc[2] = p0.z;
                        sc.gx = p0.w;
                        sc.gz = p1.x;
                        dpw = p1.y;
                    }
                    float x[CPL];
#pragma unroll
                    for (int qq = 0; qq < CPL; ++qq) x[qq] = 0.f;
                    if (sc.valid) {
                        // F = Pyz Lx + Pxz Ly + Pxy Lz, one pair at a time
#pragma unroll
                        for (int pp = 0; pp < 3; ++pp) {
                            float Pv[CPL], Lv[CPL];
                            vm_pair_read<C, CPL>(a, pslot, sc, pp, CPL * g, Pv, Lv);
#pragma unroll
                            for (int qq = 0; qq < CPL; ++qq) x[qq] = pp == 0 ? Pv[qq] * Lv[qq] : x[qq] + Pv[qq] * Lv[qq];
                        }
                    }
                    v4f h1[4], h2[4];
                    float y;
                    if constexpr (kDecBF16)
                        y = decoder_fwd_bf16<C>(sb, s, x, g, j, h1, h2);
                    else
                        y = decoder_fwd<C>(s, x, g, j, h1, h2);
                    const float dy = sc.valid ? dpw * softplus_grad_f(y, a.beta) : 0.f;
                    const v4f dx = tile_backward<C, kDecBF16, DecGradsRS, !kDecBF16>(s, sb, bufA, bufB, bufAb, bufBb,
                                                                                    x, h1, h2, dy, g, j, q);
                    if (sc.valid && 4 * g < C) {
                        // the plane and line reads again (cache hits) rather
                        // than kept live through the decoder
                        const int64_t v = v_base + base + e;
                        float4* rp = vs.plane + v * RFV + 1;
                        float4* rl = vs.line + vm_line_pos(v, vs.span) * RFV + 1;
#pragma unroll
                        for (int pp = 0; pp < 3; ++pp) {
                            float P4[4], L4[4];
                            vm_pair_read<C, 4>(a, pslot, sc, pp, 4 * g, P4, L4);
                            rp[pp * (C / 4) + g] =
                                make_float4(dx[0] * L4[0], dx[1] * L4[1], dx[2] * L4[2], dx[3] * L4[3]);
                            rl[pp * (C / 4) + g] =
                                make_float4(dx[0] * P4[0], dx[1] * P4[1], dx[2] * P4[2], dx[3] * P4[3]);
                        }
                    }
                }
                wave_lds_sync();
                continue;
            }
            for (int q4 = 0; q4 < ntile; ++q4) {
                const int e = q4 * kTile + j;
                const float4 p0 = pl[e * 3], p1 = pl[e * 3 + 1], p2 = pl[e * 3 + 2];
                const int packed = __float_as_int(p0.x);
                SampleCtx sc;
                sc.valid = packed >= 0;
                sc.tri.x0 = packed & 31;
                sc.tri.x1 = (packed >> 5) & 31;
                sc.tri.y0 = (packed >> 10) & 31;
                sc.tri.y1 = (packed >> 15) & 31;
                sc.tri.z0 = (packed >> 20) & 31;
                sc.tri.z1 = (packed >> 25) & 31;
                sc.tri.w[0] = p0.y;
                sc.tri.w[1] = p0.z;
                sc.tri.w[2] = p0.w;
                sc.tri.w[3] = p1.x;
                sc.tri.w[4] = p1.y;
                sc.tri.w[5] = p1.z;
                sc.tri.w[6] = p1.w;
                sc.tri.w[7] = p2.x;
                sc.gbase = gbase;
                float x[CPL];
                if (sc.valid) {
                    Corners<C, kBF16> cr;
                    gather_corners<C, kBF16>(a, sc, g, cr);
                    corners_combine<C, kBF16>(sc, cr, x);
                } else {
#pragma unroll
                    for (int qq = 0; qq < CPL; ++qq) x[qq] = 0.f;
                }
                v4f h1[4], h2[4];
                float y;
                if constexpr (kDecBF16)
                    y = decoder_fwd_bf16<C>(sb, s, x, g, j, h1, h2);
                else
                    y = decoder_fwd<C>(s, x, g, j, h1, h2);
                const float dy = sc.valid ? p2.y * softplus_grad_f(y, a.beta) : 0.f;
                const v4f dx =
                    tile_backward<C, kDecBF16, DecGradsRS, !kDecBF16>(s, sb, bufA, bufB, bufAb, bufBb, x, h1, h2, dy, g, j,
                                                                     q);
                if (sc.valid && 4 * g < C)
                    b.stage[(s_base + base + e) * RF + 1 + g] = make_float4(dx[0], dx[1], dx[2], dx[3]);
            }
            wave_lds_sync();
        }
    }
    // the tensorial instance runs after the explicit one and adds to its rows
    store_partials<C, kVM>(q, b.ws2 + ((int64_t)blockIdx.x * kDecWaves + wl) * NP, g, j, lane);
}

// Scatter pass.  Wave w of the workgroup owns channels 2w and 2w + 1 of the
// node gradient (C / 2 waves), so waves never touch the same LDS words and
// need no lock.  Lane l walks its own stretch of the chunk's staged samples
// (the chunk cut into 64 consecutive runs), so the 64 samples of one step
// come from distant parts of the chunk; the next step's record is loaded
// while the current one scatters.  Per step the lanes claim their floor
// cell in a per-wave owner table (one byte per cell); a lane that finds
// another lane's claim (two samples in one cell) adds with LDS atomics after
// the others.  The 8 corners go one after another: a corner of one lane can
// be another corner of a lane in an adjacent cell, and LDS runs a wave's
// operations in issue order, so each read sees the previous corners' writes.
//
// LDS layout: cell-major with C floats per cell, the channel pairs of cell
// x at (c ^ pair_swz(x)) (defined with the flush above): one wave's 64 lanes
// (same channel pair, different cells) then spread over all banks instead of
// the 2 that channels-last would give.

template <int C>
__global__ __launch_bounds__(32 * C) void project_bwd_scatter_kernel(NodeBwdArgs b)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dsm[];
    __shared__ int s_chunk;
    if (*b.mode == 0) return;
    constexpr int RF = 1 + C / 4;
    const int G = b.a.G;
    const int ncell = G * G * G;
    const int gsz = ncell * C;
    float* lgrad = reinterpret_cast<float*>(dsm);
    const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
    const int c0 = 2 * wl;  // this wave's channel pair
    unsigned char* owner = dsm + (size_t)gsz * 4 + (size_t)wl * ncell;
    for (int i = threadIdx.x; i < gsz; i += blockDim.x) lgrad[i] = 0.f;
    const int64_t total_chunks = b.chunk_off[b.exp_slots];  // explicit-grid chunks come first
    const float* stagef = reinterpret_cast<const float*>(b.stage);
    __syncthreads();
    while (true) {
        if (threadIdx.x == 0) s_chunk = atomicAdd(b.counter3, 1);
        __syncthreads();
        const int64_t c = s_chunk;
        __syncthreads();
        if (c >= total_chunks) break;
        const int4 ch = b.chunks[c];
        const int64_t gbase = (int64_t)ch.x * gsz;
        const int64_t s0 = b.soff[ch.y], s1 = b.soff[ch.z];
        const int64_t run = (s1 - s0 + 63) / 64;
        const int64_t rb = s0 + lane * run;
        const int64_t re = rb + run < s1 ? rb + run : s1;
        // three register sets, one per step modulo 3, so that the records of
        // the next two steps load while a step runs (the loop is unrolled by
        // three: no copies between sets, no wait on the later loads)
        float2 nd[3];
        float np_[3][3];
        // loads without a branch: a lane past its run reads a record of the
        // next chunk or of the stage's kStagePad padding and ignores it (it
        // is not live; its clamped cell indices stay in range), so the
        // compiler's wait counts stay exact across the unrolled steps
        auto load = [&](int u, int64_t si) {
            const float* rec = stagef + si * RF * 4;
            nd[u] = *reinterpret_cast<const float2*>(rec + 4 + c0);
            np_[u][0] = rec[0];
            np_[u][1] = rec[1];
            np_[u][2] = rec[2];
        };
        load(0, rb);
        load(1, rb + 1);
        for (int64_t st0 = 0; st0 < run; st0 += 3) {
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                const int64_t st = st0 + u;
                if (st >= run) break;
                const int64_t si = rb + st;
                const float2 dv = nd[u];
                const float px = np_[u][0], py = np_[u][1], pz = np_[u][2];
                load((u + 2) % 3, si + 2);
                const bool live = (si < re) & ((dv.x != 0.f) | (dv.y != 0.f));
                Tri t;
                tri_from_index(px, py, pz, G, G, G, t);
                const int cell = (t.z0 * G + t.y0) * G + t.x0;
                if (live) owner[cell] = (unsigned char)lane;
                wave_lds_sync();
                const bool mine = live && owner[cell] == lane;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (mine) {
                        const int cc = (int)corner_cell(t, k, G);
                        float2* gp = reinterpret_cast<float2*>(lgrad + cc * C + (c0 ^ pair_swz(cc, C)));
                        float2 v = *gp;
                        v.x = fmaf(t.w[k], dv.x, v.x);
                        v.y = fmaf(t.w[k], dv.y, v.y);
                        *gp = v;
                    }
                }
                wave_lds_sync();
                if (live && !mine) {
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int cc = (int)corner_cell(t, k, G);
                        float* gp = lgrad + cc * C + (c0 ^ pair_swz(cc, C));
                        atomicAdd(gp, t.w[k] * dv.x);
                        atomicAdd(gp + 1, t.w[k] * dv.y);
                    }
                }
                wave_lds_sync();
            }
        }
        __syncthreads();
        flush_node_grad<4, C>(b.ad, b.ad.enabled && !ch.w, ch.w != 0, lgrad, b.grad_grid + gbase, gbase, gsz >> 2);
        __syncthreads();
    }
}

// Scatter passes of the in-ROI samples of a tensorial field, in the
// explicit scatter pass's scheme: wave w owns channels 2w and 2w + 1 (LDS
// words no other wave touches) and each lane adds its own samples.  The
// three planes (or lines) of a sample are three disjoint LDS regions, so a
// step handles them together: per region a per-wave owner table (one byte
// per cell) picks one lane per floor cell, the owners add corner k of all
// three regions with plain float2 read-add-writes (corner after corner: a
// corner of one lane can be another corner of a lane in an adjacent cell),
// and the lanes that lost a claim, or whose corners are clamped together at
// the border, add with LDS atomics afterwards.  Channel pairs are swizzled
// as in the explicit pass (pair_swz of the cell index within the LDS array).
// Records are loaded two steps ahead, without a branch (the staged regions
// are padded).
//
// project_bwd_vm_plane_kernel: per in-ROI chunk, the node's three plane
// gradients (3 Gp^2 C) in LDS, w_corner * (dX L_p) into 4 corners of each
// plane; lane l walks the l-th 64th of the chunk's records; the planes are
// added to HBM once per chunk.
// project_bwd_vm_line_kernel: the three lines are shared by every node, so
// this pass is not node-major: lane l walks the l-th 64th of the whole
// in-ROI list (64 distant nodes per step: the x and z cells, which one node
// confines to a narrow window of the line, spread over the line), reading
// the interleaved line records, and each workgroup adds w_corner * (dX P_p)
// into its own LDS partials of the lines, written to its row of line
// partials at the end (rows summed in a fixed order afterwards).
template <int C, int NK>
__device__ __forceinline__ void vm_scatter3(float* lg, unsigned char* owner, const int (*cell)[NK],
                                            const float (*w)[NK], const int* f0, const int* tsz, const float2* dv,
                                            bool live, const bool* border, int c0, int lane)
{
    bool lv[3], mine[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        lv[p] = live & ((dv[p].x != 0.f) | (dv[p].y != 0.f));
        if (lv[p] && !border[p]) owner[tsz[p] + f0[p]] = (unsigned char)lane;
    }
    wave_lds_sync();
#pragma unroll
    for (int p = 0; p < 3; ++p) mine[p] = lv[p] && !border[p] && owner[tsz[p] + f0[p]] == lane;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        float2 v[3];
#pragma unroll
        for (int p = 0; p < 3; ++p)
            if (mine[p]) v[p] = *reinterpret_cast<float2*>(lg + cell[p][k] * C + (c0 ^ pair_swz(cell[p][k], C)));
#pragma unroll
        for (int p = 0; p < 3; ++p)
            if (mine[p]) {
                v[p].x = fmaf(w[p][k], dv[p].x, v[p].x);
                v[p].y = fmaf(w[p][k], dv[p].y, v[p].y);
                *reinterpret_cast<float2*>(lg + cell[p][k] * C + (c0 ^ pair_swz(cell[p][k], C))) = v[p];
            }
    }
    wave_lds_sync();
#pragma unroll
    for (int p = 0; p < 3; ++p)
        if (lv[p] && !mine[p]) {
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                float* gp = lg + cell[p][k] * C + (c0 ^ pair_swz(cell[p][k], C));
                atomicAdd(gp, w[p][k] * dv[p].x);
                atomicAdd(gp + 1, w[p][k] * dv[p].y);
            }
        }
    wave_lds_sync();
}

template <int C>
__global__ __launch_bounds__(32 * C) void project_bwd_vm_plane_kernel(NodeBwdArgs b)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dsm[];
    __shared__ int s_chunk;
    if (*b.mode == 0) return;
    constexpr int RFV = vm_record_f4<C>();
    const ProjArgs& a = b.a;
    const int Gp = a.vm.Gp;
    const int pc = Gp * Gp;
    const int pf = 3 * pc * C;
    float* lpl = reinterpret_cast<float*>(dsm);
    const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
    const int c0 = 2 * wl;
    unsigned char* owner = dsm + (size_t)pf * 4 + (size_t)wl * 3 * pc;
    for (int i = threadIdx.x; i < pf; i += blockDim.x) lpl[i] = 0.f;
    const int64_t first_chunk = b.chunk_off[b.exp_slots];
    const int64_t total_chunks = b.chunk_off[b.num_slots];
    const VMStage vs = vm_stage<C>(b);
    const float* rec = reinterpret_cast<const float*>(vs.plane);
    const int tsz[3] = {0, pc, 2 * pc};
    __syncthreads();
    while (true) {
        if (threadIdx.x == 0) s_chunk = atomicAdd(b.counter4, 1);
        __syncthreads();
        const int64_t c = first_chunk + s_chunk;
        __syncthreads();
        if (c >= total_chunks) break;
        const int4 ch = b.chunks[c];
        const int pslot = ch.x - a.vm.NG;
        const int64_t s0 = b.soff[ch.y] - vs.s_vm0, s1 = b.soff[ch.z] - vs.s_vm0;
        const int64_t run = (s1 - s0 + 63) / 64;
        const int64_t rb = s0 + lane * run;
        const int64_t re = rb + run < s1 ? rb + run : s1;
        // three register sets (the loop is unrolled by three): the records
        // of the next two steps load while a step runs; a lane past its run
        // reads a later record (of the next chunk, or the padding) and
        // ignores it
        float4 co[3];
        float2 dq[3][3];
        auto load = [&](int u, int64_t si) {
            const float* r = rec + si * RFV * 4;
            co[u] = *reinterpret_cast<const float4*>(r);
#pragma unroll
            for (int p = 0; p < 3; ++p) dq[u][p] = *reinterpret_cast<const float2*>(r + 4 + p * C + c0);
        };
        load(0, rb);
        load(1, rb + 1);
        for (int64_t st0 = 0; st0 < run; st0 += 3) {
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                const int64_t st = st0 + u;
                if (st >= run) break;
                const int64_t si = rb + st;
                const float4 A = co[u];
                float2 dv[3] = {dq[u][0], dq[u][1], dq[u][2]};
                load((u + 2) % 3, si + 2);
                const float uu[3] = {A.y, A.x, A.x}, vv[3] = {A.z, A.z, A.y};
                int cell[3][4], f0[3];
                float w[3][4];
                bool border[3];
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    Bil bb;
                    bil_setup(uu[p], vv[p], Gp, Gp, bb);
                    border[p] = bb.x0 == bb.x1 || bb.y0 == bb.y1;
                    f0[p] = bb.y0 * Gp + bb.x0;
                    cell[p][0] = p * pc + f0[p];
                    cell[p][1] = p * pc + bb.y0 * Gp + bb.x1;
                    cell[p][2] = p * pc + bb.y1 * Gp + bb.x0;
                    cell[p][3] = p * pc + bb.y1 * Gp + bb.x1;
#pragma unroll
                    for (int k = 0; k < 4; ++k) w[p][k] = bb.w[k];
                }
                vm_scatter3<C, 4>(lpl, owner, cell, w, f0, tsz, dv, si < re, border, c0, lane);
            }
        }
        __syncthreads();
        flush_node_grad<4, C>(b.ad, false, ch.w != 0, lpl, b.grad_planes + (int64_t)pslot * pf, 0, pf >> 2);
        __syncthreads();
    }
}

template <int C>
__global__ __launch_bounds__(32 * C) void project_bwd_vm_line_kernel(NodeBwdArgs b)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char dsm[];
    if (*b.mode == 0) return;  // the fused kernel writes the line partial rows
    constexpr int RFV = vm_record_f4<C>();
    const ProjArgs& a = b.a;
    const int Gl = a.vm.Gl;
    const int lf = 3 * Gl * C;
    float* lln = reinterpret_cast<float*>(dsm);
    const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
    const int c0 = 2 * wl;
    unsigned char* owner = dsm + (size_t)lf * 4 + (size_t)wl * 3 * Gl;
    for (int i = threadIdx.x; i < lf; i += blockDim.x) lln[i] = 0.f;
    const VMStage vs = vm_stage<C>(b);
    const float* rec = reinterpret_cast<const float*>(vs.line);
    const int64_t span = vs.span;
    const int64_t jb = span * blockIdx.x / gridDim.x, je = span * (blockIdx.x + 1) / gridDim.x;
    const int tsz[3] = {0, Gl, 2 * Gl};
    __syncthreads();
    float4 co[3];
    float2 dq[3][3];
    // step jj: lane l's sample l span + jj at record jj 64 + l (the records
    // past the list's end are padding, read and ignored)
    auto load = [&](int u, int64_t jj) {
        const float* r = rec + (jj * 64 + lane) * RFV * 4;
        co[u] = *reinterpret_cast<const float4*>(r);
#pragma unroll
        for (int p = 0; p < 3; ++p) dq[u][p] = *reinterpret_cast<const float2*>(r + 4 + p * C + c0);
    };
    load(0, jb);
    load(1, jb + 1);
    for (int64_t j0 = jb; j0 < je; j0 += 3) {
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int64_t jj = j0 + u;
            if (jj >= je) break;
            const float4 A = co[u];
            float2 dv[3] = {dq[u][0], dq[u][1], dq[u][2]};
            load((u + 2) % 3, jj + 2);
            const float uu[3] = {A.x, A.y, A.z};  // ROI-normalised x, local y, global z
            int cell[3][2], f0[3];
            float w[3][2];
            bool border[3];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                Lin ll;
                lin_setup(uu[p], Gl, ll);
                border[p] = ll.x0 == ll.x1;
                f0[p] = ll.x0;
                cell[p][0] = p * Gl + ll.x0;
                cell[p][1] = p * Gl + ll.x1;
                w[p][0] = ll.w0;
                w[p][1] = ll.w1;
            }
            vm_scatter3<C, 2>(lln, owner, cell, w, f0, tsz, dv, lane * span + jj < vs.n, border, c0, lane);
        }
    }
    __syncthreads();
    // this workgroup's line partials (zero if it took no sample)
    float* row = b.line_rows + (int64_t)blockIdx.x * lf;
    for (int i = threadIdx.x; i < lf / 4; i += blockDim.x) reinterpret_cast<float4*>(row)[i] = take_lds4<C>(lln, i);
}

// ---------------------------------------------------------------------------
// Node-major forward over the sorted interval records of the adjoint's
// counting sort.  Workgroups stride over the chunk table; a wave maps its
// tile samples to intervals exactly as the adjoint does, gathers (the node's
// grid stays in L2 for the whole chunk), runs the decoder, and sums
// w * softplus(y) per interval with a segmented scan over the tile's 16
// samples.  An interval belongs to one wave and its samples come in order,
// so each interval's sum is formed in a fixed order and stored once:
// ivsum[interval id].  ray_sum_kernel then adds a ray's interval sums in list
// order.  Deterministic; the ray's samples are summed per interval first, so
// the rounding differs from the ray-major forward by a few ulp.
// ---------------------------------------------------------------------------
struct NodeFwdArgs {
    ProjArgs a;
    const int32_t* sorted;
    const int4* chunks;
    const int64_t* chunk_off;
    int64_t num_slots;
    const float4* rec;
    float* ivsum;
    int* counter;  // chunk / item counter (zeroed before the launch)
    // work items of the two-pass adjoint's part table (project_fwd_part_kernel)
    const int64_t* soff;
    const int4* parts;
    const int64_t* part_off;
    int64_t max_chunks;
};

// Trilinear gather of CPL channels from a node grid staged in LDS
// (channels-last, f32 or bf16): the reads and fused multiply-add order of
// gather_channels.
template <int C, bool kBF16>
__device__ __forceinline__ void gather_channels_lds(const unsigned char* lgrid, const SampleCtx& sc, int G, int c0,
                                                    float* x)
{
    constexpr int CPL = C / 4;
#pragma unroll
    for (int q = 0; q < CPL; ++q) x[q] = 0.f;
    if (!sc.valid) return;
    float v[8][CPL];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int off = (int)corner_cell(sc.tri, k, G) * C + c0;
        if constexpr (!kBF16) {
            const float* p = reinterpret_cast<const float*>(lgrid) + off;
            if constexpr (CPL == 4) {
                const float4 f = *reinterpret_cast<const float4*>(p);
                v[k][0] = f.x;
                v[k][1] = f.y;
                v[k][2] = f.z;
                v[k][3] = f.w;
            } else {
                const float2 f = *reinterpret_cast<const float2*>(p);
                v[k][0] = f.x;
                v[k][1] = f.y;
            }
        } else {
            const uint16_t* p = reinterpret_cast<const uint16_t*>(lgrid) + off;
            if constexpr (CPL == 4) {
                const uint2 u = *reinterpret_cast<const uint2*>(p);
                v[k][0] = __uint_as_float(u.x << 16);
                v[k][1] = __uint_as_float(u.x & 0xffff0000u);
                v[k][2] = __uint_as_float(u.y << 16);
                v[k][3] = __uint_as_float(u.y & 0xffff0000u);
            } else {
                const uint32_t u = *reinterpret_cast<const uint32_t*>(p);
                v[k][0] = __uint_as_float(u << 16);
                v[k][1] = __uint_as_float(u & 0xffff0000u);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc = fmaf(v[k][q], sc.tri.w[k], acc);
        x[q] = acc;
    }
}

constexpr int kFwdLdsWaves = 16;

// kLDS: a workgroup of kNW waves takes chunks from a global counter, copies
// the chunk's node grid (G^3 C f32 or bf16) into LDS once and gathers from
// there; otherwise workgroups of kWaves waves stride over the chunk table
// and gather from HBM (the grid stays in L2 while its chunk is integrated).
template <int C, bool kBF16, bool kDecBF16, bool kLDS = false, int kNW = kWaves>
__global__ __launch_bounds__(64 * kNW, kLDS ? 1 : 4) void project_fwd_node_kernel(NodeFwdArgs b)
{
    using SM = DecoderSmem<C, kDecBF16>;
    constexpr int CPL = C / 4;
    __shared__ __attribute__((aligned(16))) float s[SM::kF32];
    __shared__ __attribute__((aligned(16))) __bf16 sb[SM::kBF16];
    extern __shared__ __attribute__((aligned(16))) unsigned char lgrid[];
    __shared__ int s_chunk;
    const ProjArgs& a = b.a;
    const float4* __restrict__ rec = b.rec;
    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds<C>(a.mlp, s);
    __syncthreads();
    const int G = a.G;
    const int gsz = G * G * G * C;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int wl = threadIdx.x >> 6;
    const int64_t total_chunks = b.chunk_off[b.num_slots];
    for (int64_t c = blockIdx.x;; c += gridDim.x) {
        if constexpr (kLDS) {
            if (threadIdx.x == 0) s_chunk = atomicAdd(b.counter, 1);
            __syncthreads();  // also: every wave is done with the previous grid
            c = s_chunk;
        }
        if (c >= total_chunks) break;
        const int4 ch = b.chunks[c];
        const int64_t gbase = (int64_t)ch.x * gsz;
        if constexpr (kLDS) {
            // the node grid, 16 bytes per thread and step
            const int n16 = gsz * (kBF16 ? 2 : 4) / 16;
            const uint4* src = reinterpret_cast<const uint4*>(reinterpret_cast<const unsigned char*>(a.grid) +
                                                              gbase * (kBF16 ? 2 : 4));
            uint4* dst = reinterpret_cast<uint4*>(lgrid);
            for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
            __syncthreads();
        }
        const PNode nd = a.nodes[a.iv_node[b.sorted[ch.y]]];
        const int n_iv = ch.z - ch.y;
        const int per = (n_iv + kNW - 1) / kNW;
        const int ib = ch.y + wl * per;
        const int ie = min(ib + per, ch.z);
        int tot = 0;
        for (int k = ib + lane; k < ie; k += 64) tot += __float_as_int(rec[3 * (int64_t)k].z);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
        const int ntiles = (tot + kTile - 1) / kTile;
        int cur = ib, cum = 0;
        const int grp = lane & 48;
        // the wave's open interval (continuing into the next tile) and its sum
        int carry_id = -1;
        float carry = 0.f;
        for (int ti = 0; ti < ntiles; ++ti) {
            const int base = ti * kTile;
            const int t = base + j;
            const int cand = cur + j;
            const int c_cnt = cand < ie ? __float_as_int(rec[3 * (int64_t)cand].z) : 0;
            const bool ones = __all(cand >= ie || c_cnt == 1);
            int c_end = c_cnt;
            int pos, posn, s_cnt, s_end, last_end = 0;
            if (ones) {
                pos = j;
                posn = min(kTile, ie - cur);
                s_cnt = 1;
                s_end = cum + j + 1;
            } else {
                c_end = scan16_inclusive(c_end);
                c_end = cand < ie ? c_end + cum : INT_MAX;
                pos = group_search(c_end, t, grp);
                posn = group_search(c_end, base + kTile, grp);
                const int src = grp | (pos < kTile ? pos : kTile - 1);
                s_cnt = __shfl(c_cnt, src, 64);
                s_end = __shfl(c_end, src, 64);
                last_end = __shfl(c_end, grp | (kTile - 1), 64);
            }
            SampleCtx sc;
            sc.valid = t < tot;
            sc.w = 0.f;
            sc.gbase = -1;
            int iv_id = -1;
            if (sc.valid) {
                int k = cur + pos, n_s = s_cnt, start = s_end - s_cnt;
                if (pos >= kTile) {
                    k = cur + kTile;
                    start = last_end;
                    n_s = __float_as_int(rec[3 * (int64_t)k].z);
                    while (start + n_s <= t) {
                        start += n_s;
                        ++k;
                        n_s = __float_as_int(rec[3 * (int64_t)k].z);
                    }
                }
                const float4 A = rec[3 * (int64_t)k];
                const float4 B = rec[3 * (int64_t)k + 1];
                const float4 D = rec[3 * (int64_t)k + 2];
                const float o[3] = {B.x, B.y, B.z};
                const float d[3] = {B.w, D.x, D.y};
                const int jj = t - start;
                make_sample_node(a, nd, gbase, __float_as_int(D.w), o, d, A.x, A.y, n_s, jj,
                                 __float_as_int(A.w) + jj, sc);
                iv_id = __float_as_int(D.z);
            }
            if (ones)
                cum += posn;
            else if (posn > 0)
                cum = __shfl(c_end, grp | (posn - 1), 64);
            cur += posn;
            float x[CPL];
            if constexpr (kLDS)
                gather_channels_lds<C, kBF16>(lgrid, sc, G, CPL * g, x);
            else
                gather_channels<C, kBF16>(a, sc, CPL * g, x);
            v4f h1[4], h2[4];
            float y;
            int opq = 0;
            asm volatile("" : "+v"(opq));  // no hoisting of the weight reads
            const float* so = s + opq;
            const __bf16* sbo = sb + opq;
            if constexpr (kDecBF16)
                y = decoder_fwd_bf16<C>(sbo, so, x, g, j, h1, h2);
            else
                y = decoder_fwd<C>(so, x, g, j, h1, h2);
            // segmented inclusive scan of w * density over the tile (lanes 0..15)
            float v = sc.valid ? softplus_f(y, a.beta) * sc.w : 0.f;
#pragma unroll
            for (int off = 1; off < kTile; off <<= 1) {
                const float ov = __shfl_up(v, off, kTile);
                const int oid = __shfl_up(iv_id, off, kTile);
                if (j >= off && oid == iv_id) v += ov;
            }
            const int next_id = __shfl_down(iv_id, 1, kTile);
            const bool seg_end = sc.valid && (j == kTile - 1 || next_id != iv_id);
            // the tile's first interval may continue the previous tile's last one
            const int first_id = __shfl(iv_id, 0, 64);
            if (carry_id >= 0 && carry_id != first_id && lane == 0) b.ivsum[carry_id] = carry;
            if (seg_end && iv_id == carry_id) v += carry;
            // the last valid segment stays open unless this is the wave's last tile
            const int last_valid = min(kTile, tot - base) - 1;
            const int last_id = __shfl(iv_id, last_valid, 64);
            const float last_v = __shfl(v, last_valid, 64);
            const bool final_tile = ti == ntiles - 1;
            if (g == 0 && seg_end && (iv_id != last_id || final_tile)) b.ivsum[iv_id] = v;
            carry_id = final_tile ? -1 : last_id;
            carry = last_v;
        }
    }
}

// Node-major forward over the part table of the adjoint's sort (a workspace
// with the staging region): waves take work items of ~kPartSamples samples
// (a run of one chunk's intervals, one node) from a counter and run them in
// steps of 64 samples.  (A) lane l maps and places sample l of the step once
// (map_tile64) and writes its packed corners, trilinear weights, integration
// weight and interval id to the wave's sample table in LDS; (B) the step's
// four tiles of 16 read the table, gather their lane's channels, run the
// decoder and sum w * density per interval with the segmented scan of
// project_fwd_node_kernel (an item holds whole intervals, so no sum crosses
// items).  ray_sum_kernel adds a ray's interval sums in list order.
constexpr int kFwdPartWaves = 4;

template <int C, bool kBF16, bool kDecBF16, bool kVM = false>
__global__ __launch_bounds__(64 * kFwdPartWaves, 4) void project_fwd_part_kernel(NodeFwdArgs b)
{
    using SM = DecoderSmem<C, kDecBF16>;
    constexpr int CPL = C / 4;
    __shared__ __attribute__((aligned(16))) float s[SM::kF32];
    __shared__ __attribute__((aligned(16))) __bf16 sb[SM::kBF16];
    __shared__ float4 s_pl[kFwdPartWaves][64 * 3];
    const ProjArgs& a = b.a;
    const float4* __restrict__ rec = b.rec;
    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds<C>(a.mlp, s);
    __syncthreads();
    const int G = a.G;
    const int gsz = G * G * G * C;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int wl = threadIdx.x >> 6;
    float4* pl = s_pl[wl];
    const int64_t n_items = b.part_off[b.max_chunks];
    while (true) {
        int item = 0;
        if (lane == 0) item = atomicAdd(b.counter, 1);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= n_items) break;
        const int4 it = b.parts[item];
        const PNode nd = a.nodes[a.iv_node[b.sorted[it.y]]];
        // a tensorial field's in-ROI items follow the explicit-grid ones
        // (slot = explicit grids + plane slot)
        const int pslot = kVM ? it.x - a.vm.NG : -1;
        const bool vmc = kVM && pslot >= 0;
        const int64_t gbase = vmc ? 0 : (int64_t)it.x * gsz;
        const int ib = it.y, ie = it.z;
        const int tot = (int)(b.soff[ie] - b.soff[ib]);
        TileCursor tc;
        tc.cur = ib;
        tc.cum = 0;
        tc.cnt = 0;
        // the wave's open interval (continuing into the next tile) and its sum
        int carry_id = -1;
        float carry = 0.f;
        const int nsteps = (tot + 63) / 64;
        for (int st = 0; st < nsteps; ++st) {
            const int base = st * 64;
            {
                TileMap m;
                m.valid = false;
                map_tile64(rec, ie, base, tot, lane, tc, m);
                int packed = -1, iv_id = -1;
                SampleCtx sc;
                sc.valid = m.valid;
                sc.w = 0.f;
                if (kVM && vmc) {
                    // an in-ROI sample keeps its local and global coordinates
                    sc.loc[0] = sc.loc[1] = sc.loc[2] = sc.gx = sc.gz = 0.f;
                    if (m.valid) {
                        const float o[3] = {m.B.x, m.B.y, m.B.z};
                        const float d[3] = {m.B.w, m.D.x, m.D.y};
                        make_sample_node<true>(a, nd, 0, __float_as_int(m.D.w), o, d, m.A.x, m.A.y, m.n_s, m.jj,
                                               __float_as_int(m.A.w) + m.jj, sc, false);
                        iv_id = __float_as_int(m.D.z);
                    }
                    pl[lane * 3] = make_float4(sc.loc[0], sc.loc[1], sc.loc[2], sc.gx);
                    pl[lane * 3 + 1] = make_float4(sc.gz, m.valid ? 1.f : 0.f, 0.f, 0.f);
                    pl[lane * 3 + 2] = make_float4(0.f, sc.w, __int_as_float(iv_id), 0.f);
                } else {
                    if (m.valid) {
                        const float o[3] = {m.B.x, m.B.y, m.B.z};
                        const float d[3] = {m.B.w, m.D.x, m.D.y};
                        make_sample_node(a, nd, gbase, __float_as_int(m.D.w), o, d, m.A.x, m.A.y, m.n_s, m.jj,
                                         __float_as_int(m.A.w) + m.jj, sc);
                        packed = sc.tri.x0 | (sc.tri.x1 << 5) | (sc.tri.y0 << 10) | (sc.tri.y1 << 15) |
                                 (sc.tri.z0 << 20) | (sc.tri.z1 << 25);
                        iv_id = __float_as_int(m.D.z);
                    }
                    pl[lane * 3] = make_float4(__int_as_float(packed), sc.tri.w[0], sc.tri.w[1], sc.tri.w[2]);
                    pl[lane * 3 + 1] = make_float4(sc.tri.w[3], sc.tri.w[4], sc.tri.w[5], sc.tri.w[6]);
                    pl[lane * 3 + 2] = make_float4(sc.tri.w[7], sc.w, __int_as_float(iv_id), 0.f);
                }
            }
            wave_lds_sync();
            const int ntile = min(4, (tot - base + kTile - 1) / kTile);
            for (int q4 = 0; q4 < ntile; ++q4) {
                const int tb = base + q4 * kTile;
                const int e = q4 * kTile + j;
                const float4 p0 = pl[e * 3], p1 = pl[e * 3 + 1], p2 = pl[e * 3 + 2];
                const int packed = __float_as_int(p0.x);
                SampleCtx sc;
                sc.valid = (kVM && vmc) ? p1.y != 0.f : packed >= 0;
                sc.tri.x0 = packed & 31;
                sc.tri.x1 = (packed >> 5) & 31;
                sc.tri.y0 = (packed >> 10) & 31;
                sc.tri.y1 = (packed >> 15) & 31;
                sc.tri.z0 = (packed >> 20) & 31;
                sc.tri.z1 = (packed >> 25) & 31;
                sc.tri.w[0] = p0.y;
                sc.tri.w[1] = p0.z;
                sc.tri.w[2] = p0.w;
                sc.tri.w[3] = p1.x;
                sc.tri.w[4] = p1.y;
                sc.tri.w[5] = p1.z;
                sc.tri.w[6] = p1.w;
                sc.tri.w[7] = p2.x;
                sc.gbase = gbase;
                const int iv_id = __float_as_int(p2.z);
                float x[CPL];
                if (kVM && vmc) {
#pragma unroll
                    for (int qq = 0; qq < CPL; ++qq) x[qq] = 0.f;
                    if (sc.valid) {
                        sc.loc[0] = p0.x;
                        sc.loc[1] = p0.y;
                        sc.loc[2] = p0.z;
                        sc.gx = p0.w;
                        sc.gz = p1.x;
                        // F = Pyz Lx + Pxz Ly + Pxy Lz, one pair at a time
#pragma unroll
                        for (int pp = 0; pp < 3; ++pp) {
                            float Pv[CPL], Lv[CPL];
                            vm_pair_read<C, CPL>(a, pslot, sc, pp, CPL * g, Pv, Lv);
#pragma unroll
                            for (int qq = 0; qq < CPL; ++qq)
                                x[qq] = pp == 0 ? Pv[qq] * Lv[qq] : x[qq] + Pv[qq] * Lv[qq];
                        }
                    }
                } else if (sc.valid) {
                    Corners<C, kBF16> cr;
                    gather_corners<C, kBF16>(a, sc, g, cr);
                    corners_combine<C, kBF16>(sc, cr, x);
                } else {
#pragma unroll
                    for (int qq = 0; qq < CPL; ++qq) x[qq] = 0.f;
                }
                v4f h1[4], h2[4];
                float y;
                int opq = 0;
                asm volatile("" : "+v"(opq));  // no hoisting of the weight reads
                const float* so = s + opq;
                const __bf16* sbo = sb + opq;
                if constexpr (kDecBF16)
                    y = decoder_fwd_bf16<C>(sbo, so, x, g, j, h1, h2);
                else
                    y = decoder_fwd<C>(so, x, g, j, h1, h2);
                // segmented inclusive scan of w * density over the tile (lanes 0..15)
                float v = sc.valid ? softplus_f(y, a.beta) * p2.y : 0.f;
#pragma unroll
                for (int off = 1; off < kTile; off <<= 1) {
                    const float ov = __shfl_up(v, off, kTile);
                    const int oid = __shfl_up(iv_id, off, kTile);
                    if (j >= off && oid == iv_id) v += ov;
                }
                const int next_id = __shfl_down(iv_id, 1, kTile);
                const bool seg_end = sc.valid && (j == kTile - 1 || next_id != iv_id);
                // the tile's first interval may continue the previous tile's last one
                const int first_id = __shfl(iv_id, 0, 64);
                if (carry_id >= 0 && carry_id != first_id && lane == 0) b.ivsum[carry_id] = carry;
                if (seg_end && iv_id == carry_id) v += carry;
                // the last valid segment stays open unless this is the item's last tile
                const int last_valid = min(kTile, tot - tb) - 1;
                const int last_id = __shfl(iv_id, last_valid, 64);
                const float last_v = __shfl(v, last_valid, 64);
                const bool final_tile = tb + kTile >= tot;
                if (g == 0 && seg_end && (iv_id != last_id || final_tile)) b.ivsum[iv_id] = v;
                carry_id = final_tile ? -1 : last_id;
                carry = last_v;
            }
            wave_lds_sync();
        }
    }
}

// A wave's 64 rays own one contiguous range of interval sums: windows of it
// are loaded coalesced into LDS, and each lane adds its own ray's sums from
// there in list order (the order of a per-ray loop over global memory, so the
// result is the same bit for bit; intervals without samples hold no sum and
// add nothing: the sums are non-negative, so adding +0 changes none).
constexpr int kRaySumWin = 512;

__global__ __launch_bounds__(256) void ray_sum_kernel(const int64_t* __restrict__ iv_off,
                                                      const int32_t* __restrict__ count,
                                                      const float* __restrict__ ivsum, int64_t R,
                                                      float* __restrict__ P)
{
    __shared__ float s_v[4][kRaySumWin];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t r0 = r - lane;
    if (r0 >= R) return;  // whole waves
    const bool valid = r < R;
    const int64_t hb = valid ? iv_off[r] : 0, he = valid ? iv_off[r + 1] : 0;
    const int64_t wb = iv_off[r0], we = iv_off[r0 + 64 < R ? r0 + 64 : R];
    float acc = 0.f;
    for (int64_t x0 = wb; x0 < we; x0 += kRaySumWin) {
        const int64_t x1 = x0 + kRaySumWin < we ? x0 + kRaySumWin : we;
        for (int64_t h = x0 + lane; h < x1; h += 64) s_v[w][h - x0] = count[h] > 0 ? ivsum[h] : 0.f;
        wave_lds_sync();
        const int64_t a = hb > x0 ? hb : x0, b = he < x1 ? he : x1;
        for (int64_t h = a; h < b; ++h) acc += s_v[w][h - x0];
        wave_lds_sync();
    }
    if (valid) P[r] = acc;
}

// ---- counting sort of intervals by grid slot, chunk table ----
__global__ __launch_bounds__(256) void fill_i32_kernel(int32_t* p, int64_t n, int32_t v)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Sort key of a node: its explicit grid slot, else NG + its plane slot
// (tensorial field), else -1 (a node without a field: its samples drop out).
struct SlotMap {
    const int32_t* n2g;
    const int32_t* n2p;  // null without a tensorial field
    int64_t NG;
};

__device__ __forceinline__ int slot_of(const SlotMap& m, int node)
{
    int s = m.n2g[node];
    if (s < 0 && m.n2p) {
        const int p = m.n2p[node];
        s = p >= 0 ? (int)(m.NG + p) : -1;
    }
    return s;
}

__global__ __launch_bounds__(256) void slot_hist_kernel(const int32_t* __restrict__ iv_node, SlotMap sm, int64_t H,
                                                        int32_t* __restrict__ cnt)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= H) return;
    int slot = slot_of(sm, iv_node[i]);
    if (slot >= 0) atomicAdd(&cnt[slot], 1);
}

__global__ __launch_bounds__(256) void slot_scatter_kernel(const int32_t* __restrict__ iv_node, SlotMap sm, int64_t H,
                                                           const int64_t* __restrict__ slot_off,
                                                           int32_t* __restrict__ cursor, int32_t* __restrict__ sorted,
                                                           RecordSrc rs)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= H) return;
    int slot = slot_of(sm, iv_node[i]);
    if (slot < 0) return;
    int pos = atomicAdd(&cursor[slot], 1);
    const int64_t k = slot_off[slot] + pos;
    sorted[k] = (int32_t)i;
    write_record(rs, i, k);
}

// Block-aggregated counting sort for lists with many intervals per slot
// (learned one-sample intervals put ~1000 on each node): each workgroup
// counts kSortItems consecutive intervals in an LDS histogram, so the global
// counters see one atomic per distinct slot per workgroup instead of one per
// interval.  The scatter pass reserves the workgroup's range of every slot
// with one atomic and ranks its intervals with LDS atomics.
constexpr int kSortItems = 8192;

__device__ __forceinline__ void block_slot_hist(const int32_t* __restrict__ iv_node, const SlotMap& sm, int64_t b0,
                                                int64_t e0, int64_t NA, int32_t* lh)
{
    for (int64_t s = threadIdx.x; s < NA; s += blockDim.x) lh[s] = 0;
    __syncthreads();
    for (int64_t i = b0 + threadIdx.x; i < e0; i += blockDim.x) {
        const int slot = slot_of(sm, iv_node[i]);
        if (slot >= 0) atomicAdd(&lh[slot], 1);
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void slot_hist_lds_kernel(const int32_t* __restrict__ iv_node, SlotMap sm,
                                                            int64_t H, int64_t NA, int32_t* __restrict__ cnt)
{
    extern __shared__ int32_t lh[];
    const int64_t b0 = blockIdx.x * (int64_t)kSortItems;
    const int64_t e0 = b0 + kSortItems < H ? b0 + kSortItems : H;
    block_slot_hist(iv_node, sm, b0, e0, NA, lh);
    for (int64_t s = threadIdx.x; s < NA; s += blockDim.x) {
        const int v = lh[s];
        if (v) atomicAdd(&cnt[s], v);
    }
}

__global__ __launch_bounds__(256) void slot_scatter_lds_kernel(const int32_t* __restrict__ iv_node, SlotMap sm,
                                                               int64_t H, int64_t NA, const int64_t* __restrict__ slot_off,
                                                               int32_t* __restrict__ cursor,
                                                               int32_t* __restrict__ sorted, RecordSrc rs)
{
    extern __shared__ int32_t lh[];
    const int64_t b0 = blockIdx.x * (int64_t)kSortItems;
    const int64_t e0 = b0 + kSortItems < H ? b0 + kSortItems : H;
    block_slot_hist(iv_node, sm, b0, e0, NA, lh);
    for (int64_t s = threadIdx.x; s < NA; s += blockDim.x) {
        const int v = lh[s];
        lh[s] = v ? (int)(slot_off[s] + atomicAdd(&cursor[s], v)) : 0;
    }
    __syncthreads();
    for (int64_t i = b0 + threadIdx.x; i < e0; i += blockDim.x) {
        const int slot = slot_of(sm, iv_node[i]);
        if (slot >= 0) {
            const int64_t k = atomicAdd(&lh[slot], 1);
            sorted[k] = (int32_t)i;
            write_record(rs, i, k);
        }
    }
}

// Total samples of the interval list (for the chunk size).
__global__ __launch_bounds__(256) void sum_counts_kernel(const int32_t* __restrict__ count, int64_t H,
                                                         unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long part[4];
    unsigned long long v = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < H; i += (int64_t)gridDim.x * blockDim.x)
        v += (unsigned long long)count[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(total, part[0] + part[1] + part[2] + part[3]);
}

__global__ __launch_bounds__(256) void chunk_count_kernel(const int64_t* __restrict__ slot_off, int64_t NA,
                                                          int64_t H, const unsigned long long* __restrict__ total,
                                                          int32_t* __restrict__ chunk_cnt)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= NA) return;
    const int64_t ch = chunk_intervals(H, (int64_t)*total);
    int64_t n = slot_off[s + 1] - slot_off[s];
    chunk_cnt[s] = (int32_t)((n + ch - 1) / ch);
}

__global__ __launch_bounds__(256) void chunk_desc_kernel(const int64_t* __restrict__ slot_off,
                                                         const int64_t* __restrict__ chunk_off, int64_t NA,
                                                         int64_t H, const unsigned long long* __restrict__ total,
                                                         int4* __restrict__ desc)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= NA) return;
    const int64_t ch = chunk_intervals(H, (int64_t)*total);
    int64_t b0 = slot_off[s], e0 = slot_off[s + 1];
    int64_t c0 = chunk_off[s], nch = chunk_off[s + 1] - c0;
    for (int64_t k = 0; k < nch; ++k) {
        int64_t b = b0 + k * ch;
        int64_t e = b + ch < e0 ? b + ch : e0;
        desc[c0 + k] = make_int4((int)s, (int)b, (int)e, nch > 1 ? 1 : 0);
    }
}

// Decoder-pass work items: each chunk's intervals cut into parts_of() runs.
__global__ __launch_bounds__(256) void part_count_kernel(const int4* __restrict__ desc,
                                                         const int64_t* __restrict__ chunk_off, int64_t NA,
                                                         int64_t max_chunks, const int64_t* __restrict__ soff,
                                                         int32_t* __restrict__ part_cnt)
{
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= max_chunks) return;
    if (c >= chunk_off[NA]) {
        part_cnt[c] = 0;
        return;
    }
    const int4 ch = desc[c];
    part_cnt[c] = parts_of(ch.z - ch.y, soff[ch.z] - soff[ch.y]);
}

__global__ __launch_bounds__(256) void part_desc_kernel(const int4* __restrict__ desc,
                                                        const int64_t* __restrict__ chunk_off, int64_t NA,
                                                        const int64_t* __restrict__ soff,
                                                        const int64_t* __restrict__ part_off, int4* __restrict__ parts)
{
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= chunk_off[NA]) return;
    const int4 ch = desc[c];
    const int n_iv = ch.z - ch.y;
    const int np = parts_of(n_iv, soff[ch.z] - soff[ch.y]);
    if (np == 0) return;
    const int per = (n_iv + np - 1) / np;
    const int64_t o = part_off[c];
    for (int p = 0; p < np; ++p) {
        const int bb = ch.y + p * per;
        const int ee = bb + per < ch.z ? bb + per : ch.z;
        parts[o + p] = make_int4(ch.x, bb, ee, (int)c);
    }
}

// mode = 1 (two-pass kernels) when the sort prepared the staging tables and
// every sample's staged record fits in the staging capacity, else 0.  With a
// tensorial field (vm_rec_bytes > 0) the in-ROI samples, sorted after the
// explicit ones from soff[slot_off[exp_slots]], take records of their own
// after the explicit region and its padding.
__global__ void stage_mode_kernel(const unsigned long long* __restrict__ stotal, const int* __restrict__ prepared,
                                  unsigned long long capacity, int rec_bytes, const int64_t* __restrict__ soff,
                                  const int64_t* __restrict__ slot_off, int64_t exp_slots, int vm_rec_bytes,
                                  int* __restrict__ mode)
{
    if (threadIdx.x != 0) return;
    if (*prepared != 1) {
        *mode = 0;
        return;
    }
    const unsigned long long total = *stotal;
    unsigned long long need = (total + kStagePad) * (unsigned long long)rec_bytes;
    if (vm_rec_bytes > 0) {
        // explicit records, then the in-ROI plane records and the line
        // records (64 span of them), each region padded (vm_stage)
        const unsigned long long s_exp = (unsigned long long)soff[slot_off[exp_slots]];
        const unsigned long long n = total - s_exp, span = (n + 63) / 64;
        need = (s_exp + kStagePad) * (unsigned long long)rec_bytes +
               (n + kStagePad + 64 * span + 2 * kStagePad) * (unsigned long long)vm_rec_bytes;
    }
    *mode = need <= capacity ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Per-node density statistics: one workgroup per active node, tiles of 16
// points per wave, FP64 sum / sum of squares and FP32 max reduced in a fixed
// order (lane, then wave).
// ---------------------------------------------------------------------------
struct StatsArgs {
    ProjArgs a;
    const int64_t* active_ids;
    int mode, n, rounds;
    uint64_t seed;
    float* node_max;
    double* node_sum;
    double* node_sumsq;
    const float* point_mask;  // nullable [num_active, points per node]
};

// Point p of the n^3 lattice of a node (UniformGlobalSamples,
// structure/HyperTree.cu:1618-1653): (idx / (n - 1)) * size + min, x fastest.
__device__ __forceinline__ void lattice_position(const PNode& nd, const float* size, int64_t p, int n, float* pos)
{
    const int64_t idx[3] = {p % n, (p / n) % n, p / ((int64_t)n * n)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lp = (float)idx[k] / (float)(n - 1);
        pos[k] = lp * size[k] + nd.mn[k];
    }
}

__global__ __launch_bounds__(256) void lattice_points_kernel(const PNode* __restrict__ nodes,
                                                             const int64_t* __restrict__ active_ids, int n,
                                                             float* __restrict__ out)
{
    const PNode nd = nodes[active_ids[blockIdx.x]];
    float size[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) size[k] = nd.mx[k] - nd.mn[k];
    const int64_t per = (int64_t)n * n * n;
    float* o = out + (int64_t)blockIdx.x * per * 3;
    for (int64_t p = threadIdx.x; p < per; p += blockDim.x) {
        float pos[3];
        lattice_position(nd, size, p, n, pos);
#pragma unroll
        for (int k = 0; k < 3; ++k) o[p * 3 + k] = pos[k];
    }
}

template <int C, bool kBF16, bool kDecBF16>
__global__ __launch_bounds__(64 * kWaves) void node_stats_kernel(StatsArgs sa)
{
    using SM = DecoderSmem<C, kDecBF16>;
    constexpr int CPL = C / 4;
    __shared__ __attribute__((aligned(16))) float s[SM::kF32];
    __shared__ __attribute__((aligned(16))) __bf16 sb[SM::kBF16];
    __shared__ double s_sum[kWaves], s_sq[kWaves];
    __shared__ float s_max[kWaves];
    const ProjArgs& a = sa.a;
    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds<C>(a.mlp, s);
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int wl = threadIdx.x >> 6;
    const int64_t act = blockIdx.x;
    const int node = (int)sa.active_ids[act];
    const PNode nd = a.nodes[node];
    float size[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) size[k] = nd.mx[k] - nd.mn[k];
    const int n = sa.n;
    const int64_t per = (int64_t)n * n * n * (sa.mode == 1 ? sa.rounds : 1);
    const int slot = a.node_to_grid[node];

    float dmax = -INFINITY;
    double s1 = 0.0, s2 = 0.0;
    for (int64_t base = (int64_t)wl * kTile; base < per; base += (int64_t)kWaves * kTile) {
        const int64_t p = base + j;
        SampleCtx sc;
        sc.valid = p < per;
        sc.w = 1.f;
        sc.gbase = -1;
        if (sc.valid) {
            float pos[3];
            if (sa.mode == 0) {
                lattice_position(nd, size, p, n, pos);
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) pos[k] = nd.mn[k] + point_uniform(sa.seed, node, p, k) * size[k];
            }
            float loc[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) loc[k] = ((pos[k] - nd.mn[k]) / size[k]) * 2 - 1;
            tri_setup(loc[0], loc[1], loc[2], a.G, a.G, a.G, sc.tri);
            if (slot >= 0) sc.gbase = (int64_t)slot * a.G * a.G * a.G * C;
        }
        float x[CPL];
        gather_channels<C, kBF16>(a, sc, CPL * g, x);
        v4f h1[4], h2[4];
        float y;
        if constexpr (kDecBF16)
            y = decoder_fwd_bf16<C>(sb, s, x, g, j, h1, h2);
        else
            y = decoder_fwd<C>(s, x, g, j, h1, h2);
        if (sc.valid && g == 0) {
            float dn = softplus_f(y, a.beta);
            if (sa.point_mask) dn = dn * sa.point_mask[act * per + p];
            dmax = fmaxf(dmax, dn);
            s1 += (double)dn;
            s2 += (double)dn * (double)dn;
        }
    }
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
        s1 += __shfl_xor(s1, off, 64);
        s2 += __shfl_xor(s2, off, 64);
        dmax = fmaxf(dmax, __shfl_xor(dmax, off, 64));
    }
    if (lane == 0) {
        s_sum[wl] = s1;
        s_sq[wl] = s2;
        s_max[wl] = dmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = 0.0, t2 = 0.0;
        float m = -INFINITY;
        for (int w = 0; w < kWaves; ++w) {
            t1 += s_sum[w];
            t2 += s_sq[w];
            m = fmaxf(m, s_max[w]);
        }
        sa.node_max[act] = m;
        sa.node_sum[act] = t1;
        sa.node_sumsq[act] = t2;
    }
}


// ---------------------------------------------------------------------------
// Field on a regular lattice (volume export): a wave takes tiles of 16
// consecutive lattice points, locates each one in the active set, gathers
// and decodes as the projection does, and writes density * mask.
// ---------------------------------------------------------------------------
struct VolumeArgs {
    ProjArgs a;
    const int64_t* active_ids;
    int mode, swap;
    int size[3];
    float roi_min[3], step[3];
    int64_t total;
    float* density;
    int64_t* node_id;
    float* valid;
    float* position;
};

// Lattice point p: index i (axis 0 fastest, or axis 2 with swap) and position
// (mode 0: i / (size - 1) * 2 - 1; mode 1: roi_min + step * i).
__device__ __forceinline__ void volume_point(const VolumeArgs& v, int64_t p, float* pos)
{
    int c[3];
    int64_t tmp = p;
    if (v.swap) {
        for (int d = 2; d >= 0; --d) {
            c[d] = (int)(tmp % v.size[d]);
            tmp /= v.size[d];
        }
    } else {
        for (int d = 0; d < 3; ++d) {
            c[d] = (int)(tmp % v.size[d]);
            tmp /= v.size[d];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (v.mode == 0) {
            const float q = (float)c[k] / ((float)v.size[k] - 1.0f);
            pos[k] = q * 2 - 1.0f;
        } else {
            pos[k] = v.roi_min[k] + v.step[k] * (float)c[k];
        }
    }
}

__global__ __launch_bounds__(256) void volume_points_kernel(VolumeArgs v)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= v.total) return;
    float pos[3];
    volume_point(v, p, pos);
    const int32_t nd = active_node_for_position(v.a.nodes, pos);
    if (v.position) {
#pragma unroll
        for (int k = 0; k < 3; ++k) v.position[p * 3 + k] = pos[k];
    }
    if (v.node_id) v.node_id[p] = nd < 0 ? v.active_ids[0] : (int64_t)nd;
    if (v.valid) v.valid[p] = nd < 0 ? 0.f : 1.f;
}

template <int C, bool kBF16, bool kDecBF16>
__global__ __launch_bounds__(64 * kWaves, 4) void volume_kernel(VolumeArgs v)
{
    using SM = DecoderSmem<C, kDecBF16>;
    constexpr int CPL = C / 4;
    __shared__ __attribute__((aligned(16))) float s[SM::kF32];
    __shared__ __attribute__((aligned(16))) __bf16 sb[SM::kBF16];
    const ProjArgs& a = v.a;
    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds<C>(a.mlp, s);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int64_t wave = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * kWaves;
    for (int64_t base = wave * kTile; base < v.total; base += nwaves * kTile) {
        const int64_t p = base + j;
        SampleCtx sc;
        sc.valid = false;
        sc.w = 1.f;
        sc.gbase = -1;
        int32_t nd = -1;
        float pos[3] = {0.f, 0.f, 0.f};
        if (p < v.total) {
            volume_point(v, p, pos);
            nd = active_node_for_position(a.nodes, pos);
            if (nd >= 0) {
                const PNode box = a.nodes[nd];
                float loc[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) loc[k] = ((pos[k] - box.mn[k]) / (box.mx[k] - box.mn[k])) * 2 - 1;
                tri_setup(loc[0], loc[1], loc[2], a.G, a.G, a.G, sc.tri);
                const int slot = a.node_to_grid[nd];
                if (slot >= 0) {
                    sc.gbase = (int64_t)slot * a.G * a.G * a.G * C;
                    sc.valid = true;
                }
            }
        }
        float x[CPL];
        gather_channels<C, kBF16>(a, sc, CPL * g, x);
        v4f h1[4], h2[4];
        float y;
        int opq = 0;
        asm volatile("" : "+v"(opq));  // no hoisting of the weight reads
        if constexpr (kDecBF16)
            y = decoder_fwd_bf16<C>(sb + opq, s + opq, x, g, j, h1, h2);
        else
            y = decoder_fwd<C>(s + opq, x, g, j, h1, h2);
        if (g == 0 && p < v.total) {
            v.density[p] = sc.valid ? softplus_f(y, a.beta) : 0.f;
            if (v.node_id) v.node_id[p] = nd < 0 ? v.active_ids[0] : (int64_t)nd;
            if (v.valid) v.valid[p] = nd < 0 ? 0.f : 1.f;
            if (v.position) {
#pragma unroll
                for (int k = 0; k < 3; ++k) v.position[p * 3 + k] = pos[k];
            }
        }
    }
}

// grad_mlp[i] += sum_w ws[w][i], summed in wave order.
// grad[i] += sum over the partial rows, in a fixed order: a workgroup owns 16
// parameters, 16 row groups (rows w = g, g + 16, ...) sum in registers, and
// the 16 group sums are added in group order.
constexpr int kReduceParams = 16;

__device__ __forceinline__ void reduce_partials(const float* __restrict__ ws, int64_t nwaves, int np,
                                                float* __restrict__ grad)
{
    __shared__ float part[16][kReduceParams + 1];
    const int pl = threadIdx.x % kReduceParams, rg = threadIdx.x / kReduceParams;
    const int i = blockIdx.x * kReduceParams + pl;
    float acc = 0.f;
    if (i < np)
        for (int64_t w = rg; w < nwaves; w += 16) acc += ws[w * np + i];
    part[rg][pl] = acc;
    __syncthreads();
    if (rg == 0 && i < np) {
        float t = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) t += part[k][pl];
        grad[i] += t;
    }
}

__global__ __launch_bounds__(256) void reduce_partials_kernel(const float* __restrict__ ws, int64_t nwaves, int np,
                                                              float* __restrict__ grad)
{
    reduce_partials(ws, nwaves, np, grad);
}

// reduce_partials over the rows of whichever adjoint ran (*mode: 1 = rows b).
__global__ __launch_bounds__(256) void reduce_partials_sel_kernel(const float* __restrict__ ws_a, int64_t rows_a,
                                                                  const float* __restrict__ ws_b, int64_t rows_b,
                                                                  const int* __restrict__ mode, int np,
                                                                  float* __restrict__ grad)
{
    if (*mode)
        reduce_partials(ws_b, rows_b, np, grad);
    else
        reduce_partials(ws_a, rows_a, np, grad);
}

// Adam for the grid slots the adjoint did not update itself: nodes split over
// several chunks (gradient in grad_grid) and nodes without samples (zero
// gradient).  Clears grad_grid of those slots.
__global__ __launch_bounds__(256) void grid_adam_rest_kernel(GridAdam ad, float* __restrict__ grad,
                                                             const int32_t* __restrict__ chunk_cnt, int blocks_per_slot,
                                                             int n4)
{
    const int64_t slot = blockIdx.x / blocks_per_slot;
    if (chunk_cnt[slot] == 1) return;
    const int i = (int)(blockIdx.x % blocks_per_slot) * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    const int64_t e = slot * (int64_t)n4 + i;
    float4 gv = reinterpret_cast<float4*>(grad)[e];
    float4 pv = reinterpret_cast<float4*>(ad.p)[e];
    float4 mv = reinterpret_cast<float4*>(ad.m)[e];
    float4 vv = reinterpret_cast<float4*>(ad.v)[e];
    float* gp = &gv.x;
    float* pp = &pv.x;
    float* mp = &mv.x;
    float* vp = &vv.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) pp[k] = adam_update(pp[k], gp[k], mp[k], vp[k], ad.lr, ad.b1, ad.b2, ad.eps, ad.bc1, ad.bc2s);
    reinterpret_cast<float4*>(ad.p)[e] = pv;
    reinterpret_cast<float4*>(ad.m)[e] = mv;
    reinterpret_cast<float4*>(ad.v)[e] = vv;
    reinterpret_cast<float4*>(grad)[e] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ad.pb) {
        ushort4 h;
        h.x = f32_to_bf16_rne(pp[0]);
        h.y = f32_to_bf16_rne(pp[1]);
        h.z = f32_to_bf16_rne(pp[2]);
        h.w = f32_to_bf16_rne(pp[3]);
        reinterpret_cast<ushort4*>(ad.pb)[e] = h;
    }
}

}  // namespace atdf

using namespace atdf;

// 16 rays per wave by default: with every wave of a 2^18-ray batch resident
// at once, 64 rays per wave left the waves holding the longest ray lists
// running alone at the end (512^3 workload, BF16 forward: 1.30 -> 1.17 ms at
// ~39 samples/ray, 3.48 -> 3.03 ms at ~128)
static int64_t rays_per_wave(const atdf_field* f) { return f->rays_per_wave > 0 ? f->rays_per_wave : 16; }

static atdf_status check_project_args(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                      int64_t R, const atdf_sampling* sp, const int64_t* iv_off,
                                      const atdf_intervals* iv, const atdf_field* f)
{
    ATDF_REQUIRE(packed_tree && num_nodes > 0, "project: empty tree");
    ATDF_REQUIRE(R >= 0, "project: negative ray count");
    ATDF_REQUIRE(sp && (sp->jitter_mode == 0 || sp->jitter_mode == 2),
                 "project: jitter_mode must be 0 (midpoint) or 2 (hash)");
    ATDF_REQUIRE(f, "project: null field");
    ATDF_REQUIRE(f->H == 64, "project: hidden width %d not built (64 is)", f->H);
    ATDF_REQUIRE(f->C == 16 || f->C == 8, "project: %d channels not built (16 and 8 are)", f->C);
    ATDF_REQUIRE(f->G >= 2, "project: grid edge must be >= 2");
    ATDF_REQUIRE(f->grid_dtype == ATDF_GRID_F32 || f->grid_dtype == ATDF_GRID_BF16, "project: unknown grid dtype");
    ATDF_REQUIRE(f->decoder_dtype == ATDF_DECODER_F32 || f->decoder_dtype == ATDF_DECODER_BF16,
                 "project: unknown decoder dtype");
    if (f->vm) {
        const atdf_vm* v = f->vm;
        ATDF_REQUIRE(f->grid_dtype == ATDF_GRID_F32, "project: a tensorial field is f32 only");
        ATDF_REQUIRE(v->planes && v->lines && v->node_to_plane && v->num_planes > 0,
                     "project: incomplete tensorial field");
        ATDF_REQUIRE(v->Gp >= 2 && v->Gl >= 2, "project: plane edge and line length must be >= 2");
        ATDF_REQUIRE(v->roi_max_x > v->roi_min_x, "project: empty ROI x range");
        ATDF_REQUIRE((int64_t)v->num_planes * 3 * v->Gp * v->Gp * f->C < ((int64_t)1 << 40),
                     "project: tensorial field too large");
        ATDF_REQUIRE(((uintptr_t)v->planes | (uintptr_t)v->lines) % 16 == 0,
                     "project: planes and lines must be 16-byte aligned");
        ATDF_REQUIRE(f->node_to_grid && f->mlp && f->num_grids >= 0 && (f->num_grids == 0 || f->grid),
                     "project: incomplete field");
    } else {
        ATDF_REQUIRE(f->grid && f->node_to_grid && f->mlp && f->num_grids > 0, "project: incomplete field");
    }
    if (R > 0) {
        ATDF_REQUIRE(O && D && iv_off && iv && iv->node && iv->tmin && iv->tmax && iv->count && iv->first,
                     "project: null ray or interval buffer");
    }
    return ATDF_OK;
}

static ProjArgs make_args(const void* packed_tree, const float* O, const float* D, int64_t R, const atdf_sampling* sp,
                          const int64_t* iv_off, const atdf_intervals* iv, const atdf_field* f)
{
    ProjArgs a{};
    a.nodes = reinterpret_cast<const PNode*>(packed_tree);
    a.O = O;
    a.D = D;
    a.R = R;
    a.jitter_mode = sp->jitter_mode;
    a.seed = sp->jitter_seed;
    a.iv_off = iv_off;
    a.iv_node = iv->node;
    a.iv_tmin = iv->tmin;
    a.iv_tmax = iv->tmax;
    a.iv_count = iv->count;
    a.iv_first = iv->first;
    a.grid = f->grid;
    a.node_to_grid = f->node_to_grid;
    a.mlp = f->mlp;
    a.G = f->G;
    a.beta = f->softplus_beta;
    a.rpw = (int)rays_per_wave(f);
    if (f->vm) {
        a.vm.planes = f->vm->planes;
        a.vm.lines = f->vm->lines;
        a.vm.node_to_plane = f->vm->node_to_plane;
        a.vm.NG = f->num_grids;
        a.vm.Gp = f->vm->Gp;
        a.vm.Gl = f->vm->Gl;
        a.vm.rx0 = f->vm->roi_min_x;
        a.vm.rx1 = f->vm->roi_max_x;
    }
    return a;
}

extern "C" {

static atdf_status project_fwd_impl(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                    int64_t R, const atdf_sampling* sp, const int64_t* iv_off,
                                    const atdf_intervals* iv, const atdf_field* f, float* P, float* bins,
                                    atdf_stream_t stream, const char* who)
{
    atdf_status st = check_project_args(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f);
    if (st) return st;
    if (R == 0) return ATDF_OK;
    ATDF_REQUIRE(P, "%s: null projection", who);
    ProjArgs a = make_args(packed_tree, O, D, R, sp, iv_off, iv, f);
    a.bins = bins;
    int64_t nwaves = div_up(R, a.rpw);
    dim3 grid((unsigned)div_up(nwaves, kWaves)), block(64 * kWaves);
    hipStream_t s = to_hip(stream);
    const bool bf = f->grid_dtype == ATDF_GRID_BF16;
    const bool bd = f->decoder_dtype == ATDF_DECODER_BF16;
    if (f->vm) {
        ATDF_REQUIRE(!bins, "%s: the binned forward does not take a tensorial field", who);
        if (f->C == 16) {
            if (bd) hipLaunchKernelGGL((project_fwd_kernel<16, false, true, false, true>), grid, block, 0, s, a, P);
            else hipLaunchKernelGGL((project_fwd_kernel<16, false, false, false, true>), grid, block, 0, s, a, P);
        } else {
            if (bd) hipLaunchKernelGGL((project_fwd_kernel<8, false, true, false, true>), grid, block, 0, s, a, P);
            else hipLaunchKernelGGL((project_fwd_kernel<8, false, false, false, true>), grid, block, 0, s, a, P);
        }
        return check_launch(who);
    }
#define ATDF_FWD(CC, BF, BD)                                                                              \
    do {                                                                                                  \
        if (bins)                                                                                         \
            hipLaunchKernelGGL((project_fwd_kernel<CC, BF, BD, true>), grid, block, 0, s, a, P);          \
        else                                                                                              \
            hipLaunchKernelGGL((project_fwd_kernel<CC, BF, BD, false>), grid, block, 0, s, a, P);         \
    } while (0)
    if (f->C == 16) {
        if (bf) { if (bd) ATDF_FWD(16, true, true); else ATDF_FWD(16, true, false); }
        else    { if (bd) ATDF_FWD(16, false, true); else ATDF_FWD(16, false, false); }
    } else {
        if (bf) { if (bd) ATDF_FWD(8, true, true); else ATDF_FWD(8, true, false); }
        else    { if (bd) ATDF_FWD(8, false, true); else ATDF_FWD(8, false, false); }
    }
#undef ATDF_FWD
    return check_launch(who);
}

atdf_status atdf_project_fwd(const void* packed_tree, int64_t num_nodes, const float* O, const float* D, int64_t R,
                             const atdf_sampling* sp, const int64_t* iv_off, const atdf_intervals* iv,
                             const atdf_field* f, float* P, atdf_stream_t stream)
{
    return project_fwd_impl(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f, P, nullptr, stream,
                            "atdf_project_fwd");
}

atdf_status atdf_project_fwd_bins(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                  int64_t R, const atdf_sampling* sp, const int64_t* iv_off,
                                  const atdf_intervals* iv, const atdf_field* f, float* P, float* bins,
                                  atdf_stream_t stream)
{
    if (R > 0) ATDF_REQUIRE(bins, "atdf_project_fwd_bins: null bins");
    return project_fwd_impl(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f, P, bins, stream,
                            "atdf_project_fwd_bins");
}

}  // extern "C"

namespace {
struct BwdWorkspace {
    size_t cnt, slot_off, sorted, chunk_cnt, chunk_off, desc, counter, stotal, scan, partials, rec, ivsum, line_part,
        total;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int persistent_blocks()
{
    static int n = 0;
    if (n == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        n = cus;
    }
    return n;
}

// workgroups (and line partial rows) of the tensorial line pass
int vm_line_blocks() { return 2 * persistent_blocks(); }

// line_floats: 3 Gl C of a tensorial field (per-workgroup line partial rows), else 0
BwdWorkspace bwd_layout(int64_t H, int64_t NA, int np, int64_t line_floats = 0)
{
    BwdWorkspace w{};
    size_t o = 0;
    auto take = [&](size_t bytes) { size_t at = o; o = align256(o + bytes); return at; };
    w.cnt = take(sizeof(int32_t) * (NA + 1));
    w.slot_off = take(sizeof(int64_t) * (NA + 1));
    w.sorted = take(sizeof(int32_t) * (H > 0 ? H : 1));
    w.chunk_cnt = take(sizeof(int32_t) * (NA + 1));
    w.chunk_off = take(sizeof(int64_t) * (NA + 1));
    w.desc = take(sizeof(int4) * (H / kMinChunk + NA + 1));
    w.counter = take(256);
    w.stotal = take(sizeof(unsigned long long));
    w.scan = take(atdf_scan_workspace_bytes(NA));
    w.partials = take(sizeof(float) * (size_t)persistent_blocks() * kMaxBwdWaves * np);
    w.rec = take(3 * sizeof(float4) * (size_t)(H > 0 ? H : 1));
    w.ivsum = take(sizeof(float) * (size_t)(H > 0 ? H : 1));
    w.line_part = take(sizeof(float) * (size_t)persistent_blocks() * (size_t)line_floats);
    w.total = o;
    return w;
}

int64_t max_chunks_of(int64_t H, int64_t NA) { return H / kMinChunk + NA + 1; }

// The two-pass adjoint's region, after the base workspace: sorted sample
// counts and offsets, the decoder-pass work items, scan scratch, the
// decoder-pass partial rows, then the staged samples (S x (1 + C/4) float4).
struct StageLayout {
    size_t scnt, soff, part_cnt, part_off, parts, scan, ws2, line_rows, stage;
};

// line_floats: 3 Gl C of a tensorial field (the line pass's partial rows), else 0
StageLayout stage_layout(size_t base, int64_t H, int64_t NA, int np, int64_t line_floats)
{
    StageLayout st{};
    size_t o = align256(base);
    auto take = [&](size_t bytes) { size_t at = o; o = align256(o + bytes); return at; };
    const int64_t mc = max_chunks_of(H, NA);
    st.scnt = take(sizeof(int32_t) * (size_t)(H > 0 ? H : 1));
    st.soff = take(sizeof(int64_t) * (size_t)(H + 1));
    st.part_cnt = take(sizeof(int32_t) * (size_t)(mc + 1));
    st.part_off = take(sizeof(int64_t) * (size_t)(mc + 1));
    st.parts = take(sizeof(int4) * (size_t)(H / kMinPartIv + mc + 1));
    st.scan = take(atdf_scan_workspace_bytes(H > mc ? H : mc));
    st.ws2 = take(sizeof(float) * (size_t)persistent_blocks() * kDecWaves * np);
    st.line_rows = take(sizeof(float) * (size_t)vm_line_blocks() * (size_t)line_floats);
    st.stage = o;
    return st;
}

// grad_floats: the node-gradient region at the start of the dynamic LDS
template <int C, bool kDecBF16>
void lds_layout(int grad_floats, int& f32_off, int& bf16_off, int& tb_off, int& total)
{
    f32_off = (grad_floats + 3) & ~3;
    int after_f32 = (f32_off + weights_f32_floats<C, kDecBF16>()) * 4;
    bf16_off = (after_f32 + 15) & ~15;
    int after_bf = bf16_off + (kDecBF16 ? BSmem<C>::kBf16Total * 2 : 0);
    tb_off = (after_bf + 15) & ~15;
    total = tb_off + (kDecBF16 ? bwd_waves<true>() * 2 * kH * kTB * 2 : kWaves * 2 * kH * kTS * 4);
}

int max_lds_per_block()
{
    static int max_lds = 0;
    if (max_lds == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) {
            (void)hipGetLastError();
            v = 64 * 1024;
        }
        max_lds = v;
    }
    return max_lds;
}

template <int C, bool kBF16, bool kDecBF16, bool kVM>
atdf_status launch_node_bwd(NodeBwdArgs args, hipStream_t s)
{
    // node-gradient region: an explicit grid (G^3 C), or for a tensorial
    // field the larger of that and three planes (3 Gp^2 C) plus the line
    // partials (3 Gl C) behind them
    const int G = args.a.G;
    int64_t grad_floats = (int64_t)G * G * G * C;
    args.lds_line_off = 0;
    if (kVM) {
        const int64_t planes = 3 * (int64_t)args.a.vm.Gp * args.a.vm.Gp * C;
        args.lds_line_off = (int)planes;
        grad_floats = std::max<int64_t>(grad_floats, planes + 3 * (int64_t)args.a.vm.Gl * C);
    }
    if (grad_floats * 4 > max_lds_per_block()) {
        set_error("atdf_project_bwd: node gradient of %lld floats does not fit in LDS", (long long)grad_floats);
        return ATDF_ERR_UNSUPPORTED;
    }
    args.lds_grad_floats = (int)grad_floats;
    int f32_off, bf16_off, tb_off, total;
    lds_layout<C, kDecBF16>((int)grad_floats, f32_off, bf16_off, tb_off, total);
    args.lds_f32_off = f32_off;
    args.lds_bf16_off = bf16_off;
    args.lds_tb_off = tb_off;
    const int max_lds = max_lds_per_block();
    if (total > max_lds) {
        set_error("atdf_project_bwd: node gradient (G=%d, C=%d) needs %d bytes of LDS per workgroup, the device "
                  "allows %d", args.a.G, C, total, max_lds);
        return ATDF_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((project_bwd_node_kernel<C, kBF16, kDecBF16, kVM>), dim3(persistent_blocks()),
                       dim3(64 * bwd_waves<kDecBF16>()), total, s, args);
    return ATDF_OK;
}

// LDS of the tensorial scatter passes: three plane gradients or the three
// lines' partials, and a byte per cell for each of the C/2 waves' three
// owner tables.
size_t vm_plane_lds(int C, int Gp)
{
    return (size_t)3 * Gp * Gp * C * sizeof(float) + (size_t)(C / 2) * 3 * Gp * Gp;
}
size_t vm_line_lds(int C, int Gl) { return (size_t)3 * Gl * C * sizeof(float) + (size_t)(C / 2) * 3 * Gl; }


// The passes of the staged adjoint; all return at once unless *args.mode.
template <int C, bool kBF16, bool kDecBF16, bool kVM = false>
atdf_status launch_two_pass(NodeBwdArgs args, hipStream_t s)
{
    const int G = args.a.G;
    const int64_t grad_floats = (int64_t)G * G * G * C;
    const int max_lds = max_lds_per_block();
    if (grad_floats * 4 > max_lds) {
        set_error("atdf_project_bwd: node gradient of %lld floats does not fit in LDS", (long long)grad_floats);
        return ATDF_ERR_UNSUPPORTED;
    }
    // decoder pass: weights (with the transposed copies on the FP32 path),
    // then kDecWaves pairs of transpose buffers
    const int wf = kDecBF16 ? BSmem<C>::kF32Total : MlpSmemT<C>::kTotal;
    args.lds_f32_off = 0;
    args.lds_bf16_off = (wf * 4 + 15) & ~15;
    args.lds_tb_off = (args.lds_bf16_off + (kDecBF16 ? BSmem<C>::kBf16Total * 2 : 0) + 15) & ~15;
    args.lds_pl_off = args.lds_tb_off + (kDecBF16 ? kDecWaves * 2 * kH * kTB * 2 : kDecWaves * 2 * kH * kTS * 4);
    const int total = args.lds_pl_off + kDecWaves * 64 * 3 * (int)sizeof(float4);
    if (total > max_lds) {
        set_error("atdf_project_bwd: the decoder pass needs %d bytes of LDS, the device allows %d", total, max_lds);
        return ATDF_ERR_UNSUPPORTED;
    }
    const size_t scatter_lds = (size_t)grad_floats * 4 + (size_t)(C / 2) * G * G * G;
    const bool explicit_pass = args.exp_slots > 0;
    if (explicit_pass && scatter_lds > (size_t)max_lds) {
        set_error("atdf_project_bwd: the scatter pass needs %zu bytes of LDS, the device allows %d", scatter_lds,
                  max_lds);
        return ATDF_ERR_UNSUPPORTED;
    }
    const size_t plane_lds = kVM ? vm_plane_lds(C, args.a.vm.Gp) : 0;
    const size_t line_lds = kVM ? vm_line_lds(C, args.a.vm.Gl) : 0;
    if (plane_lds > (size_t)max_lds || line_lds > (size_t)max_lds) {
        set_error("atdf_project_bwd: the tensorial scatter passes need %zu and %zu bytes of LDS, the device allows %d",
                  plane_lds, line_lds, max_lds);
        return ATDF_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((project_bwd_dec_kernel<C, kBF16, kDecBF16>), dim3(persistent_blocks()),
                       dim3(64 * kDecWaves), total, s, args);
    if constexpr (kVM)
        hipLaunchKernelGGL((project_bwd_dec_kernel<C, kBF16, kDecBF16, true>), dim3(persistent_blocks()),
                           dim3(64 * kDecWaves), total, s, args);
    if (explicit_pass)
        hipLaunchKernelGGL((project_bwd_scatter_kernel<C>), dim3(persistent_blocks()), dim3(32 * C), scatter_lds, s,
                           args);
    if constexpr (kVM) {
        hipLaunchKernelGGL((project_bwd_vm_plane_kernel<C>), dim3(persistent_blocks() * 4), dim3(32 * C), plane_lds,
                           s, args);
        // one workgroup per line partial row
        hipLaunchKernelGGL((project_bwd_vm_line_kernel<C>), dim3(vm_line_blocks()), dim3(32 * C), line_lds, s, args);
    }
    return ATDF_OK;
}
}  // namespace

// A tensorial field runs the two-pass adjoint when its explicit grids fit
// the decoder pass's 5-bit corner packing and its plane and line gradients
// fit the tensorial scatter pass's LDS.
bool two_pass_vm_ok(const atdf_field* f)
{
    const size_t max_lds = (size_t)max_lds_per_block();
    return f->vm && (f->num_grids == 0 || f->G <= 32) && vm_plane_lds(f->C, f->vm->Gp) <= max_lds &&
           vm_line_lds(f->C, f->vm->Gl) <= max_lds;
}

extern "C" {

size_t atdf_project_bwd_workspace_bytes(int64_t R, int64_t H, const atdf_field* f)
{
    if (!f || R <= 0 || H < 0 || f->num_grids < 0) return 0;
    const int64_t NA = f->num_grids + (f->vm ? f->vm->num_planes : 0);
    if (NA <= 0) return 0;
    int np = (int)((int64_t)f->H * f->C + f->H + (int64_t)f->H * f->H + f->H + f->H + 1);
    const int64_t lf = f->vm ? 3 * (int64_t)f->vm->Gl * f->C : 0;
    return bwd_layout(H, NA, np, lf).total;
}

size_t atdf_project_bwd_staging_bytes(int64_t H, int64_t S, const atdf_field* f)
{
    if (!f || H < 0 || S < 0 || (f->C != 8 && f->C != 16)) return 0;
    if (!(f->vm ? two_pass_vm_ok(f) : (f->num_grids > 0 && f->G <= 32))) return 0;
    const int64_t NA = f->num_grids + (f->vm ? f->vm->num_planes : 0);
    const int np = (int)((int64_t)f->H * f->C + f->H + (int64_t)f->H * f->H + f->H + f->H + 1);
    const int64_t lf = f->vm ? 3 * (int64_t)f->vm->Gl * f->C : 0;
    const size_t base = bwd_layout(H, NA, np, lf).total;
    const StageLayout st = stage_layout(base, H, NA, np, lf);
    if (f->vm) {  // every sample at the in-ROI records' size (two), bounding any split
        const size_t rf = (size_t)std::max(1 + f->C / 4, 2 * (f->C == 16 ? vm_record_f4<16>() : vm_record_f4<8>()));
        return st.stage - base + (size_t)(S + 4 * kStagePad) * rf * sizeof(float4);
    }
    return st.stage - base + (size_t)(S + kStagePad) * (1 + f->C / 4) * sizeof(float4);
}

}  // extern "C"

// The counting sort of the intervals by grid slot, their records and the
// chunk table, in the caller's workspace (shared by the node-major forward
// and the adjoint).
struct Prepared {
    BwdWorkspace L;
    StageLayout SL;
    unsigned char* w;
    size_t ws_bytes;
    bool staging;  // the workspace holds the two-pass region
    int np;
    int64_t H, NA, line_floats;
    // ints: fused, decoder-pass, scatter-pass counters, mode, prepared flag,
    // forward counter, tensorial scatter-pass and decoder-pass counters
    int* counters() const { return reinterpret_cast<int*>(w + L.counter); }
    int32_t* chunk_cnt() const { return reinterpret_cast<int32_t*>(w + L.chunk_cnt); }
    int32_t* sorted() const { return reinterpret_cast<int32_t*>(w + L.sorted); }
    int64_t* chunk_off() const { return reinterpret_cast<int64_t*>(w + L.chunk_off); }
    int4* desc() const { return reinterpret_cast<int4*>(w + L.desc); }
    float4* rec() const { return reinterpret_cast<float4*>(w + L.rec); }
};

static atdf_status prepared_view(const atdf_intervals* iv, const atdf_field* f, void* workspace,
                                 size_t workspace_bytes, int64_t R, Prepared& P, const char* who)
{
    ATDF_REQUIRE(workspace, "%s: null workspace", who);
    ATDF_REQUIRE(iv->total >= 0 && iv->ray, "%s: intervals need total and ray", who);
    ATDF_REQUIRE(iv->total < (int64_t)INT32_MAX, "%s: too many intervals", who);
    ATDF_REQUIRE(workspace_bytes >= atdf_project_bwd_workspace_bytes(R, iv->total, f), "%s: workspace too small", who);
    P.H = iv->total;
    P.NA = f->num_grids + (f->vm ? f->vm->num_planes : 0);
    P.np = f->C == 16 ? mlp_numel<16>() : mlp_numel<8>();
    P.line_floats = f->vm ? 3 * (int64_t)f->vm->Gl * f->C : 0;
    P.L = bwd_layout(P.H, P.NA, P.np, P.line_floats);
    P.w = reinterpret_cast<unsigned char*>(workspace);
    P.ws_bytes = workspace_bytes;
    P.SL = stage_layout(P.L.total, P.H, P.NA, P.np, P.line_floats);
    // (the decoder pass packs a sample's corner indices in 5 bits each: G <= 32)
    P.staging = (f->C == 8 || f->C == 16) && (f->vm ? two_pass_vm_ok(f) : (f->G <= 32 && f->num_grids > 0)) &&
                workspace_bytes > P.SL.stage;
    return ATDF_OK;
}

static atdf_status prepare_impl(const Prepared& P, const float* O, const float* D, const atdf_intervals* iv,
                                const atdf_field* f, atdf_stream_t stream)
{
    const int64_t H = P.H, NA = P.NA;
    unsigned char* w = P.w;
    const BwdWorkspace& L = P.L;
    int32_t* cnt = reinterpret_cast<int32_t*>(w + L.cnt);
    int64_t* slot_off = reinterpret_cast<int64_t*>(w + L.slot_off);
    int32_t* sorted = P.sorted();
    int32_t* chunk_cnt = P.chunk_cnt();
    int64_t* chunk_off = P.chunk_off();
    int4* desc = P.desc();
    void* scan_ws = w + L.scan;
    hipStream_t s = to_hip(stream);
    const size_t scan_bytes = atdf_scan_workspace_bytes(NA);
    const SlotMap sm{f->node_to_grid, f->vm ? f->vm->node_to_plane : nullptr, f->num_grids};

    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)div_up(NA, 256)), dim3(256), 0, s, cnt, NA, 0);
    // LDS-aggregated counting sort when a slot histogram fits in LDS
    const bool lds_sort = (int64_t)sizeof(int32_t) * NA <= (int64_t)max_lds_per_block();
    const size_t hist_lds = sizeof(int32_t) * (size_t)NA;
    const unsigned sort_blocks = (unsigned)div_up(H, kSortItems);
    if (H > 0) {
        if (lds_sort)
            hipLaunchKernelGGL(slot_hist_lds_kernel, dim3(sort_blocks), dim3(256), hist_lds, s, iv->node, sm, H,
                               NA, cnt);
        else
            hipLaunchKernelGGL(slot_hist_kernel, dim3((unsigned)div_up(H, 256)), dim3(256), 0, s, iv->node, sm, H,
                               cnt);
    }
    atdf_status st = atdf_exclusive_scan_i32(cnt, NA, slot_off, scan_ws, scan_bytes, stream);
    if (st) return st;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)div_up(NA, 256)), dim3(256), 0, s, cnt, NA, 0);
    int32_t* scnt = P.staging ? reinterpret_cast<int32_t*>(w + P.SL.scnt) : nullptr;
    if (scnt && H > 0) hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)div_up(H, 256)), dim3(256), 0, s, scnt, H, 0);
    const RecordSrc rs{iv->tmin, iv->tmax, iv->count, iv->first, iv->ray, O, D, P.rec(), scnt};
    if (H > 0) {
        if (lds_sort)
            hipLaunchKernelGGL(slot_scatter_lds_kernel, dim3(sort_blocks), dim3(256), hist_lds, s, iv->node, sm,
                               H, NA, slot_off, cnt, sorted, rs);
        else
            hipLaunchKernelGGL(slot_scatter_kernel, dim3((unsigned)div_up(H, 256)), dim3(256), 0, s, iv->node, sm,
                               H, slot_off, cnt, sorted, rs);
    }
    unsigned long long* stotal = reinterpret_cast<unsigned long long*>(w + L.stotal);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<int32_t*>(stotal), 2, 0);
    if (H > 0)
        hipLaunchKernelGGL(sum_counts_kernel, dim3((unsigned)std::min<int64_t>(div_up(H, 256), 1024)), dim3(256), 0,
                           s, iv->count, H, stotal);
    hipLaunchKernelGGL(chunk_count_kernel, dim3((unsigned)div_up(NA, 256)), dim3(256), 0, s, slot_off, NA, H, stotal,
                       chunk_cnt);
    st = atdf_exclusive_scan_i32(chunk_cnt, NA, chunk_off, scan_ws, scan_bytes, stream);
    if (st) return st;
    hipLaunchKernelGGL(chunk_desc_kernel, dim3((unsigned)div_up(NA, 256)), dim3(256), 0, s, slot_off, chunk_off, NA,
                       H, stotal, desc);
    int* prepared_flag = P.counters() + 4;
    if (P.staging) {
        const int64_t mc = max_chunks_of(H, NA);
        void* scan2 = w + P.SL.scan;
        const size_t scan2_bytes = atdf_scan_workspace_bytes(H > mc ? H : mc);
        int64_t* soff = reinterpret_cast<int64_t*>(w + P.SL.soff);
        int32_t* part_cnt = reinterpret_cast<int32_t*>(w + P.SL.part_cnt);
        int64_t* part_off = reinterpret_cast<int64_t*>(w + P.SL.part_off);
        st = atdf_exclusive_scan_i32(scnt, H, soff, scan2, scan2_bytes, stream);
        if (st) return st;
        hipLaunchKernelGGL(part_count_kernel, dim3((unsigned)div_up(mc, 256)), dim3(256), 0, s, desc, chunk_off, NA,
                           mc, soff, part_cnt);
        st = atdf_exclusive_scan_i32(part_cnt, mc, part_off, scan2, scan2_bytes, stream);
        if (st) return st;
        hipLaunchKernelGGL(part_desc_kernel, dim3((unsigned)div_up(mc, 256)), dim3(256), 0, s, desc, chunk_off, NA,
                           soff, part_off, reinterpret_cast<int4*>(w + P.SL.parts));
    }
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, s, prepared_flag, 1, P.staging ? 1 : 0);
    return check_launch("atdf_project_prepare");
}

static atdf_status bwd_node_impl(const Prepared& P, const void* packed_tree, const float* O, const float* D,
                                 int64_t R, const atdf_sampling* sp, const int64_t* iv_off, const atdf_intervals* iv,
                                 const atdf_field* f, const float* dLdP, float* grad_grid, float* grad_mlp,
                                 const atdf_grid_adam* adam, atdf_stream_t stream)
{
    const int64_t NA = P.NA;
    const int np = P.np;
    const int64_t n_grid = f->num_grids * (int64_t)f->G * f->G * f->G * f->C;
    hipStream_t s = to_hip(stream);
    int* counter = P.counters();
    float* partials = reinterpret_cast<float*>(P.w + P.L.partials);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, s, counter, 3, 0);
    NodeBwdArgs args{};
    args.a = make_args(packed_tree, O, D, R, sp, iv_off, iv, f);
    args.dLdP = dLdP;
    args.sorted = P.sorted();
    args.chunks = P.desc();
    args.chunk_off = P.chunk_off();
    args.num_slots = NA;
    args.counter = counter;
    args.iv_ray = iv->ray;
    args.rec = P.rec();
    args.grad_grid = grad_grid;
    args.ws = partials;
    args.ad = GridAdam{};
    if (adam) {
        args.ad.p = adam->param;
        args.ad.pb = adam->param_bf16;
        args.ad.m = adam->exp_avg;
        args.ad.v = adam->exp_avg_sq;
        args.ad.lr = adam->lr;
        args.ad.b1 = adam->beta1;
        args.ad.b2 = adam->beta2;
        args.ad.eps = adam->eps;
        // the bias corrections of atdf_adam_step
        args.ad.bc1 = 1.0f - (float)pow((double)adam->beta1, (double)adam->step);
        args.ad.bc2s = sqrtf(1.0f - (float)pow((double)adam->beta2, (double)adam->step));
        args.ad.enabled = 1;
    }
    atdf_status st = ATDF_OK;
    const bool bf = f->grid_dtype == ATDF_GRID_BF16;
    const bool bd = f->decoder_dtype == ATDF_DECODER_BF16;
    if (f->vm) {
        args.grad_planes = f->vm->grad_planes;
        args.line_part = reinterpret_cast<float*>(P.w + P.L.line_part);
    }
    int* mode = counter + 3;
    const int rf_bytes = (1 + f->C / 4) * (int)sizeof(float4);
    const int vm_rf_bytes =
        f->vm ? (f->C == 16 ? vm_record_f4<16>() : vm_record_f4<8>()) * (int)sizeof(float4) : 0;
    args.exp_slots = f->num_grids;
    if (P.staging) {
        const unsigned char* w = P.w;
        const unsigned long long cap = (unsigned long long)(P.ws_bytes - P.SL.stage);
        args.slot_off = reinterpret_cast<const int64_t*>(w + P.L.slot_off);
        args.soff = reinterpret_cast<const int64_t*>(w + P.SL.soff);
        args.counter4 = counter + 6;
        args.counter5 = counter + 7;
        args.line_rows = reinterpret_cast<float*>(P.w + P.SL.line_rows);
        if (f->vm) hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, s, args.counter4, 2, 0);
        hipLaunchKernelGGL(stage_mode_kernel, dim3(1), dim3(64), 0, s,
                           reinterpret_cast<const unsigned long long*>(w + P.L.stotal), counter + 4, cap, rf_bytes,
                           args.soff, args.slot_off, args.exp_slots, vm_rf_bytes, mode);
        args.mode = mode;
        args.parts = reinterpret_cast<const int4*>(w + P.SL.parts);
        args.part_off = reinterpret_cast<const int64_t*>(w + P.SL.part_off);
        args.max_chunks = max_chunks_of(P.H, NA);
        args.stage = reinterpret_cast<float4*>(P.w + P.SL.stage);
        args.ws2 = reinterpret_cast<float*>(P.w + P.SL.ws2);
        args.counter2 = counter + 1;
        args.counter3 = counter + 2;
#define ATDF_TP(CC, BF, BD) st = launch_two_pass<CC, BF, BD>(args, s)
        if (f->vm) {
            if (f->C == 16) {
                if (bd) st = launch_two_pass<16, false, true, true>(args, s);
                else st = launch_two_pass<16, false, false, true>(args, s);
            } else {
                if (bd) st = launch_two_pass<8, false, true, true>(args, s);
                else st = launch_two_pass<8, false, false, true>(args, s);
            }
        } else if (f->C == 16) {
            if (bf) { if (bd) ATDF_TP(16, true, true); else ATDF_TP(16, true, false); }
            else    { if (bd) ATDF_TP(16, false, true); else ATDF_TP(16, false, false); }
        } else {
            if (bf) { if (bd) ATDF_TP(8, true, true); else ATDF_TP(8, true, false); }
            else    { if (bd) ATDF_TP(8, false, true); else ATDF_TP(8, false, false); }
        }
#undef ATDF_TP
        if (st) return st;
    }
#define ATDF_BWD(CC, BF, BD) st = launch_node_bwd<CC, BF, BD, false>(args, s)
    if (f->vm) {
        if (f->C == 16) {
            if (bd) st = launch_node_bwd<16, false, true, true>(args, s);
            else st = launch_node_bwd<16, false, false, true>(args, s);
        } else {
            if (bd) st = launch_node_bwd<8, false, true, true>(args, s);
            else st = launch_node_bwd<8, false, false, true>(args, s);
        }
    } else if (f->C == 16) {
        if (bf) { if (bd) ATDF_BWD(16, true, true); else ATDF_BWD(16, true, false); }
        else    { if (bd) ATDF_BWD(16, false, true); else ATDF_BWD(16, false, false); }
    } else {
        if (bf) { if (bd) ATDF_BWD(8, true, true); else ATDF_BWD(8, true, false); }
        else    { if (bd) ATDF_BWD(8, false, true); else ATDF_BWD(8, false, false); }
    }
#undef ATDF_BWD
    if (st) return st;
    if (adam) {
        const int n4 = (int)(n_grid / NA / 4);
        const int bps = (int)div_up(n4, 256);
        hipLaunchKernelGGL(grid_adam_rest_kernel, dim3((unsigned)(NA * bps)), dim3(256), 0, s, args.ad, grad_grid,
                           P.chunk_cnt(), bps, n4);
    }
    const int waves = bd ? bwd_waves<true>() : bwd_waves<false>();
    if (P.staging)
        hipLaunchKernelGGL(reduce_partials_sel_kernel, dim3((unsigned)div_up(np, kReduceParams)), dim3(256), 0, s,
                           partials, (int64_t)persistent_blocks() * waves, args.ws2,
                           (int64_t)persistent_blocks() * kDecWaves, mode, np, grad_mlp);
    else
        hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)div_up(np, kReduceParams)), dim3(256), 0, s,
                           partials, (int64_t)persistent_blocks() * waves, np, grad_mlp);
    if (f->vm) {
        const int lf = (int)P.line_floats;
        if (P.staging)
            hipLaunchKernelGGL(reduce_partials_sel_kernel, dim3((unsigned)div_up(lf, kReduceParams)), dim3(256), 0, s,
                               args.line_part, (int64_t)persistent_blocks(), args.line_rows,
                               (int64_t)vm_line_blocks(), mode, lf, f->vm->grad_lines);
        else
            hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)div_up(lf, kReduceParams)), dim3(256), 0, s,
                               args.line_part, (int64_t)persistent_blocks(), lf, f->vm->grad_lines);
    }
    return check_launch("atdf_project_bwd");
}

static atdf_status fwd_node_impl(const Prepared& P, const void* packed_tree, const float* O, const float* D,
                                 int64_t R, const atdf_sampling* sp, const int64_t* iv_off, const atdf_intervals* iv,
                                 const atdf_field* f, float* out, atdf_stream_t stream)
{
    hipStream_t s = to_hip(stream);
    NodeFwdArgs args{};
    args.a = make_args(packed_tree, O, D, R, sp, iv_off, iv, f);
    args.sorted = P.sorted();
    args.chunks = P.desc();
    args.chunk_off = P.chunk_off();
    args.num_slots = P.NA;
    args.rec = P.rec();
    args.ivsum = reinterpret_cast<float*>(P.w + P.L.ivsum);
    args.counter = P.counters() + 5;
    const bool bf = f->grid_dtype == ATDF_GRID_BF16;
    const bool bd = f->decoder_dtype == ATDF_DECODER_BF16;
    // the node grid staged in LDS when it fits beside the decoder weights
    // (the weights take at most 24 KB of static LDS)
    const size_t grid_bytes = (size_t)f->G * f->G * f->G * f->C * (bf ? 2 : 4);
    const bool lds = grid_bytes + 24 * 1024 <= (size_t)max_lds_per_block() && grid_bytes % 16 == 0;
    if (lds || P.staging) hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, s, args.counter, 1, 0);
    if (P.staging && P.H > 0) {
        // the part table of the adjoint's two-pass form is in the workspace:
        // work items of ~kPartSamples samples, each sample placed once
        args.soff = reinterpret_cast<const int64_t*>(P.w + P.SL.soff);
        args.parts = reinterpret_cast<const int4*>(P.w + P.SL.parts);
        args.part_off = reinterpret_cast<const int64_t*>(P.w + P.SL.part_off);
        args.max_chunks = max_chunks_of(P.H, P.NA);
        dim3 grid_p((unsigned)persistent_blocks() * 4), block_p(64 * kFwdPartWaves);
#define ATDF_FP(CC, BF, BD) hipLaunchKernelGGL((project_fwd_part_kernel<CC, BF, BD>), grid_p, block_p, 0, s, args)
#define ATDF_FPV(CC, BD) hipLaunchKernelGGL((project_fwd_part_kernel<CC, false, BD, true>), grid_p, block_p, 0, s, args)
        if (f->vm) {  // explicit grids are f32 beside a tensorial field
            if (f->C == 16) { if (bd) ATDF_FPV(16, true); else ATDF_FPV(16, false); }
            else            { if (bd) ATDF_FPV(8, true); else ATDF_FPV(8, false); }
        } else if (f->C == 16) {
            if (bf) { if (bd) ATDF_FP(16, true, true); else ATDF_FP(16, true, false); }
            else    { if (bd) ATDF_FP(16, false, true); else ATDF_FP(16, false, false); }
        } else {
            if (bf) { if (bd) ATDF_FP(8, true, true); else ATDF_FP(8, true, false); }
            else    { if (bd) ATDF_FP(8, false, true); else ATDF_FP(8, false, false); }
        }
#undef ATDF_FP
#undef ATDF_FPV
        hipLaunchKernelGGL(ray_sum_kernel, dim3((unsigned)div_up(R, 256)), dim3(256), 0, s, iv_off, iv->count,
                           args.ivsum, R, out);
        return check_launch("atdf_project_fwd_prepared");
    }
    // chunks are similar in work; a fixed grid strides over them (or takes
    // them from the counter: up to four workgroups per CU, as LDS allows)
    const int64_t max_chunks = P.H / kMinChunk + P.NA + 1;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(max_chunks, (int64_t)persistent_blocks() * 8));
    dim3 grid(blocks), block(64 * kWaves);
    dim3 grid_l((unsigned)persistent_blocks() * 4), block_l(64 * kFwdLdsWaves);
    if (P.H > 0) {
#define ATDF_FN(CC, BF, BD)                                                                                        \
    do {                                                                                                           \
        if (lds)                                                                                                   \
            hipLaunchKernelGGL((project_fwd_node_kernel<CC, BF, BD, true, kFwdLdsWaves>), grid_l, block_l,         \
                               grid_bytes, s, args);                                                               \
        else                                                                                                       \
            hipLaunchKernelGGL((project_fwd_node_kernel<CC, BF, BD>), grid, block, 0, s, args);                    \
    } while (0)
        if (f->C == 16) {
            if (bf) { if (bd) ATDF_FN(16, true, true); else ATDF_FN(16, true, false); }
            else    { if (bd) ATDF_FN(16, false, true); else ATDF_FN(16, false, false); }
        } else {
            if (bf) { if (bd) ATDF_FN(8, true, true); else ATDF_FN(8, true, false); }
            else    { if (bd) ATDF_FN(8, false, true); else ATDF_FN(8, false, false); }
        }
#undef ATDF_FN
    }
    hipLaunchKernelGGL(ray_sum_kernel, dim3((unsigned)div_up(R, 256)), dim3(256), 0, s, iv_off, iv->count,
                       args.ivsum, R, out);
    return check_launch("atdf_project_fwd_prepared");
}

static atdf_status check_adam(const atdf_field* f, const atdf_grid_adam* adam, const float* grad_grid)
{
    const bool bf16 = f->grid_dtype == ATDF_GRID_BF16;
    ATDF_REQUIRE(adam->param && adam->exp_avg && adam->exp_avg_sq && adam->step >= 1,
                 "atdf_project_bwd_adam: incomplete Adam state");
    ATDF_REQUIRE(bf16 ? (adam->param_bf16 && (const void*)adam->param_bf16 == f->grid)
                      : (!adam->param_bf16 && (const void*)adam->param == f->grid),
                 "atdf_project_bwd_adam: the field's grid must be adam->param (f32) or adam->param_bf16 (bf16)");
    ATDF_REQUIRE(((uintptr_t)adam->param | (uintptr_t)adam->exp_avg | (uintptr_t)adam->exp_avg_sq |
                  (uintptr_t)grad_grid) % 16 == 0,
                 "atdf_project_bwd_adam: grid buffers must be 16-byte aligned");
    return ATDF_OK;
}

// no samples: every grid takes a zero-gradient step
static atdf_status adam_without_samples(const atdf_field* f, const atdf_grid_adam* adam, float* grad_grid,
                                        atdf_stream_t stream)
{
    const int64_t n_grid = f->num_grids * (int64_t)f->G * f->G * f->G * f->C;
    return f->grid_dtype == ATDF_GRID_BF16
               ? atdf_adam_step_bf16(adam->param, adam->param_bf16, grad_grid, adam->exp_avg, adam->exp_avg_sq,
                                     n_grid, adam->lr, adam->beta1, adam->beta2, adam->eps, adam->step, 1, stream)
               : atdf_adam_step(adam->param, grad_grid, adam->exp_avg, adam->exp_avg_sq, n_grid, adam->lr,
                                adam->beta1, adam->beta2, adam->eps, adam->step, 1, stream);
}

static atdf_status project_bwd_impl(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                    int64_t R, const atdf_sampling* sp, const int64_t* iv_off,
                                    const atdf_intervals* iv, const atdf_field* f, const float* dLdP,
                                    float* grad_grid, float* grad_mlp, const atdf_grid_adam* adam, void* workspace,
                                    size_t workspace_bytes, bool prepared, atdf_stream_t stream)
{
    atdf_status st = check_project_args(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f);
    if (st) return st;
    if (f->vm) {
        ATDF_REQUIRE(!adam, "atdf_project_bwd: the fused grid Adam does not take a tensorial field");
        ATDF_REQUIRE(f->vm->grad_planes && f->vm->grad_lines, "atdf_project_bwd: null tensorial gradient");
        ATDF_REQUIRE((uintptr_t)f->vm->grad_planes % 16 == 0, "atdf_project_bwd: grad_planes must be 16-byte aligned");
    }
    if (adam) {
        st = check_adam(f, adam, grad_grid);
        if (st) return st;
        if (R == 0) return adam_without_samples(f, adam, grad_grid, stream);
    }
    if (R == 0) return ATDF_OK;
    ATDF_REQUIRE(dLdP && (grad_grid || f->num_grids == 0) && grad_mlp, "atdf_project_bwd: null buffer");
    Prepared P;
    st = prepared_view(iv, f, workspace, workspace_bytes, R, P, "atdf_project_bwd");
    if (st) return st;
    if (!prepared) {
        st = prepare_impl(P, O, D, iv, f, stream);
        if (st) return st;
    }
    return bwd_node_impl(P, packed_tree, O, D, R, sp, iv_off, iv, f, dLdP, grad_grid, grad_mlp, adam, stream);
}

extern "C" {

atdf_status atdf_project_bwd(const void* packed_tree, int64_t num_nodes, const float* O, const float* D, int64_t R,
                             const atdf_sampling* sp, const int64_t* iv_off, const atdf_intervals* iv,
                             const atdf_field* f, const float* dLdP, float* grad_grid, float* grad_mlp,
                             void* workspace, size_t workspace_bytes, atdf_stream_t stream)
{
    return project_bwd_impl(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f, dLdP, grad_grid, grad_mlp, nullptr,
                            workspace, workspace_bytes, false, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The adjoint for callers that cannot allocate (HIP graph capture, serving
// loops): its counting sort, interval records, chunk table and decoder-
// gradient partials live in `scratch`, a fixed device array of the library
// (each GPU the library runs on holds its own copy until the process ends).
// In host code the array's name is only a stand-in for the symbol; its device
// address on the current device comes from hipGetSymbolAddress, looked up
// once per device and kept.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU (two streams, two host threads) race.  The
// kernels reset everything they accumulate into (counts, counters, chunk
// table) before use, so a call does not depend on what the previous one left.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchBytes = size_t(256) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

atdf_status scratch_for_current_device(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("atdf_project_bwd_noalloc: hipGetDevice: %s", hipGetErrorString(e));
        return ATDF_ERR_HIP;
    }
    ATDF_REQUIRE(dev >= 0 && dev < kMaxDevices, "atdf_project_bwd_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("atdf_project_bwd_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
            return ATDF_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return ATDF_OK;
}
}  // namespace

extern "C" {

size_t atdf_project_bwd_noalloc_max_bytes(void) { return kScratchBytes; }

atdf_status atdf_project_bwd_noalloc(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                     int64_t R, const atdf_sampling* sp, const int64_t* iv_off,
                                     const atdf_intervals* iv, const atdf_field* f, const float* dLdP,
                                     float* grad_grid, float* grad_mlp, atdf_stream_t stream)
{
    atdf_status st = check_project_args(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f);
    if (st) return st;
    if (R == 0) return ATDF_OK;
    const size_t need = atdf_project_bwd_workspace_bytes(R, iv->total, f);
    ATDF_REQUIRE(need <= kScratchBytes,
                 "atdf_project_bwd_noalloc: %lld intervals over %lld grids need %zu bytes of scratch, more than the "
                 "%zu the library holds; use atdf_project_bwd with a workspace",
                 (long long)iv->total, (long long)f->num_grids, need, kScratchBytes);
    void* ws = nullptr;
    st = scratch_for_current_device(&ws);
    if (st) return st;
    return project_bwd_impl(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f, dLdP, grad_grid, grad_mlp, nullptr,
                            ws, kScratchBytes, false, stream);
}

atdf_status atdf_project_bwd_adam(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                  int64_t R, const atdf_sampling* sp, const int64_t* iv_off, const atdf_intervals* iv,
                                  const atdf_field* f, const float* dLdP, float* grad_grid, float* grad_mlp,
                                  const atdf_grid_adam* adam, void* workspace, size_t workspace_bytes,
                                  atdf_stream_t stream)
{
    ATDF_REQUIRE(adam, "atdf_project_bwd_adam: null Adam state");
    return project_bwd_impl(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f, dLdP, grad_grid, grad_mlp, adam,
                            workspace, workspace_bytes, false, stream);
}

atdf_status atdf_project_prepare(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                 int64_t R, const atdf_sampling* sp, const int64_t* iv_off, const atdf_intervals* iv,
                                 const atdf_field* f, void* workspace, size_t workspace_bytes, atdf_stream_t stream)
{
    atdf_status st = check_project_args(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f);
    if (st) return st;
    if (R == 0) return ATDF_OK;
    Prepared P;
    st = prepared_view(iv, f, workspace, workspace_bytes, R, P, "atdf_project_prepare");
    if (st) return st;
    return prepare_impl(P, O, D, iv, f, stream);
}

atdf_status atdf_project_fwd_prepared(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                      int64_t R, const atdf_sampling* sp, const int64_t* iv_off,
                                      const atdf_intervals* iv, const atdf_field* f, float* P_out, void* workspace,
                                      size_t workspace_bytes, atdf_stream_t stream)
{
    atdf_status st = check_project_args(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f);
    if (st) return st;
    if (R == 0) return ATDF_OK;
    ATDF_REQUIRE(P_out, "atdf_project_fwd_prepared: null projection");
    Prepared P;
    st = prepared_view(iv, f, workspace, workspace_bytes, R, P, "atdf_project_fwd_prepared");
    if (st) return st;
    if (f->vm && !(P.staging && P.H > 0))  // the node-major tensorial forward needs the part table
        return atdf_project_fwd(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f, P_out, stream);
    return fwd_node_impl(P, packed_tree, O, D, R, sp, iv_off, iv, f, P_out, stream);
}

atdf_status atdf_project_bwd_prepared(const void* packed_tree, int64_t num_nodes, const float* O, const float* D,
                                      int64_t R, const atdf_sampling* sp, const int64_t* iv_off,
                                      const atdf_intervals* iv, const atdf_field* f, const float* dLdP,
                                      float* grad_grid, float* grad_mlp, const atdf_grid_adam* adam, void* workspace,
                                      size_t workspace_bytes, atdf_stream_t stream)
{
    return project_bwd_impl(packed_tree, num_nodes, O, D, R, sp, iv_off, iv, f, dLdP, grad_grid, grad_mlp, adam,
                            workspace, workspace_bytes, true, stream);
}



atdf_status atdf_sample_volume(const void* packed_tree, int64_t num_nodes, const int64_t* active_node_ids,
                               const atdf_field* f, const atdf_lattice* lattice, float* density, int64_t* node_id,
                               float* valid, float* position, atdf_stream_t stream)
{
    ATDF_REQUIRE(packed_tree && num_nodes > 0 && lattice && active_node_ids, "atdf_sample_volume: null argument");
    ATDF_REQUIRE(!f || !f->vm, "atdf_sample_volume: a tensorial field is not supported");
    ATDF_REQUIRE(lattice->mode == 0 || lattice->mode == 1, "atdf_sample_volume: mode must be 0 or 1");
    int64_t total = 1;
    for (int k = 0; k < 3; ++k) {
        ATDF_REQUIRE(lattice->size[k] >= (lattice->mode == 0 ? 2 : 1), "atdf_sample_volume: bad size[%d] = %d", k,
                     lattice->size[k]);
        total *= lattice->size[k];
    }
    VolumeArgs v{};
    v.active_ids = active_node_ids;
    v.mode = lattice->mode;
    v.swap = lattice->swap_xy != 0;
    for (int k = 0; k < 3; ++k) {
        v.size[k] = lattice->size[k];
        v.roi_min[k] = lattice->roi_min[k];
        v.step[k] = lattice->step[k];
    }
    v.total = total;
    v.density = density;
    v.node_id = node_id;
    v.valid = valid;
    v.position = position;
    hipStream_t s = to_hip(stream);
    if (!density) {  // positions / point location only
        v.a.nodes = reinterpret_cast<const PNode*>(packed_tree);
        hipLaunchKernelGGL(volume_points_kernel, dim3((unsigned)div_up(total, 256)), dim3(256), 0, s, v);
        return check_launch("atdf_sample_volume");
    }
    ATDF_REQUIRE(f && f->H == 64 && (f->C == 16 || f->C == 8), "atdf_sample_volume: field shape not built");
    ATDF_REQUIRE(f->grid && f->node_to_grid && f->mlp, "atdf_sample_volume: incomplete field");
    atdf_sampling sp{};
    atdf_intervals none{};
    v.a = make_args(packed_tree, nullptr, nullptr, 0, &sp, nullptr, &none, f);
    const int64_t tiles = div_up(total, kTile);
    const unsigned blocks = (unsigned)std::min<int64_t>(div_up(tiles, kWaves), 65536);
    const bool bf = f->grid_dtype == ATDF_GRID_BF16;
    const bool bd = f->decoder_dtype == ATDF_DECODER_BF16;
#define ATDF_VOL(CC, BF, BD) hipLaunchKernelGGL((volume_kernel<CC, BF, BD>), dim3(blocks), dim3(64 * kWaves), 0, s, v)
    if (f->C == 16) {
        if (bf) { if (bd) ATDF_VOL(16, true, true); else ATDF_VOL(16, true, false); }
        else    { if (bd) ATDF_VOL(16, false, true); else ATDF_VOL(16, false, false); }
    } else {
        if (bf) { if (bd) ATDF_VOL(8, true, true); else ATDF_VOL(8, true, false); }
        else    { if (bd) ATDF_VOL(8, false, true); else ATDF_VOL(8, false, false); }
    }
#undef ATDF_VOL
    return check_launch("atdf_sample_volume");
}

atdf_status atdf_node_lattice_points(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                     int64_t num_active, int32_t n, float* positions, atdf_stream_t stream)
{
    ATDF_REQUIRE(packed_tree && num_nodes > 0, "atdf_node_lattice_points: empty tree");
    ATDF_REQUIRE(n >= 2 && n <= 1024, "atdf_node_lattice_points: lattice size %d outside [2, 1024]", n);
    ATDF_REQUIRE(num_active >= 0 && num_active < (int64_t)INT32_MAX, "atdf_node_lattice_points: bad node count");
    if (num_active == 0) return ATDF_OK;
    ATDF_REQUIRE(active_ids && positions, "atdf_node_lattice_points: null buffer");
    hipLaunchKernelGGL(lattice_points_kernel, dim3((unsigned)num_active), dim3(256), 0, to_hip(stream),
                       reinterpret_cast<const PNode*>(packed_tree), active_ids, n, positions);
    return check_launch("atdf_node_lattice_points");
}

atdf_status atdf_node_density_stats(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                    int64_t num_active, const atdf_field* f, int32_t mode, int32_t n, int32_t rounds,
                                    uint64_t seed, float* node_max, double* node_sum, double* node_sumsq,
                                    atdf_stream_t stream)
{
    return atdf_node_density_stats_masked(packed_tree, num_nodes, active_ids, num_active, f, mode, n, rounds, seed,
                                          nullptr, node_max, node_sum, node_sumsq, stream);
}

atdf_status atdf_node_density_stats_masked(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                           int64_t num_active, const atdf_field* f, int32_t mode, int32_t n,
                                           int32_t rounds, uint64_t seed, const float* point_mask, float* node_max,
                                           double* node_sum, double* node_sumsq, atdf_stream_t stream)
{
    ATDF_REQUIRE(packed_tree && num_nodes > 0 && f, "atdf_node_density_stats: empty tree or field");
    ATDF_REQUIRE(!f->vm, "atdf_node_density_stats: a tensorial field is not supported");
    ATDF_REQUIRE(f->H == 64 && (f->C == 16 || f->C == 8), "atdf_node_density_stats: field shape not built");
    ATDF_REQUIRE(f->grid && f->node_to_grid && f->mlp, "atdf_node_density_stats: incomplete field");
    ATDF_REQUIRE(mode == 0 || mode == 1, "atdf_node_density_stats: mode must be 0 (lattice) or 1 (random)");
    ATDF_REQUIRE(mode == 1 || n >= 2, "atdf_node_density_stats: lattice needs n >= 2");
    ATDF_REQUIRE(n >= 1 && (mode == 0 || rounds >= 1), "atdf_node_density_stats: bad point count");
    if (num_active == 0) return ATDF_OK;
    ATDF_REQUIRE(active_ids && node_max && node_sum && node_sumsq, "atdf_node_density_stats: null buffer");
    StatsArgs sa{};
    atdf_sampling sp{};
    atdf_intervals none{};
    sa.a = make_args(packed_tree, nullptr, nullptr, 0, &sp, nullptr, &none, f);
    sa.active_ids = active_ids;
    sa.mode = mode;
    sa.n = n;
    sa.rounds = rounds;
    sa.seed = seed;
    sa.node_max = node_max;
    sa.node_sum = node_sum;
    sa.node_sumsq = node_sumsq;
    sa.point_mask = point_mask;
    dim3 grid((unsigned)num_active), block(64 * kWaves);
    hipStream_t s = to_hip(stream);
    const bool bf = f->grid_dtype == ATDF_GRID_BF16;
    const bool bd = f->decoder_dtype == ATDF_DECODER_BF16;
#define ATDF_ST(CC, BF, BD) hipLaunchKernelGGL((node_stats_kernel<CC, BF, BD>), grid, block, 0, s, sa)
    if (f->C == 16) {
        if (bf) { if (bd) ATDF_ST(16, true, true); else ATDF_ST(16, true, false); }
        else    { if (bd) ATDF_ST(16, false, true); else ATDF_ST(16, false, false); }
    } else {
        if (bf) { if (bd) ATDF_ST(8, true, true); else ATDF_ST(8, true, false); }
        else    { if (bd) ATDF_ST(8, false, true); else ATDF_ST(8, false, false); }
    }
#undef ATDF_ST
    return check_launch("atdf_node_density_stats");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// DecodeFeatures on given feature rows (HierarchicalNeuralGeometry::
// DecodeFeatures, geometry/geometry.cpp:176-253, softplus output):
// density[i] = softplus(MLP(x_i)), and its adjoint.  Tiles of 16 rows per
// wave with the fused kernels' lane layout (lane (g, j): row j, channels
// CPL*g ..), the same decoder code on MFMA, a fixed grid-stride assignment of
// tiles to waves, and the decoder-parameter gradient through per-wave
// partial rows and the fixed-order reduction of the projection adjoint, so
// both directions are deterministic.
// ---------------------------------------------------------------------------
namespace atdf {

struct DecodeArgs {
    const float* mlp;
    const float* x;   // [n, C]
    const float* dd;  // [n] dL/ddensity (adjoint)
    float* density;   // [n] (forward)
    float* dx;        // [n, C] (adjoint, written)
    float* partials;  // [blocks * kWaves, NP] (adjoint)
    int64_t n;
    float beta;
};

constexpr int kDecodeMaxBlocks = 1024;

template <int C>
__device__ __forceinline__ void load_feature_row(const DecodeArgs& a, int64_t row, int g, float* x)
{
    constexpr int CPL = C / 4;
#pragma unroll
    for (int q = 0; q < CPL; ++q) x[q] = row < a.n ? a.x[row * C + CPL * g + q] : 0.f;
}

template <int C, bool kDecBF16>
__global__ __launch_bounds__(64 * kWaves, 4) void decode_fwd_kernel(DecodeArgs a)
{
    using SM = DecoderSmem<C, kDecBF16>;
    constexpr int CPL = C / 4;
    __shared__ __attribute__((aligned(16))) float s[SM::kF32];
    __shared__ __attribute__((aligned(16))) __bf16 sb[SM::kBF16];
    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds<C>(a.mlp, s);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int64_t ntiles = (a.n + kTile - 1) / kTile;
    for (int64_t tile = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6); tile < ntiles;
         tile += (int64_t)gridDim.x * kWaves) {
        const int64_t row = tile * kTile + j;
        float x[CPL];
        load_feature_row<C>(a, row, g, x);
        v4f h1[4], h2[4];
        int opq = 0;
        asm volatile("" : "+v"(opq));  // no hoisting of the weight reads
        const float* so = s + opq;
        const __bf16* sbo = sb + opq;
        float y;
        if constexpr (kDecBF16)
            y = decoder_fwd_bf16<C>(sbo, so, x, g, j, h1, h2);
        else
            y = decoder_fwd<C>(so, x, g, j, h1, h2);
        if (g == 0 && row < a.n) a.density[row] = softplus_f(y, a.beta);
    }
}

template <int C, bool kDecBF16>
__global__ __launch_bounds__(64 * kWaves) void decode_bwd_kernel(DecodeArgs a)
{
    using SM = DecoderSmem<C, kDecBF16>;
    constexpr int CPL = C / 4;
    constexpr int NP = mlp_numel<C>();
    __shared__ __attribute__((aligned(16))) float s[SM::kF32];
    __shared__ __attribute__((aligned(16))) __bf16 sb[SM::kBF16];
    __shared__ __attribute__((aligned(16))) float tb[kDecBF16 ? 4 : kWaves * 2 * kH * kTS];
    __shared__ __attribute__((aligned(16))) __bf16 tbb[kDecBF16 ? kWaves * 2 * kH * kTB : 8];
    if constexpr (kDecBF16)
        load_mlp_bf16_to_lds<C>(a.mlp, sb, s);
    else
        load_mlp_to_lds<C>(a.mlp, s);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int wl = threadIdx.x >> 6;
    float* bufA = tb + (kDecBF16 ? 0 : wl * 2 * kH * kTS);
    float* bufB = bufA + (kDecBF16 ? 0 : kH * kTS);
    __bf16* bufAb = tbb + (kDecBF16 ? wl * 2 * kH * kTB : 0);
    __bf16* bufBb = bufAb + (kDecBF16 ? kH * kTB : 0);
    DecGrads q;
    zero_grads(q);
    const int64_t ntiles = (a.n + kTile - 1) / kTile;
    for (int64_t tile = (int64_t)blockIdx.x * kWaves + wl; tile < ntiles; tile += (int64_t)gridDim.x * kWaves) {
        const int64_t row = tile * kTile + j;
        float x[CPL];
        load_feature_row<C>(a, row, g, x);
        v4f h1[4], h2[4];
        float y;
        if constexpr (kDecBF16)
            y = decoder_fwd_bf16<C>(sb, s, x, g, j, h1, h2);
        else
            y = decoder_fwd<C>(s, x, g, j, h1, h2);
        const float dy = row < a.n ? a.dd[row] * softplus_grad_f(y, a.beta) : 0.f;
        const v4f dx = tile_backward<C, kDecBF16>(s, sb, bufA, bufB, bufAb, bufBb, x, h1, h2, dy, g, j, q);
        if (row < a.n && 4 * g < C) {
#pragma unroll
            for (int k = 0; k < 4; ++k) a.dx[row * C + 4 * g + k] = dx[k];
        }
    }
    store_partials<C>(q, a.partials + ((int64_t)blockIdx.x * kWaves + wl) * NP, g, j, lane);
}

int decode_blocks(int64_t n)
{
    const int64_t ntiles = div_up(n, kTile);
    return (int)std::min<int64_t>(div_up(ntiles, kWaves), kDecodeMaxBlocks);
}

atdf_status check_decoder(const atdf_field* f, const char* who)
{
    ATDF_REQUIRE(f, "%s: null field", who);
    ATDF_REQUIRE(f->H == 64 && (f->C == 16 || f->C == 8), "%s: decoder shape C=%d H=%d not built (C 16 or 8, H 64)",
                 who, f->C, f->H);
    ATDF_REQUIRE(f->mlp, "%s: null decoder parameters", who);
    ATDF_REQUIRE(f->decoder_dtype == ATDF_DECODER_F32 || f->decoder_dtype == ATDF_DECODER_BF16,
                 "%s: unknown decoder dtype %d", who, f->decoder_dtype);
    return ATDF_OK;
}

}  // namespace atdf

extern "C" {

atdf_status atdf_decode_features(const float* features, int64_t n, const atdf_field* f, float* density,
                                 atdf_stream_t stream)
{
    atdf_status st = check_decoder(f, "atdf_decode_features");
    if (st) return st;
    ATDF_REQUIRE(n >= 0, "atdf_decode_features: negative row count");
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(features && density, "atdf_decode_features: null buffer");
    DecodeArgs a{};
    a.mlp = f->mlp;
    a.x = features;
    a.density = density;
    a.n = n;
    a.beta = f->softplus_beta;
    const dim3 grid((unsigned)std::min<int64_t>(div_up(div_up(n, kTile), kWaves), 65535)), block(64 * kWaves);
    hipStream_t s = to_hip(stream);
    const bool bd = f->decoder_dtype == ATDF_DECODER_BF16;
    if (f->C == 16) {
        if (bd) hipLaunchKernelGGL((decode_fwd_kernel<16, true>), grid, block, 0, s, a);
        else    hipLaunchKernelGGL((decode_fwd_kernel<16, false>), grid, block, 0, s, a);
    } else {
        if (bd) hipLaunchKernelGGL((decode_fwd_kernel<8, true>), grid, block, 0, s, a);
        else    hipLaunchKernelGGL((decode_fwd_kernel<8, false>), grid, block, 0, s, a);
    }
    return check_launch("atdf_decode_features");
}

size_t atdf_decode_features_bwd_workspace_bytes(int64_t n, const atdf_field* f)
{
    if (!f || n <= 0 || (f->C != 16 && f->C != 8)) return 0;
    const int64_t np = (int64_t)f->H * f->C + f->H + (int64_t)f->H * f->H + f->H + f->H + 1;
    return sizeof(float) * (size_t)decode_blocks(n) * kWaves * (size_t)np;
}

atdf_status atdf_decode_features_bwd(const float* features, int64_t n, const atdf_field* f,
                                     const float* grad_density, float* grad_features, float* grad_mlp,
                                     void* workspace, size_t workspace_bytes, atdf_stream_t stream)
{
    atdf_status st = check_decoder(f, "atdf_decode_features_bwd");
    if (st) return st;
    ATDF_REQUIRE(n >= 0, "atdf_decode_features_bwd: negative row count");
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(features && grad_density && grad_features && grad_mlp, "atdf_decode_features_bwd: null buffer");
    const size_t need = atdf_decode_features_bwd_workspace_bytes(n, f);
    ATDF_REQUIRE(workspace && workspace_bytes >= need, "atdf_decode_features_bwd: workspace of %zu bytes, %zu needed",
                 workspace_bytes, need);
    DecodeArgs a{};
    a.mlp = f->mlp;
    a.x = features;
    a.dd = grad_density;
    a.dx = grad_features;
    a.partials = static_cast<float*>(workspace);
    a.n = n;
    a.beta = f->softplus_beta;
    const int blocks = decode_blocks(n);
    const int np = (int)(need / sizeof(float) / ((size_t)blocks * kWaves));
    hipStream_t s = to_hip(stream);
    const bool bd = f->decoder_dtype == ATDF_DECODER_BF16;
    const dim3 grid((unsigned)blocks), block(64 * kWaves);
    if (f->C == 16) {
        if (bd) hipLaunchKernelGGL((decode_bwd_kernel<16, true>), grid, block, 0, s, a);
        else    hipLaunchKernelGGL((decode_bwd_kernel<16, false>), grid, block, 0, s, a);
    } else {
        if (bd) hipLaunchKernelGGL((decode_bwd_kernel<8, true>), grid, block, 0, s, a);
        else    hipLaunchKernelGGL((decode_bwd_kernel<8, false>), grid, block, 0, s, a);
    }
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)div_up(np, kReduceParams)), dim3(256), 0, s,
                       a.partials, (int64_t)blocks * kWaves, np, grad_mlp);
    return check_launch("atdf_decode_features_bwd");
}

}  // extern "C"

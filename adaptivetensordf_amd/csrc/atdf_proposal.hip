// atdf_proposal.hip -- the learned sample proposal on gfx950: point-location
// table, ray features (root chord + 32-bin coarse occupancy), the proposal
// MLP 38 -> 64 -> 64 -> 64 -> 32 (inference and training), sample placement
// with the KL fallback gate, and the merge of learned and stratified
// intervals into one list for the fused projection kernels.
//
// Inference runs one ray per lane in FP32 VALU with the weights read as
// wave-uniform scalar loads (SGPR operands) and a per-wave activation buffer
// in LDS; every sum runs over the input index in increasing order with no
// contraction, so logits, KL, fallback decisions and placements match the C
// restatement bit for bit.  (Training keeps the weights in LDS.)
// The network is tiny (25 kFLOP per ray) next to the 32 decoder evaluations
// it replaces 96 of; the kernels are bound by the point-location gathers,
// not by arithmetic.
#include <hip/hip_runtime.h>

#include "atdf_device.h"
#include "atdf_internal.h"

namespace atdf {
namespace {

constexpr int NB = ATDF_PROPOSAL_BINS;
constexpr int NI = ATDF_PROPOSAL_INPUTS;
constexpr int NH = ATDF_PROPOSAL_HIDDEN;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NB * NH, NPARAM = OB4 + NB;

// LDS copy of the parameters: W1 rows padded to 40 floats so every row
// starts 16-byte aligned; the other layers keep the packed layout.
constexpr int kW1S = 40;
constexpr int LW1 = 0, LB1 = LW1 + NH * kW1S, LW2 = LB1 + NH, LB2 = LW2 + NH * NH, LW3 = LB2 + NH,
              LB3 = LW3 + NH * NH, LW4 = LB3 + NH, LB4 = LW4 + NB * NH, LTOTAL = LB4 + NB;
constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kActFloats = NH * 64;  // per-wave activation buffer [64 rows][64 lanes]

__device__ __forceinline__ void load_params(const float* __restrict__ P, float* s)
{
    for (int i = threadIdx.x; i < LTOTAL; i += blockDim.x) {
        float v;
        if (i < LB1) {
            const int o = i / kW1S, k = i - o * kW1S;
            v = k < NI ? P[OW1 + o * NI + k] : 0.f;
        } else {
            v = P[OB1 + (i - LB1)];
        }
        s[i] = v;
    }
}

// e^x and ln x of the proposal: Cody-Waite reduction with a 15-bit ln2 head
// and a degree-7 Taylor polynomial; ln via frexp and the atanh series to s^9.
__device__ __forceinline__ float prop_exp(float x)
{
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    float n = floorf(x * 1.44269504f + 0.5f);
    float r = x - n * 0.693145751953125f;
    r = r - n * 1.42860682e-6f;
    float p = 1.98412698e-4f;
    p = p * r + 1.38888889e-3f;
    p = p * r + 8.33333333e-3f;
    p = p * r + 4.16666667e-2f;
    p = p * r + 1.66666667e-1f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexpf(p, (int)n);
}

__device__ __forceinline__ float prop_log(float x)
{
    if (!(x > 0.0f)) return -INFINITY;
    if (x > 3.0e38f) return 88.7228394f;
    int e;
    float m = frexpf(x, &e);
    if (m < 0.70710678f) {
        m = m * 2.0f;
        e = e - 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 0.111111111f;
    p = p * s2 + 0.142857143f;
    p = p * s2 + 0.2f;
    p = p * s2 + 0.333333333f;
    p = p * s2 + 1.0f;
    return (float)e * 0.693147181f + 2.0f * s * p;
}

__device__ __forceinline__ int32_t locate(const PNode& root, const int32_t* __restrict__ table, int depth,
                                          const float* p)
{
    const int res = 1 << depth;
    int ci[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float u = (p[k] - root.mn[k]) / (root.mx[k] - root.mn[k]);
        int c = (int)floorf(u * (float)res);
        ci[k] = c < 0 ? 0 : (c > res - 1 ? res - 1 : c);
    }
    return table[((int64_t)ci[2] * res + ci[1]) * res + ci[0]];
}

__device__ __forceinline__ float bin_edge(float t0, float t1, float len, int b)
{
    return b == NB ? t1 : t0 + (float)b * len;
}

// Root chord; false (t0 = t1 = 0) when missing or shorter than 1e-5.
__device__ __forceinline__ bool root_chord(const PNode& root, const float* o, const float* d, float& t0, float& t1)
{
    bool hit = intersect_box_ray(root.mn, root.mx, o, d, t0, t1);
    if (!hit || t1 - t0 < 1e-5) {
        t0 = 0.f;
        t1 = 0.f;
        return false;
    }
    return true;
}

// x[0..5] = origin, direction; x[6..37] = occupancy of the 32 bins.
__device__ __forceinline__ bool ray_features(const PNode& root, const int32_t* __restrict__ table, int depth,
                                             const float* __restrict__ coarse, const float* o, const float* d,
                                             float* x, float& t0, float& t1)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = o[k];
        x[3 + k] = d[k];
    }
    const bool ok = root_chord(root, o, d, t0, t1);
    const float len = (t1 - t0) / (float)NB;
    int32_t node[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        node[b] = -1;
        if (ok) {
            float mid = (bin_edge(t0, t1, len, b) + bin_edge(t0, t1, len, b + 1)) * 0.5f;
            float pos[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) pos[k] = o[k] + d[k] * mid;
            node[b] = locate(root, table, depth, pos);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) x[6 + b] = node[b] >= 0 ? coarse[node[b]] : 0.f;
    return ok;
}

// y[o] = act(sum_k W[o][k] x[k] + b[o]) for o < NOUT into this lane's column
// (the sum as a chain of fused multiply-adds in k order, then + b)
// of buf ([o][64 lanes]); W rows of `stride` floats in LDS.
template <int NIN, int STRIDE, int NOUT, bool RELU>
__device__ __forceinline__ void dense_lane(const float* W, const float* b, const float* x, float* buf, int lane)
{
    for (int o = 0; o < NOUT; ++o) {
        const float* wr = W + o * STRIDE;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k + 4 <= NIN; k += 4) {
            const float4 w = *reinterpret_cast<const float4*>(wr + k);
            acc = fmaf(w.x, x[k], acc);
            acc = fmaf(w.y, x[k + 1], acc);
            acc = fmaf(w.z, x[k + 2], acc);
            acc = fmaf(w.w, x[k + 3], acc);
        }
#pragma unroll
        for (int k = NIN & ~3; k < NIN; ++k) acc = fmaf(wr[k], x[k], acc);
        acc = acc + b[o];
        buf[o * 64 + lane] = RELU ? (acc > 0.f ? acc : 0.f) : acc;
    }
}

// Same as dense_lane with the weights read from the packed global parameters
// (rows of NIN floats): the addresses are wave-uniform, so they become scalar
// loads, and the LDS holds only the activations.
template <int NIN, int NOUT, bool RELU>
__device__ __forceinline__ void dense_lane_s(const float* __restrict__ W, const float* __restrict__ b, const float* x,
                                             float* buf, int lane)
{
    for (int o = 0; o < NOUT; ++o) {
        const float* wr = W + o * NIN;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < NIN; ++k) acc = fmaf(wr[k], x[k], acc);
        acc = acc + b[o];
        buf[o * 64 + lane] = RELU ? (acc > 0.f ? acc : 0.f) : acc;
    }
}

template <int N>
__device__ __forceinline__ void load_column(const float* buf, int lane, float* v)
{
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = buf[k * 64 + lane];
}

// Logits z[32] of one ray from the global parameters P (buf = this wave's
// activation buffer).
__device__ __forceinline__ void mlp_forward_s(const float* __restrict__ P, float* buf, int lane, const float* x,
                                              float* z)
{
    float h[NH];
    dense_lane_s<NI, NH, true>(P + OW1, P + OB1, x, buf, lane);
    load_column<NH>(buf, lane, h);
    dense_lane_s<NH, NH, true>(P + OW2, P + OB2, h, buf, lane);
    load_column<NH>(buf, lane, h);
    dense_lane_s<NH, NH, true>(P + OW3, P + OB3, h, buf, lane);
    load_column<NH>(buf, lane, h);
    dense_lane_s<NH, NB, false>(P + OW4, P + OB4, h, buf, lane);
    load_column<NB>(buf, lane, z);
}

// Logits z[32] of one ray (s = LDS parameters, buf = this wave's buffer).
__device__ __forceinline__ void mlp_forward(const float* s, float* buf, int lane, const float* x, float* z)
{
    float h[NH];
    dense_lane<NI, kW1S, NH, true>(s + LW1, s + LB1, x, buf, lane);
    load_column<NH>(buf, lane, h);
    dense_lane<NH, NH, NH, true>(s + LW2, s + LB2, h, buf, lane);
    load_column<NH>(buf, lane, h);
    dense_lane<NH, NH, NH, true>(s + LW3, s + LB3, h, buf, lane);
    load_column<NH>(buf, lane, h);
    dense_lane<NH, NH, NB, false>(s + LW4, s + LB4, h, buf, lane);
    load_column<NB>(buf, lane, z);
}

__device__ __forceinline__ void softmax32(const float* z, float* p)
{
    float m = z[0];
#pragma unroll
    for (int b = 1; b < NB; ++b) m = z[b] > m ? z[b] : m;
    float S = 0.f;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        p[b] = prop_exp(z[b] - m);
        S = S + p[b];
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) p[b] = p[b] / S;
}

struct PropArgs {
    const PNode* nodes;
    const float* O;
    const float* D;
    int64_t R;
    const int32_t* table;
    int depth;
    const float* coarse;
    const float* params;
    float mix, kl_threshold;
    int jitter_mode;
    uint64_t jitter_seed;
    // features mode
    float* features;
    float* chord;
    // sample mode
    int32_t* fallback;
    int32_t* n_learned;
    int32_t* lnode;
    float* ltmin;
    float* ltmax;
    float* kl;
};

__global__ __launch_bounds__(256) void leaf_table_kernel(const PNode* __restrict__ nodes, int depth,
                                                         int32_t* __restrict__ table)
{
    const int res = 1 << depth;
    const int64_t n = (int64_t)res * res * res;
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const PNode root = nodes[0];
    const int ci[3] = {(int)(idx % res), (int)((idx / res) % res), (int)(idx / ((int64_t)res * res))};
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float cell = (root.mx[k] - root.mn[k]) / (float)res;
        c[k] = root.mn[k] + ((float)ci[k] + 0.5f) * cell;
    }
    int32_t node = 0;
    while (true) {
        const PNode nd = nodes[node];
        if (nd.flags & kFlagActive) break;
        if (nd.first_child < 0) {
            node = -1;
            break;
        }
        // child bit b splits axis b (octree) or axes x, z (quad tree)
        const bool quad = (nd.flags & kFlagQuad) != 0;
        int sel = 0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            if (quad && b == 2) break;
            const int k = quad && b == 1 ? 2 : b;
            float mid = (nd.mn[k] + nd.mx[k]) * 0.5f;
            if (c[k] >= mid) sel |= 1 << b;
        }
        node = nd.first_child + sel;
    }
    table[idx] = node;
}

__global__ __launch_bounds__(256) void features_kernel(PropArgs a)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= a.R) return;
    const PNode root = a.nodes[0];
    float o[3], d[3], x[NI], t0, t1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o[k] = a.O[r * 3 + k];
        d[k] = a.D[r * 3 + k];
    }
    ray_features(root, a.table, a.depth, a.coarse, o, d, x, t0, t1);
#pragma unroll
    for (int k = 0; k < NI; ++k) a.features[r * NI + k] = x[k];
    if (a.chord) {
        a.chord[r * 2] = t0;
        a.chord[r * 2 + 1] = t1;
    }
}

// One ray per lane: features, MLP, pdf, KL gate, systematic allocation and
// the one-sample intervals of a learned ray.
// The mixed pdf pm[32] of ray r over its root chord [t0, t1] and the gate's
// KL(coarse || pdf); false (kl = 0) when the ray has no root chord.  Every
// lane of the wave runs the network (its activation columns share `buf`).
__device__ __forceinline__ bool ray_pdf(const PropArgs& a, const float* __restrict__ params, float* buf, int lane,
                                        const PNode& root, const float* o, const float* d, float* pm, float& t0,
                                        float& t1, float& kl)
{
    float x[NI];
    const bool ok = ray_features(root, a.table, a.depth, a.coarse, o, d, x, t0, t1);
    float z[NB];
    mlp_forward_s(params, buf, lane, x, z);
    kl = 0.f;
    if (!(t1 > t0) || !ok) return false;
    softmax32(z, pm);
    const float a1 = 1.0f - a.mix;
    const float a0 = a.mix * (1.0f / (float)NB);
#pragma unroll
    for (int b = 0; b < NB; ++b) pm[b] = pm[b] * a1 + a0;
    float Q = 0.f;
#pragma unroll
    for (int b = 0; b < NB; ++b) Q = Q + x[6 + b];
    if (Q > 0.f) {
        bool inf = false;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const float occ = x[6 + b];
            if (!inf && occ > 0.f) {
                if (!(pm[b] > 0.f)) {
                    inf = true;
                } else {
                    const float q = occ / Q;
                    kl = kl + q * prop_log(q / pm[b]);
                }
            }
        }
        if (inf) kl = INFINITY;
    }
    return true;
}

__global__ __launch_bounds__(kThreads) void sample_kernel(PropArgs a, const float* __restrict__ params)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float* buf = sm + (threadIdx.x >> 6) * kActFloats;
    const int lane = threadIdx.x & 63;
    const int64_t r = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (r >= a.R) return;  // no barrier below
    const PNode root = a.nodes[0];
    float o[3], d[3], t0, t1, kl;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o[k] = a.O[r * 3 + k];
        d[k] = a.D[r * 3 + k];
    }
    float pm[NB];
    const bool ok = ray_pdf(a, params, buf, lane, root, o, d, pm, t0, t1, kl);
    a.fallback[r] = 0;
    a.n_learned[r] = 0;
    if (a.kl) a.kl[r] = 0.f;
    if (!ok) return;
    if (a.kl) a.kl[r] = kl;
    if (kl > a.kl_threshold) {
        a.fallback[r] = 1;
        return;
    }
    // cells between the inverse-CDF images of u = i / 32, one sample each
    const float len = (t1 - t0) / (float)NB;
    float c = 0.f, prev = t0;
    int i = 1, emitted = 0;
#pragma unroll
    for (int b = 0; b <= NB; ++b) {
        const bool last = (b == NB);
        float c_hi = 0.f, lo = 0.f, hi = 0.f, pb = 1.f;
        if (!last) {
            c_hi = (b == NB - 1) ? 2.0f : c + pm[b];
            lo = bin_edge(t0, t1, len, b);
            hi = bin_edge(t0, t1, len, b + 1);
            pb = pm[b];
        }
        while (true) {
            float T;
            if (last) {
                if (i > NB) break;
                T = t1;
            } else {
                if (!(i < NB && (float)i * (1.0f / (float)NB) < c_hi)) break;
                const float u = (float)i * (1.0f / (float)NB);
                float frac = (u - c) / pb;
                if (!(frac >= 0.0f)) frac = 0.0f;
                if (!(frac <= 1.0f)) frac = 1.0f;
                T = lo + frac * (hi - lo);
            }
            ++i;
            if (T > prev) {
                const float aj = (a.jitter_mode == 2) ? hash_jitter(a.jitter_seed, r, emitted) : 0.5f;
                const float ts = sample_t(prev, (T - prev) / 1.0f, 0, aj);
                float pos[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) pos[k] = o[k] + d[k] * ts;
                const int32_t nd = locate(root, a.table, a.depth, pos);
                if (nd >= 0) {
                    const int64_t q = (int64_t)emitted * a.R + r;
                    a.lnode[q] = nd;
                    a.ltmin[q] = prev;
                    a.ltmax[q] = T;
                    ++emitted;
                }
                prev = T;
            }
        }
        if (!last) c = c + pm[b];
    }
    a.n_learned[r] = emitted;
}

// Learned allocation over the stratified node intervals of a ray list: for a
// ray the gate keeps, interval i gets its stratified count (at least one) plus
// round(E F(tmax_i)) - round(E F(tmin_i)) samples (the extras telescope to E
// over intervals that tile the chord), F the CDF of the mixed pdf over the root
// chord, linear inside each of the 32 bins and clamped to [0, 1],
// round(x) = floor(x + 1/2); `first` becomes the running sum of the ray's
// counts.  Fallback rays and rays without a root chord keep the stratified
// counts.  The CDF and the pdf sit in the lane's own column of the wave's
// activation buffer (rows 0..31: F at the bin edges, 32..63: pm), read at the
// bin of each t.
__global__ __launch_bounds__(kThreads) void alloc_kernel(PropArgs a, const float* __restrict__ params, int extra,
                                                        const int64_t* __restrict__ iv_off,
                                                        const float* __restrict__ iv_tmin,
                                                        const float* __restrict__ iv_tmax,
                                                        int32_t* __restrict__ iv_count,
                                                        int32_t* __restrict__ iv_first, int32_t* __restrict__ ray_ns)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float* buf = sm + (threadIdx.x >> 6) * kActFloats;
    const int lane = threadIdx.x & 63;
    const int64_t r = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    const int64_t r0 = r - lane;  // the wave's first ray
    if (r0 >= a.R) return;        // whole waves: the interval pass below needs every lane
    const bool valid = r < a.R;
    float t0 = 0.f, t1 = 0.f;
    bool learn = false;
    if (valid) {
        const PNode root = a.nodes[0];
        float o[3], d[3], kl;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            o[k] = a.O[r * 3 + k];
            d[k] = a.D[r * 3 + k];
        }
        float pm[NB];
        const bool ok = ray_pdf(a, params, buf, lane, root, o, d, pm, t0, t1, kl);
        const bool fall = ok && kl > a.kl_threshold;
        a.fallback[r] = fall ? 1 : 0;
        if (a.kl) a.kl[r] = kl;
        // a gated ray allocates along its mixed pdf, a fallback ray along the
        // uniform one (stratified sampling at the same budget of extras); the
        // CDF table goes to the ray's column of the wave's buffer
        learn = ok;
        if (learn) {
            float c = 0.f;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float p = fall ? 1.f / (float)NB : pm[b];
                buf[b * 64 + lane] = c;
                buf[(NB + b) * 64 + lane] = p;
                c = c + p;
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // The interval pass: the wave's 64 rays own one contiguous range of the
    // interval list; lane l takes the l-th of each 64 intervals (coalesced),
    // finds its ray L by a binary search over the rays' first intervals, and
    // adds ray L's extras round(E F(tmax)) - round(E F(tmin)) with ray L's
    // CDF column.  first = the exclusive sum of the counts along the ray: a
    // segmented scan over the 64 intervals, carried into the next 64 for the
    // ray that continues.
    const float E = (float)extra;
    const int64_t h0 = iv_off[valid ? r : a.R];
    if (valid && ray_ns && iv_off[r + 1] == h0) ray_ns[r] = 0;
    const int64_t wbeg = __shfl(h0, 0, 64);
    const int64_t wend = iv_off[r0 + 64 < a.R ? r0 + 64 : a.R];
    int carry = 0;
    for (int64_t q0 = wbeg; q0 < wend; q0 += 64) {
        const int64_t q = q0 + lane;
        int L = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (__shfl(h0, L + step, 64) <= q) L += step;
        const int64_t hL = __shfl(h0, L, 64);
        // end of ray L's intervals (every lane takes part in the shuffle: a
        // lane reading from an inactive lane would get 0)
        const int64_t hN = __shfl(h0, L < 63 ? L + 1 : 63, 64);
        const int64_t hE = L < 63 ? hN : wend;
        const float t0L = __shfl(t0, L, 64), t1L = __shfl(t1, L, 64);
        const int learnL = __shfl(learn ? 1 : 0, L, 64);
        const bool act = q < wend;
        int n = 0;
        if (act) {
            n = iv_count[q];
            if (learnL) {
                const float len = (t1L - t0L) / (float)NB;
                auto cdf = [&](float t) {
                    int b = (int)floorf((t - t0L) / len);
                    b = b < 0 ? 0 : (b > NB - 1 ? NB - 1 : b);
                    float F = buf[b * 64 + L] + buf[(NB + b) * 64 + L] * ((t - bin_edge(t0L, t1L, len, b)) / len);
                    if (!(F >= 0.f)) F = 0.f;
                    if (!(F <= 1.f)) F = 1.f;
                    return F;
                };
                const int ka = (int)floorf(E * cdf(iv_tmin[q]) + 0.5f);
                const int kb = (int)floorf(E * cdf(iv_tmax[q]) + 0.5f);
                n += kb > ka ? kb - ka : 0;
                iv_count[q] = n;
            }
        }
        int x = n, f = (act && q == hL) ? 1 : 0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int up = __shfl_up(x, off, 64), fu = __shfl_up(f, off, 64);
            if (lane >= off) {
                if (!f) x += up;
                f |= fu;
            }
        }
        if (!f) x += carry;  // the ray continuing from the previous 64 intervals
        if (act) {
            iv_first[q] = x - n;
            if (ray_ns && q == hE - 1) ray_ns[r0 + L] = x;
        }
        carry = __shfl(x, 63, 64);
    }
}

__global__ __launch_bounds__(256) void merge_counts_kernel(int64_t R, const int32_t* __restrict__ fallback,
                                                           const int32_t* __restrict__ n_learned,
                                                           int32_t* __restrict__ niv, int32_t* __restrict__ ns)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= R || fallback[r]) return;
    niv[r] = n_learned[r];
    ns[r] = n_learned[r];
}

// Copies the learned cells of 64 rays from the slot-major staging buffers
// into the mixed interval list: coalesced reads along rays into LDS, then
// coalesced writes along each ray's contiguous run of intervals.
constexpr int kWriteRays = 64;

__global__ __launch_bounds__(256) void proposal_write_kernel(int64_t R, const int32_t* __restrict__ fallback,
                                                             const int32_t* __restrict__ n_learned,
                                                             const int32_t* __restrict__ lnode,
                                                             const float* __restrict__ ltmin,
                                                             const float* __restrict__ ltmax,
                                                             const int64_t* __restrict__ iv_off, atdf_intervals iv)
{
    __shared__ int32_t s_node[NB * kWriteRays];
    __shared__ float s_lo[NB * kWriteRays];
    __shared__ float s_hi[NB * kWriteRays];
    __shared__ int32_t s_n[kWriteRays];
    __shared__ int64_t s_off[kWriteRays];
    const int64_t r0 = blockIdx.x * (int64_t)kWriteRays;
    for (int t = threadIdx.x; t < NB * kWriteRays; t += blockDim.x) {
        const int i = t / kWriteRays, rl = t - i * kWriteRays;
        const int64_t r = r0 + rl;
        if (r < R) {
            const int64_t q = (int64_t)i * R + r;
            s_node[rl * NB + i] = lnode[q];
            s_lo[rl * NB + i] = ltmin[q];
            s_hi[rl * NB + i] = ltmax[q];
        }
    }
    for (int rl = threadIdx.x; rl < kWriteRays; rl += blockDim.x) {
        const int64_t r = r0 + rl;
        s_n[rl] = (r < R && !fallback[r]) ? n_learned[r] : 0;
        s_off[rl] = r < R ? iv_off[r] : 0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NB * kWriteRays; t += blockDim.x) {
        const int rl = t / NB, i = t - rl * NB;
        if (i >= s_n[rl]) continue;
        const int64_t q = s_off[rl] + i;
        iv.node[q] = s_node[t];
        iv.tmin[q] = s_lo[t];
        iv.tmax[q] = s_hi[t];
        iv.count[q] = 1;
        iv.first[q] = i;
        iv.ray[q] = (int32_t)(r0 + rl);
    }
}

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------
// Workspace, per ray columns ([vector][R] so lanes write consecutive
// addresses): X 38 | H1 64 | H2 64 | H3 64 | DZ 32 | G3 64 | G2 64 | G1 64 | CE 1
constexpr int kColX = 0, kColH1 = kColX + NI, kColH2 = kColH1 + NH, kColH3 = kColH2 + NH, kColDZ = kColH3 + NH,
              kColG3 = kColDZ + NB, kColG2 = kColG3 + NH, kColG1 = kColG2 + NH, kColCE = kColG1 + NH,
              kCols = kColCE + 1;
constexpr int kChunk = 64;          // rays per partial-sum chunk
constexpr int kChunkStride = kChunk + 1;

// acc[k] = sum_o W[o][k] dy[o] (o ascending), dy in this lane's buffer column
template <int NOUT, int NIN, int STRIDE>
__device__ __forceinline__ void dense_back_lane(const float* W, const float* buf, int lane, float* acc)
{
#pragma unroll
    for (int k = 0; k < NIN; ++k) acc[k] = 0.f;
    for (int o = 0; o < NOUT; ++o) {
        const float dy = buf[o * 64 + lane];
        const float* wr = W + o * STRIDE;
#pragma unroll
        for (int k = 0; k + 4 <= NIN; k += 4) {
            const float4 w = *reinterpret_cast<const float4*>(wr + k);
            acc[k] = acc[k] + w.x * dy;
            acc[k + 1] = acc[k + 1] + w.y * dy;
            acc[k + 2] = acc[k + 2] + w.z * dy;
            acc[k + 3] = acc[k + 3] + w.w * dy;
        }
    }
}

struct GradArgs {
    const float* params;
    const float* features;
    const float* bins;
    int64_t R;
    float power;
    float* cols;  // [kCols][R]
};

// One ray per lane: forward, softmax cross-entropy, backward to the hidden
// layers; writes the columns the weight-gradient pass reads.
__global__ __launch_bounds__(kThreads) void grad_rays_kernel(GradArgs a)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float* s = sm;
    float* buf = sm + LTOTAL + (threadIdx.x >> 6) * kActFloats;
    load_params(a.params, s);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t r = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    if (r >= a.R) return;  // no barrier below
    const int64_t R = a.R;
    float* col = a.cols;
    float x[NI], h[NH];
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        x[k] = a.features[r * NI + k];
        col[(kColX + k) * R + r] = x[k];
    }
    dense_lane<NI, kW1S, NH, true>(s + LW1, s + LB1, x, buf, lane);
    load_column<NH>(buf, lane, h);
#pragma unroll
    for (int k = 0; k < NH; ++k) col[(kColH1 + k) * R + r] = h[k];
    dense_lane<NH, NH, NH, true>(s + LW2, s + LB2, h, buf, lane);
    load_column<NH>(buf, lane, h);
#pragma unroll
    for (int k = 0; k < NH; ++k) col[(kColH2 + k) * R + r] = h[k];
    dense_lane<NH, NH, NH, true>(s + LW3, s + LB3, h, buf, lane);
    load_column<NH>(buf, lane, h);
#pragma unroll
    for (int k = 0; k < NH; ++k) col[(kColH3 + k) * R + r] = h[k];
    float z[NB];
    dense_lane<NH, NH, NB, false>(s + LW4, s + LB4, h, buf, lane);
    load_column<NB>(buf, lane, z);

    // softmax cross-entropy against the normalised target masses
    float q[NB];
    float Q = 0.f;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        float v = a.bins[r * NB + b];
        q[b] = v > 0.f ? (a.power == 1.0f ? v : powf(v, a.power)) : 0.f;
        Q = Q + q[b];
    }
    float p[NB];
    float m = z[0];
#pragma unroll
    for (int b = 1; b < NB; ++b) m = z[b] > m ? z[b] : m;
    float S = 0.f;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        p[b] = prop_exp(z[b] - m);
        S = S + p[b];
    }
    const float lse = m + prop_log(S);
    const bool live = Q > 0.f;
    const float invR = 1.0f / (float)R;
    float ce = 0.f;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const float qb = live ? q[b] / Q : 0.f;
        ce = ce - qb * (z[b] - lse);
        const float dz = live ? (p[b] / S - qb) * invR : 0.f;
        col[(kColDZ + b) * R + r] = dz;
        buf[b * 64 + lane] = dz;
    }
    col[kColCE * R + r] = ce;

    // backward through the hidden layers (ReLU masks from the stored columns)
    float g[NH];
    dense_back_lane<NB, NH, NH>(s + LW4, buf, lane, g);
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        g[k] = col[(kColH3 + k) * R + r] > 0.f ? g[k] : 0.f;
        col[(kColG3 + k) * R + r] = g[k];
        buf[k * 64 + lane] = g[k];
    }
    dense_back_lane<NH, NH, NH>(s + LW3, buf, lane, g);
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        g[k] = col[(kColH2 + k) * R + r] > 0.f ? g[k] : 0.f;
        col[(kColG2 + k) * R + r] = g[k];
        buf[k * 64 + lane] = g[k];
    }
    dense_back_lane<NH, NH, NH>(s + LW2, buf, lane, g);
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        g[k] = col[(kColH1 + k) * R + r] > 0.f ? g[k] : 0.f;
        col[(kColG1 + k) * R + r] = g[k];
    }
}

// Partial weight gradients of one chunk of kChunk rays for one layer
// (blockIdx.y): dW[o][k] = sum_r D[o][r] A[k][r], db[o] = sum_r D[o][r], rays
// in ascending order.  Layer 0 also sums the cross-entropy.
__global__ __launch_bounds__(256) void grad_partials_kernel(const float* __restrict__ cols, int64_t R,
                                                            float* __restrict__ partials)
{
    __shared__ float sD[NH * kChunkStride];
    __shared__ float sA[NH * kChunkStride];
    const int layer = blockIdx.y;
    const int64_t c = blockIdx.x;
    const int64_t r0 = c * kChunk;
    const int nr = (int)((R - r0) < kChunk ? (R - r0) : kChunk);
    int nin, nout, colA, colD, ow, ob;
    if (layer == 0) {
        nin = NI; nout = NH; colA = kColX; colD = kColG1; ow = OW1; ob = OB1;
    } else if (layer == 1) {
        nin = NH; nout = NH; colA = kColH1; colD = kColG2; ow = OW2; ob = OB2;
    } else if (layer == 2) {
        nin = NH; nout = NH; colA = kColH2; colD = kColG3; ow = OW3; ob = OB3;
    } else {
        nin = NH; nout = NB; colA = kColH3; colD = kColDZ; ow = OW4; ob = OB4;
    }
    for (int i = threadIdx.x; i < nout * kChunk; i += blockDim.x) {
        const int o = i / kChunk, rr = i - o * kChunk;
        sD[o * kChunkStride + rr] = rr < nr ? cols[(int64_t)(colD + o) * R + r0 + rr] : 0.f;
    }
    for (int i = threadIdx.x; i < nin * kChunk; i += blockDim.x) {
        const int k = i / kChunk, rr = i - k * kChunk;
        sA[k * kChunkStride + rr] = rr < nr ? cols[(int64_t)(colA + k) * R + r0 + rr] : 0.f;
    }
    __syncthreads();
    float* out = partials + c * (int64_t)(NPARAM + 1);
    for (int i = threadIdx.x; i < nout * nin; i += blockDim.x) {
        const int o = i / nin, k = i - o * nin;
        const float* dr = sD + o * kChunkStride;
        const float* ar = sA + k * kChunkStride;
        float acc = 0.f;
        for (int rr = 0; rr < kChunk; ++rr) acc = acc + dr[rr] * ar[rr];
        out[ow + i] = acc;
    }
    for (int o = threadIdx.x; o < nout; o += blockDim.x) {
        const float* dr = sD + o * kChunkStride;
        float acc = 0.f;
        for (int rr = 0; rr < kChunk; ++rr) acc = acc + dr[rr];
        out[ob + o] = acc;
    }
    if (layer == 0 && threadIdx.x == 0) {
        float acc = 0.f;
        for (int rr = 0; rr < nr; ++rr) acc = acc + cols[(int64_t)kColCE * R + r0 + rr];
        out[NPARAM] = acc;
    }
}

// grad[p] = sum over chunks (ascending, FP64) of the partials; loss likewise / R.
// A workgroup owns 16 entries; its 16 row groups (chunks c = g, g + 16, ...)
// sum in FP64 registers and the group sums are added in group order, so the
// order is fixed and 16 times more lanes share the chunk loop than with one
// lane per entry.
constexpr int kRedEntries = 16;

__global__ __launch_bounds__(256) void grad_reduce_kernel(const float* __restrict__ partials, int64_t nchunks,
                                                          int64_t R, float* __restrict__ grad,
                                                          double* __restrict__ loss)
{
    __shared__ double part[16][kRedEntries + 1];
    const int pl = threadIdx.x % kRedEntries, rg = threadIdx.x / kRedEntries;
    const int p = blockIdx.x * kRedEntries + pl;
    double acc = 0.0;
    if (p <= NPARAM)
        for (int64_t c = rg; c < nchunks; c += 16) acc += (double)partials[c * (NPARAM + 1) + p];
    part[rg][pl] = acc;
    __syncthreads();
    if (rg == 0 && p <= NPARAM) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) t += part[k][pl];
        if (p < NPARAM)
            grad[p] = (float)t;
        else
            loss[0] = t / (double)R;
    }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int max_lds_bytes()
{
    static int v = 0;
    if (v == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
            (void)hipGetLastError();
            v = 65536;
        }
    }
    return v;
}

constexpr size_t kMlpLds = sizeof(float) * ((size_t)LTOTAL + (size_t)kWaves * kActFloats);
constexpr size_t kSampleLds = sizeof(float) * (size_t)kWaves * kActFloats;

atdf_status check_proposal(const void* packed_tree, int64_t num_nodes, const atdf_proposal* p, const char* who)
{
    ATDF_REQUIRE(packed_tree && num_nodes > 0, "%s: empty tree", who);
    ATDF_REQUIRE(p && p->leaf_table && p->coarse && p->params, "%s: incomplete proposal", who);
    ATDF_REQUIRE(p->table_depth >= 1 && p->table_depth <= 8, "%s: table_depth must be in 1..8", who);
    ATDF_REQUIRE(p->mix >= 0.f && p->mix <= 1.f, "%s: mix must be in [0, 1]", who);
    ATDF_REQUIRE(p->jitter_mode == 0 || p->jitter_mode == 2, "%s: jitter_mode must be 0 or 2", who);
    return ATDF_OK;
}

PropArgs make_args(const void* packed_tree, const float* O, const float* D, int64_t R, const atdf_proposal* p)
{
    PropArgs a{};
    a.nodes = reinterpret_cast<const PNode*>(packed_tree);
    a.O = O;
    a.D = D;
    a.R = R;
    a.table = p->leaf_table;
    a.depth = p->table_depth;
    a.coarse = p->coarse;
    a.params = p->params;
    a.mix = p->mix;
    a.kl_threshold = p->kl_threshold;
    a.jitter_mode = p->jitter_mode;
    a.jitter_seed = p->jitter_seed;
    return a;
}

}  // namespace
}  // namespace atdf

using namespace atdf;

extern "C" {

int64_t atdf_proposal_num_params(void) { return NPARAM; }

int64_t atdf_leaf_table_numel(int32_t depth)
{
    if (depth < 1 || depth > 8) return 0;
    const int64_t res = (int64_t)1 << depth;
    return res * res * res;
}

atdf_status atdf_leaf_table(const void* packed_tree, int64_t num_nodes, int32_t depth, int32_t* table,
                            atdf_stream_t stream)
{
    ATDF_REQUIRE(packed_tree && num_nodes > 0 && table, "atdf_leaf_table: null argument");
    ATDF_REQUIRE(depth >= 1 && depth <= 8, "atdf_leaf_table: depth must be in 1..8 (got %d)", depth);
    const int64_t n = atdf_leaf_table_numel(depth);
    hipLaunchKernelGGL(leaf_table_kernel, dim3((unsigned)div_up(n, 256)), dim3(256), 0, to_hip(stream),
                       reinterpret_cast<const PNode*>(packed_tree), depth, table);
    return check_launch("atdf_leaf_table");
}

atdf_status atdf_proposal_features(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                   const float* ray_direction, int64_t num_rays, const atdf_proposal* proposal,
                                   float* features, float* chord, atdf_stream_t stream)
{
    atdf_status st = check_proposal(packed_tree, num_nodes, proposal, "atdf_proposal_features");
    if (st) return st;
    ATDF_REQUIRE(num_rays >= 0, "atdf_proposal_features: negative ray count");
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(ray_origin && ray_direction && features, "atdf_proposal_features: null buffer");
    PropArgs a = make_args(packed_tree, ray_origin, ray_direction, num_rays, proposal);
    a.features = features;
    a.chord = chord;
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)div_up(num_rays, 256)), dim3(256), 0, to_hip(stream), a);
    return check_launch("atdf_proposal_features");
}

atdf_status atdf_proposal_sample(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                 const float* ray_direction, int64_t num_rays, const atdf_proposal* proposal,
                                 int32_t* ray_fallback, int32_t* ray_num_learned, int32_t* learned_node,
                                 float* learned_tmin, float* learned_tmax, float* ray_kl, atdf_stream_t stream)
{
    atdf_status st = check_proposal(packed_tree, num_nodes, proposal, "atdf_proposal_sample");
    if (st) return st;
    ATDF_REQUIRE(num_rays >= 0, "atdf_proposal_sample: negative ray count");
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(ray_origin && ray_direction && ray_fallback && ray_num_learned && learned_node && learned_tmin &&
                     learned_tmax,
                 "atdf_proposal_sample: null buffer");
    ATDF_REQUIRE((int)kSampleLds <= max_lds_bytes(), "atdf_proposal_sample: needs %d bytes of LDS per workgroup",
                 (int)kSampleLds);
    PropArgs a = make_args(packed_tree, ray_origin, ray_direction, num_rays, proposal);
    a.fallback = ray_fallback;
    a.n_learned = ray_num_learned;
    a.lnode = learned_node;
    a.ltmin = learned_tmin;
    a.ltmax = learned_tmax;
    a.kl = ray_kl;
    hipLaunchKernelGGL(sample_kernel, dim3((unsigned)div_up(num_rays, kThreads)), dim3(kThreads), kSampleLds,
                       to_hip(stream), a, proposal->params);
    return check_launch("atdf_proposal_sample");
}

atdf_status atdf_proposal_allocate(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                   const float* ray_direction, int64_t num_rays, const atdf_proposal* proposal,
                                   int32_t extra_samples, const int64_t* ray_interval_offset,
                                   const atdf_intervals* intervals, int32_t* ray_fallback, int32_t* ray_num_samples,
                                   float* ray_kl, atdf_stream_t stream)
{
    atdf_status st = check_proposal(packed_tree, num_nodes, proposal, "atdf_proposal_allocate");
    if (st) return st;
    ATDF_REQUIRE(num_rays >= 0, "atdf_proposal_allocate: negative ray count");
    ATDF_REQUIRE(extra_samples >= 0 && extra_samples <= (1 << 20), "atdf_proposal_allocate: extra_samples out of range");
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(ray_origin && ray_direction && ray_interval_offset && intervals && ray_fallback,
                 "atdf_proposal_allocate: null buffer");
    ATDF_REQUIRE(intervals->tmin && intervals->tmax && intervals->count && intervals->first,
                 "atdf_proposal_allocate: incomplete interval buffers");
    ATDF_REQUIRE((int)kSampleLds <= max_lds_bytes(), "atdf_proposal_allocate: needs %d bytes of LDS per workgroup",
                 (int)kSampleLds);
    PropArgs a = make_args(packed_tree, ray_origin, ray_direction, num_rays, proposal);
    a.fallback = ray_fallback;
    a.kl = ray_kl;
    hipLaunchKernelGGL(alloc_kernel, dim3((unsigned)div_up(num_rays, kThreads)), dim3(kThreads), kSampleLds,
                       to_hip(stream), a, proposal->params, (int)extra_samples, ray_interval_offset, intervals->tmin,
                       intervals->tmax, intervals->count, intervals->first, ray_num_samples);
    return check_launch("atdf_proposal_allocate");
}

atdf_status atdf_proposal_merge_counts(int64_t num_rays, const int32_t* ray_fallback,
                                       const int32_t* ray_num_learned, int32_t* ray_num_intervals,
                                       int32_t* ray_num_samples, atdf_stream_t stream)
{
    ATDF_REQUIRE(num_rays >= 0, "atdf_proposal_merge_counts: negative ray count");
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(ray_fallback && ray_num_learned && ray_num_intervals && ray_num_samples,
                 "atdf_proposal_merge_counts: null buffer");
    hipLaunchKernelGGL(merge_counts_kernel, dim3((unsigned)div_up(num_rays, 256)), dim3(256), 0, to_hip(stream),
                       num_rays, ray_fallback, ray_num_learned, ray_num_intervals, ray_num_samples);
    return check_launch("atdf_proposal_merge_counts");
}

atdf_status atdf_proposal_write(int64_t num_rays, const int32_t* ray_fallback, const int32_t* ray_num_learned,
                                const int32_t* learned_node, const float* learned_tmin,
                                const float* learned_tmax, const int64_t* ray_interval_offset,
                                const atdf_intervals* intervals, atdf_stream_t stream)
{
    ATDF_REQUIRE(num_rays >= 0, "atdf_proposal_write: negative ray count");
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(ray_fallback && ray_num_learned && learned_node && learned_tmin && learned_tmax &&
                     ray_interval_offset && intervals,
                 "atdf_proposal_write: null buffer");
    ATDF_REQUIRE(intervals->node && intervals->tmin && intervals->tmax && intervals->count && intervals->first &&
                     intervals->ray,
                 "atdf_proposal_write: incomplete interval buffers");
    hipLaunchKernelGGL(proposal_write_kernel, dim3((unsigned)div_up(num_rays, kWriteRays)), dim3(256), 0,
                       to_hip(stream),
                       num_rays, ray_fallback, ray_num_learned, learned_node, learned_tmin, learned_tmax,
                       ray_interval_offset, *intervals);
    return check_launch("atdf_proposal_write");
}

size_t atdf_proposal_grad_workspace_bytes(int64_t num_rays)
{
    if (num_rays <= 0) return 0;
    const int64_t nchunks = div_up(num_rays, kChunk);
    return align256(sizeof(float) * (size_t)kCols * num_rays) + sizeof(float) * (size_t)nchunks * (NPARAM + 1);
}

atdf_status atdf_proposal_grad(const float* params, const float* features, const float* target_bins,
                               int64_t num_rays, float target_power, float* grad, double* loss, void* workspace,
                               size_t workspace_bytes, atdf_stream_t stream)
{
    ATDF_REQUIRE(target_power > 0.f && target_power <= 4.f, "atdf_proposal_grad: target_power must be in (0, 4]");
    ATDF_REQUIRE(num_rays > 0, "atdf_proposal_grad: needs at least one ray");
    ATDF_REQUIRE(params && features && target_bins && grad && loss && workspace, "atdf_proposal_grad: null buffer");
    ATDF_REQUIRE(workspace_bytes >= atdf_proposal_grad_workspace_bytes(num_rays),
                 "atdf_proposal_grad: workspace too small");
    ATDF_REQUIRE(num_rays <= ((int64_t)1 << 31) / kCols, "atdf_proposal_grad: too many rays");
    ATDF_REQUIRE((int)kMlpLds <= max_lds_bytes(), "atdf_proposal_grad: needs %d bytes of LDS per workgroup",
                 (int)kMlpLds);
    hipStream_t s = to_hip(stream);
    float* cols = reinterpret_cast<float*>(workspace);
    float* partials =
        reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(workspace) +
                                 align256(sizeof(float) * (size_t)kCols * num_rays));
    const int64_t nchunks = div_up(num_rays, kChunk);
    GradArgs a{params, features, target_bins, num_rays, target_power, cols};
    hipLaunchKernelGGL(grad_rays_kernel, dim3((unsigned)div_up(num_rays, kThreads)), dim3(kThreads), kMlpLds, s, a);
    hipLaunchKernelGGL(grad_partials_kernel, dim3((unsigned)nchunks, 4), dim3(256), 0, s, cols, num_rays, partials);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((unsigned)div_up(NPARAM + 1, kRedEntries)), dim3(256), 0, s, partials,
                       nchunks, num_rays, grad, loss);
    return check_launch("atdf_proposal_grad");
}

}  // extern "C"

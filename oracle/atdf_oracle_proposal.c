/*
 * atdf_oracle_proposal.c -- CPU oracle of the learned sample proposal.
 *
 * TEST INFRASTRUCTURE ONLY (see atdf_oracle.h).  The reference has no
 * learned proposal; this file defines the method the HIP kernels implement
 * (ray features -> MLP -> 32-bin pdf -> 32 equal-probability cells of the
 * chord, one sample each -> count-1 node intervals, KL gate against the
 * coarse occupancy) and its training loss.  The FP32 functions fix the rounding order the kernels
 * follow (compiled with -ffp-contract=off; the dense layers' dot products are
 * explicit fmaf chains, correctly rounded on both sides); atdfo_prop_grad is the FP64
 * restatement the training kernels are checked against.
 *
 * Reference anchors: the slab test of the root chord is
 * structure/HyperTree.h:338-380; the tree numbering (children contiguous,
 * bit k of the child index = upper half on axis k) is
 * structure/HyperTree.cpp:453-497; the fallback is the stratified sampler
 * of structure/HyperTree.cu:266-514.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "atdf_oracle.h"

#define NB ATDFO_PROP_BINS
#define NI ATDFO_PROP_IN
#define NH ATDFO_PROP_H

/* parameter offsets */
#define OW1 0
#define OB1 (OW1 + NH * NI)
#define OW2 (OB1 + NH)
#define OB2 (OW2 + NH * NH)
#define OW3 (OB2 + NH)
#define OB3 (OW3 + NH * NH)
#define OW4 (OB3 + NH)
#define OB4 (OW4 + NB * NH)
#define NPARAM (OB4 + NB)

int64_t atdfo_prop_num_params(void) { return NPARAM; }

/* e^x: x = n ln2 + r with ln2 split in a 15-bit head (n * head is exact for
 * |n| < 512) and a tail, |r| <= ln2 / 2, e^r by its degree-7 Taylor
 * polynomial in Horner form (truncation < 6e-9 relative), times 2^n. */
float atdfo_prop_exp(float x)
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

/* ln x for positive normal x: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m - 1) / (m + 1), series to s^9. */
float atdfo_prop_log(float x)
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

/* ------------------------------------------------------------------------ */
/* point location                                                            */
/* ------------------------------------------------------------------------ */
void atdfo_leaf_table_ns(const float* node_min, const float* node_max, const int32_t* children, int ns,
                         const int32_t* active, int depth, int32_t* table)
{
    const int res = 1 << depth;
    float cell[3];
    for (int k = 0; k < 3; ++k) cell[k] = (node_max[k] - node_min[k]) / (float)res;
    for (int64_t idx = 0; idx < (int64_t)res * res * res; ++idx) {
        int ci[3] = {(int)(idx % res), (int)((idx / res) % res), (int)(idx / ((int64_t)res * res))};
        float c[3];
        for (int k = 0; k < 3; ++k) c[k] = node_min[k] + ((float)ci[k] + 0.5f) * cell[k];
        int32_t node = 0;
        while (1) {
            if (active[node]) break;
            int32_t first = children[(int64_t)node * ns];
            if (first < 0) {
                node = -1;
                break;
            }
            /* child bit b splits axis b (octree) or axes x, z (quad tree) */
            int sel = 0;
            for (int b = 0; b < (ns == 4 ? 2 : 3); ++b) {
                int k = ns == 4 && b == 1 ? 2 : b;
                float mid = (node_min[(int64_t)node * 3 + k] + node_max[(int64_t)node * 3 + k]) * 0.5f;
                if (c[k] >= mid) sel |= 1 << b;
            }
            node = first + sel;
        }
        table[idx] = node;
    }
}

void atdfo_leaf_table(const float* node_min, const float* node_max, const int32_t* children,
                      const int32_t* active, int depth, int32_t* table)
{
    atdfo_leaf_table_ns(node_min, node_max, children, 8, active, depth, table);
}

int32_t atdfo_locate(const float* root_min, const float* root_max, const int32_t* table, int depth,
                     const float* p)
{
    const int res = 1 << depth;
    int ci[3];
    for (int k = 0; k < 3; ++k) {
        float u = (p[k] - root_min[k]) / (root_max[k] - root_min[k]);
        int c = (int)floorf(u * (float)res);
        ci[k] = c < 0 ? 0 : (c > res - 1 ? res - 1 : c);
    }
    return table[((int64_t)ci[2] * res + ci[1]) * res + ci[0]];
}

/* ------------------------------------------------------------------------ */
/* features                                                                  */
/* ------------------------------------------------------------------------ */
static int root_chord(const atdfo_proposal* p, const float* o, const float* d, float* t0, float* t1)
{
    int hit = atdfo_intersect_box_ray(p->root_min, p->root_max, o, d, t0, t1);
    if (!hit || *t1 - *t0 < 1e-5) {
        *t0 = 0;
        *t1 = 0;
        return 0;
    }
    return 1;
}

static float bin_edge(float t0, float t1, float len, int b)
{
    return b == NB ? t1 : t0 + (float)b * len;
}

static void ray_features(const atdfo_proposal* p, const float* o, const float* d, float* f, float* t0,
                         float* t1)
{
    for (int k = 0; k < 3; ++k) {
        f[k] = o[k];
        f[3 + k] = d[k];
    }
    int ok = root_chord(p, o, d, t0, t1);
    float len = (*t1 - *t0) / (float)NB;
    for (int b = 0; b < NB; ++b) {
        float occ = 0;
        if (ok) {
            float mid = (bin_edge(*t0, *t1, len, b) + bin_edge(*t0, *t1, len, b + 1)) * 0.5f;
            float pos[3];
            for (int k = 0; k < 3; ++k) pos[k] = o[k] + d[k] * mid;
            int32_t node = atdfo_locate(p->root_min, p->root_max, p->table, p->depth, pos);
            if (node >= 0) occ = p->coarse[node];
        }
        f[6 + b] = occ;
    }
}

void atdfo_prop_features(const atdfo_proposal* p, const float* o, const float* d, int64_t R, float* feat,
                         float* chord)
{
    for (int64_t r = 0; r < R; ++r) {
        float t0, t1;
        ray_features(p, o + r * 3, d + r * 3, feat + r * NI, &t0, &t1);
        if (chord) {
            chord[r * 2] = t0;
            chord[r * 2 + 1] = t1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* FP32 forward (the kernels' order)                                          */
/* ------------------------------------------------------------------------ */
static void dense_f(const float* W, const float* b, int nin, int nout, const float* x, float* y, int relu)
{
    for (int o = 0; o < nout; ++o) {
        float acc = 0;
        for (int k = 0; k < nin; ++k) acc = fmaf(W[o * nin + k], x[k], acc);  /* fused, as the kernels */
        acc = acc + b[o];
        y[o] = relu ? (acc > 0 ? acc : 0) : acc;
    }
}

static void forward_one(const float* P, const float* x, float* z, float* prob)
{
    float h1[NH], h2[NH], h3[NH];
    dense_f(P + OW1, P + OB1, NI, NH, x, h1, 1);
    dense_f(P + OW2, P + OB2, NH, NH, h1, h2, 1);
    dense_f(P + OW3, P + OB3, NH, NH, h2, h3, 1);
    dense_f(P + OW4, P + OB4, NH, NB, h3, z, 0);
    float m = z[0];
    for (int b = 1; b < NB; ++b) m = z[b] > m ? z[b] : m;
    float e[NB];
    float S = 0;
    for (int b = 0; b < NB; ++b) {
        e[b] = atdfo_prop_exp(z[b] - m);
        S = S + e[b];
    }
    for (int b = 0; b < NB; ++b) prob[b] = e[b] / S;
}

void atdfo_prop_forward(const float* params, const float* feat, int64_t R, float* logits, float* probs)
{
    for (int64_t r = 0; r < R; ++r) {
        float z[NB], pr[NB];
        forward_one(params, feat + r * NI, z, pr);
        for (int b = 0; b < NB; ++b) {
            if (logits) logits[r * NB + b] = z[b];
            if (probs) probs[r * NB + b] = pr[b];
        }
    }
}

/* Jitter of a one-sample interval: the fused kernels place sample s_local
 * of ray r at lo (1 - a) + (lo + (hi - lo)) a with a = 0.5 (mode 0) or
 * atdfo_jitter(seed, r, s_local) (mode 2). */
static float cell_sample_t(float lo, float hi, float a)
{
    float step = (hi - lo) / 1.0f;
    float t1 = lo + 0.0f * step;
    float t2 = lo + 1.0f * step;
    return t1 * (1 - a) + t2 * a;
}

/* The mixed pdf pm[32] of one ray over its root chord [t0, t1] and the
 * gate's KL(coarse || pdf); 0 (kl = 0) when the ray has no root chord. */
static int ray_pdf(const atdfo_proposal* p, const float* ro, const float* rd, float* pm, float* t0, float* t1,
                   float* kl_out)
{
    const float a1 = 1.0f - p->mix;
    const float a0 = p->mix * (1.0f / (float)NB);
    float f[NI];
    ray_features(p, ro, rd, f, t0, t1);
    *kl_out = 0;
    if (!(*t1 > *t0)) return 0;
    float z[NB], pr[NB];
    forward_one(p->params, f, z, pr);
    for (int b = 0; b < NB; ++b) pm[b] = pr[b] * a1 + a0;
    /* KL(coarse || pdf) */
    float Q = 0;
    for (int b = 0; b < NB; ++b) Q = Q + f[6 + b];
    float kl = 0;
    if (Q > 0) {
        for (int b = 0; b < NB; ++b) {
            float occ = f[6 + b];
            if (!(occ > 0)) continue;
            if (!(pm[b] > 0)) {
                kl = INFINITY;
                break;
            }
            float q = occ / Q;
            kl = kl + q * atdfo_prop_log(q / pm[b]);
        }
    }
    *kl_out = kl;
    return 1;
}

/* Learned allocation over a stratified interval list (atdf_proposal_allocate):
 * for a ray the gate keeps, interval i gets its stratified count plus
 * round(E F(tmax_i)) - round(E F(tmin_i)) samples, F(t) = cum[b] + pm[b] ((t - edge_b) / len) at
 * b = clamp(floor((t - t0) / len), 0, 31), clamped to [0, 1], cum the running
 * sum of pm, round(x) = floor(x + 1/2); first = the running sum of counts.
 * A fallback ray (KL above the threshold) does the same with pm = 1/32 in
 * every bin (the uniform pdf); rays without a root chord keep their counts
 * (first rewritten the same way). */
void atdfo_prop_allocate(const atdfo_proposal* p, const float* o, const float* d, int64_t R, int32_t extra,
                         const int64_t* iv_off, const float* tmin, const float* tmax, int32_t* count, int32_t* first,
                         int32_t* fallback, int32_t* ns, float* kl_out)
{
    for (int64_t r = 0; r < R; ++r) {
        float pm[NB], t0, t1, kl;
        int ok = ray_pdf(p, o + r * 3, d + r * 3, pm, &t0, &t1, &kl);
        int fall = ok && kl > p->kl_threshold;
        fallback[r] = fall;
        if (kl_out) kl_out[r] = kl;
        /* fallback rays allocate along the uniform pdf: stratified sampling
         * at the same budget of extras */
        int learn = ok;
        if (fall)
            for (int b = 0; b < NB; ++b) pm[b] = 1.0f / (float)NB;
        float cum[NB];
        float c = 0;
        for (int b = 0; b < NB; ++b) {
            cum[b] = c;
            c = c + pm[b];
        }
        float len = (t1 - t0) / (float)NB;
        float E = (float)extra;
        int s = 0;
        for (int64_t h = iv_off[r]; h < iv_off[r + 1]; ++h) {
            int n;
            if (learn) {
                float tt[2] = {tmin[h], tmax[h]};
                int k[2];
                for (int e = 0; e < 2; ++e) {
                    int b = (int)floorf((tt[e] - t0) / len);
                    b = b < 0 ? 0 : (b > NB - 1 ? NB - 1 : b);
                    float F = cum[b] + pm[b] * ((tt[e] - bin_edge(t0, t1, len, b)) / len);
                    if (!(F >= 0.f)) F = 0.f;
                    if (!(F <= 1.f)) F = 1.f;
                    k[e] = (int)floorf(E * F + 0.5f);
                }
                n = count[h] + (k[1] > k[0] ? k[1] - k[0] : 0);
                count[h] = n;
            } else {
                n = count[h];
            }
            first[h] = s;
            s += n;
        }
        if (ns) ns[r] = s;
    }
}

void atdfo_prop_sample(const atdfo_proposal* p, const float* o, const float* d, int64_t R, int32_t* fallback,
                       int32_t* n_learned, int32_t* node, float* tmin, float* tmax, float* kl_out)
{
    const float a1 = 1.0f - p->mix;
    const float a0 = p->mix * (1.0f / (float)NB);
    for (int64_t r = 0; r < R; ++r) {
        const float* ro = o + r * 3;
        const float* rd = d + r * 3;
        float f[NI], t0, t1;
        ray_features(p, ro, rd, f, &t0, &t1);
        fallback[r] = 0;
        n_learned[r] = 0;
        if (kl_out) kl_out[r] = 0;
        if (!(t1 > t0)) continue;
        float z[NB], pr[NB], pm[NB];
        forward_one(p->params, f, z, pr);
        for (int b = 0; b < NB; ++b) pm[b] = pr[b] * a1 + a0;
        /* KL(coarse || pdf) */
        float Q = 0;
        for (int b = 0; b < NB; ++b) Q = Q + f[6 + b];
        float kl = 0;
        if (Q > 0) {
            for (int b = 0; b < NB; ++b) {
                float occ = f[6 + b];
                if (!(occ > 0)) continue;
                if (!(pm[b] > 0)) {
                    kl = INFINITY;
                    break;
                }
                float q = occ / Q;
                kl = kl + q * atdfo_prop_log(q / pm[b]);
            }
        }
        if (kl_out) kl_out[r] = kl;
        if (kl > p->kl_threshold) {
            fallback[r] = 1;
            continue;
        }
        /* cells between the inverse-CDF images of u = i / 32: one sample each */
        float len = (t1 - t0) / (float)NB;
        float c = 0, prev = t0;
        int i = 1, emitted = 0;
        for (int b = 0; b <= NB; ++b) {
            float T;
            int last = (b == NB);
            float c_hi = 0;
            if (!last) c_hi = (b == NB - 1) ? 2.0f : c + pm[b];
            while (1) {
                if (last) {
                    if (i > NB) break;
                    T = t1;
                } else {
                    if (!(i < NB && (float)i * (1.0f / (float)NB) < c_hi)) break;
                    float u = (float)i * (1.0f / (float)NB);
                    float frac = (u - c) / pm[b];
                    if (!(frac >= 0.0f)) frac = 0.0f;
                    if (!(frac <= 1.0f)) frac = 1.0f;
                    float lo = bin_edge(t0, t1, len, b);
                    float hi = bin_edge(t0, t1, len, b + 1);
                    T = lo + frac * (hi - lo);
                }
                ++i;
                if (T > prev) {
                    float a = 0.5f;
                    if (p->jitter_mode == 2) a = atdfo_jitter(p->jitter_seed, r, emitted);
                    float ts = cell_sample_t(prev, T, a);
                    float pos[3];
                    for (int k = 0; k < 3; ++k) pos[k] = ro[k] + rd[k] * ts;
                    int32_t nd = atdfo_locate(p->root_min, p->root_max, p->table, p->depth, pos);
                    if (nd >= 0) {
                        node[(int64_t)emitted * R + r] = nd;
                        tmin[(int64_t)emitted * R + r] = prev;
                        tmax[(int64_t)emitted * R + r] = T;
                        ++emitted;
                    }
                    prev = T;
                }
            }
            if (!last) c = c + pm[b];
        }
        n_learned[r] = emitted;
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 training loss and gradient                                            */
/* ------------------------------------------------------------------------ */
static void dense_d(const float* W, const float* b, int nin, int nout, const double* x, double* y, int relu)
{
    for (int o = 0; o < nout; ++o) {
        double acc = 0;
        for (int k = 0; k < nin; ++k) acc += (double)W[o * nin + k] * x[k];
        acc += b[o];
        y[o] = relu ? (acc > 0 ? acc : 0) : acc;
    }
}

/* gW += dy x^T, gb += dy, dx = W^T dy masked by (xin > 0) when mask */
static void dense_back_d(const float* W, int nin, int nout, const double* x, const double* dy, double* gW,
                         double* gb, double* dx, const double* mask_src)
{
    for (int o = 0; o < nout; ++o) {
        gb[o] += dy[o];
        for (int k = 0; k < nin; ++k) gW[o * nin + k] += dy[o] * x[k];
    }
    if (!dx) return;
    for (int k = 0; k < nin; ++k) {
        double acc = 0;
        for (int o = 0; o < nout; ++o) acc += (double)W[o * nin + k] * dy[o];
        dx[k] = (mask_src && !(mask_src[k] > 0)) ? 0 : acc;
    }
}

double atdfo_prop_grad(const float* params, const float* feat, const float* bins, int64_t R, float power,
                       double* grad)
{
    const float* P = params;
    memset(grad, 0, sizeof(double) * NPARAM);
    double loss = 0;
    for (int64_t r = 0; r < R; ++r) {
        const float* bm = bins + r * NB;
        double qv[NB];
        double Q = 0;
        for (int b = 0; b < NB; ++b) {
            qv[b] = bm[b] > 0 ? pow((double)bm[b], (double)power) : 0;
            Q += qv[b];
        }
        if (!(Q > 0)) continue;
        double x[NI], h1[NH], h2[NH], h3[NH], z[NB];
        for (int k = 0; k < NI; ++k) x[k] = feat[r * NI + k];
        dense_d(P + OW1, P + OB1, NI, NH, x, h1, 1);
        dense_d(P + OW2, P + OB2, NH, NH, h1, h2, 1);
        dense_d(P + OW3, P + OB3, NH, NH, h2, h3, 1);
        dense_d(P + OW4, P + OB4, NH, NB, h3, z, 0);
        double m = z[0];
        for (int b = 1; b < NB; ++b) m = z[b] > m ? z[b] : m;
        double S = 0;
        for (int b = 0; b < NB; ++b) S += exp(z[b] - m);
        double lse = m + log(S);
        double dz[NB];
        for (int b = 0; b < NB; ++b) {
            double q = qv[b] / Q;
            double p = exp(z[b] - lse);
            loss += -q * (z[b] - lse);
            dz[b] = (p - q) / (double)R;
        }
        double g3[NH], g2[NH], g1[NH];
        dense_back_d(P + OW4, NH, NB, h3, dz, grad + OW4, grad + OB4, g3, h3);
        dense_back_d(P + OW3, NH, NH, h2, g3, grad + OW3, grad + OB3, g2, h2);
        dense_back_d(P + OW2, NH, NH, h1, g2, grad + OW2, grad + OB2, g1, h1);
        dense_back_d(P + OW1, NI, NH, x, g1, grad + OW1, grad + OB1, NULL, NULL);
    }
    return loss / (double)R;
}

/*
 * atdf_oracle.c -- CPU oracle (test infrastructure only; see atdf_oracle.h).
 * Built with -ffp-contract=off so every float expression below rounds in
 * the order written.
 */
#include "atdf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* tree: structure/HyperTree.cpp:298-525.  ns = 8: the octree; ns = 4: the   */
/* quad-tree representation (use_quad_tree_rep): children split x (bit 0)   */
/* and z (bit 1) at the centre, every node spans [y_min, y_max] in y         */
/* (set_node_quad, HyperTree.cpp:391-406, 461-464).                          */
/* ------------------------------------------------------------------------ */
int64_t atdfo_tree_num_nodes_ns(int max_depth, int ns)
{
    int64_t n = 0, p = 1;
    for (int i = 0; i <= max_depth; ++i) {
        n += p;
        p *= ns;
    }
    return n;
}

int64_t atdfo_tree_num_nodes(int max_depth) { return atdfo_tree_num_nodes_ns(max_depth, 8); }

void atdfo_tree_build_ns(int max_depth, int ns, float y_min, float y_max, float* node_min, float* node_max,
                         int32_t* children, int32_t* parent, int32_t* depth)
{
    const int quad = ns == 4;
    int64_t num_nodes = atdfo_tree_num_nodes_ns(max_depth, ns);
    for (int64_t i = 0; i < num_nodes * ns; ++i) children[i] = -1;
    int64_t cur = 0;
    for (int k = 0; k < 3; ++k) {
        node_min[k] = -1.0f;
        node_max[k] = 1.0f;
    }
    if (quad) {
        node_min[1] = y_min;
        node_max[1] = y_max;
    }
    parent[0] = -1;
    depth[0] = 0;
    cur = 1;
    int64_t start_layer = 0;
    for (int dpt = 1; dpt <= max_depth; ++dpt) {
        int64_t n = cur - start_layer;
        for (int64_t j = 0; j < n; ++j) {
            int64_t id = start_layer + j;
            float c[3];
            for (int k = 0; k < 3; ++k) c[k] = (node_min[id * 3 + k] + node_max[id * 3 + k]) * 0.5f;
            for (int i = 0; i < ns; ++i) {
                for (int k = 0; k < 3; ++k) {
                    node_min[cur * 3 + k] = node_min[id * 3 + k];
                    node_max[cur * 3 + k] = node_max[id * 3 + k];
                }
                for (int b = 0; b < (quad ? 2 : 3); ++b) {
                    int k = quad && b == 1 ? 2 : b; /* split axes: x, y, z or x, z */
                    if (((i >> b) & 1) == 0) {
                        node_min[cur * 3 + k] = node_min[id * 3 + k];
                        node_max[cur * 3 + k] = c[k];
                    } else {
                        node_min[cur * 3 + k] = c[k];
                        node_max[cur * 3 + k] = node_max[id * 3 + k];
                    }
                }
                if (quad) {
                    node_min[cur * 3 + 1] = y_min;
                    node_max[cur * 3 + 1] = y_max;
                }
                children[id * ns + i] = (int32_t)cur;
                parent[cur] = (int32_t)id;
                depth[cur] = dpt;
                cur++;
            }
        }
        start_layer = cur - n * ns;
    }
}

void atdfo_tree_build(int max_depth, float* node_min, float* node_max, int32_t* children, int32_t* parent,
                      int32_t* depth)
{
    atdfo_tree_build_ns(max_depth, 8, -1.f, 1.f, node_min, node_max, children, parent, depth);
}

/* ------------------------------------------------------------------------ */
/* slab test: structure/HyperTree.h:338-380                                  */
/* ------------------------------------------------------------------------ */
int atdfo_intersect_box_ray(const float* bmin, const float* bmax, const float* o, const float* d,
                            float* tmin, float* tmax)
{
    double t_near = -236436436;
    double t_far = 43637575;
    for (int i = 0; i < 3; i++) {
        if (d[i] == 0) {
            if ((o[i] < bmin[i]) || (o[i] >= bmax[i])) {
                *tmin = 0;
                *tmax = 0;
                return 0;
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
                *tmin = 0;
                *tmax = 0;
                return 0;
            }
        }
    }
    *tmin = (float)t_near;
    *tmax = (float)t_far;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* jitter: splitmix64 finaliser, 24-bit uniform in [0,1)                     */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

float atdfo_jitter(uint64_t seed, int64_t ray, int64_t local_index)
{
    uint64_t key = ((uint64_t)ray << 24) ^ (uint64_t)local_index;
    uint64_t h = splitmix64(seed ^ splitmix64(key));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

/* ------------------------------------------------------------------------ */
/* per-ray hit list: structure/HyperTree.cu:266-359                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    float tmin, tmax;
    int32_t node;
    int32_t n;
} hit_t;

static int hit_less(const hit_t* a, const hit_t* b)
{
    if (a->tmin < b->tmin) return 1;
    if (a->tmin > b->tmin) return 0;
    return a->node < b->node;
}

static int iceil_f(float v)
{
    int iv = (int)v;
    return (v > (float)iv) ? iv + 1 : iv;
}

/* Returns the number of hits (may exceed cap: then only cap are stored). */
static int ray_hits(const float* node_min, const float* node_max, const int64_t* active_ids,
                    int64_t num_active, const float* o, const float* d, int K, hit_t* hits, int cap)
{
    int nh = 0;
    for (int64_t a = 0; a < num_active; ++a) {
        int64_t id = active_ids[a];
        const float* bmin = node_min + id * 3;
        const float* bmax = node_max + id * 3;
        float tmin, tmax;
        int hit = atdfo_intersect_box_ray(bmin, bmax, o, d, &tmin, &tmax);
        if (tmax - tmin < 1e-5) continue;
        if (!hit) continue;
        if (nh < cap) {
            float dx = bmax[0] - bmin[0];
            float dy = bmax[1] - bmin[1];
            float dz = bmax[2] - bmin[2];
            float diag = sqrtf(dx * dx + dy * dy + dz * dz);
            float dis = tmax - tmin;
            float rel = dis / diag;
            hit_t h;
            h.tmin = tmin;
            h.tmax = tmax;
            h.node = (int32_t)id;
            h.n = iceil_f(rel * (float)K);
            /* insertion sort by (tmin, node) */
            int p = nh;
            while (p > 0 && hit_less(&h, &hits[p - 1])) {
                hits[p] = hits[p - 1];
                --p;
            }
            hits[p] = h;
        }
        nh++;
    }
    return nh;
}

/* t of sample j of an interval: HyperTree.cu:466-474 */
static float sample_t(float tmin, float step, int j, float a)
{
    float t1 = tmin + (float)j * step;
    float t2 = tmin + (float)(j + 1) * step;
    return t1 * (1 - a) + t2 * a;
}

static float jitter_value(int mode, const float* jitter, uint64_t seed, int64_t slot, int64_t ray,
                          int64_t local)
{
    if (mode == 1) return jitter[slot];
    if (mode == 2) return atdfo_jitter(seed, ray, local);
    return 0.5f;
}

/* midpoint-rule weight of sample j: HyperTree.cu:495-512 */
static float sample_weight(float tmin, float tmax, int n, int j, float t_prev, float t, float t_next)
{
    float lo = (j == 0) ? tmin : (t_prev + t) * 0.5f;
    float hi = (j == n - 1) ? tmax : (t_next + t) * 0.5f;
    return hi - lo;
}

#define ORACLE_HIT_CAP 4096

int64_t atdfo_traverse(const float* node_min, const float* node_max, const int64_t* active_ids,
                       int64_t num_active, const float* ray_o, const float* ray_d, int64_t num_rays,
                       int K, int max_intersections, int jitter_mode, const float* jitter, uint64_t seed,
                       int32_t* ray_num_samples, int32_t* ray_num_intervals, const int64_t* ray_sample_offset,
                       float* global_coordinate, float* weight, float* ray_t, int64_t* ray_index,
                       int64_t* local_index_in_ray, int64_t* node_id, const int64_t* ray_interval_offset,
                       int32_t* iv_node, float* iv_tmin, float* iv_tmax, int32_t* iv_count, int32_t* iv_first)
{
    hit_t* hits = (hit_t*)malloc(sizeof(hit_t) * ORACLE_HIT_CAP);
    int64_t total = 0;
    int overflow = 0;
    for (int64_t r = 0; r < num_rays; ++r) {
        const float* o = ray_o + r * 3;
        const float* d = ray_d + r * 3;
        int nh = ray_hits(node_min, node_max, active_ids, num_active, o, d, K, hits, ORACLE_HIT_CAP);
        if (nh > max_intersections || nh > ORACLE_HIT_CAP) {
            overflow = 1;
            nh = nh > ORACLE_HIT_CAP ? ORACLE_HIT_CAP : nh;
        }
        int ns = 0;
        for (int h = 0; h < nh; ++h) ns += hits[h].n;
        if (ray_num_samples) ray_num_samples[r] = ns;
        if (ray_num_intervals) ray_num_intervals[r] = nh;
        total += ns;
        if (ray_interval_offset && iv_node) {
            int scan = 0;
            for (int h = 0; h < nh; ++h) {
                int64_t q = ray_interval_offset[r] + h;
                iv_node[q] = hits[h].node;
                iv_tmin[q] = hits[h].tmin;
                iv_tmax[q] = hits[h].tmax;
                iv_count[q] = hits[h].n;
                iv_first[q] = scan;
                scan += hits[h].n;
            }
        }
        if (!ray_sample_offset || !weight) continue;
        int64_t base = ray_sample_offset[r];
        int scan = 0;
        for (int h = 0; h < nh; ++h) {
            float tmin = hits[h].tmin, tmax = hits[h].tmax;
            int n = hits[h].n;
            if (n == 0) continue;
            float step = (tmax - tmin) / (float)n;
            for (int j = 0; j < n; ++j) {
                int64_t slot = base + scan + j;
                float a = jitter_value(jitter_mode, jitter, seed, slot, r, scan + j);
                float t = sample_t(tmin, step, j, a);
                float tp = 0, tn = 0;
                if (j > 0) tp = sample_t(tmin, step, j - 1, jitter_value(jitter_mode, jitter, seed, slot - 1, r, scan + j - 1));
                if (j < n - 1) tn = sample_t(tmin, step, j + 1, jitter_value(jitter_mode, jitter, seed, slot + 1, r, scan + j + 1));
                ray_t[slot] = t;
                weight[slot] = sample_weight(tmin, tmax, n, j, tp, t, tn);
                for (int k = 0; k < 3; ++k) global_coordinate[slot * 3 + k] = o[k] + d[k] * t;
                ray_index[slot] = r;
                local_index_in_ray[slot] = scan + j;
                node_id[slot] = hits[h].node;
            }
            scan += n;
        }
    }
    free(hits);
    return overflow ? -1 : total;
}

/* structure/HyperTree.cu:237-259 */
void atdfo_local_samples(const float* node_min, const float* node_max, const float* g,
                         const int64_t* node, int64_t S, float* local)
{
    for (int64_t s = 0; s < S; ++s) {
        for (int k = 0; k < 3; ++k) {
            float mn = node_min[node[s] * 3 + k];
            float mx = node_max[node[s] * 3 + k];
            float size = mx - mn;
            local[s * 3 + k] = ((g[s * 3 + k] - mn) / size) * 2 - 1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* trilinear gather (IndirectGridSample2D.cu:22-162 pattern in 3-D)          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int x[2], y[2], z[2]; /* clipped corner indices, [0]=low, [1]=high */
    float w[8];           /* tnw tne tsw tse bnw bne bsw bse */
} tri_t;

static int clipi(int v, int lim)
{
    int a = v > 0 ? v : 0;
    return a < (lim - 1) ? a : (lim - 1);
}

static void tri_setup_f(const float* uv, int Z, int Y, int X, tri_t* t)
{
    float ix = ((uv[0] + 1) / 2) * (float)(X - 1);
    float iy = ((uv[1] + 1) / 2) * (float)(Y - 1);
    float iz = ((uv[2] + 1) / 2) * (float)(Z - 1);
    int ix0 = (int)floorf(ix), iy0 = (int)floorf(iy), iz0 = (int)floorf(iz);
    int ix1 = ix0 + 1, iy1 = iy0 + 1, iz1 = iz0 + 1;
    float fx1 = (float)ix1 - ix, fx0 = ix - (float)ix0;
    float fy1 = (float)iy1 - iy, fy0 = iy - (float)iy0;
    float fz1 = (float)iz1 - iz, fz0 = iz - (float)iz0;
    t->w[0] = fx1 * fy1 * fz1;
    t->w[1] = fx0 * fy1 * fz1;
    t->w[2] = fx1 * fy0 * fz1;
    t->w[3] = fx0 * fy0 * fz1;
    t->w[4] = fx1 * fy1 * fz0;
    t->w[5] = fx0 * fy1 * fz0;
    t->w[6] = fx1 * fy0 * fz0;
    t->w[7] = fx0 * fy0 * fz0;
    t->x[0] = clipi(ix0, X);
    t->x[1] = clipi(ix1, X);
    t->y[0] = clipi(iy0, Y);
    t->y[1] = clipi(iy1, Y);
    t->z[0] = clipi(iz0, Z);
    t->z[1] = clipi(iz1, Z);
}

typedef struct {
    int x[2], y[2], z[2];
    double w[8];
    double dw[8][3]; /* d w / d (ix, iy, iz) */
} tri_d;

static void tri_setup_d(const double* uv, int Z, int Y, int X, tri_d* t)
{
    double ix = ((uv[0] + 1) / 2) * (double)(X - 1);
    double iy = ((uv[1] + 1) / 2) * (double)(Y - 1);
    double iz = ((uv[2] + 1) / 2) * (double)(Z - 1);
    int ix0 = (int)floor(ix), iy0 = (int)floor(iy), iz0 = (int)floor(iz);
    int ix1 = ix0 + 1, iy1 = iy0 + 1, iz1 = iz0 + 1;
    double fx[2] = {(double)ix1 - ix, ix - (double)ix0};
    double fy[2] = {(double)iy1 - iy, iy - (double)iy0};
    double fz[2] = {(double)iz1 - iz, iz - (double)iz0};
    double sx[2] = {-1, 1}, sy[2] = {-1, 1}, sz[2] = {-1, 1};
    for (int c = 0; c < 8; ++c) {
        int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
        t->w[c] = fx[bx] * fy[by] * fz[bz];
        t->dw[c][0] = sx[bx] * fy[by] * fz[bz];
        t->dw[c][1] = fx[bx] * sy[by] * fz[bz];
        t->dw[c][2] = fx[bx] * fy[by] * sz[bz];
    }
    t->x[0] = clipi(ix0, X);
    t->x[1] = clipi(ix1, X);
    t->y[0] = clipi(iy0, Y);
    t->y[1] = clipi(iy1, Y);
    t->z[0] = clipi(iz0, Z);
    t->z[1] = clipi(iz1, Z);
}

void atdfo_grid_sample3d_fwd(const float* grid, int64_t N, int C, int Z, int Y, int X,
                             const int64_t* index, const float* uv, int64_t S, float* out)
{
    (void)N;
    int64_t plane = (int64_t)Z * Y * X;
    for (int64_t s = 0; s < S; ++s) {
        tri_t t;
        tri_setup_f(uv + s * 3, Z, Y, X, &t);
        const float* g = grid + index[s] * C * plane;
        for (int c = 0; c < C; ++c) {
            const float* gc = g + c * plane;
            float acc = 0;
            for (int k = 0; k < 8; ++k) {
                int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
                float v = gc[((int64_t)t.z[bz] * Y + t.y[by]) * X + t.x[bx]];
                acc += v * t.w[k];
            }
            out[s * C + c] = acc;
        }
    }
}

void atdfo_grid_sample3d_bwd(const float* grid, int64_t N, int C, int Z, int Y, int X,
                             const int64_t* index, const float* uv, int64_t S, const float* grad_out,
                             double* grad_grid, double* grad_uv)
{
    (void)N;
    int64_t plane = (int64_t)Z * Y * X;
    double scale[3] = {(X - 1) / 2.0, (Y - 1) / 2.0, (Z - 1) / 2.0};
    for (int64_t s = 0; s < S; ++s) {
        double uvd[3] = {uv[s * 3], uv[s * 3 + 1], uv[s * 3 + 2]};
        tri_d t;
        tri_setup_d(uvd, Z, Y, X, &t);
        double gu[3] = {0, 0, 0};
        for (int c = 0; c < C; ++c) {
            double go = grad_out[s * C + c];
            for (int k = 0; k < 8; ++k) {
                int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
                int64_t off = (index[s] * C + c) * plane + ((int64_t)t.z[bz] * Y + t.y[by]) * X + t.x[bx];
                grad_grid[off] += t.w[k] * go;
                if (grad_uv) {
                    double v = grid[off];
                    for (int a = 0; a < 3; ++a) gu[a] += v * t.dw[k][a] * go;
                }
            }
        }
        if (grad_uv)
            for (int a = 0; a < 3; ++a) grad_uv[s * 3 + a] += gu[a] * scale[a];
    }
}

/* ------------------------------------------------------------------------ */
/* decoder MLP: Modules.h:23-145, softplus G.cpp:210-213 (torch threshold 20) */
/* ------------------------------------------------------------------------ */
int64_t atdfo_mlp_num_params(int C, int H) { return (int64_t)H * C + H + (int64_t)H * H + H + H + 1; }

static float softplus_f(float y, float beta)
{
    float yb = y * beta;
    if (yb > 20.0f) return y;
    return log1pf(expf(yb)) / beta;
}

static double softplus_d(double y, double beta)
{
    double yb = y * beta;
    if (yb > 20.0) return y;
    return log1p(exp(yb)) / beta;
}

static double softplus_grad_d(double y, double beta)
{
    double yb = y * beta;
    if (yb > 20.0) return 1.0;
    double z = exp(yb);
    return z / (z + 1.0);
}

/* round to bfloat16 (round to nearest even), returned as double */
static double bf16r(double v)
{
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    memcpy(&f, &u, 4);
    return (double)f;
}

static double rb(int on, double v) { return on ? bf16r(v) : v; }

static float mlp_fwd_one_f(const float* P, int C, int H, float beta, const float* x, float* h1, float* h2,
                           float* y_pre)
{
    const float* W1 = P;
    const float* b1 = W1 + H * C;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* W3 = b2 + H;
    const float* b3 = W3 + H;
    for (int i = 0; i < H; ++i) {
        float acc = 0;
        for (int k = 0; k < C; ++k) acc += W1[i * C + k] * x[k];
        acc += b1[i];
        h1[i] = acc > 0 ? acc : 0;
    }
    for (int i = 0; i < H; ++i) {
        float acc = 0;
        for (int k = 0; k < H; ++k) acc += W2[i * H + k] * h1[k];
        acc += b2[i];
        h2[i] = acc > 0 ? acc : 0;
    }
    float acc = 0;
    for (int k = 0; k < H; ++k) acc += W3[k] * h2[k];
    acc += b3[0];
    if (y_pre) *y_pre = acc;
    return softplus_f(acc, beta);
}

void atdfo_mlp_fwd(const float* params, int C, int H, float beta, const float* x, int64_t S,
                   float* y_pre, float* density)
{
    float* h1 = (float*)malloc(sizeof(float) * H);
    float* h2 = (float*)malloc(sizeof(float) * H);
    for (int64_t s = 0; s < S; ++s) {
        float yp;
        float dn = mlp_fwd_one_f(params, C, H, beta, x + s * C, h1, h2, &yp);
        if (y_pre) y_pre[s] = yp;
        if (density) density[s] = dn;
    }
    free(h1);
    free(h2);
}

/* geometry.cpp:977-1019 */
void atdfo_integrate(const float* v, const float* w, const int64_t* ray, int64_t S, int C,
                     int64_t R, float* P)
{
    for (int64_t s = 0; s < S; ++s)
        for (int c = 0; c < C; ++c) P[c * R + ray[s]] += v[s * C + c] * w[s];
}

/* geometry.cpp:33-53, accumulated in FP64 */
void atdfo_accumulate_loss_per_node(const float* per_ray_loss, int C, int64_t R,
                                    const int64_t* ray_index, const float* weight,
                                    const int64_t* node_id, int64_t S, int64_t num_nodes,
                                    double* loss_sum, double* weight_sum)
{
    (void)num_nodes;
    for (int64_t s = 0; s < S; ++s) {
        int64_t r = ray_index[s];
        double l = 0;
        for (int c = 0; c < C; ++c) l += per_ray_loss[c * R + r];
        l /= C;
        loss_sum[node_id[s]] += l * (double)weight[s];
        weight_sum[node_id[s]] += (double)weight[s];
    }
}

/* ------------------------------------------------------------------------ */
/* fused per-ray path                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    float t, w;
    float pos[3];
    int32_t node;
    int32_t local;
} sample_t_;

/* Place the samples of one ray (canonical order). Returns count, -1 overflow. */
static int ray_samples(const atdfo_model* m, int64_t r, const float* o, const float* d, hit_t* hits,
                       sample_t_* smp, int smp_cap)
{
    int nh;
    if (m->iv_off) {
        nh = (int)(m->iv_off[r + 1] - m->iv_off[r]);
        if (nh > ORACLE_HIT_CAP) return -1;
        for (int h = 0; h < nh; ++h) {
            int64_t q = m->iv_off[r] + h;
            hits[h].node = m->iv_node[q];
            hits[h].tmin = m->iv_tmin[q];
            hits[h].tmax = m->iv_tmax[q];
            hits[h].n = m->iv_count[q];
        }
    } else {
        nh = ray_hits(m->node_min, m->node_max, m->active_ids, m->num_active, o, d, m->K, hits,
                      ORACLE_HIT_CAP);
        if (nh > m->max_intersections || nh > ORACLE_HIT_CAP) return -1;
    }
    int ns = 0;
    for (int h = 0; h < nh; ++h) {
        float tmin = hits[h].tmin, tmax = hits[h].tmax;
        int n = hits[h].n;
        if (n == 0) continue;
        float step = (tmax - tmin) / (float)n;
        for (int j = 0; j < n; ++j) {
            if (ns + j >= smp_cap) return -1;
            int64_t loc = ns + j;
            float a = jitter_value(m->jitter_mode, NULL, m->seed, 0, r, loc);
            float t = sample_t(tmin, step, j, a);
            float tp = 0, tn = 0;
            if (j > 0) tp = sample_t(tmin, step, j - 1, jitter_value(m->jitter_mode, NULL, m->seed, 0, r, loc - 1));
            if (j < n - 1) tn = sample_t(tmin, step, j + 1, jitter_value(m->jitter_mode, NULL, m->seed, 0, r, loc + 1));
            sample_t_* s = &smp[ns + j];
            s->t = t;
            s->w = sample_weight(tmin, tmax, n, j, tp, t, tn);
            for (int k = 0; k < 3; ++k) s->pos[k] = o[k] + d[k] * t;
            s->node = hits[h].node;
            s->local = (int32_t)loc;
        }
        ns += n;
    }
    return ns;
}

static void local_coord_f(const atdfo_model* m, const sample_t_* s, float* loc)
{
    for (int k = 0; k < 3; ++k) {
        float mn = m->node_min[(int64_t)s->node * 3 + k];
        float mx = m->node_max[(int64_t)s->node * 3 + k];
        float size = mx - mn;
        loc[k] = ((s->pos[k] - mn) / size) * 2 - 1;
    }
}

/* channels-last gather, float, corner order tnw..bse */
static void gather_cl_f(const atdfo_model* m, const sample_t_* s, float* feat)
{
    float loc[3];
    local_coord_f(m, s, loc);
    tri_t t;
    int G = m->G, C = m->C;
    tri_setup_f(loc, G, G, G, &t);
    const float* g = m->grid + (int64_t)m->node_to_grid[s->node] * G * G * G * C;
    for (int c = 0; c < C; ++c) feat[c] = 0;
    for (int c = 0; c < C; ++c) {
        float acc = 0;
        for (int k = 0; k < 8; ++k) {
            int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
            float v = g[(((int64_t)t.z[bz] * G + t.y[by]) * G + t.x[bx]) * C + c];
            acc += v * t.w[k];
        }
        feat[c] = acc;
    }
}

static void gather_cl_d(const atdfo_model* m, const sample_t_* s, double* feat, tri_d* t, int64_t* base)
{
    double loc[3];
    for (int k = 0; k < 3; ++k) {
        double mn = m->node_min[(int64_t)s->node * 3 + k];
        double mx = m->node_max[(int64_t)s->node * 3 + k];
        loc[k] = ((s->pos[k] - mn) / (mx - mn)) * 2 - 1;
    }
    int G = m->G, C = m->C;
    tri_setup_d(loc, G, G, G, t);
    *base = (int64_t)m->node_to_grid[s->node] * G * G * G * C;
    const float* g = m->grid + *base;
    for (int c = 0; c < C; ++c) {
        double acc = 0;
        for (int k = 0; k < 8; ++k) {
            int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
            acc += (double)g[(((int64_t)t->z[bz] * G + t->y[by]) * G + t->x[bx]) * C + c] * t->w[k];
        }
        feat[c] = acc;
    }
}

static double mlp_fwd_one_d(const float* P, int C, int H, double beta, const double* x, double* h1,
                            double* h2, double* y_pre, int bf)
{
    const float* W1 = P;
    const float* b1 = W1 + H * C;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* W3 = b2 + H;
    const float* b3 = W3 + H;
    for (int i = 0; i < H; ++i) {
        double acc = 0;
        for (int k = 0; k < C; ++k) acc += rb(bf, W1[i * C + k]) * rb(bf, x[k]);
        acc += b1[i];
        h1[i] = acc > 0 ? acc : 0;
    }
    for (int i = 0; i < H; ++i) {
        double acc = 0;
        for (int k = 0; k < H; ++k) acc += rb(bf, W2[i * H + k]) * rb(bf, h1[k]);
        acc += b2[i];
        h2[i] = acc > 0 ? acc : 0;
    }
    double acc = 0;
    for (int k = 0; k < H; ++k) acc += (double)W3[k] * h2[k];
    acc += b3[0];
    *y_pre = acc;
    return softplus_d(acc, beta);
}

#define ORACLE_SAMPLE_CAP 65536

static int sample_bin(float t, float t0, float t1, int nb)
{
    float span = t1 - t0;
    int b = 0;
    if (span > 0) {
        float rel = (t - t0) / span;
        b = (int)(rel * (float)nb);
    }
    return b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
}

int atdfo_project_fwd(const atdfo_model* m, const float* ray_o, const float* ray_d, int64_t R,
                      int precision, double* P, int num_threads)
{
    return atdfo_project_fwd_bins(m, ray_o, ray_d, R, precision, P, NULL, 0, num_threads);
}

int atdfo_project_fwd_bins(const atdfo_model* m, const float* ray_o, const float* ray_d, int64_t R,
                           int precision, double* P, double* bins, int nb, int num_threads)
{
    int err = 0;
#ifdef _OPENMP
    if (num_threads <= 0) num_threads = omp_get_max_threads();
#pragma omp parallel num_threads(num_threads)
#endif
    {
        hit_t* hits = (hit_t*)malloc(sizeof(hit_t) * ORACLE_HIT_CAP);
        sample_t_* smp = (sample_t_*)malloc(sizeof(sample_t_) * ORACLE_SAMPLE_CAP);
        float* featf = (float*)malloc(sizeof(float) * (m->C + 2 * m->H));
        double* featd = (double*)malloc(sizeof(double) * (m->C + 2 * m->H));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t r = 0; r < R; ++r) {
            int ns = ray_samples(m, r, ray_o + r * 3, ray_d + r * 3, hits, smp, ORACLE_SAMPLE_CAP);
            if (ns < 0) {
                err = 1;
                P[r] = 0;
                continue;
            }
            float t0 = 0, t1 = 0;
            if (bins) {
                for (int b = 0; b < nb; ++b) bins[r * nb + b] = 0;
                atdfo_intersect_box_ray(m->node_min, m->node_max, ray_o + r * 3, ray_d + r * 3, &t0, &t1);
            }
            if (precision == 0) {
                float acc = 0;
                for (int s = 0; s < ns; ++s) {
                    gather_cl_f(m, &smp[s], featf);
                    float dn = mlp_fwd_one_f(m->mlp, m->C, m->H, m->beta, featf, featf + m->C,
                                             featf + m->C + m->H, NULL);
                    acc += dn * smp[s].w;
                    if (bins) bins[r * nb + sample_bin(smp[s].t, t0, t1, nb)] += (double)(dn * smp[s].w);
                }
                P[r] = acc;
            } else {
                double acc = 0;
                for (int s = 0; s < ns; ++s) {
                    tri_d t;
                    int64_t base;
                    double yp;
                    if (m->decoder_bf16) {
                        /* the BF16 decoder rounds the FP32 gather result */
                        gather_cl_f(m, &smp[s], featf);
                        for (int c = 0; c < m->C; ++c) featd[c] = featf[c];
                    } else {
                        gather_cl_d(m, &smp[s], featd, &t, &base);
                    }
                    double dn = mlp_fwd_one_d(m->mlp, m->C, m->H, m->beta, featd, featd + m->C,
                                              featd + m->C + m->H, &yp, m->decoder_bf16);
                    acc += dn * (double)smp[s].w;
                    if (bins) bins[r * nb + sample_bin(smp[s].t, t0, t1, nb)] += dn * (double)smp[s].w;
                }
                P[r] = acc;
            }
        }
        free(hits);
        free(smp);
        free(featf);
        free(featd);
    }
    return err ? -1 : 0;
}

/* Per ray, the smallest |pre-activation| of the decoder's hidden units over
 * its samples (FP64 path).  Where it is tiny, the sign of that unit -- and
 * with it the ReLU mask and the sample's gradient direction -- can differ
 * between FP32 and FP64 evaluation; gradient comparisons leave such rays out. */
void atdfo_project_min_preact(const atdfo_model* m, const float* ray_o, const float* ray_d, int64_t R,
                              double* out)
{
    int C = m->C, H = m->H;
    hit_t* hits = (hit_t*)malloc(sizeof(hit_t) * ORACLE_HIT_CAP);
    sample_t_* smp = (sample_t_*)malloc(sizeof(sample_t_) * ORACLE_SAMPLE_CAP);
    double* x = (double*)malloc(sizeof(double) * C);
    double* h1 = (double*)malloc(sizeof(double) * H);
    float* xf = (float*)malloc(sizeof(float) * C);
    const float* W1 = m->mlp;
    const float* b1 = W1 + H * C;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    for (int64_t r = 0; r < R; ++r) {
        int ns = ray_samples(m, r, ray_o + r * 3, ray_d + r * 3, hits, smp, ORACLE_SAMPLE_CAP);
        double mn = INFINITY;
        for (int s = 0; s < ns; ++s) {
            tri_d t;
            int64_t base;
            if (m->decoder_bf16) {
                gather_cl_f(m, &smp[s], xf);
                for (int c = 0; c < C; ++c) x[c] = xf[c];
            } else {
                gather_cl_d(m, &smp[s], x, &t, &base);
            }
            for (int i = 0; i < H; ++i) {
                double acc = 0;
                for (int k = 0; k < C; ++k) acc += rb(m->decoder_bf16, W1[i * C + k]) * rb(m->decoder_bf16, x[k]);
                acc += b1[i];
                if (fabs(acc) < mn) mn = fabs(acc);
                h1[i] = acc > 0 ? acc : 0;
            }
            for (int i = 0; i < H; ++i) {
                double acc = 0;
                for (int k = 0; k < H; ++k) acc += rb(m->decoder_bf16, W2[i * H + k]) * rb(m->decoder_bf16, h1[k]);
                acc += b2[i];
                if (fabs(acc) < mn) mn = fabs(acc);
            }
        }
        out[r] = mn;
    }
    free(hits);
    free(smp);
    free(x);
    free(h1);
    free(xf);
}

/* FP64 adjoint of one ray: per sample, the grid base offset, the trilinear
 * record and dL/dfeature (dx, C values), plus the ray's decoder-parameter
 * gradient (added to gm).  Returns the sample count, or -1 on overflow. */
typedef struct {
    int64_t base;
    tri_d t;
} bwd_rec_t;

static int bwd_ray(const atdfo_model* m, int64_t r, const float* o, const float* d, double dLdP, hit_t* hits,
                   sample_t_* smp, bwd_rec_t** rec_buf, double** dx_buf, int* cap, double* gm, double* work)
{
    int C = m->C, H = m->H;
    int ns = ray_samples(m, r, o, d, hits, smp, ORACLE_SAMPLE_CAP);
    if (ns < 0) return -1;
    if (ns > *cap) {
        *cap = ns;
        *rec_buf = (bwd_rec_t*)realloc(*rec_buf, sizeof(bwd_rec_t) * (size_t)ns);
        *dx_buf = (double*)realloc(*dx_buf, sizeof(double) * (size_t)ns * C);
    }
    bwd_rec_t* rec = *rec_buf;
    double* dxs = *dx_buf;
    double* x = work;
    double* h1 = x + C;
    double* h2 = h1 + H;
    double* dh2 = h2 + H;
    double* dh1 = dh2 + H;
    const float* W1 = m->mlp;
    const float* W2 = W1 + H * C + H;
    const float* W3 = W2 + H * H + H;
    double* gW1 = gm;
    double* gb1 = gW1 + H * C;
    double* gW2 = gb1 + H;
    double* gb2 = gW2 + H * H;
    double* gW3 = gb2 + H;
    double* gb3 = gW3 + H;
    const int bf = m->decoder_bf16;
    for (int s = 0; s < ns; ++s) {
        double yp;
        gather_cl_d(m, &smp[s], x, &rec[s].t, &rec[s].base);
        if (bf) {
            float xf[64];
            gather_cl_f(m, &smp[s], xf);
            for (int c = 0; c < C; ++c) x[c] = xf[c];
        }
        mlp_fwd_one_d(m->mlp, C, H, m->beta, x, h1, h2, &yp, bf);
        double dy = dLdP * (double)smp[s].w * softplus_grad_d(yp, m->beta);
        for (int i = 0; i < H; ++i) {
            gW3[i] += dy * h2[i];
            dh2[i] = (h2[i] > 0) ? dy * (double)W3[i] : 0;
        }
        gb3[0] += dy;
        for (int i = 0; i < H; ++i) dh1[i] = 0;
        for (int i = 0; i < H; ++i) {
            if (dh2[i] == 0) continue;
            gb2[i] += dh2[i];
            for (int k = 0; k < H; ++k) {
                gW2[i * H + k] += rb(bf, dh2[i]) * rb(bf, h1[k]);
                dh1[k] += rb(bf, dh2[i]) * rb(bf, W2[i * H + k]);
            }
        }
        double* dx = dxs + (int64_t)s * C;
        for (int c = 0; c < C; ++c) dx[c] = 0;
        for (int i = 0; i < H; ++i) {
            if (!(h1[i] > 0)) continue;
            gb1[i] += dh1[i];
            for (int k = 0; k < C; ++k) {
                gW1[i * C + k] += rb(bf, dh1[i]) * rb(bf, x[k]);
                dx[k] += rb(bf, dh1[i]) * rb(bf, W1[i * C + k]);
            }
        }
    }
    return ns;
}

/* FP64 adjoint of atdfo_project_fwd(precision=1).  Rays are evaluated in
 * parallel in blocks; the grid scatter and the decoder-gradient sums run in
 * ray order afterwards, so the result does not depend on the thread count. */
#define BWD_BLOCK 64
int atdfo_project_bwd(const atdfo_model* m, const float* ray_o, const float* ray_d, int64_t R,
                      const double* dLdP, double* grad_grid, double* grad_mlp)
{
    int C = m->C, H = m->H, G = m->G;
    int64_t NP = (int64_t)H * C + H + (int64_t)H * H + H + H + 1;
    bwd_rec_t* rec[BWD_BLOCK];
    double* dxs[BWD_BLOCK];
    double* gms[BWD_BLOCK];
    int cnt[BWD_BLOCK], cap[BWD_BLOCK];
    for (int b = 0; b < BWD_BLOCK; ++b) {
        cap[b] = 1024;
        rec[b] = (bwd_rec_t*)malloc(sizeof(bwd_rec_t) * (size_t)cap[b]);
        dxs[b] = (double*)malloc(sizeof(double) * (size_t)cap[b] * C);
        gms[b] = (double*)malloc(sizeof(double) * NP);
    }
    int err = 0;
    for (int64_t r0 = 0; r0 < R; r0 += BWD_BLOCK) {
        int nb = (int)(R - r0 < BWD_BLOCK ? R - r0 : BWD_BLOCK);
#ifdef _OPENMP
#pragma omp parallel
#endif
        {
            hit_t* hits = (hit_t*)malloc(sizeof(hit_t) * ORACLE_HIT_CAP);
            sample_t_* smp = (sample_t_*)malloc(sizeof(sample_t_) * ORACLE_SAMPLE_CAP);
            double* work = (double*)malloc(sizeof(double) * (C + 4 * H));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int b = 0; b < nb; ++b) {
                int64_t r = r0 + b;
                for (int64_t i = 0; i < NP; ++i) gms[b][i] = 0;
                cnt[b] = bwd_ray(m, r, ray_o + r * 3, ray_d + r * 3, dLdP[r], hits, smp, &rec[b], &dxs[b], &cap[b],
                                 gms[b], work);
            }
            free(hits);
            free(smp);
            free(work);
        }
        for (int b = 0; b < nb; ++b) {
            if (cnt[b] < 0) {
                err = 1;
                continue;
            }
            for (int64_t i = 0; i < NP; ++i) grad_mlp[i] += gms[b][i];
            for (int s = 0; s < cnt[b]; ++s) {
                const bwd_rec_t* q = &rec[b][s];
                const double* dx = dxs[b] + (int64_t)s * C;
                for (int k = 0; k < 8; ++k) {
                    int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
                    int64_t off = q->base + (((int64_t)q->t.z[bz] * G + q->t.y[by]) * G + q->t.x[bx]) * C;
                    for (int c = 0; c < C; ++c) grad_grid[off + c] += q->t.w[k] * dx[c];
                }
            }
        }
    }
    for (int b = 0; b < BWD_BLOCK; ++b) {
        free(rec[b]);
        free(dxs[b]);
        free(gms[b]);
    }
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* per-node density statistics                                               */
/* ------------------------------------------------------------------------ */
float atdfo_point_uniform(uint64_t seed, int64_t node, int64_t point, int axis)
{
    uint64_t key = ((uint64_t)node << 34) ^ ((uint64_t)point << 2) ^ (uint64_t)axis;
    uint64_t h = splitmix64(seed ^ splitmix64(key ^ 0x5bd1e995ull));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

void atdfo_node_density_stats(const atdfo_model* m, int mode, int n, int rounds, uint64_t seed, int precision,
                              float* node_max, double* node_sum, double* node_sumsq, float* densities)
{
    int C = m->C, H = m->H;
    int64_t per = (int64_t)n * n * n * (mode == 1 ? rounds : 1);
    float* featf = (float*)malloc(sizeof(float) * (C + 2 * H));
    double* featd = (double*)malloc(sizeof(double) * (C + 2 * H));
    for (int64_t a = 0; a < m->num_active; ++a) {
        int64_t node = m->active_ids[a];
        const float* mn = m->node_min + node * 3;
        const float* mx = m->node_max + node * 3;
        float size[3];
        for (int k = 0; k < 3; ++k) size[k] = mx[k] - mn[k];
        float dmax = -INFINITY;
        double s1 = 0, s2 = 0;
        for (int64_t p = 0; p < per; ++p) {
            sample_t_ smp;
            smp.node = (int32_t)node;
            if (mode == 0) {
                int64_t idx[3] = {p % n, (p / n) % n, p / ((int64_t)n * n)};
                for (int k = 0; k < 3; ++k) {
                    float lp = (float)idx[k] / (float)(n - 1);
                    smp.pos[k] = lp * size[k] + mn[k];
                }
            } else {
                for (int k = 0; k < 3; ++k) smp.pos[k] = mn[k] + atdfo_point_uniform(seed, node, p, k) * size[k];
            }
            float dn;
            if (precision == 0) {
                gather_cl_f(m, &smp, featf);
                dn = mlp_fwd_one_f(m->mlp, C, H, m->beta, featf, featf + C, featf + C + H, NULL);
            } else {
                tri_d t;
                int64_t base;
                double yp;
                if (m->decoder_bf16) {
                    gather_cl_f(m, &smp, featf);
                    for (int c = 0; c < C; ++c) featd[c] = featf[c];
                } else {
                    gather_cl_d(m, &smp, featd, &t, &base);
                }
                dn = (float)mlp_fwd_one_d(m->mlp, C, H, m->beta, featd, featd + C, featd + C + H, &yp,
                                          m->decoder_bf16);
            }
            if (densities) densities[a * per + p] = dn;
            if (dn > dmax) dmax = dn;
            s1 += (double)dn;
            s2 += (double)dn * (double)dn;
        }
        node_max[a] = dmax;
        node_sum[a] = s1;
        node_sumsq[a] = s2;
    }
    free(featf);
    free(featd);
}

/* ------------------------------------------------------------------------ */
/* FP32 CPU training step (timed baseline)                                   */
/* ------------------------------------------------------------------------ */
static void adam_f32(float* p, float* g, float* mm, float* vv, int64_t n, int step, float lr)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float bc1 = 1.0f - powf(b1, (float)step);
    float bc2 = 1.0f - powf(b2, (float)step);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; ++i) {
        float gi = g[i];
        mm[i] = b1 * mm[i] + (1 - b1) * gi;
        vv[i] = b2 * vv[i] + (1 - b2) * gi * gi;
        float denom = sqrtf(vv[i]) / sqrtf(bc2) + eps;
        p[i] -= (lr / bc1) * mm[i] / denom;
        g[i] = 0;
    }
}

double atdfo_train_step_f32(atdfo_model* m, float* grid_mut, float* mlp_mut, int64_t grid_numel,
                            const float* ray_o, const float* ray_d, const float* target, int64_t R,
                            float* grad_grid, float* grad_mlp, float* adam_m_grid, float* adam_v_grid,
                            float* adam_m_mlp, float* adam_v_mlp, int step, float lr, int num_threads,
                            int do_adam)
{
    int C = m->C, H = m->H, G = m->G;
    int64_t nmlp = atdfo_mlp_num_params(C, H);
    m->grid = grid_mut;
    m->mlp = mlp_mut;
    double loss = 0;
#ifdef _OPENMP
    if (num_threads <= 0) num_threads = omp_get_max_threads();
#pragma omp parallel num_threads(num_threads) reduction(+ : loss)
#endif
    {
        hit_t* hits = (hit_t*)malloc(sizeof(hit_t) * ORACLE_HIT_CAP);
        sample_t_* smp = (sample_t_*)malloc(sizeof(sample_t_) * ORACLE_SAMPLE_CAP);
        float* x = (float*)malloc(sizeof(float) * C);
        float* h1 = (float*)malloc(sizeof(float) * H);
        float* h2 = (float*)malloc(sizeof(float) * H);
        float* dh1 = (float*)malloc(sizeof(float) * H);
        float* dx = (float*)malloc(sizeof(float) * C);
        float* gm = (float*)calloc((size_t)nmlp, sizeof(float));
        const float* W1 = mlp_mut;
        const float* W2 = W1 + H * C + H;
        const float* W3 = W2 + H * H + H;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t r = 0; r < R; ++r) {
            int ns = ray_samples(m, r, ray_o + r * 3, ray_d + r * 3, hits, smp, ORACLE_SAMPLE_CAP);
            if (ns < 0) continue;
            float P = 0;
            for (int s = 0; s < ns; ++s) {
                gather_cl_f(m, &smp[s], x);
                P += mlp_fwd_one_f(mlp_mut, C, H, m->beta, x, h1, h2, NULL) * smp[s].w;
            }
            float diff = P - target[r];
            loss += (double)diff * diff;
            float dP = 2.0f * diff / (float)R;
            for (int s = 0; s < ns; ++s) {
                float yp;
                gather_cl_f(m, &smp[s], x);
                mlp_fwd_one_f(mlp_mut, C, H, m->beta, x, h1, h2, &yp);
                float yb = yp * m->beta;
                float sg = yb > 20.0f ? 1.0f : 1.0f / (1.0f + expf(-yb));
                float dy = dP * smp[s].w * sg;
                float* gW1 = gm;
                float* gb1 = gW1 + H * C;
                float* gW2 = gb1 + H;
                float* gb2 = gW2 + H * H;
                float* gW3 = gb2 + H;
                float* gb3 = gW3 + H;
                gb3[0] += dy;
                for (int k = 0; k < H; ++k) dh1[k] = 0;
                for (int i = 0; i < H; ++i) {
                    gW3[i] += dy * h2[i];
                    if (!(h2[i] > 0)) continue;
                    float d2 = dy * W3[i];
                    gb2[i] += d2;
                    for (int k = 0; k < H; ++k) {
                        gW2[i * H + k] += d2 * h1[k];
                        dh1[k] += d2 * W2[i * H + k];
                    }
                }
                for (int c = 0; c < C; ++c) dx[c] = 0;
                for (int i = 0; i < H; ++i) {
                    if (!(h1[i] > 0)) continue;
                    gb1[i] += dh1[i];
                    for (int k = 0; k < C; ++k) {
                        gW1[i * C + k] += dh1[i] * x[k];
                        dx[k] += dh1[i] * W1[i * C + k];
                    }
                }
                float loc[3];
                local_coord_f(m, &smp[s], loc);
                tri_t t;
                tri_setup_f(loc, G, G, G, &t);
                int64_t base = (int64_t)m->node_to_grid[smp[s].node] * G * G * G * C;
                for (int k = 0; k < 8; ++k) {
                    int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
                    int64_t off = base + (((int64_t)t.z[bz] * G + t.y[by]) * G + t.x[bx]) * C;
                    for (int c = 0; c < C; ++c) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                        grad_grid[off + c] += t.w[k] * dx[c];
                    }
                }
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (int64_t i = 0; i < nmlp; ++i) grad_mlp[i] += gm[i];
        }
        free(hits);
        free(smp);
        free(x);
        free(h1);
        free(h2);
        free(dh1);
        free(dx);
        free(gm);
    }
    if (do_adam) {
        adam_f32(grid_mut, grad_grid, adam_m_grid, adam_v_grid, grid_numel, step, lr);
        adam_f32(mlp_mut, grad_mlp, adam_m_mlp, adam_v_mlp, nmlp, step, lr);
    }
    return loss / (double)R;
}

/* ------------------------------------------------------------------------ */
/* density of the fused field at given points (volume error)                 */
/* ------------------------------------------------------------------------ */
/* density[i] = softplus(decoder(trilinear(grid of node[i], pos[i]))), in the
 * FP32 path (precision 0) or the FP64 path (precision 1); node < 0 gives 0. */
void atdfo_field_density(const atdfo_model* m, const float* pos, const int64_t* node, int64_t n, int precision,
                         double* density)
{
    int C = m->C, H = m->H;
#pragma omp parallel
    {
        float* featf = (float*)malloc(sizeof(float) * (C + 2 * H));
        double* featd = (double*)malloc(sizeof(double) * (C + 2 * H));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            if (node[i] < 0) {
                density[i] = 0.0;
                continue;
            }
            sample_t_ smp;
            smp.node = (int32_t)node[i];
            for (int k = 0; k < 3; ++k) smp.pos[k] = pos[i * 3 + k];
            if (precision == 0) {
                gather_cl_f(m, &smp, featf);
                density[i] = mlp_fwd_one_f(m->mlp, C, H, m->beta, featf, featf + C, featf + C + H, NULL);
            } else {
                tri_d t;
                int64_t base;
                double yp;
                if (m->decoder_bf16) {
                    gather_cl_f(m, &smp, featf);
                    for (int c = 0; c < C; ++c) featd[c] = featf[c];
                } else {
                    gather_cl_d(m, &smp, featd, &t, &base);
                }
                density[i] = mlp_fwd_one_d(m->mlp, C, H, m->beta, featd, featd + C, featd + C + H, &yp,
                                           m->decoder_bf16);
            }
        }
        free(featf);
        free(featd);
    }
}

/* ------------------------------------------------------------------------ */
/* DecodeFeatures on feature rows and its adjoint (geometry.cpp:176-253,     */
/* the FCBlock of Modules.h:23-145 with the softplus last activation)        */
/* ------------------------------------------------------------------------ */
/* FP64; bf = 1 rounds the MLP inputs and the first two layers' weights to
 * BF16 as the device's BF16 decoder does.  density, grad_x nullable; dd null
 * skips the adjoint; grad_mlp is ADDED to. */
void atdfo_decode_features(const float* params, int C, int H, double beta, const float* x, int64_t S, int bf,
                           const double* dd, double* density, double* grad_x, double* grad_mlp)
{
    double* xd = (double*)malloc(sizeof(double) * C);
    double* h1 = (double*)malloc(sizeof(double) * H);
    double* h2 = (double*)malloc(sizeof(double) * H);
    double* dh2 = (double*)malloc(sizeof(double) * H);
    double* dh1 = (double*)malloc(sizeof(double) * H);
    const float* W1 = params;
    const float* W2 = W1 + H * C + H;
    const float* W3 = W2 + H * H + H;
    for (int64_t s = 0; s < S; ++s) {
        double yp;
        for (int c = 0; c < C; ++c) xd[c] = x[s * C + c];
        double dn = mlp_fwd_one_d(params, C, H, beta, xd, h1, h2, &yp, bf);
        if (density) density[s] = dn;
        if (!dd) continue;
        double dy = dd[s] * softplus_grad_d(yp, beta);
        double* gW1 = grad_mlp;
        double* gb1 = gW1 + H * C;
        double* gW2 = gb1 + H;
        double* gb2 = gW2 + H * H;
        double* gW3 = gb2 + H;
        double* gb3 = gW3 + H;
        for (int i = 0; i < H; ++i) {
            if (gW3) gW3[i] += dy * h2[i];
            dh2[i] = (h2[i] > 0) ? dy * (double)W3[i] : 0;
            dh1[i] = 0;
        }
        if (gb3) gb3[0] += dy;
        for (int i = 0; i < H; ++i) {
            if (dh2[i] == 0) continue;
            if (grad_mlp) gb2[i] += dh2[i];
            for (int k = 0; k < H; ++k) {
                if (grad_mlp) gW2[i * H + k] += rb(bf, dh2[i]) * rb(bf, h1[k]);
                dh1[k] += rb(bf, dh2[i]) * rb(bf, W2[i * H + k]);
            }
        }
        double* gx = grad_x ? grad_x + s * C : NULL;
        if (gx)
            for (int c = 0; c < C; ++c) gx[c] = 0;
        for (int i = 0; i < H; ++i) {
            if (!(h1[i] > 0)) continue;
            if (grad_mlp) gb1[i] += dh1[i];
            for (int k = 0; k < C; ++k) {
                if (grad_mlp) gW1[i * C + k] += rb(bf, dh1[i]) * rb(bf, xd[k]);
                if (gx) gx[k] += rb(bf, dh1[i]) * rb(bf, W1[i * C + k]);
            }
        }
    }
    free(xd);
    free(h1);
    free(h2);
    free(dh2);
    free(dh1);
}

/*
 * atdf_oracle_restructure.c -- CPU oracle of point location in the active
 * set and of the nearest-lattice grid transfer.
 *
 * TEST INFRASTRUCTURE ONLY (see atdf_oracle.h).  Follows
 *   ActiveNodeIdForGlobalPosition  structure/HyperTree.cu:1339-1399
 *   ComputeNodeId / NodeIdForPositionGPU  structure/HyperTree.cu:1466-1525
 *   InterpolateGridForInactiveNodes  structure/HyperTree.cu:1527-1616
 * Lattice positions and target coordinates are evaluated in the reference's
 * order with no FMA contraction (bit-exact with the kernels).  Where the
 * reference would read node -1 (a lattice point outside every active node)
 * the destination is left unchanged.  A destination node active in the
 * source set keeps its own grid (the reference skips active nodes).
 */
#include <math.h>
#include <stdint.h>

#include "atdf_oracle.h"

/* descent over the first `nchild` children of every node, in a tree of
 * `ns` children per node (stride of `children`): nchild = ns, or 4 on an
 * octree as the reference's ActiveNodeIdForGlobalPosition with use_quad
 * (HyperTree.cu:1340-1400) does on any tree */
static int32_t locate_(const float* node_min, const float* node_max, const int32_t* children, int ns,
                       const int32_t* active, const float* p, int nchild)
{
    int32_t node = 0;
    for (;;) {
        if (active[node] > 0) return node;
        int32_t next = -1;
        for (int c = 0; c < nchild; ++c) {
            int32_t ch = children[(int64_t)node * ns + c];
            if (ch < 0) return -1;
            const float* a = node_min + (int64_t)ch * 3;
            const float* b = node_max + (int64_t)ch * 3;
            if (a[0] <= p[0] && b[0] >= p[0] && a[1] <= p[1] && b[1] >= p[1] && a[2] <= p[2] && b[2] >= p[2]) {
                next = ch;
                break;
            }
        }
        if (next < 0) return -1;
        node = next;
    }
}

int32_t atdfo_active_node_for_position_ns(const float* node_min, const float* node_max, const int32_t* children,
                                          int ns, const int32_t* active, const float* p)
{
    return locate_(node_min, node_max, children, ns, active, p, ns);
}

int32_t atdfo_active_node_for_position(const float* node_min, const float* node_max, const int32_t* children,
                                       const int32_t* active, const float* p)
{
    return locate_(node_min, node_max, children, 8, active, p, 8);
}

void atdfo_node_id_for_position_ns(const float* node_min, const float* node_max, const int32_t* children, int ns,
                                   const int32_t* active, int64_t first_active, const float* pos, int64_t n,
                                   int64_t* node_id, float* mask)
{
    for (int64_t i = 0; i < n; ++i) {
        int32_t nd = locate_(node_min, node_max, children, ns, active, pos + i * 3, ns);
        node_id[i] = nd < 0 ? first_active : nd;
        mask[i] = nd < 0 ? 0.f : 1.f;
    }
}

static int64_t offset(int layout, int64_t slot, int C, int G, int x, int y, int z, int c)
{
    int64_t g3 = (int64_t)G * G * G;
    int64_t cell = ((int64_t)z * G + y) * G + x;
    return layout == 0 ? (slot * g3 + cell) * C + c : (slot * C + c) * g3 + cell;
}

void atdfo_node_id_for_position(const float* node_min, const float* node_max, const int32_t* children,
                                const int32_t* active, int64_t first_active, const float* pos, int64_t n,
                                int64_t* node_id, float* mask)
{
    atdfo_node_id_for_position_ns(node_min, node_max, children, 8, active, first_active, pos, n, node_id, mask);
}

void atdfo_grid_transfer_ns(const float* node_min, const float* node_max, const int32_t* children, int ns,
                            const int32_t* src_active, const int32_t* src_slot, const float* src_grid,
                            const int64_t* dst_nodes, int64_t n_dst, const int32_t* dst_slot, float* dst_grid, int C,
                            int G, int layout)
{
    const float gm1 = (float)(G - 1);
    for (int64_t j = 0; j < n_dst; ++j) {
        int64_t node = dst_nodes[j];
        int64_t ds = dst_slot ? dst_slot[node] : node;
        const float* mn = node_min + node * 3;
        const float* mx = node_max + node * 3;
        if (src_active[node] > 0) { /* stays active: its own grid */
            int64_t ss = src_slot ? src_slot[node] : node;
            int64_t rec = (int64_t)G * G * G * C;
            for (int64_t i = 0; i < rec; ++i) dst_grid[ds * rec + i] = src_grid[ss * rec + i];
            continue;
        }
        for (int z = 0; z < G; ++z)
            for (int y = 0; y < G; ++y)
                for (int x = 0; x < G; ++x) {
                    int res[3] = {x, y, z};
                    float g[3];
                    for (int k = 0; k < 3; ++k) {
                        float size = mx[k] - mn[k];
                        float local = (float)res[k] / gm1;
                        float c = local * size;
                        g[k] = c + mn[k];
                    }
                    int32_t t = locate_(node_min, node_max, children, ns, src_active, g, ns);
                    if (t < 0) continue;
                    int64_t ss = src_slot ? src_slot[t] : t;
                    int fi[3];
                    for (int k = 0; k < 3; ++k) {
                        float tsize = node_max[t * 3 + k] - node_min[t * 3 + k];
                        float c = g[k] - node_min[t * 3 + k];
                        c = c / tsize;
                        float f = c * gm1;
                        int r = (int)roundf(f);
                        fi[k] = r < 0 ? 0 : (r > G - 1 ? G - 1 : r);
                    }
                    for (int ch = 0; ch < C; ++ch)
                        dst_grid[offset(layout, ds, C, G, x, y, z, ch)] =
                            src_grid[offset(layout, ss, C, G, fi[0], fi[1], fi[2], ch)];
                }
    }
}

void atdfo_grid_transfer(const float* node_min, const float* node_max, const int32_t* children,
                         const int32_t* src_active, const int32_t* src_slot, const float* src_grid,
                         const int64_t* dst_nodes, int64_t n_dst, const int32_t* dst_slot, float* dst_grid, int C,
                         int G, int layout)
{
    atdfo_grid_transfer_ns(node_min, node_max, children, 8, src_active, src_slot, src_grid, dst_nodes, n_dst, dst_slot,
                           dst_grid, C, G, layout);
}

void atdfo_volume_points(int mode, int swap_xy, const int* size, const float* roi_min, const float* step,
                         float* pos)
{
    /* UniformPhantomSamplesGPU / ...bySlice (structure/HyperTree.cu:1158-1293) */
    int64_t n = (int64_t)size[0] * size[1] * size[2];
    for (int64_t p = 0; p < n; ++p) {
        int c[3];
        int64_t tmp = p;
        if (swap_xy) {
            for (int d = 2; d >= 0; --d) {
                c[d] = (int)(tmp % size[d]);
                tmp /= size[d];
            }
        } else {
            for (int d = 0; d < 3; ++d) {
                c[d] = (int)(tmp % size[d]);
                tmp /= size[d];
            }
        }
        for (int k = 0; k < 3; ++k) {
            if (mode == 0) {
                float q = (float)c[k] / ((float)size[k] - 1.0f);
                pos[p * 3 + k] = q * 2 - 1.0f;
            } else {
                float s = step[k] * (float)c[k];
                pos[p * 3 + k] = roi_min[k] + s;
            }
        }
    }
}

void atdfo_node_lattice_points(const float* node_min, const float* node_max, const int64_t* active_ids,
                               int64_t num_active, int n, float* pos)
{
    int64_t per = (int64_t)n * n * n;
    for (int64_t a = 0; a < num_active; ++a) {
        const float* mn = node_min + active_ids[a] * 3;
        const float* mx = node_max + active_ids[a] * 3;
        float size[3];
        for (int k = 0; k < 3; ++k) size[k] = mx[k] - mn[k];
        for (int64_t p = 0; p < per; ++p) {
            int64_t idx[3] = {p % n, (p / n) % n, p / ((int64_t)n * n)};
            for (int k = 0; k < 3; ++k) {
                float lp = (float)idx[k] / (float)(n - 1);
                pos[(a * per + p) * 3 + k] = lp * size[k] + mn[k];
            }
        }
    }
}

/* NodeNeighborSamples (structure/HyperTree.cu:1720-1811, variant 0): pairs
 * in candidate order (node, side, back / front, lattice index); in_roi of
 * HyperTree.cu:1667-1690 with its mid = min + max.  Variant 1 is
 * NodeNeighborSamplesQuad (HyperTree.cu:1915-2008): sides x and z only,
 * in_roi_quad (HyperTree.cu:1692-1718: the y range widened by eps / 2) and
 * the four-child descent.  Variant 2 is NodeNeighborSamples2D
 * (HyperTree.cu:1815-1913) of the quad-tree representation: on the face
 * local[plane] = 0 of each node, for tt = 0, 1 (side_a, side_b = the two
 * other axes in the order (plane + 1, plane + 2) and then swapped), back /
 * front along side_a, size[side_b] points along side_b; the four-child
 * descent; in_roi with the ROI's y range widened by eps once per candidate
 * the reference's thread has visited so far (it decrements its copy of
 * roi_min.y and increments roi_max.y in the loop body, float arithmetic,
 * 128 threads per node).  Writes up to `cap` pairs (pos [2P, 3], node [2P])
 * and returns the pair count. */
static int in_roi_(const float* mn, const float* mx, const float* lo_, const float* hi_, float widen_y)
{
    for (int k = 0; k < 3; ++k) {
        float lo = lo_[k], hi = hi_[k];
        if (k == 1) {
            lo = lo - widen_y;
            hi = hi + widen_y;
        }
        float mid = mn[k] + mx[k];
        if (!((mn[k] > lo && mn[k] < hi) || (mx[k] > lo && mx[k] < hi) || (mid > lo && mid < hi))) return 0;
    }
    return 1;
}

static int64_t neighbor_pairs_2d(const float* node_min, const float* node_max, const int32_t* children, int ns,
                                 const int32_t* active, const int64_t* active_ids, int64_t num_active, const int* size,
                                 float eps, int use_roi, const float* roi_min, const float* roi_max, int plane,
                                 float* pos, int64_t* node_out, int64_t cap)
{
    const int threads = 128;
    const int sidx[2] = {(plane + 1) % 3, (plane + 2) % 3};
    int64_t np = 0;
    for (int64_t a = 0; a < num_active; ++a) {
        int64_t node = active_ids[a];
        const float* mn = node_min + node * 3;
        const float* mx = node_max + node * 3;
        for (int tt = 0; tt < 2; ++tt) {
            const int sa = sidx[tt], sb = sidx[(tt + 1) % 2];
            for (int front = 0; front < 2; ++front) {
                for (int sid = 0; sid < size[sb]; ++sid) {
                    /* the loop iterations the reference's thread sid % 128
                     * has run up to this candidate */
                    const int t = sid % threads;
                    int m = 0;
                    for (int tp = 0; tp <= tt; ++tp)
                        for (int fp = 0; fp < 2; ++fp) {
                            if (tp == tt && fp > front) break;
                            const int nb = size[sidx[(tp + 1) % 2]];
                            const int upto = (tp == tt && fp == front) ? sid : nb - 1;
                            for (int s2 = t; s2 <= upto && s2 < nb; s2 += threads) ++m;
                        }
                    float lo[3] = {roi_min[0], roi_min[1], roi_min[2]};
                    float hi[3] = {roi_max[0], roi_max[1], roi_max[2]};
                    for (int i = 0; i < m; ++i) {
                        lo[1] = lo[1] - eps;
                        hi[1] = hi[1] + eps;
                    }
                    float local[3], g[3], q[3];
                    local[sa] = front ? 1.f : 0.f;
                    local[sb] = (float)(sid % size[sb]) / ((float)size[sb] - 1.f);
                    local[plane] = 0.f;
                    for (int k = 0; k < 3; ++k) {
                        float box = mx[k] - mn[k];
                        g[k] = local[k] * box + mn[k];
                        q[k] = g[k];
                    }
                    q[sa] = g[sa] + (front ? eps : -eps);
                    int32_t other = locate_(node_min, node_max, children, ns, active, q, 4);
                    if (other < 0 || other == node) continue;
                    if (use_roi && !(in_roi_(mn, mx, lo, hi, 0.f) &&
                                     in_roi_(node_min + (int64_t)other * 3, node_max + (int64_t)other * 3, lo, hi, 0.f)))
                        continue;
                    if (np < cap) {
                        for (int k = 0; k < 3; ++k) {
                            pos[(2 * np) * 3 + k] = g[k];
                            pos[(2 * np + 1) * 3 + k] = g[k];
                        }
                        node_out[2 * np] = node;
                        node_out[2 * np + 1] = other;
                    }
                    ++np;
                }
            }
        }
    }
    return np;
}

int64_t atdfo_neighbor_pairs_ns(const float* node_min, const float* node_max, const int32_t* children, int ns,
                                const int32_t* active, const int64_t* active_ids, int64_t num_active, const int* size,
                                float eps, int use_roi, const float* roi_min, const float* roi_max, int variant,
                                int plane, float* pos, int64_t* node_out, int64_t cap)
{
    if (variant == 2)
        return neighbor_pairs_2d(node_min, node_max, children, ns, active, active_ids, num_active, size, eps, use_roi,
                                 roi_min, roi_max, plane, pos, node_out, cap);
    int64_t np = 0;
    const int quad = variant == 1;
    const float widen = quad ? eps / 2.f : 0.f;
    for (int64_t a = 0; a < num_active; ++a) {
        int64_t node = active_ids[a];
        const float* mn = node_min + node * 3;
        const float* mx = node_max + node * 3;
        for (int side = 0; side < 3; side += quad ? 2 : 1) {
            int sa = (side + 1) % 3, sb = (side + 2) % 3;
            for (int front = 0; front < 2; ++front) {
                for (int sid = 0; sid < size[sa] * size[sb]; ++sid) {
                    float local[3], g[3], q[3];
                    local[side] = front ? 1.f : 0.f;
                    local[sa] = (float)(sid % size[sa]) / ((float)size[sa] - 1.f);
                    local[sb] = (float)(sid / size[sa]) / ((float)size[sb] - 1.f);
                    for (int k = 0; k < 3; ++k) {
                        float box = mx[k] - mn[k];
                        g[k] = local[k] * box + mn[k];
                        q[k] = g[k];
                    }
                    q[side] = g[side] + (front ? eps : -eps);
                    int32_t other = locate_(node_min, node_max, children, ns, active, q, quad ? 4 : 8);
                    if (other < 0 || other == node) continue;
                    if (use_roi && !(in_roi_(mn, mx, roi_min, roi_max, widen) &&
                                     in_roi_(node_min + (int64_t)other * 3, node_max + (int64_t)other * 3, roi_min,
                                             roi_max, widen)))
                        continue;
                    if (np < cap) {
                        for (int k = 0; k < 3; ++k) {
                            pos[(2 * np) * 3 + k] = g[k];
                            pos[(2 * np + 1) * 3 + k] = g[k];
                        }
                        node_out[2 * np] = node;
                        node_out[2 * np + 1] = other;
                    }
                    ++np;
                }
            }
        }
    }
    return np;
}

int64_t atdfo_neighbor_pairs(const float* node_min, const float* node_max, const int32_t* children,
                             const int32_t* active, const int64_t* active_ids, int64_t num_active, const int* size,
                             float eps, int use_roi, const float* roi_min, const float* roi_max, int variant,
                             float* pos, int64_t* node_out, int64_t cap)
{
    return atdfo_neighbor_pairs_ns(node_min, node_max, children, 8, active, active_ids, num_active, size, eps, use_roi,
                                   roi_min, roi_max, variant, 1, pos, node_out, cap);
}

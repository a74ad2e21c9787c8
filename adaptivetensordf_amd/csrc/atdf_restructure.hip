// atdf_restructure.hip -- point location in the active set and the transfer
// of node grids across a change of the active set.
//
// Point location follows ActiveNodeIdForGlobalPosition
// (structure/HyperTree.cu:1339-1399): from the root, while the node is not
// active, step into the first child (in child order) whose closed box holds
// the point; no such child, or a leaf, means "not in an active node".  The
// literal box tests are kept (instead of a mid-plane rule) so the result
// does not depend on how the child boxes were rounded.
//
// The grid transfer is the nearest-lattice copy of
// InterpolateGridForInactiveNodes (structure/HyperTree.cu:1527-1616): every
// lattice point (x, y, z) / (G - 1) of a destination node is mapped to the
// global position, located among the source active nodes, and takes the
// whole channel record of the nearest lattice point of that node
// (round half away from zero).  A destination node that is active in the
// source set keeps its own grid, as the reference's kernel leaves active
// nodes untouched.  Positions and local coordinates are
// evaluated in the reference's order without FMA contraction, so the chosen
// lattice points are bit-identical to a C restatement.
#include <hip/hip_runtime.h>

#include "atdf_device.h"
#include "atdf_internal.h"

namespace atdf {

// ComputeNodeId (structure/HyperTree.cu:1466-1485): a position outside every
// active node gets the first active node and mask 0.
__global__ __launch_bounds__(256) void node_id_kernel(const PNode* __restrict__ nodes,
                                                      const int64_t* __restrict__ active_ids,
                                                      const float* __restrict__ pos, int64_t n,
                                                      int64_t* __restrict__ node_id, float* __restrict__ mask)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float p[3] = {pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]};
    const int32_t nd = active_node_for_position(nodes, p);
    node_id[i] = nd < 0 ? active_ids[0] : (int64_t)nd;
    mask[i] = nd < 0 ? 0.f : 1.f;
}

struct TransferArgs {
    const PNode* src_nodes;      // source active set (boxes shared with the destination)
    const int32_t* src_slot;     // [num_nodes] node -> source grid slot, null = identity
    const float* src_grid;
    int64_t src_slots;
    const int64_t* dst_nodes;    // [n_dst]
    const int32_t* dst_slot;     // [num_nodes] node -> destination slot, null = identity
    float* dst_grid;
    int64_t dst_slots;
    int64_t num_nodes;
    int C, G;
    int layout;                  // 0: [slots, G, G, G, C]; 1: [slots, C, G, G, G]
};

__device__ __forceinline__ int64_t grid_offset(int layout, int64_t slot, int C, int G, int x, int y, int z, int c)
{
    const int64_t g3 = (int64_t)G * G * G;
    const int64_t cell = ((int64_t)z * G + y) * G + x;
    return layout == 0 ? (slot * g3 + cell) * C + c : (slot * C + c) * g3 + cell;
}

// One workgroup per destination node, one lane per lattice point.
__global__ __launch_bounds__(256) void grid_transfer_kernel(TransferArgs a)
{
    const int64_t node = a.dst_nodes[blockIdx.x];
    if (node < 0 || node >= a.num_nodes) return;
    const int64_t dslot = a.dst_slot ? (int64_t)a.dst_slot[node] : node;
    if (dslot < 0 || dslot >= a.dst_slots) return;
    const PNode box = a.src_nodes[node];
    const int G = a.G;
    if (box.flags & kFlagActive) {
        // active in the source set: the node keeps its own grid
        const int64_t sslot = a.src_slot ? (int64_t)a.src_slot[node] : node;
        if (sslot < 0 || sslot >= a.src_slots) return;
        const int64_t rec = (int64_t)G * G * G * a.C;
        for (int64_t i = threadIdx.x; i < rec; i += blockDim.x)
            a.dst_grid[dslot * rec + i] = a.src_grid[sslot * rec + i];
        return;
    }
    const float gm1 = (float)(G - 1);
    float bsize[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) bsize[k] = box.mx[k] - box.mn[k];
    const int cells = G * G * G;
    for (int cell = threadIdx.x; cell < cells; cell += blockDim.x) {
        const int res[3] = {cell % G, (cell / G) % G, cell / (G * G)};
        float g[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float local = (float)res[k] / gm1;
            float c = local * bsize[k];
            g[k] = c + box.mn[k];
        }
        const int32_t target = active_node_for_position(a.src_nodes, g);
        if (target < 0) continue;  // no source node here: the destination keeps its values
        const int64_t sslot = a.src_slot ? (int64_t)a.src_slot[target] : (int64_t)target;
        if (sslot < 0 || sslot >= a.src_slots) continue;
        const PNode tb = a.src_nodes[target];
        int fi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float tsize = tb.mx[k] - tb.mn[k];
            float c = g[k] - tb.mn[k];
            c = c / tsize;
            float f = c * gm1;
            int r = (int)roundf(f);
            fi[k] = r < 0 ? 0 : (r > G - 1 ? G - 1 : r);
        }
        for (int ch = 0; ch < a.C; ++ch)
            a.dst_grid[grid_offset(a.layout, dslot, a.C, G, res[0], res[1], res[2], ch)] =
                a.src_grid[grid_offset(a.layout, sslot, a.C, G, fi[0], fi[1], fi[2], ch)];
    }
}

// NodeNeighborSamples (structure/HyperTree.cu:1720-1811, 2010-2075): for
// every active node, every face (side s, front / back) and every point of a
// size[s_a] x size[s_b] lattice on it (s_a = s + 1, s_b = s + 2 mod 3), the
// point p = local * box + min is located at p + epsilon * (front ? +1 : -1)
// e_s; a different active node there (and, with the ROI test, both nodes in
// the ROI) makes a pair: the point evaluated in both nodes.  Candidates are
// numbered per node as (side, front, lattice index) and pairs are written in
// that order (the reference appends them with an atomic counter).
struct NeighborArgs {
    const PNode* nodes;
    const int64_t* active_ids;
    int64_t num_active;
    int size[3];
    int per_node;  // candidates per node
    float eps;
    int use_roi;
    float roi_min[3], roi_max[3];
    int quad;  // NodeNeighborSamplesQuad: sides x and z, in_roi_quad, four-child descent
    int two_d; // NodeNeighborSamples2D (quad-tree representation) on plane `plane`
    int plane;
};

// in_roi (structure/HyperTree.cu:1667-1690), including its mid = min + max;
// in_roi_quad (HyperTree.cu:1692-1718) widens the y range by widen_y
__device__ __forceinline__ bool node_in_roi(const PNode& n, const float* lo_, const float* hi_, float widen_y)
{
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo = k == 1 ? lo_[k] - widen_y : lo_[k];
        const float hi = k == 1 ? hi_[k] + widen_y : hi_[k];
        const float mid = n.mn[k] + n.mx[k];
        in = in && ((n.mn[k] > lo && n.mn[k] < hi) || (n.mx[k] > lo && n.mx[k] < hi) || (mid > lo && mid < hi));
    }
    return in;
}

// NodeNeighborSamples2D (structure/HyperTree.cu:1815-1913), candidate c of
// node a: on the face local[plane] = 0, tt = 0, 1 takes (side_a, side_b) =
// (s1, s2) then (s2, s1) with (s1, s2) = (plane + 1, plane + 2) mod 3, back
// / front along side_a, size[side_b] points along side_b; the four-child
// descent; in_roi with the ROI's y range widened by eps once for every
// candidate the reference's thread (sid % 128 of 128) visited up to this one,
// in float steps as its loop body does.
__device__ int neighbor_candidate_2d(const NeighborArgs& a, int64_t act, int c, int* node_out, float* g)
{
    constexpr int kThreads = 128;
    const int sidx[2] = {(a.plane + 1) % 3, (a.plane + 2) % 3};
    int tt, front, sid;
    const int n0 = a.size[sidx[1]];  // points per side for tt = 0 (side_b = sidx[1])
    if (c < 2 * n0) {
        tt = 0;
        front = c / n0;
        sid = c % n0;
    } else {
        const int n1 = a.size[sidx[0]];
        const int cc = c - 2 * n0;
        tt = 1;
        front = cc / n1;
        sid = cc % n1;
    }
    const int sa = sidx[tt], sb = sidx[(tt + 1) % 2];
    const int node = (int)a.active_ids[act];
    const PNode nd = a.nodes[node];
    float local[3];
    local[sa] = front ? 1.f : 0.f;
    local[sb] = (float)(sid % a.size[sb]) / ((float)a.size[sb] - 1.f);
    local[a.plane] = 0.f;
    float q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float box = nd.mx[k] - nd.mn[k];
        g[k] = local[k] * box + nd.mn[k];
        q[k] = g[k];
    }
    q[sa] = g[sa] + (front ? a.eps : -a.eps);
    const int other = active_node_for_position(a.nodes, q, 4);
    *node_out = node;
    if (other < 0 || other == node) return -1;
    if (a.use_roi) {
        const int t = sid % kThreads;
        int m = 0;
        for (int tp = 0; tp <= tt; ++tp)
            for (int fp = 0; fp < 2; ++fp) {
                if (tp == tt && fp > front) break;
                const int nb = a.size[sidx[(tp + 1) % 2]];
                const int upto = (tp == tt && fp == front) ? sid : nb - 1;
                for (int s2 = t; s2 <= upto && s2 < nb; s2 += kThreads) ++m;
            }
        float lo[3] = {a.roi_min[0], a.roi_min[1], a.roi_min[2]};
        float hi[3] = {a.roi_max[0], a.roi_max[1], a.roi_max[2]};
        for (int i = 0; i < m; ++i) {
            lo[1] = lo[1] - a.eps;
            hi[1] = hi[1] + a.eps;
        }
        if (!(node_in_roi(nd, lo, hi, 0.f) && node_in_roi(a.nodes[other], lo, hi, 0.f))) return -1;
    }
    return other;
}

// candidate c of node a: returns the partner node (or -1) and the point
__device__ int neighbor_candidate(const NeighborArgs& a, int64_t act, int c, int* node_out, float* g)
{
    if (a.two_d) return neighbor_candidate_2d(a, act, c, node_out, g);
    int side = 0, front = 0, sid = c;
    for (side = 0; side < 3; side += a.quad ? 2 : 1) {
        const int sa = (side + 1) % 3, sb = (side + 2) % 3;
        const int n = a.size[sa] * a.size[sb];
        if (sid < 2 * n) {
            front = sid >= n ? 1 : 0;
            sid -= front * n;
            break;
        }
        sid -= 2 * n;
    }
    const int sa = (side + 1) % 3, sb = (side + 2) % 3;
    const int node = (int)a.active_ids[act];
    const PNode nd = a.nodes[node];
    float local[3];
    local[side] = front ? 1.f : 0.f;
    local[sa] = (float)(sid % a.size[sa]) / ((float)a.size[sa] - 1.f);
    local[sb] = (float)(sid / a.size[sa]) / ((float)a.size[sb] - 1.f);
    float q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float box = nd.mx[k] - nd.mn[k];
        g[k] = local[k] * box + nd.mn[k];
        q[k] = g[k];
    }
    q[side] = g[side] + (front ? a.eps : -a.eps);
    const int other = active_node_for_position(a.nodes, q, a.quad ? 4 : 8);
    *node_out = node;
    if (other < 0 || other == node) return -1;
    const float widen = a.quad ? a.eps / 2.f : 0.f;
    if (a.use_roi &&
        !(node_in_roi(nd, a.roi_min, a.roi_max, widen) && node_in_roi(a.nodes[other], a.roi_min, a.roi_max, widen)))
        return -1;
    return other;
}

__global__ __launch_bounds__(256) void neighbor_flags_kernel(NeighborArgs a, int32_t* __restrict__ flags)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.num_active * a.per_node) return;
    int node;
    float g[3];
    flags[i] = neighbor_candidate(a, i / a.per_node, (int)(i % a.per_node), &node, g) >= 0 ? 1 : 0;
}

__global__ __launch_bounds__(256) void neighbor_write_kernel(NeighborArgs a, const int64_t* __restrict__ offsets,
                                                             float* __restrict__ pos, int64_t* __restrict__ node_id)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.num_active * a.per_node) return;
    if (offsets[i + 1] == offsets[i]) return;
    int node;
    float g[3];
    const int other = neighbor_candidate(a, i / a.per_node, (int)(i % a.per_node), &node, g);
    const int64_t o = 2 * offsets[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pos[o * 3 + k] = g[k];
        pos[(o + 1) * 3 + k] = g[k];
    }
    node_id[o] = node;
    node_id[o + 1] = other;
}

}  // namespace atdf

using namespace atdf;

extern "C" {

atdf_status atdf_node_id_for_position(const void* packed_tree, int64_t num_nodes, const int64_t* active_node_ids,
                                      const float* positions, int64_t num_positions, int64_t* node_id, float* mask,
                                      atdf_stream_t stream)
{
    ATDF_REQUIRE(packed_tree && num_nodes > 0, "atdf_node_id_for_position: null tree");
    ATDF_REQUIRE(num_positions >= 0, "atdf_node_id_for_position: negative count");
    if (num_positions == 0) return ATDF_OK;
    ATDF_REQUIRE(active_node_ids && positions && node_id && mask, "atdf_node_id_for_position: null buffer");
    hipLaunchKernelGGL(node_id_kernel, dim3((unsigned)div_up(num_positions, 256)), dim3(256), 0, to_hip(stream),
                       reinterpret_cast<const PNode*>(packed_tree), active_node_ids, positions, num_positions, node_id,
                       mask);
    return check_launch("atdf_node_id_for_position");
}

atdf_status atdf_grid_transfer(const void* src_packed_tree, int64_t num_nodes, const int32_t* src_slot,
                               const float* src_grid, int64_t src_slots, const int64_t* dst_nodes, int64_t num_dst,
                               const int32_t* dst_slot, float* dst_grid, int64_t dst_slots, int32_t C, int32_t G,
                               int32_t layout, atdf_stream_t stream)
{
    ATDF_REQUIRE(src_packed_tree && num_nodes > 0, "atdf_grid_transfer: null tree");
    ATDF_REQUIRE(C >= 1 && G >= 2 && (layout == 0 || layout == 1), "atdf_grid_transfer: bad C %d, G %d or layout %d",
                 C, G, layout);
    ATDF_REQUIRE(num_dst >= 0 && src_slots >= 0 && dst_slots >= 0, "atdf_grid_transfer: negative size");
    ATDF_REQUIRE(num_dst < (int64_t)INT32_MAX, "atdf_grid_transfer: too many destination nodes");
    if (num_dst == 0 || dst_slots == 0 || src_slots == 0) return ATDF_OK;
    ATDF_REQUIRE(src_grid && dst_grid && dst_nodes, "atdf_grid_transfer: null buffer");
    ATDF_REQUIRE(src_grid != dst_grid, "atdf_grid_transfer: source and destination grids alias");
    TransferArgs a;
    a.src_nodes = reinterpret_cast<const PNode*>(src_packed_tree);
    a.src_slot = src_slot;
    a.src_grid = src_grid;
    a.src_slots = src_slots;
    a.dst_nodes = dst_nodes;
    a.dst_slot = dst_slot;
    a.dst_grid = dst_grid;
    a.dst_slots = dst_slots;
    a.num_nodes = num_nodes;
    a.C = C;
    a.G = G;
    a.layout = layout;
    hipLaunchKernelGGL(grid_transfer_kernel, dim3((unsigned)num_dst), dim3(256), 0, to_hip(stream), a);
    return check_launch("atdf_grid_transfer");
}

static atdf_status make_neighbor_args(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                      int64_t num_active, const atdf_neighbor_spec* spec, NeighborArgs* a,
                                      const char* who)
{
    ATDF_REQUIRE(packed_tree && num_nodes > 0 && spec, "%s: empty tree or spec", who);
    ATDF_REQUIRE(num_active >= 0, "%s: negative node count", who);
    for (int k = 0; k < 3; ++k)
        ATDF_REQUIRE(spec->size[k] >= 2 && spec->size[k] <= 4096, "%s: face lattice size %d outside [2, 4096]", who,
                     spec->size[k]);
    a->nodes = reinterpret_cast<const PNode*>(packed_tree);
    a->active_ids = active_ids;
    a->num_active = num_active;
    for (int k = 0; k < 3; ++k) {
        a->size[k] = spec->size[k];
        a->roi_min[k] = spec->roi_min[k];
        a->roi_max[k] = spec->roi_max[k];
    }
    ATDF_REQUIRE(spec->variant >= 0 && spec->variant <= 2, "%s: variant %d (0 NodeNeighborSamples, 1 "
                 "NodeNeighborSamplesQuad, 2 NodeNeighborSamples2D)", who, spec->variant);
    ATDF_REQUIRE(spec->variant != 2 || (spec->plane >= 0 && spec->plane <= 2), "%s: plane %d outside [0, 2]", who,
                 spec->plane);
    a->quad = spec->variant == 1;
    a->two_d = spec->variant == 2;
    a->plane = spec->plane;
    a->per_node = (int)atdf_neighbor_candidates(1, spec);
    a->eps = spec->epsilon;
    a->use_roi = spec->use_roi != 0;
    ATDF_REQUIRE(num_active * a->per_node < (int64_t)INT32_MAX, "%s: too many candidates", who);
    return ATDF_OK;
}

int64_t atdf_neighbor_candidates(int64_t num_active, const atdf_neighbor_spec* spec)
{
    if (!spec || num_active <= 0) return 0;
    if (spec->variant == 2) {
        if (spec->plane < 0 || spec->plane > 2) return 0;
        return num_active * 2 * ((int64_t)spec->size[(spec->plane + 1) % 3] + spec->size[(spec->plane + 2) % 3]);
    }
    const int64_t x = (int64_t)spec->size[1] * spec->size[2], y = (int64_t)spec->size[2] * spec->size[0],
                  z = (int64_t)spec->size[0] * spec->size[1];
    return num_active * 2 * (spec->variant == 1 ? x + z : x + y + z);
}

atdf_status atdf_neighbor_flags(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                int64_t num_active, const atdf_neighbor_spec* spec, int32_t* flags,
                                atdf_stream_t stream)
{
    NeighborArgs a{};
    atdf_status st = make_neighbor_args(packed_tree, num_nodes, active_ids, num_active, spec, &a,
                                        "atdf_neighbor_flags");
    if (st) return st;
    const int64_t n = num_active * a.per_node;
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(active_ids && flags, "atdf_neighbor_flags: null buffer");
    hipLaunchKernelGGL(neighbor_flags_kernel, dim3((unsigned)div_up(n, 256)), dim3(256), 0, to_hip(stream), a, flags);
    return check_launch("atdf_neighbor_flags");
}

atdf_status atdf_neighbor_write(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                int64_t num_active, const atdf_neighbor_spec* spec, const int64_t* offsets,
                                float* positions, int64_t* node_id, atdf_stream_t stream)
{
    NeighborArgs a{};
    atdf_status st = make_neighbor_args(packed_tree, num_nodes, active_ids, num_active, spec, &a,
                                        "atdf_neighbor_write");
    if (st) return st;
    const int64_t n = num_active * a.per_node;
    if (n == 0) return ATDF_OK;
    ATDF_REQUIRE(active_ids && offsets && positions && node_id, "atdf_neighbor_write: null buffer");
    hipLaunchKernelGGL(neighbor_write_kernel, dim3((unsigned)div_up(n, 256)), dim3(256), 0, to_hip(stream), a,
                       offsets, positions, node_id);
    return check_launch("atdf_neighbor_write");
}

}  // extern "C"

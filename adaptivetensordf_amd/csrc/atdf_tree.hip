// atdf_tree.hip -- octree packing, ray traversal with per-node sample
// placement, and the exclusive scan that turns per-ray counts into offsets.
//
// Traversal is one lane per ray walking the octree depth-first in
// front-to-back child order (child c is visited at position c ^ mask, mask
// = sign bits of the direction), stackless: the return path uses the
// parent links and the contiguous child numbering.  A child box can only
// be hit when its parent box is hit (the rounded slab values of a child
// are nested inside the parent's), so pruning on the parent test visits
// exactly the active nodes the reference's brute-force loop
// (structure/HyperTree.cu:320-359) would hit.  Front-to-back order emits
// hits in increasing tmin; the kernel checks that order for every ray and
// reports a violation instead of silently emitting a different order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "atdf_device.h"
#include "atdf_internal.h"

namespace atdf {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// tree packing
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pack_nodes_kernel(const float* __restrict__ pmin, const float* __restrict__ pmax,
                                                         const int32_t* __restrict__ children,
                                                         const int32_t* __restrict__ parent,
                                                         const int32_t* __restrict__ active, int64_t n, int ns,
                                                         PNode* out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    PNode p;
    for (int k = 0; k < 3; ++k) {
        p.mn[k] = pmin[i * 3 + k];
        p.mx[k] = pmax[i * 3 + k];
    }
    p.first_child = children[i * ns];
    p.parent = parent[i];
    p.flags = (active[i] ? kFlagActive : 0) | (ns == 4 ? kFlagQuad : 0);
    p.first_sib = p.parent >= 0 ? children[(int64_t)p.parent * ns] : 0;
    p.pad[0] = p.pad[1] = 0;
    out[i] = p;
}

__global__ __launch_bounds__(256) void mark_active_below_kernel(const int32_t* __restrict__ parent,
                                                                const int32_t* __restrict__ active, int64_t n,
                                                                PNode* out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || !active[i]) return;
    int32_t p = parent[i];
    while (p >= 0) {
        int old = atomicOr(&out[p].flags, kFlagActiveBelow);
        if (old & kFlagActiveBelow) break;  // whoever set it walks further up
        p = parent[p];
    }
}

// ---------------------------------------------------------------------------
// traversal
// ---------------------------------------------------------------------------
struct TraverseArgs {
    const PNode* nodes;
    const float* O;
    const float* D;
    int64_t R;
    int K;
    int max_inter;
    int jitter_mode;
    uint64_t seed;
    const float* jitter;
    const int32_t* select;
    const int32_t* list;
    int32_t* ray_num_intervals;
    int32_t* ray_num_samples;
    int32_t* status;
    const int64_t* iv_off;
    const int64_t* smp_off;
    atdf_intervals iv;
    atdf_samples smp;
    // hit cache (dense ray lists only): the count pass records each ray's
    // hit count and its first cache_slots hits (slot-major, so a wave's
    // records are contiguous); the write pass writes a ray whose hits all
    // fit from the cache instead of walking the tree again
    int32_t* cache_n;  // [R]; more than cache_slots: walk the ray again
    int4* cache;       // [cache_slots][R]: node, tmin bits, tmax bits, samples | first sample << 12
    int cache_slots;
};

__device__ __forceinline__ float jitter_of(const TraverseArgs& a, int64_t slot, int64_t ray, int64_t loc)
{
    if (a.jitter_mode == 1) return a.jitter[slot];
    if (a.jitter_mode == 2) return hash_jitter(a.seed, ray, loc);
    return 0.5f;
}

// The write pass's outputs for one hit node (interval and, when asked, its
// samples), the nh-th hit of ray r whose samples start at ns.
__device__ __forceinline__ void write_hit(const TraverseArgs& a, int64_t r, const float* o, const float* d, int node,
                                          float tmin, float tmax, int n, int64_t ivb, int64_t sb, int nh, int ns)
{
    if (a.iv.node) {
        int64_t q = ivb + nh;
        a.iv.node[q] = node;
        a.iv.tmin[q] = tmin;
        a.iv.tmax[q] = tmax;
        a.iv.count[q] = n;
        a.iv.first[q] = ns;
        a.iv.ray[q] = (int32_t)r;
    }
    if (a.smp.weight && n > 0) {
        float step = (tmax - tmin) / (float)n;
        int64_t slot0 = sb + ns;
        float t_prev = 0.f;
        float t_cur = sample_t(tmin, step, 0, jitter_of(a, slot0, r, ns));
        for (int j = 0; j < n; ++j) {
            float t_next = 0.f;
            if (j < n - 1) t_next = sample_t(tmin, step, j + 1, jitter_of(a, slot0 + j + 1, r, ns + j + 1));
            int64_t slot = slot0 + j;
            a.smp.ray_t[slot] = t_cur;
            a.smp.weight[slot] = sample_weight(tmin, tmax, n, j, t_prev, t_cur, t_next);
            for (int k = 0; k < 3; ++k) a.smp.global_coordinate[slot * 3 + k] = o[k] + d[k] * t_cur;
            a.smp.ray_index[slot] = r;
            a.smp.local_index_in_ray[slot] = ns + j;
            a.smp.node_id[slot] = node;
            t_prev = t_cur;
            t_cur = t_next;
        }
    }
}

// A cache record packs a hit's sample count (12 bits) and its first sample
// within the ray (20 bits); a ray with a hit beyond either is walked again.
constexpr int kCacheMaxCount = 1 << 12;
constexpr int kCacheMaxFirst = 1 << 20;

template <bool kWrite>
__global__ __launch_bounds__(256) void traverse_kernel(TraverseArgs a)
{
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= a.R) return;
    if (a.list) {
        r = a.list[r];
        if (r < 0 || r >= a.R) return;
    }
    if (a.select && a.select[r] == 0) {
        if (!kWrite) {
            a.ray_num_intervals[r] = 0;
            a.ray_num_samples[r] = 0;
        }
        return;
    }
    float o[3], d[3];
    for (int k = 0; k < 3; ++k) {
        o[k] = a.O[r * 3 + k];
        d[k] = a.D[r * 3 + k];
    }
    const bool quad = (a.nodes[0].flags & kFlagQuad) != 0;
    const int mask = near_child_mask(d, quad);
    const int last_sibling = quad ? 3 : 7;

    int nh = 0, ns = 0;
    float last_t = -INFINITY;
    int last_node = -1;
    bool order_bad = false, cache_miss = false;
    const int64_t ivb = kWrite ? a.iv_off[r] : 0;
    const int64_t sb = kWrite ? a.smp_off[r] : 0;
    if (kWrite && a.cache) {
        const int nc = a.cache_n[r];
        if (nc <= a.cache_slots) {
            for (int q = 0; q < nc; ++q) {
                const int4 e = a.cache[(int64_t)q * a.R + r];
                const int n = e.w & (kCacheMaxCount - 1);
                write_hit(a, r, o, d, e.x, __int_as_float(e.y), __int_as_float(e.z), n, ivb, sb, nh, ns);
                nh++;
                ns += n;
            }
            return;
        }
    }

    int node = 0;
    while (true) {
        const PNode nd = a.nodes[node];
        float tmin, tmax;
        bool hit = intersect_box_ray(nd.mn, nd.mx, o, d, tmin, tmax);
        if (hit && (nd.flags & kFlagActive) && !(tmax - tmin < 1e-5)) {
            int n = node_sample_count(nd.mn, nd.mx, tmin, tmax, a.K);
            if (tmin < last_t || (tmin == last_t && node < last_node)) order_bad = true;
            last_t = tmin;
            last_node = node;
            if (kWrite) write_hit(a, r, o, d, node, tmin, tmax, n, ivb, sb, nh, ns);
            else if (a.cache && nh < a.cache_slots) {
                if (n < kCacheMaxCount && ns < kCacheMaxFirst)
                    a.cache[(int64_t)nh * a.R + r] =
                        make_int4(node, __float_as_int(tmin), __float_as_int(tmax), n | (ns << 12));
                else
                    cache_miss = true;
            }
            nh++;
            ns += n;
        }
        if (hit && (nd.flags & kFlagActiveBelow) && nd.first_child >= 0) {
            node = nd.first_child + mask;  // first child in front-to-back order
            continue;
        }
        // next sibling in front-to-back order, ascending while exhausted
        // (a node carries its parent's first child, so a sibling step needs
        // no load; each level ascended loads the parent's record once)
        int cur = node;
        int parent = nd.parent;
        int first = nd.first_sib;
        while (true) {
            if (parent < 0) goto done;
            int k = (cur - first) ^ mask;
            if (k < last_sibling) {
                node = first + ((k + 1) ^ mask);
                break;
            }
            cur = parent;
            const PNode& pn = a.nodes[parent];
            first = pn.first_sib;
            parent = pn.parent;
        }
    }
done:
    if (!kWrite) {
        a.ray_num_intervals[r] = nh;
        a.ray_num_samples[r] = ns;
        if (a.cache) a.cache_n[r] = cache_miss ? a.cache_slots + 1 : nh;
        if (nh > a.max_inter) atomicAdd(&a.status[0], 1);
        if (order_bad) atomicAdd(&a.status[1], 1);
        atomicMax(&a.status[2], ns);
    }
}

// Write pass of an interval list from the hit cache, coalesced: a wave owns
// 64 consecutive rays, whose intervals form one contiguous output range, and
// its lanes write that range in order (lane l the l-th of each 64 outputs),
// each finding its output's ray by a binary search over the rays' offsets.
// A ray that was not cached is walked by its own lane afterwards.  Same
// values as traverse_kernel<true>.
__global__ __launch_bounds__(256) void traverse_write_cached_kernel(TraverseArgs a)
{
    const int lane = threadIdx.x & 63;
    const int64_t r0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x - lane);
    if (r0 >= a.R) return;  // whole waves
    const int64_t r = r0 + lane;
    const bool valid = r < a.R;
    const int64_t beg = a.iv_off[valid ? r : a.R];
    const int64_t end = valid ? a.iv_off[r + 1] : beg;
    const bool selected = valid && !(a.select && a.select[r] == 0);
    const bool walk = selected && end > beg && a.cache_n[r] > a.cache_slots;
    const int64_t wbeg = __shfl(beg, 0, 64);
    const int64_t wend = a.iv_off[r0 + 64 < a.R ? r0 + 64 : a.R];
    const unsigned long long walkers = __ballot(walk);
    for (int64_t q0 = wbeg; q0 < wend; q0 += 64) {
        const int64_t q = q0 + lane;
        // the last lane whose range starts at or before q (empty ranges
        // share their start with the next range, so this one holds q)
        int L = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const int64_t b = __shfl(beg, L + step, 64);
            if (b <= q) L += step;
        }
        const int64_t bL = __shfl(beg, L, 64);
        if (q < wend && !((walkers >> L) & 1ull)) {
            const int4 e = a.cache[(int64_t)(q - bL) * a.R + r0 + L];
            a.iv.node[q] = e.x;
            a.iv.tmin[q] = __int_as_float(e.y);
            a.iv.tmax[q] = __int_as_float(e.z);
            a.iv.count[q] = e.w & (kCacheMaxCount - 1);
            a.iv.first[q] = (int32_t)((uint32_t)e.w >> 12);
            a.iv.ray[q] = (int32_t)(r0 + L);
        }
    }
    if (!walk) return;
    // a ray whose hits did not fit the cache: the walk of traverse_kernel
    float o[3], d[3];
    for (int k = 0; k < 3; ++k) {
        o[k] = a.O[r * 3 + k];
        d[k] = a.D[r * 3 + k];
    }
    const bool quad = (a.nodes[0].flags & kFlagQuad) != 0;
    const int mask = near_child_mask(d, quad);
    const int last_sibling = quad ? 3 : 7;
    int nh = 0, ns = 0;
    int node = 0;
    while (true) {
        const PNode nd = a.nodes[node];
        float tmin, tmax;
        bool hit = intersect_box_ray(nd.mn, nd.mx, o, d, tmin, tmax);
        if (hit && (nd.flags & kFlagActive) && !(tmax - tmin < 1e-5)) {
            const int n = node_sample_count(nd.mn, nd.mx, tmin, tmax, a.K);
            write_hit(a, r, o, d, node, tmin, tmax, n, beg, 0, nh, ns);
            nh++;
            ns += n;
        }
        if (hit && (nd.flags & kFlagActiveBelow) && nd.first_child >= 0) {
            node = nd.first_child + mask;
            continue;
        }
        int cur = node;
        int parent = nd.parent;
        int first = nd.first_sib;
        while (true) {
            if (parent < 0) return;
            int k = (cur - first) ^ mask;
            if (k < last_sibling) {
                node = first + ((k + 1) ^ mask);
                break;
            }
            cur = parent;
            const PNode& pn = a.nodes[parent];
            first = pn.first_sib;
            parent = pn.parent;
        }
    }
}

// ---------------------------------------------------------------------------
// Cooperative traversal for sparse ray lists (the learned proposal's fallback
// rays): eight lanes per ray.  The walk is the same front-to-back DFS as
// traverse_kernel, but the eight children of a node are loaded and
// slab-tested at once, one per lane, and kept in LDS (one row of eight per
// tree level), so a ray's chain of dependent node loads is one per visited
// inner node instead of one per visited node.  Node tests, emission order and
// every written value are those of traverse_kernel.
// ---------------------------------------------------------------------------
constexpr int kCoopThreads = 128;
constexpr int kCoopGroups = kCoopThreads / 8;
constexpr int kCoopLevels = 12;

struct CoopChild {
    float tmin, tmax;
    int32_t first_child;
    int32_t flags;  // node flags; bit 8: hit
};

__device__ __forceinline__ void coop_emit(const TraverseArgs& a, bool write, int gl, int64_t r, int node, float tmin,
                                          float tmax, const PNode& nd, int64_t ivb, int& nh, int& ns, float& last_t,
                                          int& last_node, bool& order_bad)
{
    const int n = node_sample_count(nd.mn, nd.mx, tmin, tmax, a.K);
    if (tmin < last_t || (tmin == last_t && node < last_node)) order_bad = true;
    last_t = tmin;
    last_node = node;
    if (write && gl == 0) {
        const int64_t q = ivb + nh;
        a.iv.node[q] = node;
        a.iv.tmin[q] = tmin;
        a.iv.tmax[q] = tmax;
        a.iv.count[q] = n;
        a.iv.first[q] = ns;
        a.iv.ray[q] = (int32_t)r;
    }
    nh++;
    ns += n;
}

template <bool kWrite>
__global__ __launch_bounds__(kCoopThreads) void traverse_coop_kernel(TraverseArgs a)
{
    __shared__ CoopChild s_child[kCoopGroups][kCoopLevels][8];
    __shared__ int32_t s_first[kCoopGroups][kCoopLevels];
    __shared__ int32_t s_bits[kCoopGroups][kCoopLevels];
    const int gi = threadIdx.x >> 3, gl = threadIdx.x & 7;
    int64_t r = (blockIdx.x * (int64_t)kCoopThreads + threadIdx.x) >> 3;
    if (r >= a.R) return;
    if (a.list) {
        r = a.list[r];
        if (r < 0 || r >= a.R) return;
    }
    if (a.select && a.select[r] == 0) {
        if (!kWrite && gl == 0) {
            a.ray_num_intervals[r] = 0;
            a.ray_num_samples[r] = 0;
        }
        return;
    }
    float o[3], d[3];
    for (int k = 0; k < 3; ++k) {
        o[k] = a.O[r * 3 + k];
        d[k] = a.D[r * 3 + k];
    }
    const bool quad = (a.nodes[0].flags & kFlagQuad) != 0;
    const int mask = near_child_mask(d, quad);
    const int nchild = quad ? 4 : 8;
    int nh = 0, ns = 0;
    float last_t = -INFINITY;
    int last_node = -1;
    bool order_bad = false, overflow = false;
    const int64_t ivb = kWrite ? a.iv_off[r] : 0;
    CoopChild(*ch)[8] = s_child[gi];
    int32_t* first = s_first[gi];
    int32_t* bits = s_bits[gi];
    int level = -1;

    // push the children of a node: lane gl tests the child at front-to-back
    // position gl (lanes 4-7 idle in the quad-tree representation)
    auto push = [&](int32_t fc) {
        if (level + 1 >= kCoopLevels) {
            overflow = true;
            return;
        }
        ++level;
        bool hit = false;
        if (gl < nchild) {
            const int child = fc + (gl ^ mask);
            const PNode c = a.nodes[child];
            float t0, t1;
            hit = intersect_box_ray(c.mn, c.mx, o, d, t0, t1);
            ch[level][gl] = CoopChild{t0, t1, c.first_child, c.flags | (hit ? 256 : 0)};
        }
        int b = hit ? (1 << gl) : 0;
        b |= __shfl_xor(b, 1, 8);
        b |= __shfl_xor(b, 2, 8);
        b |= __shfl_xor(b, 4, 8);
        if (gl == 0) {
            first[level] = fc;
            bits[level] = b;
        }
    };

    {
        const PNode nd = a.nodes[0];
        float tmin, tmax;
        const bool hit = intersect_box_ray(nd.mn, nd.mx, o, d, tmin, tmax);
        if (hit && (nd.flags & kFlagActive) && !(tmax - tmin < 1e-5))
            coop_emit(a, kWrite, gl, r, 0, tmin, tmax, nd, ivb, nh, ns, last_t, last_node, order_bad);
        if (hit && (nd.flags & kFlagActiveBelow) && nd.first_child >= 0) push(nd.first_child);
    }
    while (level >= 0 && !overflow) {
        const int b = bits[level];
        if (b == 0) {
            --level;
            continue;
        }
        const int k = __ffs(b) - 1;
        if (gl == 0) bits[level] = b & (b - 1);
        const CoopChild c = ch[level][k];
        const int node = first[level] + (k ^ mask);
        if ((c.flags & kFlagActive) && !(c.tmax - c.tmin < 1e-5)) {
            const PNode nd = a.nodes[node];
            coop_emit(a, kWrite, gl, r, node, c.tmin, c.tmax, nd, ivb, nh, ns, last_t, last_node, order_bad);
        }
        if ((c.flags & kFlagActiveBelow) && c.first_child >= 0) push(c.first_child);
    }
    if (!kWrite && gl == 0) {
        a.ray_num_intervals[r] = nh;
        a.ray_num_samples[r] = ns;
        if (nh > a.max_inter) atomicAdd(&a.status[0], 1);
        if (order_bad || overflow) atomicAdd(&a.status[1], 1);
        atomicMax(&a.status[2], ns);
    }
}

// ---------------------------------------------------------------------------
// exclusive scan, int32 -> int64 (n + 1 outputs)
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanTile = kScanThreads * kScanPerThread;

__device__ __forceinline__ int64_t wave_inclusive_scan(int64_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int64_t u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Inclusive scan of one value per thread across the block; returns the
// thread's inclusive prefix and writes the block total to *total.
__device__ __forceinline__ int64_t block_inclusive_scan(int64_t v, int64_t* smem, int64_t* total)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int64_t inc = wave_inclusive_scan(v);
    if (lane == 63) smem[wave] = inc;
    __syncthreads();
    int64_t off = 0;
    int64_t tot = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wave) off += smem[w];
        tot += smem[w];
    }
    __syncthreads();
    *total = tot;
    return inc + off;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_sums_kernel(const int32_t* __restrict__ in, int64_t n,
                                                                      int64_t* __restrict__ tile_sums)
{
    __shared__ int64_t smem[kScanThreads / 64];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    int64_t s = 0;
    for (int k = 0; k < kScanPerThread; ++k)
        if (base + k < n) s += in[base + k];
    int64_t tot;
    block_inclusive_scan(s, smem, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Single block: in-place exclusive scan of the tile sums.
__global__ __launch_bounds__(kScanThreads) void scan_tile_offsets_kernel(int64_t* __restrict__ tile_sums, int64_t nt)
{
    __shared__ int64_t smem[kScanThreads / 64];
    int64_t carry = 0;
    for (int64_t base = 0; base < nt; base += kScanThreads) {
        int64_t i = base + threadIdx.x;
        int64_t v = (i < nt) ? tile_sums[i] : 0;
        int64_t tot;
        int64_t inc = block_inclusive_scan(v, smem, &tot);
        if (i < nt) tile_sums[i] = carry + inc - v;
        carry += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const int32_t* __restrict__ in, int64_t n,
                                                                  const int64_t* __restrict__ tile_offsets,
                                                                  int64_t* __restrict__ out)
{
    __shared__ int64_t smem[kScanThreads / 64];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    int64_t v[kScanPerThread];
    int64_t s = 0;
    for (int k = 0; k < kScanPerThread; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int64_t tot;
    int64_t inc = block_inclusive_scan(s, smem, &tot);
    int64_t run = tile_offsets[blockIdx.x] + inc - s;
    for (int k = 0; k < kScanPerThread; ++k) {
        run += v[k];
        if (base + k < n) out[base + k + 1] = run;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

__global__ void zero_i64_kernel(int64_t* p)
{
    if (threadIdx.x == 0) p[0] = 0;
}

// ---------------------------------------------------------------------------
// ray-list compaction: one counter atomic per wave
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fill_list_kernel(int32_t* __restrict__ list, int64_t n,
                                                        int32_t* __restrict__ count)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i == 0) *count = 0;
    if (i < n) list[i] = -1;
}

__global__ __launch_bounds__(256) void compact_kernel(const int32_t* __restrict__ flags, int64_t n,
                                                      int32_t* __restrict__ list, int32_t* __restrict__ count)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool f = i < n && flags[i] != 0;
    const unsigned long long m = __ballot(f);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, 0, 64);
    if (f) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)i;
}

}  // namespace atdf

using namespace atdf;

extern "C" {

const char* atdf_last_error(void) { return g_last_error; }
const char* atdf_version(void) { return "atdf 0.1 gfx950"; }

size_t atdf_tree_packed_bytes(int64_t num_nodes) { return (size_t)(num_nodes > 0 ? num_nodes : 0) * sizeof(PNode); }

atdf_status atdf_tree_pack(const atdf_tree* t, void* packed, atdf_stream_t stream)
{
    ATDF_REQUIRE(t && packed, "atdf_tree_pack: null argument");
    ATDF_REQUIRE(t->num_nodes > 0 && t->num_nodes < (int64_t)INT32_MAX, "atdf_tree_pack: num_nodes %lld out of range",
                 (long long)t->num_nodes);
    ATDF_REQUIRE(t->node_position_min && t->node_position_max && t->node_children && t->node_parent &&
                     t->node_active,
                 "atdf_tree_pack: null tree tensor");
    const int ns = t->children_per_node == 0 ? 8 : t->children_per_node;
    ATDF_REQUIRE(ns == 8 || ns == 4, "atdf_tree_pack: children_per_node %d (8: octree, 4: quad tree)",
                 t->children_per_node);
    hipStream_t s = to_hip(stream);
    int64_t n = t->num_nodes;
    dim3 grid((unsigned)div_up(n, 256));
    PNode* out = reinterpret_cast<PNode*>(packed);
    hipLaunchKernelGGL(pack_nodes_kernel, grid, dim3(256), 0, s, t->node_position_min, t->node_position_max,
                       t->node_children, t->node_parent, t->node_active, n, ns, out);
    hipLaunchKernelGGL(mark_active_below_kernel, grid, dim3(256), 0, s, t->node_parent, t->node_active, n, out);
    return check_launch("atdf_tree_pack");
}

__global__ void pack_box_kernel(float x0, float y0, float z0, float x1, float y1, float z1, PNode* out)
{
    if (threadIdx.x != 0) return;
    PNode p;
    p.mn[0] = x0;
    p.mn[1] = y0;
    p.mn[2] = z0;
    p.mx[0] = x1;
    p.mx[1] = y1;
    p.mx[2] = z1;
    p.first_child = -1;
    p.parent = -1;
    p.flags = kFlagActive;
    p.first_sib = 0;
    p.pad[0] = p.pad[1] = 0;
    out[0] = p;
}

atdf_status atdf_tree_pack_box(const float* roi_min, const float* roi_max, void* packed, atdf_stream_t stream)
{
    ATDF_REQUIRE(roi_min && roi_max && packed, "atdf_tree_pack_box: null argument");
    for (int k = 0; k < 3; ++k)
        ATDF_REQUIRE(std::isfinite(roi_min[k]) && std::isfinite(roi_max[k]) && roi_min[k] < roi_max[k],
                     "atdf_tree_pack_box: axis %d: roi_min %g must be below roi_max %g", k, (double)roi_min[k],
                     (double)roi_max[k]);
    hipLaunchKernelGGL(pack_box_kernel, dim3(1), dim3(64), 0, to_hip(stream), roi_min[0], roi_min[1], roi_min[2],
                       roi_max[0], roi_max[1], roi_max[2], reinterpret_cast<PNode*>(packed));
    return check_launch("atdf_tree_pack_box");
}

static atdf_status check_sampling(const atdf_sampling* sp)
{
    ATDF_REQUIRE(sp, "sampling parameters are null");
    ATDF_REQUIRE(sp->max_samples_per_node >= 1, "max_samples_per_node must be >= 1 (got %d)", sp->max_samples_per_node);
    ATDF_REQUIRE(sp->max_intersections >= 1, "max_intersections must be >= 1");
    ATDF_REQUIRE(sp->jitter_mode >= 0 && sp->jitter_mode <= 2, "jitter_mode must be 0, 1 or 2");
    return ATDF_OK;
}

// Hit cache layout: [R] hit counts (padded to 16 B), then [slots][R] records.
static void cache_views(void* cache, int64_t R, int32_t*& n, int4*& rec)
{
    n = reinterpret_cast<int32_t*>(cache);
    rec = reinterpret_cast<int4*>(reinterpret_cast<unsigned char*>(cache) + ((size_t)R * 4 + 15) / 16 * 16);
}

static atdf_status traverse_count_impl(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                       const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                       int32_t* ray_num_intervals, int32_t* ray_num_samples, int32_t* status,
                                       void* cache, int32_t cache_slots, atdf_stream_t stream)
{
    atdf_status st = check_sampling(sampling);
    if (st) return st;
    ATDF_REQUIRE(packed_tree && num_nodes > 0, "atdf_traverse_count: empty tree");
    ATDF_REQUIRE(num_rays >= 0, "atdf_traverse_count: negative ray count");
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(ray_origin && ray_direction && ray_num_intervals && ray_num_samples && status,
                 "atdf_traverse_count: null buffer");
    TraverseArgs a{};
    a.nodes = reinterpret_cast<const PNode*>(packed_tree);
    a.O = ray_origin;
    a.D = ray_direction;
    a.R = num_rays;
    a.K = sampling->max_samples_per_node;
    a.max_inter = sampling->max_intersections;
    a.jitter_mode = sampling->jitter_mode;
    a.seed = sampling->jitter_seed;
    a.jitter = sampling->jitter;
    a.select = sampling->ray_select;
    a.list = sampling->ray_list;
    a.ray_num_intervals = ray_num_intervals;
    a.ray_num_samples = ray_num_samples;
    a.status = status;
    if (cache && !a.list) {
        cache_views(cache, num_rays, a.cache_n, a.cache);
        a.cache_slots = cache_slots;
    }
    if (a.list)  // a sparse selection: eight lanes per listed ray
        hipLaunchKernelGGL(traverse_coop_kernel<false>, dim3((unsigned)div_up(num_rays * 8, kCoopThreads)),
                           dim3(kCoopThreads), 0, to_hip(stream), a);
    else
        hipLaunchKernelGGL(traverse_kernel<false>, dim3((unsigned)div_up(num_rays, 256)), dim3(256), 0,
                           to_hip(stream), a);
    return check_launch("atdf_traverse_count");
}

static atdf_status traverse_write_impl(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                       const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                       const int64_t* ray_interval_offset, const int64_t* ray_sample_offset,
                                       const atdf_intervals* intervals, const atdf_samples* samples,
                                       const void* cache, int32_t cache_slots, atdf_stream_t stream)
{
    atdf_status st = check_sampling(sampling);
    if (st) return st;
    ATDF_REQUIRE(packed_tree && num_nodes > 0, "atdf_traverse_write: empty tree");
    if (num_rays == 0) return ATDF_OK;
    ATDF_REQUIRE(ray_origin && ray_direction, "atdf_traverse_write: null rays");
    ATDF_REQUIRE(intervals || samples, "atdf_traverse_write: nothing to write");
    if (intervals) {
        ATDF_REQUIRE(ray_interval_offset && intervals->node && intervals->tmin && intervals->tmax &&
                         intervals->count && intervals->first && intervals->ray,
                     "atdf_traverse_write: incomplete interval buffers");
    }
    if (samples) {
        ATDF_REQUIRE(ray_sample_offset && samples->global_coordinate && samples->weight && samples->ray_t &&
                         samples->ray_index && samples->local_index_in_ray && samples->node_id,
                     "atdf_traverse_write: incomplete sample buffers");
    }
    ATDF_REQUIRE(sampling->jitter_mode != 1 || (samples && sampling->jitter),
                 "atdf_traverse_write: jitter_mode 1 needs sample output and a jitter array");
    TraverseArgs a{};
    a.nodes = reinterpret_cast<const PNode*>(packed_tree);
    a.O = ray_origin;
    a.D = ray_direction;
    a.R = num_rays;
    a.K = sampling->max_samples_per_node;
    a.max_inter = sampling->max_intersections;
    a.jitter_mode = sampling->jitter_mode;
    a.seed = sampling->jitter_seed;
    a.jitter = sampling->jitter;
    a.select = sampling->ray_select;
    a.list = sampling->ray_list;
    a.iv_off = intervals ? ray_interval_offset : ray_sample_offset;
    a.smp_off = samples ? ray_sample_offset : ray_interval_offset;
    if (intervals) a.iv = *intervals;
    if (samples) a.smp = *samples;
    if (cache && !a.list) {
        cache_views(const_cast<void*>(cache), num_rays, a.cache_n, a.cache);
        a.cache_slots = cache_slots;
    }
    if (a.cache && !samples)  // an interval list from the hit cache, written coalesced
        hipLaunchKernelGGL(traverse_write_cached_kernel, dim3((unsigned)div_up(num_rays, 256)), dim3(256), 0,
                           to_hip(stream), a);
    else if (a.list && !samples)  // intervals of a sparse selection: eight lanes per listed ray
        hipLaunchKernelGGL(traverse_coop_kernel<true>, dim3((unsigned)div_up(num_rays * 8, kCoopThreads)),
                           dim3(kCoopThreads), 0, to_hip(stream), a);
    else
        hipLaunchKernelGGL(traverse_kernel<true>, dim3((unsigned)div_up(num_rays, 256)), dim3(256), 0,
                           to_hip(stream), a);
    return check_launch("atdf_traverse_write");
}

atdf_status atdf_traverse_count(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                int32_t* ray_num_intervals, int32_t* ray_num_samples, int32_t* status,
                                atdf_stream_t stream)
{
    return traverse_count_impl(packed_tree, num_nodes, ray_origin, ray_direction, num_rays, sampling,
                               ray_num_intervals, ray_num_samples, status, nullptr, 0, stream);
}

atdf_status atdf_traverse_write(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                const int64_t* ray_interval_offset, const int64_t* ray_sample_offset,
                                const atdf_intervals* intervals, const atdf_samples* samples,
                                atdf_stream_t stream)
{
    return traverse_write_impl(packed_tree, num_nodes, ray_origin, ray_direction, num_rays, sampling,
                               ray_interval_offset, ray_sample_offset, intervals, samples, nullptr, 0, stream);
}

size_t atdf_traverse_cache_bytes(int64_t num_rays, int32_t cache_slots)
{
    if (num_rays < 0 || cache_slots < 0) return 0;
    return ((size_t)num_rays * 4 + 15) / 16 * 16 + (size_t)cache_slots * (size_t)num_rays * sizeof(int4);
}

atdf_status atdf_traverse_count_cached(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                       const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                       int32_t* ray_num_intervals, int32_t* ray_num_samples, int32_t* status,
                                       void* cache, int32_t cache_slots, atdf_stream_t stream)
{
    ATDF_REQUIRE(cache && cache_slots > 0, "atdf_traverse_count_cached: null cache or no slots");
    return traverse_count_impl(packed_tree, num_nodes, ray_origin, ray_direction, num_rays, sampling,
                               ray_num_intervals, ray_num_samples, status, cache, cache_slots, stream);
}

atdf_status atdf_traverse_write_cached(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                       const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                       const int64_t* ray_interval_offset, const int64_t* ray_sample_offset,
                                       const atdf_intervals* intervals, const atdf_samples* samples,
                                       const void* cache, int32_t cache_slots, atdf_stream_t stream)
{
    ATDF_REQUIRE(cache && cache_slots > 0, "atdf_traverse_write_cached: null cache or no slots");
    return traverse_write_impl(packed_tree, num_nodes, ray_origin, ray_direction, num_rays, sampling,
                               ray_interval_offset, ray_sample_offset, intervals, samples, cache, cache_slots,
                               stream);
}

size_t atdf_scan_workspace_bytes(int64_t n)
{
    int64_t nt = n > 0 ? div_up(n, kScanTile) : 1;
    return (size_t)nt * sizeof(int64_t);
}

atdf_status atdf_compact_rays(const int32_t* flags, int64_t n, int32_t* list, int32_t* count, atdf_stream_t stream)
{
    ATDF_REQUIRE(n >= 0 && n < (int64_t)INT32_MAX && count, "atdf_compact_rays: bad arguments");
    ATDF_REQUIRE(n == 0 || (flags && list), "atdf_compact_rays: null buffer");
    hipStream_t s = to_hip(stream);
    const unsigned blocks = (unsigned)(n > 0 ? div_up(n, 256) : 1);
    hipLaunchKernelGGL(fill_list_kernel, dim3(blocks), dim3(256), 0, s, list, n, count);
    if (n > 0) hipLaunchKernelGGL(compact_kernel, dim3(blocks), dim3(256), 0, s, flags, n, list, count);
    return check_launch("atdf_compact_rays");
}

atdf_status atdf_exclusive_scan_i32(const int32_t* in, int64_t n, int64_t* out, void* workspace,
                                    size_t workspace_bytes, atdf_stream_t stream)
{
    ATDF_REQUIRE(n >= 0 && out, "atdf_exclusive_scan_i32: bad arguments");
    hipStream_t s = to_hip(stream);
    if (n == 0) {
        hipLaunchKernelGGL(zero_i64_kernel, dim3(1), dim3(64), 0, s, out);
        return check_launch("atdf_exclusive_scan_i32");
    }
    ATDF_REQUIRE(in && workspace, "atdf_exclusive_scan_i32: null buffer");
    ATDF_REQUIRE(workspace_bytes >= atdf_scan_workspace_bytes(n), "atdf_exclusive_scan_i32: workspace too small");
    int64_t nt = div_up(n, kScanTile);
    int64_t* tiles = reinterpret_cast<int64_t*>(workspace);
    hipLaunchKernelGGL(scan_tile_sums_kernel, dim3((unsigned)nt), dim3(kScanThreads), 0, s, in, n, tiles);
    hipLaunchKernelGGL(scan_tile_offsets_kernel, dim3(1), dim3(kScanThreads), 0, s, tiles, nt);
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)nt), dim3(kScanThreads), 0, s, in, n, tiles, out);
    return check_launch("atdf_exclusive_scan_i32");
}

}  // extern "C"

/*
 * atdf.h -- C ABI of libatdf.so, the MI355X (gfx950) reconstruction hot path.
 *
 * Plain pointers and sizes only.  Every pointer argument is DEVICE memory
 * unless stated otherwise; the caller owns every buffer (the libtorch layer
 * allocates them from torch's caching allocator).  Every launch goes on the
 * caller's stream; no entry point synchronises the device.  Status codes
 * replace the reference's glog CHECK aborts; atdf_last_error() returns the
 * message of the last failing call on the calling host thread.
 *
 * Reference interfaces (yuanhaowang1213/adaptivetensordf, hyperacorn/csrc)
 * replaced, per entry point, are cited below as file:line.
 */
#ifndef ATDF_H
#define ATDF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct ihipStream_t;
typedef struct ihipStream_t* atdf_stream_t; /* == hipStream_t */

typedef int32_t atdf_status;
#define ATDF_OK 0
#define ATDF_ERR_INVALID_ARGUMENT 1       /* bad shape / size / null pointer     */
#define ATDF_ERR_CAPACITY 2               /* output larger than caller capacity  */
#define ATDF_ERR_TOO_MANY_INTERSECTIONS 3 /* a ray hit > max_intersections nodes */
#define ATDF_ERR_ORDER 4                  /* traversal order check failed         */
#define ATDF_ERR_HIP 5                    /* HIP runtime error                    */
#define ATDF_ERR_UNSUPPORTED 6            /* combination not built                */

const char* atdf_last_error(void);
const char* atdf_version(void);

/* ======================================================================= */
/* Octree                                                                   */
/* ======================================================================= */
/* The reference tree tensors (structure/HyperTree.h:522-560), device side.
 * Children of a node are contiguous ids (BFS numbering of
 * structure/HyperTree.cpp:453-497).  Octree: child c takes the upper half on
 * axis k iff bit k of c is set.  Quad-tree representation
 * (use_quad_tree_rep, HyperTree.cpp:391-406, 461-464): four children, bit 0
 * of c selects the upper half in x and bit 1 in z; every node spans the
 * same y range. */
typedef struct {
    int64_t num_nodes;
    const float* node_position_min; /* [num_nodes, 3] */
    const float* node_position_max; /* [num_nodes, 3] */
    const int32_t* node_children;   /* [num_nodes, children_per_node], -1 = leaf */
    const int32_t* node_parent;     /* [num_nodes], -1 = root */
    const int32_t* node_active;     /* [num_nodes], 0 / 1 */
    int32_t children_per_node;      /* 8 (octree; 0 also means 8) or 4 (quad tree) */
} atdf_tree;

/* Packed traversal form of the tree (48 B per node), rebuilt after every
 * change of node_active.  Replaces the DeviceHyperTree view
 * (structure/HyperTree.cu:23-55). */
size_t atdf_tree_packed_bytes(int64_t num_nodes);
atdf_status atdf_tree_pack(const atdf_tree* tree, void* packed, atdf_stream_t stream);

/* ======================================================================= */
/* Ray traversal + per-node stratified sample placement                     */
/* ======================================================================= */
/* Replaces HyperTreeBaseImpl::CreateSamplesForRays (structure/HyperTree.h:463,
 * structure/HyperTree.cu:516-610) and its kernel ComputeRaySamples
 * (structure/HyperTree.cu:266-514).  Canonical output order: rays in input
 * order, within a ray the hit nodes sorted by (tmin, node id), samples of a
 * node in increasing t. */
typedef struct {
    int32_t max_samples_per_node; /* K: n = ceil((tmax-tmin)/diag * K)         */
    int32_t max_intersections;    /* 256 in the reference (HyperTree.cu:280)    */
    int32_t jitter_mode;          /* 0 midpoint, 1 array, 2 hash(seed,ray,idx)  */
    int32_t reserved;
    uint64_t jitter_seed;         /* mode 2                                     */
    const float* jitter;          /* mode 1: U[0,1) per canonical sample slot   */
    const int32_t* ray_select;    /* nullable [R]: rays with 0 get no intervals
                                     (the learned proposal's fallback mask)     */
    const int32_t* ray_list;      /* nullable [R]: lane i traverses ray
                                     ray_list[i] (-1: none), so a sparse
                                     selection runs in dense waves; per-ray
                                     outputs of unlisted rays are not written
                                     (atdf_compact_rays builds the list)        */
} atdf_sampling;

/* The single box of HyperTreeBaseImpl::CreateSamplesForRays_area /
 * ComputeRaySamples_area (structure/HyperTree.cu:651-883), used by
 * hyper_train for depth-0 trees with the dataset ROI
 * (samples/hyper_train.cpp:921-927): writes a packed one-node tree whose
 * node 0 is the box [roi_min, roi_max] (host arrays of 3 floats), active,
 * without children (atdf_tree_packed_bytes(1) bytes).  atdf_traverse_count /
 * atdf_traverse_write over it with num_nodes = 1 place that kernel's
 * samples: one interval per ray, node id 0, n = ceil((tmax - tmin) /
 * |roi_max - roi_min| * K) <= K, the same slab test, grazing-interval drop,
 * jitter and midpoint weights. */
atdf_status atdf_tree_pack_box(const float* roi_min, const float* roi_max, void* packed, atdf_stream_t stream);

/* status[0] += rays over max_intersections, status[1] += order-check
 * failures, status[2] = max(samples per ray).  status is device int32[4],
 * zeroed by the caller. */
atdf_status atdf_traverse_count(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                int32_t* ray_num_intervals, int32_t* ray_num_samples, int32_t* status,
                                atdf_stream_t stream);

/* list[0..m) = the indices i < n with flags[i] != 0 (in no fixed order),
 * list[m..n) = -1, *count = m (device int32, written by the call). */
atdf_status atdf_compact_rays(const int32_t* flags, int64_t n, int32_t* list, int32_t* count, atdf_stream_t stream);

/* out[0] = 0, out[i+1] = out[i] + in[i] for i < n (n+1 outputs). */
size_t atdf_scan_workspace_bytes(int64_t n);
atdf_status atdf_exclusive_scan_i32(const int32_t* in, int64_t n, int64_t* out, void* workspace,
                                    size_t workspace_bytes, atdf_stream_t stream);

/* Per-node ray intervals, indexed by interval id (ray r owns
 * [ray_interval_offset[r], ray_interval_offset[r+1])). */
typedef struct {
    int32_t* node;  /* hit node id                             */
    float* tmin;    /* entry t                                 */
    float* tmax;    /* exit t                                  */
    int32_t* count; /* samples in this node                    */
    int32_t* first; /* index of its first sample within the ray */
    int32_t* ray;   /* owning ray                              */
    int64_t total;  /* number of intervals (read by the adjoint; the write
                       pass ignores it)                         */
} atdf_intervals;

/* The reference SampleList (structure/HelperStructs.h:231-297). */
typedef struct {
    float* global_coordinate;   /* [S, 3] = origin + direction * t        */
    float* weight;              /* [S]    midpoint-rule interval weight   */
    float* ray_t;               /* [S]                                    */
    int64_t* ray_index;         /* [S]                                    */
    int64_t* local_index_in_ray;/* [S]                                    */
    int64_t* node_id;           /* [S]                                    */
} atdf_samples;

/* Second pass: writes intervals (if `intervals` is non-null) and samples
 * (if `samples` is non-null) at the scanned offsets. */
atdf_status atdf_traverse_write(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                const int64_t* ray_interval_offset, const int64_t* ray_sample_offset,
                                const atdf_intervals* intervals, const atdf_samples* samples,
                                atdf_stream_t stream);

/* The same two passes with a hit cache, for the reference's count-then-fill
 * CreateSamplesForRays (structure/HyperTree.cu:516-610) when both passes run
 * on one ray list: the count pass records each ray's hit count and its
 * first cache_slots hits (node, tmin, tmax, sample count: 16 B each,
 * slot-major in the caller's cache of atdf_traverse_cache_bytes bytes), and
 * the write pass writes a ray whose hits all fit from the cache instead of
 * walking the tree again (a ray with more hits is walked).  Outputs,
 * statuses and errors are those of atdf_traverse_count / atdf_traverse_write;
 * the write pass must see the sampling and rays its count pass saw.  With a
 * ray_list the cache is not used. */
size_t atdf_traverse_cache_bytes(int64_t num_rays, int32_t cache_slots);
atdf_status atdf_traverse_count_cached(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                       const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                       int32_t* ray_num_intervals, int32_t* ray_num_samples, int32_t* status,
                                       void* cache, int32_t cache_slots, atdf_stream_t stream);
atdf_status atdf_traverse_write_cached(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                       const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                       const int64_t* ray_interval_offset, const int64_t* ray_sample_offset,
                                       const atdf_intervals* intervals, const atdf_samples* samples,
                                       const void* cache, int32_t cache_slots, atdf_stream_t stream);

/* ======================================================================= */
/* Point location and grid transfer across a change of the active set      */
/* ======================================================================= */
/* HyperTreeBaseImpl::NodeIdForPositionGPU (structure/HyperTree.cu:1466-1525,
 * descent of ActiveNodeIdForGlobalPosition, 1339-1399): node_id[i] = the
 * active node reached from the root by stepping into the first child whose
 * closed box holds positions[i]; mask[i] = 1.  Where the descent ends
 * without one (a culled region, outside the root), node_id[i] =
 * active_node_ids[0] and mask[i] = 0.  positions [n, 3] f32. */
atdf_status atdf_node_id_for_position(const void* packed_tree, int64_t num_nodes, const int64_t* active_node_ids,
                                      const float* positions, int64_t num_positions, int64_t* node_id, float* mask,
                                      atdf_stream_t stream);

/* Nearest-lattice transfer of node grids, the copy rule of
 * HyperTreeBaseImpl::InterpolateGridForInactiveNodes
 * (structure/HyperTree.cu:1527-1616): for each destination node and each
 * lattice point (x, y, z) of its G^3 grid, the position
 * min + (x, y, z) / (G - 1) * (max - min) is located among the active nodes
 * of src_packed_tree (as atdf_node_id_for_position) and the destination
 * record takes all C channels of the nearest lattice point of that node.
 * A destination node that is active in the source set keeps its own grid
 * (the reference leaves active nodes untouched).  Lattice points outside
 * every source active node leave the destination unchanged (the reference
 * reads node -1 there).  Source and destination
 * share the tree topology; only the active sets differ.  src_slot /
 * dst_slot map node ids to grid slots (null: slot = node id).  layout 0:
 * [slots, G, G, G, C] (the fused path's channels-last grids); layout 1:
 * [slots, C, G, G, G] (the reference's [num_nodes, C, z, y, x]). */
atdf_status atdf_grid_transfer(const void* src_packed_tree, int64_t num_nodes, const int32_t* src_slot,
                               const float* src_grid, int64_t src_slots, const int64_t* dst_nodes, int64_t num_dst,
                               const int32_t* dst_slot, float* dst_grid, int64_t dst_slots, int32_t C, int32_t G,
                               int32_t layout, atdf_stream_t stream);

/* ======================================================================= */
/* Per-sample ops (the reference's operator API)                            */
/* ======================================================================= */
/* HyperTreeBaseImpl::ComputeLocalSamples (structure/HyperTree.cu:237-259):
 * local = ((g - min[node]) / (max[node] - min[node])) * 2 - 1 */
atdf_status atdf_local_samples(const float* node_position_min, const float* node_position_max,
                               const float* global_coordinate, const int64_t* node_id, int64_t num_samples,
                               float* local, atdf_stream_t stream);

/* IndirectGridSample3D(multi_grid, index, uv) (called at
 * geometry/geometry_tensor_qo.cpp:347; 2-D form modules/IndirectGridSample2D.cu:22-162):
 * out[s, c] = trilinear(grid[index[s], c], uv[s]) with align_corners=True
 * mapping and border clamping.  grid is [N, C, Z, Y, X]. */
atdf_status atdf_grid_sample3d_fwd(const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                                   const int64_t* index, const float* uv, int64_t num_samples, float* out,
                                   atdf_stream_t stream);
/* Backward (modules/IndirectGridSample2D.cu:175-393 pattern): grad_grid +=
 * scatter of grad_out; grad_uv (nullable) += analytic d/duv. */
atdf_status atdf_grid_sample3d_bwd(const float* grid, int64_t N, int32_t C, int32_t Z, int32_t Y, int32_t X,
                                   const int64_t* index, const float* uv, int64_t num_samples,
                                   const float* grad_out, float* grad_grid, float* grad_uv,
                                   atdf_stream_t stream);

/* IndirectGridSample2D(multi_grid, index, uv) (modules/IndirectGridSample2D.h:10-24,
 * IndirectGridSample2D.cu:22-492; called for the in-ROI planes at
 * geometry/geometry_tensor_qo.cpp:306-327): grid [N, C, H, W], uv [S, 2]
 * (u -> W, v -> H), corners tnw tne tsw tse, clipped indices.  Backward:
 * grad_grid += scatter of grad_out; grad_uv (nullable) += the analytic
 * derivative (the reference leaves it at zero). */
atdf_status atdf_grid_sample2d_fwd(const float* grid, int64_t N, int32_t C, int32_t H, int32_t W,
                                   const int64_t* index, const float* uv, int64_t num_samples, float* out,
                                   atdf_stream_t stream);
atdf_status atdf_grid_sample2d_bwd(const float* grid, int64_t N, int32_t C, int32_t H, int32_t W,
                                   const int64_t* index, const float* uv, int64_t num_samples,
                                   const float* grad_out, float* grad_grid, float* grad_uv, atdf_stream_t stream);
/* IndirectGridSample1D (modules/IndirectGridSample1D.cu:20-221; the in-ROI
 * lines at geometry_tensor_qo.cpp:306-327): grid [N, C, W, 1], uv [S];
 * grad_uv (nullable) += the analytic derivative (the reference sums a
 * different expression, IndirectGridSample1D.cu:128-134). */
atdf_status atdf_grid_sample1d_fwd(const float* grid, int64_t N, int32_t C, int32_t W, const int64_t* index,
                                   const float* uv, int64_t num_samples, float* out, atdf_stream_t stream);
atdf_status atdf_grid_sample1d_bwd(const float* grid, int64_t N, int32_t C, int32_t W, const int64_t* index,
                                   const float* uv, int64_t num_samples, const float* grad_out, float* grad_grid,
                                   float* grad_uv, atdf_stream_t stream);

/* GradientCorrection(sample_position, ray_direction, index) backward
 * (modules/GradientCorrection.h:11-25, GradientCorrection.cu:22-124): the
 * forward is the identity; grad_position = g - (g . d) d with d =
 * ray_direction[ray_index[s]]. */
atdf_status atdf_gradient_correction_bwd(const float* ray_direction, const int64_t* ray_index, int64_t num_samples,
                                         const float* grad_out, float* grad_position, atdf_stream_t stream);

/* HyperTreeBaseImpl::ComputeBoundWeight(node_id, in_roi_node, false)
 * (structure/HyperTree.cu:962-993): out[s] = s when node_id[s] is an in-ROI
 * node, else -1.  in_roi_flag is [num_nodes] 0/1 (one lookup per sample
 * instead of the reference's scan over the in-ROI list). */
atdf_status atdf_compute_bound_weight(const int64_t* node_id, int64_t num_samples, const int32_t* in_roi_flag,
                                      int64_t* out, atdf_stream_t stream);

/* NeuralGeometry::IntegrateSamplesXRay (geometry/geometry.h:269-270,
 * geometry/geometry.cpp:977-1019): out[c, ray[s]] += values[s, c] * weight[s]. */
atdf_status atdf_integrate_xray(const float* values, const float* weight, const int64_t* ray_index,
                                int64_t num_samples, int32_t C, int64_t num_rays, float* out,
                                atdf_stream_t stream);

/* HierarchicalNeuralGeometry::AccumulateSampleLossPerNode
 * (geometry/geometry.h:326-327, geometry/geometry.cpp:33-53), reduced in FP64:
 * loss_sum[node[s]] += mean_c(per_ray_loss[c, ray[s]]) * weight[s];
 * weight_sum[node[s]] += weight[s]. */
atdf_status atdf_accumulate_loss_per_node(const float* per_ray_loss, int32_t C, int64_t num_rays,
                                          const int64_t* ray_index, const float* weight, const int64_t* node_id,
                                          int64_t num_samples, int64_t num_nodes, double* loss_sum,
                                          double* weight_sum, atdf_stream_t stream);

/* ======================================================================= */
/* Fused projection (forward) and its hand-written adjoint                  */
/* ======================================================================= */
/* The field: per-node feature grids, channels-last [num_grids, G, G, G, C]
 * (the reference ExplicitFeatureGrid stores [N, C, G, G, G],
 * geometry/ImplicitNet.h:66-137; atdf_grid_layout converts), decoded by the
 * shared MLP C -> H -> H -> 1 with ReLU hidden activations and
 * softplus(beta) output (geometry/Modules.h:23-145, geometry/geometry.cpp:176-253).
 * mlp = [W1 (H x C), b1 (H), W2 (H x H), b2 (H), W3 (1 x H), b3 (1)]. */
#define ATDF_GRID_F32 0
#define ATDF_GRID_BF16 1
/* Decoder arithmetic: FP32 MFMA (v_mfma_f32_16x16x4_f32), or BF16 MFMA inputs
 * with FP32 accumulation (v_mfma_f32_16x16x16_bf16 / 16x16x32_bf16). */
#define ATDF_DECODER_F32 0
#define ATDF_DECODER_BF16 1
/* The tensorial ("VM") field of the in-ROI nodes
 * (GeometryTensorQO::SampleVolumeIndirect_feature, tensor_op "add",
 * geometry/geometry_tensor_qo.cpp:279-327): for a sample of node n with
 * local coordinates (lx, ly, lz) in [-1, 1] and global position (gx, gy, gz)
 *   F[c] = Pyz[c](ly, lz) Lx[c](x') + Pxz[c](lx, lz) Ly[c](ly) + Pxy[c](lx, ly) Lz[c](gz),
 *   x' = 2 (gx - roi_min_x) / (roi_max_x - roi_min_x) - 1,
 * products rounded and added in that order; P are bilinear reads (first
 * coordinate along the plane's last axis) and L linear reads of the shared
 * lines, both with align_corners and border clamping as IndirectGridSample2D
 * / 1D (modules/IndirectGridSample2D.cu:22-162, IndirectGridSample1D.cu:20-140).
 * Nodes outside the ROI read the field's explicit grids (node_to_grid).
 * F32 only (grid_dtype ATDF_GRID_F32); either decoder dtype.  Layouts are
 * channels-last: the reference stores each plane [N, C, Gp, Gp] and each
 * line [1, C, Gl, 1] (geometry/tensorfield.h:9-127, 218-335).
 * grad_planes / grad_lines are the adjoint's outputs (+=, same shapes); the
 * forward ignores them. */
typedef struct {
    const float* planes;           /* [num_planes, 3, Gp, Gp, C]: yz, xz, xy; (slot, p, v, u, c) */
    const float* lines;            /* [3, Gl, C]: x, y, z                                          */
    int64_t num_planes;
    const int32_t* node_to_plane;  /* [num_nodes] -> plane slot, -1 = not an in-ROI node           */
    int32_t Gp;                    /* plane edge                                                   */
    int32_t Gl;                    /* line length                                                  */
    float roi_min_x, roi_max_x;    /* the ROI's x range (line x coordinate)                        */
    float* grad_planes;            /* adjoint: += dL/dplanes                                       */
    float* grad_lines;             /* adjoint: += dL/dlines                                        */
} atdf_vm;

typedef struct {
    int32_t C;                   /* channels; 16 and 8 are built              */
    int32_t H;                   /* hidden width; 64 is built                 */
    int32_t G;                   /* grid edge, e.g. 11                        */
    int32_t grid_dtype;          /* ATDF_GRID_F32 / ATDF_GRID_BF16            */
    float softplus_beta;         /* 2 in the reference (Settings.h:164)       */
    int32_t rays_per_wave;       /* work unit of one wavefront (0 = 16)       */
    const void* grid;            /* [num_grids, G, G, G, C]                   */
    int64_t num_grids;
    const int32_t* node_to_grid; /* [num_nodes] -> grid slot (-1 = none)      */
    const float* mlp;            /* packed decoder parameters (f32)          */
    int32_t decoder_dtype;       /* ATDF_DECODER_F32 / ATDF_DECODER_BF16      */
    int32_t reserved;
    const atdf_vm* vm;           /* tensorial in-ROI field, or null. Honoured by
                                    atdf_project_fwd, atdf_project_bwd (and its
                                    _noalloc variant); the other entry points
                                    reject a field with vm set */
} atdf_field;

/* Forward: projection[r] = sum over the ray's samples of
 * softplus(MLP(trilinear(grid))) * weight -- the composition
 * ComputeImage -> SampleVolumeIndirect -> DecodeFeatures ->
 * IntegrateSamplesXRay (geometry/geometry.cpp:828-914) for one channel,
 * with samples recomputed from the intervals and never stored. */
atdf_status atdf_project_fwd(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                             const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                             const int64_t* ray_interval_offset, const atdf_intervals* intervals,
                             const atdf_field* field, float* projection, atdf_stream_t stream);

/* Adjoint of atdf_project_fwd with per-sample recomputation: given dL/dP per
 * ray, grad_grid (f32, channels-last, same shape as the grid) += dL/dgrid and
 * grad_mlp += dL/dmlp.  Replaces the autograd replay of
 * IndirectGridSample*::backward (modules/IndirectGridSample2D.cu:441-477),
 * the FCBlock backward and the index_add_ transpose.  Work is node-major:
 * intervals are counting-sorted by node and each node's grid gradient is
 * accumulated in LDS and written once (no per-sample global atomics).
 * intervals->total must hold the interval count. */
size_t atdf_project_bwd_workspace_bytes(int64_t num_rays, int64_t num_intervals, const atdf_field* field);

/* Extra workspace, beyond atdf_project_bwd_workspace_bytes, that switches the
 * adjoint (atdf_project_bwd, _adam, _prepared, and the sort of
 * atdf_project_prepare) to its two-pass form for a list of num_samples
 * samples: a decoder pass at two workgroups per CU writes each sample's
 * lattice coordinates and dL/dfeature (16 + 4 C bytes) in node-sorted order,
 * then a scatter pass adds them into the LDS node gradient.  Same results to
 * FP32 rounding (both forms sum the scatter in a run-dependent order).  A
 * workspace of at least the two sizes added runs the two-pass form whenever
 * the list's samples fit it; a smaller one runs the fused kernel.  With a
 * tensorial field the in-ROI samples stage their coordinates and dL/dplane,
 * dL/dline before the corner weights (32 + 24 C bytes) and a third pass adds
 * them into LDS plane gradients and line partials; the size then bounds
 * every sample at that record size.  0 for C outside {8, 16}, explicit grids
 * with G > 32, or planes and lines whose gradients exceed the LDS of one
 * workgroup.  No reference
 * counterpart (the reference scatters per sample with global atomics,
 * modules/IndirectGridSample2D.cu:175-393).  atdf_project_prepare and
 * atdf_project_bwd_prepared must be given the same workspace and size. */
size_t atdf_project_bwd_staging_bytes(int64_t num_intervals, int64_t num_samples, const atdf_field* field);

atdf_status atdf_project_bwd(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                             const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                             const int64_t* ray_interval_offset, const atdf_intervals* intervals,
                             const atdf_field* field, const float* dLdP, float* grad_grid, float* grad_mlp,
                             void* workspace, size_t workspace_bytes, atdf_stream_t stream);

/* atdf_project_bwd without the workspace argument and without allocating:
 * the counting sort, interval records, chunk table and decoder-gradient
 * partials live in a fixed device array of the library, `scratch`, of
 * atdf_project_bwd_noalloc_max_bytes() bytes (256 MiB) per GPU, held until
 * the process ends.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Results are those of atdf_project_bwd.
 * A call whose atdf_project_bwd_workspace_bytes() exceeds the maximum is
 * rejected (ATDF_ERR_INVALID_ARGUMENT) before any device work.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU race. */
size_t atdf_project_bwd_noalloc_max_bytes(void);
atdf_status atdf_project_bwd_noalloc(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                     const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                     const int64_t* ray_interval_offset, const atdf_intervals* intervals,
                                     const atdf_field* field, const float* dLdP, float* grad_grid, float* grad_mlp,
                                     atdf_stream_t stream);

/* Adam state of the node grids for atdf_project_bwd_adam. */
typedef struct {
    float* param;          /* [num_grids * G^3 * C] f32 grids; the master copy when the field is BF16 */
    uint16_t* param_bf16;  /* the field's BF16 grid (rounded from param), or null for an F32 field    */
    float* exp_avg;        /* Adam m, same shape as param                                             */
    float* exp_avg_sq;     /* Adam v                                                                  */
    float lr, beta1, beta2, eps;
    int64_t step;          /* 1-based step count of this update                                       */
} atdf_grid_adam;

/* atdf_project_bwd followed by one atdf_adam_step[_bf16] on the grids, fused:
 * a node whose samples fit one chunk gets its Adam update straight from the
 * LDS gradient at the end of its chunk (no gradient round trip through HBM);
 * the other nodes (split over chunks, or without samples) are updated from
 * grad_grid afterwards.  grad_grid must be all zero on entry and is all zero
 * on return; grad_mlp += dL/dmlp (the decoder's Adam stays with the caller).
 * The field's grid must be the memory adam->param (F32) or adam->param_bf16
 * (BF16).  No reference counterpart: the fusion of the adjoint with
 * OptimizerStep's grid update (geometry/geometry.h:203-225). */
atdf_status atdf_project_bwd_adam(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                  const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                  const int64_t* ray_interval_offset, const atdf_intervals* intervals,
                                  const atdf_field* field, const float* dLdP, float* grad_grid, float* grad_mlp,
                                  const atdf_grid_adam* adam, void* workspace, size_t workspace_bytes,
                                  atdf_stream_t stream);

/* The training step's sort shared by forward and adjoint.
 * atdf_project_prepare counting-sorts the intervals by grid slot and writes
 * their records and the chunk table into the workspace
 * (atdf_project_bwd_workspace_bytes); atdf_project_fwd_prepared is then a
 * node-major forward over that sort (each node's grid stays in L2 while its
 * chunk is integrated; per-interval sums are added per ray in list order,
 * so the rounding differs from atdf_project_fwd by a few ulp), and
 * atdf_project_bwd_prepared is atdf_project_bwd_adam (adam non-null) or
 * atdf_project_bwd (adam null) without sorting again.  The three calls must
 * see the same rays, intervals and slot map; the grid and decoder may
 * change in between.  A tensorial field (field->vm) takes the same calls:
 * its in-ROI items are integrated node-major too when the workspace holds
 * the two-pass staging region (atdf_project_bwd_staging_bytes), and
 * atdf_project_fwd_prepared falls back to atdf_project_fwd without it.  No
 * reference counterpart: the reference gathers ray-major in both directions
 * (geometry/geometry.cpp:828-914). */
atdf_status atdf_project_prepare(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                 const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                 const int64_t* ray_interval_offset, const atdf_intervals* intervals,
                                 const atdf_field* field, void* workspace, size_t workspace_bytes,
                                 atdf_stream_t stream);
atdf_status atdf_project_fwd_prepared(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                      const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                      const int64_t* ray_interval_offset, const atdf_intervals* intervals,
                                      const atdf_field* field, float* projection, void* workspace,
                                      size_t workspace_bytes, atdf_stream_t stream);
atdf_status atdf_project_bwd_prepared(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                      const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                      const int64_t* ray_interval_offset, const atdf_intervals* intervals,
                                      const atdf_field* field, const float* dLdP, float* grad_grid, float* grad_mlp,
                                      const atdf_grid_adam* adam, void* workspace, size_t workspace_bytes,
                                      atdf_stream_t stream);

/* atdf_project_fwd plus the per-ray mass histogram ray_bins[r, b] (f32,
 * [num_rays, 32]): the sum of density * weight over the samples of ray r
 * whose t falls in bin b = (int)((t - t0) / (t1 - t0) * 32), clamped to
 * [0, 31] (0 when t1 <= t0), of the ray's chord [t0, t1] through the root
 * box.  These are the learned proposal's training targets: "densities the
 * reconstruction itself produces", binned along each ray. */
atdf_status atdf_project_fwd_bins(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                  const float* ray_direction, int64_t num_rays, const atdf_sampling* sampling,
                                  const int64_t* ray_interval_offset, const atdf_intervals* intervals,
                                  const atdf_field* field, float* projection, float* ray_bins,
                                  atdf_stream_t stream);

/* Per-node density statistics for the octree optimizer, reduced in FP64:
 * for active node a (node id active_ids[a]) the field's density is evaluated
 *   mode 0: on the n^3 lattice of UniformGlobalSamples (structure/HyperTree.cu:1618-1653),
 *           as ComputeMaxDensity does (samples/hyper_train.cpp:662-707);
 *   mode 1: at rounds x n^3 uniform points of CreateSamplesRandomly
 *           (structure/HyperTree.cu:887-920), as ComputeStd does (hyper_train.cpp:709-759);
 * node_max[a] = max density (f32), node_sum[a], node_sumsq[a] = sum and sum
 * of squares (f64).  One workgroup per node, no atomics (deterministic). */
atdf_status atdf_node_density_stats(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                    int64_t num_active, const atdf_field* field, int32_t mode, int32_t n,
                                    int32_t rounds, uint64_t seed, float* node_max, double* node_sum,
                                    double* node_sumsq, atdf_stream_t stream);

/* The same with a per-point multiplier: density = density * point_mask[a *
 * points_per_node + p] before the reductions, as SampleVolumeBatched
 * multiplies by its sample mask (geometry/geometry.h:353-424) when
 * ComputeMaxDensity passes the visibility of PointInAnyImage
 * (hyper_train.cpp:686-691).  point_mask nullable. */
atdf_status atdf_node_density_stats_masked(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                           int64_t num_active, const atdf_field* field, int32_t mode, int32_t n,
                                           int32_t rounds, uint64_t seed, const float* point_mask, float* node_max,
                                           double* node_sum, double* node_sumsq, atdf_stream_t stream);

/* HyperTreeBaseImpl::UniformGlobalSamples (structure/HyperTree.cu:1618-1665):
 * positions [num_active, n, n, n, 3] f32, point (x, y, z) of node a =
 * ((x, y, z) / (n - 1)) * (max - min) + min, x fastest -- the points of
 * atdf_node_density_stats mode 0, in the same order. */
atdf_status atdf_node_lattice_points(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                     int64_t num_active, int32_t n, float* positions, atdf_stream_t stream);

/* Regular lattice of points for volume export.  Point i = (i0, i1, i2) with
 * i0 fastest (swap_xy: i2 fastest), index i0 + size0 (i1 + size1 i2):
 *   mode 0: p = (i / (size - 1)) * 2 - 1 over the unit cube
 *           (UniformPhantomSamplesGPU, structure/HyperTree.cu:1158-1219);
 *   mode 1: p = roi_min + step * i
 *           (UniformPhantomSamplesGPUbySlice, structure/HyperTree.cu:1221-1293).
 * Host memory, like atdf_sampling. */
typedef struct {
    int32_t size[3];
    int32_t mode;
    int32_t swap_xy;
    float roi_min[3];
    float step[3];
} atdf_lattice;

/* The field on the lattice: point location as atdf_node_id_for_position
 * (node_id: the active node, or active_node_ids[0]; valid: 1 / 0), then the
 * projection's gather and decoder: density[i] = softplus(decoder) for points
 * inside an active node and 0 outside (the reference multiplies by the
 * mask).  Replaces the evaluation loop of
 * HierarchicalNeuralGeometry::UniformSampledVolume (geometry/geometry.cpp:
 * 347-660, non-final path).  Every output is nullable; with density null
 * the field is not evaluated (positions / point location only, f may be
 * null).  position: [n, 3] f32. */
atdf_status atdf_sample_volume(const void* packed_tree, int64_t num_nodes, const int64_t* active_node_ids,
                               const atdf_field* field, const atdf_lattice* lattice, float* density,
                               int64_t* node_id, float* valid, float* position, atdf_stream_t stream);

/* DecodeFeatures on given feature rows: density[i] = softplus(MLP(x_i)) for
 * features [n, C] f32 (HierarchicalNeuralGeometry::DecodeFeatures,
 * geometry/geometry.cpp:176-253, with the softplus last activation; the
 * FCBlock of geometry/Modules.h:23-145).  Reads field->C, H, softplus_beta,
 * mlp and decoder_dtype only (the grid may be null).  The adjoint WRITES
 * grad_features [n, C] and ADDS dL/dmlp to grad_mlp; the decoder-parameter
 * gradient goes through per-wave partials (workspace) and a fixed-order
 * reduction, so both directions are deterministic. */
atdf_status atdf_decode_features(const float* features, int64_t n, const atdf_field* field, float* density,
                                 atdf_stream_t stream);
size_t atdf_decode_features_bwd_workspace_bytes(int64_t n, const atdf_field* field);
atdf_status atdf_decode_features_bwd(const float* features, int64_t n, const atdf_field* field,
                                     const float* grad_density, float* grad_features, float* grad_mlp,
                                     void* workspace, size_t workspace_bytes, atdf_stream_t stream);

/* Pairs of points across the faces between active nodes, for the edge
 * regulariser: HyperTreeBaseImpl::NodeNeighborSamples (structure/
 * HyperTree.cu:1720-1811, 2010-2075).  For every active node, face (side s
 * in 0..2, back then front) and point of a size[s+1] x size[s+2] lattice
 * (indices mod 3) on it, local[s] = front, local[s+1] = (i % size[s+1]) /
 * (size[s+1] - 1), local[s+2] = (i / size[s+1]) / (size[s+2] - 1), p =
 * local * (max - min) + min; p moved by epsilon * (front ? +1 : -1) along s
 * is located as atdf_node_id_for_position.  A different active node there
 * (with use_roi: both nodes pass in_roi, HyperTree.cu:1667-1690, whose
 * "mid" is min + max) gives a pair (p in the node, p in the other node).
 * variant 1 is NodeNeighborSamplesQuad (HyperTree.cu:1915-2008), which the
 * tensorial geometry's edge term uses through NodeNeighborSamplesPlane2D
 * (geometry_tensor_qo.cpp:1158-1160): sides 0 and 2 only, in_roi_quad
 * (HyperTree.cu:1692-1718: the y range widened by epsilon / 2) and point
 * location over the first four children of each node (the reference's
 * use_quad descent, applied to whatever tree it is given).  Host struct. */
typedef struct {
    int32_t size[3];
    float epsilon;
    int32_t use_roi;
    float roi_min[3];
    float roi_max[3];
    int32_t variant; /* 0 NodeNeighborSamples; 1 NodeNeighborSamplesQuad; 2 NodeNeighborSamples2D */
    int32_t plane;   /* variant 2: the plane index (the face local[plane] = 0 of every node) */
} atdf_neighbor_spec;

/* Candidates: num_active * 2 * (size1 size2 + size2 size0 + size0 size1)
 * (variant 1: without the size2 size0 term of side 1), numbered per active
 * node as (side, back / front, lattice index).  Variant 2
 * (NodeNeighborSamples2D of the quad-tree representation,
 * structure/HyperTree.cu:1815-1913): num_active * 2 * (size[s1] + size[s2])
 * with (s1, s2) = (plane + 1, plane + 2) mod 3, numbered per active node as
 * (tt, back / front, point): tt = 0 steps along s1 with points along s2,
 * tt = 1 the reverse; the four-child descent; the ROI's y range widened by
 * epsilon once per candidate the reference's thread has visited. */
int64_t atdf_neighbor_candidates(int64_t num_active, const atdf_neighbor_spec* spec);
/* flags[c] = 1 where candidate c is a pair (device int32). */
atdf_status atdf_neighbor_flags(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                int64_t num_active, const atdf_neighbor_spec* spec, int32_t* flags,
                                atdf_stream_t stream);
/* With offsets = the exclusive scan of flags (atdf_exclusive_scan_i32, n + 1
 * entries), pair k = offsets[c] of candidate c is written to samples 2k
 * (this node) and 2k + 1 (the other): positions [2P, 3] f32, node_id [2P]
 * i64 -- candidate order, so the list is deterministic (the reference
 * appends pairs with an atomic counter). */
atdf_status atdf_neighbor_write(const void* packed_tree, int64_t num_nodes, const int64_t* active_ids,
                                int64_t num_active, const atdf_neighbor_spec* spec, const int64_t* offsets,
                                float* positions, int64_t* node_id, atdf_stream_t stream);

/* ======================================================================= */
/* Scene data: detector rays, projections, random ray batches              */
/* ======================================================================= */
/* Cameras and poses of a scene (host struct, device buffers):
 *   camera_model 0 pinhole (cone beam), 1 orthographic (parallel beam)
 *     (DatasetParams::camera_model, Settings.h:300);
 *   w, h: detector size; align_corners: the uv convention of UVToPixel
 *     (1 in the reference, SceneBase.h:24);
 *   intrinsics [num_cameras, 5] f64 fx fy cx cy s (CameraBase::K,
 *     SceneBase.h:84-127);
 *   rotation [num_images, 4] f64 unit quaternion (x, y, z, w), translation
 *     [num_images, 3] f64 (CameraPoseModuleImpl, modules/CameraPose.cu:
 *     250-279): the pose translation times scene_scale minus
 *     volume_translation (SceneBase.cpp:98-117, 840); read by the pinhole
 *     model only;
 *   volume_translation: TrainParams::volume_translation (SceneBase.cpp:
 *     112-117), added by the orthographic model. */
typedef struct {
    int32_t camera_model;
    int32_t w, h;
    int32_t align_corners;
    int64_t num_cameras, num_images;
    const double* intrinsics;
    const double* rotation;
    const double* translation;
    float volume_translation[3];
} atdf_camera_rig;

/* SceneBase::GetRays (data/SceneBase.cpp:369-449) for num_rays detector
 * points: uv [n, 2] f32 in [0, 1], image_id / camera_id [n] i32 ->
 * ray_origin, ray_direction [n, 3] f32.  UVToPixel in f32
 * (modules/CameraPose.h:14-30), then in f64: u = ((px - cx) / fx,
 * (py - cy) / fy, 1) (CameraModelModuleImpl::Unproject, modules/
 * CameraPose.cu:469-481) and the rotation v' = v + w t + q x t with
 * t = 2 (q x v) (RotatePointGPUk, modules/CameraPose.cu:27-40).
 *   pinhole:      origin = translation, direction = R u / |u|;
 *   orthographic: direction = R (0, 0, 1),
 *                 origin = R (u_x, u_y, 0) + volume_translation - sqrt(2) direction.
 * Cast to f32 at the end.  An image or camera index outside the rig gives
 * NaN rays (nothing is read out of bounds). */
atdf_status atdf_generate_rays(const atdf_camera_rig* rig, const float* uv, const int32_t* image_id,
                               const int32_t* camera_id, int64_t num_rays, float* ray_origin, float* ray_direction,
                               atdf_stream_t stream);

/* X-ray log conversion of SceneBase::LoadImagesCT (data/SceneBase.cpp:
 * 939-983) over n pixels:
 *   log_space_input:  out = raw / xray_max
 *   otherwise:        out = -log(max(raw, 1e-5) / xray_max)
 *                           * (projection_factor / -log(xray_min / xray_max))
 * with log10 (log10_conversion != 0) or the natural log; the scale is
 * formed in f64 and applied in f32, like a libtorch f32 tensor times a
 * double scalar.  mask (nullable, [n] f32 in [0, 1]) multiplies the result
 * (SceneBase.cpp:976-984). */
atdf_status atdf_log_convert(const float* raw, const float* mask, int64_t n, double xray_min, double xray_max,
                             double projection_factor, int32_t log10_conversion, int32_t log_space_input,
                             float* out, atdf_stream_t stream);

/* Projections of a scene resident in device memory: projection and mask
 * (nullable) [num_images, h, w] f32, one channel (the CT case of
 * UnifiedImage::projection, SceneBase.h:58-64).  Host struct. */
typedef struct {
    const float* projection;
    const float* mask;
    int32_t w, h;
    int64_t num_images;
    int32_t align_corners;
} atdf_projection_set;

/* UnifiedImage::SampleProjection (data/SceneBase.cpp:76-97): the
 * projection (and mask) of image image_id[i] at uv[i], bilinear with border
 * clamping (libtorch grid_sample, padding_mode border, on g = uv * 2 - 1):
 *   x = (g + 1) * ((w - 1) / 2)  (align_corners)  or  (g + 1) * (w / 2) - 0.5,
 *   clamped to [0, w - 1]; the four neighbours weighted (1-a)(1-b), a(1-b),
 *   (1-a)b, ab and summed in that order, neighbours past the last pixel read 0.
 * target [n] f32; target_mask [n] f32 (nullable; 1 where the set has no mask).
 * An image index outside the set gives NaN. */
atdf_status atdf_sample_projection(const atdf_projection_set* set, const float* uv, const int32_t* image_id,
                                   int64_t n, float* target, float* target_mask, atdf_stream_t stream);

/* Random ray batch (RandomRaybasedSampleDataset::get_batch, data/
 * Dataloader.cpp:9-48): ray i draws an entry k uniformly from
 * image_indices[num_indices] (image_id = image_indices[k], camera_id =
 * image_camera[k], 0 with image_camera null) and its detector point
 *   interpolate != 0: uv uniform in [0, 1)^2 (UnifiedImage::CoordinatesRandom,
 *                     SceneBase.h:52);
 *   interpolate == 0: a uniform pixel (x, y) mapped by PixelToUV
 *                     (CoordinatesRandomNoInterpolate, SceneBase.cpp:38-45;
 *                     modules/CameraPose.h:32-48).
 * Draws come from a counter-based hash of (seed, i), so a batch is a pure
 * function of its arguments (the reference draws from libtorch's generator).
 * Then, in the same launch, the batch's targets (atdf_sample_projection;
 * set nullable) and rays (atdf_generate_rays; rig nullable).  Every output
 * except uv, image_id and camera_id is nullable. */
atdf_status atdf_random_ray_batch(const atdf_camera_rig* rig, const atdf_projection_set* set,
                                  const int32_t* image_indices, int64_t num_indices, const int32_t* image_camera,
                                  int32_t w, int32_t h, int32_t align_corners, int32_t interpolate, uint64_t seed,
                                  int64_t num_rays, float* uv, int32_t* image_id, int32_t* camera_id, float* target,
                                  float* target_mask, float* ray_origin, float* ray_direction,
                                  atdf_stream_t stream);

/* SceneBase::PointInAnyImage (data/SceneBase.cu:13-95): mask[i] = 1 if
 * points[i] lies in the closed box [roi_min, roi_max] (host arrays) and
 * projects into the detector of at least one image, else 0.  In f64: view =
 * R^T p + R^T (-t) (the inverse pose, t = rig translation, read for both
 * camera models like the reference), then with camera 0's intrinsics and
 * skew 0
 *   pinhole:      (fx (x / z) + cx, fy (y / z) + cy)   (division as x * (1 / z))
 *   orthographic: (fx x + cx, fy y + cy)
 * rounded half away from zero; visible when 0 <= px < w and 0 <= py < h.
 * points [n, 3] f32 (the reference casts its float positions to f64). */
atdf_status atdf_point_in_any_image(const atdf_camera_rig* rig, const float* roi_min, const float* roi_max,
                                    const float* points, int64_t n, float* mask, atdf_stream_t stream);

/* ======================================================================= */
/* Learned sample proposal                                                  */
/* ======================================================================= */
/* No counterpart in the reference: its sampler is the stratified per-node
 * placement of ComputeRaySamples (structure/HyperTree.cu:266-514), which
 * stays as the fallback.  A small MLP 38 -> 64 -> 64 -> 64 -> 32 (ReLU)
 * reads the ray origin (3), direction (3) and the coarse occupancy of 32
 * equal bins of the ray's chord through the root box (the coarse density of
 * the active node at each bin midpoint), and its softmax is a
 * piecewise-constant pdf over the bins.  Sampling pdf = (1 - mix) * softmax
 * + mix / 32; the inverse CDF at i / 32 (i = 1..31) cuts the chord into 32
 * equal-probability cells, narrow where the pdf is high, and each cell
 * becomes a one-sample interval (node located at the sample the fused
 * kernels will place: the midpoint, or the jittered point) that the fused
 * kernels integrate like any other interval.  The cells tile the chord, so
 * with jitter the estimate is unbiased for any pdf.  A ray falls
 * back to the stratified traversal when KL(coarse occupancy || pdf) exceeds
 * kl_threshold.  Every FP32 operation runs in a fixed order (the dense
 * layers' dot products as fused multiply-add chains, no other contraction),
 * so logits, KL, fallback decisions and sample placement are
 * reproducible bit for bit on the CPU.
 * params (torch Linear layout [out, in]): W1 [64, 38] b1 [64] W2 [64, 64]
 * b2 [64] W3 [64, 64] b3 [64] W4 [32, 64] b4 [32]. */
#define ATDF_PROPOSAL_BINS 32
#define ATDF_PROPOSAL_INPUTS 38
#define ATDF_PROPOSAL_HIDDEN 64
int64_t atdf_proposal_num_params(void);

/* Point location table: table[(z * res + y) * res + x], res = 2^depth, holds
 * the active node containing the centre of lattice cell (x, y, z) of the
 * root box, or -1 (no active node there, e.g. a culled region).  Rebuilt
 * after every change of the active set; depth = the tree's max depth
 * (1..8).  A point p maps to the cell floor((p - min) / (max - min) * res),
 * clamped.  Plays the role of NodeIdForPositionGPU
 * (structure/HyperTree.cu:1339-1525) for the proposal. */
int64_t atdf_leaf_table_numel(int32_t depth);
atdf_status atdf_leaf_table(const void* packed_tree, int64_t num_nodes, int32_t depth, int32_t* table,
                            atdf_stream_t stream);

typedef struct {
    const int32_t* leaf_table; /* atdf_leaf_table output                       */
    int32_t table_depth;
    float mix;                 /* uniform mixing weight in [0, 1]              */
    float kl_threshold;        /* fallback when KL(coarse || pdf) > threshold  */
    int32_t jitter_mode;       /* 0 / 2, as the sampling the intervals are
                                  integrated with: each cell's node is the one
                                  holding the sample the fused kernels place */
    uint64_t jitter_seed;
    const float* coarse;       /* [num_nodes] coarse density per node          */
    const float* params;       /* [atdf_proposal_num_params()]                 */
} atdf_proposal;

/* Network inputs [num_rays, 38] and the root chord [num_rays, 2] (t0, t1;
 * zero when the chord misses or is shorter than 1e-5). */
atdf_status atdf_proposal_features(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                   const float* ray_direction, int64_t num_rays, const atdf_proposal* proposal,
                                   float* features, float* chord, atdf_stream_t stream);

/* Inference + placement.  Per ray: ray_fallback (0/1), ray_num_learned
 * (0 for fallback rays) and the learned one-sample intervals, slot-major:
 * learned_node / learned_tmin / learned_tmax [32, num_rays], slot i < n of
 * ray r at i * num_rays + r.  ray_kl (nullable) = KL(coarse || pdf). */
atdf_status atdf_proposal_sample(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                 const float* ray_direction, int64_t num_rays, const atdf_proposal* proposal,
                                 int32_t* ray_fallback, int32_t* ray_num_learned, int32_t* learned_node,
                                 float* learned_tmin, float* learned_tmax, float* ray_kl, atdf_stream_t stream);

/* Learned allocation over a stratified interval list (the proposal's second
 * mode; no reference counterpart).  `intervals` is the list of
 * atdf_traverse_write for every ray (any K); for each ray the gate keeps
 * (KL(coarse || pdf) <= kl_threshold), node interval i of the ray gets
 *   count_i + round(E F(tmax_i)) - round(E F(tmin_i))
 * samples, count_i its stratified count in the list (at least one),
 * E = extra_samples, F the CDF of the mixed pdf over the ray's root chord
 * (linear within each of the 32 bins, clamped to [0, 1]),
 * round(x) = floor(x + 1/2): every node interval keeps its stratified samples
 * and the extras telescope to E over intervals that tile the chord.  A
 * fallback ray (KL above the threshold) is allocated the same way along the
 * uniform pdf (F = the chord fraction, up to FP32 rounding): stratified
 * sampling at the same budget.  count and first are rewritten in place; rays
 * without a root chord keep their stratified counts.  ray_fallback
 * [num_rays] (0/1), ray_num_samples (nullable) = the ray's new sample count,
 * ray_kl nullable. */
atdf_status atdf_proposal_allocate(const void* packed_tree, int64_t num_nodes, const float* ray_origin,
                                   const float* ray_direction, int64_t num_rays, const atdf_proposal* proposal,
                                   int32_t extra_samples, const int64_t* ray_interval_offset,
                                   const atdf_intervals* intervals, int32_t* ray_fallback, int32_t* ray_num_samples,
                                   float* ray_kl, atdf_stream_t stream);

/* Mixed interval list: run atdf_traverse_count with sampling.ray_select =
 * ray_fallback, then this replaces the counts of the learned rays by
 * ray_num_learned (in place); after atdf_exclusive_scan_i32 and
 * atdf_traverse_write (same ray_select) over the combined offsets,
 * atdf_proposal_write fills in the learned rays' intervals (count 1, first =
 * slot, ray = r). */
atdf_status atdf_proposal_merge_counts(int64_t num_rays, const int32_t* ray_fallback,
                                       const int32_t* ray_num_learned, int32_t* ray_num_intervals,
                                       int32_t* ray_num_samples, atdf_stream_t stream);
atdf_status atdf_proposal_write(int64_t num_rays, const int32_t* ray_fallback, const int32_t* ray_num_learned,
                                const int32_t* learned_node, const float* learned_tmin,
                                const float* learned_tmax, const int64_t* ray_interval_offset,
                                const atdf_intervals* intervals, atdf_stream_t stream);

/* Training: loss = (1/R) sum over rays with positive target mass of the
 * cross-entropy -sum_b q_b log softmax_b, q_b = target_bins_b^target_power
 * normalised per ray (target_bins from atdf_project_fwd_bins), and grad
 * (f32, overwritten) its gradient w.r.t. params; loss (device double[1])
 * overwritten.  Fixed-order reductions: the result does not depend on
 * scheduling. */
size_t atdf_proposal_grad_workspace_bytes(int64_t num_rays);
atdf_status atdf_proposal_grad(const float* params, const float* features, const float* target_bins,
                               int64_t num_rays, float target_power, float* grad, double* loss, void* workspace,
                               size_t workspace_bytes, atdf_stream_t stream);

/* ======================================================================= */
/* Loss, optimizer, layout                                                  */
/* ======================================================================= */
/* MSE of TrainStep (samples/hyper_train.cpp:1015-1062, loss_weighted < 0):
 * loss = scale * mean((P - T)^2) in FP64; dLdP = scale * 2 (P - T) / R.
 * loss (device double[1]) is overwritten; dLdP is nullable. */
size_t atdf_mse_workspace_bytes(int64_t num_rays);
atdf_status atdf_mse_loss(const float* projection, const float* target, int64_t num_rays, float scale,
                          float* dLdP, double* loss, void* workspace, size_t workspace_bytes,
                          atdf_stream_t stream);

/* torch::optim::Adam step (used by OptimizerStep, geometry/geometry.h:203-225),
 * one fused pass; zero_grad != 0 also clears grad (optimizer zero_grad). */
atdf_status atdf_adam_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                           float lr, float beta1, float beta2, float eps, int64_t step, int32_t zero_grad,
                           atdf_stream_t stream);
/* Same, for bf16 parameters with an f32 master copy: master is updated and
 * rounded to param_bf16 (round to nearest even). */
atdf_status atdf_adam_step_bf16(float* master, uint16_t* param_bf16, float* grad, float* exp_avg,
                                float* exp_avg_sq, int64_t n, float lr, float beta1, float beta2, float eps,
                                int64_t step, int32_t zero_grad, atdf_stream_t stream);

/* ======================================================================= */
/* Regulariser                                                              */
/* ======================================================================= */
/* A batch of D-dimensional (D = 1..3) feature grids x[b, c, s_0 .. s_{D-1}]
 * at element offset b * stride_b + c * stride_c + sum_d s_d * stride_s[d]
 * (channels-first [B, C, ...] and channels-last [B, ..., C] alike). */
typedef struct {
    int64_t B, C;
    int32_t D;
    int32_t n[3];
    int64_t stride_b, stride_c;
    int64_t stride_s[3];
} atdf_tv_shape;

/* TVLoss::forward (geometry/geometry.h:31-81), as VolumeRegularizer applies
 * it to VM planes, lines and explicit grids (geometry_tensor_qo.cpp:
 * 527-607):
 *   loss[0] = scale * C * sum_d mean_b( w_b * mean_{c,s} |x[s] - x[s + e_d]| )
 * in FP64 (weight [B] nullable = 1; every n_d >= 2).  With grad non-null
 * the gradient (subgradient 0 at equal neighbours) is ADDED to grad, same
 * layout as x.  Deterministic (fixed-order reductions). */
size_t atdf_tv_workspace_bytes(int64_t num_elements);
atdf_status atdf_tv_loss(const float* x, const atdf_tv_shape* shape, const float* weight, double scale, double* loss,
                         float* grad, void* workspace, size_t workspace_bytes, atdf_stream_t stream);

/* Ring means of the Fourier regulariser (GeometryTensorQO::VolumeRegularizer,
 * geometry/geometry_tensor_qo.cpp:784-808) on img [Y, X, Z] f32 (the
 * shifted FFT magnitudes of the density slices): out [X / 2, Y], row i the
 * ring of radius i around (X / 2, Z / 2) -- n = (int)(2 pi i) angles of step
 * s = (float)(360 / n), m = ceil(360 / s) of them (torch::arange), and the
 * mean over all m^2 (row, column) pairs of torch::meshgrid of the row list
 * i cos(th) + X / 2 and the column list i sin(th) + Z / 2 of the bilinear
 * value (grid_sample: align_corners, border; both coordinates normalised by
 * X - 1).  Row 0, which the reference leaves uninitialised (torch::empty),
 * is +inf.  FP64 sums, deterministic.  Workspace: the row and column
 * weight tables of every ring and their non-zero ranges. */
size_t atdf_fourier_ring_workspace_bytes(int32_t X, int32_t Z);
atdf_status atdf_fourier_ring_means(const float* img, int64_t Y, int32_t X, int32_t Z, float* out, void* workspace,
                                    size_t workspace_bytes, atdf_stream_t stream);

/* Slot-major backward of the indirect gathers (IndirectGridSample3D / 2D /
 * 1D backward; same results as atdf_grid_sample{3,2,1}d_bwd up to the order
 * of the float additions): samples counting-sorted by slot, each slot's
 * gradient accumulated in LDS and added to grad_grid once per chunk of
 * samples, instead of one global atomic per corner and channel.  D = 3: size
 * = {Z, Y, X}, uv [S, 3]; D = 2: size = {H, W}, uv [S, 2]; D = 1: size = {W},
 * uv [S].  grid [N, C, size...]; index in [0, N); grad_grid and grad_uv
 * (nullable) are ADDED to.  A slot's gradient (C * cells floats) must fit in
 * one workgroup's LDS.  Workspace: atdf_grid_sample_bwd_workspace_bytes(N, S). */
size_t atdf_grid_sample_bwd_workspace_bytes(int64_t num_slots, int64_t num_samples);
/* The largest slot (C * cells floats) atdf_grid_sample_bwd_slots takes on the
 * current device. */
int64_t atdf_grid_sample_bwd_max_slot_floats(void);
atdf_status atdf_grid_sample_bwd_slots(int32_t D, const float* grid, int64_t N, int32_t C, const int32_t* size,
                                       const int64_t* index, const float* uv, int64_t S, const float* grad_out,
                                       float* grad_grid, float* grad_uv, void* workspace, size_t workspace_bytes,
                                       atdf_stream_t stream);

/* [N, C, G, G, G] <-> [N, G, G, G, C] (direction 0: to channels-last). */
atdf_status atdf_grid_layout(const float* src, float* dst, int64_t N, int32_t C, int32_t G, int32_t direction,
                             atdf_stream_t stream);
/* f32 -> bf16 (round to nearest even). */
atdf_status atdf_f32_to_bf16(const float* src, uint16_t* dst, int64_t n, atdf_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif

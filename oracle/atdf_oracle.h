/*
 * atdf_oracle.h -- CPU oracle for the reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the
 * smoke() entry and the CPU-baseline leg of bench.py load it; the product
 * path (libatdf.so + the libtorch layer) never links, loads or calls it.
 *
 * It restates, in plain C, the arithmetic of the reference path
 * (yuanhaowang1213/adaptivetensordf, hyperacorn/csrc):
 *   tree build            structure/HyperTree.cpp:298-525
 *   active ids            structure/HyperTree.cpp:550-588
 *   ray/box slab test     structure/HyperTree.h:338-380
 *   ray sample placement  structure/HyperTree.cu:266-514, 516-610
 *   local coordinates     structure/HyperTree.cu:237-259
 *   3-D indirect gather   modules/IndirectGridSample2D.cu:22-162 (2-D
 *                         pattern, extended to 3-D with torch grid_sample's
 *                         corner order; the 3-D source is not shipped)
 *   decoder MLP           geometry/Modules.h:23-145, geometry.cpp:176-253
 *   line integral         geometry/geometry.cpp:977-1019
 *   per-node loss         geometry/geometry.cpp:33-53
 * Every float expression is evaluated in the order written with no FMA
 * contraction (the Makefile passes -ffp-contract=off); those orders are
 * the canonical ones the HIP kernels reproduce bit for bit where the
 * kernels promise bit-exactness (traversal, sample placement, local
 * coordinates, trilinear gather).
 */
#ifndef ATDF_ORACLE_H
#define ATDF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tree ------------------------------------------------------------- */
int64_t atdfo_tree_num_nodes(int max_depth);
/* ns = 8 (octree) or 4 (quad-tree representation: children split x and z,
 * every node spans [y_min, y_max] in y; children [num_nodes, 4]) */
int64_t atdfo_tree_num_nodes_ns(int max_depth, int ns);
void atdfo_tree_build_ns(int max_depth, int ns, float y_min, float y_max, float* node_min, float* node_max,
                         int32_t* children, int32_t* parent, int32_t* depth);
/* BFS octree over [-1,1]^3: node 0 = root, children of node i contiguous,
 * child c takes the upper half on axis k iff bit k of c is set. */
void atdfo_tree_build(int max_depth, float* node_min, float* node_max, int32_t* children,
                      int32_t* parent, int32_t* depth);

/* ---- slab test -------------------------------------------------------- */
int atdfo_intersect_box_ray(const float* bmin, const float* bmax, const float* o, const float* d,
                            float* tmin, float* tmax);

/* ---- jitter hash (shared definition with the HIP kernels) ------------- */
float atdfo_jitter(uint64_t seed, int64_t ray, int64_t local_index);

/* ---- traversal + per-node sample placement ----------------------------
 * Brute force over active_ids (as the reference kernel does), hits sorted
 * by (tmin, node id).  jitter_mode: 0 = midpoint (a = 0.5), 1 = explicit
 * per-slot array `jitter` indexed by the canonical global sample slot,
 * 2 = hash(seed, ray, local index).
 * Pass 1 (samples == NULL): fills ray_num_samples / ray_num_intervals.
 * Pass 2: fills the sample arrays at ray_sample_offset[r] (exclusive scan).
 * Returns total samples, or -1 if a ray has more than max_intersections hits. */
int64_t atdfo_traverse(const float* node_min, const float* node_max, const int64_t* active_ids,
                       int64_t num_active, const float* ray_o, const float* ray_d, int64_t num_rays,
                       int K, int max_intersections, int jitter_mode, const float* jitter, uint64_t seed,
                       int32_t* ray_num_samples, int32_t* ray_num_intervals, const int64_t* ray_sample_offset,
                       float* global_coordinate, float* weight, float* ray_t, int64_t* ray_index,
                       int64_t* local_index_in_ray, int64_t* node_id,
                       /* optional interval outputs, indexed by ray_interval_offset */
                       const int64_t* ray_interval_offset, int32_t* iv_node, float* iv_tmin, float* iv_tmax,
                       int32_t* iv_count, int32_t* iv_first);

/* local = ((g - min) / (max - min)) * 2 - 1 */
void atdfo_local_samples(const float* node_min, const float* node_max, const float* g,
                         const int64_t* node, int64_t S, float* local);

/* ---- indirect 3-D trilinear gather, grid [N,C,Z,Y,X] ------------------ */
void atdfo_grid_sample3d_fwd(const float* grid, int64_t N, int C, int Z, int Y, int X,
                             const int64_t* index, const float* uv, int64_t S, float* out);
/* grad_grid accumulated in double, samples in index order */
void atdfo_grid_sample3d_bwd(const float* grid, int64_t N, int C, int Z, int Y, int X,
                             const int64_t* index, const float* uv, int64_t S, const float* grad_out,
                             double* grad_grid, double* grad_uv);

/* ---- decoder: C -> H -> H -> 1, hidden activation relu, softplus(beta)
 * params = [W1 (H x C), b1 (H), W2 (H x H), b2 (H), W3 (1 x H), b3 (1)]
 * (torch Linear weight layout [out, in]).  Sums run over the input index
 * in increasing order, starting from the bias-free product, bias added last. */
int64_t atdfo_mlp_num_params(int C, int H);
void atdfo_mlp_fwd(const float* params, int C, int H, float beta, const float* x, int64_t S,
                   float* y_pre, float* density);

/* ---- line integral: P[c, ray[s]] += v[s,c] * w[s], samples in order --- */
void atdfo_integrate(const float* v, const float* w, const int64_t* ray, int64_t S, int C,
                     int64_t R, float* P);

/* ---- per-node loss / weight sums in FP64 ------------------------------ */
void atdfo_accumulate_loss_per_node(const float* per_ray_loss, int C, int64_t R,
                                    const int64_t* ray_index, const float* weight,
                                    const int64_t* node_id, int64_t S, int64_t num_nodes,
                                    double* loss_sum, double* weight_sum);

/* ---- fused per-ray projection (forward) and its adjoint ----------------
 * The whole hot path for one batch of rays, per ray: traversal, sample
 * placement, local coords, gather from channels-last grids
 * grid[g, z, y, x, c] with g = node_to_grid[node], decoder, line integral.
 * precision 0: FP32 restatement (reference arithmetic, sequential sums).
 * precision 1: same inputs, every operation after sample placement in FP64.
 * Returns -1 on too many intersections. */
typedef struct {
    const float* node_min;
    const float* node_max;
    const int64_t* active_ids;
    int64_t num_active;
    const int32_t* node_to_grid;
    const float* grid;      /* [num_grids, G, G, G, C] channels-last */
    int G;
    int C;
    int H;
    const float* mlp;       /* atdfo_mlp layout */
    float beta;
    int K;
    int max_intersections;
    int jitter_mode;        /* 0 or 2 */
    uint64_t seed;
    int decoder_bf16;       /* 1: MFMA inputs rounded to bf16 (RNE) in the
                               FP64 path, as the BF16 decoder kernels do */
    /* Optional explicit intervals (e.g. the learned proposal's): when iv_off
     * is non-null, ray r's samples come from intervals
     * [iv_off[r], iv_off[r+1]) -- iv_count[h] stratified samples in
     * [iv_tmin[h], iv_tmax[h]] of node iv_node[h] -- instead of the
     * traversal. */
    const int64_t* iv_off;
    const int32_t* iv_node;
    const float* iv_tmin;
    const float* iv_tmax;
    const int32_t* iv_count;
} atdfo_model;

int atdfo_project_fwd(const atdfo_model* m, const float* ray_o, const float* ray_d, int64_t R,
                      int precision, double* P, int num_threads);
/* Same, plus (bins non-null) the per-ray mass histogram bins[r, b] = sum of
 * density * weight over the samples whose t falls in bin b of the ray's
 * chord [t0, t1] through the root box: b = (int)((t - t0) / (t1 - t0) * nb)
 * clamped to [0, nb - 1] (0 when t1 <= t0). */
int atdfo_project_fwd_bins(const atdfo_model* m, const float* ray_o, const float* ray_d, int64_t R,
                           int precision, double* P, double* bins, int nb, int num_threads);
/* Per-ray smallest |hidden pre-activation| over the ray's samples (FP64):
 * rays where it is tiny sit on a ReLU kink, where FP32 and FP64 may pick
 * different masks. */
void atdfo_project_min_preact(const atdfo_model* m, const float* ray_o, const float* ray_d, int64_t R,
                              double* out);
/* dLdP given per ray; gradients accumulated (+=) in FP64. */
int atdfo_project_bwd(const atdfo_model* m, const float* ray_o, const float* ray_d, int64_t R,
                      const double* dLdP, double* grad_grid, double* grad_mlp);

/* ---- per-node density statistics (FP64 reductions) -----------------------
 * For active node a (id active_ids[a]) evaluate the field's density at
 *   mode 0: the n^3 lattice of UniformGlobalSamples (HyperTree.cu:1618-1653):
 *           point (i,j,k) = (idx / (n-1)) * size + min, x fastest;
 *   mode 1: rounds x n^3 uniform points of CreateSamplesRandomly
 *           (HyperTree.cu:887-920): min + u * size, u = atdfo_point_uniform.
 * and reduce max (FP32), sum and sum of squares (FP64).  Densities are the
 * FP32 path (precision 0) or the FP64 path (precision 1).  Feeds
 * ComputeMaxDensity (hyper_train.cpp:662-707) and ComputeStd (709-759). */
float atdfo_point_uniform(uint64_t seed, int64_t node, int64_t point, int axis);
void atdfo_node_density_stats(const atdfo_model* m, int mode, int n, int rounds, uint64_t seed, int precision,
                              float* node_max, double* node_sum, double* node_sumsq, float* densities);

/* FP32 training-step baseline (reference structure on the CPU, OpenMP over
 * rays): fwd, MSE grad, bwd with FP32 accumulation, then (do_adam) Adam on
 * grid + mlp.
 * Used only as the timed CPU baseline.  Returns the MSE loss. */
double atdfo_train_step_f32(atdfo_model* m, float* grid_mut, float* mlp_mut, int64_t grid_numel,
                            const float* ray_o, const float* ray_d, const float* target, int64_t R,
                            float* grad_grid, float* grad_mlp, float* adam_m_grid, float* adam_v_grid,
                            float* adam_m_mlp, float* adam_v_mlp, int step, float lr, int num_threads,
                            int do_adam);

/* ---- tensorial (VM) features and helpers (atdf_oracle_vm.c) -----------
 * 2-D gather: grid [N, C, H, W], uv [S, 2] (u -> W, v -> H); 1-D gather:
 * grid [N, C, W, 1], uv [S]; align_corners mapping, clipped corners. */
void atdfo_grid_sample2d_fwd(const float* grid, int64_t N, int C, int H, int W, const int64_t* index,
                             const float* uv, int64_t S, float* out);
void atdfo_grid_sample2d_bwd(const float* grid, int64_t N, int C, int H, int W, const int64_t* index,
                             const float* uv, int64_t S, const float* grad_out, double* grad_grid, double* grad_uv);
void atdfo_grid_sample1d_fwd(const float* grid, int64_t N, int C, int W, const int64_t* index, const float* uv,
                             int64_t S, float* out);
void atdfo_grid_sample1d_bwd(const float* grid, int64_t N, int C, int W, const int64_t* index, const float* uv,
                             int64_t S, const float* grad_out, double* grad_grid, double* grad_uv);
/* GradientCorrection backward: g - (g . d) d with d the sample's ray direction */
void atdfo_gradient_correction(const float* ray_direction, const int64_t* ray_index, const float* grad_out,
                               int64_t S, float* grad_pos);
/* ComputeBoundWeight: out[s] = s if node_id[s] is in in_roi_node, else -1 */
void atdfo_compute_bound_weight(const int64_t* node_id, int64_t S, const int64_t* in_roi_node, int64_t n_in,
                                int64_t* out);

/* ---- learned sample proposal (atdf_oracle_proposal.c) -----------------
 * MLP 38 -> 64 -> 64 -> 64 -> 32 (ReLU hidden), inputs ray origin (3),
 * direction (3) and the coarse occupancy of 32 equal bins of the ray's chord
 * through the root box; softmax over the 32 logits is a piecewise-constant
 * pdf over the bins.  params (torch Linear layout [out, in]):
 *   W1 [64,38] b1 [64] W2 [64,64] b2 [64] W3 [64,64] b3 [64] W4 [32,64] b4 [32]
 * Sums run over the input index in increasing order from zero, bias added
 * last, in FP32 with no contraction: the HIP kernels reproduce the logits,
 * probabilities, KL and every sample placement bit for bit. */
#define ATDFO_PROP_BINS 32
#define ATDFO_PROP_IN 38
#define ATDFO_PROP_H 64
int64_t atdfo_prop_num_params(void);
/* exp / log used by the proposal (Cody-Waite + degree-7 Taylor; atanh series) */
float atdfo_prop_exp(float x);
float atdfo_prop_log(float x);

/* Active node id (or -1) of every cell of the (2^depth)^3 lattice over the
 * root box (x fastest): descend from the root through the child whose half
 * holds the cell centre until an active node or a leaf. */
void atdfo_leaf_table_ns(const float* node_min, const float* node_max, const int32_t* children, int ns,
                         const int32_t* active, int depth, int32_t* table);
void atdfo_leaf_table(const float* node_min, const float* node_max, const int32_t* children,
                      const int32_t* active, int depth, int32_t* table);
int32_t atdfo_locate(const float* root_min, const float* root_max, const int32_t* table, int depth,
                     const float* p);

typedef struct {
    const float* root_min;   /* root box */
    const float* root_max;
    const int32_t* table;    /* atdfo_leaf_table */
    int depth;
    const float* coarse;     /* [num_nodes] coarse density of the active nodes */
    const float* params;
    float mix;               /* uniform mixing weight of the sampling pdf */
    float kl_threshold;      /* fall back when KL(coarse || pdf) exceeds it */
    int jitter_mode;         /* 0 midpoint, 2 hash: where the fused kernels put
                                each cell's sample (its node is located there) */
    uint64_t jitter_seed;
} atdfo_proposal;

/* features [R, 38]; chord [R, 2] (t0, t1; zeros when the chord is shorter
 * than 1e-5 or missing). */
void atdfo_prop_features(const atdfo_proposal* p, const float* o, const float* d, int64_t R, float* feat,
                         float* chord);
void atdfo_prop_forward(const float* params, const float* feat, int64_t R, float* logits, float* probs);
/* Per ray: fallback flag, KL, and (learned rays) up to 32 count-1 intervals,
 * slot-major: node[i * R + r], tmin[i * R + r], tmax[i * R + r], i <
 * n_learned[r].  The cells are [T_i, T_i+1] with T_0 = t0, T_32 = t1 and
 * T_i (0 < i < 32) the inverse CDF of the sampling pdf at i / 32; empty
 * cells and cells whose sample lies where no active node is are dropped. */
void atdfo_prop_sample(const atdfo_proposal* p, const float* o, const float* d, int64_t R, int32_t* fallback,
                       int32_t* n_learned, int32_t* node, float* tmin, float* tmax, float* kl);
void atdfo_prop_allocate(const atdfo_proposal* p, const float* o, const float* d, int64_t R, int32_t extra,
                         const int64_t* iv_off, const float* tmin, const float* tmax, int32_t* count, int32_t* first,
                         int32_t* fallback, int32_t* ns, float* kl_out);
/* FP64 cross-entropy of the proposal pdf against the target q_b ~
 * bins_b^power normalised per ray (bins [R, 32] the binned masses; rays with
 * zero mass skipped), divided by R; grad (overwritten, FP64) of that loss
 * w.r.t. params.  Returns the loss. */
double atdfo_prop_grad(const float* params, const float* feat, const float* bins, int64_t R, float power,
                       double* grad);

/* Point location (structure/HyperTree.cu:1339-1399): the active node reached
 * from the root through the first child whose closed box holds p, or -1. */
int32_t atdfo_active_node_for_position(const float* node_min, const float* node_max, const int32_t* children,
                                       const int32_t* active, const float* p);
/* the same over a tree of `ns` children per node (4: the quad-tree
 * representation, children [num_nodes, 4]) */
int32_t atdfo_active_node_for_position_ns(const float* node_min, const float* node_max, const int32_t* children,
                                          int ns, const int32_t* active, const float* p);
void atdfo_node_id_for_position_ns(const float* node_min, const float* node_max, const int32_t* children, int ns,
                                   const int32_t* active, int64_t first_active, const float* pos, int64_t n,
                                   int64_t* node_id, float* mask);
void atdfo_grid_transfer_ns(const float* node_min, const float* node_max, const int32_t* children, int ns,
                            const int32_t* src_active, const int32_t* src_slot, const float* src_grid,
                            const int64_t* dst_nodes, int64_t n_dst, const int32_t* dst_slot, float* dst_grid, int C,
                            int G, int layout);
/* NodeIdForPositionGPU: -1 becomes first_active with mask 0. */
void atdfo_node_id_for_position(const float* node_min, const float* node_max, const int32_t* children,
                                const int32_t* active, int64_t first_active, const float* pos, int64_t n,
                                int64_t* node_id, float* mask);
/* Nearest-lattice grid transfer (InterpolateGridForInactiveNodes copy rule)
 * into the dst_nodes' grids (a node active in the source set keeps its own
 * grid); slots null = identity; layout 0
 * [slots, G, G, G, C], layout 1 [slots, C, G, G, G]. */
void atdfo_grid_transfer(const float* node_min, const float* node_max, const int32_t* children,
                         const int32_t* src_active, const int32_t* src_slot, const float* src_grid,
                         const int64_t* dst_nodes, int64_t n_dst, const int32_t* dst_slot, float* dst_grid, int C,
                         int G, int layout);
/* Lattice points of volume export: mode 0 (i / (size - 1)) * 2 - 1, mode 1
 * roi_min + step * i; axis 0 fastest (axis 2 with swap_xy); pos [n, 3]. */
void atdfo_volume_points(int mode, int swap_xy, const int* size, const float* roi_min, const float* step,
                         float* pos);

/* Scene data path.  Quaternions (x, y, z, w); rotation v' = (v + w t) +
 * q x t with t = 2 q x v, in f64. */
void atdfo_rotate_point(const double* q, const double* v, double* out);
/* SceneBase::GetRays: camera_model 0 pinhole, 1 orthographic; outputs f32. */
void atdfo_generate_rays(int camera_model, int w, int h, int align_corners, const double* intrinsics,
                         const double* rotation, const double* translation, const float* vol_translate,
                         const float* uv, const int32_t* image_id, const int32_t* camera_id, int64_t n,
                         float* origin, float* direction);
/* LoadImagesCT log conversion; mask nullable. */
void atdfo_log_convert(const float* raw, const float* mask, int64_t n, double xray_min, double xray_max,
                       double projection_factor, int log10_conversion, int log_space_input, float* out);
/* grid_sample bilinear / border of [images, h, w] projections at uv. */
void atdfo_sample_projection(const float* projection, const float* mask, int w, int h, int align_corners,
                             const float* uv, const int32_t* image_id, int64_t n, float* target, float* target_mask);
/* PointInAnyImage visibility of points [n, 3] against a rig (camera 0). */
void atdfo_point_in_any_image(int camera_model, int w, int h, const double* intrinsics, const double* rotation,
                              const double* translation, int64_t num_images, const float* roi_min,
                              const float* roi_max, const float* points, int64_t n, float* mask);
/* UniformGlobalSamples: the n^3 lattice of each listed node, x fastest
 * (the points of atdfo_node_density_stats mode 0); pos [num_active, n^3, 3]. */
void atdfo_node_lattice_points(const float* node_min, const float* node_max, const int64_t* active_ids,
                               int64_t num_active, int n, float* pos);
/* TVLoss of a strided [B, C, n...] grid (D = 1..3): returns the FP64 loss,
 * adds the gradient to grad (nullable). */
double atdfo_tv_loss(const float* x, int64_t B, int64_t C, int D, const int* n, int64_t sb, int64_t sc,
                     const int64_t* ss, const float* weight, double scale, float* grad);
/* NodeNeighborSamples pairs in candidate order; returns the pair count and
 * writes at most cap pairs (pos [2P, 3], node [2P]). */
int64_t atdfo_neighbor_pairs_ns(const float* node_min, const float* node_max, const int32_t* children, int ns,
                                const int32_t* active, const int64_t* active_ids, int64_t num_active, const int* size,
                                float eps, int use_roi, const float* roi_min, const float* roi_max, int variant,
                                int plane, float* pos, int64_t* node_out, int64_t cap);
int64_t atdfo_neighbor_pairs(const float* node_min, const float* node_max, const int32_t* children,
                             const int32_t* active, const int64_t* active_ids, int64_t num_active, const int* size,
                             float eps, int use_roi, const float* roi_min, const float* roi_max, int variant,
                             float* pos, int64_t* node_out, int64_t cap);
/* Density of the fused field at points pos [n, 3] located in node[n]
 * (< 0: 0), FP32 path (precision 0) or FP64 path (precision 1). */
void atdfo_decode_features(const float* params, int C, int H, double beta, const float* x, int64_t S, int bf,
                           const double* dd, double* density, double* grad_x, double* grad_mlp);
void atdfo_fourier_ring_means(const float* img, int64_t Y, int X, int Z, float* out);
void atdfo_field_density(const atdfo_model* m, const float* pos, const int64_t* node, int64_t n, int precision,
                         double* density);
/* Random batch draws (image entry, uv) of the batch kernel. */
void atdfo_random_pixels(const int32_t* indices, int64_t num_indices, const int32_t* image_camera, int w, int h,
                         int align_corners, int interpolate, uint64_t seed, int64_t n, float* uv, int32_t* image_id,
                         int32_t* camera_id);

#ifdef __cplusplus
}
#endif
#endif

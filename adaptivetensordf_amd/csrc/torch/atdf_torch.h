// atdf_torch.h -- libtorch host layer over the C ABI (include/atdf.h).
//
// Recreates the reference's C++ operator API for the reconstruction hot
// path (yuanhaowang1213/adaptivetensordf, hyperacorn/csrc) so callers of
// HyperTreeBaseImpl / IndirectGridSample3D / IntegrateSamplesXRay /
// AccumulateSampleLossPerNode can switch over, and adds the fused
// projection with its hand-written adjoint behind a
// torch::autograd::Function.  All device work goes through libatdf.so; this
// layer only allocates tensors, checks shapes and orders the calls on the
// current HIP stream.
#pragma once

#include <torch/torch.h>

#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../../include/atdf.h"

namespace atdf_torch {

atdf_stream_t current_stream();
void check_status(atdf_status s, const char* what);

// structure/HelperStructs.h:103-175
struct RayList {
    torch::Tensor origin;     // [R, 3] f32
    torch::Tensor direction;  // [R, 3] f32
    int64_t size() const { return origin.defined() ? origin.size(0) : 0; }
};

// structure/HelperStructs.h:231-297
struct SampleList {
    torch::Tensor global_coordinate;   // [S, 3] f32
    torch::Tensor weight;              // [S] f32
    torch::Tensor ray_t;               // [S] f32
    torch::Tensor ray_index;           // [S] i64
    torch::Tensor local_index_in_ray;  // [S] i64
    torch::Tensor node_id;             // [S] i64
    int max_samples_per_ray = 0;
    int64_t size() const { return weight.defined() ? weight.size(0) : 0; }
};

// Per-ray node intervals: the compact traversal result the fused kernels
// expand into samples on the fly.
struct Intervals {
    torch::Tensor ray_offset;  // [R + 1] i64
    torch::Tensor node;        // [H] i32
    torch::Tensor tmin;        // [H] f32
    torch::Tensor tmax;        // [H] f32
    torch::Tensor count;       // [H] i32
    torch::Tensor first;       // [H] i32
    torch::Tensor ray;         // [H] i32
    torch::Tensor ray_num_samples;  // [R] i32
    int64_t num_intervals = 0;
    int64_t num_samples = 0;
    int max_samples_per_ray = 0;
    atdf_intervals view() const;
};

struct SamplingParams {
    int max_samples_per_node = 32;  // Settings.h:258
    int max_intersections = 256;    // HyperTree.cu:280
    int jitter_mode = 0;            // 0 midpoint, 1 array, 2 hash
    uint64_t jitter_seed = 0;
    atdf_sampling view(const float* jitter = nullptr) const;
};

// Split of the active nodes by the reconstruction ROI (tensorial VM
// features inside, explicit grids outside), geometry_tensor_qo.cpp:940-1030.
struct RoiMaps {
    torch::Tensor in_roi_node_index;   // [N_in] i64 node ids
    torch::Tensor out_roi_node_index;  // [N_out] i64
    torch::Tensor prefix_inroi;        // [Nn] i64 slot in the VM planes, -1 elsewhere
    torch::Tensor prefix_outroi;       // [Nn] i64 slot in the explicit grids, -1 elsewhere
    torch::Tensor in_roi_flag;         // [Nn] i32 0/1
    std::vector<float> roi_min, roi_max;
};

// HyperTreeBaseImpl (structure/HyperTree.h:395-563) for D = 3, octree.
class HyperTreeBase {
   public:
    HyperTreeBase(int d, int max_depth);
    // HyperTree.cpp:298-525; use_quad: the quad-tree representation (four
    // children per node splitting x and z, every node spanning [y_min, y_max])
    void settree(bool use_quad = false, float y_min = -1.f, float y_max = 1.f);
    int ChildrenPerNode() const { return (int)node_children.size(1); }
    void SetActive(int depth);           // HyperTree.cpp:544-549
    void UpdateActive();                 // HyperTree.cpp:550-588
    void to(torch::Device device);
    int64_t NumNodes() const { return node_position_min.size(0); }
    int64_t NumActiveNodes() const { return active_node_ids.size(0); }
    torch::Tensor ActiveNodeTensor() const { return active_node_ids; }
    torch::Device device() const { return node_position_min.device(); }

    // HyperTree.cu:516-610
    SampleList CreateSamplesForRays(const RayList& rays, int max_samples_per_node, bool interval_jitter);
    Intervals CreateIntervalsForRays(const RayList& rays, const SamplingParams& sp);
    // HyperTree.cu:818-883 (ComputeRaySamples_area, 651-816): one box, the
    // dataset ROI, for depth-0 trees (samples/hyper_train.cpp:921-927); every
    // sample has node id 0, so local coordinates stay relative to node 0's box
    SampleList CreateSamplesForRays_area(const RayList& rays, int max_samples_per_node, bool interval_jitter,
                                         std::vector<float> roi_min, std::vector<float> roi_max);
    Intervals CreateIntervalsForRays_area(const RayList& rays, const SamplingParams& sp, std::vector<float> roi_min,
                                          std::vector<float> roi_max);
    // HyperTree.cu:237-259
    torch::Tensor ComputeLocalSamples(torch::Tensor global_samples, torch::Tensor node_indices);
    // HyperTree.cpp:607-647
    void SetErrorForActiveNodes(torch::Tensor error, const std::string& strategy);
    // HyperTree.cpp:666-705
    void UpdateCulling();
    // in_roi test of geometry/geometry.h:426-449 over the active nodes
    RoiMaps RoiSplit(std::vector<float> roi_min, std::vector<float> roi_max, float epsilon) const;
    // NodeIdForPositionGPU (HyperTree.cu:1487-1525): (node_id [P] i64, mask [P] f32)
    // for positions [..., 3]; outside every active node: first active node, 0
    std::pair<torch::Tensor, torch::Tensor> NodeIdForPositionGPU(torch::Tensor global_samples) const;
    // NodeNeighborSamples (HyperTree.cu:1720-1811, 2010-2075): pairs of samples
    // (rows 2k, 2k + 1: one point in two face-adjacent active nodes), in
    // candidate order; weight 1, local_index_in_ray 0 / 1
    SampleList NodeNeighborSamples(std::vector<int> size, double epsilon, std::vector<float> roi_min,
                                   std::vector<float> roi_max, bool use_roi) const;
    // NodeNeighborSamplesPlane2D (HyperTree.cu:2076-2160), the pairs the
    // tensorial geometry's edge term uses (geometry_tensor_qo.cpp:1158-1160):
    // with use_quad = false the NodeNeighborSamplesQuad kernel -- faces normal
    // to x and z only, in_roi_quad (y range widened by epsilon / 2), point
    // location over the first four children of each node, as the reference
    // does on any tree; plane_index is unused.  use_quad = true: the
    // quad-tree representation's NodeNeighborSamples2D on plane plane_index
    // (HyperTree.cu:1815-1913).
    SampleList NodeNeighborSamplesPlane2D(std::vector<int> size, double epsilon, int plane_index,
                                          std::vector<float> roi_min, std::vector<float> roi_max, bool use_roi,
                                          bool use_quad) const;
    // UniformGlobalSamples (HyperTree.cu:1618-1665): [N, n, n, n, 3] lattice
    // positions of the listed nodes (x fastest)
    torch::Tensor UniformGlobalSamples(torch::Tensor node_id, int n) const;
    // InterpolateGridForInactiveNodes (HyperTree.cu:1603-1616): a copy of
    // active_grid [num_nodes, C, G, G, G] whose inactive nodes take the
    // nearest lattice values of the active nodes
    torch::Tensor InterpolateGridForInactiveNodes(torch::Tensor active_grid) const;
    // UniformPhantomSamplesGPU / UniformPhantomSamplesGPUbySlice (HyperTree.cu:
    // 1158-1293): lattice points (axis 0 fastest, axis 2 with swap_xy) as a
    // SampleList: global_coordinate, ray_index = point index, node_id and
    // weight = NodeIdForPositionGPU's (node, mask)
    SampleList UniformPhantomSamplesGPU(std::vector<int> size, bool swap_xy) const;
    SampleList UniformPhantomSamplesGPUbySlice(std::vector<int> size, bool swap_xy, std::vector<float> roi_min,
                                               std::vector<float> step_size) const;
    // packed traversal form of the current active set (replaced, not
    // modified, on every change: holding it keeps a snapshot)
    torch::Tensor PackedTree() const { return packed_; }

    const void* packed() const { return packed_.data_ptr(); }
    // point-location table of the active set (rebuilt with the packed tree)
    const int32_t* leaf_table() const { return leaf_table_.data_ptr<int32_t>(); }
    int leaf_depth() const { return max_depth < 1 ? 1 : max_depth; }
    torch::Tensor LeafTable() const { return leaf_table_; }

    int d;
    int max_depth;
    torch::Tensor node_parent, node_children, node_position_min, node_position_max, node_scale, node_error,
        node_max_density, node_active, node_culled, node_depth, node_active_prefix_sum, active_node_ids;
    torch::Tensor node_to_grid;  // i32 copy of node_active_prefix_sum for the kernels

   private:
    void repack();
    Intervals intervals_for(const void* packed, int64_t num_nodes, const RayList& rays, const SamplingParams& sp) const;
    SampleList samples_for(const void* packed, int64_t num_nodes, const RayList& rays, int max_samples_per_node,
                           bool interval_jitter) const;
    SampleList neighbor_samples(std::vector<int> size, double epsilon, std::vector<float> roi_min,
                                std::vector<float> roi_max, bool use_roi, int variant, int plane = 0) const;
    torch::Tensor packed_;
    torch::Tensor leaf_table_;
};

// ---- octree structure optimizer (structure/HyperTreeStructureOptimizer.{h,cpp}) ----
struct TreeOptimizerParams {           // HyperTreeStructureOptimizer.h:6-30
    int max_active_nodes = 512;
    double error_merge_factor = 1.1;
    double error_split_factor = 0.75;
    bool use_saved_errors = true;
    std::vector<float> tree_edge_roi_min{-1.f, -1.f, -1.f}, tree_edge_roi_max{1.f, 1.f, 1.f};
};
struct StructureDecision {
    std::vector<int> choice;           // per node: -1 not active, 0 none, 1 split, 2 merge
    double objective = 0;
    int64_t active_weight = 0;         // budget used
    int num_none = 0, num_split = 0, num_grp = 0;
    bool feasible = false;
};
// Exact optimum of the reference's structure MIP (multiple-choice knapsack DP).
StructureDecision SolveTreeStructure(const HyperTreeBase& tree, const TreeOptimizerParams& params);
// Solve, apply the split / merge decisions (ApplySplit) and UpdateActive;
// returns whether the active set changed (OptimizeTree).
bool OptimizeTreeStructure(HyperTreeBase& tree, const TreeOptimizerParams& params, StructureDecision* out = nullptr);
// NodeCulling (samples/hyper_train.cpp:761-830).
void NodeCulling(HyperTreeBase& tree, float culling_threshold, std::vector<float> roi_min, std::vector<float> roi_max,
                 float epsilon);
// Grids [N_active, G, G, G, C] f32 of the tree's current active set from the
// grids of an earlier active set (old_packed = PackedTree() and old_node_to_grid
// = node_to_grid taken before the change): nodes active in both keep their
// grid, new nodes take the nearest lattice values of the old active nodes
// (InterpolateGridForInactiveNodes' rule).  Lattice points of a new node
// that no old active node covers keep init's values (zeros if undefined).
torch::Tensor TransferNodeGrids(const HyperTreeBase& tree, torch::Tensor old_packed, torch::Tensor old_node_to_grid,
                                torch::Tensor old_grid, torch::Tensor init = {});

// IndirectGridSample3D(multi_grid, index, uv): grid [N, C, Z, Y, X]
torch::Tensor IndirectGridSample3D(torch::Tensor multi_grid, torch::Tensor index, torch::Tensor uv);

// IndirectGridSample2D / 1D (modules/IndirectGridSample2D.h:10-24,
// IndirectGridSample1D.cu): grid [N, C, H, W] with uv [S, 2]; grid
// [N, C, W, 1] with uv [S, 1].
torch::Tensor IndirectGridSample2D(torch::Tensor multi_grid, torch::Tensor index, torch::Tensor uv);
torch::Tensor IndirectGridSample1D(torch::Tensor multi_grid, torch::Tensor index, torch::Tensor uv);
// GradientCorrection (modules/GradientCorrection.h:11-25): identity forward,
// backward removes the gradient component along the sample's ray.
torch::Tensor GradientCorrection(torch::Tensor sample_position, torch::Tensor ray_direction, torch::Tensor index);
// HyperTreeBaseImpl::ComputeBoundWeight(node_id, in_roi_node, false) (HyperTree.cu:962-993)
torch::Tensor ComputeBoundWeight(torch::Tensor node_id, const RoiMaps& roi);

// Tensorial (VM) fields of the in-ROI nodes and explicit grids of the rest
// (GeometryTensorQO, geometry_tensor_qo.cpp:257-380).
struct VMFields {
    torch::Tensor plane_yz, plane_xz, plane_xy;  // [N_in, C, G, G] (u -> last axis)
    torch::Tensor line_x, line_y, line_z;        // [1, C, G, 1]
    torch::Tensor grid_out;                      // [N_out, C', G, G, G] (may be undefined)
    std::string tensor_op = "add";               // "add" or "cat"
    std::vector<float> roi_min{-1.f, -1.f, -1.f}, roi_max{1.f, 1.f, 1.f};
};
// GeometryTensorQO::SampleVolumeIndirect_feature: features [S, C] of the
// samples, VM plane x line products for in-ROI nodes, 3-D gathers for the
// others, merged back in sample order (autograd through every part).
torch::Tensor SampleVolumeIndirectFeature(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields,
                                          torch::Tensor global_coordinate, torch::Tensor node_id);

// Parameter-store initialisation.  Node grids (ExplicitFeatureGridImpl,
// geometry/ImplicitNet.h:66-100): "uniform" / "random" U(-s, s), "minus" -1,
// "zero" 0, plus "normal" N(0, s^2); shape [N, G, G, G, C] (channels_last) or
// the reference's [N, C, G, G, G].  VM planes / lines (Tensor_PlaneQuadImpl /
// Tensor_LineQuadImpl, geometry/tensorfield.h:9-127, 218-335): "uniform"
// U(-s, s) or "random" s (U[0, 1) - 1) 2, which lies in (-2s, 0].  Drawn from
// libtorch's generator seeded with `seed`, in f32 on `device`.
torch::Tensor InitNodeGrid(int64_t num_grids, int C, int G, const std::string& init, double scale, uint64_t seed,
                           bool channels_last, const std::string& device);
torch::Tensor InitTensorCoefficients(std::vector<int64_t> shape, const std::string& init, double scale, uint64_t seed,
                                     const std::string& device);

// TVLoss::forward (geometry/geometry.h:31-81) on grid [B, C, n...] (1 to 3
// spatial dims; channels_last: [B, n..., C]) with an optional per-batch weight,
// as an autograd Function (loss and gradient from one atdf_tv_loss launch).
torch::Tensor TVLoss(torch::Tensor grid, torch::Tensor weight = {}, bool channels_last = false);
struct FieldParams;
// The Fourier regulariser's switches (TrainParams, Settings.h:473-478,
// 600-605).  The term is on when loss_fourier > 0.
struct FourierParams {
    float loss_fourier = -0.0001f;
    float scale = 1.0f;    // loss_fourier_scale
    float scale_y = 1.0f;  // loss_fourier_scale_y
    float shift = 0.05f;   // loss_fourier_shift
    float rate = 0.2f;     // loss_fourier_rate: the sigmoid factor's knee, in rings
    float rate_w = 0.01f;  // loss_fourier_rate_w: half width of the central bands
};
// The Fourier term's lattice (geometry_tensor_qo.cpp:1101-1145): X x Y x Z
// points over the ROI (y inset by `shift`; X and Z made odd), x slowest, and
// their NodeIdForPositionGPU nodes.
struct FourierGrid {
    torch::Tensor coords;   // [X * Y * Z, 3]
    torch::Tensor node_id;  // [X * Y * Z] i64
    int64_t X = 0, Y = 0, Z = 0;
};
FourierGrid MakeFourierGrid(HyperTreeBase& tree, int grid_size, int tree_depth, std::vector<float> roi_min,
                            std::vector<float> roi_max, const FourierParams& fp);
// DecodeFeatures (geometry/geometry.cpp:176-253, softplus last activation) on
// features [S, C] -> [S, 1], autograd over atdf_decode_features(_bwd).
torch::Tensor DecodeFeatures(torch::Tensor features, torch::Tensor mlp, const FieldParams& fp);
// GeometryTensorQO::SampleVolumeIndirect (geometry_tensor_qo.cpp:381-395):
// DecodeFeatures(SampleVolumeIndirectFeature(...)).
torch::Tensor SampleVolumeIndirect(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields, torch::Tensor mlp,
                                   const FieldParams& fp, torch::Tensor global_coordinate, torch::Tensor node_id);
// HierarchicalNeuralGeometry::ComputeImage (geometry.cpp:828-914) for the
// tensorial geometry: SampleVolumeIndirect of the samples, then
// IntegrateSamplesXRay with their weights -> [1, num_rays] (autograd
// through the VM gathers, the decoder and the integration).
torch::Tensor ComputeImageVM(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields, torch::Tensor mlp,
                             const FieldParams& fp, const SampleList& samples, int64_t num_rays);
// The tensorial geometry through the fused kernels: the projection [R] of
// ComputeImageVM for tensor_op "add" with samples recomputed from the
// rays (CreateIntervalsForRays with sp) inside atdf_project_fwd /
// atdf_project_bwd instead of materialised; autograd to every plane, line,
// grid_out and the decoder.  tree and roi must outlive the backward pass.
torch::Tensor ProjectRaysVM(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields, torch::Tensor mlp,
                            const SamplingParams& sp, const FieldParams& fp, const RayList& rays);
// Training state of the tensorial geometry in the fused (channels-last)
// layouts: planes [N_in, 3, Gp, Gp, C], lines [3, Gl, C], explicit grids
// [N_out, G, G, G, C], decoder; gradients and Adam moments per parameter in
// that order (the gradients are zero between steps).
struct VMTrainState {
    torch::Tensor planes, lines, grid, mlp;
    std::vector<torch::Tensor> grads, m, v;
    torch::Tensor loss;  // [1] f64 on device
    int64_t step = 0;
    float lr = 1e-3f;
    float roi_min_x = -1.f, roi_max_x = 1.f;
};
VMTrainState MakeVMTrainState(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields, torch::Tensor mlp,
                              float lr, const FieldParams& fp);
// The state's fields in the reference layouts (tensor_op and ROI from `like`).
VMFields VMTrainStateFields(const VMTrainState& st, const VMFields& like);
// One step: intervals, fused forward, MSE, fused adjoint, Adam on planes,
// lines, grids and decoder (the reference's optimizer step over the
// tensorial geometry's parameters, geometry_tensor_qo.cpp:199-255).
torch::Tensor TrainStepVM(HyperTreeBase& tree, const RoiMaps& roi, VMTrainState& st, const RayList& rays,
                          torch::Tensor target, const SamplingParams& sp, const FieldParams& fp);
// Ring means [X / 2, Y] of img [Y, X, Z] (atdf_fourier_ring_means; row 0 +inf).
torch::Tensor FourierRingMeans(torch::Tensor img);
// The Fourier term on densities [X, Y, Z] (geometry_tensor_qo.cpp:758-850,
// the masked branch): per y slice the shifted, forward-normalised 2-D FFT
// magnitude over (x, z); its central band means along x and z against the
// ring profile (detached), where above it, times a sigmoid factor in the
// radius; loss_fourier * (mean_x + mean_y).  FFT through torch::fft
// (rocFFT); ring means on the device kernel.  Requires X == Z.
torch::Tensor FourierLoss(torch::Tensor densities, const FourierParams& fp);
// The regulariser switches of TrainParams (Settings.h:429-607).
struct RegularizerParams {
    double loss_tv = 0.0005;
    double loss_l2 = -0.001;
    float test_line_x = 1, test_line_y = 1, test_line_z = 1;
    bool plane_vec_only = true;
    bool use_loss_tv_3d = true;
    float test_para = -1;
    std::vector<int64_t> tv_size{16, 16, 16};  // lattice of the 3-D TV: vec_grid_size * loss_tv_scale
    double loss_edge = 0.002;
    bool use_ground_truth_volume = false;  // drops the out-of-ROI edge term
    FourierParams fourier;
};
// The edge term of VolumeRegularizer (geometry_tensor_qo.cpp:630-667): for
// neighbour pairs (NodeNeighborSamples) the features of both samples, split
// by the ROI membership of the pair's first node; per group
// mean_pairs sqrt(mean_c (f1 - f2)^2) * loss_edge * channels.
torch::Tensor EdgeLoss(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields, const SampleList& pairs,
                       const RegularizerParams& p);
// GeometryTensorQO::VolumeRegularizer (geometry_tensor_qo.cpp:527-622 and
// 913-922): the TV terms (3-D TV of the features on a lattice of each in-ROI
// node, or TV of the planes, lines and explicit grids), the L2 terms and,
// given neighbour pairs, the edge term, and given a FourierGrid with the
// decoder, the Fourier term on the decoded densities of its lattice.
// Undefined when no term is on.
torch::Tensor VolumeRegularizer(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields,
                                const RegularizerParams& p, const SampleList* neighbor_pairs = nullptr,
                                const FourierGrid* fourier = nullptr, torch::Tensor mlp = {},
                                const FieldParams* field = nullptr);

// NeuralGeometry::IntegrateSamplesXRay (geometry.cpp:977-1019)
torch::Tensor IntegrateSamplesXRay(torch::Tensor sample_values, torch::Tensor integration_weight,
                                   torch::Tensor ray_index, int num_channels, int64_t num_rays);

// HierarchicalNeuralGeometry::AccumulateSampleLossPerNode (geometry.cpp:33-53),
// FP64 sums.
std::pair<torch::Tensor, torch::Tensor> AccumulateSampleLossPerNode(const SampleList& samples,
                                                                    torch::Tensor per_ray_loss, int64_t num_nodes);

// Field parameters of the fused path.
struct FieldParams {
    int C = 16;
    int H = 64;
    int G = 11;
    float softplus_beta = 2.0f;
    int rays_per_wave = 0;
    bool bf16_decoder = false;  // BF16 MFMA inputs, FP32 accumulation
    bool node_forward = true;   // training step: node-major forward over the adjoint's sort
    bool two_pass_adjoint = true;  // workspace with the adjoint's staging region (atdf_project_bwd_staging_bytes)
};

// C-ABI view of a field (checks shapes and dtypes).
atdf_field make_field(const HyperTreeBase& tree, const torch::Tensor& grid, const torch::Tensor& mlp,
                      const FieldParams& fp);

// Fused projection with hand-written adjoint (torch::autograd::Function).
// grid: [num_grids, G, G, G, C] f32 (or bf16 for inference-only forward),
// mlp: flat decoder parameters.  Returns the projection [R].
torch::Tensor ProjectRays(HyperTreeBase& tree, const RayList& rays, torch::Tensor grid, torch::Tensor mlp,
                          const SamplingParams& sp, const FieldParams& fp);

// HierarchicalNeuralGeometry::UniformSampledVolume (geometry/geometry.cpp:
// 347-660, non-final path) for the fused field: density on the shape[0] x
// shape[1] x shape[2] lattice over the unit cube, returned as
// (volume [1, shape...] f32 holding point i at flat index i (axis 0
// fastest), node_id [shape...] i64, valid [shape...] i64).
std::tuple<torch::Tensor, torch::Tensor, torch::Tensor> UniformSampledVolume(HyperTreeBase& tree, torch::Tensor grid,
                                                                             torch::Tensor mlp, const FieldParams& fp,
                                                                             std::vector<int> shape);
// The same on the lattice roi_min + step * i (the by-slice sampler).
torch::Tensor SampleVolumeBySlice(HyperTreeBase& tree, torch::Tensor grid, torch::Tensor mlp, const FieldParams& fp,
                                  std::vector<int> shape, std::vector<float> roi_min, std::vector<float> step);
// SaveHDRImageTensor / LoadHDRImageTensor (utils/cimg_wrapper.cpp:12-52):
// a 4-D f32 volume [s0, s1, s2, s3] as an Analyze 7.5 pair (path ".hdr"
// header, ".img" raw little-endian float32, s3 fastest; dim = {4, s3, s2,
// s1, s0}); the loader returns [1, depth, height, width] of the first frame.
void SaveHDRImageTensor(torch::Tensor volume, const std::string& path);
torch::Tensor LoadHDRImageTensor(const std::string& path);

// ComputeMaxDensity (samples/hyper_train.cpp:662-707): node_max_density of the
// active nodes = max density on an n^3 lattice per node (UniformGlobalSamples),
// each point's density times its visibility in the scene's images
// (SceneBase::PointInAnyImage) when a scene is given.
class SceneBase;
void ComputeMaxDensity(HyperTreeBase& tree, torch::Tensor grid, torch::Tensor mlp, const FieldParams& fp, int n,
                       const SceneBase* scene = nullptr);
// ComputeStd (samples/hyper_train.cpp:709-759): per-node standard deviation
// (correction 1) of the density at rounds x n^3 uniform random points of each
// active node, reduced in FP64.  Returns [num_nodes] float64, 0 for inactive.
torch::Tensor ComputeStd(HyperTreeBase& tree, torch::Tensor grid, torch::Tensor mlp, const FieldParams& fp, int n,
                         int rounds, uint64_t seed);

// One training step of the plain method entirely on the device:
// intervals -> fused forward -> MSE -> fused adjoint -> Adam on grid and mlp.
// bf16 grids keep an f32 master copy (grid_master) and a bf16 copy (grid).
struct TrainState {
    torch::Tensor grid;          // [N, G, G, G, C] f32 or bf16
    torch::Tensor grid_master;   // f32 master when grid is bf16 (else undefined)
    torch::Tensor mlp;           // [P] f32
    torch::Tensor grad_grid, grad_mlp;
    torch::Tensor m_grid, v_grid, m_mlp, v_mlp;
    torch::Tensor loss;          // [1] f64 on device
    int64_t step = 0;
    float lr_grid = 1e-3f, lr_mlp = 1e-3f;
};
TrainState MakeTrainState(torch::Tensor grid, torch::Tensor mlp, bool bf16_grid, float lr_grid, float lr_mlp);
// Structure step of training (samples/hyper_train.cpp:1886-1911): optimize
// the tree structure; if the active set changed, move the grids to the new
// active set (TransferNodeGrids, through the f32 master for bf16 grids) and
// reset the Adam state of grids and decoder (ResetGeometryOptimizer,
// geometry/geometry.h:183-195).  Returns whether the structure changed.
bool RestructureTrainState(HyperTreeBase& tree, TrainState& st, const TreeOptimizerParams& p,
                           StructureDecision* out = nullptr);
torch::Tensor TrainStep(HyperTreeBase& tree, TrainState& st, const RayList& rays, torch::Tensor target,
                        const SamplingParams& sp, const FieldParams& fp);
// The same step over a given interval list (e.g. the learned proposal's).
torch::Tensor TrainStepIntervals(HyperTreeBase& tree, TrainState& st, const RayList& rays, torch::Tensor target,
                                 const SamplingParams& sp, const FieldParams& fp, const Intervals& iv);

// Projection over a given interval list (no autograd), and, with bins,
// the per-ray mass histogram [R, 32] over the root chord.
torch::Tensor ProjectIntervals(HyperTreeBase& tree, const RayList& rays, const Intervals& iv, torch::Tensor grid,
                               torch::Tensor mlp, const SamplingParams& sp, const FieldParams& fp);
// Adjoint over a given interval list: (dL/dgrid f32, dL/dmlp) for dL/dP.
std::pair<torch::Tensor, torch::Tensor> ProjectIntervalsBackward(HyperTreeBase& tree, const RayList& rays,
                                                                 const Intervals& iv, torch::Tensor grid,
                                                                 torch::Tensor mlp, const SamplingParams& sp,
                                                                 const FieldParams& fp, torch::Tensor dLdP);
std::pair<torch::Tensor, torch::Tensor> ProjectRaysBins(HyperTreeBase& tree, const RayList& rays, torch::Tensor grid,
                                                        torch::Tensor mlp, const SamplingParams& sp,
                                                        const FieldParams& fp);

// ---- learned sample proposal (no reference counterpart; the stratified
// sampler of HyperTree.cu:266-514 is its fallback) ----
struct ProposalParams {
    float mix = 0.25f;          // uniform mixing weight of the sampling pdf
    float kl_threshold = 0.5f;  // fallback when KL(coarse || pdf) exceeds it
    int coarse_lattice = 4;     // n^3 lattice per node for the coarse density
    float target_power = 0.5f;  // training target q_b ~ (binned quantity)^power
    // the binned quantity: 0 the density mass of bin b; 1 the change of the
    // bin-mean density across the bin, |m[b+1] - m[b-1]| / 2 (one-sided at
    // the ends), which is where a stratified estimate's variance sits
    int target_mode = 0;
    // 0: the proposal places 32 one-sample cells per ray.  E > 0: it allocates
    // over the stratified traversal's node intervals instead -- every node
    // interval keeps one sample and E more per ray follow the proposal's CDF
    // (atdf_proposal_allocate), fallback rays' E extras follow the uniform
    // CDF; the traversal's K only sets the list of node intervals.
    int alloc_extra = 0;
};
struct ProposalState {
    torch::Tensor params, grad, m, v;  // [P] f32
    torch::Tensor loss;                // [1] f64 (device)
    torch::Tensor coarse;              // [num_nodes] f32 coarse density
    int64_t step = 0;
    float lr = 1e-3f;
    ProposalParams pp;
    bool ready = false;                // trained at least once
};
ProposalState MakeProposalState(torch::Tensor params, float lr, const ProposalParams& pp);
// coarse[node] = mean density on the n^3 lattice of each active node (0 elsewhere)
void ProposalRefreshCoarse(HyperTreeBase& tree, ProposalState& ps, torch::Tensor grid, torch::Tensor mlp,
                           const FieldParams& fp);
// network inputs [R, 38] and root chords [R, 2]
std::pair<torch::Tensor, torch::Tensor> ProposalFeatures(HyperTreeBase& tree, const ProposalState& ps,
                                                         const RayList& rays);
// placement: {fallback [R], n_learned [R], node [32, R], tmin [32, R], tmax [32, R], kl [R]}
// for intervals integrated with the jitter of sp
std::vector<torch::Tensor> ProposalSample(HyperTreeBase& tree, const ProposalState& ps, const RayList& rays,
                                          const SamplingParams& sp);
// learned one-sample intervals for the learned rays, stratified traversal
// (fallback_sp) for the fallback rays, one list in ray order
Intervals CreateProposalIntervals(HyperTreeBase& tree, const ProposalState& ps, const RayList& rays,
                                  const SamplingParams& fallback_sp, torch::Tensor* ray_fallback = nullptr);
// loss (device [1] f64) and ps.grad for given features / target masses
torch::Tensor ProposalGrad(ProposalState& ps, torch::Tensor features, torch::Tensor target_bins);
// Online training on densities of the current reconstruction: target masses
// from the stratified binned forward over `rays`, then `steps` Adam steps.
// Refreshes the coarse densities first.  Returns the last loss.
// Proposal target of target_mode 1 from binned masses [R, B]: per bin the
// change of the bin-mean density across it, |m[b+1] - m[b-1]| / 2, and
// |m[1] - m[0]|, |m[B-1] - m[B-2]| at the ends (equal-length bins: the mass
// is the mean times a per-ray constant).
torch::Tensor BinVariation(const torch::Tensor& bins);
double ProposalTrain(HyperTreeBase& tree, ProposalState& ps, const RayList& rays, const SamplingParams& sp,
                     torch::Tensor grid, torch::Tensor mlp, const FieldParams& fp, int steps);
// Training step sampling with the proposal once it is ready (stratified before).
torch::Tensor TrainStepLearned(HyperTreeBase& tree, TrainState& st, ProposalState& ps, const RayList& rays,
                               torch::Tensor target, const SamplingParams& sp, const FieldParams& fp);

// ---- scene data (data/SceneBase.h, data/SceneBase.cpp, Settings.h:271-345) ----
// dataset.ini, section [DatasetParams] (Settings.h:271-345)
struct DatasetParams {
    std::string camera_model = "pinhole";  // "pinhole" (cone beam) or "orthographic" (parallel beam)
    std::string image_dir, mask_dir, volume_file;
    double projection_factor = 1, vis_volume_intensity_factor = 1, scene_scale = 1;
    double xray_min = 0, xray_max = 65535;
    bool log_space_input = false, use_log10_conversion = true;
    std::vector<float> roi_min{-1, -1, -1}, roi_max{1, 1, 1};
    void Load(const std::string& file);
};
// camera.ini, section [CameraBase] (SceneBase.h:84-127): detector size and
// K = fx fy cx cy s.
struct CameraBase {
    int w = 0, h = 0;
    std::vector<double> K{1, 1, 0, 0, 0};
    void Load(const std::string& file);
};
// One projection image (UnifiedImage, SceneBase.h:35-80).
struct UnifiedImage {
    int camera_id = 0;
    std::vector<double> rotation{0, 0, 0, 1};  // unit quaternion x y z w
    std::vector<double> translation{0, 0, 0};  // pose translation * scene_scale
    std::string image_file, mask_file;
    bool loaded = false;
};
// A random ray batch (PixelList of data/Dataloader.cpp:9-48) with its rays.
struct PixelList {
    torch::Tensor uv, image_id, camera_id, target, target_mask;
    RayList rays;
};
// The scene directory: dataset.ini, camera.ini, poses.txt ("qx qy qz qw tx ty
// tz" per image), images.txt, masks.txt, camera_indices.txt; projections as
// .npy / .npz (array "proj") files under image_dir.  Projections and masks
// live on the device as [num_images, h, w] f32 after LoadImagesCT.
class SceneBase {
   public:
    explicit SceneBase(const std::string& scene_dir);
    std::string scene_path, scene_name;
    DatasetParams dataset_params;
    std::vector<CameraBase> cameras;
    std::vector<UnifiedImage> frames;
    std::vector<int> train_indices, test_indices;
    std::vector<float> volume_translation{0, 0, 0};
    torch::Device device{torch::kCPU};
    torch::Tensor intrinsics, rotation, translation;  // f64 [cams, 5], [images, 4], [images, 3]
    torch::Tensor projections, masks;                 // f32 [images, h, w]; masks undefined without mask files
    bool align_corners = true;                        // uv_align_corners (SceneBase.h:24)

    // <scene>/<split_name>/train.txt and eval.txt (samples/hyper_train.cpp:93-97), sorted
    void LoadSplit(const std::string& split_name);
    // the pose and camera tensors on `device` (SceneBase::Finalize, SceneBase.cpp:98-140):
    // translation = pose translation - volume_translation
    void Finalize(const std::string& device, std::vector<float> volume_translation);
    // read, log-convert and mask the projections of `indices` (SceneBase.cpp:904-990)
    void LoadImagesCT(std::vector<int> indices);
    RayList GetRays(torch::Tensor uv, torch::Tensor image_id, torch::Tensor camera_id) const;
    std::pair<torch::Tensor, torch::Tensor> SampleProjection(torch::Tensor uv, torch::Tensor image_id) const;
    // RandomRaybasedSampleDataset::get_batch over `indices`, with the batch's rays
    PixelList RandomRayBatch(torch::Tensor indices, int64_t num_rays, uint64_t seed, bool interpolate) const;
    int NumImages() const { return (int)frames.size(); }
    // PointInAnyImage (data/SceneBase.cu:67-95): [..., 3] points -> [...] f32 0 / 1
    torch::Tensor PointInAnyImage(torch::Tensor points) const;
    int CameraModelId() const;
    static std::vector<int> ReadIndexFile(const std::string& file);  // TrainParameters.h:207-221
    atdf_camera_rig rig() const;
    atdf_projection_set projection_set() const;
};
// NPY / NPZ readers (the cnpy loaders of utils/utils.h:47-83): little- or
// big-endian float, integer and bool arrays in their own dtype; C order only.
torch::Tensor LoadNumpy(const std::string& path);
torch::Tensor LoadNumpyNpz(const std::string& path, const std::string& varname);

}  // namespace atdf_torch

// atdf_bindings.cpp -- Python bindings of the libtorch host layer.
#include <torch/extension.h>

#include "atdf_torch.h"

namespace py = pybind11;
using namespace atdf_torch;

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "libtorch host layer of the MI355X reconstruction hot path (calls libatdf.so)";

    py::class_<RayList>(m, "RayList")
        .def(py::init<>())
        .def(py::init([](torch::Tensor o, torch::Tensor d) { return RayList{o, d}; }))
        .def_readwrite("origin", &RayList::origin)
        .def_readwrite("direction", &RayList::direction)
        .def("size", &RayList::size);

    py::class_<SampleList>(m, "SampleList")
        .def(py::init<>())
        .def_readwrite("global_coordinate", &SampleList::global_coordinate)
        .def_readwrite("weight", &SampleList::weight)
        .def_readwrite("ray_t", &SampleList::ray_t)
        .def_readwrite("ray_index", &SampleList::ray_index)
        .def_readwrite("local_index_in_ray", &SampleList::local_index_in_ray)
        .def_readwrite("node_id", &SampleList::node_id)
        .def_readwrite("max_samples_per_ray", &SampleList::max_samples_per_ray)
        .def("size", &SampleList::size);

    py::class_<Intervals>(m, "Intervals")
        .def(py::init<>())
        .def_readwrite("ray_offset", &Intervals::ray_offset)
        .def_readwrite("node", &Intervals::node)
        .def_readwrite("tmin", &Intervals::tmin)
        .def_readwrite("tmax", &Intervals::tmax)
        .def_readwrite("count", &Intervals::count)
        .def_readwrite("first", &Intervals::first)
        .def_readwrite("ray", &Intervals::ray)
        .def_readwrite("ray_num_samples", &Intervals::ray_num_samples)
        .def_readwrite("num_intervals", &Intervals::num_intervals)
        .def_readwrite("num_samples", &Intervals::num_samples)
        .def_readwrite("max_samples_per_ray", &Intervals::max_samples_per_ray);

    py::class_<SamplingParams>(m, "SamplingParams")
        .def(py::init<>())
        .def_readwrite("max_samples_per_node", &SamplingParams::max_samples_per_node)
        .def_readwrite("max_intersections", &SamplingParams::max_intersections)
        .def_readwrite("jitter_mode", &SamplingParams::jitter_mode)
        .def_readwrite("jitter_seed", &SamplingParams::jitter_seed);

    py::class_<FieldParams>(m, "FieldParams")
        .def(py::init<>())
        .def_readwrite("C", &FieldParams::C)
        .def_readwrite("H", &FieldParams::H)
        .def_readwrite("G", &FieldParams::G)
        .def_readwrite("softplus_beta", &FieldParams::softplus_beta)
        .def_readwrite("rays_per_wave", &FieldParams::rays_per_wave)
        .def_readwrite("bf16_decoder", &FieldParams::bf16_decoder)
        .def_readwrite("node_forward", &FieldParams::node_forward)
        .def_readwrite("two_pass_adjoint", &FieldParams::two_pass_adjoint);

    py::class_<HyperTreeBase>(m, "HyperTreeBase")
        .def(py::init<int, int>(), py::arg("d"), py::arg("max_depth"))
        .def("settree", &HyperTreeBase::settree, py::arg("use_quad") = false, py::arg("y_min") = -1.f,
             py::arg("y_max") = 1.f)
        .def("ChildrenPerNode", &HyperTreeBase::ChildrenPerNode)
        .def("SetActive", &HyperTreeBase::SetActive)
        .def("UpdateActive", &HyperTreeBase::UpdateActive)
        .def("to", [](HyperTreeBase& t, const std::string& dev) { t.to(torch::Device(dev)); })
        .def("NumNodes", &HyperTreeBase::NumNodes)
        .def("NumActiveNodes", &HyperTreeBase::NumActiveNodes)
        .def("ActiveNodeTensor", &HyperTreeBase::ActiveNodeTensor)
        .def("CreateSamplesForRays", &HyperTreeBase::CreateSamplesForRays)
        .def("CreateIntervalsForRays", &HyperTreeBase::CreateIntervalsForRays)
        .def("CreateSamplesForRays_area", &HyperTreeBase::CreateSamplesForRays_area)
        .def("CreateIntervalsForRays_area", &HyperTreeBase::CreateIntervalsForRays_area)
        .def("ComputeLocalSamples", &HyperTreeBase::ComputeLocalSamples)
        .def("SetErrorForActiveNodes", &HyperTreeBase::SetErrorForActiveNodes)
        .def("UpdateCulling", &HyperTreeBase::UpdateCulling)
        .def("LeafTable", &HyperTreeBase::LeafTable)
        .def("RoiSplit", &HyperTreeBase::RoiSplit)
        .def("NodeIdForPositionGPU", &HyperTreeBase::NodeIdForPositionGPU)
        .def("UniformGlobalSamples", &HyperTreeBase::UniformGlobalSamples)
        .def("NodeNeighborSamples", &HyperTreeBase::NodeNeighborSamples, py::arg("size"), py::arg("epsilon"),
             py::arg("roi_min") = std::vector<float>{}, py::arg("roi_max") = std::vector<float>{},
             py::arg("use_roi") = false)
        .def("NodeNeighborSamplesPlane2D", &HyperTreeBase::NodeNeighborSamplesPlane2D, py::arg("size"),
             py::arg("epsilon"), py::arg("plane_index") = 0, py::arg("roi_min") = std::vector<float>{},
             py::arg("roi_max") = std::vector<float>{}, py::arg("use_roi") = false, py::arg("use_quad") = false)
        .def("UniformPhantomSamplesGPU", &HyperTreeBase::UniformPhantomSamplesGPU)
        .def("UniformPhantomSamplesGPUbySlice", &HyperTreeBase::UniformPhantomSamplesGPUbySlice)
        .def("InterpolateGridForInactiveNodes", &HyperTreeBase::InterpolateGridForInactiveNodes)
        .def("PackedTree", &HyperTreeBase::PackedTree)
        .def_readwrite("node_parent", &HyperTreeBase::node_parent)
        .def_readwrite("node_children", &HyperTreeBase::node_children)
        .def_readwrite("node_position_min", &HyperTreeBase::node_position_min)
        .def_readwrite("node_position_max", &HyperTreeBase::node_position_max)
        .def_readwrite("node_scale", &HyperTreeBase::node_scale)
        .def_readwrite("node_error", &HyperTreeBase::node_error)
        .def_readwrite("node_max_density", &HyperTreeBase::node_max_density)
        .def_readwrite("node_active", &HyperTreeBase::node_active)
        .def_readwrite("node_culled", &HyperTreeBase::node_culled)
        .def_readwrite("node_depth", &HyperTreeBase::node_depth)
        .def_readwrite("node_active_prefix_sum", &HyperTreeBase::node_active_prefix_sum)
        .def_readwrite("active_node_ids", &HyperTreeBase::active_node_ids)
        .def_readwrite("node_to_grid", &HyperTreeBase::node_to_grid);

    py::class_<RoiMaps>(m, "RoiMaps")
        .def_readwrite("in_roi_node_index", &RoiMaps::in_roi_node_index)
        .def_readwrite("out_roi_node_index", &RoiMaps::out_roi_node_index)
        .def_readwrite("prefix_inroi", &RoiMaps::prefix_inroi)
        .def_readwrite("prefix_outroi", &RoiMaps::prefix_outroi)
        .def_readwrite("in_roi_flag", &RoiMaps::in_roi_flag);
    py::class_<VMFields>(m, "VMFields")
        .def(py::init<>())
        .def_readwrite("plane_yz", &VMFields::plane_yz)
        .def_readwrite("plane_xz", &VMFields::plane_xz)
        .def_readwrite("plane_xy", &VMFields::plane_xy)
        .def_readwrite("line_x", &VMFields::line_x)
        .def_readwrite("line_y", &VMFields::line_y)
        .def_readwrite("line_z", &VMFields::line_z)
        .def_readwrite("grid_out", &VMFields::grid_out)
        .def_readwrite("tensor_op", &VMFields::tensor_op)
        .def_readwrite("roi_min", &VMFields::roi_min)
        .def_readwrite("roi_max", &VMFields::roi_max);
    py::class_<TreeOptimizerParams>(m, "TreeOptimizerParams")
        .def(py::init<>())
        .def_readwrite("max_active_nodes", &TreeOptimizerParams::max_active_nodes)
        .def_readwrite("error_merge_factor", &TreeOptimizerParams::error_merge_factor)
        .def_readwrite("error_split_factor", &TreeOptimizerParams::error_split_factor)
        .def_readwrite("use_saved_errors", &TreeOptimizerParams::use_saved_errors)
        .def_readwrite("tree_edge_roi_min", &TreeOptimizerParams::tree_edge_roi_min)
        .def_readwrite("tree_edge_roi_max", &TreeOptimizerParams::tree_edge_roi_max);
    py::class_<StructureDecision>(m, "StructureDecision")
        .def_readonly("choice", &StructureDecision::choice)
        .def_readonly("objective", &StructureDecision::objective)
        .def_readonly("active_weight", &StructureDecision::active_weight)
        .def_readonly("num_none", &StructureDecision::num_none)
        .def_readonly("num_split", &StructureDecision::num_split)
        .def_readonly("num_grp", &StructureDecision::num_grp)
        .def_readonly("feasible", &StructureDecision::feasible);
    m.def("SolveTreeStructure", &SolveTreeStructure);
    m.def("OptimizeTreeStructure", [](HyperTreeBase& tree, const TreeOptimizerParams& p) {
        StructureDecision d;
        bool changed = OptimizeTreeStructure(tree, p, &d);
        return std::make_pair(changed, d);
    });
    m.def("NodeCulling", &NodeCulling);
    m.def("UniformSampledVolume", &UniformSampledVolume);
    m.def("SampleVolumeBySlice", &SampleVolumeBySlice);
    py::class_<DatasetParams>(m, "DatasetParams")
        .def(py::init<>())
        .def("Load", &DatasetParams::Load)
        .def_readwrite("camera_model", &DatasetParams::camera_model)
        .def_readwrite("image_dir", &DatasetParams::image_dir)
        .def_readwrite("mask_dir", &DatasetParams::mask_dir)
        .def_readwrite("volume_file", &DatasetParams::volume_file)
        .def_readwrite("projection_factor", &DatasetParams::projection_factor)
        .def_readwrite("vis_volume_intensity_factor", &DatasetParams::vis_volume_intensity_factor)
        .def_readwrite("scene_scale", &DatasetParams::scene_scale)
        .def_readwrite("xray_min", &DatasetParams::xray_min)
        .def_readwrite("xray_max", &DatasetParams::xray_max)
        .def_readwrite("log_space_input", &DatasetParams::log_space_input)
        .def_readwrite("use_log10_conversion", &DatasetParams::use_log10_conversion)
        .def_readwrite("roi_min", &DatasetParams::roi_min)
        .def_readwrite("roi_max", &DatasetParams::roi_max);
    py::class_<CameraBase>(m, "CameraBase")
        .def(py::init<>())
        .def("Load", &CameraBase::Load)
        .def_readwrite("w", &CameraBase::w)
        .def_readwrite("h", &CameraBase::h)
        .def_readwrite("K", &CameraBase::K);
    py::class_<UnifiedImage>(m, "UnifiedImage")
        .def_readonly("camera_id", &UnifiedImage::camera_id)
        .def_readonly("rotation", &UnifiedImage::rotation)
        .def_readonly("translation", &UnifiedImage::translation)
        .def_readonly("image_file", &UnifiedImage::image_file)
        .def_readonly("mask_file", &UnifiedImage::mask_file)
        .def_readonly("loaded", &UnifiedImage::loaded);
    py::class_<PixelList>(m, "PixelList")
        .def_readonly("uv", &PixelList::uv)
        .def_readonly("image_id", &PixelList::image_id)
        .def_readonly("camera_id", &PixelList::camera_id)
        .def_readonly("target", &PixelList::target)
        .def_readonly("target_mask", &PixelList::target_mask)
        .def_readonly("rays", &PixelList::rays);
    py::class_<SceneBase>(m, "SceneBase")
        .def(py::init<const std::string&>())
        .def_readonly("scene_path", &SceneBase::scene_path)
        .def_readonly("scene_name", &SceneBase::scene_name)
        .def_readonly("dataset_params", &SceneBase::dataset_params)
        .def_readonly("cameras", &SceneBase::cameras)
        .def_readonly("frames", &SceneBase::frames)
        .def_readwrite("train_indices", &SceneBase::train_indices)
        .def_readwrite("test_indices", &SceneBase::test_indices)
        .def_readonly("volume_translation", &SceneBase::volume_translation)
        .def_readonly("intrinsics", &SceneBase::intrinsics)
        .def_readonly("rotation", &SceneBase::rotation)
        .def_readonly("translation", &SceneBase::translation)
        .def_readonly("projections", &SceneBase::projections)
        .def_readonly("masks", &SceneBase::masks)
        .def_readwrite("align_corners", &SceneBase::align_corners)
        .def("NumImages", &SceneBase::NumImages)
        .def("PointInAnyImage", &SceneBase::PointInAnyImage)
        .def("CameraModelId", &SceneBase::CameraModelId)
        .def("LoadSplit", &SceneBase::LoadSplit)
        .def("Finalize", &SceneBase::Finalize, py::arg("device"),
             py::arg("volume_translation") = std::vector<float>{0, 0, 0})
        .def("LoadImagesCT", &SceneBase::LoadImagesCT)
        .def("GetRays", &SceneBase::GetRays)
        .def("SampleProjection", &SceneBase::SampleProjection)
        .def("RandomRayBatch", &SceneBase::RandomRayBatch, py::arg("indices"), py::arg("num_rays"), py::arg("seed"),
             py::arg("interpolate") = true)
        .def_static("ReadIndexFile", &SceneBase::ReadIndexFile);
    m.def("LoadNumpy", &LoadNumpy);
    m.def("LoadNumpyNpz", &LoadNumpyNpz);
    m.def("SaveHDRImageTensor", &SaveHDRImageTensor);
    m.def("LoadHDRImageTensor", &LoadHDRImageTensor);
    m.def(
        "TransferNodeGrids",
        [](const HyperTreeBase& tree, torch::Tensor old_packed, torch::Tensor old_node_to_grid, torch::Tensor old_grid,
           std::optional<torch::Tensor> init) {
            return TransferNodeGrids(tree, old_packed, old_node_to_grid, old_grid, init ? *init : torch::Tensor());
        },
        py::arg("tree"), py::arg("old_packed"), py::arg("old_node_to_grid"), py::arg("old_grid"),
        py::arg("init") = py::none());
    m.def("RestructureTrainState", [](HyperTreeBase& tree, TrainState& st, const TreeOptimizerParams& p) {
        StructureDecision d;
        bool changed = RestructureTrainState(tree, st, p, &d);
        return std::make_pair(changed, d);
    });
    m.def("IndirectGridSample3D", &IndirectGridSample3D);
    m.def("IndirectGridSample2D", &IndirectGridSample2D);
    m.def("IndirectGridSample1D", &IndirectGridSample1D);
    m.def("GradientCorrection", &GradientCorrection);
    m.def("ComputeBoundWeight", &ComputeBoundWeight);
    m.def("SampleVolumeIndirectFeature", &SampleVolumeIndirectFeature);
    m.def(
        "TVLoss",
        [](torch::Tensor grid, std::optional<torch::Tensor> weight, bool channels_last) {
            return TVLoss(grid, weight ? *weight : torch::Tensor(), channels_last);
        },
        py::arg("grid"), py::arg("weight") = py::none(), py::arg("channels_last") = false);
    py::class_<FourierParams>(m, "FourierParams")
        .def(py::init<>())
        .def_readwrite("loss_fourier", &FourierParams::loss_fourier)
        .def_readwrite("scale", &FourierParams::scale)
        .def_readwrite("scale_y", &FourierParams::scale_y)
        .def_readwrite("shift", &FourierParams::shift)
        .def_readwrite("rate", &FourierParams::rate)
        .def_readwrite("rate_w", &FourierParams::rate_w);
    py::class_<FourierGrid>(m, "FourierGrid")
        .def(py::init<>())
        .def_readwrite("coords", &FourierGrid::coords)
        .def_readwrite("node_id", &FourierGrid::node_id)
        .def_readwrite("X", &FourierGrid::X)
        .def_readwrite("Y", &FourierGrid::Y)
        .def_readwrite("Z", &FourierGrid::Z);
    m.def("MakeFourierGrid", &MakeFourierGrid, py::arg("tree"), py::arg("grid_size"), py::arg("tree_depth"),
          py::arg("roi_min"), py::arg("roi_max"), py::arg("params"));
    m.def(
        "AdamStep",
        [](torch::Tensor param, torch::Tensor grad, torch::Tensor exp_avg, torch::Tensor exp_avg_sq, int64_t step,
           float lr, float beta1, float beta2, float eps, bool zero_grad) {
            for (const auto* t : {&param, &grad, &exp_avg, &exp_avg_sq})
                TORCH_CHECK(t->is_cuda() && t->scalar_type() == torch::kFloat32 && t->is_contiguous() &&
                                t->numel() == param.numel(),
                            "AdamStep: contiguous f32 device tensors of one size required");
            check_status(atdf_adam_step(param.data_ptr<float>(), grad.data_ptr<float>(), exp_avg.data_ptr<float>(),
                                        exp_avg_sq.data_ptr<float>(), param.numel(), lr, beta1, beta2, eps, step,
                                        zero_grad ? 1 : 0, current_stream()),
                         "atdf_adam_step");
        },
        py::arg("param"), py::arg("grad"), py::arg("exp_avg"), py::arg("exp_avg_sq"), py::arg("step"),
        py::arg("lr") = 1e-3f, py::arg("beta1") = 0.9f, py::arg("beta2") = 0.999f, py::arg("eps") = 1e-8f,
        py::arg("zero_grad") = true);
    m.def("DecodeFeatures", &DecodeFeatures);
    m.def("SampleVolumeIndirect", &SampleVolumeIndirect);
    m.def("ComputeImageVM", &ComputeImageVM);
    m.def("ProjectRaysVM", &ProjectRaysVM, py::keep_alive<0, 1>(), py::keep_alive<0, 2>());
    py::class_<VMTrainState>(m, "VMTrainState")
        .def(py::init<>())
        .def_readwrite("planes", &VMTrainState::planes)
        .def_readwrite("lines", &VMTrainState::lines)
        .def_readwrite("grid", &VMTrainState::grid)
        .def_readwrite("mlp", &VMTrainState::mlp)
        .def_readwrite("grads", &VMTrainState::grads)
        .def_readonly("loss", &VMTrainState::loss)
        .def_readonly("step", &VMTrainState::step)
        .def_readwrite("lr", &VMTrainState::lr)
        .def_readonly("roi_min_x", &VMTrainState::roi_min_x)
        .def_readonly("roi_max_x", &VMTrainState::roi_max_x);
    m.def("MakeVMTrainState", &MakeVMTrainState);
    m.def("VMTrainStateFields", &VMTrainStateFields);
    m.def("TrainStepVM", &TrainStepVM);
    m.def("FourierRingMeans", &FourierRingMeans);
    m.def("FourierLoss", &FourierLoss);
    py::class_<RegularizerParams>(m, "RegularizerParams")
        .def(py::init<>())
        .def_readwrite("loss_tv", &RegularizerParams::loss_tv)
        .def_readwrite("loss_l2", &RegularizerParams::loss_l2)
        .def_readwrite("test_line_x", &RegularizerParams::test_line_x)
        .def_readwrite("test_line_y", &RegularizerParams::test_line_y)
        .def_readwrite("test_line_z", &RegularizerParams::test_line_z)
        .def_readwrite("plane_vec_only", &RegularizerParams::plane_vec_only)
        .def_readwrite("use_loss_tv_3d", &RegularizerParams::use_loss_tv_3d)
        .def_readwrite("test_para", &RegularizerParams::test_para)
        .def_readwrite("tv_size", &RegularizerParams::tv_size)
        .def_readwrite("loss_edge", &RegularizerParams::loss_edge)
        .def_readwrite("use_ground_truth_volume", &RegularizerParams::use_ground_truth_volume)
        .def_readwrite("fourier", &RegularizerParams::fourier);
    m.def(
        "VolumeRegularizer",
        [](HyperTreeBase& tree, const RoiMaps& roi, const VMFields& f, const RegularizerParams& p,
           std::optional<SampleList> pairs, std::optional<FourierGrid> fourier, std::optional<torch::Tensor> mlp,
           std::optional<FieldParams> field) {
            return VolumeRegularizer(tree, roi, f, p, pairs ? &*pairs : nullptr, fourier ? &*fourier : nullptr,
                                     mlp ? *mlp : torch::Tensor(), field ? &*field : nullptr);
        },
        py::arg("tree"), py::arg("roi"), py::arg("fields"), py::arg("params"), py::arg("neighbor_pairs") = py::none(),
        py::arg("fourier") = py::none(), py::arg("mlp") = py::none(), py::arg("field") = py::none());
    m.def("EdgeLoss", &EdgeLoss);
    m.def("InitNodeGrid", &InitNodeGrid, py::arg("num_grids"), py::arg("C"), py::arg("G"), py::arg("init"),
          py::arg("scale"), py::arg("seed") = 0, py::arg("channels_last") = true, py::arg("device") = "cpu");
    m.def("InitTensorCoefficients", &InitTensorCoefficients, py::arg("shape"), py::arg("init"), py::arg("scale"),
          py::arg("seed") = 0, py::arg("device") = "cpu");
    m.def("IntegrateSamplesXRay", &IntegrateSamplesXRay);
    m.def("AccumulateSampleLossPerNode", &AccumulateSampleLossPerNode);
    m.def("ProjectRays", &ProjectRays);
    m.def(
        "ComputeMaxDensity",
        [](HyperTreeBase& tree, torch::Tensor grid, torch::Tensor mlp, const FieldParams& fp, int n,
           const SceneBase* scene) { ComputeMaxDensity(tree, grid, mlp, fp, n, scene); },
        py::arg("tree"), py::arg("grid"), py::arg("mlp"), py::arg("fp"), py::arg("n"), py::arg("scene") = nullptr);
    m.def("ComputeStd", &ComputeStd);

    py::class_<TrainState>(m, "TrainState")
        .def_readwrite("grid", &TrainState::grid)
        .def_readwrite("grid_master", &TrainState::grid_master)
        .def_readwrite("mlp", &TrainState::mlp)
        .def_readwrite("grad_grid", &TrainState::grad_grid)
        .def_readwrite("grad_mlp", &TrainState::grad_mlp)
        .def_readwrite("m_grid", &TrainState::m_grid)
        .def_readwrite("v_grid", &TrainState::v_grid)
        .def_readwrite("m_mlp", &TrainState::m_mlp)
        .def_readwrite("v_mlp", &TrainState::v_mlp)
        .def_readwrite("loss", &TrainState::loss)
        .def_readwrite("step", &TrainState::step)
        .def_readwrite("lr_grid", &TrainState::lr_grid)
        .def_readwrite("lr_mlp", &TrainState::lr_mlp);
    m.def("MakeTrainState", &MakeTrainState);
    m.def("TrainStep", &TrainStep);
    m.def("TrainStepIntervals", &TrainStepIntervals);
    m.def("ProjectIntervals", &ProjectIntervals);
    m.def("ProjectIntervalsBackward", &ProjectIntervalsBackward);
    m.def("ProjectRaysBins", &ProjectRaysBins);

    py::class_<ProposalParams>(m, "ProposalParams")
        .def(py::init<>())
        .def_readwrite("mix", &ProposalParams::mix)
        .def_readwrite("kl_threshold", &ProposalParams::kl_threshold)
        .def_readwrite("coarse_lattice", &ProposalParams::coarse_lattice)
        .def_readwrite("target_power", &ProposalParams::target_power)
        .def_readwrite("target_mode", &ProposalParams::target_mode)
        .def_readwrite("alloc_extra", &ProposalParams::alloc_extra);
    py::class_<ProposalState>(m, "ProposalState")
        .def_readwrite("params", &ProposalState::params)
        .def_readwrite("grad", &ProposalState::grad)
        .def_readwrite("loss", &ProposalState::loss)
        .def_readwrite("coarse", &ProposalState::coarse)
        .def_readwrite("step", &ProposalState::step)
        .def_readwrite("lr", &ProposalState::lr)
        .def_readwrite("pp", &ProposalState::pp)
        .def_readwrite("ready", &ProposalState::ready);
    m.def("MakeProposalState", &MakeProposalState);
    m.def("ProposalRefreshCoarse", &ProposalRefreshCoarse);
    m.def("ProposalFeatures", &ProposalFeatures);
    m.def("BinVariation", &BinVariation);
    m.def("ProposalSample", &ProposalSample);
    m.def("CreateProposalIntervals",
          [](HyperTreeBase& tree, const ProposalState& ps, const RayList& rays, const SamplingParams& sp) {
              torch::Tensor fb;
              Intervals iv = CreateProposalIntervals(tree, ps, rays, sp, &fb);
              return std::make_pair(iv, fb);
          });
    m.def("ProposalGrad", &ProposalGrad);
    m.def("ProposalTrain", &ProposalTrain);
    m.def("TrainStepLearned", &TrainStepLearned);
}

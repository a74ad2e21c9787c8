// The tensorial (VM) geometry through the fused projection kernels:
// GeometryTensorQO's in-ROI plane x line features and out-of-ROI explicit
// grids (geometry/geometry_tensor_qo.cpp:257-380) gathered inside
// atdf_project_fwd / atdf_project_bwd instead of being materialised per
// sample (ComputeImage -> SampleVolumeIndirect -> DecodeFeatures ->
// IntegrateSamplesXRay, geometry/geometry.cpp:828-914).
#include <array>

#include "atdf_torch.h"

namespace atdf_torch {

using torch::Tensor;

namespace {

void require_f32_cuda(const Tensor& t, const char* name)
{
    TORCH_CHECK(t.defined() && t.is_cuda() && t.scalar_type() == torch::kFloat32, name,
                " must be a float32 GPU tensor");
}

// Shapes of a VMFields in the reference layout: planes [N_in, C, Gp, Gp],
// lines [1, C, Gl, 1], grid_out [N_out, C, G, G, G] (or undefined).
void check_fields(const VMFields& f, const FieldParams& fp, int64_t n_in, int64_t n_out)
{
    TORCH_CHECK(f.tensor_op == "add", "the fused projection builds tensor_op \"add\" (\"cat\" takes ComputeImageVM)");
    for (const Tensor* p : {&f.plane_yz, &f.plane_xz, &f.plane_xy}) {
        require_f32_cuda(*p, "plane");
        TORCH_CHECK(p->dim() == 4 && p->size(0) == n_in && p->size(1) == fp.C && p->size(2) == p->size(3) &&
                        p->sizes() == f.plane_yz.sizes(),
                    "planes must all be [N_in, C, Gp, Gp] with N_in = ", n_in, " and C = ", fp.C);
    }
    for (const Tensor* l : {&f.line_x, &f.line_y, &f.line_z}) {
        require_f32_cuda(*l, "line");
        TORCH_CHECK(l->dim() == 4 && l->size(0) == 1 && l->size(1) == fp.C && l->size(3) == 1 &&
                        l->sizes() == f.line_x.sizes(),
                    "lines must all be [1, C, Gl, 1] (one shared line each) with C = ", fp.C);
    }
    if (n_out > 0) {
        require_f32_cuda(f.grid_out, "grid_out");
        TORCH_CHECK(f.grid_out.dim() == 5 && f.grid_out.size(0) >= n_out && f.grid_out.size(1) == fp.C &&
                        f.grid_out.size(2) == fp.G && f.grid_out.size(3) == fp.G && f.grid_out.size(4) == fp.G,
                    "grid_out must be [N_out, C, G, G, G] = [", n_out, ", ", fp.C, ", ", fp.G, ", ", fp.G, ", ",
                    fp.G, "]");
    }
    TORCH_CHECK(f.roi_min.size() == 3 && f.roi_max.size() == 3 && f.roi_max[0] > f.roi_min[0],
                "VMFields ROI bounds must have 3 entries with roi_max[0] > roi_min[0]");
}

// The fused layout (channels-last): planes [N_in, 3, Gp, Gp, C], lines
// [3, Gl, C], grids [N_out, G, G, G, C].
Tensor planes_cl(const Tensor& yz, const Tensor& xz, const Tensor& xy)
{
    return torch::stack({yz, xz, xy}, 1).permute({0, 1, 3, 4, 2}).contiguous();
}

Tensor lines_cl(const Tensor& x, const Tensor& y, const Tensor& z)
{
    return torch::stack({x.reshape({x.size(1), x.size(2)}), y.reshape({y.size(1), y.size(2)}),
                         z.reshape({z.size(1), z.size(2)})},
                        0)
        .permute({0, 2, 1})
        .contiguous();
}

Tensor grid_cl(const Tensor& g) { return g.permute({0, 2, 3, 4, 1}).contiguous(); }

// reference layouts back from the fused ones
std::array<Tensor, 3> planes_ref(const Tensor& p)
{
    auto t = p.permute({0, 1, 4, 2, 3});  // [N, 3, C, Gp, Gp]
    return {t.select(1, 0).contiguous(), t.select(1, 1).contiguous(), t.select(1, 2).contiguous()};
}

std::array<Tensor, 3> lines_ref(const Tensor& l)
{
    auto t = l.permute({0, 2, 1});  // [3, C, Gl]
    std::array<Tensor, 3> out;
    for (int i = 0; i < 3; ++i) out[i] = t.select(0, i).unsqueeze(0).unsqueeze(3).contiguous();
    return out;
}

Tensor grid_ref(const Tensor& g) { return g.permute({0, 4, 1, 2, 3}).contiguous(); }

// C-ABI field of a tensorial geometry in the fused layout; n2g / n2p hold
// the int32 node maps while the view is in use.
struct VMView {
    atdf_field f{};
    atdf_vm vm{};
    Tensor n2g, n2p;
};

void make_vm_view(const HyperTreeBase& tree, const RoiMaps& roi, const Tensor& planes, const Tensor& lines,
                  const Tensor& grid, const Tensor& mlp, const FieldParams& fp, float rx0, float rx1, VMView& v)
{
    TORCH_CHECK(roi.prefix_inroi.numel() == tree.NumNodes() && roi.prefix_outroi.numel() == tree.NumNodes(),
                "RoiMaps do not match the tree (run RoiSplit after the last structure change)");
    require_f32_cuda(mlp, "mlp");
    const int64_t np = (int64_t)fp.H * fp.C + fp.H + (int64_t)fp.H * fp.H + fp.H + fp.H + 1;
    TORCH_CHECK(mlp.numel() == np, "mlp has ", mlp.numel(), " parameters, expected ", np);
    v.n2p = roi.prefix_inroi.to(torch::kInt32).contiguous();
    v.n2g = roi.prefix_outroi.to(torch::kInt32).contiguous();
    v.vm.planes = planes.data_ptr<float>();
    v.vm.lines = lines.data_ptr<float>();
    v.vm.num_planes = planes.size(0);
    v.vm.node_to_plane = v.n2p.data_ptr<int32_t>();
    v.vm.Gp = (int32_t)planes.size(2);
    v.vm.Gl = (int32_t)lines.size(1);
    v.vm.roi_min_x = rx0;
    v.vm.roi_max_x = rx1;
    v.f.C = fp.C;
    v.f.H = fp.H;
    v.f.G = fp.G;
    v.f.grid_dtype = ATDF_GRID_F32;
    v.f.softplus_beta = fp.softplus_beta;
    v.f.rays_per_wave = fp.rays_per_wave;
    v.f.decoder_dtype = fp.bf16_decoder ? ATDF_DECODER_BF16 : ATDF_DECODER_F32;
    v.f.grid = grid.defined() && grid.numel() > 0 ? grid.data_ptr() : nullptr;
    v.f.num_grids = grid.defined() ? grid.size(0) : 0;
    v.f.node_to_grid = v.n2g.data_ptr<int32_t>();
    v.f.mlp = mlp.data_ptr<float>();
    v.f.vm = &v.vm;
}

// The adjoint's workspace, with the staging region of its two-pass form
// when asked for (atdf_project_bwd picks the two-pass kernels when it fits).
int64_t vm_workspace_bytes(int64_t R, const Intervals& iv, const atdf_field& f, bool two_pass)
{
    size_t n = atdf_project_bwd_workspace_bytes(R, iv.num_intervals, &f);
    if (two_pass) n += atdf_project_bwd_staging_bytes(iv.num_intervals, iv.num_samples, &f);
    return (int64_t)n;
}

struct ProjectVMFn : public torch::autograd::Function<ProjectVMFn> {
    static Tensor forward(torch::autograd::AutogradContext* ctx, Tensor yz, Tensor xz, Tensor xy, Tensor lx,
                          Tensor ly, Tensor lz, Tensor grid_out, Tensor mlp, Tensor origin, Tensor direction,
                          HyperTreeBase* tree, const RoiMaps* roi, SamplingParams sp, FieldParams fp, double rx0,
                          double rx1)
    {
        RayList rays{origin, direction};
        Intervals iv = tree->CreateIntervalsForRays(rays, sp);
        auto planes = planes_cl(yz, xz, xy);
        auto lines = lines_cl(lx, ly, lz);
        // no explicit grids: grid_out is an empty tensor (autograd inputs must be defined)
        auto grid = grid_out.numel() > 0 ? grid_cl(grid_out) : Tensor();
        VMView v;
        make_vm_view(*tree, *roi, planes, lines, grid, mlp, fp, (float)rx0, (float)rx1, v);
        atdf_sampling s = sp.view();
        atdf_intervals av = iv.view();
        auto P = torch::empty({rays.size()}, origin.options());
        check_status(atdf_project_fwd(tree->packed(), tree->NumNodes(), origin.data_ptr<float>(),
                                      direction.data_ptr<float>(), rays.size(), &s, iv.ray_offset.data_ptr<int64_t>(),
                                      &av, &v.f, P.data_ptr<float>(), current_stream()),
                     "atdf_project_fwd");
        ctx->save_for_backward({planes, lines, grid.defined() ? grid : torch::empty({0}, mlp.options()), mlp, origin,
                                direction, iv.ray_offset, iv.node, iv.tmin, iv.tmax, iv.count, iv.first, iv.ray});
        ctx->saved_data["tree"] = (int64_t)reinterpret_cast<intptr_t>(tree);
        ctx->saved_data["roi"] = (int64_t)reinterpret_cast<intptr_t>(roi);
        ctx->saved_data["K"] = (int64_t)sp.max_samples_per_node;
        ctx->saved_data["maxi"] = (int64_t)sp.max_intersections;
        ctx->saved_data["jm"] = (int64_t)sp.jitter_mode;
        ctx->saved_data["seed"] = (int64_t)sp.jitter_seed;
        ctx->saved_data["C"] = (int64_t)fp.C;
        ctx->saved_data["H"] = (int64_t)fp.H;
        ctx->saved_data["G"] = (int64_t)fp.G;
        ctx->saved_data["beta"] = (double)fp.softplus_beta;
        ctx->saved_data["rpw"] = (int64_t)fp.rays_per_wave;
        ctx->saved_data["bfdec"] = (bool)fp.bf16_decoder;
        ctx->saved_data["H_iv"] = iv.num_intervals;
        ctx->saved_data["S_iv"] = iv.num_samples;
        ctx->saved_data["twopass"] = (bool)fp.two_pass_adjoint;
        ctx->saved_data["rx0"] = rx0;
        ctx->saved_data["rx1"] = rx1;
        ctx->saved_data["has_grid"] = grid.defined();
        return P;
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_output)
    {
        auto sv = ctx->get_saved_variables();
        auto planes = sv[0], lines = sv[1], grid = sv[2], mlp = sv[3], origin = sv[4], direction = sv[5];
        const bool has_grid = ctx->saved_data["has_grid"].toBool();
        if (!has_grid) grid = Tensor();
        Intervals iv;
        iv.ray_offset = sv[6];
        iv.node = sv[7];
        iv.tmin = sv[8];
        iv.tmax = sv[9];
        iv.count = sv[10];
        iv.first = sv[11];
        iv.ray = sv[12];
        iv.num_intervals = ctx->saved_data["H_iv"].toInt();
        iv.num_samples = ctx->saved_data["S_iv"].toInt();
        auto* tree = reinterpret_cast<HyperTreeBase*>((intptr_t)ctx->saved_data["tree"].toInt());
        auto* roi = reinterpret_cast<const RoiMaps*>((intptr_t)ctx->saved_data["roi"].toInt());
        SamplingParams sp;
        sp.max_samples_per_node = (int)ctx->saved_data["K"].toInt();
        sp.max_intersections = (int)ctx->saved_data["maxi"].toInt();
        sp.jitter_mode = (int)ctx->saved_data["jm"].toInt();
        sp.jitter_seed = (uint64_t)ctx->saved_data["seed"].toInt();
        FieldParams fp;
        fp.C = (int)ctx->saved_data["C"].toInt();
        fp.H = (int)ctx->saved_data["H"].toInt();
        fp.G = (int)ctx->saved_data["G"].toInt();
        fp.softplus_beta = (float)ctx->saved_data["beta"].toDouble();
        fp.rays_per_wave = (int)ctx->saved_data["rpw"].toInt();
        fp.bf16_decoder = ctx->saved_data["bfdec"].toBool();
        fp.two_pass_adjoint = ctx->saved_data["twopass"].toBool();
        VMView v;
        make_vm_view(*tree, *roi, planes, lines, grid, mlp, fp, (float)ctx->saved_data["rx0"].toDouble(),
                     (float)ctx->saved_data["rx1"].toDouble(), v);
        auto gpl = torch::zeros_like(planes);
        auto gln = torch::zeros_like(lines);
        auto ggr = grid.defined() ? torch::zeros_like(grid) : Tensor();
        auto gml = torch::zeros_like(mlp);
        v.vm.grad_planes = gpl.data_ptr<float>();
        v.vm.grad_lines = gln.data_ptr<float>();
        atdf_sampling s = sp.view();
        atdf_intervals av = iv.view();
        auto dP = grad_output[0].to(torch::kFloat32).contiguous();
        const int64_t R = origin.size(0);
        auto ws = torch::empty({vm_workspace_bytes(R, iv, v.f, fp.two_pass_adjoint)},
                               torch::TensorOptions().dtype(torch::kUInt8).device(mlp.device()));
        check_status(atdf_project_bwd(tree->packed(), tree->NumNodes(), origin.data_ptr<float>(),
                                      direction.data_ptr<float>(), R, &s, iv.ray_offset.data_ptr<int64_t>(), &av,
                                      &v.f, dP.data_ptr<float>(), ggr.defined() ? ggr.data_ptr<float>() : nullptr,
                                      gml.data_ptr<float>(), ws.data_ptr(), ws.numel(), current_stream()),
                     "atdf_project_bwd");
        auto p3 = planes_ref(gpl);
        auto l3 = lines_ref(gln);
        return {p3[0], p3[1], p3[2], l3[0], l3[1], l3[2], ggr.defined() ? grid_ref(ggr) : Tensor(), gml,
                Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

}  // namespace

Tensor ProjectRaysVM(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& f, Tensor mlp, const SamplingParams& sp,
                     const FieldParams& fp, const RayList& rays)
{
    TORCH_CHECK(rays.origin.is_cuda() && rays.direction.is_cuda(), "rays must be on the GPU");
    const int64_t n_in = roi.in_roi_node_index.numel(), n_out = roi.out_roi_node_index.numel();
    check_fields(f, fp, n_in, n_out);
    return ProjectVMFn::apply(f.plane_yz, f.plane_xz, f.plane_xy, f.line_x, f.line_y, f.line_z,
                              n_out > 0 ? f.grid_out : torch::empty({0}, mlp.options()), mlp,
                              rays.origin.contiguous(),
                              rays.direction.contiguous(), &tree, &roi, sp, fp, (double)f.roi_min[0],
                              (double)f.roi_max[0]);
}

VMTrainState MakeVMTrainState(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& f, Tensor mlp, float lr,
                              const FieldParams& fp)
{
    const int64_t n_in = roi.in_roi_node_index.numel(), n_out = roi.out_roi_node_index.numel();
    check_fields(f, fp, n_in, n_out);
    (void)tree;
    VMTrainState st;
    torch::NoGradGuard ng;
    st.planes = planes_cl(f.plane_yz.detach(), f.plane_xz.detach(), f.plane_xy.detach());
    st.lines = lines_cl(f.line_x.detach(), f.line_y.detach(), f.line_z.detach());
    st.grid = n_out > 0 ? grid_cl(f.grid_out.detach()) : torch::empty({0, fp.G, fp.G, fp.G, fp.C}, mlp.options());
    st.mlp = mlp.detach().clone();
    for (Tensor* p : {&st.planes, &st.lines, &st.grid, &st.mlp}) {
        st.grads.push_back(torch::zeros_like(*p));
        st.m.push_back(torch::zeros_like(*p));
        st.v.push_back(torch::zeros_like(*p));
    }
    st.roi_min_x = f.roi_min[0];
    st.roi_max_x = f.roi_max[0];
    st.lr = lr;
    st.loss = torch::zeros({1}, mlp.options().dtype(torch::kFloat64));
    return st;
}

VMFields VMTrainStateFields(const VMTrainState& st, const VMFields& like)
{
    VMFields f = like;
    auto p3 = planes_ref(st.planes);
    auto l3 = lines_ref(st.lines);
    f.plane_yz = p3[0];
    f.plane_xz = p3[1];
    f.plane_xy = p3[2];
    f.line_x = l3[0];
    f.line_y = l3[1];
    f.line_z = l3[2];
    f.grid_out = st.grid.size(0) > 0 ? grid_ref(st.grid) : Tensor();
    return f;
}

Tensor TrainStepVM(HyperTreeBase& tree, const RoiMaps& roi, VMTrainState& st, const RayList& rays, Tensor target,
                   const SamplingParams& sp, const FieldParams& fp)
{
    TORCH_CHECK(rays.origin.is_cuda() && rays.direction.is_cuda(), "rays must be on the GPU");
    require_f32_cuda(target, "target");
    const int64_t R = rays.size();
    TORCH_CHECK(target.numel() == R, "target must hold one value per ray");
    auto stream = current_stream();
    Intervals iv = tree.CreateIntervalsForRays(rays, sp);
    VMView v;
    make_vm_view(tree, roi, st.planes, st.lines, st.grid.size(0) > 0 ? st.grid : Tensor(), st.mlp, fp, st.roi_min_x,
                 st.roi_max_x, v);
    atdf_sampling s = sp.view();
    atdf_intervals av = iv.view();
    const int64_t* ivo = iv.ray_offset.data_ptr<int64_t>();
    const float *O = rays.origin.data_ptr<float>(), *D = rays.direction.data_ptr<float>();
    auto P = torch::empty({R}, rays.origin.options());
    // the gradients are zero here: MakeVMTrainState allocates them zeroed and
    // each Adam pass below clears its gradient (zero_grad)
    v.vm.grad_planes = st.grads[0].data_ptr<float>();
    v.vm.grad_lines = st.grads[1].data_ptr<float>();
    auto ws = torch::empty({vm_workspace_bytes(R, iv, v.f, fp.two_pass_adjoint)},
                           torch::TensorOptions().dtype(torch::kUInt8).device(rays.origin.device()));
    if (fp.node_forward) {
        // one counting sort for both directions; node-major forward over it
        check_status(atdf_project_prepare(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &av, &v.f, ws.data_ptr(),
                                          ws.numel(), stream),
                     "atdf_project_prepare");
        check_status(atdf_project_fwd_prepared(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &av, &v.f,
                                               P.data_ptr<float>(), ws.data_ptr(), ws.numel(), stream),
                     "atdf_project_fwd_prepared");
    } else {
        check_status(atdf_project_fwd(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &av, &v.f,
                                      P.data_ptr<float>(), stream),
                     "atdf_project_fwd");
    }
    auto dP = torch::empty({R}, rays.origin.options());
    auto mse_ws = torch::empty({(int64_t)atdf_mse_workspace_bytes(R)},
                               torch::TensorOptions().dtype(torch::kUInt8).device(rays.origin.device()));
    check_status(atdf_mse_loss(P.data_ptr<float>(), target.data_ptr<float>(), R, 1.0f, dP.data_ptr<float>(),
                               st.loss.data_ptr<double>(), mse_ws.data_ptr(), mse_ws.numel(), stream),
                 "atdf_mse_loss");
    float* gg = st.grid.size(0) > 0 ? st.grads[2].data_ptr<float>() : nullptr;
    if (fp.node_forward)
        check_status(atdf_project_bwd_prepared(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &av, &v.f,
                                               dP.data_ptr<float>(), gg, st.grads[3].data_ptr<float>(), nullptr,
                                               ws.data_ptr(), ws.numel(), stream),
                     "atdf_project_bwd_prepared");
    else
        check_status(atdf_project_bwd(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &av, &v.f,
                                      dP.data_ptr<float>(), gg, st.grads[3].data_ptr<float>(), ws.data_ptr(),
                                      ws.numel(), stream),
                     "atdf_project_bwd");
    st.step += 1;
    Tensor* params[4] = {&st.planes, &st.lines, &st.grid, &st.mlp};
    for (int i = 0; i < 4; ++i) {
        if (params[i]->numel() == 0) continue;
        check_status(atdf_adam_step(params[i]->data_ptr<float>(), st.grads[i].data_ptr<float>(),
                                    st.m[i].data_ptr<float>(), st.v[i].data_ptr<float>(), params[i]->numel(), st.lr,
                                    0.9f, 0.999f, 1e-8f, st.step, 1, stream),
                     "atdf_adam_step");
    }
    return st.loss;
}

}  // namespace atdf_torch

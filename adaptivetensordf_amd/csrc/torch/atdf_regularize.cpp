// atdf_regularize.cpp -- TVLoss as an autograd Function over atdf_tv_loss,
// and the plane / line / grid regulariser of the tensorial (VM) geometry.
#include <ATen/CPUGeneratorImpl.h>

#include "atdf_torch.h"

namespace atdf_torch {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

namespace {
// Loss and gradient in one launch; the gradient is kept for backward.
class TVLossFn : public torch::autograd::Function<TVLossFn> {
   public:
    static Tensor forward(AutogradContext* ctx, Tensor x, Tensor weight, bool channels_last)
    {
        TORCH_CHECK(x.is_cuda() && x.scalar_type() == torch::kFloat32, "TVLoss: f32 device grid required");
        TORCH_CHECK(x.dim() >= 3 && x.dim() <= 5, "TVLoss: [B, C, n...] with 1 to 3 spatial dims, got ", x.dim(),
                    " dims");
        x = x.contiguous();
        atdf_tv_shape sh{};
        sh.B = x.size(0);
        sh.D = (int32_t)x.dim() - 2;
        if (channels_last) {
            sh.C = x.size(-1);
            sh.stride_c = 1;
            for (int d = 0; d < sh.D; ++d) {
                sh.n[d] = (int32_t)x.size(1 + d);
                sh.stride_s[d] = x.stride(1 + d);
            }
        } else {
            sh.C = x.size(1);
            sh.stride_c = x.stride(1);
            for (int d = 0; d < sh.D; ++d) {
                sh.n[d] = (int32_t)x.size(2 + d);
                sh.stride_s[d] = x.stride(2 + d);
            }
        }
        sh.stride_b = x.stride(0);
        const float* w = nullptr;
        if (weight.numel() > 0) {
            TORCH_CHECK(weight.numel() == sh.B, "TVLoss: one weight per batch entry");
            weight = weight.to(torch::kFloat32).contiguous();
            w = weight.data_ptr<float>();
        }
        auto loss = torch::empty({}, x.options().dtype(torch::kFloat64));
        auto grad = torch::zeros_like(x);
        auto ws = torch::empty({(int64_t)atdf_tv_workspace_bytes(x.numel())}, x.options().dtype(torch::kUInt8));
        check_status(atdf_tv_loss(x.data_ptr<float>(), &sh, w, 1.0, loss.data_ptr<double>(), grad.data_ptr<float>(),
                                  ws.data_ptr(), (size_t)ws.numel(), current_stream()),
                     "atdf_tv_loss");
        ctx->save_for_backward({grad});
        return loss.to(torch::kFloat32);
    }
    static tensor_list backward(AutogradContext* ctx, tensor_list grads)
    {
        auto g = ctx->get_saved_variables()[0];
        return {g * grads[0].to(g.dtype()), Tensor(), Tensor()};
    }
};
}  // namespace

Tensor InitNodeGrid(int64_t num_grids, int C, int G, const std::string& init, double scale, uint64_t seed,
                    bool channels_last, const std::string& device)
{
    TORCH_CHECK(num_grids >= 0 && C >= 1 && G >= 2, "InitNodeGrid: bad shape");
    std::vector<int64_t> shape = channels_last ? std::vector<int64_t>{num_grids, G, G, G, C}
                                               : std::vector<int64_t>{num_grids, C, G, G, G};
    auto opt = torch::TensorOptions().dtype(torch::kFloat32);
    auto gen = at::detail::createCPUGenerator(seed);
    Tensor g;
    if (init == "uniform" || init == "random")
        g = torch::empty(shape, opt).uniform_(-scale, scale, gen);
    else if (init == "normal")
        g = torch::empty(shape, opt).normal_(0.0, scale, gen);
    else if (init == "minus")
        g = -torch::ones(shape, opt);
    else if (init == "zero")
        g = torch::zeros(shape, opt);
    else
        TORCH_CHECK(false, "InitNodeGrid: unknown init '", init, "' (uniform, random, normal, minus, zero)");
    return g.to(torch::Device(device));
}

Tensor InitTensorCoefficients(std::vector<int64_t> shape, const std::string& init, double scale, uint64_t seed,
                              const std::string& device)
{
    auto opt = torch::TensorOptions().dtype(torch::kFloat32);
    auto gen = at::detail::createCPUGenerator(seed);
    Tensor t;
    if (init == "random")
        t = (torch::rand(shape, gen, opt) - 1.0f) * 2.0f * scale;
    else if (init == "uniform")
        t = torch::empty(shape, opt).uniform_(-scale, scale, gen);
    else
        TORCH_CHECK(false, "InitTensorCoefficients: unknown init '", init, "' (uniform, random)");
    return t.to(torch::Device(device));
}

Tensor TVLoss(Tensor grid, Tensor weight, bool channels_last)
{
    // an empty weight stands for "none": autograd Functions take defined tensors only
    if (!weight.defined()) weight = torch::empty({0}, grid.options());
    return TVLossFn::apply(grid, weight, channels_last);
}

Tensor EdgeLoss(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& f, const SampleList& pairs,
               const RegularizerParams& p)
{
    const int64_t n = pairs.node_id.numel();
    TORCH_CHECK(n % 2 == 0, "EdgeLoss: pairs come as rows 2k, 2k + 1");
    Tensor loss;
    if (n == 0) return loss;
    auto feat = SampleVolumeIndirectFeature(tree, roi, f, pairs.global_coordinate, pairs.node_id);
    auto t1 = feat.slice(0, 0, n, 2), t2 = feat.slice(0, 1, n, 2);
    auto edge = (t1 - t2).pow(2).mean(1).sqrt();
    // the pair's group: ROI membership of its first node
    auto first = pairs.node_id.slice(0, 0, n, 2).contiguous();
    auto in = ComputeBoundWeight(first, roi) >= 0;
    auto add = [&](const Tensor& t) { loss = loss.defined() ? loss + t : t; };
    const int64_t n_in = in.sum().item<int64_t>();
    if (n_in > 0) add(edge.masked_select(in).mean() * (p.loss_edge * (double)feat.size(1)));
    if (!p.use_ground_truth_volume && n_in < n / 2)
        add(edge.masked_select(in.logical_not()).mean() * (p.loss_edge * (double)feat.size(1)));
    return loss;
}

Tensor VolumeRegularizer(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& f, const RegularizerParams& p,
                         const SampleList* neighbor_pairs)
{
    Tensor loss;
    auto add = [&](const Tensor& t) { loss = loss.defined() ? loss + t : t; };
    if (p.loss_edge > 0 && neighbor_pairs) {
        auto e = EdgeLoss(tree, roi, f, *neighbor_pairs, p);
        if (e.defined()) add(e);
    }
    if (p.loss_tv > 0) {
        if (p.use_loss_tv_3d) {
            // features on a tv_size lattice of every in-ROI node, [N, x, y, z, C]
            // (geometry_tensor_qo.cpp:552-567, lattice 1059-1077)
            auto ids = roi.in_roi_node_index.to(torch::kLong).contiguous();
            const int64_t N = ids.size(0);
            TORCH_CHECK(p.tv_size.size() == 3, "VolumeRegularizer: tv_size needs 3 entries");
            if (N > 0) {
                auto opt = tree.node_position_min.options();
                auto lx = torch::linspace(-1, 1, p.tv_size[0], opt), ly = torch::linspace(-1, 1, p.tv_size[1], opt),
                     lz = torch::linspace(-1, 1, p.tv_size[2], opt);
                auto mesh = torch::meshgrid({lx, ly, lz}, "ij");
                auto cell = torch::stack({mesh[0], mesh[1], mesh[2]}, -1).unsqueeze(0);
                auto mn = tree.node_position_min.index_select(0, ids).view({N, 1, 1, 1, 3});
                auto sz = tree.node_position_max.index_select(0, ids).view({N, 1, 1, 1, 3}) - mn;
                auto pos = (cell + 1.f) / 2.f * sz + mn;
                auto nid = ids.view({N, 1, 1, 1}).expand({N, p.tv_size[0], p.tv_size[1], p.tv_size[2]});
                auto feat = SampleVolumeIndirectFeature(tree, roi, f, pos.reshape({-1, 3}), nid.reshape({-1}));
                feat = feat.reshape({N, p.tv_size[0], p.tv_size[1], p.tv_size[2], -1});
                add(TVLoss(feat, Tensor(), true) * p.loss_tv);
            }
        } else {
            // planes, lines, and the explicit grids (geometry_tensor_qo.cpp:578-607)
            add(TVLoss(f.plane_xz, Tensor(), false) * p.loss_tv);
            add(TVLoss(f.plane_yz, Tensor(), false) * p.loss_tv);
            add(TVLoss(f.plane_xy, Tensor(), false) * p.loss_tv);
            if (!p.plane_vec_only) {
                add(TVLoss(f.line_x.squeeze(-1), Tensor(), false) * (p.loss_tv * p.test_line_x));
                add(TVLoss(f.line_y.squeeze(-1), Tensor(), false) * (p.loss_tv * p.test_line_y));
                add(TVLoss(f.line_z.squeeze(-1), Tensor(), false) * (p.loss_tv * p.test_line_z));
            }
            if (p.test_para > 0 && f.grid_out.defined() && f.grid_out.size(0) > 0)
                add(TVLoss(f.grid_out, Tensor(), false) * p.loss_tv);
        }
    }
    if (p.loss_l2 > 0) {
        // geometry_tensor_qo.cpp:611-627; the explicit grids' term carries no
        // loss_l2 factor there, and here
        add(f.plane_xz.square().mean() * p.loss_l2);
        add(f.plane_yz.square().mean() * p.loss_l2);
        add(f.plane_xy.square().mean() * p.loss_l2);
        if (!p.plane_vec_only) {
            add(f.line_x.square().mean() * p.loss_l2);
            add(f.line_y.square().mean() * p.loss_l2);
            add(f.line_z.square().mean() * p.loss_l2);
        }
        if (f.grid_out.defined() && f.grid_out.numel() > 0) add(f.grid_out.square().mean());
    }
    return loss;
}

}  // namespace atdf_torch

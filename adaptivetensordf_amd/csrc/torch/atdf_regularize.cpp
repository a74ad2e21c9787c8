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

namespace {
atdf_field decoder_view(const Tensor& mlp, const FieldParams& fp)
{
    TORCH_CHECK(mlp.is_cuda() && mlp.scalar_type() == torch::kFloat32 && mlp.is_contiguous(),
                "decoder parameters must be a contiguous f32 device tensor");
    const int64_t np = (int64_t)fp.H * fp.C + fp.H + (int64_t)fp.H * fp.H + fp.H + fp.H + 1;
    TORCH_CHECK(mlp.numel() == np, "mlp has ", mlp.numel(), " parameters, expected ", np);
    atdf_field f{};
    f.C = fp.C;
    f.H = fp.H;
    f.G = fp.G;
    f.softplus_beta = fp.softplus_beta;
    f.decoder_dtype = fp.bf16_decoder ? ATDF_DECODER_BF16 : ATDF_DECODER_F32;
    f.mlp = mlp.data_ptr<float>();
    return f;
}

class DecodeFeaturesFn : public torch::autograd::Function<DecodeFeaturesFn> {
   public:
    static Tensor forward(AutogradContext* ctx, Tensor features, Tensor mlp, FieldParams fp)
    {
        TORCH_CHECK(features.is_cuda() && features.scalar_type() == torch::kFloat32,
                    "DecodeFeatures: f32 device features required");
        TORCH_CHECK(features.dim() == 2 && features.size(1) == fp.C, "DecodeFeatures: features must be [S, ", fp.C,
                    "]");
        features = features.contiguous();
        atdf_field f = decoder_view(mlp, fp);
        auto out = torch::empty({features.size(0), 1}, features.options());
        check_status(atdf_decode_features(features.data_ptr<float>(), features.size(0), &f, out.data_ptr<float>(),
                                          current_stream()),
                     "atdf_decode_features");
        ctx->save_for_backward({features, mlp});
        ctx->saved_data["C"] = (int64_t)fp.C;
        ctx->saved_data["H"] = (int64_t)fp.H;
        ctx->saved_data["beta"] = (double)fp.softplus_beta;
        ctx->saved_data["bfdec"] = (bool)fp.bf16_decoder;
        return out;
    }
    static tensor_list backward(AutogradContext* ctx, tensor_list grads)
    {
        auto sv = ctx->get_saved_variables();
        auto features = sv[0], mlp = sv[1];
        FieldParams fp;
        fp.C = (int)ctx->saved_data["C"].toInt();
        fp.H = (int)ctx->saved_data["H"].toInt();
        fp.softplus_beta = (float)ctx->saved_data["beta"].toDouble();
        fp.bf16_decoder = ctx->saved_data["bfdec"].toBool();
        atdf_field f = decoder_view(mlp, fp);
        const int64_t S = features.size(0);
        auto dd = grads[0].to(torch::kFloat32).reshape({S}).contiguous();
        auto gx = torch::empty_like(features);
        auto gm = torch::zeros_like(mlp);
        auto ws = torch::empty({(int64_t)atdf_decode_features_bwd_workspace_bytes(S, &f)},
                               features.options().dtype(torch::kUInt8));
        check_status(atdf_decode_features_bwd(features.data_ptr<float>(), S, &f, dd.data_ptr<float>(),
                                              gx.data_ptr<float>(), gm.data_ptr<float>(), ws.data_ptr(),
                                              (size_t)ws.numel(), current_stream()),
                     "atdf_decode_features_bwd");
        return {gx, gm, Tensor()};
    }
};
}  // namespace

Tensor DecodeFeatures(Tensor features, Tensor mlp, const FieldParams& fp)
{
    if (features.size(0) == 0) return torch::empty({0, 1}, features.options());
    return DecodeFeaturesFn::apply(features, mlp, fp);
}

Tensor SampleVolumeIndirect(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields, Tensor mlp,
                            const FieldParams& fp, Tensor global_coordinate, Tensor node_id)
{
    if (global_coordinate.numel() == 0) return global_coordinate;
    return DecodeFeatures(SampleVolumeIndirectFeature(tree, roi, fields, global_coordinate, node_id), mlp, fp);
}

Tensor ComputeImageVM(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& fields, Tensor mlp,
                      const FieldParams& fp, const SampleList& samples, int64_t num_rays)
{
    auto dens = SampleVolumeIndirect(tree, roi, fields, mlp, fp, samples.global_coordinate, samples.node_id);
    if (samples.global_coordinate.numel() == 0)
        return torch::zeros({1, num_rays}, samples.global_coordinate.options().dtype(torch::kFloat32));
    return IntegrateSamplesXRay(dens, samples.weight.unsqueeze(1), samples.ray_index, 1, num_rays);
}

FourierGrid MakeFourierGrid(HyperTreeBase& tree, int grid_size, int tree_depth, std::vector<float> roi_min,
                            std::vector<float> roi_max, const FourierParams& fp)
{
    TORCH_CHECK(roi_min.size() == 3 && roi_max.size() == 3, "MakeFourierGrid: 3-D ROI required");
    const int total_size = (grid_size - 1) * (1 << tree_depth);
    int x = (int)std::round(total_size * (roi_max[0] - roi_min[0]) / 2.);
    int y = (int)std::round(total_size * (roi_max[1] - roi_min[1] - 2 * fp.shift) / 2.);
    int z = (int)std::round(total_size * (roi_max[2] - roi_min[2]) / 2.);
    x = (int)(fp.scale * x);
    y = (int)(fp.scale * fp.scale_y * y);
    z = (int)(fp.scale * z);
    if (x % 2 == 0) x += 1;
    if (z % 2 == 0) z += 1;
    TORCH_CHECK(x >= 1 && y >= 1 && z >= 1, "MakeFourierGrid: empty lattice ", x, " x ", y, " x ", z);
    auto opt = torch::TensorOptions().dtype(torch::kFloat32).device(tree.device());
    auto cx = torch::linspace(roi_min[0], roi_max[0], x, opt);
    auto cy = torch::linspace(roi_min[1] + fp.shift, roi_max[1] - fp.shift, y, opt);
    auto cz = torch::linspace(roi_min[2], roi_max[2], z, opt);
    auto m = torch::meshgrid({cx, cy, cz}, "ij");
    FourierGrid g;
    g.X = x;
    g.Y = y;
    g.Z = z;
    g.coords = torch::stack({m[0], m[1], m[2]}, -1).reshape({-1, 3}).contiguous();
    g.node_id = tree.NodeIdForPositionGPU(g.coords).first;
    return g;
}

Tensor FourierRingMeans(Tensor img)
{
    TORCH_CHECK(img.is_cuda() && img.scalar_type() == torch::kFloat32 && img.dim() == 3,
                "FourierRingMeans: f32 device image [Y, X, Z] required");
    img = img.contiguous();
    const int64_t Y = img.size(0);
    const int X = (int)img.size(1), Z = (int)img.size(2);
    auto out = torch::empty({X / 2, Y}, img.options());
    auto ws = torch::empty({(int64_t)atdf_fourier_ring_workspace_bytes(X, Z)}, img.options().dtype(torch::kUInt8));
    check_status(atdf_fourier_ring_means(img.data_ptr<float>(), Y, X, Z, out.data_ptr<float>(), ws.data_ptr(),
                                         (size_t)ws.numel(), current_stream()),
                 "atdf_fourier_ring_means");
    return out;
}

Tensor FourierLoss(Tensor densities, const FourierParams& fp)
{
    TORCH_CHECK(densities.dim() == 3, "FourierLoss: densities must be [X, Y, Z]");
    auto d = densities.permute({1, 0, 2});
    const int64_t Y = d.size(0), X = d.size(1), Z = d.size(2);
    TORCH_CHECK(X == Z, "FourierLoss: the ring profile spans X values and is compared with both bands, so X (", X,
                ") must equal Z (", Z, ")");
    TORCH_CHECK(X % 2 == 1, "FourierLoss: X must be odd (the profile has 2 (X / 2) + 1 entries), got ", X);
    const int cx = (int)(X / 2), cy = (int)(Z / 2);
    const int rh = std::max((int)(fp.rate_w * cy), 1);
    const int rw = std::max((int)(fp.rate_w * cx), 1);
    const std::vector<int64_t> dims{-2, -1};
    auto F = torch::fft::fftshift(torch::fft::fftn(d, c10::nullopt, dims, "forward"), dims).abs();
    auto fx = F.slice(1, cy - rh + 1, cy + rh).mean(1);  // [Y, Z]
    auto fy = F.slice(2, cx - rw + 1, cx + rw).mean(2);  // [Y, X]
    const int64_t total = X / 2;
    auto ring = FourierRingMeans(F.detach().to(torch::kFloat32));  // [total, Y]
    auto circle = torch::cat({ring.flip({0}), torch::zeros({1, Y}, ring.options()), ring}, 0).t();
    auto xs = torch::arange(total, ring.options());
    auto yv = 1 - 1 / (torch::exp(xs - (double)(fp.rate * (float)total)) + 1);
    auto factor = torch::cat({yv.flip({0}), torch::zeros({1}, ring.options()), yv}, 0).unsqueeze(0).expand({Y, -1});
    circle = circle.to(fx.scalar_type());
    factor = factor.to(fx.scalar_type());
    auto mx = fx > circle;
    auto my = fy > circle;
    auto tx = ((torch::masked_select(fx, mx) - torch::masked_select(circle, mx)) * torch::masked_select(factor, mx))
                  .mean();
    auto ty = ((torch::masked_select(fy, my) - torch::masked_select(circle, my)) * torch::masked_select(factor, my))
                  .mean();
    return (tx + ty) * fp.loss_fourier;
}

Tensor VolumeRegularizer(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& f, const RegularizerParams& p,
                         const SampleList* neighbor_pairs, const FourierGrid* fourier, Tensor mlp,
                         const FieldParams* field)
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
    if (p.fourier.loss_fourier > 0 && fourier) {
        TORCH_CHECK(mlp.defined() && field, "VolumeRegularizer: the Fourier term needs the decoder and its FieldParams");
        auto dens = SampleVolumeIndirect(tree, roi, f, mlp, *field, fourier->coords, fourier->node_id)
                        .reshape({fourier->X, fourier->Y, fourier->Z});
        add(FourierLoss(dens, p.fourier));
    }
    return loss;
}

}  // namespace atdf_torch

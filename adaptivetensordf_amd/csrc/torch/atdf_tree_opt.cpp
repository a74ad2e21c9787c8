// atdf_tree_opt.cpp -- octree structure optimizer and node culling.
//
// The reference (structure/HyperTreeStructureOptimizer.cpp:70-392) builds a
// mixed-integer program over the active nodes and hands it to CBC:
//   variables   I_none, I_split, I_grp in {0,1} per active node
//   exactly one I_none + I_split + I_grp = 1
//   merge       I_none + I_split + sum_{s in siblings} I_grp[s] / n = 1 when
//               every sibling is active or culled (n active siblings), else
//               I_none + I_split = 1; the root never merges, leaves never split
//   budget      sum I_none + I_grp / n + I_split (NS - culled children) <= max
//   objective   min sum I_none e d + I_split e_split d_split + I_grp e_grp d_merge
// Subtracting the first constraint from the second gives I_grp = mean of the
// siblings' I_grp: a sibling group merges as a whole or not at all.  The
// program is therefore a multiple-choice knapsack over items (each
// merge-eligible sibling group, each other active node) with integer
// weights: a group either merges (weight 1) or its members choose none /
// split independently.  This file solves it exactly by dynamic programming
// over the budget, on the host like the reference.  Ties between optimal
// solutions are broken by a fixed option order (CBC's tie-breaking is not
// reproducible).
//
// One deliberate difference: the reference leaves the split objective of a
// parent outside the edge ROI uninitialised (HyperTreeStructureOptimizer.cpp:
// 244-324); here such a node may not split.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "atdf_torch.h"

namespace atdf_torch {

namespace {
// structure/HyperTreeStructureOptimizer.h:82-100 (no epsilon; mid = min + max)
bool edge_in_roi(const float* mn, const float* mx, const std::vector<float>& lo, const std::vector<float>& hi)
{
    for (int k = 0; k < 3; ++k) {
        const float mid = mn[k] + mx[k];
        if (!((mn[k] > lo[k] && mn[k] < hi[k]) || (mx[k] > lo[k] && mx[k] < hi[k]) || (mid > lo[k] && mid < hi[k])))
            return false;
    }
    return true;
}

struct Option {
    int64_t w;
    double c;
};
}  // namespace

StructureDecision SolveTreeStructure(const HyperTreeBase& tree, const TreeOptimizerParams& p)
{
    auto act_t = tree.node_active.to(torch::kCPU).contiguous();
    auto cul_t = tree.node_culled.to(torch::kCPU).contiguous();
    auto par_t = tree.node_parent.to(torch::kCPU).contiguous();
    auto ch_t = tree.node_children.to(torch::kCPU).contiguous();
    const int NS = (int)ch_t.size(1);  // 8, or 4 for the quad-tree representation
    auto err_t = tree.node_error.to(torch::kCPU).to(torch::kFloat32).contiguous();
    auto den_t = tree.node_max_density.to(torch::kCPU).to(torch::kFloat32).contiguous();
    auto dep_t = tree.node_depth.to(torch::kCPU).contiguous();
    auto mn_t = tree.node_position_min.to(torch::kCPU).contiguous();
    auto mx_t = tree.node_position_max.to(torch::kCPU).contiguous();
    const int32_t* act = act_t.data_ptr<int32_t>();
    const int32_t* cul = cul_t.data_ptr<int32_t>();
    const int32_t* par = par_t.data_ptr<int32_t>();
    const int32_t* ch = ch_t.data_ptr<int32_t>();
    const float* err = err_t.data_ptr<float>();
    const float* den = den_t.data_ptr<float>();
    const int32_t* dep = dep_t.data_ptr<int32_t>();
    const float* mn = mn_t.data_ptr<float>();
    const float* mx = mx_t.data_ptr<float>();
    const int64_t Nn = act_t.size(0);
    TORCH_CHECK(p.max_active_nodes >= 1, "max_active_nodes must be >= 1");

    // per-node costs and weights (the objective coefficients of the reference)
    std::vector<double> c_none(Nn, 0), c_split(Nn, 0), c_grp(Nn, 0);
    std::vector<int64_t> w_split(Nn, 0);
    std::vector<char> can_split(Nn, 0), eligible(Nn, 0);
    std::vector<int> n_active_sib(Nn, 0);
    for (int64_t i = 0; i < Nn; ++i) {
        if (act[i] != 1) continue;
        TORCH_CHECK(dep[i] > 0, "SolveTreeStructure: an active node at depth 0 (the root) is not supported");
        const float loss_none = err[i];
        TORCH_CHECK(std::isfinite(loss_none) && loss_none >= 0, "node_error of active node ", i, " must be finite, >= 0");
        TORCH_CHECK(den[i] >= 0, "node_max_density of active node ", i, " must be >= 0");
        c_none[i] = (double)loss_none * (double)den[i];
        // merge: siblings under the same parent
        const int32_t pa = par[i];
        if (pa >= 0) {
            int na = 0, nc = 0;
            float sum = 0.f, dmax = 0.f;
            int n_non_culled = 0;
            for (int k = 0; k < NS; ++k) {
                const int32_t s = ch[(int64_t)pa * NS + k];
                na += act[s];
                nc += cul[s];
                if (!cul[s] && act[s]) {
                    float e = err[s];
                    if (e < 0) e = 0;
                    sum += e;
                    dmax = std::max(den[s], dmax);
                    ++n_non_culled;
                }
            }
            n_active_sib[i] = na;
            eligible[i] = (na + nc == NS) ? 1 : 0;
            float loss_grp = (float)(p.error_merge_factor * (double)sum);
            loss_grp /= (float)n_non_culled;
            c_grp[i] = (double)loss_grp * (double)dmax;
        }
        // split
        if (ch[i * NS] >= 0 && edge_in_roi(mn + 3 * i, mx + 3 * i, p.tree_edge_roi_min, p.tree_edge_roi_max)) {
            int ncc = 0;
            for (int k = 0; k < NS; ++k) ncc += cul[ch[i * NS + k]];
            TORCH_CHECK(ncc < NS, "node ", i, " has all its children culled");
            float loss_split, dens_split = den[i];
            const int32_t c0 = ch[i * NS];
            if (p.use_saved_errors && err[c0] > 0) {
                dens_split = 1.f;
                loss_split = 0.f;
                for (int k = 0; k < NS; ++k) {
                    const int32_t c = ch[i * NS + k];
                    TORCH_CHECK(err[c] >= 0 && den[c] >= 0, "saved child error / density must be >= 0");
                    loss_split += err[c] * den[c];
                }
            } else {
                loss_split = (float)(p.error_split_factor * (double)loss_none);
            }
            c_split[i] = (double)loss_split * (double)dens_split;
            w_split[i] = NS - ncc;
            can_split[i] = 1;
        }
    }

    // items of the multiple-choice knapsack
    struct Item {
        std::vector<int64_t> nodes;   // members (one for a single node)
        std::vector<Option> options;  // none/split combinations..., then merge
        std::vector<uint32_t> masks;  // split mask of each combination option
        bool group = false;
    };
    std::vector<Item> items;
    std::vector<char> done(Nn, 0);
    for (int64_t i = 0; i < Nn; ++i) {
        if (act[i] != 1 || done[i]) continue;
        Item it;
        if (par[i] >= 0 && eligible[i]) {
            const int32_t pa = par[i];
            for (int k = 0; k < NS; ++k) {
                const int32_t s = ch[(int64_t)pa * NS + k];
                if (act[s]) {
                    it.nodes.push_back(s);
                    done[s] = 1;
                }
            }
            it.group = true;
        } else {
            it.nodes.push_back(i);
            done[i] = 1;
        }
        // none / split combinations, the cheapest per total weight
        const int m = (int)it.nodes.size();
        std::vector<std::pair<int64_t, std::pair<double, uint32_t>>> best;
        for (uint32_t mask = 0; mask < (1u << m); ++mask) {
            bool ok = true;
            int64_t w = 0;
            double c = 0;
            for (int q = 0; q < m; ++q) {
                const int64_t nd = it.nodes[q];
                if (mask & (1u << q)) {
                    if (!can_split[nd]) {
                        ok = false;
                        break;
                    }
                    w += w_split[nd];
                    c += c_split[nd];
                } else {
                    w += 1;
                    c += c_none[nd];
                }
            }
            if (!ok) continue;
            auto f = std::find_if(best.begin(), best.end(), [&](const auto& e) { return e.first == w; });
            if (f == best.end())
                best.push_back({w, {c, mask}});
            else if (c < f->second.first)
                f->second = {c, mask};
        }
        std::sort(best.begin(), best.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (auto& e : best) {
            it.options.push_back({e.first, e.second.first});
            it.masks.push_back(e.second.second);
        }
        if (it.group) {
            double cm = 0;
            for (int64_t nd : it.nodes) cm += c_grp[nd];
            it.options.push_back({1, cm});
            it.masks.push_back(0);
        }
        items.push_back(std::move(it));
    }

    // DP over the budget: dp[b] = min cost with total weight exactly b
    const int64_t B = p.max_active_nodes;
    const double INF = std::numeric_limits<double>::infinity();
    std::vector<double> dp(B + 1, INF), nd(B + 1);
    dp[0] = 0;
    std::vector<uint8_t> pick(items.size() * (size_t)(B + 1), 255);
    for (size_t t = 0; t < items.size(); ++t) {
        std::fill(nd.begin(), nd.end(), INF);
        uint8_t* pk = pick.data() + t * (size_t)(B + 1);
        const auto& opts = items[t].options;
        for (size_t o = 0; o < opts.size(); ++o) {
            const int64_t w = opts[o].w;
            const double c = opts[o].c;
            for (int64_t b = w; b <= B; ++b) {
                if (dp[b - w] == INF) continue;
                const double v = dp[b - w] + c;
                if (v < nd[b]) {
                    nd[b] = v;
                    pk[b] = (uint8_t)o;
                }
            }
        }
        dp.swap(nd);
    }
    StructureDecision dec;
    dec.choice.assign(Nn, -1);
    int64_t bbest = -1;
    for (int64_t b = 0; b <= B; ++b)
        if (dp[b] < INF && (bbest < 0 || dp[b] < dp[bbest])) bbest = b;
    dec.feasible = bbest >= 0;
    if (!dec.feasible) {
        for (int64_t i = 0; i < Nn; ++i)
            if (act[i] == 1) dec.choice[i] = 0;
        return dec;
    }
    dec.objective = dp[bbest];
    int64_t b = bbest;
    for (size_t t = items.size(); t-- > 0;) {
        const uint8_t o = pick[t * (size_t)(B + 1) + b];
        const Item& it = items[t];
        const bool merge = it.group && o == it.options.size() - 1;
        for (size_t q = 0; q < it.nodes.size(); ++q) {
            const int64_t ndi = it.nodes[q];
            dec.choice[ndi] = merge ? 2 : ((it.masks[o] >> q) & 1u ? 1 : 0);
        }
        b -= it.options[o].w;
    }
    dec.active_weight = bbest;
    for (int64_t i = 0; i < Nn; ++i) {
        if (dec.choice[i] == 0) dec.num_none++;
        if (dec.choice[i] == 1) dec.num_split++;
        if (dec.choice[i] == 2) dec.num_grp++;
    }
    return dec;
}

bool OptimizeTreeStructure(HyperTreeBase& tree, const TreeOptimizerParams& p, StructureDecision* out)
{
    StructureDecision dec = SolveTreeStructure(tree, p);
    if (out) *out = dec;
    if (!dec.feasible) return false;
    // ApplySplit (HyperTreeStructureOptimizer.cpp:333-392)
    auto dev = tree.node_active.device();
    auto act_t = tree.node_active.to(torch::kCPU).contiguous().clone();
    auto cul_t = tree.node_culled.to(torch::kCPU).contiguous();
    auto par_t = tree.node_parent.to(torch::kCPU).contiguous();
    auto ch_t = tree.node_children.to(torch::kCPU).contiguous();
    int32_t* act = act_t.data_ptr<int32_t>();
    const int32_t* cul = cul_t.data_ptr<int32_t>();
    const int32_t* par = par_t.data_ptr<int32_t>();
    const int32_t* ch = ch_t.data_ptr<int32_t>();
    const int64_t Nn = act_t.size(0);
    std::vector<int32_t> before(act, act + Nn);
    for (int64_t i = 0; i < Nn; ++i) {
        if (before[i] != 1) continue;
        const int c = dec.choice[i];
        if (c == 1) {
            act[i] = 0;
            const int64_t NS = ch_t.size(1);
            for (int64_t k = 0; k < NS; ++k) {
                const int32_t cc = ch[i * NS + k];
                if (cul[cc] == 0) act[cc] = 1;
            }
        } else if (c == 2) {
            TORCH_CHECK(par[i] >= 0 && cul[par[i]] == 0, "merge into a missing or culled parent");
            act[i] = 0;
            act[par[i]] = 1;
        }
    }
    const bool changed = dec.num_none != (dec.num_none + dec.num_split + dec.num_grp);
    tree.node_active = act_t.to(dev);
    tree.UpdateActive();
    return changed;
}

void NodeCulling(HyperTreeBase& tree, float culling_threshold, std::vector<float> roi_min, std::vector<float> roi_max,
                 float epsilon)
{
    // samples/hyper_train.cpp:761-830: cull active nodes whose max density is
    // below the threshold, except the nodes inside the reconstruction ROI
    auto dev = tree.node_active.device();
    auto act_t = tree.node_active.to(torch::kCPU).contiguous().clone();
    auto cul_t = tree.node_culled.to(torch::kCPU).contiguous().clone();
    auto den_t = tree.node_max_density.to(torch::kCPU).to(torch::kFloat32).contiguous();
    auto mn_t = tree.node_position_min.to(torch::kCPU).contiguous();
    auto mx_t = tree.node_position_max.to(torch::kCPU).contiguous();
    int32_t* act = act_t.data_ptr<int32_t>();
    int32_t* cul = cul_t.data_ptr<int32_t>();
    const float* den = den_t.data_ptr<float>();
    const float* mn = mn_t.data_ptr<float>();
    const float* mx = mx_t.data_ptr<float>();
    const int64_t Nn = act_t.size(0);
    int64_t n_new = 0;
    for (int64_t i = 0; i < Nn; ++i) {
        if (act[i] != 1 || !(den[i] < culling_threshold)) continue;
        bool in_roi = true;
        for (int k = 0; k < 3; ++k) {
            const float lo = k == 1 ? roi_min[k] - epsilon : roi_min[k];
            const float hi = k == 1 ? roi_max[k] + epsilon : roi_max[k];
            const float mid = mn[3 * i + k] + mx[3 * i + k];
            if (!((mn[3 * i + k] > lo && mn[3 * i + k] < hi) || (mx[3 * i + k] > lo && mx[3 * i + k] < hi) ||
                  (mid > lo && mid < hi)))
                in_roi = false;
        }
        if (in_roi) continue;
        cul[i] = 1;
        act[i] = 0;
        ++n_new;
    }
    if (n_new == 0) return;
    tree.node_culled = cul_t.to(dev);
    tree.node_active = act_t.to(dev);
    tree.UpdateActive();
    tree.UpdateCulling();
}

std::pair<torch::Tensor, torch::Tensor> HyperTreeBase::NodeIdForPositionGPU(torch::Tensor global_samples) const
{
    TORCH_CHECK(global_samples.is_cuda() && node_position_min.is_cuda(), "NodeIdForPositionGPU: GPU tensors required");
    TORCH_CHECK(global_samples.size(-1) == 3, "NodeIdForPositionGPU: positions must end in 3");
    TORCH_CHECK(NumActiveNodes() > 0, "NodeIdForPositionGPU: no active node");
    auto pos = global_samples.reshape({-1, 3}).to(torch::kFloat32).contiguous();
    const int64_t n = pos.size(0);
    auto node_id = torch::zeros({n}, pos.options().dtype(torch::kLong));
    auto mask = torch::zeros({n}, pos.options());
    check_status(atdf_node_id_for_position(packed(), NumNodes(), active_node_ids.data_ptr<int64_t>(),
                                           pos.data_ptr<float>(), n, node_id.data_ptr<int64_t>(),
                                           mask.data_ptr<float>(), current_stream()),
                 "atdf_node_id_for_position");
    return {node_id, mask};
}

SampleList HyperTreeBase::NodeNeighborSamples(std::vector<int> size, double epsilon, std::vector<float> roi_min,
                                              std::vector<float> roi_max, bool use_roi) const
{
    return neighbor_samples(size, epsilon, roi_min, roi_max, use_roi, 0);
}

SampleList HyperTreeBase::NodeNeighborSamplesPlane2D(std::vector<int> size, double epsilon, int plane_index,
                                                     std::vector<float> roi_min, std::vector<float> roi_max,
                                                     bool use_roi, bool use_quad) const
{
    // use_quad: NodeNeighborSamples2D on plane plane_index (the quad-tree
    // representation's pairs); otherwise NodeNeighborSamplesQuad, which, like
    // the reference's kernel, ignores plane_index
    TORCH_CHECK(!use_quad || (plane_index >= 0 && plane_index <= 2), "NodeNeighborSamplesPlane2D: plane_index ",
                plane_index, " outside [0, 2]");
    return neighbor_samples(size, epsilon, roi_min, roi_max, use_roi, use_quad ? 2 : 1, plane_index);
}

SampleList HyperTreeBase::neighbor_samples(std::vector<int> size, double epsilon, std::vector<float> roi_min,
                                           std::vector<float> roi_max, bool use_roi, int variant, int plane) const
{
    TORCH_CHECK(node_position_min.is_cuda(), "NodeNeighborSamples: the tree must be on the GPU");
    TORCH_CHECK(size.size() == 3, "NodeNeighborSamples: size needs 3 entries");
    TORCH_CHECK(!use_roi || (roi_min.size() == 3 && roi_max.size() == 3), "NodeNeighborSamples: ROI needs 3 entries");
    atdf_neighbor_spec spec{};
    for (int k = 0; k < 3; ++k) {
        spec.size[k] = size[k];
        spec.roi_min[k] = use_roi ? roi_min[k] : 0.f;
        spec.roi_max[k] = use_roi ? roi_max[k] : 0.f;
    }
    spec.epsilon = (float)epsilon;
    spec.use_roi = use_roi ? 1 : 0;
    spec.variant = variant;
    spec.plane = plane;
    auto dev = node_position_min.device();
    auto ids = active_node_ids.contiguous();
    const int64_t NA = ids.size(0);
    const int64_t nc = atdf_neighbor_candidates(NA, &spec);
    auto flags = torch::empty({std::max<int64_t>(nc, 1)}, torch::TensorOptions().dtype(torch::kInt32).device(dev));
    check_status(atdf_neighbor_flags(packed(), NumNodes(), ids.data_ptr<int64_t>(), NA, &spec, flags.data_ptr<int32_t>(),
                                     current_stream()),
                 "atdf_neighbor_flags");
    auto off = torch::empty({nc + 1}, torch::TensorOptions().dtype(torch::kLong).device(dev));
    auto ws = torch::empty({(int64_t)atdf_scan_workspace_bytes(nc)}, torch::TensorOptions().dtype(torch::kUInt8).device(dev));
    check_status(atdf_exclusive_scan_i32(flags.data_ptr<int32_t>(), nc, off.data_ptr<int64_t>(), ws.data_ptr(),
                                         ws.numel(), current_stream()),
                 "atdf_exclusive_scan_i32");
    const int64_t P = off[nc].item<int64_t>();
    SampleList out;
    out.global_coordinate = torch::empty({2 * P, 3}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    out.node_id = torch::empty({2 * P}, torch::TensorOptions().dtype(torch::kLong).device(dev));
    if (P > 0)
        check_status(atdf_neighbor_write(packed(), NumNodes(), ids.data_ptr<int64_t>(), NA, &spec,
                                         off.data_ptr<int64_t>(), out.global_coordinate.data_ptr<float>(),
                                         out.node_id.data_ptr<int64_t>(), current_stream()),
                     "atdf_neighbor_write");
    out.weight = torch::ones({2 * P}, out.global_coordinate.options());
    out.ray_t = torch::zeros({2 * P}, out.global_coordinate.options());
    out.ray_index = torch::arange(2 * P, out.node_id.options());
    out.local_index_in_ray = torch::arange(2 * P, out.node_id.options()).remainder(2);
    return out;
}

torch::Tensor HyperTreeBase::UniformGlobalSamples(torch::Tensor node_id, int n) const
{
    TORCH_CHECK(node_position_min.is_cuda(), "UniformGlobalSamples: the tree must be on the GPU");
    TORCH_CHECK(node_id.dim() == 1, "UniformGlobalSamples: a 1-D node id list is required");
    auto ids = node_id.to(node_position_min.device()).to(torch::kLong).contiguous();
    const int64_t N = ids.size(0);
    auto out = torch::empty({N, n, n, n, 3}, node_position_min.options());
    if (N == 0) return out;
    auto mm = torch::aminmax(ids);
    TORCH_CHECK(std::get<0>(mm).item<int64_t>() >= 0 && std::get<1>(mm).item<int64_t>() < NumNodes(),
                "UniformGlobalSamples: node id out of range");
    check_status(atdf_node_lattice_points(packed(), NumNodes(), ids.data_ptr<int64_t>(), N, n, out.data_ptr<float>(),
                                          current_stream()),
                 "atdf_node_lattice_points");
    return out;
}

torch::Tensor HyperTreeBase::InterpolateGridForInactiveNodes(torch::Tensor active_grid) const
{
    TORCH_CHECK(active_grid.is_cuda() && active_grid.scalar_type() == torch::kFloat32 && active_grid.dim() == 5,
                "InterpolateGridForInactiveNodes: float32 GPU grid [num_nodes, C, G, G, G] required");
    TORCH_CHECK(active_grid.size(0) == NumNodes(), "InterpolateGridForInactiveNodes: one grid per node required");
    const int64_t G = active_grid.size(2);
    TORCH_CHECK(active_grid.size(3) == G && active_grid.size(4) == G && G >= 2,
                "InterpolateGridForInactiveNodes: cubic grids of at least 2^3 required");
    auto src = active_grid.contiguous();
    auto out = src.clone();
    auto inactive = torch::nonzero(node_active == 0).reshape({-1}).contiguous();
    check_status(atdf_grid_transfer(packed(), NumNodes(), nullptr, src.data_ptr<float>(), NumNodes(),
                                    inactive.data_ptr<int64_t>(), inactive.numel(), nullptr, out.data_ptr<float>(),
                                    NumNodes(), (int32_t)src.size(1), (int32_t)G, 1, current_stream()),
                 "atdf_grid_transfer");
    return out;
}

torch::Tensor TransferNodeGrids(const HyperTreeBase& tree, torch::Tensor old_packed, torch::Tensor old_node_to_grid,
                                torch::Tensor old_grid, torch::Tensor init)
{
    TORCH_CHECK(old_grid.is_cuda() && old_grid.scalar_type() == torch::kFloat32 && old_grid.dim() == 5 &&
                    old_grid.is_contiguous(),
                "TransferNodeGrids: contiguous float32 GPU grid [N, G, G, G, C] required");
    TORCH_CHECK(old_packed.defined() && old_packed.is_cuda() &&
                    old_packed.numel() == (int64_t)atdf_tree_packed_bytes(tree.NumNodes()),
                "TransferNodeGrids: old_packed must be a PackedTree() of this tree");
    TORCH_CHECK(old_node_to_grid.scalar_type() == torch::kInt32 && old_node_to_grid.numel() == tree.NumNodes() &&
                    old_node_to_grid.is_cuda(),
                "TransferNodeGrids: old_node_to_grid must be the tree's int32 node_to_grid before the change");
    const int64_t G = old_grid.size(1), C = old_grid.size(4);
    TORCH_CHECK(old_grid.size(2) == G && old_grid.size(3) == G && G >= 2, "TransferNodeGrids: cubic grids required");
    const int64_t na = tree.NumActiveNodes();
    torch::Tensor out;
    if (init.defined()) {
        TORCH_CHECK(init.sizes() == torch::IntArrayRef({na, G, G, G, C}) && init.scalar_type() == torch::kFloat32,
                    "TransferNodeGrids: init must be float32 [N_active, G, G, G, C]");
        out = init.to(old_grid.device()).contiguous().clone();
    } else {
        out = torch::zeros({na, G, G, G, C}, old_grid.options());
    }
    auto ids = tree.active_node_ids.contiguous();
    auto old_slot = old_node_to_grid.contiguous();
    check_status(atdf_grid_transfer(old_packed.data_ptr(), tree.NumNodes(), old_slot.data_ptr<int32_t>(),
                                    old_grid.data_ptr<float>(), old_grid.size(0), ids.data_ptr<int64_t>(), na,
                                    tree.node_to_grid.data_ptr<int32_t>(), out.data_ptr<float>(), na, (int32_t)C,
                                    (int32_t)G, 0, current_stream()),
                 "atdf_grid_transfer");
    return out;
}

bool RestructureTrainState(HyperTreeBase& tree, TrainState& st, const TreeOptimizerParams& p, StructureDecision* out)
{
    auto old_packed = tree.PackedTree();
    auto old_slot = tree.node_to_grid;
    const bool changed = OptimizeTreeStructure(tree, p, out);
    if (!changed) return false;
    const bool bf16 = st.grid_master.defined();
    auto moved = TransferNodeGrids(tree, old_packed, old_slot, bf16 ? st.grid_master : st.grid);
    if (bf16) {
        st.grid_master = moved;
        st.grid = torch::empty(moved.sizes(), moved.options().dtype(torch::kBFloat16));
        check_status(atdf_f32_to_bf16(moved.data_ptr<float>(), reinterpret_cast<uint16_t*>(st.grid.data_ptr()),
                                      moved.numel(), current_stream()),
                     "atdf_f32_to_bf16");
    } else {
        st.grid = moved;
    }
    st.grad_grid = torch::zeros(moved.sizes(), moved.options());
    st.m_grid = torch::zeros(moved.sizes(), moved.options());
    st.v_grid = torch::zeros(moved.sizes(), moved.options());
    st.m_mlp.zero_();
    st.v_mlp.zero_();
    st.grad_mlp.zero_();
    st.step = 0;
    return true;
}

}  // namespace atdf_torch

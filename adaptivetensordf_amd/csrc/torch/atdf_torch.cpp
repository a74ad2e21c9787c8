// atdf_torch.cpp -- libtorch host layer over libatdf.so (see atdf_torch.h).
#include "atdf_torch.h"

#include <c10/hip/HIPStream.h>

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace atdf_torch {

using torch::Tensor;

atdf_stream_t current_stream()
{
    return reinterpret_cast<atdf_stream_t>(c10::hip::getCurrentHIPStream().stream());
}

void check_status(atdf_status s, const char* what)
{
    TORCH_CHECK(s == ATDF_OK, what, " failed (status ", s, "): ", atdf_last_error());
}

static void require_device_f32(const Tensor& t, const char* name)
{
    TORCH_CHECK(t.defined(), name, " is undefined");
    TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor (the reference path is GPU-only)");
    TORCH_CHECK(t.scalar_type() == torch::kFloat32, name, " must be float32");
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

atdf_intervals Intervals::view() const
{
    atdf_intervals v{};
    if (num_intervals > 0) {
        v.node = node.data_ptr<int32_t>();
        v.tmin = tmin.data_ptr<float>();
        v.tmax = tmax.data_ptr<float>();
        v.count = count.data_ptr<int32_t>();
        v.first = first.data_ptr<int32_t>();
        v.ray = ray.data_ptr<int32_t>();
    }
    v.total = num_intervals;
    return v;
}

atdf_sampling SamplingParams::view(const float* jitter) const
{
    atdf_sampling s{};
    s.max_samples_per_node = max_samples_per_node;
    s.max_intersections = max_intersections;
    s.jitter_mode = jitter_mode;
    s.jitter_seed = jitter_seed;
    s.jitter = jitter;
    return s;
}

// ---------------------------------------------------------------------------
// HyperTreeBase
// ---------------------------------------------------------------------------
HyperTreeBase::HyperTreeBase(int d_, int max_depth_) : d(d_), max_depth(max_depth_)
{
    TORCH_CHECK(d == 3, "HyperTreeBase: only D = 3 (octree) is built");
    TORCH_CHECK(max_depth >= 0 && max_depth <= 8, "HyperTreeBase: max_depth must be in [0, 8]");
}

void HyperTreeBase::settree(bool use_quad, float y_min, float y_max)
{
    TORCH_CHECK(!use_quad || y_min < y_max, "settree: the quad-tree y range needs y_min < y_max");
    const int NS = use_quad ? 4 : 8;
    int64_t num_nodes = 0, p = 1;
    for (int i = 0; i <= max_depth; ++i) {
        num_nodes += p;
        p *= NS;
    }
    node_parent = torch::empty({num_nodes}, torch::kInt32);
    node_children = -torch::ones({num_nodes, NS}, torch::kInt32);
    node_position_min = torch::empty({num_nodes, 3}, torch::kFloat32);
    node_position_max = torch::empty({num_nodes, 3}, torch::kFloat32);
    node_scale = torch::empty({num_nodes}, torch::kFloat32);
    node_error = -torch::ones({num_nodes}, torch::kFloat32);
    node_max_density = -torch::ones({num_nodes}, torch::kFloat32);
    node_active = torch::zeros({num_nodes}, torch::kInt32);
    node_culled = torch::zeros({num_nodes}, torch::kInt32);
    node_depth = torch::zeros({num_nodes}, torch::kInt32);

    float* pmin = node_position_min.data_ptr<float>();
    float* pmax = node_position_max.data_ptr<float>();
    int32_t* ch = node_children.data_ptr<int32_t>();
    int32_t* par = node_parent.data_ptr<int32_t>();
    int32_t* dep = node_depth.data_ptr<int32_t>();
    float* scl = node_scale.data_ptr<float>();
    auto set_node = [&](int64_t id, int64_t parent, const float* mn, const float* mx, int depth) {
        float scale = max_depth == 0 ? 0.f : ((float)depth / (float)max_depth) * 2 - 1;
        for (int k = 0; k < 3; ++k) {
            pmin[id * 3 + k] = mn[k];
            pmax[id * 3 + k] = mx[k];
        }
        dep[id] = depth;
        par[id] = (int32_t)parent;
        scl[id] = scale;
    };
    // the quad-tree representation (HyperTree.cpp:391-406, 461-464): children
    // split x (bit 0) and z (bit 1) at the centre, every node spans
    // [y_min, y_max] in y
    const float rmin[3] = {-1.f, use_quad ? y_min : -1.f, -1.f}, rmax[3] = {1.f, use_quad ? y_max : 1.f, 1.f};
    int64_t cur = 0;
    set_node(cur++, -1, rmin, rmax, 0);
    int64_t start_layer = 0;
    for (int depth = 1; depth <= max_depth; ++depth) {
        int64_t n = cur - start_layer;
        for (int64_t j = 0; j < n; ++j) {
            int64_t id = start_layer + j;
            float c[3], mn[3], mx[3];
            for (int k = 0; k < 3; ++k) c[k] = (pmin[id * 3 + k] + pmax[id * 3 + k]) * 0.5f;
            for (int i = 0; i < NS; ++i) {
                for (int k = 0; k < 3; ++k) {
                    mn[k] = pmin[id * 3 + k];
                    mx[k] = pmax[id * 3 + k];
                }
                for (int b = 0; b < (use_quad ? 2 : 3); ++b) {
                    const int k = use_quad && b == 1 ? 2 : b;
                    if (((i >> b) & 1) == 0) {
                        mn[k] = pmin[id * 3 + k];
                        mx[k] = c[k];
                    } else {
                        mn[k] = c[k];
                        mx[k] = pmax[id * 3 + k];
                    }
                }
                ch[id * NS + i] = (int32_t)cur;
                set_node(cur++, id, mn, mx, depth);
            }
        }
        start_layer = cur - n * NS;
    }
    active_node_ids = torch::zeros({1}, torch::kLong);
    node_active_prefix_sum = torch::zeros({num_nodes}, torch::kLong);
    SetActive(0);
}

void HyperTreeBase::SetActive(int depth)
{
    node_active = (node_depth == depth).to(torch::kInt32);
    UpdateActive();
}

void HyperTreeBase::UpdateActive()
{
    auto act = node_active.to(torch::kCPU).contiguous();
    auto cul = node_culled.to(torch::kCPU).contiguous();
    TORCH_CHECK((act * cul).sum().item<int64_t>() == 0, "UpdateActive: a culled node is not allowed to be active");
    const int32_t* a = act.data_ptr<int32_t>();
    int64_t n = act.size(0);
    std::vector<int64_t> ids;
    auto prefix = torch::empty({n}, torch::kLong);
    int64_t* pp = prefix.data_ptr<int64_t>();
    int64_t count = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (a[i] == 1) {
            ids.push_back(i);
            pp[i] = count++;
        } else {
            pp[i] = -1;
        }
    }
    auto dev = node_position_min.device();
    active_node_ids = torch::from_blob(ids.data(), {(int64_t)ids.size()}, torch::kLong).clone().to(dev);
    node_active_prefix_sum = prefix.to(dev);
    node_to_grid = prefix.to(torch::kInt32).to(dev);
    node_active = node_active.to(dev);
    if (dev.is_cuda()) repack();
}

void HyperTreeBase::to(torch::Device device)
{
    node_parent = node_parent.to(device);
    node_children = node_children.to(device);
    node_position_min = node_position_min.to(device);
    node_position_max = node_position_max.to(device);
    node_scale = node_scale.to(device);
    node_error = node_error.to(device);
    node_max_density = node_max_density.to(device);
    node_active = node_active.to(device);
    node_culled = node_culled.to(device);
    node_depth = node_depth.to(device);
    node_active_prefix_sum = node_active_prefix_sum.to(device);
    active_node_ids = active_node_ids.to(device);
    node_to_grid = node_to_grid.to(device);
    if (device.is_cuda()) repack();
}

void HyperTreeBase::repack()
{
    int64_t n = NumNodes();
    packed_ = torch::empty({(int64_t)atdf_tree_packed_bytes(n)},
                           torch::TensorOptions().dtype(torch::kUInt8).device(node_position_min.device()));
    atdf_tree t{};
    t.num_nodes = n;
    t.node_position_min = node_position_min.data_ptr<float>();
    t.node_position_max = node_position_max.data_ptr<float>();
    t.node_children = node_children.data_ptr<int32_t>();
    t.node_parent = node_parent.data_ptr<int32_t>();
    t.node_active = node_active.data_ptr<int32_t>();
    t.children_per_node = (int32_t)node_children.size(1);
    check_status(atdf_tree_pack(&t, packed_.data_ptr(), current_stream()), "atdf_tree_pack");
    const int depth = leaf_depth();
    leaf_table_ = torch::empty({atdf_leaf_table_numel(depth)},
                               torch::TensorOptions().dtype(torch::kInt32).device(node_position_min.device()));
    check_status(atdf_leaf_table(packed_.data_ptr(), n, depth, leaf_table_.data_ptr<int32_t>(), current_stream()),
                 "atdf_leaf_table");
}

namespace {
struct TraversalCounts {
    Tensor num_iv, num_s, iv_off, smp_off;
    Tensor cache;  // the count pass's hits, replayed by the write pass (dense ray lists)
    int64_t total_iv = 0, total_s = 0;
    int max_per_ray = 0;
};

// Hit-cache slots per ray: rays with more hits are walked again by the write
// pass (the adaptive 512^3 octrees give about 22 per ray).
constexpr int32_t kTraverseCacheSlots = 64;

TraversalCounts count_and_scan(const void* packed, int64_t num_nodes, torch::Device dev, const RayList& rays,
                               const atdf_sampling& s)
{
    TraversalCounts tc;
    const int64_t R = rays.size();
    auto i32 = torch::TensorOptions().dtype(torch::kInt32).device(dev);
    auto i64 = torch::TensorOptions().dtype(torch::kLong).device(dev);
    tc.num_iv = torch::empty({R}, i32);
    tc.num_s = torch::empty({R}, i32);
    tc.iv_off = torch::empty({R + 1}, i64);
    tc.smp_off = torch::empty({R + 1}, i64);
    auto status = torch::zeros({4}, i32);
    auto st = current_stream();
    if (R > 0 && !s.ray_list) {
        tc.cache = torch::empty({(int64_t)atdf_traverse_cache_bytes(R, kTraverseCacheSlots)},
                                torch::TensorOptions().dtype(torch::kUInt8).device(dev));
        check_status(atdf_traverse_count_cached(packed, num_nodes, rays.origin.data_ptr<float>(),
                                                rays.direction.data_ptr<float>(), R, &s, tc.num_iv.data_ptr<int32_t>(),
                                                tc.num_s.data_ptr<int32_t>(), status.data_ptr<int32_t>(),
                                                tc.cache.data_ptr(), kTraverseCacheSlots, st),
                     "atdf_traverse_count_cached");
    } else if (R > 0) {
        check_status(atdf_traverse_count(packed, num_nodes, rays.origin.data_ptr<float>(),
                                         rays.direction.data_ptr<float>(), R, &s, tc.num_iv.data_ptr<int32_t>(),
                                         tc.num_s.data_ptr<int32_t>(), status.data_ptr<int32_t>(), st),
                     "atdf_traverse_count");
    }
    auto ws = torch::empty({(int64_t)atdf_scan_workspace_bytes(R)},
                           torch::TensorOptions().dtype(torch::kUInt8).device(dev));
    check_status(atdf_exclusive_scan_i32(tc.num_iv.data_ptr<int32_t>(), R, tc.iv_off.data_ptr<int64_t>(),
                                         ws.data_ptr(), ws.numel(), st),
                 "atdf_exclusive_scan_i32");
    check_status(atdf_exclusive_scan_i32(tc.num_s.data_ptr<int32_t>(), R, tc.smp_off.data_ptr<int64_t>(),
                                         ws.data_ptr(), ws.numel(), st),
                 "atdf_exclusive_scan_i32");
    // one device -> host copy: status, totals (the reference reads its sample
    // count back the same way, HyperTree.cu:550)
    auto info = torch::cat({status.to(torch::kLong), tc.iv_off.slice(0, R, R + 1), tc.smp_off.slice(0, R, R + 1)})
                    .to(torch::kCPU);
    const int64_t* ip = info.data_ptr<int64_t>();
    TORCH_CHECK(ip[0] == 0, "CreateSamplesForRays: ", ip[0], " ray(s) hit more than ", s.max_intersections,
                " nodes (max_intersections)");
    TORCH_CHECK(ip[1] == 0, "CreateSamplesForRays: traversal order check failed on ", ip[1], " ray(s)");
    tc.max_per_ray = (int)ip[2];
    tc.total_iv = ip[4];
    tc.total_s = ip[5];
    return tc;
}

// The write pass over the rays count_and_scan counted (from its hit cache
// when it kept one).
void traverse_write(const void* packed, int64_t num_nodes, const RayList& rays, const atdf_sampling& s,
                    const TraversalCounts& tc, const atdf_intervals* iv, const atdf_samples* smp)
{
    if (tc.cache.defined())
        check_status(atdf_traverse_write_cached(packed, num_nodes, rays.origin.data_ptr<float>(),
                                                rays.direction.data_ptr<float>(), rays.size(), &s,
                                                tc.iv_off.data_ptr<int64_t>(), tc.smp_off.data_ptr<int64_t>(), iv, smp,
                                                tc.cache.data_ptr(), kTraverseCacheSlots, current_stream()),
                     "atdf_traverse_write_cached");
    else
        check_status(atdf_traverse_write(packed, num_nodes, rays.origin.data_ptr<float>(),
                                         rays.direction.data_ptr<float>(), rays.size(), &s,
                                         tc.iv_off.data_ptr<int64_t>(), tc.smp_off.data_ptr<int64_t>(), iv, smp,
                                         current_stream()),
                     "atdf_traverse_write");
}

void check_rays(const RayList& rays)
{
    require_device_f32(rays.origin, "rays.origin");
    require_device_f32(rays.direction, "rays.direction");
    TORCH_CHECK(rays.origin.dim() == 2 && rays.origin.size(1) == 3, "rays.origin must be [R, 3]");
    TORCH_CHECK(rays.direction.sizes() == rays.origin.sizes(), "rays.direction must match rays.origin");
}
}  // namespace

Intervals HyperTreeBase::intervals_for(const void* packed, int64_t num_nodes, const RayList& rays,
                                       const SamplingParams& sp) const
{
    check_rays(rays);
    TORCH_CHECK(sp.jitter_mode != 1, "CreateIntervalsForRays: jitter_mode 1 needs materialised samples");
    atdf_sampling s = sp.view();
    auto tc = count_and_scan(packed, num_nodes, device(), rays, s);
    Intervals iv;
    auto dev = device();
    auto i32 = torch::TensorOptions().dtype(torch::kInt32).device(dev);
    auto f32 = torch::TensorOptions().dtype(torch::kFloat32).device(dev);
    iv.ray_offset = tc.iv_off;
    iv.ray_num_samples = tc.num_s;
    iv.num_intervals = tc.total_iv;
    iv.num_samples = tc.total_s;
    iv.max_samples_per_ray = tc.max_per_ray;
    const int64_t H = tc.total_iv;
    iv.node = torch::empty({H}, i32);
    iv.tmin = torch::empty({H}, f32);
    iv.tmax = torch::empty({H}, f32);
    iv.count = torch::empty({H}, i32);
    iv.first = torch::empty({H}, i32);
    iv.ray = torch::empty({H}, i32);
    if (H > 0) {
        atdf_intervals v = iv.view();
        traverse_write(packed, num_nodes, rays, s, tc, &v, nullptr);
    }
    return iv;
}

Intervals HyperTreeBase::CreateIntervalsForRays(const RayList& rays, const SamplingParams& sp)
{
    return intervals_for(packed(), NumNodes(), rays, sp);
}

namespace {
Tensor pack_box(const std::vector<float>& roi_min, const std::vector<float>& roi_max, torch::Device dev)
{
    TORCH_CHECK(roi_min.size() == 3 && roi_max.size() == 3, "roi_min and roi_max need 3 entries");
    auto packed = torch::empty({(int64_t)atdf_tree_packed_bytes(1)},
                               torch::TensorOptions().dtype(torch::kUInt8).device(dev));
    check_status(atdf_tree_pack_box(roi_min.data(), roi_max.data(), packed.data_ptr(), current_stream()),
                 "atdf_tree_pack_box");
    return packed;
}
}  // namespace

Intervals HyperTreeBase::CreateIntervalsForRays_area(const RayList& rays, const SamplingParams& sp,
                                                     std::vector<float> roi_min, std::vector<float> roi_max)
{
    auto box = pack_box(roi_min, roi_max, device());
    return intervals_for(box.data_ptr(), 1, rays, sp);
}

SampleList HyperTreeBase::samples_for(const void* packed, int64_t num_nodes, const RayList& rays,
                                      int max_samples_per_node, bool interval_jitter) const
{
    check_rays(rays);
    SamplingParams sp;
    sp.max_samples_per_node = max_samples_per_node;
    sp.jitter_mode = 0;
    atdf_sampling s = sp.view();
    auto tc = count_and_scan(packed, num_nodes, device(), rays, s);
    const int64_t S = tc.total_s;
    auto dev = device();
    auto f32 = torch::TensorOptions().dtype(torch::kFloat32).device(dev);
    auto i64 = torch::TensorOptions().dtype(torch::kLong).device(dev);
    SampleList list;
    list.global_coordinate = torch::empty({S, 3}, f32);
    list.weight = torch::empty({S}, f32);
    list.ray_t = torch::empty({S}, f32);
    list.ray_index = torch::empty({S}, i64);
    list.local_index_in_ray = torch::empty({S}, i64);
    list.node_id = torch::empty({S}, i64);
    list.max_samples_per_ray = tc.max_per_ray;
    Tensor rnd;
    if (interval_jitter) {
        // the reference draws one U[0,1) per sample slot (HyperTree.cu:533)
        rnd = torch::rand({S}, f32);
        s.jitter_mode = 1;
        s.jitter = rnd.data_ptr<float>();
    }
    if (S > 0) {
        atdf_samples v{list.global_coordinate.data_ptr<float>(), list.weight.data_ptr<float>(),
                       list.ray_t.data_ptr<float>(),              list.ray_index.data_ptr<int64_t>(),
                       list.local_index_in_ray.data_ptr<int64_t>(), list.node_id.data_ptr<int64_t>()};
        traverse_write(packed, num_nodes, rays, s, tc, nullptr, &v);
    }
    return list;
}

SampleList HyperTreeBase::CreateSamplesForRays(const RayList& rays, int max_samples_per_node, bool interval_jitter)
{
    return samples_for(packed(), NumNodes(), rays, max_samples_per_node, interval_jitter);
}

SampleList HyperTreeBase::CreateSamplesForRays_area(const RayList& rays, int max_samples_per_node, bool interval_jitter,
                                                    std::vector<float> roi_min, std::vector<float> roi_max)
{
    auto box = pack_box(roi_min, roi_max, device());
    return samples_for(box.data_ptr(), 1, rays, max_samples_per_node, interval_jitter);
}

Tensor HyperTreeBase::ComputeLocalSamples(Tensor global_samples, Tensor node_indices)
{
    require_device_f32(global_samples, "global_samples");
    TORCH_CHECK(node_indices.is_cuda() && node_indices.scalar_type() == torch::kLong, "node_indices must be GPU int64");
    auto g = global_samples.reshape({-1, 3}).contiguous();
    auto nd = node_indices.reshape({-1}).contiguous();
    TORCH_CHECK(g.size(0) == nd.size(0), "ComputeLocalSamples: ", g.size(0), " samples vs ", nd.size(0), " node ids");
    auto local = torch::empty_like(g);
    check_status(atdf_local_samples(node_position_min.data_ptr<float>(), node_position_max.data_ptr<float>(),
                                    g.data_ptr<float>(), nd.data_ptr<int64_t>(), g.size(0), local.data_ptr<float>(),
                                    current_stream()),
                 "atdf_local_samples");
    return local.reshape(global_samples.sizes());
}

void HyperTreeBase::SetErrorForActiveNodes(Tensor error, const std::string& strategy)
{
    TORCH_CHECK(error.sizes() == node_error.sizes(), "SetErrorForActiveNodes: size mismatch");
    TORCH_CHECK(error.device() == node_error.device(), "SetErrorForActiveNodes: device mismatch");
    auto active_float = node_active.to(torch::kFloat32);
    auto inactive = 1 - active_float;
    auto invalid_errors = (node_error < 0).to(torch::kFloat32);
    error = active_float * error.to(torch::kFloat32);
    if (strategy == "override") {
        node_error = node_error * inactive + error;
    } else if (strategy == "min") {
        error = error + inactive * 1000000;
        auto filtered = node_error + (invalid_errors * active_float) * 1000000;
        node_error = torch::min(filtered, error);
    } else {
        TORCH_CHECK(false, "SetErrorForActiveNodes: unknown strategy ", strategy);
    }
}

void HyperTreeBase::UpdateCulling()
{
    auto culled = node_culled.to(torch::kCPU).contiguous();
    auto children = node_children.to(torch::kCPU).contiguous();
    auto active = node_active.to(torch::kCPU).contiguous();
    int32_t* cp = culled.data_ptr<int32_t>();
    const int32_t* chp = children.data_ptr<int32_t>();
    const int32_t* ap = active.data_ptr<int32_t>();
    const int64_t NS = children.size(1);
    bool changed = true;
    while (changed) {
        changed = false;
        for (int64_t i = 0; i < culled.size(0); ++i) {
            if (cp[i]) continue;
            if (chp[i * NS] < 0) continue;  // leaves have no children to inherit culling from
            bool all = true;
            for (int64_t c = 0; c < NS; ++c)
                if (!cp[chp[i * NS + c]]) all = false;
            if (all) {
                TORCH_CHECK(!ap[i], "UpdateCulling: active node with all children culled");
                cp[i] = 1;
                changed = true;
            }
        }
    }
    node_culled = culled.to(node_culled.device());
    node_error = node_error * (1 - node_culled);
}

// ---------------------------------------------------------------------------
// IndirectGridSample3D
// ---------------------------------------------------------------------------
namespace {
// The slot-major backward (samples sorted by slot, each slot's gradient in
// LDS) when a slot's gradient fits one workgroup's LDS; false otherwise, and
// the caller takes the per-sample backward.
bool slot_backward(int D, const Tensor& g, const Tensor& idx, const Tensor& u, const Tensor& go, Tensor& grad_grid,
                   Tensor& grad_uv)
{
    const int64_t S = idx.size(0), N = g.size(0);
    if (S == 0 || N == 0) return true;
    if (g[0].numel() > atdf_grid_sample_bwd_max_slot_floats()) return false;
    int32_t size[3] = {1, 1, 1};
    for (int d = 0; d < D; ++d) size[d] = (int32_t)g.size(2 + d);
    auto ws = torch::empty({(int64_t)atdf_grid_sample_bwd_workspace_bytes(N, S)}, g.options().dtype(torch::kUInt8));
    check_status(atdf_grid_sample_bwd_slots(D, g.data_ptr<float>(), N, (int)g.size(1), size, idx.data_ptr<int64_t>(),
                                            u.data_ptr<float>(), S, go.data_ptr<float>(), grad_grid.data_ptr<float>(),
                                            grad_uv.data_ptr<float>(), ws.data_ptr(), (size_t)ws.numel(),
                                            current_stream()),
                 "atdf_grid_sample_bwd_slots");
    return true;
}

struct IndirectGridSample3DFn : public torch::autograd::Function<IndirectGridSample3DFn> {
    static Tensor forward(torch::autograd::AutogradContext* ctx, Tensor grid, Tensor index, Tensor uv)
    {
        require_device_f32(grid, "multi_grid");
        require_device_f32(uv, "uv");
        TORCH_CHECK(grid.dim() == 5, "multi_grid must be [N, C, Z, Y, X]");
        TORCH_CHECK(index.dim() == 1 && index.scalar_type() == torch::kLong && index.is_cuda(), "index must be GPU int64 [S]");
        TORCH_CHECK(uv.dim() == 2 && uv.size(1) == 3 && uv.size(0) == index.size(0), "uv must be [S, 3]");
        auto idx = index.contiguous();
        const int64_t S = idx.size(0), C = grid.size(1);
        auto out = torch::empty({S, C}, grid.options());
        check_status(atdf_grid_sample3d_fwd(grid.data_ptr<float>(), grid.size(0), (int)C, (int)grid.size(2),
                                            (int)grid.size(3), (int)grid.size(4), idx.data_ptr<int64_t>(),
                                            uv.data_ptr<float>(), S, out.data_ptr<float>(), current_stream()),
                     "atdf_grid_sample3d_fwd");
        ctx->save_for_backward({grid, idx, uv});
        return out;
    }
    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_output)
    {
        auto saved = ctx->get_saved_variables();
        auto grid = saved[0], idx = saved[1], uv = saved[2];
        auto go = grad_output[0].contiguous();
        auto grad_grid = torch::zeros_like(grid);
        auto grad_uv = torch::zeros_like(uv);
        if (slot_backward(3, grid, idx, uv.contiguous(), go, grad_grid, grad_uv)) return {grad_grid, Tensor(), grad_uv};
        check_status(atdf_grid_sample3d_bwd(grid.data_ptr<float>(), grid.size(0), (int)grid.size(1), (int)grid.size(2),
                                            (int)grid.size(3), (int)grid.size(4), idx.data_ptr<int64_t>(),
                                            uv.data_ptr<float>(), idx.size(0), go.data_ptr<float>(),
                                            grad_grid.data_ptr<float>(), grad_uv.data_ptr<float>(), current_stream()),
                     "atdf_grid_sample3d_bwd");
        return {grad_grid, Tensor(), grad_uv};
    }
};

struct IndirectGridSample2DFn : public torch::autograd::Function<IndirectGridSample2DFn> {
    static Tensor forward(torch::autograd::AutogradContext* ctx, Tensor grid, Tensor index, Tensor uv)
    {
        require_device_f32(grid.contiguous(), "multi_grid");
        require_device_f32(uv.contiguous(), "uv");
        TORCH_CHECK(grid.dim() == 4, "multi_grid must be [N, C, H, W]");
        TORCH_CHECK(index.dim() == 1 && index.scalar_type() == torch::kLong && index.is_cuda(), "index must be GPU int64 [S]");
        TORCH_CHECK(uv.dim() == 2 && uv.size(1) == 2 && uv.size(0) == index.size(0), "uv must be [S, 2]");
        auto g = grid.contiguous();
        auto idx = index.contiguous();
        auto u = uv.contiguous();
        const int64_t S = idx.size(0), C = g.size(1);
        auto out = torch::empty({S, C}, g.options());
        check_status(atdf_grid_sample2d_fwd(g.data_ptr<float>(), g.size(0), (int)C, (int)g.size(2), (int)g.size(3),
                                            idx.data_ptr<int64_t>(), u.data_ptr<float>(), S, out.data_ptr<float>(),
                                            current_stream()),
                     "atdf_grid_sample2d_fwd");
        ctx->save_for_backward({g, idx, u});
        return out;
    }
    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_output)
    {
        auto saved = ctx->get_saved_variables();
        auto g = saved[0], idx = saved[1], u = saved[2];
        auto go = grad_output[0].contiguous();
        auto grad_grid = torch::zeros_like(g);
        auto grad_uv = torch::zeros_like(u);
        if (slot_backward(2, g, idx, u, go, grad_grid, grad_uv)) return {grad_grid, Tensor(), grad_uv};
        check_status(atdf_grid_sample2d_bwd(g.data_ptr<float>(), g.size(0), (int)g.size(1), (int)g.size(2),
                                            (int)g.size(3), idx.data_ptr<int64_t>(), u.data_ptr<float>(), idx.size(0),
                                            go.data_ptr<float>(), grad_grid.data_ptr<float>(),
                                            grad_uv.data_ptr<float>(), current_stream()),
                     "atdf_grid_sample2d_bwd");
        return {grad_grid, Tensor(), grad_uv};
    }
};

struct IndirectGridSample1DFn : public torch::autograd::Function<IndirectGridSample1DFn> {
    static Tensor forward(torch::autograd::AutogradContext* ctx, Tensor grid, Tensor index, Tensor uv)
    {
        require_device_f32(grid.contiguous(), "multi_grid");
        require_device_f32(uv.contiguous(), "uv");
        TORCH_CHECK(grid.dim() == 4 && grid.size(3) == 1, "multi_grid must be [N, C, W, 1]");
        TORCH_CHECK(index.dim() == 1 && index.scalar_type() == torch::kLong && index.is_cuda(), "index must be GPU int64 [S]");
        TORCH_CHECK(uv.numel() == index.size(0), "uv must hold one coordinate per sample");
        auto g = grid.contiguous();
        auto idx = index.contiguous();
        auto u = uv.contiguous();
        const int64_t S = idx.size(0), C = g.size(1);
        auto out = torch::empty({S, C}, g.options());
        check_status(atdf_grid_sample1d_fwd(g.data_ptr<float>(), g.size(0), (int)C, (int)g.size(2),
                                            idx.data_ptr<int64_t>(), u.data_ptr<float>(), S, out.data_ptr<float>(),
                                            current_stream()),
                     "atdf_grid_sample1d_fwd");
        ctx->save_for_backward({g, idx, u});
        return out;
    }
    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_output)
    {
        auto saved = ctx->get_saved_variables();
        auto g = saved[0], idx = saved[1], u = saved[2];
        auto go = grad_output[0].contiguous();
        auto grad_grid = torch::zeros_like(g);
        auto grad_uv = torch::zeros_like(u);
        if (slot_backward(1, g, idx, u, go, grad_grid, grad_uv)) return {grad_grid, Tensor(), grad_uv};
        check_status(atdf_grid_sample1d_bwd(g.data_ptr<float>(), g.size(0), (int)g.size(1), (int)g.size(2),
                                            idx.data_ptr<int64_t>(), u.data_ptr<float>(), idx.size(0),
                                            go.data_ptr<float>(), grad_grid.data_ptr<float>(),
                                            grad_uv.data_ptr<float>(), current_stream()),
                     "atdf_grid_sample1d_bwd");
        return {grad_grid, Tensor(), grad_uv};
    }
};

struct GradientCorrectionFn : public torch::autograd::Function<GradientCorrectionFn> {
    static Tensor forward(torch::autograd::AutogradContext* ctx, Tensor sample_position, Tensor ray_direction,
                          Tensor index)
    {
        require_device_f32(sample_position, "sample_position");
        require_device_f32(ray_direction, "ray_direction");
        TORCH_CHECK(index.scalar_type() == torch::kLong && index.is_cuda(), "index must be GPU int64");
        ctx->save_for_backward({ray_direction.contiguous(), index.contiguous()});
        return sample_position.view_as(sample_position);
    }
    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_output)
    {
        auto saved = ctx->get_saved_variables();
        auto dir = saved[0], idx = saved[1];
        auto go = grad_output[0].contiguous();
        auto grad = torch::empty_like(go);
        check_status(atdf_gradient_correction_bwd(dir.data_ptr<float>(), idx.data_ptr<int64_t>(), idx.numel(),
                                                  go.data_ptr<float>(), grad.data_ptr<float>(), current_stream()),
                     "atdf_gradient_correction_bwd");
        return {grad, Tensor(), Tensor()};
    }
};

struct IntegrateFn : public torch::autograd::Function<IntegrateFn> {
    static Tensor forward(torch::autograd::AutogradContext* ctx, Tensor values, Tensor weight, Tensor ray_index,
                          int64_t C, int64_t R)
    {
        auto v = values.reshape({-1, C}).contiguous();
        auto w = weight.reshape({-1}).contiguous();
        auto ri = ray_index.reshape({-1}).contiguous();
        require_device_f32(v, "sample_values");
        require_device_f32(w, "integration_weight");
        TORCH_CHECK(ri.scalar_type() == torch::kLong && ri.is_cuda(), "ray_index must be GPU int64");
        TORCH_CHECK(v.size(0) == w.size(0) && w.size(0) == ri.size(0), "IntegrateSamplesXRay: size mismatch");
        auto out = torch::zeros({C, R}, v.options());
        check_status(atdf_integrate_xray(v.data_ptr<float>(), w.data_ptr<float>(), ri.data_ptr<int64_t>(), v.size(0),
                                         (int)C, R, out.data_ptr<float>(), current_stream()),
                     "atdf_integrate_xray");
        ctx->save_for_backward({v, w, ri});
        ctx->saved_data["vshape"] = values.sizes().vec();
        ctx->saved_data["wshape"] = weight.sizes().vec();
        return out;
    }
    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_output)
    {
        auto saved = ctx->get_saved_variables();
        auto v = saved[0], w = saved[1], ri = saved[2];
        // d out[c, ray[s]] / d v[s, c] = w[s];  / d w[s] = sum_c v[s, c] * g
        auto g = grad_output[0].t().index_select(0, ri);  // [S, C]
        auto gv = (g * w.unsqueeze(1)).reshape(ctx->saved_data["vshape"].toIntVector());
        auto gw = (g * v).sum(1).reshape(ctx->saved_data["wshape"].toIntVector());
        return {gv, gw, Tensor(), Tensor(), Tensor()};
    }
};
}  // namespace

Tensor IndirectGridSample3D(Tensor multi_grid, Tensor index, Tensor uv)
{
    return IndirectGridSample3DFn::apply(multi_grid, index, uv);
}

Tensor IndirectGridSample2D(Tensor multi_grid, Tensor index, Tensor uv)
{
    return IndirectGridSample2DFn::apply(multi_grid, index, uv);
}

Tensor IndirectGridSample1D(Tensor multi_grid, Tensor index, Tensor uv)
{
    return IndirectGridSample1DFn::apply(multi_grid, index, uv);
}

Tensor GradientCorrection(Tensor sample_position, Tensor ray_direction, Tensor index)
{
    return GradientCorrectionFn::apply(sample_position, ray_direction, index);
}

Tensor ComputeBoundWeight(Tensor node_id, const RoiMaps& roi)
{
    TORCH_CHECK(node_id.scalar_type() == torch::kLong && node_id.is_cuda(), "node_id must be GPU int64");
    auto nid = node_id.reshape({-1}).contiguous();
    auto out = torch::empty({nid.size(0)}, nid.options());
    check_status(atdf_compute_bound_weight(nid.data_ptr<int64_t>(), nid.size(0), roi.in_roi_flag.data_ptr<int32_t>(),
                                           out.data_ptr<int64_t>(), current_stream()),
                 "atdf_compute_bound_weight");
    return out;
}

namespace {
// geometry/geometry.h:426-449: a node is in the ROI when, on every axis, its
// min, its max or its "mid" (min + max, as the reference computes it) lies
// strictly inside the ROI interval; the y interval is widened by epsilon.
bool node_in_roi(const float* mn, const float* mx, const float* roi_min, const float* roi_max, float eps)
{
    for (int k = 0; k < 3; ++k) {
        const float lo = k == 1 ? roi_min[k] - eps : roi_min[k];
        const float hi = k == 1 ? roi_max[k] + eps : roi_max[k];
        const float mid = mn[k] + mx[k];
        if (!((mn[k] > lo && mn[k] < hi) || (mx[k] > lo && mx[k] < hi) || (mid > lo && mid < hi))) return false;
    }
    return true;
}
}  // namespace

RoiMaps HyperTreeBase::RoiSplit(std::vector<float> roi_min, std::vector<float> roi_max, float epsilon) const
{
    TORCH_CHECK(roi_min.size() == 3 && roi_max.size() == 3, "RoiSplit: ROI bounds must have 3 entries");
    auto pmin = node_position_min.to(torch::kCPU).contiguous();
    auto pmax = node_position_max.to(torch::kCPU).contiguous();
    auto act = node_active.to(torch::kCPU).contiguous();
    const float* mn = pmin.data_ptr<float>();
    const float* mx = pmax.data_ptr<float>();
    const int32_t* a = act.data_ptr<int32_t>();
    const int64_t n = act.size(0);
    std::vector<int64_t> in_ids, out_ids, pin(n, -1), pout(n, -1);
    std::vector<int32_t> flag(n, 0);
    for (int64_t i = 0; i < n; ++i) {
        if (a[i] != 1) continue;
        if (node_in_roi(mn + 3 * i, mx + 3 * i, roi_min.data(), roi_max.data(), epsilon)) {
            pin[i] = (int64_t)in_ids.size();
            in_ids.push_back(i);
            flag[i] = 1;
        } else {
            pout[i] = (int64_t)out_ids.size();
            out_ids.push_back(i);
        }
    }
    auto dev = node_position_min.device();
    auto to_t = [&](std::vector<int64_t>& v) {
        return torch::from_blob(v.data(), {(int64_t)v.size()}, torch::kLong).clone().to(dev);
    };
    RoiMaps m;
    m.in_roi_node_index = to_t(in_ids);
    m.out_roi_node_index = to_t(out_ids);
    m.prefix_inroi = to_t(pin);
    m.prefix_outroi = to_t(pout);
    m.in_roi_flag = torch::from_blob(flag.data(), {n}, torch::kInt32).clone().to(dev);
    m.roi_min = roi_min;
    m.roi_max = roi_max;
    return m;
}

Tensor SampleVolumeIndirectFeature(HyperTreeBase& tree, const RoiMaps& roi, const VMFields& f,
                                   Tensor global_coordinate, Tensor node_id)
{
    if (global_coordinate.numel() == 0) return global_coordinate;
    auto g = global_coordinate.reshape({-1, 3});
    auto nid = node_id.reshape({-1}).contiguous();
    const int64_t S = nid.size(0);
    auto dev = g.device();
    // route samples by the ROI membership of their node
    auto bound = ComputeBoundWeight(nid, roi);
    auto all = torch::arange(S, torch::TensorOptions().dtype(torch::kLong).device(dev));
    auto in_idx = torch::masked_select(all, bound >= 0);
    auto out_idx = torch::masked_select(all, bound < 0);
    auto nid_in = nid.index_select(0, in_idx);
    auto nid_out = nid.index_select(0, out_idx);
    auto g_in = g.index_select(0, in_idx);
    auto g_out = g.index_select(0, out_idx);
    Tensor feat_in;
    {
        // in-ROI nodes: plane x line features (geometry_tensor_qo.cpp:279-327)
        auto local = tree.ComputeLocalSamples(g_in, nid_in);
        auto slot = roi.prefix_inroi.index_select(0, nid_in).contiguous();
        auto lx = local.slice(1, 0, 1), ly = local.slice(1, 1, 2), lz = local.slice(1, 2, 3);
        auto plane_yz = torch::cat({ly, lz}, 1);
        auto plane_xz = torch::cat({lx, lz}, 1);
        auto plane_xy = torch::cat({lx, ly}, 1);
        auto vec_id = torch::zeros({local.size(0)}, torch::TensorOptions().dtype(torch::kLong).device(dev));
        auto line_x = 2. * (g_in.slice(1, 0, 1) - f.roi_min[0]) / (f.roi_max[0] - f.roi_min[0]) - 1;
        auto line_y = ly;
        auto line_z = g_in.slice(1, 2, 3);
        if (f.tensor_op == "cat") {
            auto mat = torch::cat({IndirectGridSample2D(f.plane_yz, slot, plane_yz),
                                   IndirectGridSample2D(f.plane_xz, slot, plane_xz),
                                   IndirectGridSample2D(f.plane_xy, slot, plane_xy)},
                                  1);
            auto vec = torch::cat({IndirectGridSample1D(f.line_x, vec_id, line_x),
                                   IndirectGridSample1D(f.line_y, vec_id, line_y),
                                   IndirectGridSample1D(f.line_z, vec_id, line_z)},
                                  1);
            feat_in = mat * vec;
        } else {
            TORCH_CHECK(f.tensor_op == "add", "tensor_op must be add or cat");
            feat_in = IndirectGridSample2D(f.plane_yz, slot, plane_yz) * IndirectGridSample1D(f.line_x, vec_id, line_x);
            feat_in = feat_in + IndirectGridSample2D(f.plane_xz, slot, plane_xz) *
                                    IndirectGridSample1D(f.line_y, vec_id, line_y);
            feat_in = feat_in + IndirectGridSample2D(f.plane_xy, slot, plane_xy) *
                                    IndirectGridSample1D(f.line_z, vec_id, line_z);
        }
    }
    Tensor feat_out;
    if (f.grid_out.defined()) {
        // out-of-ROI nodes: explicit 3-D grids (geometry_tensor_qo.cpp:336-347)
        auto local = tree.ComputeLocalSamples(g_out, nid_out);
        auto slot = roi.prefix_outroi.index_select(0, nid_out).contiguous();
        feat_out = IndirectGridSample3D(f.grid_out, slot, local);
    }
    auto feats = torch::zeros({S, feat_in.size(1)}, feat_in.options());
    feats = feats.index_add(0, in_idx, feat_in);
    if (feat_out.defined()) feats = feats.index_add(0, out_idx, feat_out);
    return feats;
}

Tensor IntegrateSamplesXRay(Tensor sample_values, Tensor integration_weight, Tensor ray_index, int num_channels,
                            int64_t num_rays)
{
    if (sample_values.numel() == 0)
        return torch::zeros({num_channels, num_rays}, sample_values.options());
    return IntegrateFn::apply(sample_values, integration_weight, ray_index, (int64_t)num_channels, num_rays);
}

std::pair<Tensor, Tensor> AccumulateSampleLossPerNode(const SampleList& samples, Tensor per_ray_loss,
                                                      int64_t num_nodes)
{
    auto prl = per_ray_loss.to(torch::kFloat32).contiguous();
    TORCH_CHECK(prl.dim() == 2, "per_ray_loss must be [C, R]");
    auto opts = torch::TensorOptions().dtype(torch::kFloat64).device(prl.device());
    auto loss_sum = torch::zeros({num_nodes}, opts);
    auto weight_sum = torch::zeros({num_nodes}, opts);
    check_status(atdf_accumulate_loss_per_node(prl.data_ptr<float>(), (int)prl.size(0), prl.size(1),
                                               samples.ray_index.data_ptr<int64_t>(), samples.weight.data_ptr<float>(),
                                               samples.node_id.data_ptr<int64_t>(), samples.size(), num_nodes,
                                               loss_sum.data_ptr<double>(), weight_sum.data_ptr<double>(),
                                               current_stream()),
                 "atdf_accumulate_loss_per_node");
    return {loss_sum, weight_sum};
}

// ---------------------------------------------------------------------------
// fused projection
// ---------------------------------------------------------------------------
// Adjoint workspace: atdf_project_bwd's, plus the staging region of its
// two-pass form when enabled (the list's sample count sizes it).
static int64_t adjoint_workspace_bytes(int64_t R, const Intervals& iv, const atdf_field& f, bool two_pass)
{
    size_t n = atdf_project_bwd_workspace_bytes(R, iv.num_intervals, &f);
    if (two_pass) n += atdf_project_bwd_staging_bytes(iv.num_intervals, iv.num_samples, &f);
    return (int64_t)n;
}

atdf_field make_field(const HyperTreeBase& tree, const Tensor& grid, const Tensor& mlp, const FieldParams& fp)
{
    TORCH_CHECK(grid.is_cuda() && grid.is_contiguous(), "grid must be a contiguous GPU tensor");
    TORCH_CHECK(grid.dim() == 5 && grid.size(1) == fp.G && grid.size(2) == fp.G && grid.size(3) == fp.G &&
                    grid.size(4) == fp.C,
                "grid must be [N, G, G, G, C] = [N, ", fp.G, ", ", fp.G, ", ", fp.G, ", ", fp.C, "]");
    TORCH_CHECK(grid.scalar_type() == torch::kFloat32 || grid.scalar_type() == torch::kBFloat16,
                "grid must be float32 or bfloat16");
    require_device_f32(mlp, "mlp");
    const int64_t np = (int64_t)fp.H * fp.C + fp.H + (int64_t)fp.H * fp.H + fp.H + fp.H + 1;
    TORCH_CHECK(mlp.numel() == np, "mlp has ", mlp.numel(), " parameters, expected ", np);
    TORCH_CHECK(grid.size(0) >= tree.NumActiveNodes(), "grid has ", grid.size(0), " node grids for ",
                tree.NumActiveNodes(), " active nodes");
    atdf_field f{};
    f.C = fp.C;
    f.H = fp.H;
    f.G = fp.G;
    f.grid_dtype = grid.scalar_type() == torch::kBFloat16 ? ATDF_GRID_BF16 : ATDF_GRID_F32;
    f.softplus_beta = fp.softplus_beta;
    f.rays_per_wave = fp.rays_per_wave;
    f.decoder_dtype = fp.bf16_decoder ? ATDF_DECODER_BF16 : ATDF_DECODER_F32;
    f.grid = grid.data_ptr();
    f.num_grids = grid.size(0);
    f.node_to_grid = tree.node_to_grid.data_ptr<int32_t>();
    f.mlp = mlp.data_ptr<float>();
    return f;
}

namespace {

struct ProjectRaysFn : public torch::autograd::Function<ProjectRaysFn> {
    static Tensor forward(torch::autograd::AutogradContext* ctx, Tensor grid, Tensor mlp, Tensor origin,
                          Tensor direction, HyperTreeBase* tree, SamplingParams sp, FieldParams fp)
    {
        RayList rays{origin, direction};
        Intervals iv = tree->CreateIntervalsForRays(rays, sp);
        atdf_field f = make_field(*tree, grid, mlp, fp);
        atdf_sampling s = sp.view();
        auto P = torch::empty({rays.size()}, origin.options());
        atdf_intervals v = iv.view();
        check_status(atdf_project_fwd(tree->packed(), tree->NumNodes(), origin.data_ptr<float>(),
                                      direction.data_ptr<float>(), rays.size(), &s, iv.ray_offset.data_ptr<int64_t>(),
                                      &v, &f, P.data_ptr<float>(), current_stream()),
                     "atdf_project_fwd");
        ctx->save_for_backward({grid, mlp, origin, direction, iv.ray_offset, iv.node, iv.tmin, iv.tmax, iv.count,
                                iv.first, iv.ray});
        ctx->saved_data["tree"] = (int64_t)reinterpret_cast<intptr_t>(tree);
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
        ctx->saved_data["twopass"] = (bool)fp.two_pass_adjoint;
        ctx->saved_data["H_iv"] = iv.num_intervals;
        return P;
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_output)
    {
        auto sv = ctx->get_saved_variables();
        auto grid = sv[0], mlp = sv[1], origin = sv[2], direction = sv[3];
        Intervals iv;
        iv.ray_offset = sv[4];
        iv.node = sv[5];
        iv.tmin = sv[6];
        iv.tmax = sv[7];
        iv.count = sv[8];
        iv.first = sv[9];
        iv.ray = sv[10];
        iv.num_intervals = ctx->saved_data["H_iv"].toInt();
        auto* tree = reinterpret_cast<HyperTreeBase*>((intptr_t)ctx->saved_data["tree"].toInt());
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
        atdf_field f = make_field(*tree, grid, mlp, fp);
        atdf_sampling s = sp.view();
        auto dP = grad_output[0].to(torch::kFloat32).contiguous();
        const int64_t R = origin.size(0);
        auto grad_grid = torch::zeros(grid.sizes(), grid.options().dtype(torch::kFloat32));
        auto grad_mlp = torch::zeros_like(mlp);
        auto ws = torch::empty({adjoint_workspace_bytes(R, iv, f, fp.two_pass_adjoint)},
                               torch::TensorOptions().dtype(torch::kUInt8).device(grid.device()));
        atdf_intervals v = iv.view();
        check_status(atdf_project_bwd(tree->packed(), tree->NumNodes(), origin.data_ptr<float>(),
                                      direction.data_ptr<float>(), R, &s, iv.ray_offset.data_ptr<int64_t>(), &v, &f,
                                      dP.data_ptr<float>(), grad_grid.data_ptr<float>(), grad_mlp.data_ptr<float>(),
                                      ws.data_ptr(), ws.numel(), current_stream()),
                     "atdf_project_bwd");
        if (grid.scalar_type() != torch::kFloat32) grad_grid = grad_grid.to(grid.scalar_type());
        return {grad_grid, grad_mlp, Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};
}  // namespace

Tensor ProjectRays(HyperTreeBase& tree, const RayList& rays, Tensor grid, Tensor mlp, const SamplingParams& sp,
                   const FieldParams& fp)
{
    check_rays(rays);
    return ProjectRaysFn::apply(grid, mlp, rays.origin, rays.direction, &tree, sp, fp);
}

namespace {
std::tuple<Tensor, Tensor, Tensor> node_stats(HyperTreeBase& tree, const Tensor& grid, const Tensor& mlp,
                                              const FieldParams& fp, int mode, int n, int rounds, uint64_t seed,
                                              const Tensor& point_mask = Tensor())
{
    atdf_field f = make_field(tree, grid, mlp, fp);
    const int64_t NA = tree.NumActiveNodes();
    auto dev = grid.device();
    auto mx = torch::empty({NA}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    auto s1 = torch::empty({NA}, torch::TensorOptions().dtype(torch::kFloat64).device(dev));
    auto s2 = torch::empty({NA}, torch::TensorOptions().dtype(torch::kFloat64).device(dev));
    auto ids = tree.active_node_ids.contiguous();
    check_status(atdf_node_density_stats_masked(
                     tree.packed(), tree.NumNodes(), ids.data_ptr<int64_t>(), NA, &f, mode, n, rounds, seed,
                     point_mask.defined() ? point_mask.data_ptr<float>() : nullptr, mx.data_ptr<float>(),
                     s1.data_ptr<double>(), s2.data_ptr<double>(), current_stream()),
                 "atdf_node_density_stats_masked");
    return {mx, s1, s2};
}
}  // namespace

void ComputeMaxDensity(HyperTreeBase& tree, Tensor grid, Tensor mlp, const FieldParams& fp, int n,
                       const SceneBase* scene)
{
    Tensor mask;
    if (scene) {
        auto pos = tree.UniformGlobalSamples(tree.active_node_ids, n);
        mask = scene->PointInAnyImage(pos).reshape({-1}).contiguous();
    }
    auto [mx, s1, s2] = node_stats(tree, grid, mlp, fp, 0, n, 1, 0, mask);
    (void)s1;
    (void)s2;
    tree.node_max_density = tree.node_max_density.to(mx.device()).index_copy(0, tree.active_node_ids, mx);
}

Tensor ComputeStd(HyperTreeBase& tree, Tensor grid, Tensor mlp, const FieldParams& fp, int n, int rounds,
                  uint64_t seed)
{
    auto [mx, s1, s2] = node_stats(tree, grid, mlp, fp, 1, n, rounds, seed);
    (void)mx;
    const double cnt = (double)n * n * n * rounds;
    auto var = (s2 - s1 * s1 / cnt) / (cnt - 1.0);
    auto sd = torch::sqrt(torch::clamp_min(var, 0.0));
    auto out = torch::zeros({tree.NumNodes()}, sd.options());
    return out.index_copy(0, tree.active_node_ids, sd);
}

TrainState MakeTrainState(Tensor grid, Tensor mlp, bool bf16_grid, float lr_grid, float lr_mlp)
{
    require_device_f32(grid, "grid");
    require_device_f32(mlp, "mlp");
    TrainState st;
    if (bf16_grid) {
        st.grid_master = grid.clone();
        st.grid = torch::empty(grid.sizes(), grid.options().dtype(torch::kBFloat16));
        check_status(atdf_f32_to_bf16(st.grid_master.data_ptr<float>(),
                                      reinterpret_cast<uint16_t*>(st.grid.data_ptr()), grid.numel(), current_stream()),
                     "atdf_f32_to_bf16");
    } else {
        st.grid = grid.clone();
    }
    st.mlp = mlp.clone();
    st.grad_grid = torch::zeros(grid.sizes(), grid.options());
    st.grad_mlp = torch::zeros_like(mlp);
    st.m_grid = torch::zeros(grid.sizes(), grid.options());
    st.v_grid = torch::zeros(grid.sizes(), grid.options());
    st.m_mlp = torch::zeros_like(mlp);
    st.v_mlp = torch::zeros_like(mlp);
    st.loss = torch::zeros({1}, grid.options().dtype(torch::kFloat64));
    st.lr_grid = lr_grid;
    st.lr_mlp = lr_mlp;
    return st;
}

Tensor TrainStep(HyperTreeBase& tree, TrainState& st, const RayList& rays, Tensor target, const SamplingParams& sp,
                 const FieldParams& fp)
{
    check_rays(rays);
    Intervals iv = tree.CreateIntervalsForRays(rays, sp);
    return TrainStepIntervals(tree, st, rays, target, sp, fp, iv);
}

Tensor TrainStepIntervals(HyperTreeBase& tree, TrainState& st, const RayList& rays, Tensor target,
                          const SamplingParams& sp, const FieldParams& fp, const Intervals& iv)
{
    check_rays(rays);
    require_device_f32(target, "target");
    const int64_t R = rays.size();
    TORCH_CHECK(target.numel() == R, "target must hold one value per ray");
    auto stream = current_stream();
    atdf_field f = make_field(tree, st.grid, st.mlp, fp);
    atdf_sampling s = sp.view();
    atdf_intervals v = iv.view();
    auto P = torch::empty({R}, rays.origin.options());
    auto ws = torch::empty({adjoint_workspace_bytes(R, iv, f, fp.two_pass_adjoint)},
                           torch::TensorOptions().dtype(torch::kUInt8).device(rays.origin.device()));
    const int64_t* ivo = iv.ray_offset.data_ptr<int64_t>();
    const float *O = rays.origin.data_ptr<float>(), *D = rays.direction.data_ptr<float>();
    if (fp.node_forward) {
        // one counting sort for both directions; node-major forward over it
        check_status(atdf_project_prepare(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &v, &f, ws.data_ptr(),
                                          ws.numel(), stream),
                     "atdf_project_prepare");
        check_status(atdf_project_fwd_prepared(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &v, &f,
                                               P.data_ptr<float>(), ws.data_ptr(), ws.numel(), stream),
                     "atdf_project_fwd_prepared");
    } else {
        check_status(atdf_project_fwd(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &v, &f, P.data_ptr<float>(),
                                      stream),
                     "atdf_project_fwd");
    }
    auto dP = torch::empty({R}, rays.origin.options());
    auto mse_ws = torch::empty({(int64_t)atdf_mse_workspace_bytes(R)},
                               torch::TensorOptions().dtype(torch::kUInt8).device(rays.origin.device()));
    check_status(atdf_mse_loss(P.data_ptr<float>(), target.data_ptr<float>(), R, 1.0f, dP.data_ptr<float>(),
                               st.loss.data_ptr<double>(), mse_ws.data_ptr(), mse_ws.numel(), stream),
                 "atdf_mse_loss");
    st.step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    // adjoint and the grids' Adam step in one pass (grad_grid stays zero)
    atdf_grid_adam ga{};
    const bool bf16 = st.grid_master.defined();
    ga.param = bf16 ? st.grid_master.data_ptr<float>() : st.grid.data_ptr<float>();
    ga.param_bf16 = bf16 ? reinterpret_cast<uint16_t*>(st.grid.data_ptr()) : nullptr;
    ga.exp_avg = st.m_grid.data_ptr<float>();
    ga.exp_avg_sq = st.v_grid.data_ptr<float>();
    ga.lr = st.lr_grid;
    ga.beta1 = b1;
    ga.beta2 = b2;
    ga.eps = eps;
    ga.step = st.step;
    if (fp.node_forward)
        check_status(atdf_project_bwd_prepared(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &v, &f,
                                               dP.data_ptr<float>(), st.grad_grid.data_ptr<float>(),
                                               st.grad_mlp.data_ptr<float>(), &ga, ws.data_ptr(), ws.numel(), stream),
                     "atdf_project_bwd_prepared");
    else
        check_status(atdf_project_bwd_adam(tree.packed(), tree.NumNodes(), O, D, R, &s, ivo, &v, &f,
                                           dP.data_ptr<float>(), st.grad_grid.data_ptr<float>(),
                                           st.grad_mlp.data_ptr<float>(), &ga, ws.data_ptr(), ws.numel(), stream),
                     "atdf_project_bwd_adam");
    check_status(atdf_adam_step(st.mlp.data_ptr<float>(), st.grad_mlp.data_ptr<float>(), st.m_mlp.data_ptr<float>(),
                                st.v_mlp.data_ptr<float>(), st.mlp.numel(), st.lr_mlp, b1, b2, eps, st.step, 1, stream),
                 "atdf_adam_step");
    return st.loss;
}

namespace {
Tensor project_intervals(HyperTreeBase& tree, const RayList& rays, const Intervals& iv, const Tensor& grid,
                         const Tensor& mlp, const SamplingParams& sp, const FieldParams& fp, Tensor* bins)
{
    check_rays(rays);
    const int64_t R = rays.size();
    atdf_field f = make_field(tree, grid, mlp, fp);
    atdf_sampling s = sp.view();
    atdf_intervals v = iv.view();
    auto P = torch::empty({R}, rays.origin.options());
    if (bins) {
        *bins = torch::empty({R, ATDF_PROPOSAL_BINS}, rays.origin.options());
        check_status(atdf_project_fwd_bins(tree.packed(), tree.NumNodes(), rays.origin.data_ptr<float>(),
                                           rays.direction.data_ptr<float>(), R, &s, iv.ray_offset.data_ptr<int64_t>(),
                                           &v, &f, P.data_ptr<float>(), bins->data_ptr<float>(), current_stream()),
                     "atdf_project_fwd_bins");
    } else {
        check_status(atdf_project_fwd(tree.packed(), tree.NumNodes(), rays.origin.data_ptr<float>(),
                                      rays.direction.data_ptr<float>(), R, &s, iv.ray_offset.data_ptr<int64_t>(), &v,
                                      &f, P.data_ptr<float>(), current_stream()),
                     "atdf_project_fwd");
    }
    return P;
}

atdf_proposal proposal_view(const HyperTreeBase& tree, const ProposalState& ps, const SamplingParams* sp = nullptr)
{
    TORCH_CHECK(ps.params.defined() && ps.params.is_cuda(), "proposal parameters must be on the device");
    TORCH_CHECK(ps.coarse.defined() && ps.coarse.numel() == tree.NumNodes(),
                "proposal coarse densities missing (run ProposalRefreshCoarse)");
    atdf_proposal p{};
    p.leaf_table = tree.leaf_table();
    p.table_depth = tree.leaf_depth();
    p.mix = ps.pp.mix;
    p.kl_threshold = ps.pp.kl_threshold;
    p.coarse = ps.coarse.data_ptr<float>();
    p.params = ps.params.data_ptr<float>();
    if (sp) {
        p.jitter_mode = sp->jitter_mode;
        p.jitter_seed = sp->jitter_seed;
    }
    return p;
}
}  // namespace

Tensor ProjectIntervals(HyperTreeBase& tree, const RayList& rays, const Intervals& iv, Tensor grid, Tensor mlp,
                        const SamplingParams& sp, const FieldParams& fp)
{
    return project_intervals(tree, rays, iv, grid, mlp, sp, fp, nullptr);
}

std::pair<Tensor, Tensor> ProjectIntervalsBackward(HyperTreeBase& tree, const RayList& rays, const Intervals& iv,
                                                   Tensor grid, Tensor mlp, const SamplingParams& sp,
                                                   const FieldParams& fp, Tensor dLdP)
{
    check_rays(rays);
    const int64_t R = rays.size();
    require_device_f32(dLdP, "dLdP");
    TORCH_CHECK(dLdP.numel() == R, "dLdP must hold one value per ray");
    atdf_field f = make_field(tree, grid, mlp, fp);
    atdf_sampling s = sp.view();
    atdf_intervals v = iv.view();
    auto grad_grid = torch::zeros(grid.sizes(), grid.options().dtype(torch::kFloat32));
    auto grad_mlp = torch::zeros_like(mlp);
    auto ws = torch::empty({adjoint_workspace_bytes(R, iv, f, fp.two_pass_adjoint)},
                           torch::TensorOptions().dtype(torch::kUInt8).device(grid.device()));
    check_status(atdf_project_bwd(tree.packed(), tree.NumNodes(), rays.origin.data_ptr<float>(),
                                  rays.direction.data_ptr<float>(), R, &s, iv.ray_offset.data_ptr<int64_t>(), &v, &f,
                                  dLdP.data_ptr<float>(), grad_grid.data_ptr<float>(), grad_mlp.data_ptr<float>(),
                                  ws.data_ptr(), ws.numel(), current_stream()),
                 "atdf_project_bwd");
    return {grad_grid, grad_mlp};
}

std::pair<Tensor, Tensor> ProjectRaysBins(HyperTreeBase& tree, const RayList& rays, Tensor grid, Tensor mlp,
                                          const SamplingParams& sp, const FieldParams& fp)
{
    Intervals iv = tree.CreateIntervalsForRays(rays, sp);
    Tensor bins;
    Tensor P = project_intervals(tree, rays, iv, grid, mlp, sp, fp, &bins);
    return {P, bins};
}

ProposalState MakeProposalState(Tensor params, float lr, const ProposalParams& pp)
{
    require_device_f32(params, "proposal params");
    TORCH_CHECK(params.numel() == atdf_proposal_num_params(), "proposal params must hold ",
                atdf_proposal_num_params(), " values");
    ProposalState ps;
    ps.params = params.clone();
    ps.grad = torch::zeros_like(params);
    ps.m = torch::zeros_like(params);
    ps.v = torch::zeros_like(params);
    ps.loss = torch::zeros({1}, params.options().dtype(torch::kFloat64));
    ps.lr = lr;
    ps.pp = pp;
    return ps;
}

void ProposalRefreshCoarse(HyperTreeBase& tree, ProposalState& ps, Tensor grid, Tensor mlp, const FieldParams& fp)
{
    const int n = ps.pp.coarse_lattice;
    TORCH_CHECK(n >= 2, "coarse_lattice must be >= 2");
    auto [mx, s1, s2] = node_stats(tree, grid, mlp, fp, 0, n, 1, 0);
    (void)mx;
    (void)s2;
    auto mean = (s1 / (double)((int64_t)n * n * n)).to(torch::kFloat32);
    ps.coarse = torch::zeros({tree.NumNodes()}, mean.options()).index_copy(0, tree.active_node_ids, mean);
}

std::pair<Tensor, Tensor> ProposalFeatures(HyperTreeBase& tree, const ProposalState& ps, const RayList& rays)
{
    check_rays(rays);
    const int64_t R = rays.size();
    atdf_proposal p = proposal_view(tree, ps);
    auto f = torch::empty({R, ATDF_PROPOSAL_INPUTS}, rays.origin.options());
    auto ch = torch::empty({R, 2}, rays.origin.options());
    check_status(atdf_proposal_features(tree.packed(), tree.NumNodes(), rays.origin.data_ptr<float>(),
                                        rays.direction.data_ptr<float>(), R, &p, f.data_ptr<float>(),
                                        ch.data_ptr<float>(), current_stream()),
                 "atdf_proposal_features");
    return {f, ch};
}

std::vector<Tensor> ProposalSample(HyperTreeBase& tree, const ProposalState& ps, const RayList& rays,
                                   const SamplingParams& sp)
{
    check_rays(rays);
    TORCH_CHECK(sp.jitter_mode == 0 || sp.jitter_mode == 2, "ProposalSample: jitter_mode must be 0 or 2");
    const int64_t R = rays.size();
    atdf_proposal p = proposal_view(tree, ps, &sp);
    auto i32 = rays.origin.options().dtype(torch::kInt32);
    auto f32 = rays.origin.options();
    auto fb = torch::empty({R}, i32), nl = torch::empty({R}, i32);
    auto node = torch::empty({ATDF_PROPOSAL_BINS, R}, i32);
    auto lo = torch::empty({ATDF_PROPOSAL_BINS, R}, f32), hi = torch::empty({ATDF_PROPOSAL_BINS, R}, f32);
    auto kl = torch::empty({R}, f32);
    check_status(atdf_proposal_sample(tree.packed(), tree.NumNodes(), rays.origin.data_ptr<float>(),
                                      rays.direction.data_ptr<float>(), R, &p, fb.data_ptr<int32_t>(),
                                      nl.data_ptr<int32_t>(), node.data_ptr<int32_t>(), lo.data_ptr<float>(),
                                      hi.data_ptr<float>(), kl.data_ptr<float>(), current_stream()),
                 "atdf_proposal_sample");
    return {fb, nl, node, lo, hi, kl};
}

Intervals CreateProposalIntervals(HyperTreeBase& tree, const ProposalState& ps, const RayList& rays,
                                  const SamplingParams& fallback_sp, Tensor* ray_fallback)
{
    check_rays(rays);
    TORCH_CHECK(fallback_sp.jitter_mode != 1, "CreateProposalIntervals: jitter_mode 1 needs materialised samples");
    const int64_t R = rays.size();
    if (ps.pp.alloc_extra > 0) {
        // allocation mode: the stratified list of every ray, then the counts
        // along the proposal's CDF (the uniform one for the fallback rays)
        Intervals iv = tree.CreateIntervalsForRays(rays, fallback_sp);
        auto dev = tree.device();
        auto i32 = torch::TensorOptions().dtype(torch::kInt32).device(dev);
        auto fb = torch::empty({R}, i32), ns = torch::empty({R}, i32);
        if (R > 0) {
            atdf_proposal p = proposal_view(tree, ps, &fallback_sp);
            atdf_intervals w = iv.view();
            check_status(atdf_proposal_allocate(tree.packed(), tree.NumNodes(), rays.origin.data_ptr<float>(),
                                                rays.direction.data_ptr<float>(), R, &p, ps.pp.alloc_extra,
                                                iv.ray_offset.data_ptr<int64_t>(), &w, fb.data_ptr<int32_t>(),
                                                ns.data_ptr<int32_t>(), nullptr, current_stream()),
                         "atdf_proposal_allocate");
            auto info = torch::stack({ns.to(torch::kLong).sum(), ns.max().to(torch::kLong)}).to(torch::kCPU);
            iv.num_samples = info[0].item<int64_t>();
            iv.max_samples_per_ray = (int)info[1].item<int64_t>();
        }
        iv.ray_num_samples = ns;
        if (ray_fallback) *ray_fallback = fb;
        return iv;
    }
    auto v = ProposalSample(tree, ps, rays, fallback_sp);
    Tensor fb = v[0], nl = v[1];
    atdf_sampling s = fallback_sp.view();
    s.ray_select = fb.data_ptr<int32_t>();
    auto dev = tree.device();
    auto i32 = torch::TensorOptions().dtype(torch::kInt32).device(dev);
    auto i64 = torch::TensorOptions().dtype(torch::kLong).device(dev);
    auto f32 = torch::TensorOptions().dtype(torch::kFloat32).device(dev);
    auto st = current_stream();
    // the fallback rays (a few percent) traverse in dense waves
    auto list = torch::empty({R}, i32), list_count = torch::empty({1}, i32);
    check_status(atdf_compact_rays(fb.data_ptr<int32_t>(), R, list.data_ptr<int32_t>(), list_count.data_ptr<int32_t>(),
                                   st),
                 "atdf_compact_rays");
    s.ray_list = list.data_ptr<int32_t>();
    auto num_iv = torch::empty({R}, i32), num_s = torch::empty({R}, i32);
    auto status = torch::zeros({4}, i32);
    Intervals iv;
    iv.ray_offset = torch::empty({R + 1}, i64);
    auto smp_off = torch::empty({R + 1}, i64);
    if (R > 0) {
        check_status(atdf_traverse_count(tree.packed(), tree.NumNodes(), rays.origin.data_ptr<float>(),
                                         rays.direction.data_ptr<float>(), R, &s, num_iv.data_ptr<int32_t>(),
                                         num_s.data_ptr<int32_t>(), status.data_ptr<int32_t>(), st),
                     "atdf_traverse_count");
        check_status(atdf_proposal_merge_counts(R, fb.data_ptr<int32_t>(), nl.data_ptr<int32_t>(),
                                                num_iv.data_ptr<int32_t>(), num_s.data_ptr<int32_t>(), st),
                     "atdf_proposal_merge_counts");
    }
    auto ws = torch::empty({(int64_t)atdf_scan_workspace_bytes(R)}, torch::TensorOptions().dtype(torch::kUInt8).device(dev));
    check_status(atdf_exclusive_scan_i32(num_iv.data_ptr<int32_t>(), R, iv.ray_offset.data_ptr<int64_t>(),
                                         ws.data_ptr(), ws.numel(), st),
                 "atdf_exclusive_scan_i32");
    check_status(atdf_exclusive_scan_i32(num_s.data_ptr<int32_t>(), R, smp_off.data_ptr<int64_t>(), ws.data_ptr(),
                                         ws.numel(), st),
                 "atdf_exclusive_scan_i32");
    auto max_s = num_s.numel() > 0 ? num_s.max().to(torch::kLong).reshape({1}) : torch::zeros({1}, i64);
    auto info = torch::cat({status.to(torch::kLong), iv.ray_offset.slice(0, R, R + 1), smp_off.slice(0, R, R + 1),
                            max_s, fb.sum().to(torch::kLong).reshape({1})})
                    .to(torch::kCPU);
    const int64_t* ip = info.data_ptr<int64_t>();
    TORCH_CHECK(ip[0] == 0, "CreateProposalIntervals: ", ip[0], " fallback ray(s) hit more than ",
                s.max_intersections, " nodes");
    TORCH_CHECK(ip[1] == 0, "CreateProposalIntervals: traversal order check failed on ", ip[1], " ray(s)");
    iv.ray_num_samples = num_s;
    iv.num_intervals = ip[4];
    iv.num_samples = ip[5];
    iv.max_samples_per_ray = (int)ip[6];
    const int64_t H = iv.num_intervals;
    iv.node = torch::empty({H}, i32);
    iv.tmin = torch::empty({H}, f32);
    iv.tmax = torch::empty({H}, f32);
    iv.count = torch::empty({H}, i32);
    iv.first = torch::empty({H}, i32);
    iv.ray = torch::empty({H}, i32);
    if (H > 0) {
        atdf_intervals w = iv.view();
        if (ip[7] > 0)
            check_status(atdf_traverse_write(tree.packed(), tree.NumNodes(), rays.origin.data_ptr<float>(),
                                             rays.direction.data_ptr<float>(), R, &s, iv.ray_offset.data_ptr<int64_t>(),
                                             smp_off.data_ptr<int64_t>(), &w, nullptr, st),
                         "atdf_traverse_write");
        check_status(atdf_proposal_write(R, fb.data_ptr<int32_t>(), nl.data_ptr<int32_t>(), v[2].data_ptr<int32_t>(),
                                         v[3].data_ptr<float>(), v[4].data_ptr<float>(),
                                         iv.ray_offset.data_ptr<int64_t>(), &w, st),
                     "atdf_proposal_write");
    }
    if (ray_fallback) *ray_fallback = fb;
    return iv;
}

Tensor ProposalGrad(ProposalState& ps, Tensor features, Tensor target_bins)
{
    require_device_f32(features, "features");
    require_device_f32(target_bins, "target_bins");
    const int64_t R = features.size(0);
    TORCH_CHECK(features.dim() == 2 && features.size(1) == ATDF_PROPOSAL_INPUTS, "features must be [R, 38]");
    TORCH_CHECK(target_bins.dim() == 2 && target_bins.size(0) == R && target_bins.size(1) == ATDF_PROPOSAL_BINS,
                "target_bins must be [R, 32]");
    auto ws = torch::empty({(int64_t)atdf_proposal_grad_workspace_bytes(R)},
                           torch::TensorOptions().dtype(torch::kUInt8).device(features.device()));
    check_status(atdf_proposal_grad(ps.params.data_ptr<float>(), features.data_ptr<float>(),
                                    target_bins.data_ptr<float>(), R, ps.pp.target_power, ps.grad.data_ptr<float>(),
                                    ps.loss.data_ptr<double>(), ws.data_ptr(), ws.numel(), current_stream()),
                 "atdf_proposal_grad");
    return ps.loss;
}

Tensor BinVariation(const Tensor& bins)
{
    TORCH_CHECK(bins.dim() == 2 && bins.size(1) >= 2, "BinVariation: bins must be [R, B] with B >= 2");
    const int64_t B = bins.size(1);
    auto v = torch::empty_like(bins);
    using torch::indexing::Slice;
    v.index_put_({Slice(), Slice(1, B - 1)},
                 (bins.index({Slice(), Slice(2, B)}) - bins.index({Slice(), Slice(0, B - 2)})).abs() * 0.5f);
    v.index_put_({Slice(), 0}, (bins.index({Slice(), 1}) - bins.index({Slice(), 0})).abs());
    v.index_put_({Slice(), B - 1}, (bins.index({Slice(), B - 1}) - bins.index({Slice(), B - 2})).abs());
    return v;
}

double ProposalTrain(HyperTreeBase& tree, ProposalState& ps, const RayList& rays, const SamplingParams& sp,
                     Tensor grid, Tensor mlp, const FieldParams& fp, int steps)
{
    TORCH_CHECK(steps >= 1, "ProposalTrain: steps must be >= 1");
    ProposalRefreshCoarse(tree, ps, grid, mlp, fp);
    auto pb = ProjectRaysBins(tree, rays, grid, mlp, sp, fp);
    if (ps.pp.target_mode == 1) pb.second = BinVariation(pb.second);
    auto feat = ProposalFeatures(tree, ps, rays).first;
    const int64_t R = rays.size();
    auto ws = torch::empty({(int64_t)atdf_proposal_grad_workspace_bytes(R)},
                           torch::TensorOptions().dtype(torch::kUInt8).device(feat.device()));
    auto st = current_stream();
    for (int i = 0; i < steps; ++i) {
        check_status(atdf_proposal_grad(ps.params.data_ptr<float>(), feat.data_ptr<float>(),
                                        pb.second.data_ptr<float>(), R, ps.pp.target_power, ps.grad.data_ptr<float>(),
                                        ps.loss.data_ptr<double>(), ws.data_ptr(), ws.numel(), st),
                     "atdf_proposal_grad");
        ps.step += 1;
        check_status(atdf_adam_step(ps.params.data_ptr<float>(), ps.grad.data_ptr<float>(), ps.m.data_ptr<float>(),
                                    ps.v.data_ptr<float>(), ps.params.numel(), ps.lr, 0.9f, 0.999f, 1e-8f, ps.step, 1,
                                    st),
                     "atdf_adam_step");
    }
    ps.ready = true;
    return ps.loss.item<double>();
}

Tensor TrainStepLearned(HyperTreeBase& tree, TrainState& st, ProposalState& ps, const RayList& rays, Tensor target,
                        const SamplingParams& sp, const FieldParams& fp)
{
    if (!ps.ready) return TrainStep(tree, st, rays, target, sp, fp);
    Intervals iv = CreateProposalIntervals(tree, ps, rays, sp);
    return TrainStepIntervals(tree, st, rays, target, sp, fp, iv);
}

}  // namespace atdf_torch

// atdf_volume.cpp -- volume export: the field on a regular lattice
// (UniformPhantomSamplesGPU / UniformSampledVolume) and the Analyze 7.5
// volume files of SaveHDRImageTensor / LoadHDRImageTensor.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "atdf_torch.h"

namespace atdf_torch {

using torch::Tensor;

namespace {
atdf_lattice make_lattice(const std::vector<int>& size, int mode, bool swap_xy, const std::vector<float>& roi_min,
                          const std::vector<float>& step)
{
    TORCH_CHECK(size.size() == 3, "lattice size must have 3 entries");
    atdf_lattice l{};
    for (int k = 0; k < 3; ++k) {
        l.size[k] = size[k];
        l.roi_min[k] = roi_min.empty() ? 0.f : roi_min[k];
        l.step[k] = step.empty() ? 0.f : step[k];
    }
    l.mode = mode;
    l.swap_xy = swap_xy ? 1 : 0;
    return l;
}

int64_t lattice_points(const atdf_lattice& l) { return (int64_t)l.size[0] * l.size[1] * l.size[2]; }

SampleList lattice_samples(const HyperTreeBase& tree, const atdf_lattice& l)
{
    TORCH_CHECK(tree.node_position_min.is_cuda(), "the tree must be on the GPU");
    TORCH_CHECK(tree.NumActiveNodes() > 0, "no active node");
    const int64_t n = lattice_points(l);
    auto dev = tree.device();
    SampleList out;
    out.global_coordinate = torch::empty({n, 3}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    out.weight = torch::empty({n}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    out.node_id = torch::empty({n}, torch::TensorOptions().dtype(torch::kLong).device(dev));
    out.ray_index = torch::arange(n, torch::TensorOptions().dtype(torch::kLong).device(dev));
    check_status(atdf_sample_volume(tree.packed(), tree.NumNodes(), tree.active_node_ids.data_ptr<int64_t>(), nullptr,
                                    &l, nullptr, out.node_id.data_ptr<int64_t>(), out.weight.data_ptr<float>(),
                                    out.global_coordinate.data_ptr<float>(), current_stream()),
                 "atdf_sample_volume");
    return out;
}
}  // namespace

SampleList HyperTreeBase::UniformPhantomSamplesGPU(std::vector<int> size, bool swap_xy) const
{
    return lattice_samples(*this, make_lattice(size, 0, swap_xy, {}, {}));
}

SampleList HyperTreeBase::UniformPhantomSamplesGPUbySlice(std::vector<int> size, bool swap_xy,
                                                          std::vector<float> roi_min,
                                                          std::vector<float> step_size) const
{
    TORCH_CHECK(roi_min.size() == 3 && step_size.size() == 3, "roi_min and step_size need 3 entries");
    return lattice_samples(*this, make_lattice(size, 1, swap_xy, roi_min, step_size));
}

std::tuple<Tensor, Tensor, Tensor> UniformSampledVolume(HyperTreeBase& tree, Tensor grid, Tensor mlp,
                                                        const FieldParams& fp, std::vector<int> shape)
{
    atdf_lattice l = make_lattice(shape, 0, false, {}, {});
    atdf_field f = make_field(tree, grid, mlp, fp);
    const int64_t n = lattice_points(l);
    auto dev = tree.device();
    auto density = torch::empty({1, shape[0], shape[1], shape[2]},
                                torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    auto node = torch::empty({shape[0], shape[1], shape[2]}, torch::TensorOptions().dtype(torch::kLong).device(dev));
    auto valid = torch::empty({n}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
    check_status(atdf_sample_volume(tree.packed(), tree.NumNodes(), tree.active_node_ids.data_ptr<int64_t>(), &f, &l,
                                    density.data_ptr<float>(), node.data_ptr<int64_t>(), valid.data_ptr<float>(),
                                    nullptr, current_stream()),
                 "atdf_sample_volume");
    return {density, node, valid.to(torch::kLong).reshape({shape[0], shape[1], shape[2]})};
}

Tensor SampleVolumeBySlice(HyperTreeBase& tree, Tensor grid, Tensor mlp, const FieldParams& fp, std::vector<int> shape,
                           std::vector<float> roi_min, std::vector<float> step)
{
    TORCH_CHECK(roi_min.size() == 3 && step.size() == 3, "roi_min and step need 3 entries");
    atdf_lattice l = make_lattice(shape, 1, false, roi_min, step);
    atdf_field f = make_field(tree, grid, mlp, fp);
    auto density = torch::empty({1, shape[0], shape[1], shape[2]},
                                torch::TensorOptions().dtype(torch::kFloat32).device(tree.device()));
    check_status(atdf_sample_volume(tree.packed(), tree.NumNodes(), tree.active_node_ids.data_ptr<int64_t>(), &f, &l,
                                    density.data_ptr<float>(), nullptr, nullptr, nullptr, current_stream()),
                 "atdf_sample_volume");
    return density;
}

// ---------------------------------------------------------------------------
// Analyze 7.5: a 348-byte header (header_key 40 B, image_dimension 108 B,
// data_history 200 B) in "<name>.hdr" and the raw voxels in "<name>.img".
// ---------------------------------------------------------------------------
namespace {
constexpr int kHdrBytes = 348;
constexpr int kOffDim = 40;       // short dim[8]
constexpr int kOffDatatype = 70;  // short
constexpr int kOffBitpix = 72;    // short
constexpr int kOffPixdim = 76;    // float pixdim[8]
constexpr int kOffVoxOffset = 108;
constexpr int kOffFunused1 = 112; // scale factor (SPM convention)
constexpr int kDtFloat = 16;

std::string stem_of(const std::string& path)
{
    const size_t dot = path.rfind('.');
    const size_t slash = path.rfind('/');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) return path.substr(0, dot);
    return path;
}

template <typename T>
void put(unsigned char* h, int off, T v)
{
    std::memcpy(h + off, &v, sizeof(T));
}

template <typename T>
T get(const unsigned char* h, int off, bool swap)
{
    unsigned char b[sizeof(T)];
    std::memcpy(b, h + off, sizeof(T));
    if (swap)
        for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
}
}  // namespace

void SaveHDRImageTensor(Tensor volume, const std::string& path)
{
    TORCH_CHECK(volume.dim() == 4, "SaveHDRImageTensor: a 4-D volume is required, got ", volume.dim(), " dims");
    auto v = volume.detach().to(torch::kCPU).to(torch::kFloat32).contiguous();
    unsigned char h[kHdrBytes];
    std::memset(h, 0, sizeof(h));
    put<int32_t>(h, 0, kHdrBytes);
    put<int32_t>(h, 32, 16384);  // extents
    h[38] = 'r';                 // regular
    const int16_t dim[8] = {4, (int16_t)v.size(3), (int16_t)v.size(2), (int16_t)v.size(1), (int16_t)v.size(0), 1, 1, 1};
    for (int k = 0; k < 4; ++k) TORCH_CHECK(v.size(k) <= 32767, "SaveHDRImageTensor: dimension too large for Analyze");
    for (int k = 0; k < 8; ++k) put<int16_t>(h, kOffDim + 2 * k, dim[k]);
    put<int16_t>(h, kOffDatatype, kDtFloat);
    put<int16_t>(h, kOffBitpix, 32);
    for (int k = 0; k < 8; ++k) put<float>(h, kOffPixdim + 4 * k, k == 0 ? 0.f : 1.f);
    put<float>(h, kOffVoxOffset, 0.f);
    put<float>(h, kOffFunused1, 1.f);
    const std::string stem = stem_of(path);
    std::ofstream hf(stem + ".hdr", std::ios::binary);
    TORCH_CHECK(hf.good(), "SaveHDRImageTensor: cannot write ", stem, ".hdr");
    hf.write(reinterpret_cast<const char*>(h), kHdrBytes);
    std::ofstream imf(stem + ".img", std::ios::binary);
    TORCH_CHECK(imf.good(), "SaveHDRImageTensor: cannot write ", stem, ".img");
    imf.write(reinterpret_cast<const char*>(v.data_ptr<float>()), (std::streamsize)(v.numel() * sizeof(float)));
    TORCH_CHECK(hf.good() && imf.good(), "SaveHDRImageTensor: write failed for ", stem);
}

Tensor LoadHDRImageTensor(const std::string& path)
{
    const std::string stem = stem_of(path);
    std::ifstream hf(stem + ".hdr", std::ios::binary);
    TORCH_CHECK(hf.good(), "LoadHDRImageTensor: cannot open ", stem, ".hdr");
    unsigned char h[kHdrBytes];
    hf.read(reinterpret_cast<char*>(h), kHdrBytes);
    TORCH_CHECK(hf.gcount() == kHdrBytes, "LoadHDRImageTensor: short header in ", stem, ".hdr");
    bool swap = false;
    if (get<int32_t>(h, 0, false) != kHdrBytes) {
        swap = true;
        TORCH_CHECK(get<int32_t>(h, 0, true) == kHdrBytes, "LoadHDRImageTensor: not an Analyze header: ", stem);
    }
    int64_t dims[4] = {1, 1, 1, 1};
    const int nd = get<int16_t>(h, kOffDim, swap);
    TORCH_CHECK(nd >= 1 && nd <= 7, "LoadHDRImageTensor: bad dimension count ", nd);
    for (int k = 0; k < 4 && k < nd; ++k) dims[k] = std::max<int16_t>(1, get<int16_t>(h, kOffDim + 2 * (k + 1), swap));
    const int dt = get<int16_t>(h, kOffDatatype, swap);
    const int64_t offset = (int64_t)get<float>(h, kOffVoxOffset, swap);
    const int64_t w = dims[0], hh = dims[1], d = dims[2];
    const int64_t n = w * hh * d;  // the first frame
    int bytes = 0;
    switch (dt) {
        case 2: bytes = 1; break;    // unsigned char
        case 4: bytes = 2; break;    // signed short
        case 8: bytes = 4; break;    // signed int
        case 16: bytes = 4; break;   // float
        case 64: bytes = 8; break;   // double
        default: TORCH_CHECK(false, "LoadHDRImageTensor: unsupported Analyze datatype ", dt);
    }
    std::ifstream imf(stem + ".img", std::ios::binary);
    TORCH_CHECK(imf.good(), "LoadHDRImageTensor: cannot open ", stem, ".img");
    imf.seekg(offset);
    std::vector<unsigned char> raw((size_t)(n * bytes));
    imf.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)raw.size());
    TORCH_CHECK(imf.gcount() == (std::streamsize)raw.size(), "LoadHDRImageTensor: short image data in ", stem, ".img");
    auto out = torch::empty({1, d, hh, w}, torch::kFloat32);
    float* o = out.data_ptr<float>();
    for (int64_t i = 0; i < n; ++i) {
        const unsigned char* p = raw.data() + i * bytes;
        switch (dt) {
            case 2: o[i] = (float)p[0]; break;
            case 4: o[i] = (float)get<int16_t>(p, 0, swap); break;
            case 8: o[i] = (float)get<int32_t>(p, 0, swap); break;
            case 16: o[i] = get<float>(p, 0, swap); break;
            default: o[i] = (float)get<double>(p, 0, swap); break;
        }
    }
    return out;
}

}  // namespace atdf_torch

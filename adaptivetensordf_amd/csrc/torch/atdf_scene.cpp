// atdf_scene.cpp -- the scene directory (SceneBase, data/SceneBase.cpp) and the
// NPY / NPZ projection files, feeding the device-side ray and batch kernels.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "atdf_torch.h"

namespace atdf_torch {

using torch::Tensor;
namespace fs = std::filesystem;

// ---------------------------------------------------------------------------
// INI files (Saiga's SimpleIni layout: "[Section]" headers, "key = value",
// full-line comments starting with ';' or '#').
// ---------------------------------------------------------------------------
namespace {
std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// key -> value of `section` (keys before any section header count too)
std::map<std::string, std::string> read_ini(const std::string& file, const std::string& section)
{
    std::ifstream in(file);
    TORCH_CHECK(in.good(), "cannot open ", file);
    std::map<std::string, std::string> kv;
    std::string line, cur;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (line[0] == '[') {
            const size_t e = line.find(']');
            TORCH_CHECK(e != std::string::npos, file, ": bad section header '", line, "'");
            cur = trim(line.substr(1, e - 1));
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        if (!cur.empty() && cur != section) continue;
        kv[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return kv;
}

double to_double(const std::string& key, const std::string& v)
{
    try {
        size_t pos = 0;
        double d = std::stod(v, &pos);
        TORCH_CHECK(trim(v.substr(pos)).empty(), "trailing characters");
        return d;
    } catch (const std::exception&) {
        TORCH_CHECK(false, "ini key ", key, ": '", v, "' is not a number");
    }
    return 0;
}

bool to_bool(const std::string& key, const std::string& v)
{
    std::string l;
    for (char c : v) l += (char)std::tolower((unsigned char)c);
    if (l == "true" || l == "yes" || l == "on" || l == "1" || l == "t" || l == "y") return true;
    if (l == "false" || l == "no" || l == "off" || l == "0" || l == "f" || l == "n") return false;
    TORCH_CHECK(false, "ini key ", key, ": '", v, "' is not a boolean");
    return false;
}

std::vector<double> to_list(const std::string& key, const std::string& v)
{
    std::vector<double> out;
    std::stringstream ss(v);
    std::string tok;
    while (ss >> tok) out.push_back(to_double(key, tok));
    return out;
}
}  // namespace

void DatasetParams::Load(const std::string& file)
{
    auto kv = read_ini(file, "DatasetParams");
    auto get = [&](const char* k) -> const std::string* {
        auto it = kv.find(k);
        return it == kv.end() ? nullptr : &it->second;
    };
    if (auto v = get("camera_model")) camera_model = *v;
    if (auto v = get("image_dir")) image_dir = *v;
    if (auto v = get("mask_dir")) mask_dir = *v;
    if (auto v = get("volume_file")) volume_file = *v;
    if (auto v = get("projection_factor")) projection_factor = to_double("projection_factor", *v);
    if (auto v = get("vis_volume_intensity_factor"))
        vis_volume_intensity_factor = to_double("vis_volume_intensity_factor", *v);
    if (auto v = get("scene_scale")) scene_scale = to_double("scene_scale", *v);
    if (auto v = get("xray_min")) xray_min = to_double("xray_min", *v);
    if (auto v = get("xray_max")) xray_max = to_double("xray_max", *v);
    if (auto v = get("log_space_input")) log_space_input = to_bool("log_space_input", *v);
    if (auto v = get("use_log10_conversion")) use_log10_conversion = to_bool("use_log10_conversion", *v);
    for (auto [key, dst] : {std::make_pair("roi_min", &roi_min), std::make_pair("roi_max", &roi_max)}) {
        if (auto v = get(key)) {
            auto l = to_list(key, *v);
            TORCH_CHECK(l.size() == 3, file, ": ", key, " needs 3 values");
            for (int k = 0; k < 3; ++k) (*dst)[k] = (float)l[k];
        }
    }
    TORCH_CHECK(camera_model == "pinhole" || camera_model == "orthographic", file,
                ": camera_model must be pinhole or orthographic, got '", camera_model, "'");
}

void CameraBase::Load(const std::string& file)
{
    auto kv = read_ini(file, "CameraBase");
    TORCH_CHECK(kv.count("w") && kv.count("h") && kv.count("K"), file, ": w, h and K are required");
    w = (int)to_double("w", kv["w"]);
    h = (int)to_double("h", kv["h"]);
    K = to_list("K", kv["K"]);
    TORCH_CHECK(K.size() == 5, file, ": K needs 5 values (fx fy cx cy s), got ", K.size());
    TORCH_CHECK(w > 0 && h > 0, file, ": bad detector size ", w, " x ", h);
}

// ---------------------------------------------------------------------------
// NPY (format 1.0 - 3.0): magic, version, header length, a Python dict
// literal {'descr': ..., 'fortran_order': ..., 'shape': (...)}, raw data.
// ---------------------------------------------------------------------------
namespace {
Tensor parse_npy(const unsigned char* buf, size_t size, const std::string& what)
{
    static const unsigned char magic[6] = {0x93, 'N', 'U', 'M', 'P', 'Y'};
    TORCH_CHECK(size >= 10 && std::memcmp(buf, magic, 6) == 0, what, ": not an NPY file");
    const int major = buf[6];
    size_t hlen, hoff;
    if (major == 1) {
        hlen = (size_t)buf[8] | ((size_t)buf[9] << 8);
        hoff = 10;
    } else {
        TORCH_CHECK(major == 2 || major == 3, what, ": unsupported NPY version ", major);
        TORCH_CHECK(size >= 12, what, ": truncated NPY header");
        hlen = (size_t)buf[8] | ((size_t)buf[9] << 8) | ((size_t)buf[10] << 16) | ((size_t)buf[11] << 24);
        hoff = 12;
    }
    TORCH_CHECK(hoff + hlen <= size, what, ": truncated NPY header");
    const std::string header(reinterpret_cast<const char*>(buf + hoff), hlen);
    auto field = [&](const char* key) {
        const size_t k = header.find(std::string("'") + key + "'");
        TORCH_CHECK(k != std::string::npos, what, ": NPY header has no '", key, "'");
        const size_t c = header.find(':', k);
        TORCH_CHECK(c != std::string::npos, what, ": bad NPY header");
        return c + 1;
    };
    // descr: '<f4', '>f8', '|u1', ...
    size_t d = header.find('\'', field("descr"));
    TORCH_CHECK(d != std::string::npos, what, ": bad descr");
    const size_t d2 = header.find('\'', d + 1);
    const std::string descr = header.substr(d + 1, d2 - d - 1);
    TORCH_CHECK(descr.size() >= 3, what, ": unsupported dtype '", descr, "'");
    const char order = descr[0], kind = descr[1];
    const int bytes = std::atoi(descr.c_str() + 2);
    const bool swap = order == '>' && bytes > 1;
    TORCH_CHECK(order == '<' || order == '>' || order == '|' || order == '=', what, ": bad byte order in '", descr,
                "'");
    // fortran_order
    const size_t f = field("fortran_order");
    TORCH_CHECK(header.compare(header.find_first_not_of(' ', f), 5, "False") == 0, what,
                ": Fortran-order arrays are not supported");
    // shape
    const size_t s0 = header.find('(', field("shape")), s1 = header.find(')', s0);
    TORCH_CHECK(s0 != std::string::npos && s1 != std::string::npos, what, ": bad shape");
    std::vector<int64_t> shape;
    {
        std::string t = header.substr(s0 + 1, s1 - s0 - 1);
        std::replace(t.begin(), t.end(), ',', ' ');
        std::stringstream ss(t);
        int64_t v;
        while (ss >> v) shape.push_back(v);
    }
    TORCH_CHECK(bytes >= 1 && bytes <= 8, what, ": unsupported dtype '", descr, "'");
    const size_t data_off = hoff + hlen;
    // element count with overflow checks: every dimension >= 0 and the total
    // byte count within what the file holds
    const size_t avail = size - data_off;
    int64_t n = 1;
    for (auto v : shape) {
        TORCH_CHECK(v >= 0, what, ": negative dimension ", v, " in the NPY shape");
        TORCH_CHECK(v == 0 || n <= (int64_t)(avail / (size_t)bytes) / v, what, ": data shorter than the shape");
        n *= v;
    }
    TORCH_CHECK((size_t)n * (size_t)bytes <= avail, what, ": data shorter than the shape ", n, " x ", bytes, " B");
    const unsigned char* src = buf + data_off;
    c10::ScalarType dt;
    if (kind == 'f' && bytes == 4)
        dt = torch::kFloat32;
    else if (kind == 'f' && bytes == 8)
        dt = torch::kFloat64;
    else if (kind == 'b' && bytes == 1)
        dt = torch::kBool;
    else if (kind == 'u' && bytes == 1)
        dt = torch::kUInt8;
    else if (kind == 'u' && bytes == 2)
        dt = torch::kUInt16;
    else if (kind == 'u' && bytes == 4)
        dt = torch::kUInt32;
    else if (kind == 'i' && bytes == 1)
        dt = torch::kInt8;
    else if (kind == 'i' && bytes == 2)
        dt = torch::kInt16;
    else if (kind == 'i' && bytes == 4)
        dt = torch::kInt32;
    else if (kind == 'i' && bytes == 8)
        dt = torch::kInt64;
    else
        TORCH_CHECK(false, what, ": unsupported dtype '", descr, "'");
    Tensor t = torch::empty(shape, dt);
    auto* o = static_cast<unsigned char*>(t.data_ptr());
    if (!swap) {
        std::memcpy(o, src, (size_t)n * bytes);
    } else {
        for (int64_t i = 0; i < n; ++i)
            for (int b = 0; b < bytes; ++b) o[i * bytes + b] = src[i * bytes + bytes - 1 - b];
    }
    return t;
}

std::vector<unsigned char> read_file(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    TORCH_CHECK(in.good(), "cannot open ", path);
    in.seekg(0, std::ios::end);
    const std::streamoff len = in.tellg();
    in.seekg(0);
    std::vector<unsigned char> buf((size_t)len);
    in.read(reinterpret_cast<char*>(buf.data()), len);
    TORCH_CHECK(in.gcount() == len, "short read of ", path);
    return buf;
}

uint32_t rd16(const unsigned char* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
uint32_t rd32(const unsigned char* p) { return rd16(p) | (rd16(p + 2) << 16); }
uint64_t rd64(const unsigned char* p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }
}  // namespace

Tensor LoadNumpy(const std::string& path)
{
    auto buf = read_file(path);
    return parse_npy(buf.data(), buf.size(), path);
}

// ZIP archive (PKWARE APPNOTE): end-of-central-directory record (ZIP64 when
// its fields are saturated), central directory entries, local headers;
// members stored (method 0) or deflated (method 8, raw zlib inflate).
Tensor LoadNumpyNpz(const std::string& path, const std::string& varname)
{
    auto z = read_file(path);
    const size_t n = z.size();
    TORCH_CHECK(n >= 22, path, ": not a ZIP archive");
    size_t eocd = std::string::npos;
    for (size_t i = n - 22 + 1; i-- > 0 && n - i <= 22 + 65535;)
        if (rd32(&z[i]) == 0x06054b50u) {
            eocd = i;
            break;
        }
    TORCH_CHECK(eocd != std::string::npos, path, ": no ZIP end-of-central-directory record");
    uint64_t entries = rd16(&z[eocd + 10]);
    uint64_t cd_off = rd32(&z[eocd + 16]);
    if (cd_off == 0xffffffffu || entries == 0xffffu) {
        TORCH_CHECK(eocd >= 20 && rd32(&z[eocd - 20]) == 0x07064b50u, path, ": missing ZIP64 locator");
        const uint64_t z64 = rd64(&z[eocd - 20 + 8]);
        TORCH_CHECK(z64 < n && n - z64 >= 56 && rd32(&z[z64]) == 0x06064b50u, path, ": bad ZIP64 end record");
        entries = rd64(&z[z64 + 32]);
        cd_off = rd64(&z[z64 + 48]);
    }
    const std::string want = varname + ".npy";
    size_t p = (size_t)cd_off;
    TORCH_CHECK(cd_off <= n, path, ": central directory outside the archive");
    for (uint64_t e = 0; e < entries; ++e) {
        TORCH_CHECK(p <= n && n - p >= 46 && rd32(&z[p]) == 0x02014b50u, path, ": bad central directory entry");
        const uint32_t method = rd16(&z[p + 10]);
        uint64_t csize = rd32(&z[p + 20]), usize = rd32(&z[p + 24]);
        const uint32_t name_len = rd16(&z[p + 28]), extra_len = rd16(&z[p + 30]), comment_len = rd16(&z[p + 32]);
        uint64_t local = rd32(&z[p + 42]);
        TORCH_CHECK(n - p - 46 >= (size_t)name_len + extra_len + comment_len, path,
                    ": central directory entry runs past the archive");
        const std::string name(reinterpret_cast<const char*>(&z[p + 46]), name_len);
        // ZIP64 extra field: the saturated sizes / offset, in that order
        const size_t xend = p + 46 + name_len + extra_len;
        for (size_t x = p + 46 + name_len; x + 4 <= xend;) {
            const uint32_t id = rd16(&z[x]), len = rd16(&z[x + 2]);
            TORCH_CHECK(x + 4 + len <= xend, path, ": bad extra field in the entry of ", name);
            if (id == 0x0001) {
                size_t q = x + 4;
                const size_t qend = x + 4 + len;
                if (usize == 0xffffffffu) {
                    TORCH_CHECK(q + 8 <= qend, path, ": short ZIP64 extra field");
                    usize = rd64(&z[q]), q += 8;
                }
                if (csize == 0xffffffffu) {
                    TORCH_CHECK(q + 8 <= qend, path, ": short ZIP64 extra field");
                    csize = rd64(&z[q]), q += 8;
                }
                if (local == 0xffffffffu) {
                    TORCH_CHECK(q + 8 <= qend, path, ": short ZIP64 extra field");
                    local = rd64(&z[q]);
                }
            }
            x += 4 + len;
        }
        p += 46 + (size_t)name_len + extra_len + comment_len;
        if (name != want) continue;
        TORCH_CHECK(local < n && n - local >= 30 && rd32(&z[local]) == 0x04034b50u, path, ": bad local header of ",
                    name);
        const size_t data = (size_t)local + 30 + rd16(&z[local + 26]) + rd16(&z[local + 28]);
        TORCH_CHECK(data <= n && csize <= n - data, path, ": member ", name, " is truncated");
        TORCH_CHECK(usize <= ((uint64_t)1 << 40), path, ": member ", name, " is implausibly large");
        if (method == 0) return parse_npy(&z[data], (size_t)csize, path + ":" + name);
        TORCH_CHECK(method == 8, path, ": member ", name, " uses unsupported compression method ", method);
        std::vector<unsigned char> out((size_t)usize);
        z_stream s{};
        TORCH_CHECK(inflateInit2(&s, -MAX_WBITS) == Z_OK, "inflateInit2 failed");
        s.next_in = &z[data];
        s.avail_in = (uInt)csize;
        s.next_out = out.data();
        s.avail_out = (uInt)usize;
        const int rc = inflate(&s, Z_FINISH);
        inflateEnd(&s);
        TORCH_CHECK(rc == Z_STREAM_END && s.total_out == usize, path, ": inflate of ", name, " failed (", rc, ")");
        return parse_npy(out.data(), out.size(), path + ":" + name);
    }
    TORCH_CHECK(false, path, ": no array '", varname, "'");
    return {};
}

// ---------------------------------------------------------------------------
// SceneBase
// ---------------------------------------------------------------------------
namespace {
std::vector<std::string> read_lines(const std::string& file)
{
    std::vector<std::string> out;
    std::ifstream in(file);
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        out.push_back(line);
    }
    return out;
}
}  // namespace

SceneBase::SceneBase(const std::string& scene_dir)
{
    TORCH_CHECK(fs::exists(scene_dir), "scene directory ", scene_dir, " does not exist");
    scene_path = fs::canonical(scene_dir).string();
    scene_name = fs::path(scene_path).filename().string();
    TORCH_CHECK(fs::exists(scene_path + "/dataset.ini"), scene_path, "/dataset.ini is missing");
    dataset_params.Load(scene_path + "/dataset.ini");
    CameraBase cam;
    cam.Load(scene_path + "/camera.ini");
    cameras.push_back(cam);

    // poses.txt: one "qx qy qz qw tx ty tz" per image; the quaternion is
    // normalised (as the SE3 constructor does) and the translation scaled by
    // scene_scale (SceneBase.cpp:781-845).
    std::vector<std::vector<double>> poses;
    if (fs::exists(scene_path + "/poses.txt")) {
        for (auto& line : read_lines(scene_path + "/poses.txt")) {
            if (trim(line).empty()) continue;
            std::stringstream ss(line);
            std::vector<double> v(7);
            for (int k = 0; k < 7; ++k) TORCH_CHECK((bool)(ss >> v[k]), "poses.txt: bad line '", line, "'");
            poses.push_back(v);
        }
    }
    TORCH_CHECK(!poses.empty(), scene_path, "/poses.txt holds no pose");
    std::vector<std::string> images, masks;
    std::vector<int> cam_idx;
    if (fs::exists(scene_path + "/images.txt")) images = read_lines(scene_path + "/images.txt");
    if (fs::exists(scene_path + "/masks.txt")) masks = read_lines(scene_path + "/masks.txt");
    if (fs::exists(scene_path + "/camera_indices.txt"))
        for (auto& l : read_lines(scene_path + "/camera_indices.txt")) cam_idx.push_back(std::stoi(l));
    const size_t n_frames = std::max(images.size(), poses.size());
    TORCH_CHECK(poses.size() == n_frames, "poses.txt has ", poses.size(), " poses for ", n_frames, " images");
    TORCH_CHECK(masks.empty() || masks.size() == n_frames, "masks.txt has ", masks.size(), " lines for ", n_frames,
                " images");
    TORCH_CHECK(cam_idx.empty() || cam_idx.size() == n_frames, "camera_indices.txt has ", cam_idx.size(),
                " lines for ", n_frames, " images");
    frames.resize(n_frames);
    for (size_t i = 0; i < n_frames; ++i) {
        UnifiedImage& f = frames[i];
        f.camera_id = cam_idx.empty() ? 0 : cam_idx[i];
        TORCH_CHECK(f.camera_id >= 0 && f.camera_id < (int)cameras.size(), "image ", i, ": camera ", f.camera_id,
                    " does not exist");
        const auto& v = poses[i];
        const double norm = std::sqrt((v[0] * v[0] + v[2] * v[2]) + (v[1] * v[1] + v[3] * v[3]));
        TORCH_CHECK(norm > 1e-10, "image ", i, ": zero quaternion");
        for (int k = 0; k < 4; ++k) f.rotation[k] = v[k] / norm;
        for (int k = 0; k < 3; ++k) f.translation[k] = v[4 + k] * dataset_params.scene_scale;
        if (!images.empty()) f.image_file = images[i];
        if (!masks.empty()) f.mask_file = masks[i];
    }
}

std::vector<int> SceneBase::ReadIndexFile(const std::string& file)
{
    TORCH_CHECK(fs::exists(file), "index file ", file, " does not exist");
    std::vector<int> out;
    for (auto& l : read_lines(file)) {
        if (trim(l).empty()) continue;
        out.push_back(std::stoi(l));
    }
    return out;
}

void SceneBase::LoadSplit(const std::string& split_name)
{
    train_indices = ReadIndexFile(scene_path + "/" + split_name + "/train.txt");
    test_indices = ReadIndexFile(scene_path + "/" + split_name + "/eval.txt");
    for (auto* v : {&train_indices, &test_indices}) {
        std::sort(v->begin(), v->end());
        for (int i : *v) TORCH_CHECK(i >= 0 && i < NumImages(), "split ", split_name, ": image ", i, " out of range");
    }
}

int SceneBase::CameraModelId() const { return dataset_params.camera_model == "orthographic" ? 1 : 0; }

void SceneBase::Finalize(const std::string& dev, std::vector<float> vol_translation)
{
    TORCH_CHECK(vol_translation.size() == 3, "volume_translation needs 3 values");
    volume_translation = vol_translation;
    device = torch::Device(dev);
    const int64_t ni = NumImages(), nc = (int64_t)cameras.size();
    auto rot = torch::empty({ni, 4}, torch::kFloat64), tr = torch::empty({ni, 3}, torch::kFloat64);
    auto K = torch::empty({nc, 5}, torch::kFloat64);
    for (int64_t i = 0; i < ni; ++i) {
        for (int k = 0; k < 4; ++k) rot[i][k] = frames[i].rotation[k];
        for (int k = 0; k < 3; ++k) tr[i][k] = frames[i].translation[k] - (double)volume_translation[k];
    }
    for (int64_t c = 0; c < nc; ++c)
        for (int k = 0; k < 5; ++k) K[c][k] = cameras[c].K[k];
    rotation = rot.to(device);
    translation = tr.to(device);
    intrinsics = K.to(device);
}

atdf_camera_rig SceneBase::rig() const
{
    TORCH_CHECK(rotation.defined(), "SceneBase: Finalize() first");
    atdf_camera_rig r{};
    r.camera_model = CameraModelId();
    r.w = cameras.front().w;
    r.h = cameras.front().h;
    r.align_corners = align_corners ? 1 : 0;
    r.num_cameras = (int64_t)cameras.size();
    r.num_images = NumImages();
    r.intrinsics = intrinsics.data_ptr<double>();
    r.rotation = rotation.data_ptr<double>();
    r.translation = translation.data_ptr<double>();
    for (int k = 0; k < 3; ++k) r.volume_translation[k] = volume_translation[k];
    return r;
}

atdf_projection_set SceneBase::projection_set() const
{
    TORCH_CHECK(projections.defined(), "SceneBase: no projection loaded (LoadImagesCT)");
    atdf_projection_set s{};
    s.projection = projections.data_ptr<float>();
    s.mask = masks.defined() ? masks.data_ptr<float>() : nullptr;
    s.w = cameras.front().w;
    s.h = cameras.front().h;
    s.num_images = NumImages();
    s.align_corners = align_corners ? 1 : 0;
    return s;
}

namespace {
std::string resolve(const std::string& dir, const std::string& file, const std::string& scene_path)
{
    const std::string p = dir + "/" + file;
    if (fs::exists(p) || fs::path(dir).is_absolute()) return p;
    return scene_path + "/" + p;  // relative image_dir: also looked up in the scene directory
}
}  // namespace

void SceneBase::LoadImagesCT(std::vector<int> indices)
{
    TORCH_CHECK(device.is_cuda(), "SceneBase::LoadImagesCT: Finalize() on a GPU device first");
    const int w = cameras.front().w, h = cameras.front().h;
    const auto opt = torch::TensorOptions().dtype(torch::kFloat32).device(device);
    if (!projections.defined()) projections = torch::zeros({NumImages(), h, w}, opt);
    const bool use_masks = !dataset_params.mask_dir.empty();
    const auto& dp = dataset_params;
    for (int id : indices) {
        TORCH_CHECK(id >= 0 && id < NumImages(), "LoadImagesCT: image ", id, " out of range");
        UnifiedImage& f = frames[id];
        if (f.loaded) continue;
        TORCH_CHECK(!f.image_file.empty(), "LoadImagesCT: image ", id, " has no file in images.txt");
        const std::string file = resolve(dp.image_dir, f.image_file, scene_path);
        const std::string ext = file.substr(file.find_last_of('.') + 1);
        Tensor raw;
        if (ext == "npz")
            raw = LoadNumpyNpz(file, "proj");
        else if (ext == "npy")
            raw = LoadNumpy(file);
        else
            TORCH_CHECK(false, "LoadImagesCT: ", file, ": only .npy and .npz projections are read");
        raw = raw.to(torch::kFloat32).reshape({-1});
        TORCH_CHECK(raw.numel() == (int64_t)w * h, "LoadImagesCT: ", file, " holds ", raw.numel(),
                    " values for a ", w, " x ", h, " detector");
        Tensor mask;
        if (use_masks && !f.mask_file.empty()) {
            const std::string mfile = resolve(dp.mask_dir, f.mask_file, scene_path);
            TORCH_CHECK(fs::exists(mfile), "LoadImagesCT: mask ", mfile, " does not exist");
            const std::string mext = mfile.substr(mfile.find_last_of('.') + 1);
            TORCH_CHECK(mext == "npy", "LoadImagesCT: ", mfile, ": masks are read from .npy files");
            // 8-bit masks are normalised to [0, 1] like an 8-bit mask image
            mask = LoadNumpy(mfile);
            const bool u8 = mask.scalar_type() == torch::kUInt8;
            mask = mask.to(torch::kFloat32).reshape({-1});
            if (u8) mask = mask / 255.0f;
            TORCH_CHECK(mask.numel() == (int64_t)w * h, "LoadImagesCT: mask ", mfile, " has the wrong size");
            if (!masks.defined()) masks = torch::ones({NumImages(), h, w}, opt);
            masks[id].copy_(mask.reshape({h, w}));
            mask = mask.to(device);
        }
        raw = raw.to(device);
        Tensor out = projections[id];
        check_status(atdf_log_convert(raw.data_ptr<float>(), mask.defined() ? mask.data_ptr<float>() : nullptr,
                                      raw.numel(), dp.xray_min, dp.xray_max, dp.projection_factor,
                                      dp.use_log10_conversion ? 1 : 0, dp.log_space_input ? 1 : 0,
                                      out.data_ptr<float>(), current_stream()),
                     "atdf_log_convert");
        TORCH_CHECK(out.isfinite().all().item<bool>(), "LoadImagesCT: projection ", f.image_file,
                    " is not finite after the log conversion");
        f.loaded = true;
    }
}

namespace {
void check_ids(const Tensor& ids, int64_t bound, const char* what)
{
    TORCH_CHECK(ids.is_cuda() && ids.scalar_type() == torch::kInt32 && ids.dim() == 1, what,
                " must be a 1-D int32 device tensor");
    if (ids.numel() == 0) return;
    auto mm = torch::aminmax(ids);
    const int64_t lo = std::get<0>(mm).item<int64_t>(), hi = std::get<1>(mm).item<int64_t>();
    TORCH_CHECK(lo >= 0 && hi < bound, what, " out of range [0, ", bound, "): ", lo, " .. ", hi);
}
}  // namespace

Tensor SceneBase::PointInAnyImage(Tensor points) const
{
    TORCH_CHECK(points.is_cuda() && points.size(-1) == 3, "PointInAnyImage: [..., 3] device points required");
    auto sizes = points.sizes().vec();
    sizes.pop_back();
    auto p = points.reshape({-1, 3}).to(torch::kFloat32).contiguous();
    auto mask = torch::empty({p.size(0)}, p.options());
    atdf_camera_rig r = rig();
    const auto& dp = dataset_params;
    check_status(atdf_point_in_any_image(&r, dp.roi_min.data(), dp.roi_max.data(), p.data_ptr<float>(), p.size(0),
                                         mask.data_ptr<float>(), current_stream()),
                 "atdf_point_in_any_image");
    return mask.reshape(sizes);
}

RayList SceneBase::GetRays(Tensor uv, Tensor image_id, Tensor camera_id) const
{
    TORCH_CHECK(uv.is_cuda() && uv.scalar_type() == torch::kFloat32 && uv.dim() == 2 && uv.size(1) == 2,
                "GetRays: uv must be [n, 2] f32 on the device");
    uv = uv.contiguous();
    const int64_t n = uv.size(0);
    TORCH_CHECK(image_id.numel() == n && camera_id.numel() == n, "GetRays: one image and camera id per ray");
    image_id = image_id.contiguous();
    camera_id = camera_id.contiguous();
    check_ids(image_id, NumImages(), "GetRays: image_id");
    check_ids(camera_id, (int64_t)cameras.size(), "GetRays: camera_id");
    atdf_camera_rig r = rig();
    RayList out;
    out.origin = torch::empty({n, 3}, uv.options());
    out.direction = torch::empty({n, 3}, uv.options());
    check_status(atdf_generate_rays(&r, uv.data_ptr<float>(), image_id.data_ptr<int32_t>(),
                                    camera_id.data_ptr<int32_t>(), n, out.origin.data_ptr<float>(),
                                    out.direction.data_ptr<float>(), current_stream()),
                 "atdf_generate_rays");
    return out;
}

std::pair<Tensor, Tensor> SceneBase::SampleProjection(Tensor uv, Tensor image_id) const
{
    TORCH_CHECK(uv.is_cuda() && uv.scalar_type() == torch::kFloat32 && uv.dim() == 2 && uv.size(1) == 2,
                "SampleProjection: uv must be [n, 2] f32 on the device");
    uv = uv.contiguous();
    image_id = image_id.contiguous();
    const int64_t n = uv.size(0);
    TORCH_CHECK(image_id.numel() == n, "SampleProjection: one image id per point");
    check_ids(image_id, NumImages(), "SampleProjection: image_id");
    atdf_projection_set s = projection_set();
    Tensor t = torch::empty({n}, uv.options()), m = torch::empty({n}, uv.options());
    check_status(atdf_sample_projection(&s, uv.data_ptr<float>(), image_id.data_ptr<int32_t>(), n,
                                        t.data_ptr<float>(), m.data_ptr<float>(), current_stream()),
                 "atdf_sample_projection");
    return {t, m};
}

PixelList SceneBase::RandomRayBatch(Tensor indices, int64_t num_rays, uint64_t seed, bool interpolate) const
{
    TORCH_CHECK(indices.dim() == 1 && indices.numel() > 0, "RandomRayBatch: a non-empty image index list");
    auto idx_cpu = indices.to(torch::kCPU).to(torch::kInt32).contiguous();
    std::vector<int32_t> cams((size_t)idx_cpu.numel());
    for (int64_t k = 0; k < idx_cpu.numel(); ++k) {
        const int32_t id = idx_cpu.data_ptr<int32_t>()[k];
        TORCH_CHECK(id >= 0 && id < NumImages(), "RandomRayBatch: image ", id, " out of range");
        cams[(size_t)k] = frames[id].camera_id;
    }
    auto iopt = torch::TensorOptions().dtype(torch::kInt32).device(device);
    auto fopt = torch::TensorOptions().dtype(torch::kFloat32).device(device);
    Tensor idx = idx_cpu.to(device);
    Tensor cam = torch::from_blob(cams.data(), {(int64_t)cams.size()}, torch::kInt32).to(device);
    atdf_camera_rig r = rig();
    atdf_projection_set s{};
    const bool with_targets = projections.defined();
    if (with_targets) s = projection_set();
    PixelList out;
    out.uv = torch::empty({num_rays, 2}, fopt);
    out.image_id = torch::empty({num_rays}, iopt);
    out.camera_id = torch::empty({num_rays}, iopt);
    if (with_targets) {
        out.target = torch::empty({num_rays}, fopt);
        out.target_mask = torch::empty({num_rays}, fopt);
    }
    out.rays.origin = torch::empty({num_rays, 3}, fopt);
    out.rays.direction = torch::empty({num_rays, 3}, fopt);
    check_status(atdf_random_ray_batch(&r, with_targets ? &s : nullptr, idx.data_ptr<int32_t>(), idx.numel(),
                                       cam.data_ptr<int32_t>(), r.w, r.h, r.align_corners, interpolate ? 1 : 0, seed,
                                       num_rays, out.uv.data_ptr<float>(), out.image_id.data_ptr<int32_t>(),
                                       out.camera_id.data_ptr<int32_t>(),
                                       with_targets ? out.target.data_ptr<float>() : nullptr,
                                       with_targets ? out.target_mask.data_ptr<float>() : nullptr,
                                       out.rays.origin.data_ptr<float>(), out.rays.direction.data_ptr<float>(),
                                       current_stream()),
                 "atdf_random_ray_batch");
    return out;
}

}  // namespace atdf_torch

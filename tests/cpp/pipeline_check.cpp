// pipeline_check.cpp -- renders cases through the C++ mirror (include/euchip.hpp)
// the way a euc user would, for tests/test_gpu_cpp.py to compare with the oracle.
//
// usage: pipeline_check CASE_DIR...
// Each CASE_DIR holds case.txt ("key value" lines) and raw little-endian input
// arrays; the program writes pixel.u32 / depth.f32 there, or panic.txt with the
// message of the reference panic the render raised.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

#include "euchip.hpp"

namespace {

std::map<std::string, std::string> read_case(const std::string &dir) {
    std::map<std::string, std::string> kv;
    std::ifstream f(dir + "/case.txt");
    std::string k, v;
    while (f >> k >> v) kv[k] = v;
    return kv;
}

template <class T>
std::vector<T> read_raw(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return {};
    const std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<T> v((size_t)n / sizeof(T));
    f.read(reinterpret_cast<char *>(v.data()), n);
    return v;
}

template <class T>
void write_raw(const std::string &path, const std::vector<T> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

euc::DepthMode depth_mode_of(const std::string &s) {
    if (s == "less_write") return euc::DepthMode::less_write();
    if (s == "greater_write") return euc::DepthMode::greater_write();
    if (s == "less_pass") return euc::DepthMode::less_pass();
    if (s == "greater_pass") return euc::DepthMode::greater_pass();
    return euc::DepthMode::none();
}

euc::CoordinateMode coords_of(const std::string &s) {
    if (s == "opengl") return euc::CoordinateMode::opengl();
    if (s == "metal") return euc::CoordinateMode::metal();
    if (s == "directx") return euc::CoordinateMode::directx();
    if (s == "vulkan_noclip") return euc::CoordinateMode::vulkan().without_z_clip();
    return euc::CoordinateMode::vulkan();
}

euc::CullMode cull_of(const std::string &s) {
    if (s == "none") return euc::CullMode::None;
    if (s == "front") return euc::CullMode::Front;
    return euc::CullMode::Back;
}

// One `impl Pipeline` configured from the case file.
class CasePipeline : public euc::Pipeline {
public:
    explicit CasePipeline(const std::map<std::string, std::string> &kv, const euc::Uniforms &u) : kv_(kv), u_(u) {}
    euc::Shader shader() const override { return (euc::Shader)std::stoul(kv_.at("shader")); }
    euc::Uniforms uniforms() const override { return u_; }
    bool clip_input() const override { return kv_.at("clip") == "1"; }
    euc::DepthMode depth_mode() const override { return depth_mode_of(kv_.at("depth_mode")); }
    euc::PixelMode pixel_mode() const override {
        return kv_.at("pixel_write") == "1" ? euc::PixelMode::write_mode() : euc::PixelMode::pass();
    }
    euc::CoordinateMode coordinate_mode() const override { return coords_of(kv_.at("coords")); }
    euc::CullMode rasterizer_config() const override { return cull_of(kv_.at("cull")); }
    euc::AaMode aa_mode() const override {
        const int l = std::stoi(kv_.at("msaa"));
        return l < 0 ? euc::AaMode::none() : euc::AaMode::msaa_level(l);
    }
    euc::BandPolicy band_policy() const override {
        return kv_.at("bands") == "seq" ? euc::BandPolicy::Seq : euc::BandPolicy::Par;
    }
    euchip_pixel_format pixel_format() const override { return (euchip_pixel_format)opt("pixel_format"); }
    euchip_geometry geometry() const override { return (euchip_geometry)opt("geometry"); }
    euchip_primitives primitives() const override { return (euchip_primitives)opt("prims"); }

private:
    unsigned long opt(const char *key) const {
        const auto it = kv_.find(key);
        return it == kv_.end() ? 0ul : std::stoul(it->second);
    }
    std::map<std::string, std::string> kv_;
    euc::Uniforms u_;
};

int run_case(euc::Context &ctx, const std::string &dir) {
    const auto kv = read_case(dir);
    euc::MeshData md;
    md.pos = read_raw<float>(dir + "/pos.f32");
    md.pos_comps = (uint32_t)std::stoul(kv.at("pos_comps"));
    md.col = read_raw<float>(dir + "/col.f32");
    md.rgba8 = read_raw<uint32_t>(dir + "/rgba8.u32");
    md.nrm = read_raw<float>(dir + "/nrm.f32");
    md.indices = read_raw<uint32_t>(dir + "/idx.u32");
    md.models = read_raw<float>(dir + "/models.f32");
    const std::vector<float> uv = read_raw<float>(dir + "/uniforms.f32");
    euc::Uniforms u{};
    for (size_t i = 0; i < u.size() && i < uv.size(); i++) u[i] = uv[i];
    euc::Mesh mesh(ctx, md);
    const uint32_t pw = (uint32_t)std::stoul(kv.at("pw")), ph = (uint32_t)std::stoul(kv.at("ph"));
    const uint32_t dw = (uint32_t)std::stoul(kv.at("dw")), dh = (uint32_t)std::stoul(kv.at("dh"));
    const uint32_t clear_color = (uint32_t)std::stoul(kv.at("clear_color"), nullptr, 0);
    const float clear_depth = std::stof(kv.at("clear_depth"));
    std::unique_ptr<euc::Buffer2d<uint32_t>> pixel;
    std::unique_ptr<euc::Buffer2d<float>> depth;
    if (kv.at("pixel") == "1")
        pixel.reset(new euc::Buffer2d<uint32_t>(euc::Buffer2d<uint32_t>::fill(ctx, {pw, ph}, clear_color)));
    if (kv.at("depth") == "1")
        depth.reset(new euc::Buffer2d<float>(euc::Buffer2d<float>::fill(ctx, {dw, dh}, clear_depth)));
    CasePipeline pipe(kv, u);
    try {
        pipe.render(mesh, pixel.get(), depth.get());
    } catch (const euc::Panic &p) {
        std::ofstream(dir + "/panic.txt") << p.what();
    }
    if (pixel) write_raw(dir + "/pixel.u32", pixel->raw());
    if (depth) write_raw(dir + "/depth.f32", depth->raw());
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        euc::Context ctx(0);
        for (int i = 1; i < argc; i++) {
            std::cout << "case " << argv[i] << std::endl;
            run_case(ctx, argv[i]);
        }
    } catch (const std::exception &e) {
        std::cerr << "pipeline_check: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

// euchip.hpp -- C++17 host mirror of euc 0.6.0's rendering surface over the C ABI
// (euchip.h).  Header-only; link against libeuchip.so.
//
// Names and meaning follow the reference:
//   DepthMode, PixelMode        src/pipeline.rs:11-73
//   CoordinateMode              src/pipeline.rs:93-163 (OPENGL / VULKAN / METAL / DIRECTX)
//   AaMode                      src/pipeline.rs:102-113
//   CullMode                    src/rasterizer/mod.rs:9-18
//   Buffer2d, Target::clear     src/buffer.rs:19-22, 213-218; Empty src/texture.rs:290-319
//   Pipeline::render            src/pipeline.rs:248-300
// Shaders are the fixed device set (Shader::*) with a 64-float uniform block.
// Where the reference panics, render() throws euc::Panic with the reference's
// message; other failures throw euc::Error with the library's message.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "euchip.h"

namespace euc {

enum class Ordering { Less, Equal, Greater };

struct DepthMode {
    std::optional<Ordering> test;
    bool write = false;
    bool uses_depth() const { return test.has_value() || write; }
    static DepthMode none() { return {std::nullopt, false}; }
    static DepthMode less_write() { return {Ordering::Less, true}; }
    static DepthMode greater_write() { return {Ordering::Greater, true}; }
    static DepthMode less_pass() { return {Ordering::Less, false}; }
    static DepthMode greater_pass() { return {Ordering::Greater, false}; }
};

struct PixelMode {
    bool write = true;
    static PixelMode write_mode() { return {true}; }
    static PixelMode pass() { return {false}; }
};

enum class Handedness { Left, Right };
enum class YAxisDirection { Down, Up };

struct CoordinateMode {
    Handedness handedness = Handedness::Left;
    YAxisDirection y_axis_direction = YAxisDirection::Down;
    std::optional<std::pair<float, float>> z_clip_range = std::make_pair(0.0f, 1.0f);
    static CoordinateMode opengl() { return {Handedness::Right, YAxisDirection::Up, std::make_pair(-1.0f, 1.0f)}; }
    static CoordinateMode vulkan() { return {Handedness::Left, YAxisDirection::Down, std::make_pair(0.0f, 1.0f)}; }
    static CoordinateMode metal() { return {Handedness::Right, YAxisDirection::Down, std::make_pair(0.0f, 1.0f)}; }
    static CoordinateMode directx() { return {Handedness::Left, YAxisDirection::Up, std::make_pair(0.0f, 1.0f)}; }
    CoordinateMode without_z_clip() const {
        CoordinateMode m = *this;
        m.z_clip_range.reset();
        return m;
    }
};

struct AaMode {
    bool msaa = false;
    int32_t level = 0;
    static AaMode none() { return {false, 0}; }
    static AaMode msaa_level(int32_t level) { return {true, level}; }
};

enum class CullMode { None, Back, Front };

enum class Shader : uint32_t {
    Flat = EUCHIP_SHADER_FLAT, Color = EUCHIP_SHADER_COLOR, Gouraud = EUCHIP_SHADER_GOURAUD,
    Phong = EUCHIP_SHADER_PHONG, DepthOnly = EUCHIP_SHADER_DEPTH_ONLY, Textured = EUCHIP_SHADER_TEXTURED,
    PhongShadow = EUCHIP_SHADER_PHONG_SHADOW, TexShadow = EUCHIP_SHADER_TEX_SHADOW,
};

enum class BandPolicy { Par = EUCHIP_BANDS_PAR, Seq = EUCHIP_BANDS_SEQ };

using Uniforms = std::array<float, EUCHIP_UNIFORM_FLOATS>;

// A reference panic (assertion or arithmetic), with its message.
struct Panic : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Error : std::runtime_error {
    euchip_status status;
    Error(euchip_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

class Context {
public:
    explicit Context(int device = 0) {
        const euchip_status s = euchip_ctx_create(device, &ctx_);
        if (s) throw Error(s, std::string("euchip_ctx_create: ") + euchip_status_string(s));
    }
    ~Context() {
        if (ctx_) euchip_ctx_destroy(ctx_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    euchip_ctx *raw() const { return ctx_; }
    void check(euchip_status s, const char *what) const {
        if (s) throw Error(s, std::string(what) + ": " + euchip_last_error(ctx_));
    }

private:
    euchip_ctx *ctx_ = nullptr;
};

// Device Buffer2d<T>: T = uint32_t (RGBA8 packed as in [u8; 4]) or float.
template <class T>
class Buffer2d {
    static_assert(std::is_same<T, uint32_t>::value || std::is_same<T, float>::value, "RGBA8 (uint32_t) or f32");

public:
    // Buffer2d::fill (src/buffer.rs:62-70)
    static Buffer2d fill(Context &ctx, std::array<uint32_t, 2> size, T value) {
        Buffer2d b(ctx, size);
        b.clear(value);
        return b;
    }
    Buffer2d(Context &ctx, std::array<uint32_t, 2> size) : ctx_(&ctx), size_(size) {
        const euchip_format f = std::is_same<T, float>::value ? EUCHIP_FORMAT_F32 : EUCHIP_FORMAT_RGBA8;
        ctx.check(euchip_buffer_create(ctx.raw(), f, size[0], size[1], &buf_), "euchip_buffer_create");
    }
    ~Buffer2d() {
        if (buf_) euchip_buffer_destroy(buf_);
    }
    Buffer2d(Buffer2d &&o) noexcept : ctx_(o.ctx_), size_(o.size_), buf_(o.buf_) { o.buf_ = nullptr; }
    Buffer2d(const Buffer2d &) = delete;
    Buffer2d &operator=(const Buffer2d &) = delete;

    std::array<uint32_t, 2> size() const { return size_; }
    // Target::clear (src/buffer.rs:213-218)
    void clear(T value) { ctx_->check(euchip_buffer_clear(buf_, &value), "euchip_buffer_clear"); }
    // Texels in row-major order, x + W * y (src/buffer.rs:145-150).
    std::vector<T> raw() const {
        std::vector<T> v((size_t)size_[0] * size_[1]);
        ctx_->check(euchip_buffer_download(buf_, v.data(), v.size() * sizeof(T)), "euchip_buffer_download");
        return v;
    }
    void write(const std::vector<T> &v) {
        ctx_->check(euchip_buffer_upload(buf_, v.data(), v.size() * sizeof(T)), "euchip_buffer_upload");
    }
    euchip_buffer *handle() const { return buf_; }

private:
    Context *ctx_;
    std::array<uint32_t, 2> size_;
    euchip_buffer *buf_ = nullptr;
};

// A resident vertex stream: TriangleList over an optional IndexedVertices index
// buffer (src/primitives.rs:23-44, src/index.rs:20-55), with per-instance model
// matrices (row-major, column vectors).
struct MeshData {
    std::vector<float> pos;       // 3 (object space) or 4 (clip space) floats per vertex
    uint32_t pos_comps = 3;
    std::vector<float> nrm, col, uv;
    std::vector<uint32_t> rgba8;
    std::vector<uint32_t> indices;  // empty: the vertices in order
    std::vector<float> models;      // 16 per instance; empty: one identity instance
};

class Mesh {
public:
    Mesh(Context &ctx, const MeshData &d) : ctx_(&ctx) {
        euchip_mesh_desc m{};
        m.pos = d.pos.data();
        m.pos_comps = d.pos_comps;
        m.nrm = d.nrm.empty() ? nullptr : d.nrm.data();
        m.col = d.col.empty() ? nullptr : d.col.data();
        m.uv = d.uv.empty() ? nullptr : d.uv.data();
        m.rgba8 = d.rgba8.empty() ? nullptr : d.rgba8.data();
        m.n_vertices = d.pos.size() / d.pos_comps;
        m.idx = d.indices.empty() ? nullptr : d.indices.data();
        m.count = d.indices.empty() ? m.n_vertices : d.indices.size();
        m.n_instances = d.models.empty() ? 1u : (uint32_t)(d.models.size() / 16);
        m.models = d.models.empty() ? nullptr : d.models.data();
        ctx.check(euchip_mesh_create(ctx.raw(), &m, &mesh_), "euchip_mesh_create");
    }
    ~Mesh() {
        if (mesh_) euchip_mesh_destroy(mesh_);
    }
    Mesh(const Mesh &) = delete;
    Mesh &operator=(const Mesh &) = delete;
    euchip_mesh *handle() const { return mesh_; }
    Context &context() const { return *ctx_; }

private:
    Context *ctx_;
    euchip_mesh *mesh_ = nullptr;
};

// The reference's `Pipeline` trait with a fixed device shader.
class Pipeline {
public:
    virtual ~Pipeline() = default;
    virtual Shader shader() const = 0;
    virtual Uniforms uniforms() const = 0;
    // vertex positions are transformed by the uniforms' view-projection (MVP)
    // or taken as clip-space positions (clip_input)
    virtual bool clip_input() const { return false; }
    virtual PixelMode pixel_mode() const { return PixelMode::write_mode(); }
    virtual DepthMode depth_mode() const { return DepthMode::none(); }
    virtual CoordinateMode coordinate_mode() const { return CoordinateMode::vulkan(); }
    virtual AaMode aa_mode() const { return AaMode::none(); }
    virtual CullMode rasterizer_config() const { return CullMode::Back; }
    virtual BandPolicy band_policy() const { return BandPolicy::Par; }
    virtual uint32_t threads() const { return 8; }
    virtual const Buffer2d<uint32_t> *texture() const { return nullptr; }
    virtual const Buffer2d<float> *shadow() const { return nullptr; }
    // the blend's u32 layout, the texture sampler, the fixed geometry stage and
    // the PrimitiveKind (TriangleList by default, as every reference example but
    // wireframes.rs)
    virtual euchip_pixel_format pixel_format() const { return EUCHIP_PIXEL_RGBA8; }
    virtual euchip_sampler sampler() const { return EUCHIP_SAMPLER_DEFAULT; }
    virtual euchip_geometry geometry() const { return EUCHIP_GEOMETRY_IDENTITY; }
    virtual euchip_primitives primitives() const { return EUCHIP_PRIMS_TRIANGLE_LIST; }

    euchip_pipeline_desc desc() const {
        euchip_pipeline_desc d{};
        d.shader = (euchip_shader)shader();
        d.xform = clip_input() ? EUCHIP_XFORM_CLIP : EUCHIP_XFORM_MVP;
        d.pixel_write = pixel_mode().write;
        const DepthMode dm = depth_mode();
        d.depth_test = !dm.test ? EUCHIP_DEPTH_NONE
                     : *dm.test == Ordering::Less ? EUCHIP_DEPTH_LESS
                     : *dm.test == Ordering::Equal ? EUCHIP_DEPTH_EQUAL : EUCHIP_DEPTH_GREATER;
        d.depth_write = dm.write;
        const CoordinateMode cm = coordinate_mode();
        d.y_up = cm.y_axis_direction == YAxisDirection::Up;
        d.z_clip = cm.z_clip_range.has_value();
        if (cm.z_clip_range) d.z_min = cm.z_clip_range->first, d.z_max = cm.z_clip_range->second;
        const CullMode c = rasterizer_config();
        d.cull = c == CullMode::None ? EUCHIP_CULL_NONE : c == CullMode::Back ? EUCHIP_CULL_BACK : EUCHIP_CULL_FRONT;
        const AaMode aa = aa_mode();
        d.aa_msaa = aa.msaa;
        d.aa_level = aa.level < 0 ? 0u : (uint32_t)aa.level;
        d.band_policy = (euchip_band_policy)band_policy();
        d.threads = threads();
        d.texture = texture() ? texture()->handle() : nullptr;
        d.shadow = shadow() ? shadow()->handle() : nullptr;
        d.pixel_format = pixel_format();
        d.sampler = sampler();
        d.geometry = geometry();
        d.primitives = primitives();
        return d;
    }

    // Pipeline::render (src/pipeline.rs:248-300).  nullptr plays Empty.
    void render(const Mesh &vertices, Buffer2d<uint32_t> *pixel, Buffer2d<float> *depth) const {
        const euchip_pipeline_desc d = desc();
        const Uniforms u = uniforms();
        const euchip_status s = euchip_render(vertices.context().raw(), &d, u.data(), vertices.handle(),
                                              pixel ? pixel->handle() : nullptr, depth ? depth->handle() : nullptr);
        if (s == EUCHIP_SIZE_MISMATCH)  // src/pipeline.rs:262-266
            throw Panic("assertion `left == right` failed: Pixel target size is compatible with depth target size");
        if (s == EUCHIP_BAND_ZERO)  // src/pipeline.rs:328-330 (G = 0)
            throw Panic("attempt to divide by zero");
        vertices.context().check(s, "euchip_render");
    }
};

}  // namespace euc

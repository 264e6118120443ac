// C++ host components of the MI355X path-tracing hot path, over the lmgpu C ABI.
//
// Mirrors the reference's component interfaces for this path so that a
// caller written against them switches by changing the component:
//   lm::Accel (build / intersect / Hit)        include/lm/accel.h:26-71
//   lm::Renderer::render -> {processed, elapsed} include/lm/renderer.h:20-26,
//                                               src/renderer/renderer_pt.cpp:51-77,314
//   renderer::raycast                          src/renderer/renderer_raycast.cpp:19-101
//   lm::Film (film::bitmap)                    include/lm/film.h:54-143, src/film/film_bitmap.cpp
//   Scene::traverse_primitive_nodes            src/scene.cpp:216-230
//   lm::Transform                              include/lm/math.h:724-746
//   mesh::raw triangles                        src/mesh/mesh_raw.cpp:60-97
//   LM_THROW_EXCEPTION(Error::...)             include/lm/exception.h:142-150
// The scene description is the subset of components on the path: raw
// meshes, material::{diffuse, glossy, mirror, glass}, light::{area,
// envconst}, camera::pinhole, in a scene graph of groups with optional
// local transforms. Vectors are FP64 as the reference's Float; device work is
// FP32 (FP64 film accumulation).
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

struct lmgpu_accel;
struct lmgpu_scene;
struct lmgpu_guide;
struct lmgpu_ray;
struct lmgpu_hit;

namespace lm_mi355x {

using Float = double;

struct Vec2 {
    Float x = 0, y = 0;
};

struct Vec3 {
    Float x = 0, y = 0, z = 0;
    Float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator-(Vec3 a) { return {-a.x, -a.y, -a.z}; }
inline Vec3 operator*(Vec3 a, Float s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator*(Float s, Vec3 a) { return a * s; }
inline Vec3 operator*(Vec3 a, Vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline Vec3 operator/(Vec3 a, Float s) { return {a.x / s, a.y / s, a.z / s}; }
Float dot(Vec3 a, Vec3 b);
Vec3 cross(Vec3 a, Vec3 b);
Float length(Vec3 a);
Vec3 normalize(Vec3 a);

// 4x4 matrix acting on column vectors, m[row][col].
struct Mat4 {
    Float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    static Mat4 identity() { return Mat4{}; }
    static Mat4 translate(Vec3 t);
    static Mat4 scale(Vec3 s);
    static Mat4 rotate(Float angle_rad, Vec3 axis);
};
Mat4 operator*(const Mat4& a, const Mat4& b);

struct Mat3 {
    Float m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
};

// lm::Transform: M, the normal matrix transpose(inverse(M)) and |det M_lin|.
struct Transform {
    Mat4 M;
    Mat3 normal_M;
    Float J = 1;
    Transform() = default;
    explicit Transform(const Mat4& M);
    Vec3 point(Vec3 p) const;
    Vec3 normal(Vec3 n) const;   // normalised
};

struct Ray {
    Vec3 o, d;
};

// ---- errors (lm::Error, include/lm/exception.h:28-37) ----------------------
enum class Error { None, Unknown, Uninitialized, InvalidArgument, Unsupported, IOError, FailedToRender };

class Exception : public std::runtime_error {
public:
    Exception(Error e, const std::string& msg) : std::runtime_error(msg), err_(e) {}
    Error error() const { return err_; }

private:
    Error err_;
};

// ---- scene description -------------------------------------------------------

// mesh::raw: positions, normals and per-corner indices (n < 0: no vertex
// normal, the geometric normal is used as in Mesh_Raw::surface_point).
struct Mesh {
    struct Index {
        int p = 0, n = -1;
    };
    struct Point {
        Vec3 p, n;
    };
    struct Tri {
        Point p1, p2, p3;
    };
    std::vector<Vec3> ps, ns;
    std::vector<std::array<Index, 3>> fs;

    int num_triangles() const { return int(fs.size()); }
    Tri triangle_at(int face) const;
    void foreach_triangle(const std::function<void(int face, const Tri& tri)>& f) const;
};

enum class MaterialType { Diffuse = 0, Glossy = 1, Mirror = 2, Glass = 3 };

struct Material {
    MaterialType type = MaterialType::Diffuse;
    Vec3 color{1, 1, 1};   // Kd (diffuse) or Ks (glossy)
    Float ax = 0, ay = 0;  // glossy roughness
    Float Ni = 1;          // glass index of refraction
    static Material diffuse(Vec3 Kd) { return {MaterialType::Diffuse, Kd, 0, 0, 1}; }
    static Material glossy(Vec3 Ks, Float ax, Float ay) { return {MaterialType::Glossy, Ks, ax, ay, 1}; }
    static Material mirror() { return {MaterialType::Mirror, {}, 0, 0, 1}; }
    static Material glass(Float Ni) { return {MaterialType::Glass, {}, 0, 0, Ni}; }
    // Material::reflectance (Kd, Ks, 0 for mirror and glass)
    Vec3 reflectance() const;
};

struct AreaLight {   // light::area (one-sided emitter of radiance Ke)
    Vec3 Ke;
};

struct Camera {      // camera::pinhole (src/camera/camera_pinhole.cpp:86-132)
    Vec3 position{0, 0, 1}, center{0, 0, 0}, up{0, 1, 0};
    Float vfov = 40;     // degrees
    Float aspect = 1;    // width / height
    Ray primary_ray(Vec2 rp) const;
};

enum class SceneNodeType { Primitive, Group };

struct SceneNode {
    SceneNodeType type = SceneNodeType::Group;
    int index = 0;
    struct Primitive {
        std::shared_ptr<const Mesh> mesh;
        Material material;
        std::optional<AreaLight> light;
    } primitive;
    struct Group {
        std::vector<int> children;
        std::optional<Mat4> local_transform;
    } group;
};

// Flattened world-space triangles in traversal order (global triangle i is
// the i-th triangle met by traverse_primitive_nodes, as Accel_SAHBVH::build,
// src/accel/accel_sahbvh.cpp:128-151).
struct FlatScene {
    std::vector<float> verts;     // [N][9] positions
    std::vector<float> normals;   // [N][9] vertex normals (world space, normalised)
    std::vector<int32_t> tri_node, tri_face;
    std::vector<int32_t> node_prim;   // scene node -> primitive slot (-1: none)
    std::vector<int32_t> prim_node;   // primitive slot -> scene node
    std::vector<Transform> node_transform;   // per scene node: its global transform
};

class Scene {
public:
    Scene();   // node 0 is the root group
    int create_group_node(std::optional<Mat4> local_transform = {});
    int create_primitive_node(std::shared_ptr<const Mesh> mesh, const Material& material,
                              std::optional<AreaLight> light = {});
    void add_child(int parent, int child);
    // primitive under the root, optionally inside its own transform group
    int add_primitive(std::shared_ptr<const Mesh> mesh, const Material& material,
                      std::optional<AreaLight> light = {}, std::optional<Mat4> transform = {});
    void set_env(Vec3 Le) { env_ = Le; }
    const std::optional<Vec3>& env() const { return env_; }
    void set_camera(const Camera& c) { camera_ = c; }
    const Camera& camera() const { return camera_; }
    const SceneNode& node_at(int index) const;
    int num_nodes() const { return int(nodes_.size()); }
    // Scene::traverse_primitive_nodes: DFS from the root with accumulated transforms
    void traverse_primitive_nodes(const std::function<void(const SceneNode&, const Mat4&)>& f) const;
    FlatScene flatten() const;
    // bumped by every modification, so renderers can rebuild their device copy
    uint64_t version() const { return version_; }

private:
    std::vector<SceneNode> nodes_;
    std::optional<Vec3> env_;
    Camera camera_;
    uint64_t version_ = 0;
};

// ---- accel -------------------------------------------------------------------

// lm::Accel::Hit
struct Hit {
    Float t;
    Vec2 uv;
    Transform global_transform;
    int primitive;   // scene node index
    int face;
};

class Accel {
public:
    virtual ~Accel() = default;
    virtual void build(const Scene& scene) = 0;
    virtual std::optional<Hit> intersect(Ray ray, Float tmin, Float tmax) const = 0;
};

struct AccelOptions {
    int width = 4;        // 2, 4 or 8
    int leaf_size = 2;    // 1..15
    int refill = 0;       // traversal wave refill threshold (0: the library default, 8)
    bool ploc = false;    // PLOC instead of LBVH
    int ploc_radius = 16;
};

// accel::mi355x
class Accel_MI355X final : public Accel {
public:
    explicit Accel_MI355X(AccelOptions opts = {});
    ~Accel_MI355X() override;
    Accel_MI355X(const Accel_MI355X&) = delete;
    Accel_MI355X& operator=(const Accel_MI355X&) = delete;
    void build(const Scene& scene) override;
    // single ray, thread-safe (a batch of one on the device)
    std::optional<Hit> intersect(Ray ray, Float tmin, Float tmax) const override;
    // many rays from host memory in one launch
    std::vector<std::optional<Hit>> intersect_many(const std::vector<Ray>& rays, Float tmin, Float tmax) const;
    // device-pointer batch (asynchronous on stream); hits carry global triangle indices
    void intersect_device(const lmgpu_ray* rays, lmgpu_hit* hits, int64_t n, void* stream) const;
    std::optional<Hit> map_hit(int32_t tri, float t, float u, float v) const;
    lmgpu_accel* handle() const { return h_; }

private:
    AccelOptions opts_;
    lmgpu_accel* h_ = nullptr;
    FlatScene flat_;
};

// ---- film --------------------------------------------------------------------

// film::bitmap: FP64 RGB, row-major, y = 0 the bottom row.
class Film {
public:
    struct Size {
        int w, h;
    };
    Film(int w, int h);
    Size size() const { return {w_, h_}; }
    long long num_pixels() const { return (long long)w_ * h_; }
    void set_pixel(int x, int y, Vec3 v);
    void splat_pixel(int x, int y, Vec3 v);
    Vec3 pixel(int x, int y) const;
    void rescale(Float s);
    void clear();
    // contiguous w * h * 3 doubles (Film::buffer)
    const std::vector<double>& buffer() const { return data_; }
    std::vector<double>& data() { return data_; }

private:
    int w_, h_;
    std::vector<double> data_;
};

// ---- renderers ------------------------------------------------------------------

// JSON property subset of the reference's construct() calls.
using PropValue = std::variant<bool, int, double, std::string, Vec3>;
using Props = std::map<std::string, PropValue>;

struct RenderResult {
    long long processed = 0;   // samples per pixel ("processed")
    Float elapsed = 0;         // seconds ("elapsed")
    int64_t extension_rays = 0, shadow_rays = 0, guide_records = 0;
};

class Renderer {
public:
    virtual ~Renderer() = default;
    virtual RenderResult render() = 0;
};

// renderer::pt_mi355x. Props (src/renderer/renderer_pt.cpp:51-77):
//   max_verts (int, required), sampling_mode ("mis" | "nee" | "naive"),
//   primary_ray_sampling_mode ("pixel" | "image"), scheduler ("sample" |
//   "time"): spp (pixel mode, scheduler::spp::sample), num_samples (image
//   mode, scheduler::spi::sample) or render_time (seconds),
//   seed (int), rr_depth (int; 5, or max_verts when guiding), guiding
//   (bool) with guide_alpha (0.65), guide_lr (1e-2), guide_batch (2^14),
//   guide_steps (4), guide_depth (8);
//   accel_width, accel_leaf_size (BVH of the renderer's device scene).
class Renderer_PT_MI355X final : public Renderer {
public:
    Renderer_PT_MI355X(const Scene* scene, Film* film, const Props& props);
    ~Renderer_PT_MI355X() override;
    RenderResult render() override;
    // per-pixel sum and sum of squares of the per-sample estimate (FP64) of the last render
    // (no sums of squares in image mode: film_sumsq() is empty)
    const std::vector<double>& film_sum() const { return sum_; }
    const std::vector<double>& film_sumsq() const { return sumsq_; }

private:
    const Scene* scene_;
    Film* film_;
    int max_verts_, mode_, rr_depth_, spp_;
    bool image_;
    int64_t num_samples_;
    double render_time_;
    uint64_t seed_;
    bool guiding_;
    float g_alpha_, g_lr_;
    int g_batch_, g_steps_, g_depth_;
    int acc_width_, acc_leaf_;
    lmgpu_scene* gs_ = nullptr;
    lmgpu_guide* guide_ = nullptr;
    uint64_t built_version_ = ~0ull;
    std::vector<double> sum_, sumsq_;
};

// renderer::lt_mi355x (src/renderer/renderer_lt.cpp:22-170: light tracing with
// NEE to the camera). Props: max_verts (int, required), scheduler ("sample":
// num_samples light paths, or "time": render_time seconds), seed (int),
// accel_width, accel_leaf_size. processed = light paths traced; the film is
// the splat sum times w*h / processed (renderer_lt.cpp:166).
class Renderer_LT_MI355X final : public Renderer {
public:
    Renderer_LT_MI355X(const Scene* scene, Film* film, const Props& props);
    ~Renderer_LT_MI355X() override;
    RenderResult render() override;

private:
    const Scene* scene_;
    Film* film_;
    int max_verts_;
    long long num_samples_;
    double render_time_;
    uint64_t seed_;
    int acc_width_, acc_leaf_;
    lmgpu_scene* gs_ = nullptr;
    uint64_t built_version_ = ~0ull;
};

// renderer::bdpt_mi355x (src/renderer/renderer_bdpt.cpp:339-411, bidirectional
// path tracing with power-heuristic MIS). Props as Renderer_Path_Base
// (:43-55): max_verts (int, required), min_verts (2), scheduler ("sample":
// num_samples, or "time": render_time), seed; accel_width, accel_leaf_size.
// Area lights only (an environment light is Error::Unsupported).
class Renderer_BDPT_MI355X final : public Renderer {
public:
    Renderer_BDPT_MI355X(const Scene* scene, Film* film, const Props& props);
    ~Renderer_BDPT_MI355X() override;
    RenderResult render() override;

private:
    const Scene* scene_;
    Film* film_;
    int max_verts_, min_verts_;
    long long num_samples_;
    double render_time_;
    uint64_t seed_;
    int acc_width_, acc_leaf_;
    lmgpu_scene* gs_ = nullptr;
    uint64_t built_version_ = ~0ull;
};

// renderer::raycast_mi355x (src/renderer/renderer_raycast.cpp:48-97). Props:
//   bg_color (Vec3), use_constant_color (bool), visualize_normal (bool), color (Vec3).
class Renderer_Raycast_MI355X final : public Renderer {
public:
    Renderer_Raycast_MI355X(const Scene* scene, Film* film, const Props& props);
    ~Renderer_Raycast_MI355X() override;
    RenderResult render() override;

private:
    const Scene* scene_;
    Film* film_;
    Vec3 bg_;
    bool constant_, normal_;
    std::optional<Vec3> color_;
    lmgpu_scene* gs_ = nullptr;
    uint64_t built_version_ = ~0ull;
};

}  // namespace lm_mi355x

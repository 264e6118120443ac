// C++ host components over the lmgpu C ABI (see lm_mi355x.h).
#include "lm_mi355x.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "lmgpu.h"

namespace lm_mi355x {

// ---- math -------------------------------------------------------------------

Float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Float length(Vec3 a) { return std::sqrt(dot(a, a)); }
Vec3 normalize(Vec3 a) { return a / length(a); }

Mat4 operator*(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            Float s = 0;
            for (int k = 0; k < 4; k++) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}

Mat4 Mat4::translate(Vec3 t) {
    Mat4 r;
    r.m[0][3] = t.x;
    r.m[1][3] = t.y;
    r.m[2][3] = t.z;
    return r;
}

Mat4 Mat4::scale(Vec3 s) {
    Mat4 r;
    r.m[0][0] = s.x;
    r.m[1][1] = s.y;
    r.m[2][2] = s.z;
    return r;
}

Mat4 Mat4::rotate(Float a, Vec3 axis) {
    // Rodrigues' rotation about a unit axis
    const Vec3 u = normalize(axis);
    const Float c = std::cos(a), s = std::sin(a), t = 1 - c;
    Mat4 r;
    r.m[0][0] = c + u.x * u.x * t;
    r.m[0][1] = u.x * u.y * t - u.z * s;
    r.m[0][2] = u.x * u.z * t + u.y * s;
    r.m[1][0] = u.y * u.x * t + u.z * s;
    r.m[1][1] = c + u.y * u.y * t;
    r.m[1][2] = u.y * u.z * t - u.x * s;
    r.m[2][0] = u.z * u.x * t - u.y * s;
    r.m[2][1] = u.z * u.y * t + u.x * s;
    r.m[2][2] = c + u.z * u.z * t;
    return r;
}

Transform::Transform(const Mat4& M_) : M(M_) {
    // normal_M = transpose(inverse(M_lin)) = cofactor(M_lin) / det(M_lin)
    const auto& a = M.m;
    const Float c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    const Float c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    const Float c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    const Float c10 = a[0][2] * a[2][1] - a[0][1] * a[2][2];
    const Float c11 = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    const Float c12 = a[0][1] * a[2][0] - a[0][0] * a[2][1];
    const Float c20 = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    const Float c21 = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    const Float c22 = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    const Float det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    J = std::abs(det);
    const Float inv = det != 0 ? 1 / det : 0;
    const Float c[3][3] = {{c00, c01, c02}, {c10, c11, c12}, {c20, c21, c22}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) normal_M.m[i][j] = c[i][j] * inv;
}

Vec3 Transform::point(Vec3 p) const {
    const auto& a = M.m;
    return {a[0][0] * p.x + a[0][1] * p.y + a[0][2] * p.z + a[0][3],
            a[1][0] * p.x + a[1][1] * p.y + a[1][2] * p.z + a[1][3],
            a[2][0] * p.x + a[2][1] * p.y + a[2][2] * p.z + a[2][3]};
}

Vec3 Transform::normal(Vec3 n) const {
    const auto& a = normal_M.m;
    return normalize({a[0][0] * n.x + a[0][1] * n.y + a[0][2] * n.z, a[1][0] * n.x + a[1][1] * n.y + a[1][2] * n.z,
                      a[2][0] * n.x + a[2][1] * n.y + a[2][2] * n.z});
}

// ---- errors -------------------------------------------------------------------

static void check(int rc, const char* what) {
    if (rc == LMGPU_OK) return;
    const char* msg = lmgpu_last_error();
    const Error e = rc == LMGPU_ERR_INVALID_ARGUMENT ? Error::InvalidArgument
                    : rc == LMGPU_ERR_UNSUPPORTED     ? Error::Unsupported
                                                      : Error::FailedToRender;
    throw Exception(e, std::string(what) + ": " + (msg ? msg : ""));
}

// ---- scene ------------------------------------------------------------------------

Mesh::Tri Mesh::triangle_at(int face) const {
    const auto& f = fs.at(size_t(face));
    const Vec3 p1 = ps.at(size_t(f[0].p)), p2 = ps.at(size_t(f[1].p)), p3 = ps.at(size_t(f[2].p));
    // Mesh_Raw::surface_point (src/mesh/mesh_raw.cpp:87-97): vertex normals,
    // or the geometric normal when the mesh has none
    const Vec3 gn = normalize(cross(p2 - p1, p3 - p1));
    auto nrm = [&](const Index& ix) { return ix.n >= 0 ? ns.at(size_t(ix.n)) : gn; };
    return Tri{{p1, nrm(f[0])}, {p2, nrm(f[1])}, {p3, nrm(f[2])}};
}

void Mesh::foreach_triangle(const std::function<void(int, const Tri&)>& f) const {
    for (int i = 0; i < num_triangles(); i++) f(i, triangle_at(i));
}

Vec3 Material::reflectance() const {
    if (type == MaterialType::Diffuse || type == MaterialType::Glossy) return color;
    return {};
}

Ray Camera::primary_ray(Vec2 rp) const {
    // camera::pinhole construct + primary_ray (src/camera/camera_pinhole.cpp:86-96,128-132)
    const Vec3 w = normalize(position - center);
    const Vec3 u = normalize(cross(up, w));
    const Vec3 v = cross(w, u);
    const Float tf = std::tan(vfov * 3.14159265358979323846 / 180.0 * .5);
    const Float x = 2 * rp.x - 1, y = 2 * rp.y - 1;
    const Vec3 d = normalize(Vec3{aspect * tf * x, tf * y, -1});
    return {position, u * d.x + v * d.y + w * d.z};
}

Scene::Scene() {
    SceneNode root;
    root.type = SceneNodeType::Group;
    root.index = 0;
    nodes_.push_back(root);
}

int Scene::create_group_node(std::optional<Mat4> local_transform) {
    SceneNode n;
    n.type = SceneNodeType::Group;
    n.index = int(nodes_.size());
    n.group.local_transform = local_transform;
    nodes_.push_back(n);
    version_++;
    return n.index;
}

int Scene::create_primitive_node(std::shared_ptr<const Mesh> mesh, const Material& material,
                                 std::optional<AreaLight> light) {
    if (!mesh) throw Exception(Error::InvalidArgument, "primitive without a mesh");
    if (material.type == MaterialType::Glossy && !(material.ax > 0 && material.ay > 0))
        throw Exception(Error::InvalidArgument, "glossy roughness must be positive");
    if (material.type == MaterialType::Glass && !(material.Ni > 0))
        throw Exception(Error::InvalidArgument, "glass Ni must be positive");
    SceneNode n;
    n.type = SceneNodeType::Primitive;
    n.index = int(nodes_.size());
    n.primitive.mesh = std::move(mesh);
    n.primitive.material = material;
    n.primitive.light = light;
    nodes_.push_back(n);
    version_++;
    return n.index;
}

void Scene::add_child(int parent, int child) {
    if (parent < 0 || parent >= num_nodes() || nodes_[size_t(parent)].type != SceneNodeType::Group)
        throw Exception(Error::InvalidArgument, "parent must be a group node");
    if (child <= 0 || child >= num_nodes()) throw Exception(Error::InvalidArgument, "invalid child node");
    // the graph stays acyclic: the parent must not be reachable from the child
    std::function<bool(int)> reaches = [&](int n) {
        if (n == parent) return true;
        const SceneNode& node = nodes_[size_t(n)];
        if (node.type != SceneNodeType::Group) return false;
        for (int c : node.group.children)
            if (reaches(c)) return true;
        return false;
    };
    if (reaches(child)) throw Exception(Error::InvalidArgument, "add_child would create a cycle");
    nodes_[size_t(parent)].group.children.push_back(child);
    version_++;
}

int Scene::add_primitive(std::shared_ptr<const Mesh> mesh, const Material& material, std::optional<AreaLight> light,
                         std::optional<Mat4> transform) {
    const int p = create_primitive_node(std::move(mesh), material, light);
    if (transform) {
        const int g = create_group_node(transform);
        add_child(0, g);
        add_child(g, p);
    } else {
        add_child(0, p);
    }
    return p;
}

const SceneNode& Scene::node_at(int index) const {
    if (index < 0 || index >= num_nodes()) throw Exception(Error::InvalidArgument, "node index out of range");
    return nodes_[size_t(index)];
}

void Scene::traverse_primitive_nodes(const std::function<void(const SceneNode&, const Mat4&)>& f) const {
    std::function<void(int, const Mat4&)> visit = [&](int index, const Mat4& global) {
        const SceneNode& node = nodes_[size_t(index)];
        f(node, global);
        if (node.type == SceneNodeType::Group) {
            const Mat4 M = node.group.local_transform ? global * *node.group.local_transform : global;
            for (int c : node.group.children) visit(c, M);
        }
    };
    visit(0, Mat4::identity());
}

FlatScene Scene::flatten() const {
    FlatScene fl;
    fl.node_prim.assign(nodes_.size(), -1);
    fl.node_transform.assign(nodes_.size(), Transform());
    traverse_primitive_nodes([&](const SceneNode& node, const Mat4& global) {
        if (node.type != SceneNodeType::Primitive) return;
        const Transform T(global);
        fl.node_transform[size_t(node.index)] = T;
        if (fl.node_prim[size_t(node.index)] < 0) {
            fl.node_prim[size_t(node.index)] = int32_t(fl.prim_node.size());
            fl.prim_node.push_back(node.index);
        }
        node.primitive.mesh->foreach_triangle([&](int face, const Mesh::Tri& tri) {
            for (const Mesh::Point* pt : {&tri.p1, &tri.p2, &tri.p3}) {
                const Vec3 w = T.point(pt->p);
                const Vec3 n = T.normal(pt->n);
                fl.verts.insert(fl.verts.end(), {float(w.x), float(w.y), float(w.z)});
                fl.normals.insert(fl.normals.end(), {float(n.x), float(n.y), float(n.z)});
            }
            fl.tri_node.push_back(node.index);
            fl.tri_face.push_back(face);
        });
    });
    return fl;
}

// Device scene (geometry, accel, materials, lights, camera) of a Scene.
static lmgpu_scene* make_device_scene(const Scene& scene, int width, int leaf, int w, int h) {
    const FlatScene fl = scene.flatten();
    if (fl.tri_node.empty()) throw Exception(Error::InvalidArgument, "scene has no triangles");
    const int P = int(fl.prim_node.size());
    std::vector<int32_t> tri_prim(fl.tri_node.size());
    for (size_t t = 0; t < tri_prim.size(); t++) tri_prim[t] = fl.node_prim[size_t(fl.tri_node[t])];
    std::vector<int32_t> mat_type(static_cast<size_t>(P));
    std::vector<float> mat_param(size_t(P) * 8, 0.f), Ke(size_t(P) * 3, 0.f);
    std::vector<int32_t> light_prims;
    for (int i = 0; i < P; i++) {
        const SceneNode& n = scene.node_at(fl.prim_node[size_t(i)]);
        const Material& m = n.primitive.material;
        mat_type[size_t(i)] = int32_t(m.type);
        float* mp = &mat_param[size_t(i) * 8];
        mp[0] = float(m.color.x);
        mp[1] = float(m.color.y);
        mp[2] = float(m.color.z);
        mp[3] = float(m.ax);
        mp[4] = float(m.ay);
        mp[5] = float(m.Ni);
        if (n.primitive.light) {
            light_prims.push_back(i);
            Ke[size_t(i) * 3 + 0] = float(n.primitive.light->Ke.x);
            Ke[size_t(i) * 3 + 1] = float(n.primitive.light->Ke.y);
            Ke[size_t(i) * 3 + 2] = float(n.primitive.light->Ke.z);
        }
    }
    lmgpu_scene_desc d{};
    d.num_tris = int64_t(fl.tri_node.size());
    d.verts = fl.verts.data();
    d.normals = fl.normals.data();
    d.tri_prim = tri_prim.data();
    d.num_prims = P;
    d.mat_type = mat_type.data();
    d.mat_param = mat_param.data();
    d.prim_Ke = Ke.data();
    d.num_area_lights = int32_t(light_prims.size());
    d.light_prims = light_prims.empty() ? nullptr : light_prims.data();
    d.has_env = scene.env() ? 1 : 0;
    if (scene.env()) {
        d.env_Le[0] = float(scene.env()->x);
        d.env_Le[1] = float(scene.env()->y);
        d.env_Le[2] = float(scene.env()->z);
    }
    const Camera& c = scene.camera();
    const Vec3 cv[3] = {c.position, c.center, c.up};
    for (int k = 0; k < 3; k++) {
        d.cam_pos[k] = float(cv[0][k]);
        d.cam_center[k] = float(cv[1][k]);
        d.cam_up[k] = float(cv[2][k]);
    }
    d.vfov = float(c.vfov);
    d.aspect = float(c.aspect);
    d.w = w;
    d.h = h;
    lmgpu_build_opts o{};
    o.width = width;
    o.leaf_size = leaf;
    o.builder = LMGPU_BUILDER_LBVH;
    lmgpu_scene* s = nullptr;
    check(lmgpu_scene_create(&s, &d, &o), "lmgpu_scene_create");
    return s;
}

// ---- accel --------------------------------------------------------------------------

Accel_MI355X::Accel_MI355X(AccelOptions opts) : opts_(opts) {
    if (opts_.width != 2 && opts_.width != 4 && opts_.width != 8)
        throw Exception(Error::InvalidArgument, "accel width must be 2, 4 or 8");
    check(lmgpu_accel_create(&h_), "lmgpu_accel_create");
}

Accel_MI355X::~Accel_MI355X() {
    if (h_) lmgpu_accel_destroy(h_);
}

void Accel_MI355X::build(const Scene& scene) {
    flat_ = scene.flatten();
    lmgpu_build_opts o{};
    o.width = opts_.width;
    o.leaf_size = opts_.leaf_size;
    o.refill = opts_.refill;
    o.builder = opts_.ploc ? LMGPU_BUILDER_PLOC : LMGPU_BUILDER_LBVH;
    o.ploc_radius = opts_.ploc_radius;
    check(lmgpu_accel_build(h_, flat_.verts.data(), int64_t(flat_.tri_node.size()), 0, &o, nullptr),
          "lmgpu_accel_build");
}

std::optional<Hit> Accel_MI355X::map_hit(int32_t tri, float t, float u, float v) const {
    if (tri < 0) return std::nullopt;
    if (size_t(tri) >= flat_.tri_node.size()) throw Exception(Error::FailedToRender, "hit index out of range");
    const int node = flat_.tri_node[size_t(tri)];
    return Hit{Float(t), Vec2{Float(u), Float(v)}, flat_.node_transform[size_t(node)], node,
               flat_.tri_face[size_t(tri)]};
}

static lmgpu_ray to_ray(const Ray& r, Float tmin, Float tmax) {
    lmgpu_ray g;
    g.org[0] = float(r.o.x);
    g.org[1] = float(r.o.y);
    g.org[2] = float(r.o.z);
    g.tmin = float(tmin);
    g.dir[0] = float(r.d.x);
    g.dir[1] = float(r.d.y);
    g.dir[2] = float(r.d.z);
    g.tmax = float(std::min<Float>(tmax, 3.0e38));
    return g;
}

std::optional<Hit> Accel_MI355X::intersect(Ray ray, Float tmin, Float tmax) const {
    const lmgpu_ray r = to_ray(ray, tmin, tmax);
    lmgpu_hit h;
    check(lmgpu_accel_intersect_host(h_, &r, &h, 1), "lmgpu_accel_intersect_host");
    return map_hit(h.prim, h.t, h.u, h.v);
}

std::vector<std::optional<Hit>> Accel_MI355X::intersect_many(const std::vector<Ray>& rays, Float tmin,
                                                             Float tmax) const {
    std::vector<lmgpu_ray> r(rays.size());
    for (size_t i = 0; i < rays.size(); i++) r[i] = to_ray(rays[i], tmin, tmax);
    std::vector<lmgpu_hit> h(rays.size());
    if (!rays.empty())
        check(lmgpu_accel_intersect_host(h_, r.data(), h.data(), int64_t(r.size())), "lmgpu_accel_intersect_host");
    std::vector<std::optional<Hit>> out(rays.size());
    for (size_t i = 0; i < rays.size(); i++) out[i] = map_hit(h[i].prim, h[i].t, h[i].u, h[i].v);
    return out;
}

void Accel_MI355X::intersect_device(const lmgpu_ray* rays, lmgpu_hit* hits, int64_t n, void* stream) const {
    check(lmgpu_accel_intersect(h_, rays, hits, n, stream), "lmgpu_accel_intersect");
}

// ---- film ------------------------------------------------------------------------------

Film::Film(int w, int h) : w_(w), h_(h) {
    if (w <= 0 || h <= 0) throw Exception(Error::InvalidArgument, "film size must be positive");
    data_.assign(size_t(w) * size_t(h) * 3, 0.0);
}

void Film::set_pixel(int x, int y, Vec3 v) {
    double* p = &data_[(size_t(y) * size_t(w_) + size_t(x)) * 3];
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

void Film::splat_pixel(int x, int y, Vec3 v) {
    double* p = &data_[(size_t(y) * size_t(w_) + size_t(x)) * 3];
    p[0] += v.x;
    p[1] += v.y;
    p[2] += v.z;
}

Vec3 Film::pixel(int x, int y) const {
    const double* p = &data_[(size_t(y) * size_t(w_) + size_t(x)) * 3];
    return {p[0], p[1], p[2]};
}

void Film::rescale(Float s) {
    for (double& v : data_) v *= s;
}

void Film::clear() { std::fill(data_.begin(), data_.end(), 0.0); }

// ---- props -----------------------------------------------------------------------------

template <class T>
static std::optional<T> prop(const Props& p, const std::string& key) {
    const auto it = p.find(key);
    if (it == p.end()) return std::nullopt;
    if (const T* v = std::get_if<T>(&it->second)) return *v;
    if constexpr (std::is_same_v<T, double>)
        if (const int* v = std::get_if<int>(&it->second)) return double(*v);
    throw Exception(Error::InvalidArgument, "property '" + key + "' has the wrong type");
}

template <class T>
static T prop_or(const Props& p, const std::string& key, T def) {
    const auto v = prop<T>(p, key);
    return v ? *v : def;
}

template <class T>
static T prop_req(const Props& p, const std::string& key) {
    const auto v = prop<T>(p, key);
    if (!v) throw Exception(Error::InvalidArgument, "missing property '" + key + "'");
    return *v;
}

// ---- renderer::pt ---------------------------------------------------------------------

Renderer_PT_MI355X::Renderer_PT_MI355X(const Scene* scene, Film* film, const Props& p) : scene_(scene), film_(film) {
    if (!scene_ || !film_) throw Exception(Error::InvalidArgument, "scene and output film are required");
    max_verts_ = prop_req<int>(p, "max_verts");
    const std::string mode = prop_or<std::string>(p, "sampling_mode", "mis");
    if (mode == "naive") mode_ = LMGPU_SAMPLING_NAIVE;
    else if (mode == "nee") mode_ = LMGPU_SAMPLING_NEE;
    else if (mode == "mis") mode_ = LMGPU_SAMPLING_MIS;
    else throw Exception(Error::InvalidArgument, "invalid sampling_mode '" + mode + "'");
    guiding_ = prop_or<bool>(p, "guiding", false);
    // Roulette on throughput alone cuts the low-throughput paths that a
    // guide steers toward the light, so guided rendering defaults to none
    // (paths end at max_verts); plain rendering keeps renderer_pt.cpp:281.
    rr_depth_ = prop_or<int>(p, "rr_depth", guiding_ ? std::max(max_verts_, 5) : 5);
    seed_ = uint64_t(prop_or<int>(p, "seed", 0));
    const std::string prm = prop_or<std::string>(p, "primary_ray_sampling_mode", "pixel");
    if (prm != "pixel" && prm != "image")
        throw Exception(Error::InvalidArgument, "invalid primary_ray_sampling_mode '" + prm + "'");
    image_ = prm == "image";
    if (image_ && guiding_) throw Exception(Error::Unsupported, "guided rendering uses the pixel sampling mode");
    const std::string sched = prop_or<std::string>(p, "scheduler", "sample");
    spp_ = 0;
    num_samples_ = 0;
    render_time_ = 0;
    if (sched == "sample" && !image_) {
        spp_ = prop_req<int>(p, "spp");
        if (spp_ <= 0) throw Exception(Error::InvalidArgument, "spp must be positive");
    } else if (sched == "sample") {
        // an int, or a whole double for counts beyond 2^31
        const double ns = prop_req<double>(p, "num_samples");
        if (!(ns >= 1) || ns != std::floor(ns) || ns > 9.0e15)
            throw Exception(Error::InvalidArgument, "num_samples must be a positive whole number");
        num_samples_ = int64_t(ns);
    } else if (sched == "time") {
        render_time_ = prop_req<double>(p, "render_time");
        if (!(render_time_ > 0)) throw Exception(Error::InvalidArgument, "render_time must be positive");
    } else {
        throw Exception(Error::InvalidArgument, "invalid scheduler '" + sched + "'");
    }
    g_alpha_ = float(prop_or<double>(p, "guide_alpha", 0.65));
    g_lr_ = float(prop_or<double>(p, "guide_lr", 1e-2));
    g_batch_ = prop_or<int>(p, "guide_batch", 1 << 14);
    g_steps_ = prop_or<int>(p, "guide_steps", 4);
    g_depth_ = prop_or<int>(p, "guide_depth", 8);
    acc_width_ = prop_or<int>(p, "accel_width", 4);
    acc_leaf_ = prop_or<int>(p, "accel_leaf_size", 2);
}

Renderer_PT_MI355X::~Renderer_PT_MI355X() {
    if (guide_) lmgpu_guide_destroy(guide_);
    if (gs_) lmgpu_scene_destroy(gs_);
}

RenderResult Renderer_PT_MI355X::render() {
    const auto size = film_->size();
    if (!gs_ || built_version_ != scene_->version()) {
        if (gs_) lmgpu_scene_destroy(gs_);
        gs_ = nullptr;
        gs_ = make_device_scene(*scene_, acc_width_, acc_leaf_, size.w, size.h);
        built_version_ = scene_->version();
        if (guide_) {
            lmgpu_guide_destroy(guide_);
            guide_ = nullptr;
        }
    }
    if (guiding_ && !guide_) {
        // guide box: the scene bounds, slightly enlarged
        const FlatScene fl = scene_->flatten();
        float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
        for (size_t i = 0; i < fl.verts.size(); i++) {
            lo[i % 3] = std::min(lo[i % 3], fl.verts[i]);
            hi[i % 3] = std::max(hi[i % 3], fl.verts[i]);
        }
        for (int a = 0; a < 3; a++) {
            lo[a] -= 1e-3f;
            hi[a] += 1e-3f;
        }
        check(lmgpu_guide_create(&guide_, lo, hi, seed_ + 1), "lmgpu_guide_create");
    }
    lmgpu_pt_params pp{};
    pp.spp = spp_;
    pp.max_verts = max_verts_;
    pp.sampling_mode = mode_;
    pp.rr_depth = rr_depth_;
    pp.seed = seed_;
    pp.render_time = render_time_;
    pp.primary_mode = image_ ? LMGPU_PRIMARY_IMAGE : LMGPU_PRIMARY_PIXEL;
    pp.num_samples = num_samples_;
    const size_t n = size_t(size.w) * size_t(size.h) * 3;
    sum_.assign(n, 0.0);
    if (image_) sumsq_.clear();
    else sumsq_.assign(n, 0.0);
    lmgpu_pt_stats st{};
    if (guiding_) {
        lmgpu_guide_opts go{};
        go.alpha = g_alpha_;
        go.learning_rate = g_lr_;
        go.batch = g_batch_;
        go.steps_per_pass = g_steps_;
        go.record_depth = g_depth_;
        check(lmgpu_pt_render_guided(gs_, &pp, guide_, &go, sum_.data(), sumsq_.data(), 0, &st, nullptr),
              "lmgpu_pt_render_guided");
    } else {
        check(lmgpu_pt_render(gs_, &pp, sum_.data(), image_ ? nullptr : sumsq_.data(), 0, &st, nullptr),
              "lmgpu_pt_render");
    }
    // film value = sum / spp in pixel mode, sum * w*h / processed in image mode
    // (src/renderer/renderer_pt.cpp:307-312)
    film_->clear();
    auto& data = film_->data();
    const int64_t processed = image_ ? st.samples : st.spp;
    const double scale = processed <= 0 ? 0.0 : (image_ ? double(size.w) * size.h : 1.0) / double(processed);
    for (size_t i = 0; i < n; i++) data[i] = sum_[i] * scale;
    RenderResult r;
    r.processed = processed;
    r.elapsed = st.elapsed_s;
    r.extension_rays = st.extension_rays;
    r.shadow_rays = st.shadow_rays;
    r.guide_records = st.guide_records;
    return r;
}

// ---- renderer::lt ----------------------------------------------------------------------

Renderer_LT_MI355X::Renderer_LT_MI355X(const Scene* scene, Film* film, const Props& p) : scene_(scene), film_(film) {
    if (!scene_ || !film_) throw Exception(Error::InvalidArgument, "scene and output film are required");
    max_verts_ = prop_req<int>(p, "max_verts");
    seed_ = uint64_t(prop_or<int>(p, "seed", 0));
    const std::string sched = prop_req<std::string>(p, "scheduler");
    num_samples_ = 0;
    render_time_ = 0;
    if (sched == "sample") {
        num_samples_ = prop_req<int>(p, "num_samples");
        if (num_samples_ <= 0) throw Exception(Error::InvalidArgument, "num_samples must be positive");
    } else if (sched == "time") {
        render_time_ = prop_req<double>(p, "render_time");
        if (!(render_time_ > 0)) throw Exception(Error::InvalidArgument, "render_time must be positive");
    } else {
        throw Exception(Error::InvalidArgument, "invalid scheduler '" + sched + "'");
    }
    acc_width_ = prop_or<int>(p, "accel_width", 4);
    acc_leaf_ = prop_or<int>(p, "accel_leaf_size", 2);
}

Renderer_LT_MI355X::~Renderer_LT_MI355X() {
    if (gs_) lmgpu_scene_destroy(gs_);
}

RenderResult Renderer_LT_MI355X::render() {
    const auto size = film_->size();
    if (!gs_ || built_version_ != scene_->version()) {
        if (gs_) lmgpu_scene_destroy(gs_);
        gs_ = nullptr;
        gs_ = make_device_scene(*scene_, acc_width_, acc_leaf_, size.w, size.h);
        built_version_ = scene_->version();
    }
    lmgpu_lt_params lp{};
    lp.num_samples = num_samples_;
    lp.max_verts = max_verts_;
    lp.seed = seed_;
    lp.render_time = render_time_;
    film_->clear();
    auto& data = film_->data();
    lmgpu_pt_stats st{};
    check(lmgpu_lt_render(gs_, &lp, data.data(), 0, &st, nullptr), "lmgpu_lt_render");
    // Film::rescale(w*h / processed) (src/renderer/renderer_lt.cpp:166)
    const double s = st.samples > 0 ? double(size.w) * double(size.h) / double(st.samples) : 0.0;
    for (auto& v : data) v *= s;
    RenderResult r;
    r.processed = st.samples;
    r.elapsed = st.elapsed_s;
    r.extension_rays = st.extension_rays;
    r.shadow_rays = st.shadow_rays;
    return r;
}

// ---- renderer::bdpt --------------------------------------------------------------------

Renderer_BDPT_MI355X::Renderer_BDPT_MI355X(const Scene* scene, Film* film, const Props& p)
    : scene_(scene), film_(film) {
    if (!scene_ || !film_) throw Exception(Error::InvalidArgument, "scene and output film are required");
    max_verts_ = prop_req<int>(p, "max_verts");
    min_verts_ = prop_or<int>(p, "min_verts", 2);
    seed_ = uint64_t(prop_or<int>(p, "seed", 0));
    const std::string sched = prop_req<std::string>(p, "scheduler");
    num_samples_ = 0;
    render_time_ = 0;
    if (sched == "sample") {
        num_samples_ = prop_req<int>(p, "num_samples");
        if (num_samples_ <= 0) throw Exception(Error::InvalidArgument, "num_samples must be positive");
    } else if (sched == "time") {
        render_time_ = prop_req<double>(p, "render_time");
        if (!(render_time_ > 0)) throw Exception(Error::InvalidArgument, "render_time must be positive");
    } else {
        throw Exception(Error::InvalidArgument, "invalid scheduler '" + sched + "'");
    }
    acc_width_ = prop_or<int>(p, "accel_width", 4);
    acc_leaf_ = prop_or<int>(p, "accel_leaf_size", 2);
}

Renderer_BDPT_MI355X::~Renderer_BDPT_MI355X() {
    if (gs_) lmgpu_scene_destroy(gs_);
}

RenderResult Renderer_BDPT_MI355X::render() {
    const auto size = film_->size();
    if (!gs_ || built_version_ != scene_->version()) {
        if (gs_) lmgpu_scene_destroy(gs_);
        gs_ = nullptr;
        gs_ = make_device_scene(*scene_, acc_width_, acc_leaf_, size.w, size.h);
        built_version_ = scene_->version();
    }
    lmgpu_lt_params lp{};
    lp.num_samples = num_samples_;
    lp.max_verts = max_verts_;
    lp.min_verts = min_verts_;
    lp.seed = seed_;
    lp.render_time = render_time_;
    film_->clear();
    auto& data = film_->data();
    lmgpu_pt_stats st{};
    check(lmgpu_bdpt_render(gs_, &lp, data.data(), 0, &st, nullptr), "lmgpu_bdpt_render");
    // Film::rescale(w*h / processed) (src/renderer/renderer_bdpt.cpp:398-399)
    const double s = st.samples > 0 ? double(size.w) * double(size.h) / double(st.samples) : 0.0;
    for (auto& v : data) v *= s;
    RenderResult r;
    r.processed = st.samples;
    r.elapsed = st.elapsed_s;
    r.extension_rays = st.extension_rays;
    r.shadow_rays = st.shadow_rays;
    return r;
}

// ---- renderer::raycast ------------------------------------------------------------------

Renderer_Raycast_MI355X::Renderer_Raycast_MI355X(const Scene* scene, Film* film, const Props& p)
    : scene_(scene), film_(film) {
    if (!scene_ || !film_) throw Exception(Error::InvalidArgument, "scene and output film are required");
    bg_ = prop_or<Vec3>(p, "bg_color", Vec3{});
    constant_ = prop_or<bool>(p, "use_constant_color", false);
    normal_ = prop_or<bool>(p, "visualize_normal", false);
    color_ = prop<Vec3>(p, "color");
}

Renderer_Raycast_MI355X::~Renderer_Raycast_MI355X() {
    if (gs_) lmgpu_scene_destroy(gs_);
}

RenderResult Renderer_Raycast_MI355X::render() {
    const auto t0 = std::chrono::steady_clock::now();
    const auto size = film_->size();
    if (!gs_ || built_version_ != scene_->version()) {
        if (gs_) lmgpu_scene_destroy(gs_);
        gs_ = nullptr;
        gs_ = make_device_scene(*scene_, 4, 2, size.w, size.h);
        built_version_ = scene_->version();
    }
    lmgpu_raycast_params rp{};
    rp.bg_color[0] = float(bg_.x);
    rp.bg_color[1] = float(bg_.y);
    rp.bg_color[2] = float(bg_.z);
    if (color_) {
        rp.use_color = 1;
        rp.color[0] = float(color_->x);
        rp.color[1] = float(color_->y);
        rp.color[2] = float(color_->z);
    }
    rp.use_constant_color = constant_ ? 1 : 0;
    rp.visualize_normal = normal_ ? 1 : 0;
    film_->clear();
    check(lmgpu_raycast_render(gs_, &rp, film_->data().data(), 0, nullptr), "lmgpu_raycast_render");
    RenderResult r;
    r.processed = 1;
    r.elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

}  // namespace lm_mi355x

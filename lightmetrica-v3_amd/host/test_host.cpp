// Tests of the C++ host components on the GPU (run as a program; exit code 0
// when every check passes). Modelled on the reference's C++ unit tests
// (test/test_*.cpp, doctest): the accel against a brute-force FP64
// Moller-Trumbore closest hit, hit-to-primitive mapping with transforms,
// renderer::pt against the furnace closed form, renderer::raycast against a
// CPU ray cast of the same scene, property and error handling.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "lm_mi355x.h"

using namespace lm_mi355x;

static int g_failed = 0, g_checks = 0;
#define CHECK(cond)                                                                        \
    do {                                                                                   \
        g_checks++;                                                                        \
        if (!(cond)) {                                                                     \
            g_failed++;                                                                    \
            std::printf("  CHECK FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);           \
        }                                                                                  \
    } while (0)

template <class F>
static void run(const char* name, F f) {
    const int before = g_failed;
    std::printf("[ RUN ] %s\n", name);
    std::fflush(stdout);
    try {
        f();
    } catch (const Exception& e) {
        g_failed++;
        std::printf("  unexpected lm exception: %s\n", e.what());
    } catch (const std::exception& e) {
        g_failed++;
        std::printf("  unexpected exception: %s\n", e.what());
    }
    std::printf("[ %s ] %s\n", g_failed == before ? " OK " : "FAIL", name);
    std::fflush(stdout);
}

// ---- scene helpers ----------------------------------------------------------------

static std::shared_ptr<Mesh> quad_mesh(Vec3 a, Vec3 b, Vec3 c, Vec3 d) {
    auto m = std::make_shared<Mesh>();
    m->ps = {a, b, c, d};
    m->fs = {{{{0, -1}, {1, -1}, {2, -1}}}, {{{0, -1}, {2, -1}, {3, -1}}}};
    return m;
}

// axis-aligned unit cube [-0.5, 0.5]^3, outward faces, flat normals
static std::shared_ptr<Mesh> cube_mesh() {
    auto m = std::make_shared<Mesh>();
    for (int i = 0; i < 8; i++) m->ps.push_back({(i & 1) ? .5 : -.5, (i & 2) ? .5 : -.5, (i & 4) ? .5 : -.5});
    const int f[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
    for (auto& q : f) {
        m->fs.push_back({{{q[0], -1}, {q[1], -1}, {q[2], -1}}});
        m->fs.push_back({{{q[0], -1}, {q[2], -1}, {q[3], -1}}});
    }
    return m;
}

// closed box [-1,1]^3 with inward-facing walls
static std::shared_ptr<Mesh> closed_box() {
    auto m = std::make_shared<Mesh>();
    for (int i = 0; i < 8; i++) m->ps.push_back({(i & 1) ? 1. : -1., (i & 2) ? 1. : -1., (i & 4) ? 1. : -1.});
    const int f[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
    for (auto& q : f) {
        m->fs.push_back({{{q[0], -1}, {q[1], -1}, {q[2], -1}}});
        m->fs.push_back({{{q[0], -1}, {q[2], -1}, {q[3], -1}}});
    }
    return m;
}

// room (open front) + two transformed cubes + a ceiling light
static Scene room_scene(int w, int h) {
    Scene s;
    auto room = std::make_shared<Mesh>(*closed_box());
    room->fs.erase(room->fs.begin() + 2, room->fs.begin() + 4);   // drop the two z = +1 triangles: open front
    s.add_primitive(room, Material::diffuse({0.7, 0.7, 0.7}));
    s.add_primitive(cube_mesh(), Material::diffuse({0.8, 0.3, 0.2}), {},
                    Mat4::translate({-0.4, -0.7, -0.2}) * Mat4::rotate(0.5, {0, 1, 0}) * Mat4::scale({0.6, 0.6, 0.6}));
    s.add_primitive(cube_mesh(), Material::glossy({0.5, 0.6, 0.8}, 0.2, 0.2), {},
                    Mat4::translate({0.45, -0.75, 0.3}) * Mat4::rotate(-0.3, {0, 1, 0}) * Mat4::scale({0.5, 0.5, 0.5}));
    s.add_primitive(quad_mesh({-.25, .999, -.25}, {.25, .999, -.25}, {.25, .999, .25}, {-.25, .999, .25}),
                    Material::diffuse({0, 0, 0}), AreaLight{{17, 17, 17}});
    Camera c;
    c.position = {0, 0, 3.5};
    c.center = {0, 0, 0};
    c.up = {0, 1, 0};
    c.vfov = 40;
    c.aspect = double(w) / h;
    s.set_camera(c);
    return s;
}

// FP64 Moller-Trumbore closest hit over the flattened triangles
// (src/accel/accel_sahbvh.cpp:58-75, brute force); also reports how close
// the best hit is to an edge or to another candidate (ambiguous rays).
struct BruteHit {
    int tri = -1;
    double t = 0, u = 0, v = 0;
    bool ambiguous = false;
};

static BruteHit brute(const FlatScene& fl, Ray r, double tmin, double tmax) {
    BruteHit best;
    double second = 1e300;
    const size_t N = fl.tri_node.size();
    for (size_t i = 0; i < N; i++) {
        const float* v = &fl.verts[i * 9];
        const Vec3 p1{v[0], v[1], v[2]}, p2{v[3], v[4], v[5]}, p3{v[6], v[7], v[8]};
        const Vec3 e1 = p2 - p1, e2 = p3 - p1;
        const Vec3 p = cross(r.d, e2);
        const double det = dot(e1, p);
        if (std::abs(det) < 1e-8) continue;
        const Vec3 tv = r.o - p1;
        const double u = dot(tv, p) / det;
        const Vec3 q = cross(tv, e1);
        const double vv = dot(r.d, q) / det;
        const double t = dot(e2, q) / det;
        const double m = 1e-6;
        if (u < -m || vv < -m || u + vv > 1 + m || t < tmin || t > tmax) continue;
        const bool edge = u < m || vv < m || u + vv > 1 - m;
        if (best.tri < 0 || t < best.t) {
            second = best.tri < 0 ? second : best.t;
            best.tri = int(i);
            best.t = t;
            best.u = u;
            best.v = vv;
            best.ambiguous = edge;
        } else if (t < second) {
            second = t;
        }
    }
    if (best.tri >= 0 && second < 1e300 && std::abs(second - best.t) <= 1e-5 * (1 + best.t)) best.ambiguous = true;
    return best;
}

int main() {
    run("transform: normal matrix is the inverse transpose", [] {
        const Mat4 M = Mat4::translate({1, 2, 3}) * Mat4::rotate(0.7, {1, 1, 0}) * Mat4::scale({2, 0.5, 3});
        const Transform T(M);
        const Vec3 a{0.3, -0.2, 0.9}, b{-0.5, 0.4, 0.1};
        const Vec3 n = T.normal(cross(a, b));
        const Vec3 ta = T.point(a) - T.point({0, 0, 0}), tb = T.point(b) - T.point({0, 0, 0});
        CHECK(std::abs(dot(n, ta)) < 1e-12 && std::abs(dot(n, tb)) < 1e-12);
        CHECK(std::abs(T.J - 3.0) < 1e-12);
    });

    run("scene: traverse_primitive_nodes composes group transforms", [] {
        Scene s;
        const int g1 = s.create_group_node(Mat4::translate({1, 0, 0}));
        const int g2 = s.create_group_node(Mat4::scale({2, 2, 2}));
        const int p = s.create_primitive_node(quad_mesh({0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}),
                                              Material::diffuse({.5, .5, .5}));
        s.add_child(0, g1);
        s.add_child(g1, g2);
        s.add_child(g2, p);
        const FlatScene fl = s.flatten();
        CHECK(fl.tri_node.size() == 2 && fl.tri_node[0] == p && fl.tri_face[1] == 1);
        // vertex (1,1,0) of face 0 -> 2*(1,1,0) + (1,0,0) = (3,2,0)
        CHECK(fl.verts[6] == 3.f && fl.verts[7] == 2.f && fl.verts[8] == 0.f);
        bool threw = false;
        try {
            s.add_child(p, g1);
        } catch (const Exception& e) {
            threw = e.error() == Error::InvalidArgument;
        }
        CHECK(threw);
    });

    run("accel::mi355x closest hit = brute force (W 2/4/8)", [] {
        Scene s = room_scene(64, 64);
        const FlatScene fl = s.flatten();
        std::mt19937_64 rng(7);
        std::uniform_real_distribution<double> U(-0.95, 0.95), D(-1, 1);
        std::vector<Ray> rays(20000);
        for (auto& r : rays) {
            r.o = {U(rng), U(rng), U(rng)};
            Vec3 d;
            do d = {D(rng), D(rng), D(rng)};
            while (length(d) < 1e-3 || length(d) > 1);
            r.d = normalize(d);
        }
        for (int W : {2, 4, 8}) {
            AccelOptions o;
            o.width = W;
            Accel_MI355X acc(o);
            acc.build(s);
            const auto hits = acc.intersect_many(rays, 1e-3, 1e10);
            int mism = 0, amb = 0, terr = 0;
            for (size_t i = 0; i < rays.size(); i++) {
                // the accel sees the ray rounded to FP32 at the boundary
                const Ray& r0 = rays[i];
                const Ray rf{{float(r0.o.x), float(r0.o.y), float(r0.o.z)}, {float(r0.d.x), float(r0.d.y), float(r0.d.z)}};
                const BruteHit b = brute(fl, rf, double(float(1e-3)), 1e10);
                if (b.ambiguous) {
                    amb++;
                    continue;
                }
                if (b.tri < 0) {
                    if (hits[i]) mism++;
                    continue;
                }
                if (!hits[i] || hits[i]->primitive != fl.tri_node[size_t(b.tri)] || hits[i]->face != fl.tri_face[size_t(b.tri)]) {
                    mism++;
                    continue;
                }
                if (std::abs(hits[i]->t - b.t) > 1e-5 * b.t) terr++;
            }
            std::printf("  W%d: %zu rays, %d ambiguous, %d mismatches, %d t errors\n", W, rays.size(), amb, mism, terr);
            CHECK(mism == 0);
            CHECK(terr == 0);
        }
    });

    run("accel::mi355x Hit maps to (primitive, face) and the transformed surface point", [] {
        Scene s = room_scene(64, 64);
        Accel_MI355X acc;
        acc.build(s);
        int checked = 0, bad = 0;
        std::mt19937_64 rng(11);
        std::uniform_real_distribution<double> U(-0.9, 0.9);
        for (int i = 0; i < 2000; i++) {
            const Ray r{{U(rng), U(rng), 2.5}, normalize(Vec3{U(rng) * 0.3, U(rng) * 0.3, -1})};
            const auto h = acc.intersect(r, 1e-3, 1e10);
            if (!h) continue;
            const SceneNode& node = s.node_at(h->primitive);
            CHECK(node.type == SceneNodeType::Primitive);
            const Mesh::Tri tri = node.primitive.mesh->triangle_at(h->face);
            // surface point as Scene_::intersect: barycentric mix of the
            // local vertices, then the hit's global transform
            const Vec3 pl = tri.p1.p * (1 - h->uv.x - h->uv.y) + tri.p2.p * h->uv.x + tri.p3.p * h->uv.y;
            const Vec3 pw = h->global_transform.point(pl);
            const Vec3 pr = r.o + r.d * h->t;
            checked++;
            if (length(pw - pr) > 1e-4) bad++;
        }
        std::printf("  %d hits checked, %d off the ray\n", checked, bad);
        CHECK(checked > 1000 && bad == 0);
    });

    run("renderer::pt_mi355x furnace closed form (mis, nee, naive, guided)", [] {
        const double rho = 0.5, Le = 1.0;
        const int M = 8, spp = 64;
        Scene s;
        s.add_primitive(closed_box(), Material::diffuse({rho, rho, rho}), AreaLight{{Le, Le, Le}});
        Camera c;
        c.position = {0, 0, 0.5};
        c.center = {0, 0, -1};
        c.vfov = 60;
        s.set_camera(c);
        double expected = 0;
        for (int k = 0; k < M - 1; k++) expected += Le * std::pow(rho, k);
        for (const char* mode : {"mis", "nee", "naive", "guided"}) {
            Film film(32, 32);
            Props p{{"max_verts", M}, {"spp", spp}, {"seed", 5}, {"scheduler", std::string("sample")}};
            const bool guided = std::string(mode) == "guided";
            p["sampling_mode"] = std::string(guided ? "mis" : mode);
            if (guided) {
                p["guiding"] = true;
                p["guide_batch"] = 4096;
            }
            Renderer_PT_MI355X r(&s, &film, p);
            const RenderResult res = r.render();
            CHECK(res.processed == spp);
            const auto& sum = r.film_sum();
            const auto& sq = r.film_sumsq();
            double S = 0, Q = 0;
            for (size_t i = 0; i < sum.size(); i += 3) {
                S += sum[i];
                Q += sq[i];
            }
            const double n = double(film.num_pixels()) * spp;
            const double mean = S / n, var = Q / n - mean * mean;
            const double z = (mean - expected) / std::sqrt(var / n);
            std::printf("  %s: mean %.6f expected %.6f z %.2f (film(0,0) %.4f)\n", mode, mean, expected, z,
                        film.pixel(0, 0).x);
            CHECK(std::abs(z) < 5);
            if (guided) CHECK(res.guide_records > 0);
        }
        // primary_ray_sampling_mode = "image" (scheduler::spi::sample, a last partial pass):
        // film value sum * w*h / num_samples, the same furnace constant
        for (const char* mode : {"mis", "nee", "naive"}) {
            Film film(32, 32);
            const int64_t ns = int64_t(film.num_pixels()) * spp + 100;
            Props p{{"max_verts", M}, {"seed", 6}, {"scheduler", std::string("sample")},
                    {"primary_ray_sampling_mode", std::string("image")}, {"num_samples", double(ns)},
                    {"sampling_mode", std::string(mode)}};
            Renderer_PT_MI355X r(&s, &film, p);
            const RenderResult res = r.render();
            CHECK(res.processed == ns && r.film_sumsq().empty());
            double m = 0;
            for (int y = 0; y < 32; y++)
                for (int x = 0; x < 32; x++) m += film.pixel(x, y).x;
            m /= film.num_pixels();
            std::printf("  image mode %s: film mean %.6f expected %.6f\n", mode, m, expected);
            CHECK(std::abs(m - expected) < 0.03 * expected);
        }
    });

    run("renderer::pt_mi355x properties and errors", [] {
        Scene s = room_scene(16, 16);
        Film film(16, 16);
        auto expect = [&](const Props& p, Error e) {
            bool ok = false;
            try {
                Renderer_PT_MI355X r(&s, &film, p);
                r.render();
            } catch (const Exception& ex) {
                ok = ex.error() == e;
            }
            CHECK(ok);
        };
        expect({{"spp", 4}}, Error::InvalidArgument);                                       // no max_verts
        expect({{"max_verts", 8}, {"spp", 4}, {"sampling_mode", std::string("bdpt")}}, Error::InvalidArgument);
        expect({{"max_verts", 8}, {"scheduler", std::string("time")}}, Error::InvalidArgument);   // no render_time
        expect({{"max_verts", 1}, {"spp", 4}}, Error::InvalidArgument);                      // lmgpu rejects
        expect({{"max_verts", std::string("8")}, {"spp", 4}}, Error::InvalidArgument);      // wrong type
        expect({{"max_verts", 8}, {"spp", 4}, {"primary_ray_sampling_mode", std::string("film")}},
               Error::InvalidArgument);
        expect({{"max_verts", 8}, {"spp", 4}, {"primary_ray_sampling_mode", std::string("image")}},
               Error::InvalidArgument);                                                     // no num_samples
        expect({{"max_verts", 8}, {"num_samples", 100}, {"primary_ray_sampling_mode", std::string("image")},
                {"guiding", true}}, Error::Unsupported);
        // time scheduler
        Renderer_PT_MI355X r(&s, &film, {{"max_verts", 8}, {"scheduler", std::string("time")}, {"render_time", 0.2}});
        const RenderResult res = r.render();
        CHECK(res.processed >= 1 && res.elapsed >= 0.2);
        // rebuild after a scene change
        Renderer_PT_MI355X r2(&s, &film, {{"max_verts", 4}, {"spp", 1}});
        r2.render();
        s.add_primitive(cube_mesh(), Material::mirror(), {}, Mat4::translate({0, 0.5, 0}) * Mat4::scale({.2, .2, .2}));
        CHECK(r2.render().processed == 1);
    });

    run("renderer::lt_mi355x furnace (closed form) and properties", [] {
        // light paths of 1 .. M-2 reflections reach the camera: E = Le sum_{k=1}^{M-2} rho^k
        const double rho = 0.5, Le = 1.0;
        const int M = 8;
        Scene s;
        s.add_primitive(closed_box(), Material::diffuse({rho, rho, rho}), AreaLight{{Le, Le, Le}});
        Camera c;
        c.position = {0, 0, 0.5};
        c.center = {0, 0, -1};
        c.vfov = 60;
        s.set_camera(c);
        double expected = 0;
        for (int k = 1; k < M - 1; k++) expected += Le * std::pow(rho, k);
        Film film(32, 32);
        const int K = 16;
        std::vector<double> means;
        for (int k = 0; k < K; k++) {
            Renderer_LT_MI355X r(&s, &film, {{"max_verts", M}, {"scheduler", std::string("sample")},
                                             {"num_samples", 16 * 1024}, {"seed", 100 + k}});
            const RenderResult res = r.render();
            CHECK(res.processed == 16 * 1024);
            double m = 0;
            for (int y = 0; y < 32; y++)
                for (int x = 0; x < 32; x++) m += film.pixel(x, y).x;
            means.push_back(m / 1024);
        }
        double mu = 0, var = 0;
        for (double m : means) mu += m / K;
        for (double m : means) var += (m - mu) * (m - mu) / (K - 1);
        const double z = (mu - expected) / std::sqrt(var / K);
        std::printf("  LT furnace: mean %.6f expected %.6f z %.2f\n", mu, expected, z);
        CHECK(std::abs(z) < 5);
        auto expect = [&](const Props& p, Error e) {
            bool ok = false;
            try {
                Renderer_LT_MI355X r(&s, &film, p);
                r.render();
            } catch (const Exception& ex) {
                ok = ex.error() == e;
            }
            CHECK(ok);
        };
        expect({{"max_verts", 8}, {"num_samples", 100}}, Error::InvalidArgument);   // scheduler is required
        expect({{"max_verts", 8}, {"scheduler", std::string("sample")}}, Error::InvalidArgument);
        expect({{"max_verts", 1}, {"scheduler", std::string("sample")}, {"num_samples", 100}}, Error::InvalidArgument);
        Renderer_LT_MI355X rt(&s, &film, {{"max_verts", 8}, {"scheduler", std::string("time")}, {"render_time", 0.2}});
        const RenderResult res = rt.render();
        CHECK(res.processed >= 1024 && res.processed % 1024 == 0 && res.elapsed >= 0.2);
    });

    run("renderer::bdpt_mi355x furnace (closed form) and properties", [] {
        // all path lengths 2..M: E = Le sum_{k=0}^{M-2} rho^k
        const double rho = 0.5, Le = 1.0;
        const int M = 8;
        Scene s;
        s.add_primitive(closed_box(), Material::diffuse({rho, rho, rho}), AreaLight{{Le, Le, Le}});
        Camera c;
        c.position = {0, 0, 0.5};
        c.center = {0, 0, -1};
        c.vfov = 60;
        s.set_camera(c);
        double expected = 0;
        for (int k = 0; k < M - 1; k++) expected += Le * std::pow(rho, k);
        Film film(32, 32);
        const int K = 16;
        std::vector<double> means;
        for (int k = 0; k < K; k++) {
            Renderer_BDPT_MI355X r(&s, &film, {{"max_verts", M}, {"scheduler", std::string("sample")},
                                               {"num_samples", 4 * 1024}, {"seed", 200 + k}});
            CHECK(r.render().processed == 4 * 1024);
            double m = 0;
            for (int y = 0; y < 32; y++)
                for (int x = 0; x < 32; x++) m += film.pixel(x, y).x;
            means.push_back(m / 1024);
        }
        double mu = 0, var = 0;
        for (double m : means) mu += m / K;
        for (double m : means) var += (m - mu) * (m - mu) / (K - 1);
        const double z = (mu - expected) / std::sqrt(var / K);
        std::printf("  BDPT furnace: mean %.6f expected %.6f z %.2f\n", mu, expected, z);
        CHECK(std::abs(z) < 5);
        bool ok = false;
        try {
            Renderer_BDPT_MI355X r(&s, &film, {{"max_verts", 8}, {"num_samples", 100}});
        } catch (const Exception& ex) {
            ok = ex.error() == Error::InvalidArgument;   // scheduler is required
        }
        CHECK(ok);
    });

    run("renderer::raycast_mi355x = CPU ray cast of the same scene", [] {
        const int w = 96, h = 64;
        Scene s = room_scene(w, h);
        for (int mode = 0; mode < 3; mode++) {
            Film film(w, h);
            Props p{{"bg_color", Vec3{0.1, 0.2, 0.3}}};
            if (mode == 1) p["visualize_normal"] = true;
            if (mode == 2) {
                p["use_constant_color"] = true;
                p["color"] = Vec3{0.4, 0.5, 0.6};
            }
            Renderer_Raycast_MI355X r(&s, &film, p);
            r.render();
            const FlatScene fl = s.flatten();
            int bad = 0, amb = 0;
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    const Ray ray = s.camera().primary_ray({(x + .5) / w, (y + .5) / h});
                    const BruteHit b = brute(fl, ray, 1e-3, 1e10);
                    if (b.ambiguous) {
                        amb++;
                        continue;
                    }
                    Vec3 C{0.1, 0.2, 0.3};
                    if (b.tri >= 0) {
                        const float* nv = &fl.normals[size_t(b.tri) * 9];
                        const Vec3 n = normalize(Vec3{nv[0], nv[1], nv[2]} * (1 - b.u - b.v) +
                                                 Vec3{nv[3], nv[4], nv[5]} * b.u + Vec3{nv[6], nv[7], nv[8]} * b.v);
                        const Material& m = s.node_at(fl.tri_node[size_t(b.tri)]).primitive.material;
                        if (mode == 1) C = {std::abs(n.x), std::abs(n.y), std::abs(n.z)};
                        else if (mode == 2) C = {0.4, 0.5, 0.6};
                        else C = m.reflectance() * (.2 + .8 * std::abs(dot(n, -ray.d)));
                    }
                    const Vec3 g = film.pixel(x, y);
                    if (length(g - C) > 1e-4) bad++;
                }
            std::printf("  mode %d: %d ambiguous, %d differing pixels of %d\n", mode, amb, bad, w * h);
            CHECK(bad == 0);
        }
    });

    std::printf("host tests: %d checks, %d failed\n", g_checks, g_failed);
    return g_failed == 0 ? 0 : 1;
}

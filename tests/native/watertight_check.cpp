// Host check of the watertight ray/triangle test (csrc/bvh.h, ray_prepare +
// tri_intersect): rays fired from outside a closed mesh toward its centre,
// exactly through shared edges (FP32 points on each edge, its end points
// included) and through shared vertices, must hit at least one triangle.
// The mesh is an icosphere (level 3, 1280 triangles) under several random
// rotations, scales and offsets, so the edges are neither axis-aligned nor
// at convenient magnitudes. The program is compiled with FMA instructions
// enabled and HIP's default contraction (fused across statements unless a
// pragma says otherwise), as the device code is, so a fused a*b - c*d in
// the edge functions (which breaks the antisymmetry between neighbouring
// triangles) shows up as misses: without the contraction pragma in
// tri_intersect a ray through an edge of the 0.05-scale sphere misses both
// triangles. Prints "ok <rays>" or the first miss and exits non-zero.
#include <array>
#include <cmath>
#include <cstdio>
#include <map>
#include <random>
#include <vector>

#include "../../lightmetrica-v3_amd/csrc/bvh.h"

using lmgpu::f3;
using lmgpu::mk;

struct Mesh {
    std::vector<std::array<double, 3>> v;
    std::vector<std::array<int, 3>> f;
};

static Mesh icosphere(int levels) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    Mesh m;
    m.v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
           {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    m.f = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
           {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
           {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (auto& p : m.v) {
        const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (double& c : p) c /= l;
    }
    for (int it = 0; it < levels; it++) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto found = mid.find(key);
            if (found != mid.end()) return found->second;
            std::array<double, 3> p;
            for (int c = 0; c < 3; c++) p[c] = 0.5 * (m.v[a][c] + m.v[b][c]);
            const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            for (double& c : p) c /= l;
            m.v.push_back(p);
            return mid[key] = int(m.v.size()) - 1;
        };
        std::vector<std::array<int, 3>> nf;
        for (const auto& tri : m.f) {
            const int a = midpoint(tri[0], tri[1]), b = midpoint(tri[1], tri[2]), c = midpoint(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        m.f = nf;
    }
    return m;
}

int main() {
    const Mesh base = icosphere(3);
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double scales[] = {1.0, 0.05, 250.0};
    const double offsets[] = {0.0, 7.0, -1500.0};
    long rays = 0;
    for (int trial = 0; trial < 9; trial++) {
        const double scale = scales[trial % 3], offset = offsets[trial / 3];
        // random rotation (normalised quaternion)
        double q[4];
        double ql = 0;
        for (double& c : q) { c = U(rng) - 0.5; ql += c * c; }
        for (double& c : q) c /= std::sqrt(ql);
        const double w = q[0], x = q[1], y = q[2], z = q[3];
        const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                                {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                                {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
        const double centre[3] = {offset, 0.5 * offset, -0.25 * offset};
        std::vector<std::array<float, 3>> v(base.v.size());
        for (size_t i = 0; i < base.v.size(); i++)
            for (int r = 0; r < 3; r++)
                v[i][r] = float(centre[r] + scale * (R[r][0] * base.v[i][0] + R[r][1] * base.v[i][1] +
                                                     R[r][2] * base.v[i][2]));
        std::vector<float4> tv(3 * base.f.size());
        for (size_t i = 0; i < base.f.size(); i++)
            for (int k = 0; k < 3; k++) {
                const auto& p = v[base.f[i][k]];
                tv[3 * i + k] = make_float4(p[0], p[1], p[2], 0.f);
            }
        // targets: the vertices, and FP32 points along every edge (end points
        // of the parameter range included, where the edge meets the vertex)
        std::vector<std::array<float, 3>> targets;
        for (const auto& p : v) targets.push_back(p);
        const float ss[] = {0.5f, 0.25f, 0.125f, 1e-3f, 1e-6f, 1.0f / 3.0f};
        for (const auto& tri : base.f)
            for (int k = 0; k < 3; k++) {
                const auto& a = v[tri[k]];
                const auto& b = v[tri[(k + 1) % 3]];
                if (tri[k] > tri[(k + 1) % 3]) continue;   // each edge once
                for (float s : ss) {
                    std::array<float, 3> p;
                    for (int c = 0; c < 3; c++) p[c] = a[c] + s * (b[c] - a[c]);
                    targets.push_back(p);
                }
            }
        for (const auto& p : targets) {
            // origin outside, on the line from the centre through the target;
            // the direction points back at the centre through the target
            double dir[3], dl = 0;
            for (int c = 0; c < 3; c++) { dir[c] = double(p[c]) - centre[c]; dl += dir[c] * dir[c]; }
            dl = std::sqrt(dl);
            f3 o, d;
            o = mk(float(double(p[0]) + 2.0 * scale * dir[0] / dl), float(double(p[1]) + 2.0 * scale * dir[1] / dl),
                   float(double(p[2]) + 2.0 * scale * dir[2] / dl));
            double e[3] = {double(p[0]) - double(o.x), double(p[1]) - double(o.y), double(p[2]) - double(o.z)};
            const double el = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            d = mk(float(e[0] / el), float(e[1] / el), float(e[2] / el));
            const lmgpu::RayPre r = lmgpu::ray_prepare(o, d);
            bool hit = false;
            for (size_t i = 0; i < base.f.size() && !hit; i++) {
                float t, uu, vv;
                hit = lmgpu::tri_intersect(r, tv[3 * i], tv[3 * i + 1], tv[3 * i + 2], 0.f, 1e30f, t, uu, vv);
            }
            rays++;
            if (!hit) {
                std::printf("miss: trial %d (scale %g, offset %g) target %.9g %.9g %.9g\n", trial, scale, offset,
                            p[0], p[1], p[2]);
                return 1;
            }
        }
    }
    std::printf("ok %ld\n", rays);
    return 0;
}

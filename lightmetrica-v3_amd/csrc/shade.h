// Device (FP32) restatement of the reference's surface, camera, material and
// light routines used by renderer::pt.
//
//   PointGeometry::make_on_surface, opposite, orthonormal_basis_twosided
//                                   include/lm/surface.h:124-195
//   surface::geometry_term, convert_pdf_SA_to_projSA, shading_normal_correction
//                                   include/lm/surface.h:329-438
//   math::orthonormal_basis (Duff et al. 2017), reflection, refraction,
//   sample_cosine_weighted, sample_uniform_sphere, balance_heuristic
//                                   include/lm/math.h:497-702
//   camera::pinhole primary_ray     src/camera/camera_pinhole.cpp:128-132
//   material::diffuse               src/material/material_diffuse.cpp:60-95
//   material::glossy (aniso GGX)    src/material/material_glossy.cpp:43-113
//   material::mirror                src/material/material_mirror.cpp:33-62
//   material::glass (Schlick)       src/material/material_glass.cpp:85-286
//   material::mixture_wavefrontobj  src/model/model_wavefrontobj.cpp:340-513
//   light::area, light::envconst    src/light/light_area.cpp:40-187,
//                                   src/light/light_envconst.cpp:100-141
//   Scene::sample_light_selection   src/scene.cpp:374-387 (uniform)
#pragma once

#include "common.h"

namespace lmgpu {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kEps = 1e-3f;   // include/lm/math.h:55
constexpr float kInf = 1e10f;   // include/lm/math.h:54

struct Geom {
    f3 p, n, gn, u, v;
};

__device__ __forceinline__ void onb(f3 n, f3& u, f3& v) {
    const float s = copysignf(1.0f, n.z);
    const float a = -1.0f / (s + n.z);
    const float b = n.x * n.y * a;
    u = mk(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
    v = mk(b, s + n.y * n.y * a, -n.y);
}

__device__ __forceinline__ Geom make_on_surface(f3 p, f3 n, f3 gn) {
    Geom g;
    g.p = p; g.n = n; g.gn = gn;
    onb(n, g.u, g.v);
    return g;
}

__device__ __forceinline__ bool opposite(const Geom& g, f3 a, f3 b) { return dot(a, g.n) * dot(b, g.n) <= 0.f; }

__device__ __forceinline__ void twosided(const Geom& g, f3 wi, f3& n, f3& u, f3& v) {
    const bool i = dot(wi, g.n) > 0.f;
    n = i ? g.n : -g.n;
    u = g.u;
    v = i ? g.v : -g.v;
}

// TransDir::EL: 0 if wi or wo sees the geometric and shading normals with
// different signs, else 1.
__device__ __forceinline__ float sn_correction(const Geom& g, f3 wi, f3 wo) {
    const float wi_ng = dot(wi, g.gn), wo_ng = dot(wo, g.gn);
    const float wi_ns = dot(wi, g.n), wo_ns = dot(wo, g.n);
    return (wi_ng * wi_ns <= 0.f || wo_ng * wo_ns <= 0.f) ? 0.f : 1.f;
}

__device__ __forceinline__ f3 reflection(f3 w, f3 n) { return 2.0f * dot(w, n) * n - w; }

__device__ __forceinline__ bool refraction(f3 wi, f3 n, float eta, f3& wt) {
    const float t = dot(wi, n);
    const float t2 = 1.f - eta * eta * (1.f - t * t);
    if (t2 <= 0.f) return false;
    wt = eta * (n * t - wi) - n * sqrtf(fmaxf(0.f, t2));
    return true;
}

__device__ __forceinline__ f3 sample_cosine_weighted(float u0, float u1) {
    const float r = sqrtf(fmaxf(0.f, u0));
    const float t = 2.f * kPi * u1;
    float s, c;
    sincosf(t, &s, &c);
    const float x = r * c, y = r * s;
    return mk(x, y, sqrtf(fmaxf(0.f, 1.f - x * x - y * y)));
}

__device__ __forceinline__ f3 sample_uniform_sphere(float u0, float u1) {
    const float z = 1.f - 2.f * u0;
    const float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    const float t = 2.f * kPi * u1;
    float s, c;
    sincosf(t, &s, &c);
    return mk(r * c, r * s, z);
}

__device__ __forceinline__ float balance_heuristic(float p1, float p2) {
    if (p1 == 0.f && p2 == 0.f) return 0.f;
    return p1 / (p1 + p2);
}

// ---- materials ---------------------------------------------------------------

struct Mat {
    int type;       // LMGPU_MAT_*
    f3 c;           // Kd (diffuse, mixture: map_Kd's texel when textured) or Ks (glossy)
    float ax, ay, ni;
    f3 ks;          // mixture: Ks of the glossy lobe
    float alpha;    // mixture: alpha mask value at the hit (1 without a mask)
};

__device__ __forceinline__ float sq(float x) { return x * x; }

__device__ __forceinline__ float ggx_D(const Mat& m, f3 wh, f3 u, f3 v, f3 n) {
    return 1.f / (kPi * m.ax * m.ay * sq(sq(dot(wh, u) / m.ax) + sq(dot(wh, v) / m.ay) + sq(dot(wh, n))));
}

__device__ __forceinline__ float ggx_G1(const Mat& m, f3 w, f3 u, f3 v, f3 n) {
    const float c = dot(w, n);
    const float s = fmaxf(kEps, sqrtf(fmaxf(0.f, 1.f - c * c)));
    const float cp = dot(w, u) / s;
    const float cs = dot(w, v) / s;
    const float a2 = sq(cp * m.ax) + sq(cs * m.ay);
    return c == 0.f ? 0.f : 2.f / (1.f + sqrtf(fmaxf(0.f, 1.f + a2 * sq(s / c))));
}

__device__ __forceinline__ float glossy_pdf(const Mat& m, const Geom& g, f3 wi, f3 wo) {
    if (opposite(g, wi, wo)) return 0.f;
    const f3 wh = normalize(wi + wo);
    f3 n, u, v;
    twosided(g, wi, n, u, v);
    return ggx_D(m, wh, u, v, n) * dot(wh, n) / (4.f * dot(wo, wh) * dot(wo, n));
}

__device__ __forceinline__ f3 glossy_eval(const Mat& m, const Geom& g, f3 wi, f3 wo) {
    if (opposite(g, wi, wo)) return mk(0, 0, 0);
    const f3 wh = normalize(wi + wo);
    f3 n, u, v;
    twosided(g, wi, n, u, v);
    const float x = 1.f - dot(wo, wh);
    const float f5 = x * x * x * x * x;
    const f3 Fr = m.c + (mk(1, 1, 1) - m.c) * f5;
    // Ks appears twice (material_glossy.cpp:107-108)
    return m.c * Fr *
           (ggx_D(m, wh, u, v, n) * ggx_G1(m, wi, u, v, n) * ggx_G1(m, wo, u, v, n) / (4.f * dot(wi, n) * dot(wo, n)));
}

__device__ __forceinline__ float glass_Fr(const Mat& m, const Geom& g, f3 wi, f3& wt, float& eta) {
    const bool in = dot(wi, g.n) > 0.f;
    const f3 n = in ? g.n : -g.n;
    eta = in ? 1.f / m.ni : m.ni;
    if (!refraction(wi, n, eta, wt)) return 1.f;
    const float cs = in ? dot(wi, g.n) : dot(wt, g.n);
    const float r = (1.f - m.ni) / (1.f + m.ni);
    const float r2 = r * r;
    const float x = 1.f - cs;
    return r2 + (1.f - r2) * (x * x * x * x * x);
}

// Material::is_specular_component: mirror and glass always, the mixture's
// alpha pass-through component (1).
__device__ __forceinline__ bool is_specular(const Mat& m, int comp) {
    return m.type == LMGPU_MAT_MIRROR || m.type == LMGPU_MAT_GLASS || (m.type == LMGPU_MAT_MIXTURE && comp != 0);
}

// The glossy lobe of a mixture as a material::glossy.
__device__ __forceinline__ Mat mixture_glossy(const Mat& m) {
    Mat gm = m;
    gm.type = LMGPU_MAT_GLOSSY;
    gm.c = m.ks;
    return gm;
}

__device__ __forceinline__ f3 diffuse_eval(const Mat& m, const Geom& g, f3 wi, f3 wo) {
    return opposite(g, wi, wo) ? mk(0, 0, 0) : m.c * (1.f / kPi);
}

// Probability of the diffuse lobe of a mixture (diffuse_selection_prob):
// max(Kd) / (max(Kd) + max(Ks)), 1 when both are 0.
__device__ __forceinline__ float mixture_diffuse_prob(const Mat& m) {
    const float wd = maxc(m.c), wg = maxc(m.ks);
    return (wd == 0.f && wg == 0.f) ? 1.f : wd / (wd + wg);
}

// Material::eval of component `comp` with eval_delta = true (renderer_pt.cpp:162).
__device__ __forceinline__ f3 mat_eval(const Mat& m, const Geom& g, f3 wi, f3 wo, int comp) {
    if (m.type == LMGPU_MAT_DIFFUSE) return diffuse_eval(m, g, wi, wo);
    if (m.type == LMGPU_MAT_GLOSSY) return glossy_eval(m, g, wi, wo);
    if (m.type == LMGPU_MAT_MIXTURE && comp == 0)
        return (diffuse_eval(m, g, wi, wo) + glossy_eval(mixture_glossy(m), g, wi, wo)) * m.alpha;
    return mk(0, 0, 0);
}

// Material::pdf_direction of component `comp` with eval_delta = true,
// projected solid angle (the mixture's is the lobe-selection marginal).
__device__ __forceinline__ float mat_pdf(const Mat& m, const Geom& g, f3 wi, f3 wo, int comp) {
    if (m.type == LMGPU_MAT_DIFFUSE) return opposite(g, wi, wo) ? 0.f : 1.f / kPi;
    if (m.type == LMGPU_MAT_GLOSSY) return glossy_pdf(m, g, wi, wo);
    if (m.type == LMGPU_MAT_MIXTURE && comp == 0) {
        const float w = mixture_diffuse_prob(m);
        return w * (opposite(g, wi, wo) ? 0.f : 1.f / kPi) + (1.f - w) * glossy_pdf(mixture_glossy(m), g, wi, wo);
    }
    return 0.f;
}

// Material::sample_component: glass picks reflection (0) / refraction (1)
// with probability Fr / 1-Fr, the mixture its diffuse+glossy group (0) /
// alpha pass-through (1) with probability alpha / 1-alpha, each with weight
// 1/p; the others {0, 1}. `urest` is uc rescaled to [0,1) within the chosen
// interval, independent of the choice (the mixture's lobe selection).
__device__ __forceinline__ int mat_sample_component(const Mat& m, const Geom& g, f3 wi, float uc, float& w,
                                                    float& urest) {
    w = 1.f;
    urest = uc;
    if (m.type == LMGPU_MAT_MIXTURE) {
        if (uc < m.alpha) {
            w = 1.f / m.alpha;
            urest = uc / m.alpha;
            return 0;
        }
        w = 1.f / (1.f - m.alpha);
        urest = (uc - m.alpha) / (1.f - m.alpha);
        return 1;
    }
    if (m.type != LMGPU_MAT_GLASS) return 0;
    f3 wt;
    float eta;
    const float Fr = glass_Fr(m, g, wi, wt, eta);
    const int comp = uc < Fr ? 0 : 1;
    w = 1.f / (comp == 0 ? Fr : 1.f - Fr);
    return comp;
}

// Material::sample_direction (without the shading-normal correction). `ul`
// selects the mixture's lobe.
__device__ __forceinline__ bool mat_sample_direction(const Mat& m, const Geom& g, f3 wi, int comp, float u0, float u1,
                                                     float ul, f3& wo, f3& weight);

__device__ __forceinline__ bool mat_sample_direction_basic(const Mat& m, const Geom& g, f3 wi, int comp, float u0,
                                                           float u1, f3& wo, f3& weight) {
    if (m.type == LMGPU_MAT_DIFFUSE) {
        f3 n, u, v;
        twosided(g, wi, n, u, v);
        const f3 d = sample_cosine_weighted(u0, u1);
        wo = u * d.x + v * d.y + n * d.z;
        weight = m.c;
        return true;
    }
    if (m.type == LMGPU_MAT_GLOSSY) {
        f3 n, u, v;
        twosided(g, wi, n, u, v);
        const float a = u0 * 2.f * kPi;
        float s, c;
        sincosf(a, &s, &c);
        const float r = sqrtf(fmaxf(0.f, u1 / (1.f - u1)));
        const f3 wh = normalize(r * (m.ax * c * u + m.ay * s * v) + n);
        wo = reflection(wi, wh);
        if (opposite(g, wi, wo)) return false;
        const f3 f = glossy_eval(m, g, wi, wo);
        const float p = glossy_pdf(m, g, wi, wo);
        weight = f / p;
        return true;
    }
    if (m.type == LMGPU_MAT_MIRROR) {
        wo = reflection(wi, g.n);
        weight = mk(1, 1, 1);
        return true;
    }
    // glass
    f3 wt;
    float eta;
    const float Fr = glass_Fr(m, g, wi, wt, eta);
    if (comp == 0) {
        wo = reflection(wi, g.n);
        weight = mk(Fr, Fr, Fr);
    } else {
        wo = wt;
        const float w = (1.f - Fr) * eta * eta;
        weight = mk(w, w, w);
    }
    return true;
}

__device__ __forceinline__ bool mat_sample_direction(const Mat& m, const Geom& g, f3 wi, int comp, float u0, float u1,
                                                     float ul, f3& wo, f3& weight) {
    if (m.type != LMGPU_MAT_MIXTURE) return mat_sample_direction_basic(m, g, wi, comp, u0, u1, wo, weight);
    if (comp != 0) {
        // alpha pass-through; the reference returns weight 1 here
        wo = -wi;
        weight = mk(1, 1, 1);
        return true;
    }
    // pick the diffuse or the glossy lobe, sample it, weight = f / marginal pdf
    Mat lobe = m;
    if (ul < mixture_diffuse_prob(m)) lobe.type = LMGPU_MAT_DIFFUSE;
    else lobe = mixture_glossy(m);
    f3 lw;
    if (!mat_sample_direction_basic(lobe, g, wi, 0, u0, u1, wo, lw)) return false;
    const float p = mat_pdf(m, g, wi, wo, 0);
    weight = p > 0.f ? mat_eval(m, g, wi, wo, 0) / p : mk(0, 0, 0);
    return true;
}

// Material::eval with eval_delta = false (path-space estimators): specular
// components evaluate to their sampled weight (mirror 1, glass Fr or
// (1 - Fr) with eta^2 for radiance transport only, material_glass.cpp:85-87,
// 231-286), the mixture's pass-through to 1 - alpha.
__device__ __forceinline__ f3 mat_eval_full(const Mat& m, const Geom& g, f3 wi, f3 wo, int comp, bool le) {
    if (m.type == LMGPU_MAT_MIRROR) return mk(1, 1, 1);
    if (m.type == LMGPU_MAT_GLASS) {
        f3 wt;
        float eta;
        const float Fr = glass_Fr(m, g, wi, wt, eta);
        const float w = comp == 0 ? Fr : (1.f - Fr) * (le ? 1.f : eta * eta);
        return mk(w, w, w);
    }
    if (m.type == LMGPU_MAT_MIXTURE && comp != 0) return mk(1.f - m.alpha, 1.f - m.alpha, 1.f - m.alpha);
    return mat_eval(m, g, wi, wo, comp);
}

// Material::pdf_direction with eval_delta = false: 1 for specular components.
__device__ __forceinline__ float mat_pdf_full(const Mat& m, const Geom& g, f3 wi, f3 wo, int comp) {
    if (is_specular(m, comp)) return 1.f;
    return mat_pdf(m, g, wi, wo, comp);
}

// ---- camera ------------------------------------------------------------------

struct Camera {
    f3 pos, u, v, w;
    float tf, aspect;
};

__device__ __forceinline__ f3 primary_dir(const Camera& c, float rx, float ry) {
    rx = 2.f * rx - 1.f;
    ry = 2.f * ry - 1.f;
    const f3 d = normalize(mk(c.aspect * c.tf * rx, c.tf * ry, -1.f));
    return c.u * d.x + c.v * d.y + c.w * d.z;
}

// camera::pinhole: raster position of the direction `woc` (camera -> point)
// and its importance J (camera_pinhole.cpp:102-109, 207-226).
__device__ __forceinline__ bool cam_raster(const Camera& cam, f3 woc, float& rx, float& ry) {
    const float ez = dot(woc, cam.w);
    if (!(ez < 0.f)) return false;
    rx = -dot(woc, cam.u) / ez / cam.tf / cam.aspect * .5f + .5f;
    ry = -dot(woc, cam.v) / ez / cam.tf * .5f + .5f;
    return rx >= 0.f && rx <= 1.f && ry >= 0.f && ry <= 1.f;
}
__device__ __forceinline__ float cam_J(const Camera& cam, f3 woc) {
    const float ct = -dot(woc, cam.w);
    return 1.f / (ct * ct * ct * (4.f * cam.tf * cam.tf * cam.aspect));
}

// Film::raster_to_pixel (film.h:166-171): pixel index y * w + x.
__device__ __forceinline__ int raster_pixel(int w, int h, float rx, float ry) {
    const int x = min(max(int(rx * float(w)), 0), w - 1);
    const int y = min(max(int(ry * float(h)), 0), h - 1);
    return y * w + x;
}

}  // namespace lmgpu

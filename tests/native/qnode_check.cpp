// Host check of the quantised BVH nodes (csrc/bvh.h, quantise_node /
// qplane / child_keys): for random 4- and 8-wide nodes over many scales and
// offsets, every decoded child box contains the original one, the grid is
// tight (within one step plus rounding of the exact planes), and stack
// entries and empty slots are encoded as the traversal expects. Then the
// quantised slab test (child_keys, the same FP32 code the traversal runs)
// is checked to be conservative: for rays aimed at points of a child box,
// grazing its faces, edges and corners, starting on its faces or running
// parallel to them, every child whose exact box the ray meets in
// [tmin, tmax] (decided in exact arithmetic: __float128 products of the FP32
// inputs) is hit. Prints "ok <nodes> <rays>" or the first failure and exits
// non-zero.
#include <cstdio>
#include <random>

#include "../../lightmetrica-v3_amd/csrc/bvh.h"

using lmgpu::QNode;
using lmgpu::WNode;

template <int W>
int check(std::mt19937_64& rng, int n_nodes) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (int it = 0; it < n_nodes; it++) {
        WNode<W> n;
        const double offset = (U(rng) - 0.5) * std::pow(10.0, int(U(rng) * 9) - 4);   // up to +-5e4
        const double extent = std::pow(10.0, U(rng) * 10 - 7);                          // 1e-7 .. 1e3
        const int nvalid = 1 + int(U(rng) * W);
        for (int c = 0; c < W; c++) {
            float* lo[3] = {&n.lox[c], &n.loy[c], &n.loz[c]};
            float* hi[3] = {&n.hix[c], &n.hiy[c], &n.hiz[c]};
            for (int a = 0; a < 3; a++) {
                double l = offset + U(rng) * extent, h = l + U(rng) * extent * (U(rng) < 0.1 ? 0.0 : 1.0);
                *lo[a] = float(l);
                *hi[a] = std::max(float(h), *lo[a]);
            }
            if (c < nvalid) {
                n.count[c] = U(rng) < 0.5 ? 0 : 1 + int(U(rng) * 15);
                n.child[c] = int(U(rng) * 1e6);
            } else {
                n.count[c] = -1;
                n.child[c] = 0;
            }
        }
        const QNode<W> q = lmgpu::quantise_node<W>(n);
        const float org[3] = {q.ox, q.oy, q.oz};
        for (int c = 0; c < W; c++) {
            const int want = n.count[c] < 0 ? lmgpu::kQEmpty
                             : n.count[c] == 0 ? n.child[c] : lmgpu::encode_leaf(n.child[c], n.count[c]);
            if (q.ent[c] != want) {
                std::printf("W%d node %d child %d: entry %d, expected %d\n", W, it, c, q.ent[c], want);
                return 1;
            }
            if (n.count[c] < 0) continue;
            const float lo[3] = {n.lox[c], n.loy[c], n.loz[c]}, hi[3] = {n.hix[c], n.hiy[c], n.hiz[c]};
            for (int a = 0; a < 3; a++) {
                const float step = std::ldexp(1.f, int((q.e >> (8 * a)) & 0xffu) - 127);
                const float dl = lmgpu::qplane(q.qlo[a][c / 4], c % 4, step, org[a]);
                const float dh = lmgpu::qplane(q.qhi[a][c / 4], c % 4, step, org[a]);
                if (!(dl <= lo[a] && dh >= hi[a])) {
                    std::printf("W%d node %d child %d axis %d: decoded [%.9g, %.9g] misses [%.9g, %.9g]\n", W, it, c,
                                a, dl, dh, lo[a], hi[a]);
                    return 1;
                }
                // tight: at most one grid step (plus the rounding of the decoded plane) outside
                const double slack = double(step) + 4.0 * std::fabs(double(org[a])) * 1.2e-7;
                if (double(lo[a]) - dl > slack || dh - double(hi[a]) > slack) {
                    std::printf("W%d node %d child %d axis %d: loose decode [%.9g, %.9g] of [%.9g, %.9g], step %g\n",
                                W, it, c, a, dl, dh, lo[a], hi[a], step);
                    return 1;
                }
            }
        }
    }
    return 0;
}

// Exact slab test of the box [lo, hi] for the ray o + t d, t in [tmin, tmax]
// (d as the traversal sees it: components below 1e-30 in magnitude replaced
// by +-1e-30). Entry and exit distances are fractions n / d of exact FP32
// differences; t1 <= t2 is decided by the sign of n1 d2 - n2 d1 (times the
// sign of d1 d2), whose two products are exact in __float128.
static bool exact_hit(const float o[3], const float d[3], const float lo[3], const float hi[3], float tmin,
                      float tmax) {
    using Q = __float128;
    Q nn[3], nf[3], dd[3];
    for (int a = 0; a < 3; a++) {
        const float da = std::fabs(d[a]) < 1e-30f ? std::copysign(1e-30f, d[a]) : d[a];
        dd[a] = Q(da);
        const float pn = da > 0 ? lo[a] : hi[a], pf = da > 0 ? hi[a] : lo[a];
        nn[a] = Q(pn) - Q(o[a]);
        nf[a] = Q(pf) - Q(o[a]);
    }
    // t1 = n1 / d1 <= t2 = n2 / d2
    auto le = [](Q n1, Q d1, Q n2, Q d2) {
        const Q x = n1 * d2 - n2 * d1;
        return (d1 * d2 > 0) ? x <= 0 : x >= 0;
    };
    for (int a = 0; a < 3; a++) {
        for (int b = 0; b < 3; b++)
            if (!le(nn[a], dd[a], nf[b], dd[b])) return false;
        // tmin <= far_a and near_a <= tmax
        if (!le(Q(tmin), Q(1), nf[a], dd[a])) return false;
        if (tmax < __builtin_inff() && !le(nn[a], dd[a], Q(tmax), Q(1))) return false;
    }
    return true;
}

template <int W>
int check_slab(std::mt19937_64& rng, int n_nodes, int rays_per_node, long long& n_rays, long long& n_exact,
               long long& n_extra) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (int it = 0; it < n_nodes; it++) {
        WNode<W> n;
        const double offset = (U(rng) - 0.5) * std::pow(10.0, int(U(rng) * 9) - 4);
        const double extent = std::pow(10.0, U(rng) * 10 - 7);
        for (int c = 0; c < W; c++) {
            float* lo[3] = {&n.lox[c], &n.loy[c], &n.loz[c]};
            float* hi[3] = {&n.hix[c], &n.hiy[c], &n.hiz[c]};
            for (int a = 0; a < 3; a++) {
                double l = offset + U(rng) * extent, h = l + U(rng) * extent * (U(rng) < 0.1 ? 0.0 : 1.0);
                *lo[a] = float(l);
                *hi[a] = std::max(float(h), *lo[a]);
            }
            n.count[c] = U(rng) < 0.5 ? 0 : 1 + int(U(rng) * 15);
            n.child[c] = int(U(rng) * 1e6);
        }
        const QNode<W> q = lmgpu::quantise_node<W>(n);
        for (int k = 0; k < rays_per_node; k++) {
            const int c = int(U(rng) * W) % W;
            const float lo[3] = {n.lox[c], n.loy[c], n.loz[c]}, hi[3] = {n.hix[c], n.hiy[c], n.hiz[c]};
            float target[3], o[3], d[3];
            for (int a = 0; a < 3; a++) {
                // a face, edge or corner of the box, or a point inside
                const double v = U(rng);
                target[a] = v < 0.3 ? lo[a] : v < 0.6 ? hi[a] : float(lo[a] + U(rng) * (double(hi[a]) - lo[a]));
            }
            const double mode = U(rng);
            const double dist = extent * std::pow(10.0, U(rng) * 6 - 2);
            for (int a = 0; a < 3; a++) o[a] = float(target[a] + (U(rng) - 0.5) * 2 * dist);
            if (mode < 0.2) {
                // the origin on one of the box's planes
                const int a = int(U(rng) * 3) % 3;
                o[a] = U(rng) < 0.5 ? lo[a] : hi[a];
            }
            for (int a = 0; a < 3; a++) d[a] = target[a] - o[a];
            if (mode > 0.8) {
                // parallel to an axis plane, inside, on or outside the slab
                const int a = int(U(rng) * 3) % 3;
                d[a] = 0.f;
                const double v = U(rng);
                o[a] = v < 0.3 ? lo[a] : v < 0.6 ? hi[a] : target[a];
            }
            if (d[0] == 0.f && d[1] == 0.f && d[2] == 0.f) continue;
            const float tmin = U(rng) < 0.5 ? 0.f : 1e-4f;
            const float tmax = U(rng) < 0.7 ? __builtin_inff() : float(1.0 + U(rng));
            const lmgpu::RayPre r = lmgpu::ray_prepare(lmgpu::mk(o[0], o[1], o[2]), lmgpu::mk(d[0], d[1], d[2]));
            float key[W];
            int ent[W];
            lmgpu::child_keys<W>(q, r, tmin, tmax, key, ent);
            n_rays++;
            const bool ex = exact_hit(o, d, lo, hi, tmin, tmax);
            const bool got = key[c] != __builtin_inff();
            n_exact += ex;
            n_extra += got && !ex;
            if (ex && !got) {
                std::printf("W%d node %d child %d: ray o (%.9g, %.9g, %.9g) d (%.9g, %.9g, %.9g) tmin %g tmax %g "
                            "meets the box [%.9g %.9g %.9g]-[%.9g %.9g %.9g] but the quantised test misses it\n",
                            W, it, c, o[0], o[1], o[2], d[0], d[1], d[2], tmin, tmax, lo[0], lo[1], lo[2], hi[0],
                            hi[1], hi[2]);
                return 1;
            }
        }
    }
    return 0;
}

int main() {
    std::mt19937_64 rng(12345);
    const int kNodes = 200000;
    if (check<4>(rng, kNodes) || check<8>(rng, kNodes)) return 1;
    long long n_rays = 0, n_exact = 0, n_extra = 0;
    if (check_slab<4>(rng, 100000, 16, n_rays, n_exact, n_extra) ||
        check_slab<8>(rng, 50000, 16, n_rays, n_exact, n_extra))
        return 1;
    std::printf("ok %d %lld (exact hits %lld, accepted beyond the exact boxes %lld)\n", 2 * kNodes, n_rays, n_exact,
                n_extra);
    return 0;
}

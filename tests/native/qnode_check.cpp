// Host check of the quantised BVH nodes (csrc/bvh.h, quantise_node /
// qplane): for random 4- and 8-wide nodes over many scales and offsets, every
// decoded child box contains the original one, the grid is tight (within
// one step plus rounding of the exact planes), and stack entries and empty
// slots are encoded as the traversal expects. Prints "ok <nodes>" or the
// first failure and exits non-zero.
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

int main() {
    std::mt19937_64 rng(12345);
    const int kNodes = 200000;
    if (check<4>(rng, kNodes) || check<8>(rng, kNodes)) return 1;
    std::printf("ok %d\n", 2 * kNodes);
    return 0;
}

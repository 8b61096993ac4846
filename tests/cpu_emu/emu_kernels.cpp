// Runs the batched 2D solver kernels on the CPU (see hip/hip_runtime.h in this
// folder) and checks that one two-level launch gives bit for bit what two
// single-level launches give: levels, bfloat16 low parts, traces, adjoint
// fields and the Neumaier gradient accumulators.  Buffers carry the same guard
// zones as the library's, so under AddressSanitizer any access outside them
// aborts.  Usage: emu_kernels <seed> [case-filter]; exit status 0 = all equal.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "usct_kernels.hpp"

#ifndef GUARD_CELLS
#error "compile with -DGUARD_CELLS=<kGuardCells of the library>"
#endif

using namespace usct;

namespace {

std::vector<std::unique_ptr<char[]>> g_keep;

template <class T>
T *alloc(int64_t count, int64_t guard_stride) {
    const int64_t g = guard_stride > 0 ? GUARD_CELLS * guard_stride : 0;
    const int64_t total = count + 2 * g;
    char *raw = new char[sizeof(T) * (size_t)std::max<int64_t>(total, 1)]();
    g_keep.emplace_back(raw);
    return reinterpret_cast<T *>(raw) + g;
}

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct Problem {
    int ny, nx, Bp, LB, CX, RY, ng, n_rec;
    int64_t N, FE;
    int nxs, nyc;
    unsigned blocks;
    float2 *sp2;
    int *src_cell, *rec_head, *rec_next, *row_rec;
    float *sinc;  // [3][Bp]
    std::mt19937 rng;

    Problem(int ny_, int nx_, int LB_, int ng_, int CX_, int RY_, uint32_t seed)
        : ny(ny_), nx(nx_), Bp(LB_ * ng_), LB(LB_), CX(CX_), RY(RY_), ng(ng_), rng(seed) {
        N = (int64_t)ny * nx;
        FE = N * Bp;
        const int W = 64 / LB;
        nxs = (int)cdiv(nx, (int64_t)W * CX);
        nyc = (int)cdiv(ny, RY);
        const int64_t waves = (int64_t)ng * nxs * nyc;
        blocks = (unsigned)cdiv(waves, 4);
        std::uniform_real_distribution<float> u(0.05f, 0.24f), t(-1e-9f, 1e-9f);
        sp2 = alloc<float2>(N, 1);
        for (int64_t c = 0; c < N; ++c) sp2[c] = make_float2(u(rng), t(rng));
        // sources: every lane group gets cells at borders, strip edges and random
        // interior cells; the last shot of the pass is padding (-1)
        src_cell = alloc<int>(Bp, 0);
        std::uniform_int_distribution<int> cy(0, ny - 1), cx(0, nx - 1);
        for (int b = 0; b < Bp; ++b) {
            int y = cy(rng), x = cx(rng);
            switch (b % 7) {
                case 0: x = 0; break;
                case 1: x = std::min(CX - 1, nx - 1); break;
                case 2: x = std::min(CX, nx - 1); break;
                case 3: y = 0; break;
                case 4: y = std::min(RY, ny - 1); x = std::min(2 * CX, nx - 1); break;
                default: break;
            }
            src_cell[b] = y * nx + x;
        }
        src_cell[Bp - 1] = -1;
        sinc = alloc<float>(3 * (int64_t)Bp, 0);
        std::uniform_real_distribution<float> w(-2.0f, 2.0f);
        for (int i = 0; i < 3 * Bp; ++i) sinc[i] = w(rng);
        // receivers: random cells, a duplicate pair, the four corners, strip edges
        std::vector<int> rc;
        for (int i = 0; i < 12; ++i) rc.push_back(cy(rng) * nx + cx(rng));
        rc.push_back(rc[3]);
        rc.push_back(0);
        rc.push_back(nx - 1);
        rc.push_back((ny - 1) * nx);
        rc.push_back((int)N - 1);
        rc.push_back(std::min(RY, ny - 1) * nx + std::min(CX, nx - 1));
        rc.push_back(std::min(RY - 1, ny - 1) * nx + std::min(CX - 1, nx - 1));
        n_rec = (int)rc.size();
        rec_head = alloc<int>(N, 0);
        rec_next = alloc<int>(n_rec, 0);
        row_rec = alloc<int>(ny, 0);
        for (int64_t c = 0; c < N; ++c) rec_head[c] = -1;
        for (int j = n_rec - 1; j >= 0; --j) {
            rec_next[j] = rec_head[rc[j]];
            rec_head[rc[j]] = j;
        }
        for (int j = 0; j < n_rec; ++j) row_rec[rc[j] / nx]++;
    }

    float *field(float scale) {
        float *p = alloc<float>(FE, Bp);
        std::uniform_real_distribution<float> d(-scale, scale);
        for (int64_t i = 0; i < FE; ++i) p[i] = d(rng);
        return p;
    }
    uint16_t *low(float scale) {
        uint16_t *p = alloc<uint16_t>(FE, Bp);
        std::uniform_real_distribution<float> d(-scale, scale);
        for (int64_t i = 0; i < FE; ++i) p[i] = f2bf(d(rng));
        return p;
    }
    float *zeros() { return alloc<float>(FE, Bp); }
    uint16_t *zeros_lo() { return alloc<uint16_t>(FE, Bp); }
    float *traces() { return alloc<float>((int64_t)n_rec * Bp, 0); }
};

template <class T>
bool same(const char *what, const T *a, const T *b, int64_t n, const char *tag) {
    if (std::memcmp(a, b, sizeof(T) * (size_t)n) == 0) return true;
    for (int64_t i = 0; i < n; ++i)
        if (std::memcmp(a + i, b + i, sizeof(T)) != 0) {
            std::printf("MISMATCH %s %s at element %lld\n", tag, what, (long long)i);
            break;
        }
    return false;
}

template <int LB, int CX, bool REC>
bool check_fwd(int ny, int nx, int ng, int RY, uint32_t seed) {
    Problem P(ny, nx, LB, ng, CX, RY, seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "fwd ny=%d nx=%d LB=%d ng=%d CX=%d RY=%d rec=%d", ny, nx, LB, ng,
                  CX, RY, (int)REC);
    float *pc = P.field(1.0f), *pp = P.field(1.0f);
    uint16_t *pcl = P.low(1e-7f), *ppl = P.low(1e-7f);
    float *a1 = P.zeros(), *b1 = P.zeros(), *a2 = P.zeros(), *b2 = P.zeros();
    uint16_t *a1l = P.zeros_lo(), *b1l = P.zeros_lo(), *a2l = P.zeros_lo(), *b2l = P.zeros_lo();
    float *ra1 = P.traces(), *ral1 = P.traces(), *rb1 = P.traces(), *rbl1 = P.traces();
    float *ra2 = P.traces(), *ral2 = P.traces(), *rb2 = P.traces(), *rbl2 = P.traces();
    const float *sa = P.sinc + P.Bp, *sb = P.sinc + 2 * P.Bp;
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d<LB, CX, false, REC>(pc, pp, a1, pcl, ppl, a1l, P.sp2, nullptr, P.src_cell, sa,
                                    P.rec_head, P.rec_next, P.row_rec, ra1, ral1, nullptr, P.ny,
                                    P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
    });
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d<LB, CX, false, REC>(a1, pc, b1, a1l, pcl, b1l, P.sp2, nullptr, P.src_cell, sb,
                                    P.rec_head, P.rec_next, P.row_rec, rb1, rbl1, nullptr, P.ny,
                                    P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
    });
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d_x2<LB, CX, REC>(pc, pp, a2, b2, pcl, ppl, a2l, b2l, P.sp2, P.src_cell, sa, sb,
                                P.rec_head, P.rec_next, P.row_rec, ra2, ral2, rb2, rbl2, P.ny,
                                P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
    });
    const int64_t nr = (int64_t)P.n_rec * P.Bp;
    bool ok = same("level n", a1, a2, P.FE, tag) & same("level n low", a1l, a2l, P.FE, tag) &
              same("level n+1", b1, b2, P.FE, tag) & same("level n+1 low", b1l, b2l, P.FE, tag);
    if (REC)
        ok = ok & same("traces n", ra1, ra2, nr, tag) & same("traces n low", ral1, ral2, nr, tag) &
             same("traces n+1", rb1, rb2, nr, tag) & same("traces n+1 low", rbl1, rbl2, nr, tag);
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

template <int LB, int CX>
bool fwd_cases(uint32_t seed) {
    bool ok = true;
    const int shapes[][2] = {{29, 37}, {9, 13}, {40, 33}, {3, 5}};
    for (auto &s : shapes)
        for (int RY : {4, 8, 64}) {
            ok &= check_fwd<LB, CX, true>(s[0], s[1], 2, RY, seed++);
        }
    ok &= check_fwd<LB, CX, false>(29, 37, 1, 4, seed++);
    return ok;
}

template <int LB, int CX>
bool check_adj(int ny, int nx, int ng, int RY, uint32_t seed) {
    Problem P(ny, nx, LB, ng, CX, RY, seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "adj ny=%d nx=%d LB=%d ng=%d CX=%d RY=%d", ny, nx, LB, ng, CX,
                  RY);
    float *l1 = P.field(1.0f), *l2 = P.field(1.0f), *pa1 = P.field(1.0f), *pa2 = P.field(1.0f);
    float *u1 = P.zeros(), *v1 = P.zeros(), *u2 = P.zeros(), *v2 = P.zeros();
    const int64_t nr = (int64_t)P.n_rec * P.Bp;
    float *res = alloc<float>(2 * nr, 0);
    std::uniform_real_distribution<float> d(-3.0f, 3.0f);
    for (int64_t i = 0; i < 2 * nr; ++i) res[i] = d(P.rng);
    const int64_t gn = (int64_t)P.ng * P.N;
    float *gh1 = alloc<float>(gn, 0), *gl1 = alloc<float>(gn, 0);
    float *gh2 = alloc<float>(gn, 0), *gl2 = alloc<float>(gn, 0);
    std::uniform_real_distribution<float> gd(-10.0f, 10.0f), ld(-1e-6f, 1e-6f);
    for (int64_t i = 0; i < gn; ++i) {
        gh1[i] = gh2[i] = gd(P.rng);
        gl1[i] = gl2[i] = ld(P.rng);
    }
    const float *su = P.sinc + 2 * P.Bp, *sv = P.sinc + P.Bp;  // levels k and k-1
    const float *ru = res + nr, *rv = res;
    emu::launch(P.blocks, kBlock, [&] {
        k_adj2d<LB, CX, false>(l1, l2, u1, pa1, nullptr, nullptr, P.sp2, nullptr, P.src_cell, su,
                               P.rec_head, P.rec_next, P.row_rec, ru, nullptr, gh1, gl1, P.ny,
                               P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
    });
    emu::launch(P.blocks, kBlock, [&] {
        k_adj2d<LB, CX, false>(u1, l1, v1, pa2, nullptr, nullptr, P.sp2, nullptr, P.src_cell, sv,
                               P.rec_head, P.rec_next, P.row_rec, rv, nullptr, gh1, gl1, P.ny,
                               P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
    });
    emu::launch(P.blocks, kBlock, [&] {
        k_adj2d_x2<LB, CX>(l1, l2, u2, v2, pa1, pa2, P.sp2, P.src_cell, su, sv, P.rec_head,
                           P.rec_next, P.row_rec, ru, rv, gh2, gl2, P.ny, P.nx, P.Bp, P.RY, P.nxs,
                           P.nyc, P.ng);
    });
    const bool ok = same("lam^k", u1, u2, P.FE, tag) & same("lam^{k-1}", v1, v2, P.FE, tag) &
                    same("g_hi", gh1, gh2, gn, tag) & same("g_lo", gl1, gl2, gn, tag);
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

template <int LB, int CX>
bool adj_cases(uint32_t seed) {
    bool ok = true;
    const int shapes[][2] = {{29, 37}, {9, 13}, {40, 33}, {3, 5}};
    for (auto &s : shapes)
        for (int RY : {4, 8, 64}) ok &= check_adj<LB, CX>(s[0], s[1], 2, RY, seed++);
    ok &= check_adj<LB, CX>(29, 37, 1, 4, seed++);
    return ok;
}

}  // namespace

int main(int argc, char **argv) {
    const uint32_t seed = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 1u;
    const std::string only = argc > 2 ? argv[2] : "";
    bool ok = true;
    if (only.empty() || only == "fwd") {
        ok &= fwd_cases<16, 4>(seed);
        ok &= fwd_cases<16, 8>(seed + 100);
        ok &= fwd_cases<32, 4>(seed + 200);
        ok &= fwd_cases<64, 4>(seed + 300);
        ok &= fwd_cases<64, 8>(seed + 400);
    }
    if (only.empty() || only == "adj") {
        ok &= adj_cases<16, 4>(seed + 500);
        ok &= adj_cases<16, 8>(seed + 600);
        ok &= adj_cases<32, 4>(seed + 700);
        ok &= adj_cases<64, 4>(seed + 800);
        ok &= adj_cases<64, 8>(seed + 900);
    }
    std::printf(ok ? "ALL EQUAL\n" : "DIFFERENCES FOUND\n");
    return ok ? 0 : 1;
}

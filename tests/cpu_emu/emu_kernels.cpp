// Runs the batched 2D solver kernels on the CPU (see hip/hip_runtime.h in this
// folder) and checks that one two-level launch gives bit for bit what two
// single-level launches give: levels, bfloat16 low parts, traces, adjoint
// fields and the Neumaier gradient accumulators.  Buffers carry the same guard
// zones as the library's, so under AddressSanitizer any access outside them
// aborts.  Usage: emu_kernels <seed> [case-filter]; exit status 0 = all equal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
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

// strip order of the 3D patch kernels in the current check (tasks3d / strip3d)
int g_blk = kBlk3dDefault;

// A low part from its value: bfloat16 bits (round to nearest even) or FP32.
template <class L>
L mk_lo(float r);
template <>
uint16_t mk_lo<uint16_t>(float r) { return f2bf(r); }
template <>
float mk_lo<float>(float r) { return r; }

struct Problem {
    int ny, nx, Bp, LB, CX, RY, ng, n_rec;
    int64_t N, FE;
    int nxs, nyc;
    unsigned blocks;
    float2 *sp2;
    int *src_cell, *rec_head, *rec_next, *row_rec;
    float *sinc;    // [5][Bp]
    double *sincd;  // [5][Bp], the same values in FP64
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
        sinc = alloc<float>(5 * (int64_t)Bp, 0);
        sincd = alloc<double>(5 * (int64_t)Bp, 0);
        std::uniform_real_distribution<float> w(-2.0f, 2.0f);
        for (int i = 0; i < 5 * Bp; ++i) {
            sincd[i] = (double)w(rng) * 1.0000001234567;
            sinc[i] = (float)sincd[i];
        }
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
    // low parts of either type (bfloat16 bits or FP32)
    template <class L>
    L *lowT(float scale) {
        L *p = alloc<L>(FE, Bp);
        std::uniform_real_distribution<float> d(-scale, scale);
        for (int64_t i = 0; i < FE; ++i) p[i] = mk_lo<L>(d(rng));
        return p;
    }
    template <class L>
    L *zeros_loT() { return alloc<L>(FE, Bp); }
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

// One 2D forward level by a direct loop over cells: the FP64 update the
// kernels implement (full hi + lo state, FP64 S', source and sponge).
template <class L>
bool direct_fwd2d(const Problem &P, const float *pc, const L *pcl, const float *pp,
                  const L *ppl, const double *sincd, const float *sponge, const float *pn,
                  const L *pnl, const char *tag) {
    auto Pv = [&](int64_t c, int b) { return hl(pc[c * P.Bp + b], pcl[c * P.Bp + b]); };
    for (int y = 0; y < P.ny; ++y)
        for (int x = 0; x < P.nx; ++x)
            for (int b = 0; b < P.Bp; ++b) {
                const int64_t c = (int64_t)y * P.nx + x;
                double lu = 0.0;
                if (y >= 1 && y <= P.ny - 2 && x >= 1 && x <= P.nx - 2)
                    lu = lap5d(Pv(c, b), Pv(c + P.nx, b), Pv(c - P.nx, b), Pv(c + 1, b), Pv(c - 1, b));
                double inc = s2d(P.sp2[c]) * lu;
                if (c == P.src_cell[b]) inc = inc + sincd[b];
                double nv = (2.0 * Pv(c, b) - hl(pp[c * P.Bp + b], ppl[c * P.Bp + b])) + inc;
                if (sponge) nv = nv * (double)sponge[c];
                L lb;
                const float v = split_hl(nv, lb);
                if (std::memcmp(&v, &pn[c * P.Bp + b], 4) != 0 || lb != pnl[c * P.Bp + b]) {
                    std::printf("MISMATCH %s direct FP64 update at cell %lld shot %d\n", tag,
                                (long long)c, b);
                    return false;
                }
            }
    return true;
}

template <int LB, int CX, bool SPONGE, class L = uint16_t>
bool check_fwd_direct(int ny, int nx, int ng, int RY, uint32_t seed) {
    Problem P(ny, nx, LB, ng, CX, RY, seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "fwd-direct ny=%d nx=%d LB=%d ng=%d CX=%d RY=%d sponge=%d lo=%d", ny,
                  nx, LB, ng, CX, RY, (int)SPONGE, (int)sizeof(L));
    float *pc = P.field(1.0f), *pp = P.field(1.0f), *pn = P.zeros();
    L *pcl = P.template lowT<L>(1e-7f), *ppl = P.template lowT<L>(1e-7f), *pnl = P.template zeros_loT<L>();
    float *sp = alloc<float>(P.N, 0);
    std::uniform_real_distribution<float> d(0.5f, 1.0f);
    for (int64_t c = 0; c < P.N; ++c) sp[c] = d(P.rng);
    float *ra = P.traces(), *rl = P.traces();
    const double *sa = P.sincd + P.Bp;
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d<LB, CX, SPONGE, true, L>(pc, pp, pn, pcl, ppl, pnl, P.sp2, sp, P.src_cell, sa,
                                      P.rec_head, P.rec_next, P.row_rec, ra, rl, P.ny, P.nx, P.Bp,
                                      P.RY, P.nxs, P.nyc, P.ng);
    });
    const bool ok = direct_fwd2d(P, pc, pcl, pp, ppl, sa, SPONGE ? sp : nullptr, pn, pnl, tag);
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

template <int LB, int CX, bool REC, class L = uint16_t>
bool check_fwd(int ny, int nx, int ng, int RY, uint32_t seed, int sched) {
    Problem P(ny, nx, LB, ng, CX, RY, seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "fwd ny=%d nx=%d LB=%d ng=%d CX=%d RY=%d rec=%d sched=%d lo=%d", ny,
                  nx, LB, ng, CX, RY, (int)REC, sched, (int)sizeof(L));
    float *pc = P.field(1.0f), *pp = P.field(1.0f);
    L *pcl = P.template lowT<L>(1e-7f), *ppl = P.template lowT<L>(1e-7f);
    float *a1 = P.zeros(), *b1 = P.zeros(), *a2 = P.zeros(), *b2 = P.zeros();
    L *a1l = P.template zeros_loT<L>(), *b1l = P.template zeros_loT<L>(), *a2l = P.template zeros_loT<L>(), *b2l = P.template zeros_loT<L>();
    float *ra1 = P.traces(), *ral1 = P.traces(), *rb1 = P.traces(), *rbl1 = P.traces();
    float *ra2 = P.traces(), *ral2 = P.traces(), *rb2 = P.traces(), *rbl2 = P.traces();
    const double *sa = P.sincd + P.Bp, *sb = P.sincd + 2 * P.Bp;
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d<LB, CX, false, REC, L>(pc, pp, a1, pcl, ppl, a1l, P.sp2, nullptr, P.src_cell, sa,
                                    P.rec_head, P.rec_next, P.row_rec, ra1, ral1, P.ny, P.nx, P.Bp,
                                    P.RY, P.nxs, P.nyc, P.ng);
    });
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d<LB, CX, false, REC, L>(a1, pc, b1, a1l, pcl, b1l, P.sp2, nullptr, P.src_cell, sb,
                                    P.rec_head, P.rec_next, P.row_rec, rb1, rbl1, P.ny, P.nx, P.Bp,
                                    P.RY, P.nxs, P.nyc, P.ng);
    });
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d_x2<LB, CX, REC, L>(pc, pp, a2, b2, pcl, ppl, a2l, b2l, P.sp2, P.src_cell, sa, sb,
                                P.rec_head, P.rec_next, P.row_rec, ra2, ral2, rb2, rbl2, P.ny,
                                P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng, sched);
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
    // every wavefront order and march direction (sched 0..7)
    for (auto &s : shapes)
        for (int RY : {4, 8, 64})
            for (int sched = 0; sched < 8; ++sched)
                ok &= check_fwd<LB, CX, true>(s[0], s[1], 2, RY, seed++, sched);
    ok &= check_fwd<LB, CX, false>(29, 37, 1, 4, seed++, 3);
    ok &= check_fwd_direct<LB, CX, false>(29, 37, 2, 8, seed++);
    ok &= check_fwd_direct<LB, CX, true>(9, 13, 1, 4, seed++);
    return ok;
}

template <int LB, int CX>
bool check_adj(int ny, int nx, int ng, int RY, uint32_t seed, int sched) {
    Problem P(ny, nx, LB, ng, CX, RY, seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "adj ny=%d nx=%d LB=%d ng=%d CX=%d RY=%d sched=%d", ny, nx, LB,
                  ng, CX, RY, sched);
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
                           P.nyc, P.ng, sched);
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
        for (int RY : {4, 8, 64})
            for (int sched = 0; sched < 8; ++sched)
                ok &= check_adj<LB, CX>(s[0], s[1], 2, RY, seed++, sched);
    ok &= check_adj<LB, CX>(29, 37, 1, 4, seed++, 3);
    return ok;
}

// LV-level forward launch (k_fwd2d_xn) against LV single-level launches of
// k_fwd2d.  sparse: the launch stores only what a gradient's forward asks of
// it (hi of every level but the second, lo of the last two), the rest must be
// left untouched.
template <int LB, int CX, int LV, bool REC, class L = uint16_t>
bool check_fwdn(int ny, int nx, int ng, int RY, uint32_t seed, int sched, bool sparse) {
    Problem P(ny, nx, LB, ng, CX, RY, seed);
    char tag[200];
    std::snprintf(tag, sizeof tag,
                  "fwd-x%d ny=%d nx=%d LB=%d ng=%d CX=%d RY=%d rec=%d sched=%d lo=%d sparse=%d", LV,
                  ny, nx, LB, ng, CX, RY, (int)REC, sched, (int)sizeof(L), (int)sparse);
    float *pc = P.field(1.0f), *pp = P.field(1.0f);
    L *pcl = P.template lowT<L>(1e-7f), *ppl = P.template lowT<L>(1e-7f);
    float *ref[LV], *out[LV], *rr[LV], *rrl[LV], *ro[LV], *rol[LV];
    L *refl[LV], *outl[LV];
    const float *c = pc, *p = pp;
    const L *cl = pcl, *pl = ppl;
    for (int i = 0; i < LV; ++i) {
        ref[i] = P.zeros();
        refl[i] = P.template zeros_loT<L>();
        out[i] = P.zeros();
        outl[i] = P.template zeros_loT<L>();
        rr[i] = P.traces();
        rrl[i] = P.traces();
        ro[i] = P.traces();
        rol[i] = P.traces();
        const double *sd = P.sincd + (int64_t)(i + 1) * P.Bp;
        float *a = ref[i];
        L *al = refl[i];
        emu::launch(P.blocks, kBlock, [&] {
            k_fwd2d<LB, CX, false, REC, L>(c, p, a, cl, pl, al, P.sp2, nullptr, P.src_cell, sd,
                                           P.rec_head, P.rec_next, P.row_rec, rr[i], rrl[i], P.ny,
                                           P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
        });
        p = c;
        pl = cl;
        c = ref[i];
        cl = refl[i];
    }
    FwdLevels<LV, L> o{};
    for (int i = 0; i < LV; ++i) {
        const bool last2 = i >= LV - 2;
        o.hi[i] = sparse && !last2 && i == 1 ? nullptr : out[i];
        o.lo[i] = sparse && !last2 ? nullptr : outl[i];
        o.sincd[i] = P.sincd + (int64_t)(i + 1) * P.Bp;
        o.rec[i] = ro[i];
        o.rec_lo[i] = rol[i];
    }
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d_xn<LB, CX, LV, REC, L>(pc, pp, pcl, ppl, o, P.sp2, P.src_cell, P.rec_head,
                                       P.rec_next, P.row_rec, P.ny, P.nx, P.Bp, P.RY, P.nxs, P.nyc,
                                       P.ng, sched);
    });
    const int64_t nr = (int64_t)P.n_rec * P.Bp;
    float *zf = P.zeros();
    L *zl = P.template zeros_loT<L>();
    bool ok = true;
    for (int i = 0; i < LV; ++i) {
        char what[64];
        std::snprintf(what, sizeof what, "level n+%d", i);
        ok &= o.hi[i] ? same(what, ref[i], out[i], P.FE, tag) : same(what, zf, out[i], P.FE, tag);
        std::snprintf(what, sizeof what, "level n+%d low", i);
        ok &= o.lo[i] ? same(what, refl[i], outl[i], P.FE, tag) : same(what, zl, outl[i], P.FE, tag);
        if (REC) {
            std::snprintf(what, sizeof what, "traces n+%d", i);
            ok &= same(what, rr[i], ro[i], nr, tag);
            std::snprintf(what, sizeof what, "traces n+%d low", i);
            ok &= same(what, rrl[i], rol[i], nr, tag);
        }
    }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

// LV-level adjoint launch (k_adj2d_xn) against LV single-level launches of
// k_adj2d: adjoint levels and the Neumaier gradient accumulators.
template <int LB, int CX, int LV>
bool check_adjn(int ny, int nx, int ng, int RY, uint32_t seed, int sched, bool sparse) {
    Problem P(ny, nx, LB, ng, CX, RY, seed);
    char tag[200];
    std::snprintf(tag, sizeof tag, "adj-x%d ny=%d nx=%d LB=%d ng=%d CX=%d RY=%d sched=%d sparse=%d",
                  LV, ny, nx, LB, ng, CX, RY, sched, (int)sparse);
    float *l1 = P.field(1.0f), *l2 = P.field(1.0f);
    float *pf[LV], *ref[LV], *out[LV];
    for (int i = 0; i < LV; ++i) {
        pf[i] = P.field(1.0f);
        ref[i] = P.zeros();
        out[i] = P.zeros();
    }
    const int64_t nr = (int64_t)P.n_rec * P.Bp;
    float *res = alloc<float>(LV * nr, 0);
    std::uniform_real_distribution<float> d(-3.0f, 3.0f);
    for (int64_t i = 0; i < LV * nr; ++i) res[i] = d(P.rng);
    const int64_t gn = (int64_t)P.ng * P.N;
    float *gh1 = alloc<float>(gn, 0), *gl1 = alloc<float>(gn, 0);
    float *gh2 = alloc<float>(gn, 0), *gl2 = alloc<float>(gn, 0);
    std::uniform_real_distribution<float> gd(-10.0f, 10.0f), ld(-1e-6f, 1e-6f);
    for (int64_t i = 0; i < gn; ++i) {
        gh1[i] = gh2[i] = gd(P.rng);
        gl1[i] = gl2[i] = ld(P.rng);
    }
    const float *a = l1, *b = l2;
    for (int i = 0; i < LV; ++i) {
        float *u = ref[i];
        const float *si = P.sinc + (int64_t)(i + 1) * P.Bp, *ri = res + i * nr;
        emu::launch(P.blocks, kBlock, [&] {
            k_adj2d<LB, CX, false>(a, b, u, pf[i], nullptr, nullptr, P.sp2, nullptr, P.src_cell,
                                   si, P.rec_head, P.rec_next, P.row_rec, ri, nullptr, gh1, gl1,
                                   P.ny, P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
        });
        b = a;
        a = ref[i];
    }
    AdjLevels<LV> o{};
    for (int i = 0; i < LV; ++i) {
        o.lam[i] = sparse && i < LV - 2 ? nullptr : out[i];
        o.pf[i] = pf[i];
        o.sinc[i] = P.sinc + (int64_t)(i + 1) * P.Bp;
        o.res[i] = res + i * nr;
    }
    emu::launch(P.blocks, kBlock, [&] {
        k_adj2d_xn<LB, CX, LV>(l1, l2, o, P.sp2, P.src_cell, P.rec_head, P.rec_next, P.row_rec,
                               gh2, gl2, P.ny, P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng, sched);
    });
    float *zf = P.zeros();
    bool ok = same("g_hi", gh1, gh2, gn, tag) & same("g_lo", gl1, gl2, gn, tag);
    for (int i = 0; i < LV; ++i) {
        char what[64];
        std::snprintf(what, sizeof what, "lam^{k-%d}", i);
        ok &= o.lam[i] ? same(what, ref[i], out[i], P.FE, tag) : same(what, zf, out[i], P.FE, tag);
    }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

// all: every wavefront order and march direction (sched 0..7) over five grids and
// four row counts; else sched 0, 5, 6 over three grids and two row counts
template <int LB, int CX, int LV>
bool xn_cases(uint32_t seed, bool all) {
    bool ok = true;
    const int shapes[][2] = {{29, 37}, {3, 5}, {7, 4}, {9, 13}, {40, 33}};
    for (int si = 0; si < (all ? 5 : 3); ++si)
        for (int RY : {1, 5, 4, 64}) {
            if (!all && RY != 1 && RY != 5) continue;
            const int *s = shapes[si];
            for (int sched = 0; sched < 8; ++sched) {
                if (!all && sched != 0 && sched != 5 && sched != 6) continue;
                const bool sparse = (sched & 1) != 0;
                ok &= check_fwdn<LB, CX, LV, true>(s[0], s[1], 2, RY, seed++, sched, sparse);
                ok &= check_adjn<LB, CX, LV>(s[0], s[1], 2, RY, seed++, sched, sparse);
            }
        }
    ok &= check_fwdn<LB, CX, LV, false>(29, 37, 1, 4, seed++, 3, true);
    ok &= check_fwdn<LB, CX, LV, true, float>(29, 37, 2, 8, seed++, 1, false);
    ok &= check_fwdn<LB, CX, LV, true, float>(9, 13, 1, 3, seed++, 6, true);
    ok &= check_adjn<LB, CX, LV>(29, 37, 1, 4, seed++, 3, false);
    return ok;
}

// 3D forward: k_fwd3d against a direct loop over cells of the same FP64 update
// (full hi + lo state, FP64 S' and source term, split into FP32 hi + bf16 lo).
// KIND 0: k_fwd3d with CX x CY patches per lane; KIND 1: k_fwd3d_lds with a
// TX x TY tile per work-group (CX, CY = TX, TY) of PX x PY patches per lane;
// KIND 2: k_fwd3d in work-groups of y-stacked strips sharing edge rows (YSH).
template <int LB, int CX, int CY, int KIND = 0, int PX = 1, int PY = 1, class L = uint16_t>
bool check_fwd3d(int nz, int ny, int nx, int ng, int RZ, uint32_t seed) {
    const int Bp = LB * ng, W = 64 / LB;
    const int64_t N = (int64_t)nz * ny * nx, FE = N * Bp;
    std::mt19937 rng(seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "%s nz=%d ny=%d nx=%d LB=%d ng=%d C/T=%dx%d P=%dx%d RZ=%d",
                  KIND == 2 ? (sizeof(L) == 4 ? "fwd3d-ysh-lo32" : "fwd3d-ysh")
                            : (KIND ? "fwd3d-lds" : (sizeof(L) == 4 ? "fwd3d-lo32" : "fwd3d")),
                  nz, ny, nx, LB, ng, CX, CY, PX, PY, RZ);
    std::uniform_real_distribution<float> u(0.02f, 0.1f), tt(-1e-9f, 1e-9f), d(-1.0f, 1.0f),
        dl(-1e-7f, 1e-7f);
    float2 *sp2 = alloc<float2>(N, 1);
    for (int64_t c = 0; c < N; ++c) sp2[c] = make_float2(u(rng), tt(rng));
    float *pc = alloc<float>(FE, Bp), *pp = alloc<float>(FE, Bp), *pn = alloc<float>(FE, Bp);
    L *pcl = alloc<L>(FE, Bp), *ppl = alloc<L>(FE, Bp),
             *pnl = alloc<L>(FE, Bp);
    for (int64_t i = 0; i < FE; ++i) {
        pc[i] = d(rng);
        pp[i] = d(rng);
        pcl[i] = mk_lo<L>(dl(rng));
        ppl[i] = mk_lo<L>(dl(rng));
    }
    int *src = alloc<int>(Bp, 0);
    std::uniform_int_distribution<int> cz(0, nz - 1), cy(0, ny - 1), cx(0, nx - 1);
    for (int b = 0; b < Bp; ++b) src[b] = (cz(rng) * ny + cy(rng)) * nx + (b % 3 == 0 ? 0 : cx(rng));
    src[Bp - 1] = -1;
    double *sincd = alloc<double>(Bp, 0);
    for (int b = 0; b < Bp; ++b) sincd[b] = (double)d(rng) * 1.2345678901;
    float *srcd2 = alloc<float>(Bp, 0);
    std::vector<int> rc = {0, (int)N - 1, (int)(N / 2), (int)(N / 3), (int)(N / 3)};
    const int n_rec = (int)rc.size();
    int *rec_head = alloc<int>(N, 0), *rec_next = alloc<int>(n_rec, 0);
    int *row_rec = alloc<int>((int64_t)nz * ny, 0);
    for (int64_t c = 0; c < N; ++c) rec_head[c] = -1;
    for (int j = n_rec - 1; j >= 0; --j) {
        rec_next[j] = rec_head[rc[j]];
        rec_head[rc[j]] = j;
    }
    for (int j = 0; j < n_rec; ++j) row_rec[rc[j] / nx]++;
    float *rec = alloc<float>((int64_t)n_rec * Bp, 0), *rec_lo = alloc<float>((int64_t)n_rec * Bp, 0);
    if constexpr (KIND == 0) {
        const int nxs = (int)cdiv(nx, (int64_t)W * CX), nys = (int)cdiv(ny, CY);
        const int nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)cdiv(ng * tasks3d(nxs, nys, nzc, g_blk), 4);
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d<LB, CX, CY, true, L>(pc, pp, pn, pcl, ppl, pnl, sp2, src, sincd, rec_head,
                                      rec_next, row_rec, rec, rec_lo, srcd2, nz, ny, nx, Bp, RZ,
                                      nxs, nys, nzc, ng, g_blk);
        });
    } else if constexpr (KIND == 2) {
        const int nxs = (int)cdiv(nx, (int64_t)W * CX), nys = (int)cdiv(ny, CY);
        const int nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)(ng * nxs * cdiv(nys, kYshF) * nzc);
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d<LB, CX, CY, true, L, true>(pc, pp, pn, pcl, ppl, pnl, sp2, src, sincd,
                                               rec_head, rec_next, row_rec, rec, rec_lo, srcd2,
                                               nz, ny, nx, Bp, RZ, nxs, nys, nzc, ng, g_blk);
        });
    } else {
        const int ntx = (int)cdiv(nx, CX), nty = (int)cdiv(ny, CY), nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)((int64_t)ng * ntx * nty * nzc);
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d_lds<LB, CX, CY, PX, PY, true>(pc, pp, pn, pcl, ppl, pnl, sp2, src, sincd,
                                                  rec_head, rec_next, row_rec, rec, rec_lo, srcd2,
                                                  nz, ny, nx, Bp, RZ, ntx, nty, nzc, ng);
        });
    }
    bool ok = true;
    auto Pv = [&](int64_t c, int b) { return (double)pc[c * Bp + b] + (double)lo_f(pcl[c * Bp + b]); };
    for (int z = 0; z < nz && ok; ++z)
        for (int y = 0; y < ny && ok; ++y)
            for (int x = 0; x < nx && ok; ++x)
                for (int b = 0; b < Bp; ++b) {
                    const int64_t c = ((int64_t)z * ny + y) * nx + x;
                    double lu = 0.0;
                    if (z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2)
                        lu = lap7d(Pv(c, b), Pv(c + (int64_t)nx * ny, b), Pv(c - (int64_t)nx * ny, b),
                                   Pv(c + nx, b), Pv(c - nx, b), Pv(c + 1, b), Pv(c - 1, b));
                    double inc = ((double)sp2[c].x + (double)sp2[c].y) * lu;
                    if (c == src[b]) inc = inc + sincd[b];
                    const double q = (double)pp[c * Bp + b] + (double)lo_f(ppl[c * Bp + b]);
                    const double nv = (2.0 * Pv(c, b) - q) + inc;
                    const float v = (float)nv;
                    const L lb = mk_lo<L>((float)(nv - (double)v));
                    if (std::memcmp(&v, &pn[c * Bp + b], 4) != 0 || lb != pnl[c * Bp + b]) {
                        std::printf("MISMATCH %s cell %lld shot %d\n", tag, (long long)c, b);
                        ok = false;
                        break;
                    }
                    if (c == src[b] && srcd2[b] != (float)inc) {
                        std::printf("MISMATCH %s srcd2 shot %d\n", tag, b);
                        ok = false;
                    }
                    for (int r = rec_head[c]; r >= 0; r = rec_next[r])
                        if (rec[(int64_t)r * Bp + b] != v || rec_lo[(int64_t)r * Bp + b] != lo_f(lb)) {
                            std::printf("MISMATCH %s trace %d shot %d\n", tag, r, b);
                            ok = false;
                        }
                }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

// 3D adjoint: lam^k against a direct loop (bit for bit: the same FP32 ops in
// the same order), and the imaging sums against a direct FP64 sum (the kernel
// reduces over a lane group first, in another order).
template <int LB, int CX, int CY, int KIND = 0, int PX = 1, int PY = 1>
bool check_adj3d(int nz, int ny, int nx, int ng, int RZ, uint32_t seed) {
    const int Bp = LB * ng, W = 64 / LB;
    const int64_t N = (int64_t)nz * ny * nx, FE = N * Bp;
    std::mt19937 rng(seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "%s nz=%d ny=%d nx=%d LB=%d ng=%d C/T=%dx%d P=%dx%d RZ=%d",
                  KIND ? "adj3d-lds" : "adj3d", nz, ny, nx, LB, ng, CX, CY, PX, PY, RZ);
    std::uniform_real_distribution<float> u(0.02f, 0.1f), tt(-1e-9f, 1e-9f), d(-1.0f, 1.0f);
    float2 *sp2 = alloc<float2>(N, 1);
    for (int64_t c = 0; c < N; ++c) sp2[c] = make_float2(u(rng), tt(rng));
    float *l1 = alloc<float>(FE, Bp), *l2 = alloc<float>(FE, Bp), *pa = alloc<float>(FE, Bp),
          *lk = alloc<float>(FE, Bp);
    for (int64_t i = 0; i < FE; ++i) {
        l1[i] = d(rng);
        l2[i] = d(rng);
        pa[i] = d(rng);
    }
    int *src = alloc<int>(Bp, 0);
    std::uniform_int_distribution<int> cz(0, nz - 1), cy(0, ny - 1), cx(0, nx - 1);
    for (int b = 0; b < Bp; ++b) src[b] = (cz(rng) * ny + cy(rng)) * nx + cx(rng);
    src[Bp - 1] = -1;
    float *sinc = alloc<float>(Bp, 0), *srcd2 = alloc<float>(Bp, 0);
    for (int b = 0; b < Bp; ++b) { sinc[b] = d(rng); srcd2[b] = d(rng); }
    std::vector<int> rc = {0, (int)N - 1, (int)(N / 2), (int)(N / 3), (int)(N / 3)};
    const int n_rec = (int)rc.size();
    int *rec_head = alloc<int>(N, 0), *rec_next = alloc<int>(n_rec, 0);
    int *row_rec = alloc<int>((int64_t)nz * ny, 0);
    for (int64_t c = 0; c < N; ++c) rec_head[c] = -1;
    for (int j = n_rec - 1; j >= 0; --j) {
        rec_next[j] = rec_head[rc[j]];
        rec_head[rc[j]] = j;
    }
    for (int j = 0; j < n_rec; ++j) row_rec[rc[j] / nx]++;
    float *res = alloc<float>((int64_t)n_rec * Bp, 0);
    for (int64_t i = 0; i < (int64_t)n_rec * Bp; ++i) res[i] = d(rng);
    float *gh = alloc<float>(ng * N, 0), *gl = alloc<float>(ng * N, 0);
    if constexpr (KIND == 0) {
        const int nxs = (int)cdiv(nx, (int64_t)W * CX), nys = (int)cdiv(ny, CY);
        const int nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)cdiv(ng * tasks3d(nxs, nys, nzc, g_blk), 4);
        emu::launch(blocks, kBlock, [&] {
            k_adj3d<LB, CX, CY>(l1, l2, lk, pa, sp2, src, sinc, rec_head, rec_next, row_rec, res,
                                srcd2, gh, gl, nz, ny, nx, Bp, RZ, nxs, nys, nzc, ng, g_blk);
        });
    } else {
        const int ntx = (int)cdiv(nx, CX), nty = (int)cdiv(ny, CY), nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)((int64_t)ng * ntx * nty * nzc);
        emu::launch(blocks, kBlock, [&] {
            k_adj3d_lds<LB, CX, CY, PX, PY>(l1, l2, lk, pa, sp2, src, rec_head, rec_next, row_rec,
                                            res, srcd2, gh, gl, nz, ny, nx, Bp, RZ, ntx, nty, nzc,
                                            ng);
        });
    }
    bool ok = true;
    const int64_t sz = (int64_t)nx * ny;
    std::vector<double> gref(ng * N, 0.0);
    for (int64_t c = 0; c < N && ok; ++c) {
        const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / sz);
        const bool in = z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2;
        for (int b = 0; b < Bp; ++b) {
            auto q = [&](int64_t e) { return smul(sp2[e], l1[e * Bp + b]); };
            auto f = [&](int64_t e) { return pa[e * Bp + b]; };
            const float lq = in ? lap7(q(c), q(c + sz), q(c - sz), q(c + nx), q(c - nx), q(c + 1),
                                       q(c - 1))
                                : 0.0f;
            float v = leap(l1[c * Bp + b], l2[c * Bp + b], lq);
            for (int r = rec_head[c]; r >= 0; r = rec_next[r]) v = v - res[(int64_t)r * Bp + b];
            if (std::memcmp(&v, &lk[c * Bp + b], 4) != 0) {
                std::printf("MISMATCH %s lam cell %lld shot %d\n", tag, (long long)c, b);
                ok = false;
                break;
            }
            const float lp = in ? lap7(f(c), f(c + sz), f(c - sz), f(c + nx), f(c - nx), f(c + 1),
                                       f(c - 1))
                                : 0.0f;
            float d2 = smul(sp2[c], lp);
            if (c == src[b]) d2 = srcd2[b];
            gref[(b / LB) * N + c] += (double)(v * d2);
        }
    }
    double num = 0, den = 0;
    for (int64_t i = 0; i < ng * N; ++i) {
        const double e = ((double)gh[i] + (double)gl[i]) - gref[i];
        num += e * e;
        den += gref[i] * gref[i];
    }
    if (ok && !(num <= 1e-12 * den)) {
        std::printf("MISMATCH %s imaging sums rel err %.3e\n", tag, std::sqrt(num / den));
        ok = false;
    }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

// 3D adjoint with the forward run backwards (k_radj3d) against direct loops:
// lam^k bit for bit (the k_adj3d update), level k-2 hi + lo bit for bit (the
// FP64 leapfrog with p^k in the place of p^{n-2}), the imaging sums of
// lam^k * (float)inc against a direct FP64 sum; with wr = 0 level k-2 is left
// as it was.
template <int LB, int CX, int CY, class L = uint16_t, bool YSH = false>
bool check_radj3d(int nz, int ny, int nx, int ng, int RZ, int wr, uint32_t seed) {
    const int Bp = LB * ng, W = 64 / LB;
    const int64_t N = (int64_t)nz * ny * nx, FE = N * Bp;
    std::mt19937 rng(seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, sizeof(L) == 4 ? "radj3d-lo32 nz=%d ny=%d nx=%d LB=%d ng=%d C=%dx%d RZ=%d wr=%d ysh=%d" : "radj3d nz=%d ny=%d nx=%d LB=%d ng=%d C=%dx%d RZ=%d wr=%d ysh=%d", nz,
                  ny, nx, LB, ng, CX, CY, RZ, wr, (int)YSH);
    std::uniform_real_distribution<float> u(0.02f, 0.1f), tt(-1e-9f, 1e-9f), d(-1.0f, 1.0f),
        dl(-1e-7f, 1e-7f);
    float2 *sp2 = alloc<float2>(N, 1);
    for (int64_t c = 0; c < N; ++c) sp2[c] = make_float2(u(rng), tt(rng));
    float *l1 = alloc<float>(FE, Bp), *l2 = alloc<float>(FE, Bp), *lk = alloc<float>(FE, Bp);
    float *pc = alloc<float>(FE, Bp), *pk = alloc<float>(FE, Bp), *po = alloc<float>(FE, Bp);
    L *pcl = alloc<L>(FE, Bp), *pkl = alloc<L>(FE, Bp),
             *pol = alloc<L>(FE, Bp);
    for (int64_t i = 0; i < FE; ++i) {
        l1[i] = d(rng);
        l2[i] = d(rng);
        pc[i] = d(rng);
        pk[i] = d(rng);
        pcl[i] = mk_lo<L>(dl(rng));
        pkl[i] = mk_lo<L>(dl(rng));
        po[i] = 7.0f;
        pol[i] = 0x1234;
    }
    int *src = alloc<int>(Bp, 0);
    std::uniform_int_distribution<int> cz(0, nz - 1), cy(0, ny - 1), cx(0, nx - 1);
    for (int b = 0; b < Bp; ++b) src[b] = (cz(rng) * ny + cy(rng)) * nx + (b % 3 == 0 ? 0 : cx(rng));
    src[Bp - 1] = -1;
    double *sincd = alloc<double>(Bp, 0);
    for (int b = 0; b < Bp; ++b) sincd[b] = (double)d(rng) * 1.2345678901;
    std::vector<int> rc = {0, (int)N - 1, (int)(N / 2), (int)(N / 3), (int)(N / 3)};
    const int n_rec = (int)rc.size();
    int *rec_head = alloc<int>(N, 0), *rec_next = alloc<int>(n_rec, 0);
    int *row_rec = alloc<int>((int64_t)nz * ny, 0);
    for (int64_t c = 0; c < N; ++c) rec_head[c] = -1;
    for (int j = n_rec - 1; j >= 0; --j) {
        rec_next[j] = rec_head[rc[j]];
        rec_head[rc[j]] = j;
    }
    for (int j = 0; j < n_rec; ++j) row_rec[rc[j] / nx]++;
    float *res = alloc<float>((int64_t)n_rec * Bp, 0);
    for (int64_t i = 0; i < (int64_t)n_rec * Bp; ++i) res[i] = d(rng);
    float *gh = alloc<float>(ng * N, 0), *gl = alloc<float>(ng * N, 0);
    const int nxs = (int)cdiv(nx, (int64_t)W * CX), nys = (int)cdiv(ny, CY);
    const int nzc = (int)cdiv(nz, RZ);
    // YSH: one work-group per shot group, x strip, block of kYsh y strips and z chunk
    const unsigned blocks = YSH ? (unsigned)(ng * nxs * cdiv(nys, kYsh) * nzc)
                                : (unsigned)cdiv(ng * tasks3d(nxs, nys, nzc, g_blk), 4);
    emu::launch(blocks, kBlock, [&] {
        k_radj3d<LB, CX, CY, L, YSH>(l1, l2, lk, pc, pcl, pk, pkl, po, pol, wr, sp2, src, sincd,
                                     rec_head, rec_next, row_rec, res, gh, gl, nz, ny, nx, Bp, RZ,
                                     nxs, nys, nzc, ng, g_blk);
    });
    bool ok = true;
    const int64_t sz = (int64_t)nx * ny;
    std::vector<double> gref(ng * N, 0.0);
    auto Pv = [&](int64_t c, int b) { return (double)pc[c * Bp + b] + (double)lo_f(pcl[c * Bp + b]); };
    for (int64_t c = 0; c < N && ok; ++c) {
        const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / sz);
        const bool in = z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2;
        for (int b = 0; b < Bp; ++b) {
            auto q = [&](int64_t e) { return smul(sp2[e], l1[e * Bp + b]); };
            const float lq = in ? lap7(q(c), q(c + sz), q(c - sz), q(c + nx), q(c - nx), q(c + 1),
                                       q(c - 1))
                                : 0.0f;
            float v = leap(l1[c * Bp + b], l2[c * Bp + b], lq);
            for (int r = rec_head[c]; r >= 0; r = rec_next[r]) v = v - res[(int64_t)r * Bp + b];
            if (std::memcmp(&v, &lk[c * Bp + b], 4) != 0) {
                std::printf("MISMATCH %s lam cell %lld shot %d\n", tag, (long long)c, b);
                ok = false;
                break;
            }
            double lu = 0.0;
            if (in)
                lu = lap7d(Pv(c, b), Pv(c + sz, b), Pv(c - sz, b), Pv(c + nx, b), Pv(c - nx, b),
                           Pv(c + 1, b), Pv(c - 1, b));
            double inc = ((double)sp2[c].x + (double)sp2[c].y) * lu;
            if (c == src[b]) inc = inc + sincd[b];
            const double qk = (double)pk[c * Bp + b] + (double)lo_f(pkl[c * Bp + b]);
            const double nv = (2.0 * Pv(c, b) - qk) + inc;
            float pv = 7.0f;
            L lb = 0x1234;
            if (wr) {
                pv = (float)nv;
                lb = mk_lo<L>((float)(nv - (double)pv));
            }
            if (std::memcmp(&pv, &po[c * Bp + b], 4) != 0 || lb != pol[c * Bp + b]) {
                std::printf("MISMATCH %s level k-2 cell %lld shot %d\n", tag, (long long)c, b);
                ok = false;
                break;
            }
            gref[(b / LB) * N + c] += (double)(v * (float)inc);
        }
    }
    double num = 0, den = 0;
    for (int64_t i = 0; i < ng * N; ++i) {
        const double e = ((double)gh[i] + (double)gl[i]) - gref[i];
        num += e * e;
        den += gref[i] * gref[i];
    }
    if (ok && !(num <= 1e-12 * den)) {
        std::printf("MISMATCH %s imaging sums rel err %.3e\n", tag, std::sqrt(num / den));
        ok = false;
    }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

// Fused 3D Gauss-Newton level (k_born3d) against k_fwd3d<REC=false> followed by
// k_born<false> on the same inputs: background and Born levels (hi and lo),
// srcd2 and the Born traces, bit for bit.
template <int LB, int CX, int CY>
bool check_born3d(int nz, int ny, int nx, int ng, int RZ, uint32_t seed) {
    const int Bp = LB * ng, W = 64 / LB;
    const int64_t N = (int64_t)nz * ny * nx, FE = N * Bp;
    std::mt19937 rng(seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "born3d nz=%d ny=%d nx=%d LB=%d ng=%d C=%dx%d RZ=%d", nz, ny, nx,
                  LB, ng, CX, CY, RZ);
    std::uniform_real_distribution<float> u(0.02f, 0.1f), tt(-1e-9f, 1e-9f), d(-1.0f, 1.0f),
        dl(-1e-7f, 1e-7f);
    float2 *sp2 = alloc<float2>(N, 1), *vm2 = alloc<float2>(N, 1);
    for (int64_t c = 0; c < N; ++c) {
        sp2[c] = make_float2(u(rng), tt(rng));
        vm2[c] = make_float2(d(rng), tt(rng));
    }
    auto fl = [&](bool rnd) {
        float *p = alloc<float>(FE, Bp);
        if (rnd) for (int64_t i = 0; i < FE; ++i) p[i] = d(rng);
        return p;
    };
    auto lo = [&](bool rnd) {
        uint16_t *p = alloc<uint16_t>(FE, Bp);
        if (rnd) for (int64_t i = 0; i < FE; ++i) p[i] = f2bf(dl(rng));
        return p;
    };
    float *pc = fl(true), *pp = fl(true), *dc = fl(true), *dp = fl(true);
    uint16_t *pcl = lo(true), *ppl = lo(true), *dcl = lo(true), *dpl = lo(true);
    float *pn_a = fl(false), *pn_b = fl(false), *dn_a = fl(false), *dn_b = fl(false);
    uint16_t *pnl_a = lo(false), *pnl_b = lo(false), *dnl_a = lo(false), *dnl_b = lo(false);
    int *src = alloc<int>(Bp, 0);
    std::uniform_int_distribution<int> cz(0, nz - 1), cy(0, ny - 1), cx(0, nx - 1);
    for (int b = 0; b < Bp; ++b) src[b] = (cz(rng) * ny + cy(rng)) * nx + (b % 3 == 0 ? 0 : cx(rng));
    src[Bp - 1] = -1;
    double *sincd = alloc<double>(Bp, 0);
    for (int b = 0; b < Bp; ++b) sincd[b] = (double)d(rng) * 1.2345678901;
    std::vector<int> rc = {0, (int)N - 1, (int)(N / 2), (int)(N / 3), (int)(N / 3)};
    const int n_rec = (int)rc.size();
    int *rec_head = alloc<int>(N, 0), *rec_next = alloc<int>(n_rec, 0);
    int *row_rec = alloc<int>((int64_t)nz * ny, 0);
    for (int64_t c = 0; c < N; ++c) rec_head[c] = -1;
    for (int j = n_rec - 1; j >= 0; --j) {
        rec_next[j] = rec_head[rc[j]];
        rec_head[rc[j]] = j;
    }
    for (int j = 0; j < n_rec; ++j) row_rec[rc[j] / nx]++;
    float *res_a = alloc<float>((int64_t)n_rec * Bp, 0), *res_b = alloc<float>((int64_t)n_rec * Bp, 0);
    float *s2_a = alloc<float>(Bp, 0), *s2_b = alloc<float>(Bp, 0);
    const int nxs = (int)cdiv(nx, (int64_t)W * CX), nys = (int)cdiv(ny, CY);
    const int nzc = (int)cdiv(nz, RZ);
    const unsigned blocks = (unsigned)cdiv(ng * tasks3d(nxs, nys, nzc, g_blk), 4);
    emu::launch(blocks, kBlock, [&] {
        k_born3d<LB, CX, CY>(pc, pp, pn_a, pcl, ppl, pnl_a, dc, dp, dn_a, dcl, dpl, dnl_a, sp2, vm2,
                             src, sincd, rec_head, rec_next, row_rec, res_a, s2_a, nz, ny, nx, Bp,
                             RZ, nxs, nys, nzc, ng, g_blk);
    });
    emu::launch(blocks, kBlock, [&] {
        k_fwd3d<LB, CX, CY, false>(pc, pp, pn_b, pcl, ppl, pnl_b, sp2, src, sincd, rec_head, rec_next,
                                   row_rec, nullptr, nullptr, s2_b, nz, ny, nx, Bp, RZ, nxs, nys, nzc,
                                   ng, g_blk);
    });
    emu::launch((unsigned)cdiv(FE, 256), 256, [&] {
        k_born<false>(pn_b, pnl_b, pc, pcl, pp, ppl, dc, dcl, dp, dpl, dn_b, dnl_b, sp2, vm2, nullptr,
                      rec_head, rec_next, res_b, nz, ny, nx, 3, Bp);
    });
    bool ok = same("background hi", pn_a, pn_b, FE, tag) && same("background lo", pnl_a, pnl_b, FE, tag) &&
              same("Born hi", dn_a, dn_b, FE, tag) && same("Born lo", dnl_a, dnl_b, FE, tag) &&
              same("Born traces", res_a, res_b, (int64_t)n_rec * Bp, tag);
    for (int b = 0; b < Bp && ok; ++b)
        if (src[b] >= 0 && std::memcmp(&s2_a[b], &s2_b[b], 4) != 0) {
            std::printf("MISMATCH %s srcd2 shot %d\n", tag, b);
            ok = false;
        }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

template <int LB, int CX, int CY>
bool born3d_cases(uint32_t seed) {
    bool ok = true;
    ok &= check_born3d<LB, CX, CY>(9, 7, 11, 2, 4, seed++);
    ok &= check_born3d<LB, CX, CY>(5, 6, 13, 1, 64, seed++);
    ok &= check_born3d<LB, CX, CY>(12, 9, 5, 2, 8, seed++);
    ok &= check_born3d<LB, CX, CY>(3, 3, 3, 1, 2, seed++);
    return ok;
}

template <int LB, int CX, int CY, bool YSH = false>
bool radj3d_cases(uint32_t seed) {
    bool ok = true;
    ok &= check_radj3d<LB, CX, CY, uint16_t, YSH>(9, 7, 11, 2, 4, 1, seed++);
    ok &= check_radj3d<LB, CX, CY, uint16_t, YSH>(5, 6, 13, 1, 64, 1, seed++);
    ok &= check_radj3d<LB, CX, CY, uint16_t, YSH>(12, 9, 5, 2, 8, 0, seed++);
    ok &= check_radj3d<LB, CX, CY, uint16_t, YSH>(3, 3, 3, 1, 2, 1, seed++);
    if (YSH) {
        // y extents that leave the last block of y strips part-empty, and more
        // than one block of strips along y
        ok &= check_radj3d<LB, CX, CY, uint16_t, YSH>(6, 21, 7, 1, 3, 1, seed++);
        ok &= check_radj3d<LB, CX, CY, uint16_t, YSH>(7, 18, 9, 2, 5, 1, seed++);
        ok &= check_radj3d<LB, CX, CY, float, YSH>(8, 17, 6, 1, 4, 1, seed++);
    }
    return ok;
}

template <int LB, int CX, int CY>
bool fwd3d_cases(uint32_t seed) {
    bool ok = true;
    ok &= check_fwd3d<LB, CX, CY>(9, 7, 11, 2, 4, seed++);
    ok &= check_fwd3d<LB, CX, CY>(5, 6, 13, 1, 64, seed++);
    ok &= check_fwd3d<LB, CX, CY>(12, 9, 5, 2, 8, seed++);
    if (CX * CY <= LB) {  // the adjoint reduces a lane's cells over its lane group
        ok &= check_adj3d<LB, CX, CY>(9, 7, 11, 2, 4, seed++);
        ok &= check_adj3d<LB, CX, CY>(5, 6, 13, 1, 64, seed++);
    }
    return ok;
}

// Two-level 3D forward on LDS tiles against two single-level k_fwd3d launches:
// levels n and n+1 (hi and lo), both levels' traces and recorded source
// increments, bit for bit.
template <int LB, int TX, int TY, int PX, int PY>
bool check_fwd3d_x2(int nz, int ny, int nx, int ng, int RZ, uint32_t seed) {
    const int Bp = LB * ng, W = 64 / LB;
    const int64_t N = (int64_t)nz * ny * nx, FE = N * Bp;
    std::mt19937 rng(seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "fwd3d-x2-lds nz=%d ny=%d nx=%d LB=%d ng=%d T=%dx%d P=%dx%d RZ=%d",
                  nz, ny, nx, LB, ng, TX, TY, PX, PY, RZ);
    std::uniform_real_distribution<float> u(0.02f, 0.1f), tt(-1e-9f, 1e-9f), d(-1.0f, 1.0f),
        dl(-1e-7f, 1e-7f);
    float2 *sp2 = alloc<float2>(N, 1);
    for (int64_t c = 0; c < N; ++c) sp2[c] = make_float2(u(rng), tt(rng));
    auto field = [&](float *&h, uint16_t *&l, bool fill) {
        h = alloc<float>(FE, Bp);
        l = alloc<uint16_t>(FE, Bp);
        if (fill)
            for (int64_t i = 0; i < FE; ++i) {
                h[i] = d(rng);
                l[i] = f2bf(dl(rng));
            }
    };
    float *A, *B, *C1, *D1, *C2, *D2;
    uint16_t *Al, *Bl, *C1l, *D1l, *C2l, *D2l;
    field(A, Al, true);
    field(B, Bl, true);
    field(C1, C1l, false);
    field(D1, D1l, false);
    field(C2, C2l, false);
    field(D2, D2l, false);
    int *src = alloc<int>(Bp, 0);
    std::uniform_int_distribution<int> cz(0, nz - 1), cy(0, ny - 1), cx(0, nx - 1);
    for (int b = 0; b < Bp; ++b) {
        int x = cx(rng), y = cy(rng);
        if (b % 5 == 0) x = 0;
        if (b % 5 == 1) x = std::min(TX, nx - 1);       // a tile edge
        if (b % 5 == 2) y = std::min(TY - 1, ny - 1);
        src[b] = (cz(rng) * ny + y) * nx + x;
    }
    src[Bp - 1] = -1;
    double *sa = alloc<double>(Bp, 0), *sb = alloc<double>(Bp, 0);
    for (int b = 0; b < Bp; ++b) {
        sa[b] = (double)d(rng) * 1.2345678901;
        sb[b] = (double)d(rng) * 0.9876543211;
    }
    float *da1 = alloc<float>(Bp, 0), *db1 = alloc<float>(Bp, 0), *da2 = alloc<float>(Bp, 0),
          *db2 = alloc<float>(Bp, 0);
    std::vector<int> rc = {0, (int)N - 1, (int)(N / 2), (int)(N / 3), (int)(N / 3),
                           (int)std::min<int64_t>(N - 1, (int64_t)TX)};
    const int n_rec = (int)rc.size();
    int *rec_head = alloc<int>(N, 0), *rec_next = alloc<int>(n_rec, 0);
    int *row_rec = alloc<int>((int64_t)nz * ny, 0);
    for (int64_t c = 0; c < N; ++c) rec_head[c] = -1;
    for (int j = n_rec - 1; j >= 0; --j) {
        rec_next[j] = rec_head[rc[j]];
        rec_head[rc[j]] = j;
    }
    for (int j = 0; j < n_rec; ++j) row_rec[rc[j] / nx]++;
    const int64_t NR = (int64_t)n_rec * Bp;
    float *ra1 = alloc<float>(NR, 0), *ral1 = alloc<float>(NR, 0), *rb1 = alloc<float>(NR, 0),
          *rbl1 = alloc<float>(NR, 0), *ra2 = alloc<float>(NR, 0), *ral2 = alloc<float>(NR, 0),
          *rb2 = alloc<float>(NR, 0), *rbl2 = alloc<float>(NR, 0);
    // reference: two single-level launches (2 x 2 patches)
    {
        constexpr int CXr = 2, CYr = 2;
        const int nxs = (int)cdiv(nx, (int64_t)W * CXr), nys = (int)cdiv(ny, CYr);
        const int nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)cdiv(ng * tasks3d(nxs, nys, nzc, g_blk), 4);
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d<LB, CXr, CYr, true>(A, B, C1, Al, Bl, C1l, sp2, src, sa, rec_head, rec_next,
                                        row_rec, ra1, ral1, da1, nz, ny, nx, Bp, RZ, nxs, nys, nzc,
                                        ng, g_blk);
        });
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d<LB, CXr, CYr, true>(C1, A, D1, C1l, Al, D1l, sp2, src, sb, rec_head, rec_next,
                                        row_rec, rb1, rbl1, db1, nz, ny, nx, Bp, RZ, nxs, nys, nzc,
                                        ng, g_blk);
        });
    }
    {
        const int ntx = (int)cdiv(nx, TX), nty = (int)cdiv(ny, TY), nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)((int64_t)ng * ntx * nty * nzc);
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d_x2_lds<LB, TX, TY, PX, PY, true>(
                A, B, C2, D2, Al, Bl, C2l, D2l, sp2, src, sa, sb, rec_head, rec_next, row_rec, ra2,
                ral2, rb2, rbl2, da2, db2, nz, ny, nx, Bp, RZ, ntx, nty, nzc, ng);
        });
    }
    bool ok = std::memcmp(C1, C2, sizeof(float) * FE) == 0 &&
              std::memcmp(C1l, C2l, sizeof(uint16_t) * FE) == 0;
    if (!ok) {
        for (int64_t i = 0; i < FE; ++i)
            if (std::memcmp(&C1[i], &C2[i], 4) || C1l[i] != C2l[i]) {
                std::printf("MISMATCH %s level n at cell %lld shot %lld\n", tag, (long long)(i / Bp),
                            (long long)(i % Bp));
                break;
            }
    }
    if (ok && (std::memcmp(D1, D2, sizeof(float) * FE) || std::memcmp(D1l, D2l, sizeof(uint16_t) * FE))) {
        ok = false;
        for (int64_t i = 0; i < FE; ++i)
            if (std::memcmp(&D1[i], &D2[i], 4) || D1l[i] != D2l[i]) {
                std::printf("MISMATCH %s level n+1 at cell %lld shot %lld\n", tag,
                            (long long)(i / Bp), (long long)(i % Bp));
                break;
            }
    }
    if (ok && (std::memcmp(ra1, ra2, 4 * NR) || std::memcmp(ral1, ral2, 4 * NR) ||
               std::memcmp(rb1, rb2, 4 * NR) || std::memcmp(rbl1, rbl2, 4 * NR))) {
        std::printf("MISMATCH %s traces\n", tag);
        ok = false;
    }
    if (ok && (std::memcmp(da1, da2, 4 * Bp) || std::memcmp(db1, db2, 4 * Bp))) {
        std::printf("MISMATCH %s recorded source increments\n", tag);
        ok = false;
    }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

// Two-level 3D forward on lane patches (k_fwd3d_x2, z marched in registers) against
// two single-level k_fwd3d launches with 2 x 2 patches:
// levels n and n+1 (hi and lo), both levels' traces and recorded source
// increments, bit for bit.
template <int LB, int CX, int CY>
bool check_fwd3d_x2p(int nz, int ny, int nx, int ng, int RZ, uint32_t seed) {
    const int Bp = LB * ng, W = 64 / LB;
    const int64_t N = (int64_t)nz * ny * nx, FE = N * Bp;
    std::mt19937 rng(seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "fwd3d-x2-patch nz=%d ny=%d nx=%d LB=%d ng=%d P=%dx%d RZ=%d", nz,
                  ny, nx, LB, ng, CX, CY, RZ);
    constexpr int TX = 2 * CX, TY = 2 * CY;  // source and receiver cells on patch edges
    std::uniform_real_distribution<float> u(0.02f, 0.1f), tt(-1e-9f, 1e-9f), d(-1.0f, 1.0f),
        dl(-1e-7f, 1e-7f);
    float2 *sp2 = alloc<float2>(N, 1);
    for (int64_t c = 0; c < N; ++c) sp2[c] = make_float2(u(rng), tt(rng));
    auto field = [&](float *&h, uint16_t *&l, bool fill) {
        h = alloc<float>(FE, Bp);
        l = alloc<uint16_t>(FE, Bp);
        if (fill)
            for (int64_t i = 0; i < FE; ++i) {
                h[i] = d(rng);
                l[i] = f2bf(dl(rng));
            }
    };
    float *A, *B, *C1, *D1, *C2, *D2;
    uint16_t *Al, *Bl, *C1l, *D1l, *C2l, *D2l;
    field(A, Al, true);
    field(B, Bl, true);
    field(C1, C1l, false);
    field(D1, D1l, false);
    field(C2, C2l, false);
    field(D2, D2l, false);
    int *src = alloc<int>(Bp, 0);
    std::uniform_int_distribution<int> cz(0, nz - 1), cy(0, ny - 1), cx(0, nx - 1);
    for (int b = 0; b < Bp; ++b) {
        int x = cx(rng), y = cy(rng);
        if (b % 5 == 0) x = 0;
        if (b % 5 == 1) x = std::min(TX, nx - 1);       // a tile edge
        if (b % 5 == 2) y = std::min(TY - 1, ny - 1);
        src[b] = (cz(rng) * ny + y) * nx + x;
    }
    src[Bp - 1] = -1;
    double *sa = alloc<double>(Bp, 0), *sb = alloc<double>(Bp, 0);
    for (int b = 0; b < Bp; ++b) {
        sa[b] = (double)d(rng) * 1.2345678901;
        sb[b] = (double)d(rng) * 0.9876543211;
    }
    float *da1 = alloc<float>(Bp, 0), *db1 = alloc<float>(Bp, 0), *da2 = alloc<float>(Bp, 0),
          *db2 = alloc<float>(Bp, 0);
    std::vector<int> rc = {0, (int)N - 1, (int)(N / 2), (int)(N / 3), (int)(N / 3),
                           (int)std::min<int64_t>(N - 1, (int64_t)TX)};
    const int n_rec = (int)rc.size();
    int *rec_head = alloc<int>(N, 0), *rec_next = alloc<int>(n_rec, 0);
    int *row_rec = alloc<int>((int64_t)nz * ny, 0);
    for (int64_t c = 0; c < N; ++c) rec_head[c] = -1;
    for (int j = n_rec - 1; j >= 0; --j) {
        rec_next[j] = rec_head[rc[j]];
        rec_head[rc[j]] = j;
    }
    for (int j = 0; j < n_rec; ++j) row_rec[rc[j] / nx]++;
    const int64_t NR = (int64_t)n_rec * Bp;
    float *ra1 = alloc<float>(NR, 0), *ral1 = alloc<float>(NR, 0), *rb1 = alloc<float>(NR, 0),
          *rbl1 = alloc<float>(NR, 0), *ra2 = alloc<float>(NR, 0), *ral2 = alloc<float>(NR, 0),
          *rb2 = alloc<float>(NR, 0), *rbl2 = alloc<float>(NR, 0);
    // reference: two single-level launches (2 x 2 patches)
    {
        constexpr int CXr = 2, CYr = 2;
        const int nxs = (int)cdiv(nx, (int64_t)W * CXr), nys = (int)cdiv(ny, CYr);
        const int nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)cdiv(ng * tasks3d(nxs, nys, nzc, g_blk), 4);
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d<LB, CXr, CYr, true>(A, B, C1, Al, Bl, C1l, sp2, src, sa, rec_head, rec_next,
                                        row_rec, ra1, ral1, da1, nz, ny, nx, Bp, RZ, nxs, nys, nzc,
                                        ng, g_blk);
        });
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d<LB, CXr, CYr, true>(C1, A, D1, C1l, Al, D1l, sp2, src, sb, rec_head, rec_next,
                                        row_rec, rb1, rbl1, db1, nz, ny, nx, Bp, RZ, nxs, nys, nzc,
                                        ng, g_blk);
        });
    }
    {
        const int nxs = (int)cdiv(nx, (int64_t)W * CX), nys = (int)cdiv(ny, CY);
        const int nzc = (int)cdiv(nz, RZ);
        const unsigned blocks = (unsigned)cdiv(ng * tasks3d(nxs, nys, nzc, g_blk), 4);
        emu::launch(blocks, kBlock, [&] {
            k_fwd3d_x2<LB, CX, CY, true>(A, B, C2, D2, Al, Bl, C2l, D2l, sp2, src, sa, sb, rec_head,
                                         rec_next, row_rec, ra2, ral2, rb2, rbl2, da2, db2, nz, ny,
                                         nx, Bp, RZ, nxs, nys, nzc, ng, g_blk);
        });
    }
    bool ok = std::memcmp(C1, C2, sizeof(float) * FE) == 0 &&
              std::memcmp(C1l, C2l, sizeof(uint16_t) * FE) == 0;
    if (!ok) {
        for (int64_t i = 0; i < FE; ++i)
            if (std::memcmp(&C1[i], &C2[i], 4) || C1l[i] != C2l[i]) {
                std::printf("MISMATCH %s level n at cell %lld shot %lld\n", tag, (long long)(i / Bp),
                            (long long)(i % Bp));
                break;
            }
    }
    if (ok && (std::memcmp(D1, D2, sizeof(float) * FE) || std::memcmp(D1l, D2l, sizeof(uint16_t) * FE))) {
        ok = false;
        for (int64_t i = 0; i < FE; ++i)
            if (std::memcmp(&D1[i], &D2[i], 4) || D1l[i] != D2l[i]) {
                std::printf("MISMATCH %s level n+1 at cell %lld shot %lld\n", tag,
                            (long long)(i / Bp), (long long)(i % Bp));
                break;
            }
    }
    if (ok && (std::memcmp(ra1, ra2, 4 * NR) || std::memcmp(ral1, ral2, 4 * NR) ||
               std::memcmp(rb1, rb2, 4 * NR) || std::memcmp(rbl1, rbl2, 4 * NR))) {
        std::printf("MISMATCH %s traces\n", tag);
        ok = false;
    }
    if (ok && (std::memcmp(da1, da2, 4 * Bp) || std::memcmp(db1, db2, 4 * Bp))) {
        std::printf("MISMATCH %s recorded source increments\n", tag);
        ok = false;
    }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

// FP32 low parts (USCT_FLAG_FP32_LO): the 2D forward against its direct FP64 loop,
// the two-level 2D forward against two single-level launches, the 3D forward and
// the reversing 3D adjoint against their direct loops.
bool lo32_cases(uint32_t seed) {
    bool ok = true;
    ok &= check_fwd_direct<64, 4, false, float>(29, 37, 2, 8, seed++);
    ok &= check_fwd_direct<16, 8, true, float>(9, 13, 1, 4, seed++);
    for (int sched = 0; sched < 8; ++sched) {
        ok &= check_fwd<64, 4, true, float>(29, 37, 2, 8, seed++, sched);
        ok &= check_fwd<16, 8, true, float>(40, 33, 2, 4, seed++, sched);
    }
    ok &= check_fwd<32, 4, true, float>(3, 5, 2, 64, seed++, 0);
    ok &= check_fwd3d<64, 2, 2, 0, 1, 1, float>(9, 7, 11, 2, 4, seed++);
    ok &= check_fwd3d<16, 1, 2, 0, 1, 1, float>(5, 6, 13, 1, 64, seed++);
    ok &= check_radj3d<64, 2, 2, float>(9, 7, 11, 2, 4, 1, seed++);
    ok &= check_radj3d<32, 1, 1, float>(12, 9, 5, 2, 8, 0, seed++);
    ok &= check_radj3d<16, 2, 1, float>(5, 6, 13, 1, 64, 1, seed++);
    return ok;
}

bool x2p_3d_cases(uint32_t seed) {
    bool ok = true;
    ok &= check_fwd3d_x2p<64, 1, 1>(9, 7, 11, 2, 4, seed++);
    ok &= check_fwd3d_x2p<64, 1, 1>(12, 17, 19, 1, 64, seed++);
    ok &= check_fwd3d_x2p<64, 2, 1>(7, 9, 21, 2, 3, seed++);
    ok &= check_fwd3d_x2p<64, 1, 2>(5, 6, 13, 1, 2, seed++);
    ok &= check_fwd3d_x2p<32, 1, 1>(3, 3, 3, 1, 1, seed++);
    ok &= check_fwd3d_x2p<32, 2, 1>(10, 16, 8, 1, 5, seed++);
    ok &= check_fwd3d_x2p<16, 1, 1>(8, 5, 6, 2, 4, seed++);
    ok &= check_fwd3d_x2p<16, 1, 2>(6, 9, 7, 1, 8, seed++);
    ok &= check_fwd3d_x2p<16, 2, 1>(11, 4, 9, 2, 16, seed++);
    ok &= check_fwd3d_x2p<64, 2, 1>(4, 4, 4, 1, 1, seed++);
    return ok;
}

bool x2_3d_cases(uint32_t seed) {
    bool ok = true;
    ok &= check_fwd3d_x2<16, 8, 8, 2, 2>(9, 7, 11, 2, 4, seed++);
    ok &= check_fwd3d_x2<16, 8, 8, 2, 2>(12, 17, 19, 1, 64, seed++);
    ok &= check_fwd3d_x2<16, 8, 8, 2, 2>(7, 9, 21, 2, 3, seed++);
    ok &= check_fwd3d_x2<16, 8, 8, 2, 2>(5, 6, 13, 1, 2, seed++);
    ok &= check_fwd3d_x2<16, 8, 8, 2, 2>(3, 3, 3, 1, 1, seed++);
    ok &= check_fwd3d_x2<16, 8, 8, 2, 2>(10, 16, 8, 1, 5, seed++);
    return ok;
}

bool lds_cases(uint32_t seed) {
    bool ok = true;
    // tiles that do and do not divide the grid, shot groups of 16 / 32 / 64
    ok &= check_fwd3d<16, 8, 8, 1, 2, 2>(9, 7, 11, 2, 4, seed++);
    ok &= check_fwd3d<16, 8, 8, 1, 2, 2>(12, 17, 19, 1, 64, seed++);
    ok &= check_fwd3d<16, 16, 8, 1, 4, 2>(7, 9, 21, 2, 3, seed++);
    ok &= check_fwd3d<32, 8, 8, 1, 4, 2>(6, 10, 9, 2, 8, seed++);
    ok &= check_fwd3d<64, 8, 4, 1, 4, 2>(5, 6, 13, 1, 4, seed++);
    ok &= check_fwd3d<16, 8, 8, 1, 2, 2>(3, 3, 3, 1, 2, seed++);
    ok &= check_adj3d<16, 8, 8, 1, 2, 2>(9, 7, 11, 2, 4, seed++);
    ok &= check_adj3d<16, 8, 8, 1, 2, 2>(12, 17, 19, 1, 64, seed++);
    ok &= check_adj3d<16, 16, 8, 1, 4, 2>(7, 9, 21, 2, 3, seed++);
    ok &= check_adj3d<32, 8, 8, 1, 4, 2>(6, 10, 9, 2, 8, seed++);
    ok &= check_adj3d<64, 8, 4, 1, 4, 2>(5, 6, 13, 1, 4, seed++);
    ok &= check_adj3d<16, 8, 8, 1, 2, 2>(3, 3, 3, 1, 2, seed++);
    return ok;
}

// Born step: every shot-cell of a grid (2D: nz = 1) against a direct FP64 loop
// of its update; fields carry no guard zone, so under AddressSanitizer any
// neighbour read outside the field aborts.
bool check_born(int nz, int ny, int nx, int Bp, int ndim, bool sponge, uint32_t seed) {
    std::mt19937 rng(seed);
    const int64_t N = (int64_t)nz * ny * nx, FE = N * Bp;
    std::uniform_real_distribution<float> val(-1.0f, 1.0f), u(0.05f, 0.24f), t(-1e-9f, 1e-9f);
    auto field = [&](float *&h, uint16_t *&l) {
        h = alloc<float>(FE, 0);
        l = alloc<uint16_t>(FE, 0);
        for (int64_t i = 0; i < FE; ++i) {
            h[i] = val(rng);
            l[i] = f2bf(1e-8f * val(rng));
        }
    };
    float *u0, *u1, *u2, *d1, *d2;
    uint16_t *u0l, *u1l, *u2l, *d1l, *d2l;
    field(u0, u0l);
    field(u1, u1l);
    field(u2, u2l);
    field(d1, d1l);
    field(d2, d2l);
    float *dn = alloc<float>(FE, 0);
    uint16_t *dnl = alloc<uint16_t>(FE, 0);
    float2 *sp2 = alloc<float2>(N, 0), *vm2 = alloc<float2>(N, 0);
    float *spg = alloc<float>(N, 0);
    for (int64_t c = 0; c < N; ++c) {
        sp2[c] = make_float2(u(rng), t(rng));
        vm2[c] = make_float2(val(rng), t(rng));
        spg[c] = sponge ? 0.5f + 0.5f * u(rng) : 1.0f;
    }
    // receivers: the first and last cell, a duplicate, random cells
    const int n_rec = 9;
    std::vector<int64_t> rc(n_rec);
    std::uniform_int_distribution<int64_t> cell(0, N - 1);
    for (int j = 0; j < n_rec; ++j) rc[j] = cell(rng);
    rc[0] = 0;
    rc[1] = N - 1;
    rc[2] = rc[3];
    int *rec_head = alloc<int>(N, 0), *rec_next = alloc<int>(n_rec, 0);
    for (int64_t c = 0; c < N; ++c) rec_head[c] = -1;
    for (int j = n_rec - 1; j >= 0; --j) {
        rec_next[j] = rec_head[rc[j]];
        rec_head[rc[j]] = j;
    }
    float *res = alloc<float>((int64_t)n_rec * Bp, 0);
    emu::launch((unsigned)cdiv(FE, 256), 256, [&] {
        if (sponge)
            k_born<true>(u0, u0l, u1, u1l, u2, u2l, d1, d1l, d2, d2l, dn, dnl, sp2, vm2, spg,
                         rec_head, rec_next, res, nz, ny, nx, ndim, Bp);
        else
            k_born<false>(u0, u0l, u1, u1l, u2, u2l, d1, d1l, d2, d2l, dn, dnl, sp2, vm2, spg,
                          rec_head, rec_next, res, nz, ny, nx, ndim, Bp);
    });
    char tag[96];
    std::snprintf(tag, sizeof tag, "born %dD %dx%dx%d Bp=%d sponge=%d", ndim, nz, ny, nx, Bp,
                  (int)sponge);
    bool ok = true;
    const int64_t sy = (int64_t)nx, sz = (int64_t)nx * ny;
    for (int64_t c = 0; c < N && ok; ++c) {
        const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / sz);
        for (int b = 0; b < Bp; ++b) {
            auto D = [&](int64_t e) { return hl(d1[e * Bp + b], d1l[e * Bp + b]); };
            auto U = [&](const float *h, const uint16_t *l) {
                return hl(h[c * Bp + b], l[c * Bp + b]);
            };
            double lu = 0.0;
            if (ndim == 3) {
                if (z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2)
                    lu = lap7d(D(c), D(c + sz), D(c - sz), D(c + sy), D(c - sy), D(c + 1), D(c - 1));
            } else if (y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2) {
                lu = lap5d(D(c), D(c + sy), D(c - sy), D(c + 1), D(c - 1));
            }
            const double d2u = (U(u0, u0l) - 2.0 * U(u1, u1l)) + U(u2, u2l);
            double nv = (2.0 * D(c) - U(d2, d2l)) + s2d(sp2[c]) * lu - s2d(vm2[c]) * d2u;
            if (sponge) nv = nv * (double)spg[c];
            uint16_t lb;
            const float v = split_hl(nv, lb);
            if (std::memcmp(&v, &dn[c * Bp + b], 4) != 0 || lb != dnl[c * Bp + b]) {
                std::printf("MISMATCH %s cell %lld shot %d\n", tag, (long long)c, b);
                ok = false;
                break;
            }
        }
    }
    // each receiver holds its cell's new level (duplicates both)
    for (int j = 0; j < n_rec && ok; ++j)
        for (int b = 0; b < Bp; ++b)
            if (std::memcmp(&res[(int64_t)j * Bp + b], &dn[rc[j] * Bp + b], 4) != 0) {
                std::printf("MISMATCH %s receiver %d shot %d\n", tag, j, b);
                ok = false;
                break;
            }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

bool born_cases(uint32_t seed) {
    bool ok = true;
    ok &= check_born(1, 13, 11, 16, 2, false, seed++);
    ok &= check_born(1, 9, 7, 3, 2, true, seed++);
    ok &= check_born(1, 5, 6, 1, 2, true, seed++);  // plain layout
    ok &= check_born(7, 9, 11, 4, 3, false, seed++);
    ok &= check_born(5, 6, 3, 1, 3, false, seed++);
    return ok;
}

// Gauss-Newton two-level launch: background and Born levels n, n+1 of one
// k_born2d_x2 launch against k_fwd2d, k_born, k_fwd2d, k_born (fields, low
// parts and the Born traces of both levels), fields with the library's guard
// zones.
template <int LB, int CX>
bool check_born_x2(int ny, int nx, int ng, int RY, uint32_t seed) {
    Problem P(ny, nx, LB, ng, CX, RY, seed);
    char tag[160];
    std::snprintf(tag, sizeof tag, "born-x2 ny=%d nx=%d LB=%d ng=%d CX=%d RY=%d", ny, nx, LB, ng, CX,
                  RY);
    float *pc = P.field(1.0f), *pp = P.field(1.0f), *dc = P.field(0.3f), *dp = P.field(0.3f);
    uint16_t *pcl = P.low(1e-7f), *ppl = P.low(1e-7f), *dcl = P.low(1e-8f), *dpl = P.low(1e-8f);
    float2 *vm2 = alloc<float2>(P.N, 1);
    std::uniform_real_distribution<float> v(-2.0f, 2.0f), t(-1e-9f, 1e-9f);
    for (int64_t c = 0; c < P.N; ++c) vm2[c] = make_float2(v(P.rng), t(P.rng));
    float *a1 = P.zeros(), *b1 = P.zeros(), *a2 = P.zeros(), *b2 = P.zeros();
    float *ea1 = P.zeros(), *eb1 = P.zeros(), *ea2 = P.zeros(), *eb2 = P.zeros();
    uint16_t *a1l = P.zeros_lo(), *b1l = P.zeros_lo(), *a2l = P.zeros_lo(), *b2l = P.zeros_lo();
    uint16_t *ea1l = P.zeros_lo(), *eb1l = P.zeros_lo(), *ea2l = P.zeros_lo(), *eb2l = P.zeros_lo();
    float *ra1 = P.traces(), *rb1 = P.traces(), *ra2 = P.traces(), *rb2 = P.traces();
    const double *sa = P.sincd + P.Bp, *sb = P.sincd + 2 * P.Bp;
    const unsigned gb = (unsigned)cdiv(P.FE, 256);
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d<LB, CX, false, false>(pc, pp, a1, pcl, ppl, a1l, P.sp2, nullptr, P.src_cell, sa,
                                      P.rec_head, P.rec_next, P.row_rec, nullptr, nullptr, P.ny,
                                      P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
    });
    emu::launch(gb, 256, [&] {
        k_born<false>(a1, a1l, pc, pcl, pp, ppl, dc, dcl, dp, dpl, ea1, ea1l, P.sp2, vm2, nullptr,
                      P.rec_head, P.rec_next, ra1, 1, P.ny, P.nx, 2, P.Bp);
    });
    emu::launch(P.blocks, kBlock, [&] {
        k_fwd2d<LB, CX, false, false>(a1, pc, b1, a1l, pcl, b1l, P.sp2, nullptr, P.src_cell, sb,
                                      P.rec_head, P.rec_next, P.row_rec, nullptr, nullptr, P.ny,
                                      P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
    });
    emu::launch(gb, 256, [&] {
        k_born<false>(b1, b1l, a1, a1l, pc, pcl, ea1, ea1l, dc, dcl, eb1, eb1l, P.sp2, vm2,
                      nullptr, P.rec_head, P.rec_next, rb1, 1, P.ny, P.nx, 2, P.Bp);
    });
    emu::launch(P.blocks, kBlock, [&] {
        k_born2d_x2<LB, CX>(pc, pp, a2, b2, pcl, ppl, a2l, b2l, dc, dp, ea2, eb2, dcl, dpl, ea2l,
                            eb2l, P.sp2, vm2, P.src_cell, sa, sb, P.rec_head, P.rec_next,
                            P.row_rec, ra2, rb2, P.ny, P.nx, P.Bp, P.RY, P.nxs, P.nyc, P.ng);
    });
    const int64_t nr = (int64_t)P.n_rec * P.Bp;
    bool ok = same("background n", a1, a2, P.FE, tag) & same("background n low", a1l, a2l, P.FE, tag) &
              same("background n+1", b1, b2, P.FE, tag) & same("background n+1 low", b1l, b2l, P.FE, tag) &
              same("Born n", ea1, ea2, P.FE, tag) & same("Born n low", ea1l, ea2l, P.FE, tag) &
              same("Born n+1", eb1, eb2, P.FE, tag) & same("Born n+1 low", eb1l, eb2l, P.FE, tag) &
              same("Born traces n", ra1, ra2, nr, tag) & same("Born traces n+1", rb1, rb2, nr, tag);
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

template <int LB, int CX>
bool born_x2_cases(uint32_t seed) {
    bool ok = true;
    const int shapes[][2] = {{29, 37}, {9, 13}, {40, 33}, {3, 5}};
    for (auto &s : shapes)
        for (int RY : {4, 8, 64}) ok &= check_born_x2<LB, CX>(s[0], s[1], 2, RY, seed++);
    ok &= check_born_x2<LB, CX>(29, 37, 1, 4, seed++);
    return ok;
}

// GSOT assignment (k_gsot): per trace the returned assignment is a
// permutation and is optimal, certified by the dual potentials the kernel
// leaves (reduced costs C - u - v >= 0 everywhere, 0 on the assignment); the
// trace's misfit and adjoint source equal a direct evaluation; for nt <= 7 the
// misfit equals the brute-force minimum over all permutations.
bool check_gsot(int n_tx, int nt, int n_rx, int pattern, uint32_t seed) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> g(0.0, 1.0);
    const int ntr = n_tx * n_rx;
    const int64_t NE = (int64_t)n_tx * nt * n_rx;
    double *ds = alloc<double>(NE, 0), *dob = alloc<double>(NE, 0);
    for (int64_t i = 0; i < NE; ++i) {
        ds[i] = g(rng);
        dob[i] = 0.8 * ds[i] + 0.5 * g(rng);
        if (pattern == 1 && (i / n_rx) % nt < nt / 3) ds[i] = dob[i] = 0.0;  // quiet onsets
    }
    if (pattern == 2)  // trace 0: observed all zero; trace 1: both zero
        for (int k = 0; k < nt; ++k) {
            dob[(int64_t)k * n_rx] = 0.0;
            if (n_rx > 1) ds[(int64_t)k * n_rx + 1] = dob[(int64_t)k * n_rx + 1] = 0.0;
        }
    const int64_t W = nt + 1;
    double *wsd = alloc<double>(5 * W * ntr, 0), *Jt = alloc<double>(ntr, 0), *adj = alloc<double>(NE, 0);
    int *wsi = alloc<int>(3 * W * ntr, 0), *status = alloc<int>(1, 0);
    emu::launch((unsigned)ntr, 256, [&] {
        k_gsot(ds, dob, nt, n_rx, ntr, 1.0, wsd, wsi, Jt, adj, status);
    });
    char tag[96];
    std::snprintf(tag, sizeof tag, "gsot n_tx=%d nt=%d n_rx=%d pattern=%d", n_tx, nt, n_rx, pattern);
    bool ok = *status == 0;
    if (!ok) std::printf("MISMATCH %s status %d\n", tag, *status);
    for (int tr = 0; tr < ntr && ok; ++tr) {
        const int t = tr / n_rx, r = tr % n_rx;
        const double *u = wsd + (int64_t)tr * 5 * W, *v = u + W, *sh = v + 2 * W, *oh = sh + W;
        const int *pp = wsi + (int64_t)tr * 3 * W;
        double A = 0.0;
        for (int k = 0; k < nt; ++k) A = std::max(A, std::fabs(dob[((int64_t)t * nt + k) * n_rx + r]));
        A += 1e-30;
        std::vector<int> col(nt + 1, 0), seen(nt + 1, 0);
        for (int j = 1; j <= nt; ++j) {
            if (pp[j] < 1 || pp[j] > nt || seen[pp[j]]++) { ok = false; break; }
            col[pp[j]] = j;
        }
        if (!ok) { std::printf("MISMATCH %s trace %d: not a permutation\n", tag, tr); break; }
        double cmax = 0.0, jd = 0.0;
        for (int i = 1; i <= nt; ++i)
            for (int j = 1; j <= nt; ++j) cmax = std::max(cmax, gsot_cost(sh, oh, i, j, nt, 1.0));
        const double tol = 1e-9 * (cmax + 1.0);
        for (int i = 1; i <= nt && ok; ++i) {
            for (int j = 1; j <= nt; ++j) {
                const double rc = gsot_cost(sh, oh, i, j, nt, 1.0) - u[i] - v[j];
                if (rc < -tol || (j == col[i] && std::fabs(rc) > tol)) {
                    std::printf("MISMATCH %s trace %d: reduced cost %g at (%d,%d)\n", tag, tr, rc, i, j);
                    ok = false;
                    break;
                }
            }
            const double c = gsot_cost(sh, oh, i, col[i], nt, 1.0);
            jd += c;
            const double s_i = ds[((int64_t)t * nt + i - 1) * n_rx + r] / A;
            const double o_c = dob[((int64_t)t * nt + col[i] - 1) * n_rx + r] / A;
            const double want = 2.0 * (s_i - o_c) / A;
            if (std::memcmp(&want, &adj[((int64_t)t * nt + i - 1) * n_rx + r], 8) != 0) {
                std::printf("MISMATCH %s trace %d: adjoint source at %d\n", tag, tr, i);
                ok = false;
            }
        }
        if (ok && std::fabs(jd - Jt[tr]) > 1e-12 * (std::fabs(jd) + 1e-300)) {
            std::printf("MISMATCH %s trace %d: J %g vs %g\n", tag, tr, Jt[tr], jd);
            ok = false;
        }
        if (ok && nt <= 7) {  // brute force
            std::vector<int> perm(nt);
            for (int k = 0; k < nt; ++k) perm[k] = k + 1;
            double best = INFINITY;
            do {
                double c = 0.0;
                for (int k = 0; k < nt; ++k) c += gsot_cost(sh, oh, k + 1, perm[k], nt, 1.0);
                best = std::min(best, c);
            } while (std::next_permutation(perm.begin(), perm.end()));
            if (std::fabs(best - jd) > 1e-12 * (std::fabs(best) + 1e-300)) {
                std::printf("MISMATCH %s trace %d: brute force %g vs %g\n", tag, tr, best, jd);
                ok = false;
            }
        }
    }
    std::printf("%s %s\n", ok ? "ok" : "FAIL", tag);
    return ok;
}

// A trace with a NaN sample ends with the status flag set and a NaN misfit;
// the other traces are unaffected.
bool check_gsot_nan() {
    const int nt = 9, n_rx = 2;
    double *ds = alloc<double>(nt * n_rx, 0), *dob = alloc<double>(nt * n_rx, 0);
    for (int i = 0; i < nt * n_rx; ++i) {
        ds[i] = std::sin(0.7 * i);
        dob[i] = std::cos(0.3 * i);
    }
    ds[4 * n_rx + 1] = NAN;
    const int64_t W = nt + 1;
    double *wsd = alloc<double>(5 * W * n_rx, 0), *Jt = alloc<double>(n_rx, 0), *adj = alloc<double>(nt * n_rx, 0);
    int *wsi = alloc<int>(3 * W * n_rx, 0), *status = alloc<int>(1, 0);
    emu::launch((unsigned)n_rx, 256, [&] { k_gsot(ds, dob, nt, n_rx, n_rx, 1.0, wsd, wsi, Jt, adj, status); });
    const bool ok = *status == 1 && std::isnan(Jt[1]) && std::isfinite(Jt[0]) && adj[4 * n_rx + 1] == 0.0;
    std::printf("%s gsot NaN trace flagged\n", ok ? "ok" : "FAIL");
    return ok;
}

bool gsot_cases(uint32_t seed) {
    bool ok = check_gsot_nan();
    for (int nt : {1, 2, 5, 7})
        for (int pattern : {0, 1}) ok &= check_gsot(2, nt, 3, pattern, seed++);
    ok &= check_gsot(1, 7, 2, 2, seed++);
    ok &= check_gsot(2, 64, 3, 0, seed++);
    ok &= check_gsot(1, 300, 2, 1, seed++);
    ok &= check_gsot(1, 257, 2, 2, seed++);
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
    if (only.empty() || only == "xn") {
        ok &= xn_cases<64, 4, 3>(seed + 11500, true);
        ok &= xn_cases<64, 4, 2>(seed + 11000, false);
        ok &= xn_cases<16, 4, 3>(seed + 12000, false);
        ok &= xn_cases<32, 4, 3>(seed + 12500, false);
        ok &= xn_cases<16, 8, 3>(seed + 13000, false);
        ok &= xn_cases<64, 4, 4>(seed + 13500, false);
    }
    if (only.empty() || only == "lds") {
        ok &= lds_cases(seed + 2000);
    }
    if (only.empty() || only == "blk3d") {
        // the 3D patch kernels with strip blocks other than the default pairs
        for (int blk : {(2 << 8) | 2, (4 << 8) | 4, (8 << 8) | 8, (16 << 8) | 8, (4 << 8) | 16}) {
            g_blk = blk;
            ok &= fwd3d_cases<64, 2, 2>(seed + 9700 + blk);
            ok &= fwd3d_cases<16, 4, 2>(seed + 9800 + blk);
            ok &= radj3d_cases<64, 2, 2>(seed + 9900 + blk);
            ok &= radj3d_cases<32, 1, 1>(seed + 10000 + blk);
            ok &= born3d_cases<64, 2, 1>(seed + 10100 + blk);
            ok &= check_fwd3d_x2p<64, 1, 1>(9, 7, 11, 2, 4, seed + 10200 + blk);
            ok &= check_fwd3d_x2p<16, 2, 1>(11, 4, 9, 2, 16, seed + 10300 + blk);
        }
        g_blk = kBlk3dDefault;
    }
    if (only.empty() || only == "lo32") {
        ok &= lo32_cases(seed + 9500);
    }
    if (only.empty() || only == "x2p-3d") {
        ok &= x2p_3d_cases(seed + 9000);
    }
    if (only.empty() || only == "x2-3d") {
        ok &= x2_3d_cases(seed + 4000);
    }
    if (only.empty() || only == "born") {
        ok &= born_cases(seed + 3000);
    }
    if (only.empty() || only == "gsot") {
        ok &= gsot_cases(seed + 6000);
    }
    if (only.empty() || only == "born-x2") {
        ok &= born_x2_cases<16, 4>(seed + 5000);
        ok &= born_x2_cases<32, 4>(seed + 5100);
        ok &= born_x2_cases<64, 4>(seed + 5200);
        ok &= born_x2_cases<16, 8>(seed + 5300);
    }
    if (only.empty() || only == "fwd3d") {
        ok &= fwd3d_cases<64, 4, 2>(seed + 1000);
        ok &= fwd3d_cases<64, 4, 4>(seed + 1100);
        ok &= fwd3d_cases<32, 8, 2>(seed + 1200);
        ok &= fwd3d_cases<16, 4, 2>(seed + 1300);
        ok &= fwd3d_cases<16, 2, 2>(seed + 1400);
        ok &= fwd3d_cases<16, 2, 4>(seed + 1500);
        ok &= fwd3d_cases<32, 2, 4>(seed + 1600);
        ok &= fwd3d_cases<16, 2, 1>(seed + 1700);
        ok &= fwd3d_cases<16, 1, 2>(seed + 1800);
        ok &= fwd3d_cases<16, 1, 1>(seed + 1900);
    }
    if (only.empty() || only == "born3d") {
        ok &= born3d_cases<16, 1, 1>(seed + 8000);
        ok &= born3d_cases<16, 2, 1>(seed + 8100);
        ok &= born3d_cases<16, 1, 2>(seed + 8200);
        ok &= born3d_cases<64, 1, 1>(seed + 8300);
        ok &= born3d_cases<32, 2, 2>(seed + 8400);
    }
    if (only.empty() || only == "fwd3d_ysh") {
        uint32_t sd = seed + 8300;
        for (auto c : {std::array<int, 5>{9, 7, 11, 2, 4}, {5, 6, 13, 1, 64}, {12, 9, 5, 2, 8},
                       {3, 3, 3, 1, 2}, {6, 21, 7, 1, 3}, {7, 18, 9, 2, 5}}) {
            ok &= check_fwd3d<64, 2, 2, 2>(c[0], c[1], c[2], c[3], c[4], sd++);
            ok &= check_fwd3d<64, 2, 1, 2>(c[0], c[1], c[2], c[3], c[4], sd++);
        }
        ok &= check_fwd3d<64, 2, 2, 2, 1, 1, float>(8, 17, 6, 1, 4, sd++);
    }
    if (only.empty() || only == "radj3d_ysh") {
        ok &= radj3d_cases<64, 2, 2, true>(seed + 7700);
        ok &= radj3d_cases<64, 2, 1, true>(seed + 7800);
        ok &= radj3d_cases<32, 2, 2, true>(seed + 7900);
        ok &= radj3d_cases<32, 2, 1, true>(seed + 8000);
        ok &= radj3d_cases<16, 2, 2, true>(seed + 8100);
        ok &= radj3d_cases<16, 2, 1, true>(seed + 8200);
    }
    if (only.empty() || only == "radj3d") {
        ok &= radj3d_cases<16, 2, 2>(seed + 7000);
        ok &= radj3d_cases<16, 2, 1>(seed + 7100);
        ok &= radj3d_cases<16, 1, 2>(seed + 7200);
        ok &= radj3d_cases<16, 1, 1>(seed + 7300);
        ok &= radj3d_cases<32, 2, 2>(seed + 7400);
        ok &= radj3d_cases<64, 2, 2>(seed + 7500);
        ok &= radj3d_cases<64, 1, 1>(seed + 7600);
    }
    std::printf(ok ? "ALL EQUAL\n" : "DIFFERENCES FOUND\n");
    return ok ? 0 : 1;
}

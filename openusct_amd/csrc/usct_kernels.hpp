// usct_kernels.hpp -- CDNA4 (gfx950) kernels of the acoustic leapfrog solver and
// its exact discrete adjoint.
//
// Two field layouts:
//   plain    p[cell]                 one shot, one thread per cell
//   batched  p[cell * Bp + shot]     all shots of a pass, shot index fastest.
//            A wavefront's lanes hold LB consecutive shots of 64/LB adjacent
//            cell strips, so every load/store is a contiguous run of LB floats,
//            and per-cell data (S', sponge, receiver map) is shared by the LB
//            lanes.  Each lane marches a strip of CX cells (2D: along y,
//            3D: a CX x CY patch along z) and keeps the three stencil rows or
//            planes in registers; only the x (and y) halo is re-read.
//
// Arithmetic (per cell, FP32 fields, identical in both layouts):
//   lap_u   = interior-only 5/7-point sum, centre term first, then axes
//             0..ndim-1, "+1" before "-1" (simulation/ringfwi/solver.py:45-62),
//             without the 1/h^2 factor;
//   S'      = dt^2 / (m h^2) computed in FP64 and carried as an FP32 hi/lo pair,
//             so the FP32 update sees S' to ~2^-48 (the FP32 rounding of S is
//             a systematic velocity bias otherwise);
//   forward p^n = (2 p^{n-1} - p^{n-2}) + S' lap_u(p^{n-1}) + S_src w[n-1]
//             (fwi.py:63-67, lap[src] += w before the S product);
//   adjoint lam^k = (2 lam^{k+1} - lam^{k+2}) + lap_u(S' lam^{k+1}) - R^T r^k
//             (fwi.py:310-322); sponge multiplies p^n and lam^k (fwi.py:68,313);
//   imaging d2p^k = p^k - 2p^{k-1} + p^{k-2}; without a sponge the recursion
//             makes this exactly S' lap_u(p^{k-1}) + S_src w[k-1], which reads one
//             stored level instead of three and avoids a cancellation
//             (fwi.py:325-326);  g += lam^k d2p^k with Neumaier compensation.
// The forward state is stored as FP32 hi + bfloat16 lo.  2D updates it with
// FP32 TwoSum arithmetic; 3D forms the whole update in FP64 (see k_fwd3d).
// Adjoint fields are FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace usct {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // 4 wavefronts

// Minimum wavefronts per SIMD the compiler must leave room for (caps the
// registers of a kernel family); 0 = the compiler's choice.  The 2D kernels
// run 3 per SIMD (measured on the 512^2, 128-shot gradient: 3% faster than the
// compiler's choice; 4 spills and is 1.8x slower); the 3D kernels gain nothing.
#ifndef USCT_OCC2D
#define USCT_OCC2D 3
#endif
#ifndef USCT_OCC3D
#define USCT_OCC3D 0
#endif
#if USCT_OCC2D > 0
#define USCT_ATTR2D __attribute__((amdgpu_waves_per_eu(USCT_OCC2D)))
#else
#define USCT_ATTR2D
#endif
#if USCT_OCC3D > 0
#define USCT_ATTR3D __attribute__((amdgpu_waves_per_eu(USCT_OCC3D)))
#else
#define USCT_ATTR3D
#endif
// The two-level 3D forward needs 256+ registers; 2 per SIMD (a 28-byte spill)
// runs as fast as 1, 3 spills heavily and runs 2.3x slower.
#ifndef USCT_OCC3DX2
#define USCT_OCC3DX2 2
#endif
#define USCT_ATTR3DX2 __attribute__((amdgpu_waves_per_eu(USCT_OCC3DX2)))
// Wavefronts per SIMD of the 2D kernels that advance LV levels per launch.
#ifndef USCT_OCCXN
#define USCT_OCCXN(LV) ((LV) <= 2 ? 3 : 2)
#endif

// Blocks are dispatched round-robin over the 8 XCDs (block i -> XCD i % 8).
// Give each XCD a contiguous range of logical blocks so neighbouring strips,
// which share halo rows, share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
    const int per = nblk >> 3, rem = nblk & 7;
    const int xcd = bid & 7, slot = bid >> 3;
    return xcd * per + (xcd < rem ? xcd : rem) + slot;
}

// Strip order of the 3D patch kernels (k_fwd3d, k_adj3d, k_radj3d, k_born3d,
// k_fwd3d_x2): consecutive wavefronts take the strips of a block of bx x by
// strips (y fastest), blocks follow along x, then along y, then z chunks, so
// strips whose halo cells neighbour each other run at about the same time on
// the same XCD and the halo rows and columns they share are fetched from L2.
// blk = (bx << 8) | by, both in [1, 255]; bx = 1, by = 2 pairs two y strips of
// one x strip.  The grid covers the strips rounded up to whole blocks.
constexpr int kBlk3dDefault = (1 << 8) | 2;
__host__ __device__ inline int64_t tasks3d(int nxs, int nys, int nzc, int blk) {
    const int by = blk & 255, bx = blk >> 8;
    return (int64_t)((nxs + bx - 1) / bx * bx) * ((nys + by - 1) / by * by) * nzc;
}
__device__ __forceinline__ bool strip3d(int t, int nxs, int nys, int nzc, int blk, int &xs,
                                        int &ys, int &zc) {
    const int by = blk & 255, bx = blk >> 8;
    const int q = t % (bx * by);
    t /= bx * by;
    const int nbx = (nxs + bx - 1) / bx, nby = (nys + by - 1) / by;
    xs = (t % nbx) * bx + q / by;
    t /= nbx;
    ys = (t % nby) * by + q % by;
    zc = t / nby;
    return xs < nxs && ys < nys && zc < nzc;
}

// Global wavefront index (XCD-remapped block, then wave within the block),
// marked wave-uniform so everything derived from it lives in scalar registers.
__device__ __forceinline__ int wave_id() {
    const int w = (int)(threadIdx.x >> 6);
    return xcd_remap(blockIdx.x, gridDim.x) * (kBlock / kWave) + __builtin_amdgcn_readfirstlane(w);
}

__device__ __forceinline__ float lap5(float c, float yp, float ym, float xp, float xm) {
    float a = -4.0f * c;
    a = a + yp;
    a = a + ym;
    a = a + xp;
    a = a + xm;
    return a;
}

__device__ __forceinline__ float lap7(float c, float zp, float zm, float yp, float ym,
                                      float xp, float xm) {
    float a = -6.0f * c;
    a = a + zp;
    a = a + zm;
    a = a + yp;
    a = a + ym;
    a = a + xp;
    a = a + xm;
    return a;
}

// S' * v with S' = hi + lo.
__device__ __forceinline__ float smul(float2 s, float v) { return fmaf(s.x, v, s.y * v); }

__device__ __forceinline__ float leap(float cur, float prev, float inc) {
    return (2.0f * cur - prev) + inc;
}

// Forward state p = hi + lo with lo kept in bfloat16 (round to nearest even):
// the FP32 rounding of p at every step, integrated over thousands of steps, is
// what separates FP32 traces from FP64 ones; carrying it in 16 extra bits
// removes most of that drift (the update itself is FP64, see lap5d below).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ void neumaier(float &s, float &c, float x) {
    const float t = s + x;
    if (fabsf(s) >= fabsf(x)) c = c + ((s - t) + x);
    else c = c + ((x - t) + s);
    s = t;
}

// The value of lane (lane ^ OFF) of the wavefront.  On the device: quad
// permutations of DPP for OFF = 1, 2 (no LDS trip), the bit-mask swizzle for
// OFF = 4, 8, 16 (within each half of the wavefront) and a permute through the
// LDS crossbar for OFF = 32; none needs per-call index arithmetic.
template <int OFF>
__device__ __forceinline__ float lane_xor(float v, int lane) {
    static_assert(OFF >= 1 && OFF <= 32 && (OFF & (OFF - 1)) == 0, "pow2 lane offset");
#if defined(__HIP_DEVICE_COMPILE__)
    const int x = __float_as_int(v);
    if constexpr (OFF == 1) return __int_as_float(__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true));
    else if constexpr (OFF == 2) return __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true));
    else if constexpr (OFF < 32) return __int_as_float(__builtin_amdgcn_ds_swizzle(x, (OFF << 10) | 0x1F));
    else return __int_as_float(__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, x));
#else
    (void)lane;
    return __shfl_xor(v, OFF, kWave);
#endif
}

// Reduce CX per-lane values over each aligned group of LB lanes and scatter the
// CX sums: afterwards lane l holds the sum for cell index sum_cell<LB,CX>(l)
// (recursive halving: CX/2 + CX/4 + ... + log2(LB/CX) shuffles instead of
// CX*log2(LB)).  Order is fixed, so results do not depend on scheduling.  The
// two candidates of each exchange are read as values before the lane picks one
// (a pick between array elements would index the array at run time).
template <int LB, int CX, int OFF = LB / 2>
__device__ __forceinline__ void halve_stages(float (&v)[CX], int lane) {
    if constexpr (OFF >= LB / CX) {
        constexpr int half = CX * OFF / LB;  // CX/2 at OFF = LB/2, then halving
        const bool up = (lane & OFF) != 0;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const float lo_v = v[i], hi_v = v[i + half];
            const float send = up ? lo_v : hi_v;
            const float keep = up ? hi_v : lo_v;
            v[i] = keep + lane_xor<OFF>(send, lane);
        }
        halve_stages<LB, CX, OFF / 2>(v, lane);
    }
}
template <int LB, int OFF>
__device__ __forceinline__ float sum_stages(float s, int lane) {
    if constexpr (OFF >= 1) return sum_stages<LB, OFF / 2>(s + lane_xor<OFF>(s, lane), lane);
    else return s;
}
template <int LB, int CX>
__device__ __forceinline__ float reduce_scatter(float (&v)[CX], int lane) {
    static_assert((LB & (LB - 1)) == 0 && (CX & (CX - 1)) == 0 && CX <= LB, "pow2");
    if constexpr (CX > 1) halve_stages<LB, CX>(v, lane);
    return sum_stages<LB, LB / (2 * CX)>(v[0], lane);
}

template <int LB, int CX>
__device__ __forceinline__ int sum_cell(int lane) {
    int ci = 0;
#pragma unroll
    for (int half = CX / 2, off = LB / 2; half >= 1; half >>= 1, off >>= 1)
        if (lane & off) ci += half;
    return ci;
}

// Element at a byte offset from a wave-uniform base: lets the compiler use the
// "scalar base + 32-bit vector offset" form of global loads and stores.
template <class T>
__device__ __forceinline__ T ldb(const T *base, unsigned boff) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}
template <class T>
__device__ __forceinline__ void stb(T *base, unsigned boff, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff) = v;
}
// The same for operands read or written exactly once per launch (no stencil
// reuse): non-temporal loads (NT bit 1) and stores (NT bit 2) bypass the L2's
// normal retention, leaving it to the rows other wavefronts re-read.  Measured
// on the 512^2, 128-shot gradient the 2D two-level kernels gain 9 % with both;
// the 3D adjoint that reads stored levels (k_adj3d, 16-lane tiles) loses 6 % with
// both (the non-temporal loads give up the scalar-base addressing form).
#ifndef USCT_NT3D
#define USCT_NT3D 0
#endif
#ifndef USCT_NT2D
#define USCT_NT2D 3
#endif
// The reversing 3D adjoint (k_radj3d) gains with both: on the 256^3, 64-shot
// gradient 6.80 ms per level against 7.42 (loads only 7.33, stores only 7.04).
#ifndef USCT_NTR3D
#define USCT_NTR3D 3
#endif
// The 3D forward kernels (k_fwd3d, k_born3d) on the same gradient, whose tuned
// tiles are 64-lane groups: 4.00 ms per level with both against 4.31.
#ifndef USCT_NTF3D
#define USCT_NTF3D 3
#endif
template <int NT, class T>
__device__ __forceinline__ T ldn(const T *base, unsigned boff) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr ((NT & 1) != 0)
        return __builtin_nontemporal_load(
            reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff));
#endif
    return ldb(base, boff);
}
template <class T>
__device__ __forceinline__ T ld_nt(const T *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
template <class T>
__device__ __forceinline__ void st_nt(T *p, T v) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
template <int NT, class T>
__device__ __forceinline__ void stn(T *base, unsigned boff, T v) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr ((NT & 2) != 0) {
        __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff));
        return;
    }
#endif
    stb(base, boff, v);
}

// ---------------------------------------------------------------------------
// Forward update in FP64 (2D and 3D).  The state is stored as FP32 hi +
// bfloat16 lo (6 bytes); the update reads the full P = hi + lo of every
// stencil point, forms lap_u(P), S' lap_u(P), the source term S_src w[n-1]
// and 2P^{n-1} - P^{n-2} + increment in FP64, and splits the result back into
// hi + lo.  The gradient sums lam^k d2p^k over many levels with strong
// cancellation: an increment rounded to FP32 at every step (or a Laplacian of
// the high parts only, or an FP32 source term) leaves a wavefield error of a
// few 1e-7 that becomes a gradient error of 1e-5 .. 1e-4 on coarse or
// high-contrast grids; with the FP64 update the gradient is within ~1e-6 of
// the FP64 reference.  CDNA FP64 vector arithmetic runs at the rate of
// (unpacked) FP32, and the storage, hence the HBM traffic, is unchanged.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lap5d(double c, double yp, double ym, double xp, double xm) {
    double a = -4.0 * c;
    a = a + yp;
    a = a + ym;
    a = a + xp;
    a = a + xm;
    return a;
}
__device__ __forceinline__ double lap7d(double c, double zp, double zm, double yp, double ym,
                                        double xp, double xm) {
    double a = -6.0 * c;
    a = a + zp;
    a = a + zm;
    a = a + yp;
    a = a + ym;
    a = a + xp;
    a = a + xm;
    return a;
}
__device__ __forceinline__ double hl(float h, uint16_t l) { return (double)h + (double)bf2f(l); }
__device__ __forceinline__ double s2d(float2 s) { return (double)s.x + (double)s.y; }
// Split an FP64 value into the stored FP32 hi and bfloat16 lo.
__device__ __forceinline__ float split_hl(double v, uint16_t &lo) {
    const float h = (float)v;
    lo = f2bf((float)(v - (double)h));
    return h;
}
// The same with an FP32 low part (plans created with USCT_FLAG_FP32_LO): the
// state then carries about 48 significant bits instead of 32.  The kernels
// that read or write the forward state take the low part's type as a
// template parameter L (uint16_t for bfloat16, float).
__device__ __forceinline__ double hl(float h, float l) { return (double)h + (double)l; }
__device__ __forceinline__ float split_hl(double v, float &lo) {
    const float h = (float)v;
    lo = (float)(v - (double)h);
    return h;
}
__device__ __forceinline__ float lo_f(uint16_t b) { return bf2f(b); }
__device__ __forceinline__ float lo_f(float f) { return f; }

// ---------------------------------------------------------------------------
// Batched 2D forward step: level n from levels n-1 (pc) and n-2 (pp) into pn.
// sincd[b] = S_src(b) * w[n-1] in FP64; rec: [n_rec][Bp] row of level n;
// row_rec[y] = receivers on grid row y.  Grid: ng * nxs * nyc wavefronts, 4
// per block.  Sponge: the new level is multiplied by the mask (fwi.py:68).
//
// Addressing: each lane keeps one 32-bit byte offset, that of its strip's first
// cell and shot; row starts (clamped to the grid, wave-uniform) and the column
// steps +-i*Bp go into the scalar base address.  Columns are not clamped: the
// x halo of the first and last column reads the neighbouring row (or the
// guard zone every field carries before and after it), values that only feed
// border cells, whose Laplacian is zero.
// ---------------------------------------------------------------------------
template <int LB, int CX, bool SPONGE, bool REC, class L = uint16_t>
__global__ void __launch_bounds__(kBlock) USCT_ATTR2D
k_fwd2d(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pn,
        const L *__restrict__ pcl, const L *__restrict__ ppl,
        L *__restrict__ pnl, const float2 *__restrict__ sp2,
        const float *__restrict__ sponge, const int *__restrict__ src_cell,
        const double *__restrict__ sincd, const int *__restrict__ rec_head,
        const int *__restrict__ rec_next, const int *__restrict__ row_rec,
        float *__restrict__ rec, float *__restrict__ rec_lo, int ny, int nx, int Bp, int RY,
        int nxs, int nyc, int ng) {
    constexpr int W = kWave / LB;
    const int wv = wave_id();
    const int lane = threadIdx.x & 63;
    const int sg = wv % ng;
    const int t = wv / ng;
    const int xs = t % nxs, yc = t / nxs;
    if (yc >= nyc) return;
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    if (x0 >= nx) return;
    const int y0 = yc * RY;
    const int y1 = min(y0 + RY, ny);
    const int my_src = src_cell[b];
    const double my_inc = sincd[b];
    const size_t rowE = (size_t)nx * Bp;
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;  // lane byte offset, float fields
    const unsigned o2 = (unsigned)(x0 * Bp + b) * (unsigned)sizeof(L);  // bf16 fields
    const unsigned oc = (unsigned)x0 * 8u;             // S' (float2 per cell)
    auto P = [&](size_t e) -> double { return hl(ldb(pc + e, o4), ldb(pcl + e, o2)); };
    double rm[CX], rc[CX + 2], rp[CX];
    {
        const size_t r0 = (size_t)max(y0 - 1, 0) * rowE, r1 = (size_t)y0 * rowE;
#pragma unroll
        for (int i = 0; i < CX; ++i) rm[i] = P(r0 + (ptrdiff_t)i * Bp);
#pragma unroll
        for (int i = 0; i < CX + 2; ++i) rc[i] = P(r1 + (ptrdiff_t)(i - 1) * Bp);
    }
    for (int y = y0; y < y1; ++y) {
        const size_t rn = (size_t)min(y + 1, ny - 1) * rowE;
#pragma unroll
        for (int i = 0; i < CX; ++i) rp[i] = P(rn + (ptrdiff_t)i * Bp);
        const bool yint = (y >= 1) && (y <= ny - 2);
        const size_t ro = (size_t)y * rowE;
        const float2 *spr = sp2 + (size_t)y * nx;
        const bool has_rec = REC && row_rec[y] > 0;
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            const int x = x0 + i;
            if (x < nx) {
                double lu = 0.0;
                if (yint && x >= 1 && x <= nx - 2) lu = lap5d(rc[i + 1], rp[i], rm[i], rc[i + 2], rc[i]);
                double inc = s2d(ldb(spr + i, oc)) * lu;
                if (y * nx + x == my_src) inc = inc + my_inc;
                const size_t ei = ro + (size_t)i * Bp;
                double nv = (2.0 * rc[i + 1] - hl(ldb(pp + ei, o4), ldb(ppl + ei, o2))) + inc;
                if (SPONGE) nv = nv * (double)sponge[(size_t)y * nx + x];
                L lb;
                const float v = split_hl(nv, lb);
                stb(pn + ei, o4, v);
                stb(pnl + ei, o2, lb);
                if (has_rec) {
                    for (int j = rec_head[y * nx + x]; j >= 0; j = rec_next[j]) {
                        rec[(size_t)j * Bp + b] = v;
                        rec_lo[(size_t)j * Bp + b] = lo_f(lb);
                    }
                }
            }
        }
        if (y + 1 < y1) {
#pragma unroll
            for (int i = 0; i < CX; ++i) { rm[i] = rc[i + 1]; rc[i + 1] = rp[i]; }
            rc[0] = P(rn - (ptrdiff_t)Bp);
            rc[CX + 1] = P(rn + (ptrdiff_t)CX * Bp);
        }
    }
}

// ---------------------------------------------------------------------------
// Batched 2D forward, two levels per launch (temporal blocking, no sponge):
// level n (A) from levels n-1 (pc) and n-2 (pp), then level n+1 (B) from A and
// pc, both in FP64 as k_fwd2d.  A wavefront marches its rows with B one row
// behind A, so the three A rows around B's row sit in registers (as the values
// A's stored hi + lo represent, which is what a separate launch would read).
// Each lane computes A on its CX cells plus one halo column on either side,
// and on one halo row above and below its row range: the same operations on
// the same operands as the owner of those cells, hence the same values.  Both
// levels and their traces are bit-identical to two single-level launches; per
// shot-cell the pair reads 12 bytes and writes 12 instead of reading 24 and
// writing 12.  The next row's loads are issued one iteration ahead.
//   sincd_a[b] = S_src w[n-1], sincd_b[b] = S_src w[n] (FP64);
//   rec_a / rec_b: trace rows of levels n and n+1.
// Halo columns read one or two cells beyond the strip; the guard zones of the
// operands cover the reads before the first and after the last row.
//
// Row chunks and their order (sched, wave-uniform):
//   bit 0  odd row chunks march upwards (from their last row to their first),
//          so that two neighbouring chunks reach the rows they both read (the
//          halo rows of their common boundary) at the same time, while those
//          rows are still in L2: downward chunks meet at their ends, upward
//          ones at their starts;
//   bit 1  consecutive wavefronts take consecutive row chunks of one strip
//          (chunk index fastest) instead of consecutive strips of one chunk,
//          so that vertically neighbouring chunks run on the same XCD;
//   bit 2  shot groups slowest instead of fastest: the wavefronts of a
//          work-group take neighbouring strips (or chunks) of one shot group
//          and read each other's halo columns on one CU.
// Only the order in which rows are visited changes: every value is formed
// from the same operands in the same order, so the results do not depend on
// sched.  Rows are named relative to the front row yf and the march
// direction dir: row(r) = yf + r * dir; "ahead" rows are r > 0.
// ---------------------------------------------------------------------------
// Row index clamped to the grid (halo rows of a wavefront read the edge row).
__device__ __forceinline__ size_t clamp_row(int y, int ny) { return (size_t)min(max(y, 0), ny - 1); }

// Shot group, strip and row chunk of 2D wavefront wv, with the march direction;
// false for the wavefronts that round the grid up to whole blocks.  Shot groups
// are fastest (the wavefronts of a work-group share strips) unless sched bit 2
// is set (a work-group holds neighbouring strips of one shot group).
__device__ __forceinline__ bool chunk2d(int wv, int ng, int nxs, int nyc, int sched, int &sg,
                                        int &xs, int &yc, int &dir) {
    int t;
    if (sched & 4) {
        sg = wv / (nxs * nyc);
        t = wv % (nxs * nyc);
    } else {
        sg = wv % ng;
        t = wv / ng;
    }
    if (sched & 2) {
        yc = t % nyc;
        xs = t / nyc;
    } else {
        xs = t % nxs;
        yc = t / nxs;
    }
    dir = ((sched & 1) && (yc & 1)) ? -1 : 1;
    return sg < ng && xs < nxs && yc < nyc;
}

// The march direction is a template parameter of the body (the kernel picks
// the instance per wavefront), so row selections by direction cost nothing.
template <int LB, int CX, bool REC, class L, int DIR>
__device__ __forceinline__ void fwd2d_x2_body(
    const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pa,
    float *__restrict__ pb, const L *__restrict__ pcl, const L *__restrict__ ppl,
    L *__restrict__ pal, L *__restrict__ pbl, const float2 *__restrict__ sp2,
    const int *__restrict__ src_cell, const double *__restrict__ sincd_a,
    const double *__restrict__ sincd_b, const int *__restrict__ rec_head,
    const int *__restrict__ rec_next, const int *__restrict__ row_rec, float *__restrict__ rec_a,
    float *__restrict__ rec_lo_a, float *__restrict__ rec_b, float *__restrict__ rec_lo_b, int ny,
    int nx, int Bp, int RY, int sg, int xs, int yc) {
    constexpr int W = kWave / LB;
    constexpr int CA = CX + 2;  // A columns x0-1 .. x0+CX
    constexpr int dir = DIR;
    const int lane = threadIdx.x & 63;
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    if (x0 >= nx) return;
    const int y0 = yc * RY;
    const int y1 = min(y0 + RY, ny);
    const int my_src = src_cell[b];
    const double inc_a = sincd_a[b], inc_b = sincd_b[b];
    const size_t rowE = (size_t)nx * Bp;
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned o2 = (unsigned)(x0 * Bp + b) * (unsigned)sizeof(L);
    const unsigned oc = (unsigned)x0 * 8u;
    // first front row: one halo row outside the chunk, on the side the march
    // starts from; rows y0-1 .. y1 are visited (RY + 2 fronts)
    const int yfs = dir > 0 ? y0 - 1 : y1;
    auto rowE_at = [&](int yf, int r) { return clamp_row(yf + r * dir, ny) * rowE; };
    // Full state (FP64) around the front row yf (index j <-> column x0-2+j for
    // uc/un, x0-1+j for um): um = pc row(-1), uc = row(0), un = row(+1).  un and
    // q (level n-2 on row(0), columns x0-1 ..) come from raw loads issued one
    // iteration ahead (nh/nl, qh/ql).
    double um[CA], uc[CA + 2], un[CA + 2];
    float nh[CA + 2], qh[CA];
    L nl[CA + 2], ql[CA];
    // A rows(-2), (-1), (0) (index j <-> column x0-1+j), as stored
    double am[CA], ac[CA], an[CA];
    {
        const size_t r0 = rowE_at(yfs, -1), r1 = rowE_at(yfs, 0), r2 = rowE_at(yfs, 1);
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            const ptrdiff_t c = (ptrdiff_t)(j - 1) * Bp;
            um[j] = hl(ldb(pc + r0 + c, o4), ldb(pcl + r0 + c, o2));
            qh[j] = ldn<USCT_NT2D>(pp + r1 + c, o4);
            ql[j] = ldn<USCT_NT2D>(ppl + r1 + c, o2);
        }
#pragma unroll
        for (int j = 0; j < CA + 2; ++j) {
            const ptrdiff_t c = (ptrdiff_t)(j - 2) * Bp;
            uc[j] = hl(ldb(pc + r1 + c, o4), ldb(pcl + r1 + c, o2));
            nh[j] = ldb(pc + r2 + c, o4);
            nl[j] = ldb(pcl + r2 + c, o2);
        }
#pragma unroll
        for (int j = 0; j < CA; ++j) { am[j] = 0.0; ac[j] = 0.0; an[j] = 0.0; }
    }
    const int n_front = y1 - y0 + 2;
    for (int it = 0, yf = yfs; it < n_front; ++it, yf += dir) {
        const int yr = (int)clamp_row(yf, ny);  // row actually read for halo rows
#pragma unroll
        for (int j = 0; j < CA + 2; ++j) un[j] = hl(nh[j], nl[j]);
        double q[CA];
#pragma unroll
        for (int j = 0; j < CA; ++j) q[j] = hl(qh[j], ql[j]);
        // ---- loads for the next iteration: pc row(+2), pp row(+1)
        {
            const size_t rnn = rowE_at(yf, 2);
            const size_t en = rowE_at(yf, 1);
#pragma unroll
            for (int j = 0; j < CA + 2; ++j) {
                const ptrdiff_t c = (ptrdiff_t)(j - 2) * Bp;
                nh[j] = ldb(pc + rnn + c, o4);
                nl[j] = ldb(pcl + rnn + c, o2);
            }
#pragma unroll
            for (int j = 0; j < CA; ++j) {
                const ptrdiff_t c = (ptrdiff_t)(j - 1) * Bp;
                qh[j] = ldn<USCT_NT2D>(pp + en + c, o4);
                ql[j] = ldn<USCT_NT2D>(ppl + en + c, o2);
            }
        }
        // ---- level n on row yf, columns x0-1 .. x0+CX
        const bool own_a = yf >= y0 && yf < y1;
        const bool yint = yf >= 1 && yf <= ny - 2;
        const size_t ro = (size_t)yr * rowE;
        const float2 *spr = sp2 + (size_t)yr * nx;
        const bool rec_a_row = REC && own_a && row_rec[yr] > 0;
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            const int x = x0 - 1 + j;
            double lu = 0.0;
            if (yint && x >= 1 && x <= nx - 2) {
                // rows y+1 and y-1 of level n-1
                const double yp = dir > 0 ? un[j + 1] : um[j];
                const double ym = dir > 0 ? um[j] : un[j + 1];
                lu = lap5d(uc[j + 1], yp, ym, uc[j + 2], uc[j]);
            }
            double inc = s2d(ldb(spr + (j - 1), oc)) * lu;
            if (yf * nx + x == my_src) inc = inc + inc_a;
            const double nv = (2.0 * uc[j + 1] - q[j]) + inc;
            L lb;
            const float v = split_hl(nv, lb);
            am[j] = ac[j];
            ac[j] = an[j];
            an[j] = hl(v, lb);
            if (j >= 1 && j <= CX && own_a && x < nx) {
                const ptrdiff_t e = (ptrdiff_t)ro + (ptrdiff_t)(j - 1) * Bp;
                stn<USCT_NT2D>(pa + e, o4, v);
                stn<USCT_NT2D>(pal + e, o2, lb);
                if (rec_a_row) {
                    for (int r = rec_head[yr * nx + x]; r >= 0; r = rec_next[r]) {
                        rec_a[(size_t)r * Bp + b] = v;
                        rec_lo_a[(size_t)r * Bp + b] = lo_f(lb);
                    }
                }
            }
        }
        // ---- level n+1 on row y = row(-1), own columns
        const int y = yf - dir;
        if (y >= y0 && y < y1) {
            const bool yint_b = y >= 1 && y <= ny - 2;
            const size_t rob = (size_t)y * rowE;
            const float2 *spb = sp2 + (size_t)y * nx;
            const bool rec_b_row = REC && row_rec[y] > 0;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                if (x < nx) {
                    double lu = 0.0;
                    if (yint_b && x >= 1 && x <= nx - 2) {
                        const double yp = dir > 0 ? an[i + 1] : am[i + 1];
                        const double ym = dir > 0 ? am[i + 1] : an[i + 1];
                        lu = lap5d(ac[i + 1], yp, ym, ac[i + 2], ac[i]);
                    }
                    double inc = s2d(ldb(spb + i, oc)) * lu;
                    if (y * nx + x == my_src) inc = inc + inc_b;
                    const double nv = (2.0 * ac[i + 1] - um[i + 1]) + inc;
                    L lb;
                    const float v = split_hl(nv, lb);
                    const size_t e = rob + (size_t)i * Bp;
                    stn<USCT_NT2D>(pb + e, o4, v);
                    stn<USCT_NT2D>(pbl + e, o2, lb);
                    if (rec_b_row) {
                        for (int r = rec_head[y * nx + x]; r >= 0; r = rec_next[r]) {
                            rec_b[(size_t)r * Bp + b] = v;
                            rec_lo_b[(size_t)r * Bp + b] = lo_f(lb);
                        }
                    }
                }
            }
        }
        // ---- advance
#pragma unroll
        for (int j = 0; j < CA; ++j) um[j] = uc[j + 1];
#pragma unroll
        for (int j = 0; j < CA + 2; ++j) uc[j] = un[j];
    }
}

template <int LB, int CX, bool REC, class L = uint16_t>
__global__ void __launch_bounds__(kBlock) USCT_ATTR2D
k_fwd2d_x2(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pa,
           float *__restrict__ pb, const L *__restrict__ pcl,
           const L *__restrict__ ppl, L *__restrict__ pal,
           L *__restrict__ pbl, const float2 *__restrict__ sp2,
           const int *__restrict__ src_cell, const double *__restrict__ sincd_a,
           const double *__restrict__ sincd_b, const int *__restrict__ rec_head,
           const int *__restrict__ rec_next, const int *__restrict__ row_rec,
           float *__restrict__ rec_a, float *__restrict__ rec_lo_a, float *__restrict__ rec_b,
           float *__restrict__ rec_lo_b, int ny, int nx, int Bp, int RY, int nxs, int nyc,
           int ng, int sched) {
    int sg, xs, yc, dir;
    if (!chunk2d(wave_id(), ng, nxs, nyc, sched, sg, xs, yc, dir)) return;  // wave-uniform
    if (dir > 0)
        fwd2d_x2_body<LB, CX, REC, L, 1>(pc, pp, pa, pb, pcl, ppl, pal, pbl, sp2, src_cell,
                                         sincd_a, sincd_b, rec_head, rec_next, row_rec, rec_a,
                                         rec_lo_a, rec_b, rec_lo_b, ny, nx, Bp, RY, sg, xs, yc);
    else
        fwd2d_x2_body<LB, CX, REC, L, -1>(pc, pp, pa, pb, pcl, ppl, pal, pbl, sp2, src_cell,
                                          sincd_a, sincd_b, rec_head, rec_next, row_rec, rec_a,
                                          rec_lo_a, rec_b, rec_lo_b, ny, nx, Bp, RY, sg, xs, yc);
}

// ---------------------------------------------------------------------------
// Batched 2D forward, LV time levels per launch (temporal blocking along the
// row march, no sponge): levels n .. n+LV-1 from levels n-1 (pc) and n-2 (pp).
// Level l of the launch (1-based) trails level 1 by l-1 rows in the direction
// of the march; it is computed on LV-l halo columns either side of the lane's
// CX columns and on LV-l halo rows either side of the row chunk, so what it
// hands to level l+1 equals the owners' values.  Every level is the FP64
// update of k_fwd2d with the same operand order, so levels and traces are
// bit-identical to LV single-level launches.  Per shot-cell a launch reads
// levels n-1 and n-2 (hi + lo) once and writes the parts whose pointers `out`
// sets: the next launch needs hi + lo of the last two levels and a gradient
// keeps the hi of every level, so a gradient's forward moves 12 bytes per
// shot-cell and level at LV = 2, 28/3 at LV = 3 and 8 at LV = 4.
// Row chunks, their order and march direction: sched, as k_fwd2d_x2.
// ---------------------------------------------------------------------------
template <int LV, class L>
struct FwdLevels {
    float *hi[LV];                // level n+i (index i), FP32 part; nullptr: not stored
    L *lo[LV];                    // its low part; nullptr: not stored
    const double *sincd[LV];      // S_src w[n+i-1] per shot (FP64)
    float *rec[LV], *rec_lo[LV];  // trace rows of level n+i (REC)
};

template <int LB, int CX, int LV, bool REC, class L = uint16_t>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(USCT_OCCXN(LV))))
k_fwd2d_xn(const float *__restrict__ pc, const float *__restrict__ pp, const L *__restrict__ pcl,
           const L *__restrict__ ppl, const FwdLevels<LV, L> out, const float2 *__restrict__ sp2,
           const int *__restrict__ src_cell, const int *__restrict__ rec_head,
           const int *__restrict__ rec_next, const int *__restrict__ row_rec, int ny, int nx,
           int Bp, int RY, int nxs, int nyc, int ng, int sched) {
    static_assert(LV >= 2 && LV <= 4, "levels per launch");
    constexpr int W = kWave / LB;
    constexpr int NC = CX + 2 * LV;  // columns x0-LV .. x0+CX+LV-1 (index g)
    const int wv = wave_id();
    const int lane = threadIdx.x & 63;
    int sg, xs, yc, dir;
    if (!chunk2d(wv, ng, nxs, nyc, sched, sg, xs, yc, dir)) return;
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    if (x0 >= nx) return;
    const int y0 = yc * RY;
    const int y1 = min(y0 + RY, ny);
    const int my_src = src_cell[b];
    double inc[LV];
#pragma unroll
    for (int i = 0; i < LV; ++i) inc[i] = out.sincd[i][b];
    const size_t rowE = (size_t)nx * Bp;
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned o2 = (unsigned)(x0 * Bp + b) * (unsigned)sizeof(L);
    const unsigned oc = (unsigned)x0 * 8u;
    // level 1's first row: LV-1 halo rows outside the chunk on the side the
    // march starts from; RY + 2(LV-1) fronts
    const int yfs = dir > 0 ? y0 - (LV - 1) : y1 + (LV - 2);
    const int n_front = y1 - y0 + 2 * (LV - 1);
    auto rowE_at = [&](int yf, int r) { return clamp_row(yf + r * dir, ny) * rowE; };
    // S[l][k][g]: level l of the launch (l = 0: level n-1 as loaded) in FP64.
    // With front row yf (the row level 1 computes at front t), level l's row
    // is yf - (l-1)*dir (level 0's: yf + dir), and the row a level takes at
    // front t sits in slot k = t % 3: with the front loop unrolled by three the
    // slots of the newest, middle and oldest rows are constants and no row is
    // ever copied.  q: level n-2 on row yf.  nh/nl and qh/ql: raw loads issued
    // one front ahead (level n-1 on row yf+2dir, level n-2 on yf+dir).
    double S[LV][3][NC];
    double q[NC];
    float nh[NC], qh[NC];
    L nl[NC], ql[NC];
    {
        // level 0's rows of fronts -1 (slot 2) and -2 (slot 1)
        const size_t r0 = rowE_at(yfs, -1), r1 = rowE_at(yfs, 0), r2 = rowE_at(yfs, 1);
#pragma unroll
        for (int g = 0; g < NC; ++g) {
            const ptrdiff_t c = (ptrdiff_t)(g - LV) * Bp;
            S[0][2][g] = hl(ldb(pc + r1 + c, o4), ldb(pcl + r1 + c, o2));
            nh[g] = ldb(pc + r2 + c, o4);
            nl[g] = ldb(pcl + r2 + c, o2);
            if (g >= 1 && g < NC - 1) {
                S[0][1][g] = hl(ldb(pc + r0 + c, o4), ldb(pcl + r0 + c, o2));
                qh[g] = ldn<USCT_NT2D>(pp + r1 + c, o4);
                ql[g] = ldn<USCT_NT2D>(ppl + r1 + c, o2);
            } else {
                S[0][1][g] = 0.0;
            }
        }
#pragma unroll
        for (int l = 1; l < LV; ++l)
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int g = 0; g < NC; ++g) S[l][k][g] = 0.0;
    }
    for (int t0 = 0; t0 < n_front; t0 += 3) {
#pragma unroll
        for (int ph = 0; ph < 3; ++ph) {
            const int t = t0 + ph;
            if (t >= n_front) break;  // uniform over the wavefront
            const int yf = yfs + t * dir;
            const int kn = ph, km = (ph + 2) % 3, ko = (ph + 1) % 3;  // newest, middle, oldest
#pragma unroll
            for (int g = 0; g < NC; ++g) S[0][kn][g] = hl(nh[g], nl[g]);
#pragma unroll
            for (int g = 1; g < NC - 1; ++g) q[g] = hl(qh[g], ql[g]);
            // ---- loads for the next front (none after the last)
            if (t + 1 < n_front) {
                const size_t rnn = rowE_at(yf, 2), en = rowE_at(yf, 1);
#pragma unroll
                for (int g = 0; g < NC; ++g) {
                    const ptrdiff_t c = (ptrdiff_t)(g - LV) * Bp;
                    nh[g] = ldb(pc + rnn + c, o4);
                    nl[g] = ldb(pcl + rnn + c, o2);
                }
#pragma unroll
                for (int g = 1; g < NC - 1; ++g) {
                    const ptrdiff_t c = (ptrdiff_t)(g - LV) * Bp;
                    qh[g] = ldn<USCT_NT2D>(pp + en + c, o4);
                    ql[g] = ldn<USCT_NT2D>(ppl + en + c, o2);
                }
            }
            // ---- level l on row yf - (l-1)*dir, columns x0-LV+l .. x0+CX+LV-l-1
#pragma unroll
            for (int l = 1; l <= LV; ++l) {
                const int y = yf - (l - 1) * dir;
                const bool act = y >= y0 - (LV - l) && y < y1 + (LV - l);
                const bool own = y >= y0 && y < y1;
                const bool yint = y >= 1 && y <= ny - 2;
                const int yr = (int)clamp_row(y, ny);
                const size_t ro = (size_t)yr * rowE;
                const float2 *spr = sp2 + (size_t)yr * nx;
                const bool rec_row = REC && own && row_rec[yr] > 0;
                float *const oh = out.hi[l - 1];
                L *const ol = out.lo[l - 1];
#pragma unroll
                for (int g = l; g < NC - l; ++g) {
                    const int x = x0 - LV + g;
                    double nw = 0.0;
                    if (act) {
                        // level l-1 on rows y+dir (newest), y, y-dir; level l-2 on row y
                        const double cur = S[l - 1][km][g];
                        const double prev = l == 1 ? q[g] : S[l - 2][ko][g];
                        double lu = 0.0;
                        if (yint && x >= 1 && x <= nx - 2) {
                            const double yp = dir > 0 ? S[l - 1][kn][g] : S[l - 1][ko][g];
                            const double ym = dir > 0 ? S[l - 1][ko][g] : S[l - 1][kn][g];
                            lu = lap5d(cur, yp, ym, S[l - 1][km][g + 1], S[l - 1][km][g - 1]);
                        }
                        double incr = s2d(ldb(spr + (g - LV), oc)) * lu;
                        if (y * nx + x == my_src) incr = incr + inc[l - 1];
                        const double nv = (2.0 * cur - prev) + incr;
                        L lb;
                        const float v = split_hl(nv, lb);
                        nw = hl(v, lb);
                        if (own && g >= LV && g < LV + CX && x < nx) {
                            const ptrdiff_t e = (ptrdiff_t)ro + (ptrdiff_t)(g - LV) * Bp;
                            if (oh) stn<USCT_NT2D>(oh + e, o4, v);
                            if (ol) stn<USCT_NT2D>(ol + e, o2, lb);
                            if (rec_row) {
                                for (int r = rec_head[yr * nx + x]; r >= 0; r = rec_next[r]) {
                                    out.rec[l - 1][(size_t)r * Bp + b] = v;
                                    out.rec_lo[l - 1][(size_t)r * Bp + b] = lo_f(lb);
                                }
                            }
                        }
                    }
                    if (l < LV) S[l][kn][g] = nw;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Gauss-Newton, 2D, no sponge: background levels n and n+1 (exactly as
// k_fwd2d_x2, without trace recording) and the Born levels n and n+1 that use
// them (exactly as k_born), in one launch.  The Born source -(v/m)(U^l -
// 2U^{l-1} + U^{l-2}) takes the three background levels from the registers
// that hold them as stored (hi + lo), so per shot-cell the launch reads two
// background and two Born levels and writes two of each (48 B per two levels
// against 24 + 2 x 36 for k_fwd2d_x2 and two k_born launches).  Born traces
// of level n go to res_a, of level n+1 to res_b.  The Born state rotates over
// four buffers so that db (level n+1) is never dp (level n-2), which
// neighbouring wavefronts still read.  With four cells per lane the two
// fields want 256 VGPRs: two wavefronts per SIMD spill 14 of them and still
// run faster than one wavefront per SIMD without spills.
// ---------------------------------------------------------------------------
#ifndef USCT_OCCGN
#define USCT_OCCGN 2
#endif
template <int LB, int CX>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(USCT_OCCGN)))
k_born2d_x2(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pa,
            float *__restrict__ pb, const uint16_t *__restrict__ pcl,
            const uint16_t *__restrict__ ppl, uint16_t *__restrict__ pal,
            uint16_t *__restrict__ pbl, const float *__restrict__ dc, const float *__restrict__ dp,
            float *__restrict__ da, float *__restrict__ db, const uint16_t *__restrict__ dcl,
            const uint16_t *__restrict__ dpl, uint16_t *__restrict__ dal,
            uint16_t *__restrict__ dbl, const float2 *__restrict__ sp2,
            const float2 *__restrict__ vm2, const int *__restrict__ src_cell,
            const double *__restrict__ sincd_a, const double *__restrict__ sincd_b,
            const int *__restrict__ rec_head, const int *__restrict__ rec_next,
            const int *__restrict__ row_rec, float *__restrict__ res_a,
            float *__restrict__ res_b, int ny, int nx, int Bp, int RY, int nxs, int nyc, int ng) {
    constexpr int W = kWave / LB;
    constexpr int CA = CX + 2;  // columns x0-1 .. x0+CX
    const int wv = wave_id();
    const int lane = threadIdx.x & 63;
    const int sg = wv % ng;
    const int t = wv / ng;
    const int xs = t % nxs, yc = t / nxs;
    if (yc >= nyc) return;
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    if (x0 >= nx) return;
    const int y0 = yc * RY;
    const int y1 = min(y0 + RY, ny);
    const int my_src = src_cell[b];
    const double inc_a = sincd_a[b], inc_b = sincd_b[b];
    const size_t rowE = (size_t)nx * Bp;
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned o2 = (unsigned)(x0 * Bp + b) * 2u;
    const unsigned oc = (unsigned)x0 * 8u;
    // Background as in k_fwd2d_x2: um / uc / un = level n-1 on rows yf-1, yf,
    // yf+1; q = level n-2 on row yf; am / ac / an = level n on rows yf-2 .. yf.
    // The Born field has the same roles: em / ec / en, r, bm / bc / bn.
    double um[CA], uc[CA + 2], un[CA + 2], am[CA], ac[CA], an[CA];
    double em[CA], ec[CA + 2], en[CA + 2], bm[CA], bc[CA], bn[CA];
    float nh[CA + 2], qh[CA], mh[CA + 2], rh[CA];
    uint16_t nl[CA + 2], ql[CA], ml[CA + 2], rl[CA];
    {
        const size_t r0 = clamp_row(y0 - 2, ny) * rowE, r1 = clamp_row(y0 - 1, ny) * rowE;
        const size_t r2 = clamp_row(y0, ny) * rowE;
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            const ptrdiff_t c = (ptrdiff_t)(j - 1) * Bp;
            um[j] = hl(ldb(pc + r0 + c, o4), ldb(pcl + r0 + c, o2));
            em[j] = hl(ldb(dc + r0 + c, o4), ldb(dcl + r0 + c, o2));
            qh[j] = ldn<USCT_NT2D>(pp + r1 + c, o4);
            ql[j] = ldn<USCT_NT2D>(ppl + r1 + c, o2);
            rh[j] = ldn<USCT_NT2D>(dp + r1 + c, o4);
            rl[j] = ldn<USCT_NT2D>(dpl + r1 + c, o2);
        }
#pragma unroll
        for (int j = 0; j < CA + 2; ++j) {
            const ptrdiff_t c = (ptrdiff_t)(j - 2) * Bp;
            uc[j] = hl(ldb(pc + r1 + c, o4), ldb(pcl + r1 + c, o2));
            ec[j] = hl(ldb(dc + r1 + c, o4), ldb(dcl + r1 + c, o2));
            nh[j] = ldb(pc + r2 + c, o4);
            nl[j] = ldb(pcl + r2 + c, o2);
            mh[j] = ldb(dc + r2 + c, o4);
            ml[j] = ldb(dcl + r2 + c, o2);
        }
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            am[j] = 0.0; ac[j] = 0.0; an[j] = 0.0;
            bm[j] = 0.0; bc[j] = 0.0; bn[j] = 0.0;
        }
    }
    for (int yf = y0 - 1; yf <= y1; ++yf) {
        const int yr = (int)clamp_row(yf, ny);
#pragma unroll
        for (int j = 0; j < CA + 2; ++j) {
            un[j] = hl(nh[j], nl[j]);
            en[j] = hl(mh[j], ml[j]);
        }
        double q[CA], r[CA];
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            q[j] = hl(qh[j], ql[j]);
            r[j] = hl(rh[j], rl[j]);
        }
        {
            const size_t rnn = clamp_row(yf + 2, ny) * rowE;
            const size_t enx = clamp_row(yf + 1, ny) * rowE;
#pragma unroll
            for (int j = 0; j < CA + 2; ++j) {
                const ptrdiff_t c = (ptrdiff_t)(j - 2) * Bp;
                nh[j] = ldb(pc + rnn + c, o4);
                nl[j] = ldb(pcl + rnn + c, o2);
                mh[j] = ldb(dc + rnn + c, o4);
                ml[j] = ldb(dcl + rnn + c, o2);
            }
#pragma unroll
            for (int j = 0; j < CA; ++j) {
                const ptrdiff_t c = (ptrdiff_t)(j - 1) * Bp;
                qh[j] = ldn<USCT_NT2D>(pp + enx + c, o4);
                ql[j] = ldn<USCT_NT2D>(ppl + enx + c, o2);
                rh[j] = ldn<USCT_NT2D>(dp + enx + c, o4);
                rl[j] = ldn<USCT_NT2D>(dpl + enx + c, o2);
            }
        }
        // ---- levels n on row yf, columns x0-1 .. x0+CX
        const bool own_a = yf >= y0 && yf < y1;
        const bool yint = yf >= 1 && yf <= ny - 2;
        const size_t ro = (size_t)yr * rowE;
        const float2 *spr = sp2 + (size_t)yr * nx;
        const float2 *vmr = vm2 + (size_t)yr * nx;
        const bool rec_a_row = own_a && row_rec[yr] > 0;
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            const int x = x0 - 1 + j;
            const bool in = yint && x >= 1 && x <= nx - 2;
            const double sv = s2d(ldb(spr + (j - 1), oc));
            double lu = 0.0;
            if (in) lu = lap5d(uc[j + 1], un[j + 1], um[j], uc[j + 2], uc[j]);
            double inc = sv * lu;
            if (yf * nx + x == my_src) inc = inc + inc_a;
            const double nv = (2.0 * uc[j + 1] - q[j]) + inc;
            uint16_t lb;
            const float v = split_hl(nv, lb);
            am[j] = ac[j];
            ac[j] = an[j];
            an[j] = hl(v, lb);
            double ld = 0.0;
            if (in) ld = lap5d(ec[j + 1], en[j + 1], em[j], ec[j + 2], ec[j]);
            const double d2u = (an[j] - 2.0 * uc[j + 1]) + q[j];
            const double dv = (2.0 * ec[j + 1] - r[j]) + sv * ld - s2d(ldb(vmr + (j - 1), oc)) * d2u;
            uint16_t dlb;
            const float w = split_hl(dv, dlb);
            bm[j] = bc[j];
            bc[j] = bn[j];
            bn[j] = hl(w, dlb);
            if (j >= 1 && j <= CX && own_a && x < nx) {
                const ptrdiff_t e = (ptrdiff_t)ro + (ptrdiff_t)(j - 1) * Bp;
                stn<USCT_NT2D>(pa + e, o4, v);
                stn<USCT_NT2D>(pal + e, o2, lb);
                stn<USCT_NT2D>(da + e, o4, w);
                stn<USCT_NT2D>(dal + e, o2, dlb);
                if (rec_a_row)
                    for (int rr = rec_head[yr * nx + x]; rr >= 0; rr = rec_next[rr])
                        res_a[(size_t)rr * Bp + b] = w;
            }
        }
        // ---- levels n+1 on row y = yf-1, own columns
        const int y = yf - 1;
        if (y >= y0 && y < y1) {
            const bool yint_b = y >= 1 && y <= ny - 2;
            const size_t rob = (size_t)y * rowE;
            const float2 *spb = sp2 + (size_t)y * nx;
            const float2 *vmb = vm2 + (size_t)y * nx;
            const bool rec_b_row = row_rec[y] > 0;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                if (x < nx) {
                    const bool in = yint_b && x >= 1 && x <= nx - 2;
                    const double sv = s2d(ldb(spb + i, oc));
                    double lu = 0.0;
                    if (in) lu = lap5d(ac[i + 1], an[i + 1], am[i + 1], ac[i + 2], ac[i]);
                    double inc = sv * lu;
                    if (y * nx + x == my_src) inc = inc + inc_b;
                    const double nv = (2.0 * ac[i + 1] - um[i + 1]) + inc;
                    uint16_t lb;
                    const float v = split_hl(nv, lb);
                    double ld = 0.0;
                    if (in) ld = lap5d(bc[i + 1], bn[i + 1], bm[i + 1], bc[i + 2], bc[i]);
                    const double d2u = (hl(v, lb) - 2.0 * ac[i + 1]) + um[i + 1];
                    const double dv = (2.0 * bc[i + 1] - em[i + 1]) + sv * ld - s2d(ldb(vmb + i, oc)) * d2u;
                    uint16_t dlb;
                    const float w = split_hl(dv, dlb);
                    const size_t e = rob + (size_t)i * Bp;
                    stn<USCT_NT2D>(pb + e, o4, v);
                    stn<USCT_NT2D>(pbl + e, o2, lb);
                    stn<USCT_NT2D>(db + e, o4, w);
                    stn<USCT_NT2D>(dbl + e, o2, dlb);
                    if (rec_b_row)
                        for (int rr = rec_head[y * nx + x]; rr >= 0; rr = rec_next[rr])
                            res_b[(size_t)rr * Bp + b] = w;
                }
            }
        }
        // ---- advance
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            um[j] = uc[j + 1];
            em[j] = ec[j + 1];
        }
#pragma unroll
        for (int j = 0; j < CA + 2; ++j) {
            uc[j] = un[j];
            ec[j] = en[j];
        }
    }
}

// ---------------------------------------------------------------------------
// Batched 2D adjoint step with the imaging condition.
//   l1 = lam^{k+1}, l2 = lam^{k+2}, lk = lam^k (output)
//   no sponge: pa = p^{k-1};  sponge: pa = p^k, pb = p^{k-1}, pc2 = p^{k-2}
//   res = [n_rec][Bp] residual row of level k; sinc[b] = S_src w[k-1]
//   g_hi/g_lo: [ng][N] Neumaier accumulators of sum_shots sum_k lam^k d2p^k
// Addressing as k_fwd2d (lanes past the last column still load, from the
// guard zone, and join the reductions with zero contributions).
// ---------------------------------------------------------------------------
template <int LB, int CX, bool SPONGE>
__global__ void __launch_bounds__(kBlock) USCT_ATTR2D
k_adj2d(const float *__restrict__ l1, const float *__restrict__ l2, float *__restrict__ lk,
        const float *__restrict__ pa, const float *__restrict__ pb, const float *__restrict__ pc2,
        const float2 *__restrict__ sp2, const float *__restrict__ sponge,
        const int *__restrict__ src_cell, const float *__restrict__ sinc,
        const int *__restrict__ rec_head, const int *__restrict__ rec_next,
        const int *__restrict__ row_rec, const float *__restrict__ res,
        const float *__restrict__ srcd2, float *__restrict__ g_hi, float *__restrict__ g_lo,
        int ny, int nx, int Bp, int RY, int nxs, int nyc, int ng) {
    constexpr int W = kWave / LB;
    const int wv = wave_id();
    const int lane = threadIdx.x & 63;
    const int sg = wv % ng;
    const int t = wv / ng;
    const int xs = t % nxs, yc = t / nxs;
    if (yc >= nyc) return;  // uniform over the wavefront
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    const int y0 = yc * RY;
    const int y1 = min(y0 + RY, ny);
    const size_t rowE = (size_t)nx * Bp;
    const size_t N = (size_t)ny * nx;
    (void)srcd2;
    const int my_src = src_cell[b];
    const float my_inc = sinc[b];
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned oc = (unsigned)x0 * 8u;
    // q = S' lam^{k+1} rows (qm, qc with halo, qp); lc: raw centre row of lam^{k+1}.
    float qm[CX], qc[CX + 2], qp[CX], lc[CX], lpn[CX];
    float fm[CX], fc[CX + 2], fp[CX];  // p^{k-1} rows (no sponge)
    {
        const int ym = max(y0 - 1, 0);
        const float *a0 = l1 + (size_t)ym * rowE;
        const float *a1 = l1 + (size_t)y0 * rowE;
        const float2 *s0 = sp2 + (size_t)ym * nx;
        const float2 *s1 = sp2 + (size_t)y0 * nx;
#pragma unroll
        for (int i = 0; i < CX; ++i) qm[i] = smul(ldb(s0 + i, oc), ldb(a0 + (ptrdiff_t)i * Bp, o4));
#pragma unroll
        for (int i = 0; i < CX + 2; ++i) {
            const float v = ldb(a1 + (ptrdiff_t)(i - 1) * Bp, o4);
            qc[i] = smul(ldb(s1 + (i - 1), oc), v);
            if (i >= 1 && i <= CX) lc[i - 1] = v;
        }
        if (!SPONGE) {
            const float *f0 = pa + (size_t)ym * rowE;
            const float *f1 = pa + (size_t)y0 * rowE;
#pragma unroll
            for (int i = 0; i < CX; ++i) fm[i] = ldb(f0 + (ptrdiff_t)i * Bp, o4);
#pragma unroll
            for (int i = 0; i < CX + 2; ++i) fc[i] = ldb(f1 + (ptrdiff_t)(i - 1) * Bp, o4);
        }
    }
    for (int y = y0; y < y1; ++y) {
        const int yp = min(y + 1, ny - 1);
        const float *lnr = l1 + (size_t)yp * rowE;
        const float2 *snr = sp2 + (size_t)yp * nx;
        const float *fnr = pa + (size_t)yp * rowE;
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            const float v = ldb(lnr + (ptrdiff_t)i * Bp, o4);
            lpn[i] = v;
            qp[i] = smul(ldb(snr + i, oc), v);
            if (!SPONGE) fp[i] = ldb(fnr + (ptrdiff_t)i * Bp, o4);
        }
        const bool yint = (y >= 1) && (y <= ny - 2);
        const size_t ro = (size_t)y * rowE;
        const float2 *spr = sp2 + (size_t)y * nx;
        const bool has_rec = row_rec[y] > 0;
        float tv[CX];
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            const int x = x0 + i;
            tv[i] = 0.0f;
            if (x < nx) {
                const bool in = yint && x >= 1 && x <= nx - 2;
                const float lq = in ? lap5(qc[i + 1], qp[i], qm[i], qc[i + 2], qc[i]) : 0.0f;
                const size_t ei = ro + (size_t)i * Bp;
                float v = leap(lc[i], ldb(l2 + ei, o4), lq);
                if (SPONGE) v = v * sponge[(size_t)y * nx + x];
                if (has_rec) {
                    for (int j = rec_head[y * nx + x]; j >= 0; j = rec_next[j])
                        v = v - res[(size_t)j * Bp + b];
                }
                stb(lk + ei, o4, v);
                float d2;
                if (SPONGE) {
                    d2 = (ldb(pa + ei, o4) - 2.0f * ldb(pb + ei, o4)) + ldb(pc2 + ei, o4);
                } else {
                    const float lp = in ? lap5(fc[i + 1], fp[i], fm[i], fc[i + 2], fc[i]) : 0.0f;
                    d2 = smul(ldb(spr + i, oc), lp);
                    if (y * nx + x == my_src) d2 = d2 + my_inc;
                }
                tv[i] = v * d2;
            }
        }
        const float s = reduce_scatter<LB, CX>(tv, lane);
        if ((sb & (LB / CX - 1)) == 0) {
            const int x = x0 + sum_cell<LB, CX>(lane);
            if (x < nx) {
                float *gh = g_hi + (size_t)sg * N + (size_t)y * nx;
                float *gl = g_lo + (size_t)sg * N + (size_t)y * nx;
                float hi = gh[x], lo = gl[x];
                neumaier(hi, lo, s);
                gh[x] = hi;
                gl[x] = lo;
            }
        }
        if (y + 1 < y1) {
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                qm[i] = qc[i + 1];
                qc[i + 1] = qp[i];
                lc[i] = lpn[i];
            }
            qc[0] = smul(ldb(snr - 1, oc), ldb(lnr - Bp, o4));
            qc[CX + 1] = smul(ldb(snr + CX, oc), ldb(lnr + (ptrdiff_t)CX * Bp, o4));
            if (!SPONGE) {
#pragma unroll
                for (int i = 0; i < CX; ++i) { fm[i] = fc[i + 1]; fc[i + 1] = fp[i]; }
                fc[0] = ldb(fnr - Bp, o4);
                fc[CX + 1] = ldb(fnr + (ptrdiff_t)CX * Bp, o4);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Batched 2D adjoint, two levels per launch (temporal blocking, no sponge):
// lam^k (U) from lam^{k+1} (l1) and lam^{k+2} (l2), then lam^{k-1} (V) from U
// and l1, each imaged against its own stored forward level:
//   g += U d2p^k     with d2p^k     = S' lap_u(p^{k-1}) + S_src w[k-1]  (pa1)
//   g += V d2p^{k-1} with d2p^{k-1} = S' lap_u(p^{k-2}) + S_src w[k-2]  (pa2)
// V trails U by one row; U is computed on one halo column either side and one
// halo row above and below the wavefront's rows, with the receiver residuals
// of those cells, so it equals the owners' values.  Each accumulator cell gets
// the two Neumaier updates in the order two single-level launches apply them,
// so fields and gradient are bit-identical to those launches.  Per shot-cell
// the pair reads 16 bytes and writes 8 instead of reading 24 and writing 8.
//   sinc_u/sinc_v: S_src w[k-1], S_src w[k-2]; res_u/res_v: residual rows k, k-1.
// Row chunks, their order and march direction: sched, as for k_fwd2d_x2 (V
// trails U by one row in the direction of the march).
// ---------------------------------------------------------------------------
template <int LB, int CX, int DIR>
__device__ __forceinline__ void adj2d_x2_body(
    const float *__restrict__ l1, const float *__restrict__ l2, float *__restrict__ lu,
    float *__restrict__ lv, const float *__restrict__ pa1, const float *__restrict__ pa2,
    const float2 *__restrict__ sp2, const int *__restrict__ src_cell,
    const float *__restrict__ sinc_u, const float *__restrict__ sinc_v,
    const int *__restrict__ rec_head, const int *__restrict__ rec_next,
    const int *__restrict__ row_rec, const float *__restrict__ res_u,
    const float *__restrict__ res_v, float *__restrict__ g_hi, float *__restrict__ g_lo, int ny,
    int nx, int Bp, int RY, int sg, int xs, int yc) {
    constexpr int W = kWave / LB;
    constexpr int CA = CX + 2;  // U columns x0-1 .. x0+CX
    constexpr int dir = DIR;
    const int lane = threadIdx.x & 63;
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    const int y0 = yc * RY;
    const int y1 = min(y0 + RY, ny);
    const size_t rowE = (size_t)nx * Bp;
    const size_t N = (size_t)ny * nx;
    const int my_src = src_cell[b];
    const float inc_u = sinc_u[b], inc_v = sinc_v[b];
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned oc = (unsigned)x0 * 8u;
    const int yfs = dir > 0 ? y0 - 1 : y1;  // first front row (as k_fwd2d_x2)
    auto row_at = [&](int yf, int r) { return clamp_row(yf + r * dir, ny); };
    // Registers at the top of the iteration for front row yf, rows named
    // relative to it (row(r) = yf + r * dir; index j <-> column x0-2+j for the
    // CA+2-wide rows, x0-1+j for CA-wide, x0+j for CX-wide):
    //   qm, qc       S' lam^{k+1} rows (-1) (CA), (0) (CA+2)
    //   lc           lam^{k+1} row (0) (CA)
    //   lr           lam^{k+1} row (+1) (CA+2), loaded one iteration ahead
    //   l2c          lam^{k+2} row (0) (CA), loaded one iteration ahead
    //   fm, fc, fr   p^{k-1} rows (-1) (CX), (0) (CX+2), (+1) (CX+2, ahead)
    //   hm, hc, hr   p^{k-2} rows (-2) (CX), (-1) (CX+2), (0) (CX+2, ahead)
    //   wm, wc, wn   S' U rows (-2), (-1), (0) (CA; shifted as U is made)
    //   l1p, up      row (-1), own columns: lam^{k+1} and U
    float qm[CA], qc[CA + 2], lc[CA], lr[CA + 2], l2c[CA];
    float fm[CX], fc[CX + 2], fr[CX + 2];
    float hm[CX], hc[CX + 2], hr[CX + 2];
    float wm[CA], wc[CA], wn[CA];
    float l1p[CX], up[CX];
    float s_prev = 0.0f;  // U imaging sum of row (-1) (this lane's cell)
    {
        const size_t r0 = row_at(yfs, -1), r1 = row_at(yfs, 0);
        const size_t r2 = row_at(yfs, 1);
        const float *a0 = l1 + r0 * rowE, *a1 = l1 + r1 * rowE, *a2 = l1 + r2 * rowE;
        const float2 *s0 = sp2 + r0 * nx, *s1 = sp2 + r1 * nx;
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            qm[j] = smul(ldb(s0 + (j - 1), oc), ldb(a0 + (ptrdiff_t)(j - 1) * Bp, o4));
            l2c[j] = ldn<USCT_NT2D>(l2 + r1 * rowE + (ptrdiff_t)(j - 1) * Bp, o4);
        }
#pragma unroll
        for (int j = 0; j < CA + 2; ++j) {
            const float v = ldb(a1 + (ptrdiff_t)(j - 2) * Bp, o4);
            qc[j] = smul(ldb(s1 + (j - 2), oc), v);
            if (j >= 1 && j <= CA) lc[j - 1] = v;
            lr[j] = ldb(a2 + (ptrdiff_t)(j - 2) * Bp, o4);
        }
        const float *h0 = pa2 + row_at(yfs, -2) * rowE;
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            fm[i] = ldb(pa1 + r0 * rowE + (ptrdiff_t)i * Bp, o4);
            hm[i] = ldb(h0 + (ptrdiff_t)i * Bp, o4);
        }
#pragma unroll
        for (int i = 0; i < CX + 2; ++i) {
            fc[i] = ldb(pa1 + r1 * rowE + (ptrdiff_t)(i - 1) * Bp, o4);
            fr[i] = ldb(pa1 + r2 * rowE + (ptrdiff_t)(i - 1) * Bp, o4);
            hc[i] = ldb(pa2 + r0 * rowE + (ptrdiff_t)(i - 1) * Bp, o4);
            hr[i] = ldb(pa2 + r1 * rowE + (ptrdiff_t)(i - 1) * Bp, o4);
        }
#pragma unroll
        for (int j = 0; j < CA; ++j) { wm[j] = 0.0f; wc[j] = 0.0f; wn[j] = 0.0f; }
#pragma unroll
        for (int i = 0; i < CX; ++i) { l1p[i] = 0.0f; up[i] = 0.0f; }
    }
    // the lane's reduced cell (the same on every row); its accumulators of the row
    // updated at the end of an iteration are loaded at the iteration's start
    const int gx = x0 + sum_cell<LB, CX>(lane);
    const bool gval = (sb & (LB / CX - 1)) == 0 && gx < nx;
    float *const ghb = g_hi + (size_t)sg * N + gx;
    float *const glb = g_lo + (size_t)sg * N + gx;
    const int n_front = y1 - y0 + 2;
    for (int it = 0, yf = yfs; it < n_front; ++it, yf += dir) {
        const size_t yr = clamp_row(yf, ny), yn = row_at(yf, 1);
        const int y = yf - dir;  // row (-1): the V level's row
        float gh0 = 0.0f, gl0 = 0.0f;
        if (gval && y >= y0 && y < y1) {
            gh0 = ghb[(size_t)y * nx];
            gl0 = glb[(size_t)y * nx];
        }
        // ---- loads for the next iteration
        float ln2[CA + 2], l2n[CA], fn2[CX + 2], hn[CX + 2];
        {
            const size_t r2 = row_at(yf, 2);
            const float *a = l1 + r2 * rowE;
            const float *c = l2 + yn * rowE;
            const float *f = pa1 + r2 * rowE;
            const float *h = pa2 + yn * rowE;
#pragma unroll
            for (int j = 0; j < CA + 2; ++j) ln2[j] = ldb(a + (ptrdiff_t)(j - 2) * Bp, o4);
#pragma unroll
            for (int j = 0; j < CA; ++j) l2n[j] = ldn<USCT_NT2D>(c + (ptrdiff_t)(j - 1) * Bp, o4);
#pragma unroll
            for (int i = 0; i < CX + 2; ++i) {
                fn2[i] = ldb(f + (ptrdiff_t)(i - 1) * Bp, o4);
                hn[i] = ldb(h + (ptrdiff_t)(i - 1) * Bp, o4);
            }
        }
        // ---- S' lam^{k+1} on row (+1)
        float qn[CA + 2];
        {
            const float2 *snr = sp2 + yn * nx;
#pragma unroll
            for (int j = 0; j < CA + 2; ++j) qn[j] = smul(ldb(snr + (j - 2), oc), lr[j]);
        }
        // ---- U = lam^k on row yf, columns x0-1 .. x0+CX
        const bool own_u = yf >= y0 && yf < y1;
        const bool yint = yf >= 1 && yf <= ny - 2;
        const bool rec_u = yf >= 0 && yf < ny && row_rec[yr] > 0;
        const float2 *spr = sp2 + yr * nx;
        float uo[CX], tv[CX];
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            const int x = x0 - 1 + j;
            const bool in = yint && x >= 1 && x <= nx - 2;
            // rows y+1 and y-1 of S' lam^{k+1}
            const float yp = dir > 0 ? qn[j + 1] : qm[j];
            const float ym = dir > 0 ? qm[j] : qn[j + 1];
            const float lq = in ? lap5(qc[j + 1], yp, ym, qc[j + 2], qc[j]) : 0.0f;
            float v = leap(lc[j], l2c[j], lq);
            if (rec_u && x >= 0 && x < nx) {
                for (int r = rec_head[(int)yr * nx + x]; r >= 0; r = rec_next[r])
                    v = v - res_u[(size_t)r * Bp + b];
            }
            wm[j] = wc[j];
            wc[j] = wn[j];
            wn[j] = smul(ldb(spr + (j - 1), oc), v);
            if (j >= 1 && j <= CX) {
                uo[j - 1] = v;
                if (own_u && x < nx) stn<USCT_NT2D>(lu + yr * rowE + (size_t)(j - 1) * Bp, o4, v);
            }
        }
        float s_u = 0.0f;
        if (own_u) {
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                tv[i] = 0.0f;
                if (x < nx) {
                    const bool in = yint && x >= 1 && x <= nx - 2;
                    const float yp = dir > 0 ? fr[i + 1] : fm[i];
                    const float ym = dir > 0 ? fm[i] : fr[i + 1];
                    const float lp = in ? lap5(fc[i + 1], yp, ym, fc[i + 2], fc[i]) : 0.0f;
                    float d2 = smul(ldb(spr + i, oc), lp);
                    if ((int)yr * nx + x == my_src) d2 = d2 + inc_u;
                    tv[i] = uo[i] * d2;
                }
            }
            s_u = reduce_scatter<LB, CX>(tv, lane);
        }
        // ---- V = lam^{k-1} on row y = row(-1), own columns
        if (y >= y0 && y < y1) {
            const bool yint_v = y >= 1 && y <= ny - 2;
            const size_t rov = (size_t)y * rowE;
            const float2 *spv = sp2 + (size_t)y * nx;
            const bool rec_v = row_rec[y] > 0;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                tv[i] = 0.0f;
                if (x < nx) {
                    const bool in = yint_v && x >= 1 && x <= nx - 2;
                    const float wp = dir > 0 ? wn[i + 1] : wm[i + 1];
                    const float wq = dir > 0 ? wm[i + 1] : wn[i + 1];
                    const float lw = in ? lap5(wc[i + 1], wp, wq, wc[i + 2], wc[i]) : 0.0f;
                    float v = leap(up[i], l1p[i], lw);
                    if (rec_v) {
                        for (int r = rec_head[y * nx + x]; r >= 0; r = rec_next[r])
                            v = v - res_v[(size_t)r * Bp + b];
                    }
                    stn<USCT_NT2D>(lv + rov + (size_t)i * Bp, o4, v);
                    const float hp = dir > 0 ? hr[i + 1] : hm[i];
                    const float hq = dir > 0 ? hm[i] : hr[i + 1];
                    const float lp = in ? lap5(hc[i + 1], hp, hq, hc[i + 2], hc[i]) : 0.0f;
                    float d2 = smul(ldb(spv + i, oc), lp);
                    if (y * nx + x == my_src) d2 = d2 + inc_v;
                    tv[i] = v * d2;
                }
            }
            const float s_v = reduce_scatter<LB, CX>(tv, lane);
            if (gval) {
                float hi = gh0, lo = gl0;
                neumaier(hi, lo, s_prev);
                neumaier(hi, lo, s_v);
                ghb[(size_t)y * nx] = hi;
                glb[(size_t)y * nx] = lo;
            }
        }
        // ---- advance
#pragma unroll
        for (int i = 0; i < CX; ++i) { l1p[i] = lc[i + 1]; up[i] = uo[i]; }
        s_prev = s_u;
#pragma unroll
        for (int j = 0; j < CA; ++j) {
            qm[j] = qc[j + 1];
            lc[j] = lr[j + 1];
            l2c[j] = l2n[j];
        }
#pragma unroll
        for (int j = 0; j < CA + 2; ++j) { qc[j] = qn[j]; lr[j] = ln2[j]; }
#pragma unroll
        for (int i = 0; i < CX; ++i) { fm[i] = fc[i + 1]; hm[i] = hc[i + 1]; }
#pragma unroll
        for (int i = 0; i < CX + 2; ++i) {
            fc[i] = fr[i];
            fr[i] = fn2[i];
            hc[i] = hr[i];
            hr[i] = hn[i];
        }
    }
}

template <int LB, int CX>
__global__ void __launch_bounds__(kBlock) USCT_ATTR2D
k_adj2d_x2(const float *__restrict__ l1, const float *__restrict__ l2, float *__restrict__ lu,
           float *__restrict__ lv, const float *__restrict__ pa1, const float *__restrict__ pa2,
           const float2 *__restrict__ sp2, const int *__restrict__ src_cell,
           const float *__restrict__ sinc_u, const float *__restrict__ sinc_v,
           const int *__restrict__ rec_head, const int *__restrict__ rec_next,
           const int *__restrict__ row_rec, const float *__restrict__ res_u,
           const float *__restrict__ res_v, float *__restrict__ g_hi, float *__restrict__ g_lo,
           int ny, int nx, int Bp, int RY, int nxs, int nyc, int ng, int sched) {
    int sg, xs, yc, dir;
    if (!chunk2d(wave_id(), ng, nxs, nyc, sched, sg, xs, yc, dir)) return;  // wave-uniform
    if (dir > 0)
        adj2d_x2_body<LB, CX, 1>(l1, l2, lu, lv, pa1, pa2, sp2, src_cell, sinc_u, sinc_v,
                                 rec_head, rec_next, row_rec, res_u, res_v, g_hi, g_lo, ny, nx,
                                 Bp, RY, sg, xs, yc);
    else
        adj2d_x2_body<LB, CX, -1>(l1, l2, lu, lv, pa1, pa2, sp2, src_cell, sinc_u, sinc_v,
                                  rec_head, rec_next, row_rec, res_u, res_v, g_hi, g_lo, ny, nx,
                                  Bp, RY, sg, xs, yc);
}

// ---------------------------------------------------------------------------
// Batched 2D adjoint with the imaging condition, LV levels per launch (no
// sponge): lam^k .. lam^{k-LV+1} from lam^{k+1} (l1) and lam^{k+2} (l2).
// Level l of the launch (1-based, lam^{k-l+1}) trails level 1 by l-1 rows,
// on LV-l halo columns and rows (with the receiver residuals of those cells),
// as k_fwd2d_xn; it is imaged against its own stored forward level:
//   g += lam^{k-l+1} d2p^{k-l+1},  d2p^{k-l+1} = S' lap_u(p^{k-l}) + S_src w[k-l].
// A lane keeps each level's reduced imaging sum until the last level reaches
// the same row, and each accumulator cell takes its LV Neumaier updates in
// the order LV single-level launches apply them: fields and gradient are
// bit-identical to those launches.  Per shot-cell a launch reads lam^{k+1},
// lam^{k+2} and LV forward levels and writes the levels `out` asks for (the
// next launch reads the last two): 24 bytes per two levels, 28 per three, 32
// per four.
// ---------------------------------------------------------------------------
template <int LV>
struct AdjLevels {
    float *lam[LV];         // lam^{k-i} (index i); nullptr: not stored
    const float *pf[LV];    // p^{k-i-1}, FP32 part (the level d2p^{k-i} differentiates)
    const float *sinc[LV];  // S_src w[k-i-1] per shot
    const float *res[LV];   // residual row k-i
};

template <int LB, int CX, int LV>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(USCT_OCCXN(LV))))
k_adj2d_xn(const float *__restrict__ l1, const float *__restrict__ l2, const AdjLevels<LV> lv,
           const float2 *__restrict__ sp2, const int *__restrict__ src_cell,
           const int *__restrict__ rec_head, const int *__restrict__ rec_next,
           const int *__restrict__ row_rec, float *__restrict__ g_hi, float *__restrict__ g_lo,
           int ny, int nx, int Bp, int RY, int nxs, int nyc, int ng, int sched) {
    static_assert(LV >= 2 && LV <= 4, "levels per launch");
    constexpr int W = kWave / LB;
    constexpr int NC = CX + 2 * LV;  // columns x0-LV .. x0+CX+LV-1 (index g)
    const int wv = wave_id();
    const int lane = threadIdx.x & 63;
    int sg, xs, yc, dir;
    if (!chunk2d(wv, ng, nxs, nyc, sched, sg, xs, yc, dir)) return;  // uniform over the wavefront
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    const int y0 = yc * RY;
    const int y1 = min(y0 + RY, ny);
    const size_t rowE = (size_t)nx * Bp;
    const size_t N = (size_t)ny * nx;
    const int my_src = src_cell[b];
    float inc[LV];
#pragma unroll
    for (int i = 0; i < LV; ++i) inc[i] = lv.sinc[i][b];
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned oc = (unsigned)x0 * 8u;
    const int yfs = dir > 0 ? y0 - (LV - 1) : y1 + (LV - 2);
    const int n_front = y1 - y0 + 2 * (LV - 1);
    auto row_at = [&](int yf, int r) { return clamp_row(yf + r * dir, ny); };
    // Q[l][k][g] = S' lam and R[l][k][g] = lam of level l of the launch (l = 0:
    // lam^{k+1} as loaded).  With front row yf (the row level 1 computes at
    // front t), level l's row is yf - (l-1)*dir (level 0's: yf + dir), and the
    // row a level takes at front t sits in slot k = t % 3: with the front loop
    // unrolled by three no row is ever copied.  ln: lam^{k+1} on row yf+2dir,
    // c2 / c2n: lam^{k+2} on rows yf / yf+dir (raw loads one front ahead).
    // F[l][k][i]: forward level p^{k-l} (index l-1) around level l's row y,
    // columns x0-1 .. x0+CX (index i): row y+dir of front t in slot t % 3; the
    // row for the next front is loaded into the oldest slot once this front's
    // imaging has read it.
    float Q[LV][3][NC], R[LV][3][NC];
    float ln[NC], c2[NC], c2n[NC];
    float F[LV][3][CX + 2];
    float hist[LV][LV];  // imaging sums of levels 1 .. LV-1, newest first
    {
        // level 0's rows of fronts -1 (slot 2) and -2 (slot 1)
        const size_t r0 = row_at(yfs, -1), r1 = row_at(yfs, 0), r2 = row_at(yfs, 1);
        const float *a0 = l1 + r0 * rowE, *a1 = l1 + r1 * rowE, *a2 = l1 + r2 * rowE;
        const float2 *s0 = sp2 + r0 * nx, *s1 = sp2 + r1 * nx;
#pragma unroll
        for (int g = 0; g < NC; ++g) {
            const ptrdiff_t c = (ptrdiff_t)(g - LV) * Bp;
            ln[g] = ldb(a2 + c, o4);
            R[0][2][g] = ldb(a1 + c, o4);
            Q[0][2][g] = smul(ldb(s1 + (g - LV), oc), R[0][2][g]);
            if (g >= 1 && g < NC - 1) {
                R[0][1][g] = ldb(a0 + c, o4);
                Q[0][1][g] = smul(ldb(s0 + (g - LV), oc), R[0][1][g]);
                c2n[g] = ldn<USCT_NT2D>(l2 + r1 * rowE + c, o4);
            } else {
                R[0][1][g] = 0.0f;
                Q[0][1][g] = 0.0f;
            }
        }
#pragma unroll
        for (int l = 1; l < LV; ++l)
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int g = 0; g < NC; ++g) { Q[l][k][g] = 0.0f; R[l][k][g] = 0.0f; }
#pragma unroll
        for (int l = 1; l <= LV; ++l) {
            // rows y+dir, y, y-dir of front 0 in slots 0, 2, 1
            const int y = yfs - (l - 1) * dir;
            const float *f = lv.pf[l - 1];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const float *fr = f + row_at(y, 1 - r) * rowE;
#pragma unroll
                for (int i = 0; i < CX + 2; ++i)
                    F[l - 1][(3 - r) % 3][i] = ldb(fr + (ptrdiff_t)(i - 1) * Bp, o4);
            }
        }
#pragma unroll
        for (int l = 0; l < LV; ++l)
#pragma unroll
            for (int d = 0; d < LV; ++d) hist[l][d] = 0.0f;
    }
    // the lane's reduced cell (the same on every row)
    const int gx = x0 + sum_cell<LB, CX>(lane);
    const bool gval = (sb & (LB / CX - 1)) == 0 && gx < nx;
    float *const ghb = g_hi + (size_t)sg * N + gx;
    float *const glb = g_lo + (size_t)sg * N + gx;
    for (int t0 = 0; t0 < n_front; t0 += 3) {
#pragma unroll
        for (int ph = 0; ph < 3; ++ph) {
            const int t = t0 + ph;
            if (t >= n_front) break;  // uniform over the wavefront
            const int yf = yfs + t * dir;
            const int kn = ph, km = (ph + 2) % 3, ko = (ph + 1) % 3;  // newest, middle, oldest
            const int yg = yf - (LV - 1) * dir;  // the last level's row: accumulators updated there
            const bool gown = yg >= y0 && yg < y1;
            float gh0 = 0.0f, gl0 = 0.0f;
            if (gval && gown) {
                gh0 = ghb[(size_t)yg * nx];
                gl0 = glb[(size_t)yg * nx];
            }
            // ---- level 0's newest row (S' lam^{k+1} on row yf+dir), lam^{k+2} on row yf
            {
                const float2 *snr = sp2 + row_at(yf, 1) * nx;
#pragma unroll
                for (int g = 0; g < NC; ++g) {
                    R[0][kn][g] = ln[g];
                    Q[0][kn][g] = smul(ldb(snr + (g - LV), oc), ln[g]);
                }
#pragma unroll
                for (int g = 1; g < NC - 1; ++g) c2[g] = c2n[g];
            }
            // ---- loads for the next front: lam^{k+1} row yf+2dir, lam^{k+2} row yf+dir
            const bool more = t + 1 < n_front;
            if (more) {
                const float *a = l1 + row_at(yf, 2) * rowE;
                const float *c = l2 + row_at(yf, 1) * rowE;
#pragma unroll
                for (int g = 0; g < NC; ++g) ln[g] = ldb(a + (ptrdiff_t)(g - LV) * Bp, o4);
#pragma unroll
                for (int g = 1; g < NC - 1; ++g)
                    c2n[g] = ldn<USCT_NT2D>(c + (ptrdiff_t)(g - LV) * Bp, o4);
            }
            float s_cur[LV];
#pragma unroll
            for (int l = 1; l <= LV; ++l) {
                const int y = yf - (l - 1) * dir;
                const bool act = y >= y0 - (LV - l) && y < y1 + (LV - l);
                const bool own = y >= y0 && y < y1;
                const bool yint = y >= 1 && y <= ny - 2;
                const size_t yr = clamp_row(y, ny);
                const bool rec = act && y >= 0 && y < ny && row_rec[yr] > 0;
                const float2 *spr = sp2 + yr * nx;
                float *const out = lv.lam[l - 1];
                const float *const res = lv.res[l - 1];
                float uo[CX];
#pragma unroll
                for (int g = l; g < NC - l; ++g) {
                    const int x = x0 - LV + g;
                    float v = 0.0f;
                    if (act) {
                        // level l-1 on rows y+dir (newest), y, y-dir; level l-2 on row y
                        const float cur = R[l - 1][km][g];
                        const float prev = l == 1 ? c2[g] : R[l - 2][ko][g];
                        const bool in = yint && x >= 1 && x <= nx - 2;
                        const float yp = dir > 0 ? Q[l - 1][kn][g] : Q[l - 1][ko][g];
                        const float ym = dir > 0 ? Q[l - 1][ko][g] : Q[l - 1][kn][g];
                        const float lq =
                            in ? lap5(Q[l - 1][km][g], yp, ym, Q[l - 1][km][g + 1], Q[l - 1][km][g - 1])
                               : 0.0f;
                        v = leap(cur, prev, lq);
                        if (rec && x >= 0 && x < nx) {
                            for (int r = rec_head[(int)yr * nx + x]; r >= 0; r = rec_next[r])
                                v = v - res[(size_t)r * Bp + b];
                        }
                        if (own && out && g >= LV && g < LV + CX && x < nx)
                            stn<USCT_NT2D>(out + yr * rowE + (size_t)(g - LV) * Bp, o4, v);
                    }
                    if (g >= LV && g < LV + CX) uo[g - LV] = v;
                    if (l < LV) {
                        R[l][kn][g] = v;
                        Q[l][kn][g] = act ? smul(ldb(spr + (g - LV), oc), v) : 0.0f;
                    }
                }
                // imaging of level l on its row (own rows only)
                s_cur[l - 1] = 0.0f;
                if (own) {
                    float tv[CX];
#pragma unroll
                    for (int i = 0; i < CX; ++i) {
                        const int x = x0 + i;
                        tv[i] = 0.0f;
                        if (x < nx) {
                            const bool in = yint && x >= 1 && x <= nx - 2;
                            const float fp = dir > 0 ? F[l - 1][kn][i + 1] : F[l - 1][ko][i + 1];
                            const float fq = dir > 0 ? F[l - 1][ko][i + 1] : F[l - 1][kn][i + 1];
                            const float lp = in ? lap5(F[l - 1][km][i + 1], fp, fq, F[l - 1][km][i + 2],
                                                       F[l - 1][km][i])
                                                : 0.0f;
                            float d2 = smul(ldb(spr + i, oc), lp);
                            if ((int)yr * nx + x == my_src) d2 = d2 + inc[l - 1];
                            tv[i] = uo[i] * d2;
                        }
                    }
                    s_cur[l - 1] = reduce_scatter<LB, CX>(tv, lane);
                }
                // the forward level's row for the next front, while it is within one
                // row of the chunk (the oldest slot has been read)
                const int yn = y + 2 * dir;
                if (more && yn >= y0 - 1 && yn <= y1) {
                    const float *f = lv.pf[l - 1] + clamp_row(yn, ny) * rowE;
#pragma unroll
                    for (int i = 0; i < CX + 2; ++i) F[l - 1][ko][i] = ldb(f + (ptrdiff_t)(i - 1) * Bp, o4);
                }
            }
            // ---- accumulators of row yg: levels 1 .. LV in order (as LV single-level launches)
            if (gval && gown) {
                float hi = gh0, lo = gl0;
#pragma unroll
                for (int l = 1; l < LV; ++l) neumaier(hi, lo, hist[l - 1][LV - l - 1]);
                neumaier(hi, lo, s_cur[LV - 1]);
                ghb[(size_t)yg * nx] = hi;
                glb[(size_t)yg * nx] = lo;
            }
#pragma unroll
            for (int l = 1; l < LV; ++l) {
#pragma unroll
                for (int d = LV - l - 1; d >= 1; --d) hist[l - 1][d] = hist[l - 1][d - 1];
                hist[l - 1][0] = s_cur[l - 1];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Batched 3D forward step: each lane owns a CX x CY patch and marches RZ
// planes along z.  Row starts (y clamped, wave-uniform) and column steps go
// into the scalar base; a lane keeps one byte offset (addressing as k_fwd2d).
// The update is the FP64 one of k_fwd2d (seven-point Laplacian of the full
// hi + lo state); the own cells of the next plane are loaded one plane ahead.
// srcd2[b] records the increment applied at shot b's source cell for the
// adjoint's imaging.
// YSH (64-lane groups): the work-group's kYsh wavefronts take strips stacked
// along y and hand each other the edge rows of the next plane through LDS at
// each plane (one barrier per plane), as k_radj3d's YSH; only the group's outer
// y halo rows and the x halo are loaded a plane ahead.  Bit-identical to YSH =
// false (the same cells, the same operations).
// ---------------------------------------------------------------------------

constexpr int kYshF = kBlock / kWave;
template <int LB, int CX, int CY, bool REC, class L = uint16_t, bool YSH = false>
__global__ void __launch_bounds__(kBlock) USCT_ATTR3D
k_fwd3d(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pn,
        const L *__restrict__ pcl, const L *__restrict__ ppl,
        L *__restrict__ pnl, const float2 *__restrict__ sp2,
        const int *__restrict__ src_cell, const double *__restrict__ sincd,
        const int *__restrict__ rec_head, const int *__restrict__ rec_next,
        const int *__restrict__ row_rec, float *__restrict__ rec, float *__restrict__ rec_lo,
        float *__restrict__ srcd2, int nz, int ny, int nx, int Bp, int RZ, int nxs, int nys,
        int nzc, int ng, int blk) {
    constexpr int W = kWave / LB;
    static_assert(!YSH || LB == kWave, "YSH: one strip per wavefront");
    const int lane = threadIdx.x & 63;
    int sg, xs, ys, zc;
    const int wy = YSH ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    if constexpr (YSH) {
        int t = xcd_remap(blockIdx.x, gridDim.x);
        sg = t % ng;
        t /= ng;
        xs = t % nxs;
        t /= nxs;
        const int nyb = (nys + kYshF - 1) / kYshF;
        ys = (t % nyb) * kYshF + wy;
        zc = t / nyb;
        if (zc >= nzc) return;  // uniform over the work-group
    } else {
        const int wv = wave_id();
        sg = wv % ng;
        if (!strip3d(wv / ng, nxs, nys, nzc, blk, xs, ys, zc)) return;
    }
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    const int y0 = ys * CY;
    const int z0 = zc * RZ;
    const int z1 = min(z0 + RZ, nz);
    if (x0 >= nx) return;
    const int my_src = src_cell[b];
    const double my_inc = sincd[b];
    const size_t rowE = (size_t)nx * Bp;
    const size_t planeE = rowE * ny;
    const size_t planeC = (size_t)nx * ny;
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned o2 = (unsigned)(x0 * Bp + b) * (unsigned)sizeof(L);
    const unsigned oc = (unsigned)x0 * 8u;
    // row j of the patch (j = -1..CY), clamped into the grid; wave-uniform
    auto row = [&](int j) -> size_t { return (size_t)min(max(y0 + j, 0), ny - 1); };
    // full state hi + lo of cell (row j, column i) of a plane, in FP64
    auto P = [&](size_t plane, int j, int i) -> double {
        const size_t e = plane + row(j) * rowE + (size_t)((ptrdiff_t)i * Bp);
        return (double)ldb(pc + e, o4) + (double)lo_f(ldb(pcl + e, o2));
    };
    double zm[CY][CX], zcu[CY + 2][CX + 2], zp[CY][CX];
    // raw loads issued one plane ahead: own cells of plane z+1 (nh/nl) and of
    // level n-2 on plane z (qh/ql)
    float nh[CY][CX], qh[CY][CX];
    L nl[CY][CX], ql[CY][CX];
    auto load_halo = [&](size_t pl) {
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            zcu[0][i + 1] = P(pl, -1, i);
            zcu[CY + 1][i + 1] = P(pl, CY, i);
        }
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            zcu[j + 1][0] = P(pl, j, -1);
            zcu[j + 1][CX + 1] = P(pl, j, CX);
        }
    };
    // The halo of the next plane (the cells next to the patch, which neighbouring
    // wavefronts own) is also loaded a whole plane ahead, as raw hi / lo: a
    // plane of work covers its latency, and its lines are requested while the
    // owners still hold them in L2 (on the 256^3, 64-shot gradient the reads
    // fell from 19.0 to 15.9 bytes per shot-cell).  hh/hq[k] for k = the CX
    // cells of row -1, the CX of row CY, then rows 0..CY-1 at columns -1 and CX.
    constexpr int NH = 2 * (CX + CY);
    float hh[NH];
    L hq[NH];
    float yoh[YSH ? CX : 1];
    L yoq[YSH ? CX : 1];
    auto load_halo_raw = [&](size_t pl) {
        auto one = [&](int k, int j, int i) {
            const size_t e = pl + row(j) * rowE + (size_t)((ptrdiff_t)i * Bp);
            hh[k] = ldb(pc + e, o4);
            hq[k] = ldb(pcl + e, o2);
        };
        if constexpr (YSH) {
            // the group's outer y halo row (row -1 above the first wavefront, row
            // CY below the last); selected by value, so nothing is indexed by wy
            if (wy == 0 || wy == kYshF - 1) {
                const int jo = wy == 0 ? -1 : CY;
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    const size_t e = pl + row(jo) * rowE + (size_t)((ptrdiff_t)i * Bp);
                    yoh[i] = ldb(pc + e, o4);
                    yoq[i] = ldb(pcl + e, o2);
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                one(i, -1, i);
                one(CX + i, CY, i);
            }
        }
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            one(2 * CX + 2 * j, j, -1);
            one(2 * CX + 2 * j + 1, j, CX);
        }
    };
    // YSH: edge rows [plane parity][wavefront][top, bottom][cell][lane] of the
    // plane the patch moves to
    constexpr int kEdge = YSH ? 2 * kYshF * 2 * CX * 64 : 1;
    __shared__ double ez[kEdge];
    int par = 0;
    auto share_rows = [&]() {
        const int e0 = par * (kYshF * 2 * CX * 64) + lane;
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            ez[e0 + ((wy * 2 + 0) * CX + i) * 64] = zcu[1][i + 1];
            ez[e0 + ((wy * 2 + 1) * CX + i) * 64] = zcu[CY][i + 1];
        }
        __syncthreads();
        if (wy > 0)
#pragma unroll
            for (int i = 0; i < CX; ++i) zcu[0][i + 1] = ez[e0 + (((wy - 1) * 2 + 1) * CX + i) * 64];
        if (wy < kYshF - 1)
#pragma unroll
            for (int i = 0; i < CX; ++i) zcu[CY + 1][i + 1] = ez[e0 + (((wy + 1) * 2 + 0) * CX + i) * 64];
        par ^= 1;
    };
    auto halo_from_raw = [&]() {
        if constexpr (YSH) {
            if (wy == 0 || wy == kYshF - 1) {
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    const double v = hl(yoh[i], yoq[i]);
                    zcu[0][i + 1] = wy == 0 ? v : zcu[0][i + 1];
                    zcu[CY + 1][i + 1] = wy == 0 ? zcu[CY + 1][i + 1] : v;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                zcu[0][i + 1] = hl(hh[i], hq[i]);
                zcu[CY + 1][i + 1] = hl(hh[CX + i], hq[CX + i]);
            }
        }
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            zcu[j + 1][0] = hl(hh[2 * CX + 2 * j], hq[2 * CX + 2 * j]);
            zcu[j + 1][CX + 1] = hl(hh[2 * CX + 2 * j + 1], hq[2 * CX + 2 * j + 1]);
        }
        if constexpr (YSH) share_rows();
    };
    auto load_raw = [&](size_t pl, size_t ppl_, float (&h)[CY][CX], L (&l)[CY][CX],
                        float (&qh_)[CY][CX], L (&ql_)[CY][CX]) {
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const size_t e = pl + row(j) * rowE + (size_t)((ptrdiff_t)i * Bp);
                h[j][i] = ldb(pc + e, o4);
                l[j][i] = ldb(pcl + e, o2);
                const size_t eq = ppl_ + row(j) * rowE + (size_t)((ptrdiff_t)i * Bp);
                qh_[j][i] = ldn<USCT_NTF3D>(pp + eq, o4);
                ql_[j][i] = ldn<USCT_NTF3D>(ppl + eq, o2);
            }
    };
    {
        const size_t pm = (size_t)max(z0 - 1, 0) * planeE;
        const size_t p0 = (size_t)z0 * planeE;
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                zm[j][i] = P(pm, j, i);
                zcu[j + 1][i + 1] = P(p0, j, i);
            }
        load_halo(p0);
        if constexpr (YSH) share_rows();  // the loaded inner y halo rows are overwritten
        load_raw((size_t)min(z0 + 1, nz - 1) * planeE, p0, nh, nl, qh, ql);
        if (z0 + 1 < z1) load_halo_raw((size_t)(z0 + 1) * planeE);
    }
    for (int z = z0; z < z1; ++z) {
        const size_t pnz = (size_t)min(z + 1, nz - 1) * planeE;
        double q[CY][CX];
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                zp[j][i] = hl(nh[j][i], nl[j][i]);
                q[j][i] = hl(qh[j][i], ql[j][i]);
            }
        if (z + 1 < z1) load_raw((size_t)min(z + 2, nz - 1) * planeE, pnz, nh, nl, qh, ql);
        const bool zint = (z >= 1) && (z <= nz - 2);
        // The arithmetic of every cell of the patch first, with no branch between
        // cells, so their FP64 chains interleave; then the stores, source increment
        // and traces of the cells inside the grid.  Cells outside the grid (strips
        // rounded up) compute on clamped rows and are not stored.
        float vv[CY][CX];
        double incv[CY][CX];
        L lbv[CY][CX];
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            const int y = y0 + j;
            const bool yint = zint && (y >= 1) && (y <= ny - 2);
            const float2 *spr = sp2 + (size_t)z * planeC + row(j) * nx;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                const double lu = (yint && x >= 1 && x <= nx - 2)
                                      ? lap7d(zcu[j + 1][i + 1], zp[j][i], zm[j][i], zcu[j + 2][i + 1],
                                              zcu[j][i + 1], zcu[j + 1][i + 2], zcu[j + 1][i])
                                      : 0.0;
                double inc = s2d(ldb(spr + i, oc)) * lu;
                if ((z * ny + y) * nx + x == my_src) inc = inc + my_inc;
                incv[j][i] = inc;
                const double nv = (2.0 * zcu[j + 1][i + 1] - q[j][i]) + inc;
                vv[j][i] = split_hl(nv, lbv[j][i]);
            }
        }
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            const int y = y0 + j;
            const bool has_rec = REC && y < ny && row_rec[z * ny + y] > 0;
            const size_t ro = (size_t)z * planeE + (size_t)y * rowE;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                if (x < nx && y < ny) {
                    const int c = (z * ny + y) * nx + x;
                    const size_t ei = ro + (size_t)i * Bp;
                    const float v = vv[j][i];
                    stn<USCT_NTF3D>(pn + ei, o4, v);
                    stn<USCT_NTF3D>(pnl + ei, o2, lbv[j][i]);
                    if (c == my_src) srcd2[b] = (float)incv[j][i];
                    if (has_rec) {
                        for (int r = rec_head[c]; r >= 0; r = rec_next[r]) {
                            rec[(size_t)r * Bp + b] = v;
                            rec_lo[(size_t)r * Bp + b] = lo_f(lbv[j][i]);
                        }
                    }
                }
            }
        }
        if (z + 1 < z1) {
#pragma unroll
            for (int j = 0; j < CY; ++j)
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    zm[j][i] = zcu[j + 1][i + 1];
                    zcu[j + 1][i + 1] = zp[j][i];
                }
            halo_from_raw();  // plane z+1, loaded one iteration ago
            if (z + 2 < z1) load_halo_raw((size_t)(z + 2) * planeE);
        }
    }
}

// ---------------------------------------------------------------------------
// Batched 3D forward, two levels per launch (temporal blocking along z, no
// sponge): level n (A) from levels n-1 (P, pc) and n-2 (pp), then level n+1
// (B) from A and P, both with k_fwd3d's FP64 update.  Each lane owns a CX x CY
// patch of cells of LB shots and marches planes along z with B one plane
// behind A.  A is formed on the patch and the cells next to it in x and y (the
// "A region": what B's stencil needs on B's own plane), P is held on the A
// region and, on the front plane, on the cells next to the A region (the "P
// region"); B's planes above and below need A at the patch only.  Halo cells
// of A are computed with the same operations on the same operands as their
// owners (source term included), and B reads A as its stored hi + lo, so both
// levels, their traces and the recorded source increments are bit-identical to
// two k_fwd3d launches.  Per shot-cell the pair reads 12 bytes and writes 12
// instead of reading 24 and writing 12; the x and y halo reads of P and of
// level n-2 come from neighbouring lanes' cells, through L1 and L2.
// Everything stays in registers: no LDS, no barriers.  The next plane's P
// region and level n-2 on the A region are loaded one plane ahead.
//   sincd_a / sincd_b: S_src w[n-1], S_src w[n]; rec_a / rec_b: trace rows of
//   levels n and n+1; srcd2_a / srcd2_b: increments at each shot's source cell.
// Rows are clamped to the grid, columns read up to two cells beyond the strip
// (the guard zones cover them), planes are clamped.
// ---------------------------------------------------------------------------
template <int CX, int CY>
struct Reg3 {
    // cells relative to the patch origin: the patch (indices 0 .. NO-1), then
    // the rest of the A region (plus-shaped ring), then the rest of the P region
    // (corner cells of the ring and the cells two steps out along x and y)
    static constexpr int NO = CX * CY;
    static constexpr int NA = NO + 2 * (CX + CY);
    static constexpr int NP = NA + 4 + 2 * (CX + CY);
    int dx[NP], dy[NP];
    int nb[NA][4];   // P-region indices of an A cell's neighbours: y+1, y-1, x+1, x-1
    int anb[NO][4];  // A-region indices of a patch cell's neighbours, the same order
    constexpr Reg3() : dx(), dy(), nb(), anb() {
        int k = 0;
        for (int j = 0; j < CY; ++j)
            for (int i = 0; i < CX; ++i) { dx[k] = i; dy[k] = j; ++k; }
        for (int i = 0; i < CX; ++i) { dx[k] = i; dy[k] = -1; ++k; dx[k] = i; dy[k] = CY; ++k; }
        for (int j = 0; j < CY; ++j) { dx[k] = -1; dy[k] = j; ++k; dx[k] = CX; dy[k] = j; ++k; }
        dx[k] = -1; dy[k] = -1; ++k;
        dx[k] = CX; dy[k] = -1; ++k;
        dx[k] = -1; dy[k] = CY; ++k;
        dx[k] = CX; dy[k] = CY; ++k;
        for (int i = 0; i < CX; ++i) { dx[k] = i; dy[k] = -2; ++k; dx[k] = i; dy[k] = CY + 1; ++k; }
        for (int j = 0; j < CY; ++j) { dx[k] = -2; dy[k] = j; ++k; dx[k] = CX + 1; dy[k] = j; ++k; }
        for (int a = 0; a < NA; ++a) {
            const int ox[4] = {0, 0, 1, -1}, oy[4] = {1, -1, 0, 0};
            for (int d = 0; d < 4; ++d) {
                nb[a][d] = -1;
                for (int q = 0; q < NP; ++q)
                    if (dx[q] == dx[a] + ox[d] && dy[q] == dy[a] + oy[d]) nb[a][d] = q;
                if (a < NO) anb[a][d] = nb[a][d];
            }
        }
    }
};

template <int LB, int CX, int CY, bool REC>
__global__ void __launch_bounds__(kBlock) USCT_ATTR3D
k_fwd3d_x2(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pa,
           float *__restrict__ pb, const uint16_t *__restrict__ pcl,
           const uint16_t *__restrict__ ppl, uint16_t *__restrict__ pal,
           uint16_t *__restrict__ pbl, const float2 *__restrict__ sp2,
           const int *__restrict__ src_cell, const double *__restrict__ sincd_a,
           const double *__restrict__ sincd_b, const int *__restrict__ rec_head,
           const int *__restrict__ rec_next, const int *__restrict__ row_rec,
           float *__restrict__ rec_a, float *__restrict__ rec_lo_a, float *__restrict__ rec_b,
           float *__restrict__ rec_lo_b, float *__restrict__ srcd2_a, float *__restrict__ srcd2_b,
           int nz, int ny, int nx, int Bp, int RZ, int nxs, int nys, int nzc, int ng, int blk) {
    using R = Reg3<CX, CY>;
    constexpr R T{};
    constexpr int NO = R::NO, NA = R::NA, NP = R::NP;
    constexpr int W = kWave / LB;
    const int wv = wave_id();
    const int lane = threadIdx.x & 63;
    const int sg = wv % ng;
    int xs, ys, zc;
    if (!strip3d(wv / ng, nxs, nys, nzc, blk, xs, ys, zc)) return;
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    const int y0 = ys * CY;
    const int z0 = zc * RZ;
    const int z1 = min(z0 + RZ, nz);
    if (x0 >= nx) return;
    const int my_src = src_cell[b];
    const double inc_a = sincd_a[b], inc_b = sincd_b[b];
    const size_t rowE = (size_t)nx * Bp;
    const size_t planeE = rowE * ny;
    const size_t planeC = (size_t)nx * ny;
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned o2 = (unsigned)(x0 * Bp + b) * 2u;
    const unsigned oc = (unsigned)x0 * 8u;
    auto rowc = [&](int q) -> size_t { return (size_t)min(max(y0 + T.dy[q], 0), ny - 1); };
    auto plane = [&](int z) -> size_t { return (size_t)min(max(z, 0), nz - 1) * planeE; };
    auto el = [&](size_t pl, int q) -> size_t {
        return pl + rowc(q) * rowE + (size_t)((ptrdiff_t)T.dx[q] * Bp);
    };
    // P (level n-1) as FP64: um = front plane - 1 (A region), uc = front plane
    // (P region), un = front plane + 1 (A region).  Raw loads run ahead: the A
    // region of plane front+1 (nh/nl, converted into un at the front) and the
    // rest of the P region of plane front (rh/rl, completing uc), both issued
    // one front earlier; level n-2 on the A region of the front plane (qh/ql).
    constexpr int NR = NP - NA;
    double um[NA], uc[NP], un[NA];
    float nh[NA], rh[NR], qh[NA];
    uint16_t nl[NA], rl[NR], ql[NA];
    // A (level n) as stored: planes front-2 (patch), front-1 and front (A region)
    double am[NO], ac[NA], an[NA];
    {
        const size_t pm = plane(z0 - 2), p0 = plane(z0 - 1), p1 = plane(z0);
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            um[q] = hl(ldb(pc + el(pm, q), o4), ldb(pcl + el(pm, q), o2));
            nh[q] = ldb(pc + el(p1, q), o4);
            nl[q] = ldb(pcl + el(p1, q), o2);
            qh[q] = ldn<USCT_NTF3D>(pp + el(p0, q), o4);
            ql[q] = ldn<USCT_NTF3D>(ppl + el(p0, q), o2);
        }
#pragma unroll
        for (int q = 0; q < NA; ++q) uc[q] = hl(ldb(pc + el(p0, q), o4), ldb(pcl + el(p0, q), o2));
#pragma unroll
        for (int q = NA; q < NP; ++q) {
            rh[q - NA] = ldb(pc + el(p0, q), o4);
            rl[q - NA] = ldb(pcl + el(p0, q), o2);
        }
#pragma unroll
        for (int q = 0; q < NA; ++q) { ac[q] = 0.0; an[q] = 0.0; }
#pragma unroll
        for (int q = 0; q < NO; ++q) am[q] = 0.0;
    }
    for (int zf = z0 - 1; zf <= z1; ++zf) {
#pragma unroll
        for (int q = 0; q < NA; ++q) un[q] = hl(nh[q], nl[q]);
#pragma unroll
        for (int q = NA; q < NP; ++q) uc[q] = hl(rh[q - NA], rl[q - NA]);
        double qq[NA];
#pragma unroll
        for (int q = 0; q < NA; ++q) qq[q] = hl(qh[q], ql[q]);
        // ---- loads for the next front: the A region of plane zf+2, the rest of
        // plane zf+1's P region, level n-2 on plane zf+1
        if (zf < z1) {
            const size_t pn2 = plane(zf + 2), pn1 = plane(zf + 1);
#pragma unroll
            for (int q = 0; q < NA; ++q) {
                nh[q] = ldb(pc + el(pn2, q), o4);
                nl[q] = ldb(pcl + el(pn2, q), o2);
                qh[q] = ldn<USCT_NTF3D>(pp + el(pn1, q), o4);
                ql[q] = ldn<USCT_NTF3D>(ppl + el(pn1, q), o2);
            }
#pragma unroll
            for (int q = NA; q < NP; ++q) {
                rh[q - NA] = ldb(pc + el(pn1, q), o4);
                rl[q - NA] = ldb(pcl + el(pn1, q), o2);
            }
        }
        // ---- level n on plane zf, A region
        const bool own_a = zf >= z0 && zf < z1;
        const bool zint = zf >= 1 && zf <= nz - 2;
        const int zr = min(max(zf, 0), nz - 1);
        double anew[NA];
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int x = x0 + T.dx[q], y = y0 + T.dy[q];
            const int yr = min(max(y, 0), ny - 1);
            double lu = 0.0;
            if (zint && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2)
                lu = lap7d(uc[q], un[q], um[q], uc[T.nb[q][0]], uc[T.nb[q][1]], uc[T.nb[q][2]],
                           uc[T.nb[q][3]]);
            double inc = s2d(ldb(sp2 + (size_t)zr * planeC + (size_t)yr * nx + T.dx[q], oc)) * lu;
            const int c = (zf * ny + y) * nx + x;
            if (c == my_src) inc = inc + inc_a;
            const double nv = (2.0 * uc[q] - qq[q]) + inc;
            uint16_t lb;
            const float v = split_hl(nv, lb);
            anew[q] = hl(v, lb);
            if (q < NO && own_a && x < nx && y < ny) {
                const size_t e = (size_t)zf * planeE + (size_t)y * rowE + (size_t)T.dx[q] * Bp;
                stn<USCT_NTF3D>(pa + e, o4, v);
                stn<USCT_NTF3D>(pal + e, o2, lb);
                if (c == my_src) srcd2_a[b] = (float)inc;
                if (REC && row_rec[zf * ny + y] > 0) {
                    for (int r = rec_head[c]; r >= 0; r = rec_next[r]) {
                        rec_a[(size_t)r * Bp + b] = v;
                        rec_lo_a[(size_t)r * Bp + b] = bf2f(lb);
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NO; ++q) am[q] = ac[q];
#pragma unroll
        for (int q = 0; q < NA; ++q) { ac[q] = an[q]; an[q] = anew[q]; }
        // ---- level n+1 on plane z = zf-1, the patch
        const int z = zf - 1;
        if (z >= z0 && z < z1) {
            const bool zint_b = z >= 1 && z <= nz - 2;
#pragma unroll
            for (int q = 0; q < NO; ++q) {
                const int x = x0 + T.dx[q], y = y0 + T.dy[q];
                if (x < nx && y < ny) {
                    double lu = 0.0;
                    if (zint_b && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2)
                        lu = lap7d(ac[q], an[q], am[q], ac[T.anb[q][0]], ac[T.anb[q][1]],
                                   ac[T.anb[q][2]], ac[T.anb[q][3]]);
                    double inc = s2d(ldb(sp2 + (size_t)z * planeC + (size_t)y * nx + T.dx[q], oc)) * lu;
                    const int c = (z * ny + y) * nx + x;
                    if (c == my_src) inc = inc + inc_b;
                    const double nv = (2.0 * ac[q] - um[q]) + inc;
                    uint16_t lb;
                    const float v = split_hl(nv, lb);
                    const size_t e = (size_t)z * planeE + (size_t)y * rowE + (size_t)T.dx[q] * Bp;
                    stn<USCT_NTF3D>(pb + e, o4, v);
                    stn<USCT_NTF3D>(pbl + e, o2, lb);
                    if (c == my_src) srcd2_b[b] = (float)inc;
                    if (REC && row_rec[z * ny + y] > 0) {
                        for (int r = rec_head[c]; r >= 0; r = rec_next[r]) {
                            rec_b[(size_t)r * Bp + b] = v;
                            rec_lo_b[(size_t)r * Bp + b] = bf2f(lb);
                        }
                    }
                }
            }
        }
        // ---- advance P (the rest of the new front plane's P region arrives in rh/rl)
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            um[q] = uc[q];
            uc[q] = un[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Gauss-Newton, 3D without sponge or footprints: background level n (the
// k_fwd3d update, without trace recording) and the Born level n that uses it
// (the k_born update) in one launch, on k_fwd3d's patch march with a second
// field.  The Born source reads the background's second difference
// U^n - 2U^{n-1} + U^{n-2} from the registers that hold the stored levels, so
// both levels, srcd2 and the Born traces (row n of the adjoint source res) are
// bit-identical to k_fwd3d<REC=false> followed by k_born<false>.  Per
// shot-cell it reads 24 bytes and writes 12 instead of reading 18 + 30 and
// writing 6 + 6.
//   pc/pp/pn, dc/dp/dn: levels n-1, n-2 (read) and n (written) of the
//   background and the Born field; sincd/srcd2: row n.
// ---------------------------------------------------------------------------
template <int LB, int CX, int CY>
__global__ void __launch_bounds__(kBlock) USCT_ATTR3D
k_born3d(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pn,
         const uint16_t *__restrict__ pcl, const uint16_t *__restrict__ ppl,
         uint16_t *__restrict__ pnl, const float *__restrict__ dc, const float *__restrict__ dp,
         float *__restrict__ dn, const uint16_t *__restrict__ dcl, const uint16_t *__restrict__ dpl,
         uint16_t *__restrict__ dnl, const float2 *__restrict__ sp2, const float2 *__restrict__ vm2,
         const int *__restrict__ src_cell, const double *__restrict__ sincd,
         const int *__restrict__ rec_head, const int *__restrict__ rec_next,
         const int *__restrict__ row_rec, float *__restrict__ res, float *__restrict__ srcd2,
         int nz, int ny, int nx, int Bp, int RZ, int nxs, int nys, int nzc, int ng, int blk) {
    constexpr int W = kWave / LB;
    const int wv = wave_id();
    const int lane = threadIdx.x & 63;
    const int sg = wv % ng;
    int xs, ys, zc;
    if (!strip3d(wv / ng, nxs, nys, nzc, blk, xs, ys, zc)) return;
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    const int y0 = ys * CY;
    const int z0 = zc * RZ;
    const int z1 = min(z0 + RZ, nz);
    if (x0 >= nx) return;
    const int my_src = src_cell[b];
    const double my_inc = sincd[b];
    const size_t rowE = (size_t)nx * Bp;
    const size_t planeE = rowE * ny;
    const size_t planeC = (size_t)nx * ny;
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned o2 = (unsigned)(x0 * Bp + b) * 2u;
    const unsigned oc = (unsigned)x0 * 8u;
    auto row = [&](int j) -> size_t { return (size_t)min(max(y0 + j, 0), ny - 1); };
    auto at = [&](size_t plane, int j, int i) -> size_t {
        return plane + row(j) * rowE + (size_t)((ptrdiff_t)i * Bp);
    };
    // full state hi + lo of the background (P) and the Born field (D), in FP64
    auto P = [&](size_t plane, int j, int i) -> double {
        const size_t e = at(plane, j, i);
        return (double)ldb(pc + e, o4) + (double)bf2f(ldb(pcl + e, o2));
    };
    auto D = [&](size_t plane, int j, int i) -> double {
        const size_t e = at(plane, j, i);
        return (double)ldb(dc + e, o4) + (double)bf2f(ldb(dcl + e, o2));
    };
    double zm[CY][CX], zcu[CY + 2][CX + 2], zp[CY][CX];
    double em[CY][CX], ecu[CY + 2][CX + 2], ep[CY][CX];
    auto load_halo = [&](size_t pl) {
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            zcu[0][i + 1] = P(pl, -1, i);
            zcu[CY + 1][i + 1] = P(pl, CY, i);
            ecu[0][i + 1] = D(pl, -1, i);
            ecu[CY + 1][i + 1] = D(pl, CY, i);
        }
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            zcu[j + 1][0] = P(pl, j, -1);
            zcu[j + 1][CX + 1] = P(pl, j, CX);
            ecu[j + 1][0] = D(pl, j, -1);
            ecu[j + 1][CX + 1] = D(pl, j, CX);
        }
    };
    // raw loads issued one plane ahead: own cells of level n-1 on plane z+1 and
    // of level n-2 on plane z, for both fields
    float nh[CY][CX], qh[CY][CX], mh[CY][CX], rh[CY][CX];
    uint16_t nl[CY][CX], ql[CY][CX], ml[CY][CX], rl[CY][CX];
    auto load_raw = [&](size_t pl, size_t pl2) {
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const size_t e = at(pl, j, i), e2 = at(pl2, j, i);
                nh[j][i] = ldb(pc + e, o4);
                nl[j][i] = ldb(pcl + e, o2);
                mh[j][i] = ldb(dc + e, o4);
                ml[j][i] = ldb(dcl + e, o2);
                qh[j][i] = ldn<USCT_NTF3D>(pp + e2, o4);
                ql[j][i] = ldn<USCT_NTF3D>(ppl + e2, o2);
                rh[j][i] = ldn<USCT_NTF3D>(dp + e2, o4);
                rl[j][i] = ldn<USCT_NTF3D>(dpl + e2, o2);
            }
    };
    {
        const size_t pm = (size_t)max(z0 - 1, 0) * planeE;
        const size_t p0 = (size_t)z0 * planeE;
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                zm[j][i] = P(pm, j, i);
                zcu[j + 1][i + 1] = P(p0, j, i);
                em[j][i] = D(pm, j, i);
                ecu[j + 1][i + 1] = D(p0, j, i);
            }
        load_halo(p0);
        load_raw((size_t)min(z0 + 1, nz - 1) * planeE, p0);
    }
    for (int z = z0; z < z1; ++z) {
        const size_t pnz = (size_t)min(z + 1, nz - 1) * planeE;
        double q[CY][CX], r[CY][CX];
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                zp[j][i] = hl(nh[j][i], nl[j][i]);
                ep[j][i] = hl(mh[j][i], ml[j][i]);
                q[j][i] = hl(qh[j][i], ql[j][i]);
                r[j][i] = hl(rh[j][i], rl[j][i]);
            }
        if (z + 1 < z1) load_raw((size_t)min(z + 2, nz - 1) * planeE, pnz);
        const bool zint = (z >= 1) && (z <= nz - 2);
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            const int y = y0 + j;
            const bool yint = zint && (y >= 1) && (y <= ny - 2);
            const bool has_rec = y < ny && row_rec[z * ny + y] > 0;
            const size_t ro = (size_t)z * planeE + (size_t)y * rowE;
            const float2 *spr = sp2 + (size_t)z * planeC + (size_t)y * nx;
            const float2 *vmr = vm2 + (size_t)z * planeC + (size_t)y * nx;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                if (x < nx && y < ny) {
                    const bool in = yint && x >= 1 && x <= nx - 2;
                    const double s = s2d(ldb(spr + i, oc));
                    // background level n (k_fwd3d)
                    double lu = 0.0;
                    if (in)
                        lu = lap7d(zcu[j + 1][i + 1], zp[j][i], zm[j][i], zcu[j + 2][i + 1],
                                   zcu[j][i + 1], zcu[j + 1][i + 2], zcu[j + 1][i]);
                    double inc = s * lu;
                    const int c = (z * ny + y) * nx + x;
                    if (c == my_src) inc = inc + my_inc;
                    const size_t ei = ro + (size_t)i * Bp;
                    const double nv = (2.0 * zcu[j + 1][i + 1] - q[j][i]) + inc;
                    uint16_t lb;
                    const float v = split_hl(nv, lb);
                    stn<USCT_NTF3D>(pn + ei, o4, v);
                    stn<USCT_NTF3D>(pnl + ei, o2, lb);
                    if (c == my_src) srcd2[b] = (float)inc;
                    // Born level n (k_born): the stored background levels' second difference
                    double ld = 0.0;
                    if (in)
                        ld = lap7d(ecu[j + 1][i + 1], ep[j][i], em[j][i], ecu[j + 2][i + 1],
                                   ecu[j][i + 1], ecu[j + 1][i + 2], ecu[j + 1][i]);
                    const double d2u = (hl(v, lb) - 2.0 * zcu[j + 1][i + 1]) + q[j][i];
                    const double bv = (2.0 * ecu[j + 1][i + 1] - r[j][i]) + s * ld -
                                      s2d(ldb(vmr + i, oc)) * d2u;
                    uint16_t bl;
                    const float w = split_hl(bv, bl);
                    stn<USCT_NTF3D>(dn + ei, o4, w);
                    stn<USCT_NTF3D>(dnl + ei, o2, bl);
                    if (has_rec) {
                        for (int rj = rec_head[c]; rj >= 0; rj = rec_next[rj])
                            res[(size_t)rj * Bp + b] = w;
                    }
                }
            }
        }
        if (z + 1 < z1) {
#pragma unroll
            for (int j = 0; j < CY; ++j)
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    zm[j][i] = zcu[j + 1][i + 1];
                    zcu[j + 1][i + 1] = zp[j][i];
                    em[j][i] = ecu[j + 1][i + 1];
                    ecu[j + 1][i + 1] = ep[j][i];
                }
            load_halo(pnz);
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-tiled 3D kernels.  A work-group of 256 lanes owns a TX x TY tile of cells
// (for LB shots) and marches RZ planes along z; lane = (shot, patch), each lane
// a PX x PY patch of the tile.  The centre plane of the tile, with a one-cell
// halo ring (no corners), is staged in LDS (double-buffered: plane z+1 is
// written while plane z is read; one barrier per plane), so a lane keeps only
// its own cells of the three planes in registers and the tile's halo is read
// from HBM/L2 once per plane, not once per lane.  The own cells and the ring of
// the next plane are loaded one plane ahead.
// ---------------------------------------------------------------------------
template <int TX, int TY>
struct Ring {  // halo cells of a TX x TY tile, without corners: r -> (hx, hy) in the padded tile
    static constexpr int N = 2 * TX + 2 * TY;
    static __device__ __forceinline__ void at(int r, int &hx, int &hy) {
        if (r < TX) { hx = r + 1; hy = 0; }
        else if (r < 2 * TX) { hx = r - TX + 1; hy = TY + 1; }
        else if (r < 2 * TX + TY) { hx = 0; hy = r - 2 * TX + 1; }
        else { hx = TX + 1; hy = r - 2 * TX - TY + 1; }
    }
};

// 3D forward level (FP64 update of k_fwd2d / k_fwd3d) with an LDS-staged tile.
// Grid: ng * ntx * nty * nzc work-groups of 256 lanes.
template <int LB, int TX, int TY, int PX, int PY, bool REC>
__global__ void __launch_bounds__(kBlock) USCT_ATTR3D
k_fwd3d_lds(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pn,
            const uint16_t *__restrict__ pcl, const uint16_t *__restrict__ ppl,
            uint16_t *__restrict__ pnl, const float2 *__restrict__ sp2,
            const int *__restrict__ src_cell, const double *__restrict__ sincd,
            const int *__restrict__ rec_head, const int *__restrict__ rec_next,
            const int *__restrict__ row_rec, float *__restrict__ rec, float *__restrict__ rec_lo,
            float *__restrict__ srcd2, int nz, int ny, int nx, int Bp, int RZ, int ntx, int nty,
            int nzc, int ng) {
    constexpr int TCX = TX / PX, TCY = TY / PY;
    static_assert(TCX * PX == TX && TCY * PY == TY && TCX * TCY * LB == kBlock, "tile shape");
    constexpr int HX = TX + 2, HY = TY + 2;
    constexpr int RN = Ring<TX, TY>::N * LB;            // ring entries (cell, shot)
    constexpr int RPL = (RN + kBlock - 1) / kBlock;     // ring entries per lane
    __shared__ double sh[2][HY * HX * LB];
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int sg = blk % ng;
    int t = blk / ng;
    const int txi = t % ntx;
    t /= ntx;
    const int tyi = t % nty, zc = t / nty;
    if (zc >= nzc) return;  // uniform over the work-group
    const int tid = threadIdx.x;
    const int bl = tid % LB, col = tid / LB;
    const int cx = col % TCX, cy = col / TCX;
    const int b = sg * LB + bl;
    const int X0 = txi * TX, Y0 = tyi * TY;
    const int x0 = X0 + cx * PX, y0 = Y0 + cy * PY;
    const int z0 = zc * RZ, z1 = min(z0 + RZ, nz);
    const int my_src = src_cell[b];
    const double my_inc = sincd[b];
    const size_t planeE = (size_t)nx * ny * Bp;
    // Per-lane 32-bit byte offsets from the plane base shifted back by one cell
    // (so the halo column x = -1 is offset 0): own cells (y clamped) and ring entries.
    auto off = [&](int y, int x, int shot) -> unsigned {
        return (unsigned)(((min(max(y, 0), ny - 1) * nx + x + 1) * Bp + shot));
    };
    unsigned oo[PY][PX], ro[RPL];
    int r_hx[RPL], r_hy[RPL];
    bool r_on[RPL];
#pragma unroll
    for (int jj = 0; jj < PY; ++jj)
#pragma unroll
        for (int ii = 0; ii < PX; ++ii) oo[jj][ii] = off(y0 + jj, x0 + ii, b);
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        const int e = tid + q * kBlock;
        r_on[q] = e < RN;
        Ring<TX, TY>::at(r_on[q] ? e / LB : 0, r_hx[q], r_hy[q]);
        ro[q] = off(Y0 + r_hy[q] - 1, X0 + r_hx[q] - 1, sg * LB + e % LB);
    }
    auto sidx = [&](int hy, int hx, int shot) { return (hy * HX + hx) * LB + shot; };
    // plane bases (wave-uniform), one cell before the plane
    auto F4 = [&](const float *f, int z) { return f + (ptrdiff_t)z * (ptrdiff_t)planeE - Bp; };
    auto F2 = [&](const uint16_t *f, int z) { return f + (ptrdiff_t)z * (ptrdiff_t)planeE - Bp; };
    auto W4 = [&](float *f, int z) { return f + (ptrdiff_t)z * (ptrdiff_t)planeE - Bp; };
    auto W2 = [&](uint16_t *f, int z) { return f + (ptrdiff_t)z * (ptrdiff_t)planeE - Bp; };
    double zm[PY][PX], zcn[PY][PX], zp[PY][PX];
    float nh[PY][PX], qh[PY][PX], rh[RPL];
    uint16_t nl[PY][PX], ql[PY][PX], rl[RPL];
    // raw loads of plane zn (own cells, ring) and of level n-2 on plane zq (own cells)
    auto load_raw = [&](int zn, int zq) {
        const float *a = F4(pc, zn), *aq = F4(pp, zq);
        const uint16_t *al = F2(pcl, zn), *aql = F2(ppl, zq);
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                nh[jj][ii] = ldb(a, oo[jj][ii] * 4u);
                nl[jj][ii] = ldb(al, oo[jj][ii] * 2u);
                qh[jj][ii] = ldb(aq, oo[jj][ii] * 4u);
                ql[jj][ii] = ldb(aql, oo[jj][ii] * 2u);
            }
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            if (!r_on[q]) continue;
            rh[q] = ldb(a, ro[q] * 4u);
            rl[q] = ldb(al, ro[q] * 2u);
        }
    };
    // put plane (own cells, ring from rh/rl) into buffer s
    auto stage = [&](int s, const double (&own)[PY][PX]) {
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii)
                sh[s][sidx(cy * PY + jj + 1, cx * PX + ii + 1, bl)] = own[jj][ii];
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            if (!r_on[q]) continue;
            const int e = tid + q * kBlock;
            sh[s][sidx(r_hy[q], r_hx[q], e % LB)] = hl(rh[q], rl[q]);
        }
    };
    {
        const float *am = F4(pc, max(z0 - 1, 0));
        const uint16_t *aml = F2(pcl, max(z0 - 1, 0));
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii)
                zm[jj][ii] = hl(ldb(am, oo[jj][ii] * 4u), ldb(aml, oo[jj][ii] * 2u));
        load_raw(z0, z0);  // plane z0 (staged now) and level n-2 on plane z0
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) zcn[jj][ii] = hl(nh[jj][ii], nl[jj][ii]);
        stage(0, zcn);
        load_raw(min(z0 + 1, nz - 1), z0);
        __syncthreads();
    }
    for (int z = z0; z < z1; ++z) {
        const int cur = (z - z0) & 1;
        double q[PY][PX];
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                zp[jj][ii] = hl(nh[jj][ii], nl[jj][ii]);
                q[jj][ii] = hl(qh[jj][ii], ql[jj][ii]);
            }
        if (z + 1 < z1) {
            stage(cur ^ 1, zp);
            load_raw(min(z + 2, nz - 1), z + 1);
        }
        const bool zint = (z >= 1) && (z <= nz - 2);
        float *on = W4(pn, z);
        uint16_t *onl = W2(pnl, z);
        const float2 *spz = sp2 + (size_t)z * nx * ny;
#pragma unroll
        for (int jj = 0; jj < PY; ++jj) {
            const int y = y0 + jj;
            const bool yint = zint && (y >= 1) && (y <= ny - 2);
            const bool has_rec = REC && y < ny && row_rec[z * ny + y] > 0;
            const int ly = cy * PY + jj + 1;
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                const int x = x0 + ii;
                const int lx = cx * PX + ii + 1;
                if (x < nx && y < ny) {
                    double lu = 0.0;
                    if (yint && x >= 1 && x <= nx - 2) {
                        const double ym = jj > 0 ? zcn[jj - 1][ii] : sh[cur][sidx(ly - 1, lx, bl)];
                        const double yp = jj < PY - 1 ? zcn[jj + 1][ii] : sh[cur][sidx(ly + 1, lx, bl)];
                        const double xm = ii > 0 ? zcn[jj][ii - 1] : sh[cur][sidx(ly, lx - 1, bl)];
                        const double xp = ii < PX - 1 ? zcn[jj][ii + 1] : sh[cur][sidx(ly, lx + 1, bl)];
                        lu = lap7d(zcn[jj][ii], zp[jj][ii], zm[jj][ii], yp, ym, xp, xm);
                    }
                    const int c = (z * ny + y) * nx + x;
                    double inc = s2d(spz[y * nx + x]) * lu;
                    if (c == my_src) inc = inc + my_inc;
                    const double nv = (2.0 * zcn[jj][ii] - q[jj][ii]) + inc;
                    uint16_t lb;
                    const float v = split_hl(nv, lb);
                    stb(on, oo[jj][ii] * 4u, v);
                    stb(onl, oo[jj][ii] * 2u, lb);
                    if (c == my_src) srcd2[b] = (float)inc;
                    if (has_rec) {
                        for (int r = rec_head[c]; r >= 0; r = rec_next[r]) {
                            rec[(size_t)r * Bp + b] = v;
                            rec_lo[(size_t)r * Bp + b] = bf2f(lb);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                zm[jj][ii] = zcn[jj][ii];
                zcn[jj][ii] = zp[jj][ii];
            }
        __syncthreads();
    }
}

// Cells of a (TX+4) x (TY+4) square around a TX x TY tile (padded coordinates,
// tile at 2..TX+1), outside the tile: two rows above, four per tile row, two below.
template <int TX, int TY>
struct Halo2 {
    static constexpr int GX = TX + 4;
    static constexpr int N = 4 * GX + 4 * TY;
    static __device__ __forceinline__ void at(int r, int &gx, int &gy) {
        if (r < 2 * GX) {
            gy = r / GX;
            gx = r % GX;
        } else if (r < 2 * GX + 4 * TY) {
            const int q = r - 2 * GX, k = q % 4;
            gy = 2 + q / 4;
            gx = k < 2 ? k : TX + k;
        } else {
            const int q = r - 2 * GX - 4 * TY;
            gy = TY + 2 + q / GX;
            gx = q % GX;
        }
    }
};

// ---------------------------------------------------------------------------
// Two 3D forward levels per launch (temporal blocking) on LDS-staged tiles.
// A work-group of 256 lanes owns a TX x TY tile for LB shots (lane = shot x
// PX x PY patch) and marches planes zz = z0-1 .. z1.  On plane zz it stages
// level n-1 (A) of the tile with a two-cell halo, computes level n (C) on the
// tile and its one-cell ring (no corners; each ring (cell, shot) has an owner
// lane that keeps A of its cell on planes zz-1 .. zz+1) and keeps C, as
// stored (hi + lo), in LDS; then level n+1 (D) of plane zz-1 follows from C on
// planes zz-2 .. zz.  Every C and D cell takes the operations of the
// single-level kernels on the same operands, so the two levels equal two
// k_fwd3d launches bit for bit.  The planes z0-1 and z1 of C are re-computed
// from the neighbouring chunks (not stored).  Per shot-cell the launch reads
// A and B (level n-2) once and writes C and D: 24 B for two levels.
// ---------------------------------------------------------------------------
template <int LB, int TX, int TY, int PX, int PY, bool REC>
__global__ void __launch_bounds__(kBlock) USCT_ATTR3DX2
k_fwd3d_x2_lds(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pa,
               float *__restrict__ pb, const uint16_t *__restrict__ pcl,
               const uint16_t *__restrict__ ppl, uint16_t *__restrict__ pal,
               uint16_t *__restrict__ pbl, const float2 *__restrict__ sp2,
               const int *__restrict__ src_cell, const double *__restrict__ sinc_a,
               const double *__restrict__ sinc_b, const int *__restrict__ rec_head,
               const int *__restrict__ rec_next, const int *__restrict__ row_rec,
               float *__restrict__ rec_a, float *__restrict__ rec_a_lo, float *__restrict__ rec_b,
               float *__restrict__ rec_b_lo, float *__restrict__ srcd2_a,
               float *__restrict__ srcd2_b, int nz, int ny, int nx, int Bp, int RZ, int ntx,
               int nty, int nzc, int ng) {
    constexpr int TCX = TX / PX, TCY = TY / PY;
    static_assert(TCX * PX == TX && TCY * PY == TY && TCX * TCY * LB == kBlock, "tile shape");
    constexpr int HX = TX + 2, HY = TY + 2;     // C: tile + ring, padded 0 .. TX+1
    constexpr int GX = TX + 4, GY = TY + 4;     // A: tile + two-cell halo, padded 0 .. TX+3
    constexpr int RN = Ring<TX, TY>::N * LB;    // C ring entries (cell, shot)
    constexpr int RPL = (RN + kBlock - 1) / kBlock;
    constexpr int AN = Halo2<TX, TY>::N * LB;   // A halo entries
    constexpr int APL = (AN + kBlock - 1) / kBlock;
    __shared__ double shA[GY * GX * LB];
    __shared__ double shC[2][HY * HX * LB];
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int sg = blk % ng;
    int t = blk / ng;
    const int txi = t % ntx;
    t /= ntx;
    const int tyi = t % nty, zc = t / nty;
    if (zc >= nzc) return;  // uniform over the work-group
    const int tid = threadIdx.x;
    const int bl = tid % LB, col = tid / LB;
    const int cx = col % TCX, cy = col / TCX;
    const int b = sg * LB + bl;
    const int X0 = txi * TX, Y0 = tyi * TY;
    const int x0 = X0 + cx * PX, y0 = Y0 + cy * PY;
    const int z0 = zc * RZ, z1 = min(z0 + RZ, nz);
    const int my_src = src_cell[b];
    const double inc_a = sinc_a[b], inc_b = sinc_b[b];
    const size_t planeE = (size_t)nx * ny * Bp, planeC = (size_t)nx * ny;
    // element offsets from a plane base shifted back by two cells (x = -2 is 0);
    // rows clamped into the grid (values of out-of-grid rows feed only out-of-grid cells)
    auto off = [&](int y, int x, int shot) -> unsigned {
        return (unsigned)(((min(max(y, 0), ny - 1) * nx + x + 2) * Bp + shot));
    };
    unsigned oo[PY][PX], ro[RPL], ao[APL];
    int r_hx[RPL], r_hy[RPL], a_gx[APL], a_gy[APL];
    bool r_on[RPL], a_on[APL];
#pragma unroll
    for (int jj = 0; jj < PY; ++jj)
#pragma unroll
        for (int ii = 0; ii < PX; ++ii) oo[jj][ii] = off(y0 + jj, x0 + ii, b);
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        const int e = tid + q * kBlock;  // shot e % LB == bl (256 is a multiple of LB)
        r_on[q] = e < RN;
        Ring<TX, TY>::at(r_on[q] ? e / LB : 0, r_hx[q], r_hy[q]);
        ro[q] = off(Y0 + r_hy[q] - 1, X0 + r_hx[q] - 1, b);
    }
#pragma unroll
    for (int q = 0; q < APL; ++q) {
        const int e = tid + q * kBlock;
        a_on[q] = e < AN;
        Halo2<TX, TY>::at(a_on[q] ? e / LB : 0, a_gx[q], a_gy[q]);
        ao[q] = off(Y0 + a_gy[q] - 2, X0 + a_gx[q] - 2, b);
    }
    auto aidx = [&](int gy, int gx) { return (gy * GX + gx) * LB + bl; };
    auto cidx = [&](int hy, int hx) { return (hy * HX + hx) * LB + bl; };
    auto zcl = [&](int z) { return min(max(z, 0), nz - 1); };
    auto R4 = [&](const float *f, int z) { return f + (ptrdiff_t)zcl(z) * (ptrdiff_t)planeE - 2 * Bp; };
    auto R2 = [&](const uint16_t *f, int z) {
        return f + (ptrdiff_t)zcl(z) * (ptrdiff_t)planeE - 2 * Bp;
    };
    auto Wf = [&](float *f, int z) { return f + (ptrdiff_t)z * (ptrdiff_t)planeE - 2 * Bp; };
    auto Wh = [&](uint16_t *f, int z) { return f + (ptrdiff_t)z * (ptrdiff_t)planeE - 2 * Bp; };
    auto ld = [&](const float *h, const uint16_t *l, unsigned o) {
        return hl(ldb(h, o * 4u), ldb(l, o * 2u));
    };
    // own cells: A on planes zz-1, zz, zz+1; B on zz; C on zz-2, zz-1, zz
    double aM[PY][PX], aC[PY][PX], aP[PY][PX], bC[PY][PX], cM2[PY][PX], cM1[PY][PX];
    // ring cells: A on zz-1, zz, zz+1; B on zz
    double raM[RPL], raC[RPL], raP[RPL], rbC[RPL];
    // raw loads one plane ahead: A own and ring on zz+2, B own and ring on zz+1,
    // A halo on zz+1
    float nah[PY][PX], nbh[PY][PX], nrah[RPL], nrbh[RPL], nhh[APL];
    uint16_t nal[PY][PX], nbl[PY][PX], nral[RPL], nrbl[RPL], nhl[APL];
    float hh[APL];
    uint16_t hlo[APL];
    auto load_raw = [&](int za, int zb, int zh) {
        const float *a = R4(pc, za), *bb = R4(pp, zb), *h = R4(pc, zh);
        const uint16_t *al = R2(pcl, za), *bbl = R2(ppl, zb), *hl2 = R2(pcl, zh);
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                nah[jj][ii] = ldb(a, oo[jj][ii] * 4u);
                nal[jj][ii] = ldb(al, oo[jj][ii] * 2u);
                nbh[jj][ii] = ldb(bb, oo[jj][ii] * 4u);
                nbl[jj][ii] = ldb(bbl, oo[jj][ii] * 2u);
            }
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            if (!r_on[q]) continue;
            nrah[q] = ldb(a, ro[q] * 4u);
            nral[q] = ldb(al, ro[q] * 2u);
            nrbh[q] = ldb(bb, ro[q] * 4u);
            nrbl[q] = ldb(bbl, ro[q] * 2u);
        }
#pragma unroll
        for (int q = 0; q < APL; ++q) {
            if (!a_on[q]) continue;
            nhh[q] = ldb(h, ao[q] * 4u);
            nhl[q] = ldb(hl2, ao[q] * 2u);
        }
    };
    const int zs = z0 - 1, ze = z1;
    {
        const float *am = R4(pc, zs - 1), *a0 = R4(pc, zs), *ap = R4(pc, zs + 1), *b0 = R4(pp, zs);
        const uint16_t *aml = R2(pcl, zs - 1), *a0l = R2(pcl, zs), *apl = R2(pcl, zs + 1),
                       *b0l = R2(ppl, zs);
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                aM[jj][ii] = ld(am, aml, oo[jj][ii]);
                aC[jj][ii] = ld(a0, a0l, oo[jj][ii]);
                aP[jj][ii] = ld(ap, apl, oo[jj][ii]);
                bC[jj][ii] = ld(b0, b0l, oo[jj][ii]);
                cM2[jj][ii] = 0.0;
                cM1[jj][ii] = 0.0;
            }
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            raM[q] = raC[q] = raP[q] = rbC[q] = 0.0;
            if (!r_on[q]) continue;
            raM[q] = ld(am, aml, ro[q]);
            raC[q] = ld(a0, a0l, ro[q]);
            raP[q] = ld(ap, apl, ro[q]);
            rbC[q] = ld(b0, b0l, ro[q]);
        }
#pragma unroll
        for (int q = 0; q < APL; ++q) {
            hh[q] = 0.0f;
            hlo[q] = 0;
            if (!a_on[q]) continue;
            hh[q] = ldb(a0, ao[q] * 4u);
            hlo[q] = ldb(a0l, ao[q] * 2u);
        }
    }
    for (int zz = zs; zz <= ze; ++zz) {
        const int cur = zz & 1;
        // 1. stage A on plane zz: own cells and the two-cell halo
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii)
                shA[aidx(cy * PY + jj + 2, cx * PX + ii + 2)] = aC[jj][ii];
#pragma unroll
        for (int q = 0; q < APL; ++q)
            if (a_on[q]) shA[aidx(a_gy[q], a_gx[q])] = hl(hh[q], hlo[q]);
        if (zz < ze) load_raw(zz + 2, zz + 1, zz + 1);
        __syncthreads();
        // 2. C on plane zz: own cells (stored inside the chunk) and ring cells
        const bool cval = zz >= 0 && zz <= nz - 1;
        const bool cstore = zz >= z0 && zz < z1;
        const bool zint = zz >= 1 && zz <= nz - 2;
        double cC[PY][PX];
        if (cval) {
            const float2 *spz = sp2 + (size_t)zz * planeC;
#pragma unroll
            for (int jj = 0; jj < PY; ++jj) {
                const int y = y0 + jj;
                const bool yint = zint && y >= 1 && y <= ny - 2;
                const bool has_rec = REC && cstore && y < ny && row_rec[zz * ny + y] > 0;
                const int gy = cy * PY + jj + 2;
#pragma unroll
                for (int ii = 0; ii < PX; ++ii) {
                    const int x = x0 + ii;
                    const int gx = cx * PX + ii + 2;
                    cC[jj][ii] = 0.0;
                    if (x < nx && y < ny) {
                        double lu = 0.0;
                        if (yint && x >= 1 && x <= nx - 2)
                            lu = lap7d(aC[jj][ii], aP[jj][ii], aM[jj][ii], shA[aidx(gy + 1, gx)],
                                       shA[aidx(gy - 1, gx)], shA[aidx(gy, gx + 1)],
                                       shA[aidx(gy, gx - 1)]);
                        const int c = (zz * ny + y) * nx + x;
                        double inc = s2d(spz[y * nx + x]) * lu;
                        if (c == my_src) inc = inc + inc_a;
                        const double nv = (2.0 * aC[jj][ii] - bC[jj][ii]) + inc;
                        uint16_t lb;
                        const float v = split_hl(nv, lb);
                        cC[jj][ii] = hl(v, lb);
                        if (cstore) {
                            stb(Wf(pa, zz), oo[jj][ii] * 4u, v);
                            stb(Wh(pal, zz), oo[jj][ii] * 2u, lb);
                            if (c == my_src) srcd2_a[b] = (float)inc;
                            if (has_rec) {
                                for (int r = rec_head[c]; r >= 0; r = rec_next[r]) {
                                    rec_a[(size_t)r * Bp + b] = v;
                                    rec_a_lo[(size_t)r * Bp + b] = bf2f(lb);
                                }
                            }
                        }
                    }
                    shC[cur][cidx(cy * PY + jj + 1, cx * PX + ii + 1)] = cC[jj][ii];
                }
            }
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                if (!r_on[q]) continue;
                const int x = X0 + r_hx[q] - 1, y = Y0 + r_hy[q] - 1;
                double cv = 0.0;
                if (x >= 0 && x < nx && y >= 0 && y < ny) {
                    const int gx = r_hx[q] + 1, gy = r_hy[q] + 1;
                    double lu = 0.0;
                    if (zint && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2)
                        lu = lap7d(raC[q], raP[q], raM[q], shA[aidx(gy + 1, gx)],
                                   shA[aidx(gy - 1, gx)], shA[aidx(gy, gx + 1)],
                                   shA[aidx(gy, gx - 1)]);
                    const int c = (zz * ny + y) * nx + x;
                    double inc = s2d(spz[y * nx + x]) * lu;
                    if (c == my_src) inc = inc + inc_a;
                    const double nv = (2.0 * raC[q] - rbC[q]) + inc;
                    uint16_t lb;
                    const float v = split_hl(nv, lb);
                    cv = hl(v, lb);
                }
                shC[cur][cidx(r_hy[q], r_hx[q])] = cv;
            }
        }
        __syncthreads();
        // 3. D on plane zd = zz - 1 (own cells) from C on planes zz-2 .. zz
        const int zd = zz - 1;
        if (zd >= z0 && zd < z1) {
            const bool dint = zd >= 1 && zd <= nz - 2;
            const float2 *spz = sp2 + (size_t)zd * planeC;
#pragma unroll
            for (int jj = 0; jj < PY; ++jj) {
                const int y = y0 + jj;
                const bool yint = dint && y >= 1 && y <= ny - 2;
                const bool has_rec = REC && y < ny && row_rec[zd * ny + y] > 0;
                const int hy = cy * PY + jj + 1;
#pragma unroll
                for (int ii = 0; ii < PX; ++ii) {
                    const int x = x0 + ii;
                    const int hx = cx * PX + ii + 1;
                    if (x < nx && y < ny) {
                        double lu = 0.0;
                        if (yint && x >= 1 && x <= nx - 2)
                            lu = lap7d(cM1[jj][ii], cC[jj][ii], cM2[jj][ii],
                                       shC[cur ^ 1][cidx(hy + 1, hx)], shC[cur ^ 1][cidx(hy - 1, hx)],
                                       shC[cur ^ 1][cidx(hy, hx + 1)], shC[cur ^ 1][cidx(hy, hx - 1)]);
                        const int c = (zd * ny + y) * nx + x;
                        double inc = s2d(spz[y * nx + x]) * lu;
                        if (c == my_src) inc = inc + inc_b;
                        const double nv = (2.0 * cM1[jj][ii] - aM[jj][ii]) + inc;
                        uint16_t lb;
                        const float v = split_hl(nv, lb);
                        stb(Wf(pb, zd), oo[jj][ii] * 4u, v);
                        stb(Wh(pbl, zd), oo[jj][ii] * 2u, lb);
                        if (c == my_src) srcd2_b[b] = (float)inc;
                        if (has_rec) {
                            for (int r = rec_head[c]; r >= 0; r = rec_next[r]) {
                                rec_b[(size_t)r * Bp + b] = v;
                                rec_b_lo[(size_t)r * Bp + b] = bf2f(lb);
                            }
                        }
                    }
                }
            }
        }
        // 4. advance one plane
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                cM2[jj][ii] = cM1[jj][ii];
                cM1[jj][ii] = cval ? cC[jj][ii] : 0.0;
                aM[jj][ii] = aC[jj][ii];
                aC[jj][ii] = aP[jj][ii];
                aP[jj][ii] = hl(nah[jj][ii], nal[jj][ii]);
                bC[jj][ii] = hl(nbh[jj][ii], nbl[jj][ii]);
            }
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            raM[q] = raC[q];
            raC[q] = raP[q];
            raP[q] = r_on[q] ? hl(nrah[q], nral[q]) : 0.0;
            rbC[q] = r_on[q] ? hl(nrbh[q], nrbl[q]) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < APL; ++q) {
            hh[q] = nhh[q];
            hlo[q] = nhl[q];
        }
    }
}

// 3D adjoint level with the imaging condition (as k_adj3d), LDS-staged tiles of
// S' lam^{k+1} and p^{k-1}.  Grid as k_fwd3d_lds.
template <int LB, int TX, int TY, int PX, int PY>
__global__ void __launch_bounds__(kBlock) USCT_ATTR3D
k_adj3d_lds(const float *__restrict__ l1, const float *__restrict__ l2, float *__restrict__ lk,
            const float *__restrict__ pa, const float2 *__restrict__ sp2,
            const int *__restrict__ src_cell, const int *__restrict__ rec_head,
            const int *__restrict__ rec_next, const int *__restrict__ row_rec,
            const float *__restrict__ res, const float *__restrict__ srcd2,
            float *__restrict__ g_hi, float *__restrict__ g_lo, int nz, int ny, int nx, int Bp,
            int RZ, int ntx, int nty, int nzc, int ng) {
    constexpr int TCX = TX / PX, TCY = TY / PY, PC = PX * PY;
    static_assert(TCX * PX == TX && TCY * PY == TY && TCX * TCY * LB == kBlock, "tile shape");
    static_assert(PC <= LB, "a patch is reduced over its lane group");
    constexpr int HX = TX + 2, HY = TY + 2;
    constexpr int RN = Ring<TX, TY>::N * LB;
    constexpr int RPL = (RN + kBlock - 1) / kBlock;
    __shared__ float shq[2][HY * HX * LB], shf[2][HY * HX * LB];
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int sg = blk % ng;
    int t = blk / ng;
    const int txi = t % ntx;
    t /= ntx;
    const int tyi = t % nty, zc = t / nty;
    if (zc >= nzc) return;  // uniform over the work-group
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int bl = tid % LB, col = tid / LB;
    const int cx = col % TCX, cy = col / TCX;
    const int b = sg * LB + bl;
    const int X0 = txi * TX, Y0 = tyi * TY;
    const int x0 = X0 + cx * PX, y0 = Y0 + cy * PY;
    const int z0 = zc * RZ, z1 = min(z0 + RZ, nz);
    const int my_src = src_cell[b];
    const float my_d2 = srcd2[b];
    const size_t planeE = (size_t)nx * ny * Bp, planeC = (size_t)nx * ny;
    const size_t N = planeC * nz;
    // per-lane 32-bit offsets (elements) from a plane base shifted back by one
    // cell: own cells, ring entries; cell offsets likewise for S'
    auto off = [&](int y, int x, int shot) -> unsigned {
        return (unsigned)(((min(max(y, 0), ny - 1) * nx + x + 1) * Bp + shot));
    };
    auto coff = [&](int y, int x) -> unsigned {
        return (unsigned)(min(max(y, 0), ny - 1) * nx + x + 1);
    };
    unsigned oo[PY][PX], oc[PY][PX], ro[RPL], rc[RPL];
    int r_hx[RPL], r_hy[RPL];
    bool r_on[RPL];
#pragma unroll
    for (int jj = 0; jj < PY; ++jj)
#pragma unroll
        for (int ii = 0; ii < PX; ++ii) {
            oo[jj][ii] = off(y0 + jj, x0 + ii, b);
            oc[jj][ii] = coff(y0 + jj, x0 + ii);
        }
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        const int e = tid + q * kBlock;
        r_on[q] = e < RN;
        Ring<TX, TY>::at(r_on[q] ? e / LB : 0, r_hx[q], r_hy[q]);
        ro[q] = off(Y0 + r_hy[q] - 1, X0 + r_hx[q] - 1, sg * LB + e % LB);
        rc[q] = coff(Y0 + r_hy[q] - 1, X0 + r_hx[q] - 1);
    }
    auto sidx = [&](int hy, int hx, int shot) { return (hy * HX + hx) * LB + shot; };
    auto F4 = [&](const float *f, int z) { return f + (ptrdiff_t)z * (ptrdiff_t)planeE - Bp; };
    auto W4 = [&](float *f, int z) { return f + (ptrdiff_t)z * (ptrdiff_t)planeE - Bp; };
    auto S2 = [&](int z) { return sp2 + (ptrdiff_t)z * (ptrdiff_t)planeC - 1; };
    float qm[PY][PX], qcn[PY][PX], qp[PY][PX], lcn[PY][PX];
    float fm[PY][PX], fcn[PY][PX], fp[PY][PX];
    // raw loads one plane ahead: lam^{k+1} and p^{k-1} of plane zn (own cells and
    // ring), lam^{k+2} of plane z2 (own cells)
    float ln[PY][PX], fnr[PY][PX], l2n[PY][PX], rlam[RPL], rf[RPL];
    auto load_raw = [&](int zn, int z2) {
        const float *a = F4(l1, zn), *f = F4(pa, zn), *a2 = F4(l2, z2);
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                ln[jj][ii] = ldb(a, oo[jj][ii] * 4u);
                fnr[jj][ii] = ldb(f, oo[jj][ii] * 4u);
                l2n[jj][ii] = ldb(a2, oo[jj][ii] * 4u);
            }
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            if (!r_on[q]) continue;
            rlam[q] = ldb(a, ro[q] * 4u);
            rf[q] = ldb(f, ro[q] * 4u);
        }
    };
    // stage plane zn (S' lam and p, own cells and ring) into buffer s
    auto stage = [&](int s, int zn, const float (&qo)[PY][PX], const float (&fo)[PY][PX]) {
        const float2 *sz = S2(zn);
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) {
                const int si = sidx(cy * PY + jj + 1, cx * PX + ii + 1, bl);
                shq[s][si] = qo[jj][ii];
                shf[s][si] = fo[jj][ii];
            }
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            if (!r_on[q]) continue;
            const int e = tid + q * kBlock;
            const int si = sidx(r_hy[q], r_hx[q], e % LB);
            shq[s][si] = smul(ldb(sz, rc[q] * 8u), rlam[q]);
            shf[s][si] = rf[q];
        }
    };
    auto own_q = [&](int zn, float (&qo)[PY][PX]) {
        const float2 *sz = S2(zn);
#pragma unroll
        for (int jj = 0; jj < PY; ++jj)
#pragma unroll
            for (int ii = 0; ii < PX; ++ii) qo[jj][ii] = smul(ldb(sz, oc[jj][ii] * 8u), ln[jj][ii]);
    };
    {
        const int zmi = max(z0 - 1, 0);
        const float *am = F4(l1, zmi), *fmz = F4(pa, zmi);
        const float2 *sm = S2(zmi);
#pragma unroll
        for (int j = 0; j < PY; ++j)
#pragma unroll
            for (int i = 0; i < PX; ++i) {
                qm[j][i] = smul(ldb(sm, oc[j][i] * 8u), ldb(am, oo[j][i] * 4u));
                fm[j][i] = ldb(fmz, oo[j][i] * 4u);
            }
        load_raw(z0, z0);
        own_q(z0, qcn);
#pragma unroll
        for (int j = 0; j < PY; ++j)
#pragma unroll
            for (int i = 0; i < PX; ++i) {
                lcn[j][i] = ln[j][i];
                fcn[j][i] = fnr[j][i];
            }
        stage(0, z0, qcn, fcn);
        load_raw(min(z0 + 1, nz - 1), z0);
        __syncthreads();
    }
    // the lane's reduced cell (the same on every plane) and its gradient
    // accumulators, loaded one plane ahead so the update does not wait on them
    const bool gown = (bl & (LB / PC - 1)) == 0;
    const int gci = sum_cell<LB, PC>(lane);
    const int gx = x0 + gci % PX, gy = y0 + gci / PX;
    const bool gval = gown && gx < nx && gy < ny;
    const size_t gbase = (size_t)sg * N + (size_t)gy * nx + gx;
    float ghn = 0.0f, gln = 0.0f;
    if (gval) {
        ghn = g_hi[gbase + (size_t)z0 * planeC];
        gln = g_lo[gbase + (size_t)z0 * planeC];
    }
    for (int z = z0; z < z1; ++z) {
        const int cur = (z - z0) & 1;
        const int zpi = min(z + 1, nz - 1);
        float l2c[PY][PX], lnx[PY][PX];
        const float gh0 = ghn, gl0 = gln;
        if (gval && z + 1 < z1) {
            ghn = g_hi[gbase + (size_t)(z + 1) * planeC];
            gln = g_lo[gbase + (size_t)(z + 1) * planeC];
        }
        own_q(zpi, qp);
#pragma unroll
        for (int j = 0; j < PY; ++j)
#pragma unroll
            for (int i = 0; i < PX; ++i) {
                fp[j][i] = fnr[j][i];
                l2c[j][i] = l2n[j][i];
                lnx[j][i] = ln[j][i];
            }
        if (z + 1 < z1) {
            stage(cur ^ 1, zpi, qp, fp);
            load_raw(min(z + 2, nz - 1), z + 1);
        }
        const bool zint = (z >= 1) && (z <= nz - 2);
        float tv[PC];
#pragma unroll
        for (int j = 0; j < PY; ++j) {
            const int y = y0 + j;
            const bool yint = zint && (y >= 1) && (y <= ny - 2);
            const bool has_rec = y < ny && row_rec[z * ny + y] > 0;
            const int ly = cy * PY + j + 1;
#pragma unroll
            for (int i = 0; i < PX; ++i) {
                const int x = x0 + i;
                const int lx = cx * PX + i + 1;
                tv[j * PX + i] = 0.0f;
                if (x < nx && y < ny) {
                    const bool in = yint && x >= 1 && x <= nx - 2;
                    float lq = 0.0f, lp = 0.0f;
                    if (in) {
                        const float qym = j > 0 ? qcn[j - 1][i] : shq[cur][sidx(ly - 1, lx, bl)];
                        const float qyp = j < PY - 1 ? qcn[j + 1][i] : shq[cur][sidx(ly + 1, lx, bl)];
                        const float qxm = i > 0 ? qcn[j][i - 1] : shq[cur][sidx(ly, lx - 1, bl)];
                        const float qxp = i < PX - 1 ? qcn[j][i + 1] : shq[cur][sidx(ly, lx + 1, bl)];
                        lq = lap7(qcn[j][i], qp[j][i], qm[j][i], qyp, qym, qxp, qxm);
                        const float fym = j > 0 ? fcn[j - 1][i] : shf[cur][sidx(ly - 1, lx, bl)];
                        const float fyp = j < PY - 1 ? fcn[j + 1][i] : shf[cur][sidx(ly + 1, lx, bl)];
                        const float fxm = i > 0 ? fcn[j][i - 1] : shf[cur][sidx(ly, lx - 1, bl)];
                        const float fxp = i < PX - 1 ? fcn[j][i + 1] : shf[cur][sidx(ly, lx + 1, bl)];
                        lp = lap7(fcn[j][i], fp[j][i], fm[j][i], fyp, fym, fxp, fxm);
                    }
                    float v = leap(lcn[j][i], l2c[j][i], lq);
                    const int c = (z * ny + y) * nx + x;
                    if (has_rec) {
                        for (int r = rec_head[c]; r >= 0; r = rec_next[r])
                            v = v - res[(size_t)r * Bp + b];
                    }
                    stb(W4(lk, z), oo[j][i] * 4u, v);
                    float d2 = smul(ldb(S2(z), oc[j][i] * 8u), lp);
                    if (c == my_src) d2 = my_d2;
                    tv[j * PX + i] = v * d2;
                }
            }
        }
        const float s = reduce_scatter<LB, PC>(tv, lane);
        if (gval) {
            float hi = gh0, lo = gl0;
            neumaier(hi, lo, s);
            g_hi[gbase + (size_t)z * planeC] = hi;
            g_lo[gbase + (size_t)z * planeC] = lo;
        }
#pragma unroll
        for (int j = 0; j < PY; ++j)
#pragma unroll
            for (int i = 0; i < PX; ++i) {
                qm[j][i] = qcn[j][i];
                qcn[j][i] = qp[j][i];
                lcn[j][i] = lnx[j][i];
                fm[j][i] = fcn[j][i];
                fcn[j][i] = fp[j][i];
            }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Batched 3D adjoint step with the imaging condition (no sponge in 3D:
// the reference sponge is 2D only, solver.py:30).  Addressing as k_fwd3d.
// ---------------------------------------------------------------------------
template <int LB, int CX, int CY>
__global__ void __launch_bounds__(kBlock) USCT_ATTR3D
k_adj3d(const float *__restrict__ l1, const float *__restrict__ l2, float *__restrict__ lk,
        const float *__restrict__ pa, const float2 *__restrict__ sp2,
        const int *__restrict__ src_cell, const float *__restrict__ sinc,
        const int *__restrict__ rec_head, const int *__restrict__ rec_next,
        const int *__restrict__ row_rec, const float *__restrict__ res,
        const float *__restrict__ srcd2, float *__restrict__ g_hi, float *__restrict__ g_lo,
        int nz, int ny, int nx, int Bp, int RZ, int nxs, int nys, int nzc, int ng, int blk) {
    constexpr int W = kWave / LB;
    constexpr int CC = CX * CY;
    const int wv = wave_id();
    const int lane = threadIdx.x & 63;
    const int sg = wv % ng;
    int xs, ys, zc;
    if (!strip3d(wv / ng, nxs, nys, nzc, blk, xs, ys, zc)) return;  // uniform over the wavefront
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    const int y0 = ys * CY;
    const int z0 = zc * RZ;
    const int z1 = min(z0 + RZ, nz);
    (void)sinc;
    const size_t rowE = (size_t)nx * Bp;
    const size_t planeE = rowE * ny;
    const size_t planeC = (size_t)nx * ny;
    const size_t N = planeC * nz;
    const int my_src = src_cell[b];
    const float my_d2 = srcd2[b];
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned oc = (unsigned)x0 * 8u;
    auto row = [&](int j) -> size_t { return (size_t)min(max(y0 + j, 0), ny - 1); };
    auto F = [&](const float *plane, int j, int i) -> float {
        return ldb(plane + row(j) * rowE + (ptrdiff_t)i * Bp, o4);
    };
    auto Q = [&](const float *plane, const float2 *splane, int j, int i) -> float {
        return smul(ldb(splane + row(j) * nx + i, oc), F(plane, j, i));
    };
    float qm[CY][CX], qc[CY + 2][CX + 2], qp[CY][CX], lc[CY][CX], lpn[CY][CX];
    float fm[CY][CX], fc[CY + 2][CX + 2], fp[CY][CX];
    auto load_halo = [&](const float *lz, const float2 *sz, const float *fz) {
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            qc[0][i + 1] = Q(lz, sz, -1, i);
            qc[CY + 1][i + 1] = Q(lz, sz, CY, i);
            fc[0][i + 1] = F(fz, -1, i);
            fc[CY + 1][i + 1] = F(fz, CY, i);
        }
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            qc[j + 1][0] = Q(lz, sz, j, -1);
            qc[j + 1][CX + 1] = Q(lz, sz, j, CX);
            fc[j + 1][0] = F(fz, j, -1);
            fc[j + 1][CX + 1] = F(fz, j, CX);
        }
    };
    // raw loads issued one plane ahead: own cells of lam^{k+1} and p^{k-1} on
    // plane z+1 (an, fn) and lam^{k+2} on plane z (l2n)
    float an[CY][CX], fn[CY][CX], l2n[CY][CX];
    auto load_raw = [&](size_t pl, size_t pl2) {
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                an[j][i] = F(l1 + pl, j, i);
                fn[j][i] = F(pa + pl, j, i);
                l2n[j][i] = ldn<USCT_NT3D>(l2 + pl2 + row(j) * rowE + (ptrdiff_t)i * Bp, o4);
            }
    };
    {
        const int zmi = max(z0 - 1, 0);
        const float *lm = l1 + (size_t)zmi * planeE, *l0 = l1 + (size_t)z0 * planeE;
        const float2 *sm = sp2 + (size_t)zmi * planeC, *s0 = sp2 + (size_t)z0 * planeC;
        const float *fmz = pa + (size_t)zmi * planeE, *f0 = pa + (size_t)z0 * planeE;
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                qm[j][i] = Q(lm, sm, j, i);
                fm[j][i] = F(fmz, j, i);
                const float v = F(l0, j, i);
                lc[j][i] = v;
                qc[j + 1][i + 1] = smul(ldb(s0 + row(j) * nx + i, oc), v);
                fc[j + 1][i + 1] = F(f0, j, i);
            }
        load_halo(l0, s0, f0);
        load_raw((size_t)min(z0 + 1, nz - 1) * planeE, (size_t)z0 * planeE);
    }
    // the lane's reduced cell (the same on every plane) and its gradient
    // accumulators, loaded one plane ahead so the update does not wait on them
    const bool gown = (sb & (LB / CC - 1)) == 0;
    const int gci = sum_cell<LB, CC>(lane);
    const int gx = x0 + gci % CX, gy = y0 + gci / CX;
    const bool gval = gown && gx < nx && gy < ny;
    const size_t gbase = (size_t)sg * N + (size_t)gy * nx + gx;
    float ghn = 0.0f, gln = 0.0f;
    if (gval) {
        ghn = g_hi[gbase + (size_t)z0 * planeC];
        gln = g_lo[gbase + (size_t)z0 * planeC];
    }
    for (int z = z0; z < z1; ++z) {
        const int zpi = min(z + 1, nz - 1);
        const float *lnz = l1 + (size_t)zpi * planeE;
        const float2 *snz = sp2 + (size_t)zpi * planeC;
        const float *fnz = pa + (size_t)zpi * planeE;
        float l2c[CY][CX];
        const float gh0 = ghn, gl0 = gln;
        if (gval && z + 1 < z1) {
            ghn = g_hi[gbase + (size_t)(z + 1) * planeC];
            gln = g_lo[gbase + (size_t)(z + 1) * planeC];
        }
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const float v = an[j][i];
                lpn[j][i] = v;
                qp[j][i] = smul(ldb(snz + row(j) * nx + i, oc), v);
                fp[j][i] = fn[j][i];
                l2c[j][i] = l2n[j][i];
            }
        if (z + 1 < z1) load_raw((size_t)min(z + 2, nz - 1) * planeE, (size_t)zpi * planeE);
        const bool zint = (z >= 1) && (z <= nz - 2);
        float tv[CC];
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            const int y = y0 + j;
            const bool yint = zint && (y >= 1) && (y <= ny - 2);
            const bool has_rec = y < ny && row_rec[z * ny + y] > 0;
            const size_t ro = (size_t)z * planeE + (size_t)y * rowE;
            const float2 *spr = sp2 + (size_t)z * planeC + (size_t)y * nx;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                tv[j * CX + i] = 0.0f;
                if (x < nx && y < ny) {
                    const bool in = yint && x >= 1 && x <= nx - 2;
                    const float lq = in ? lap7(qc[j + 1][i + 1], qp[j][i], qm[j][i], qc[j + 2][i + 1],
                                               qc[j][i + 1], qc[j + 1][i + 2], qc[j + 1][i])
                                        : 0.0f;
                    const size_t ei = ro + (size_t)i * Bp;
                    float v = leap(lc[j][i], l2c[j][i], lq);
                    const int c = (z * ny + y) * nx + x;
                    if (has_rec) {
                        for (int r = rec_head[c]; r >= 0; r = rec_next[r])
                            v = v - res[(size_t)r * Bp + b];
                    }
                    stn<USCT_NT3D>(lk + ei, o4, v);
                    const float lp = in ? lap7(fc[j + 1][i + 1], fp[j][i], fm[j][i], fc[j + 2][i + 1],
                                               fc[j][i + 1], fc[j + 1][i + 2], fc[j + 1][i])
                                        : 0.0f;
                    float d2 = smul(ldb(spr + i, oc), lp);
                    if (c == my_src) d2 = my_d2;
                    tv[j * CX + i] = v * d2;
                }
            }
        }
        const float s = reduce_scatter<LB, CC>(tv, lane);
        if (gval) {
            float hi = gh0, lo = gl0;
            neumaier(hi, lo, s);
            g_hi[gbase + (size_t)z * planeC] = hi;
            g_lo[gbase + (size_t)z * planeC] = lo;
        }
        if (z + 1 < z1) {
#pragma unroll
            for (int j = 0; j < CY; ++j)
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    qm[j][i] = qc[j + 1][i + 1];
                    qc[j + 1][i + 1] = qp[j][i];
                    lc[j][i] = lpn[j][i];
                    fm[j][i] = fc[j + 1][i + 1];
                    fc[j + 1][i + 1] = fp[j][i];
                }
            load_halo(lnz, snz, fnz);
        }
    }
}

// ---------------------------------------------------------------------------
// Batched 3D adjoint level k with the forward run backwards alongside it.
// Without a sponge the leapfrog recursion is its own inverse:
//   p^{k-2} = 2 p^{k-1} - p^k + inc,  inc = S'(L p^{k-1}) [+ S_src w[k-1]],
// the forward's FP64 update with the roles of p^n and p^{n-2} exchanged, and
// inc is exactly the second difference p^k - 2p^{k-1} + p^{k-2} the imaging
// condition multiplies by lam^k (fwi.py:323-326).  One launch therefore
// advances lam (the k_adj3d update, bit for bit), re-creates the forward level
// k-2 (hi + lo, FP64 from the full state, as k_fwd3d) and images with inc, so
// the gradient needs neither stored levels nor a re-computed forward sweep.
//   pc/pcl  level k-1 (stencil), pk/pkl level k (own cells),
//   po/pol  level k-2 (written when wr != 0; the caller passes wr = 0 for
//           k <= 2, whose levels 0 and -1 are zero by definition)
//   sincd   row k of the FP64 source increments.  Addressing as k_fwd3d.
// YSH: the kYsh wavefronts of a work-group take strips stacked along y (one x
// strip, z chunk and shot group: work-group t -> shot group t % ng, then x
// strip, then block of kYsh y strips, then z chunk) and march together; at each
// plane they hand the edge rows of their centre plane (S' lam^{k+1} and level
// k-1 in FP64) to each other through LDS, one barrier per plane, so only the
// group's outer y halo rows are loaded from memory.  The values are the ones
// the loads would give (the same cells, the same operations): bit-identical to
// YSH = false.  Strips past the grid's last y strip compute on clamped rows and
// store nothing, so that every wavefront of the group reaches every barrier.
// ---------------------------------------------------------------------------
constexpr int kYsh = kBlock / kWave;
template <int LB, int CX, int CY, class L = uint16_t, bool YSH = false>
__global__ void __launch_bounds__(kBlock) USCT_ATTR3D
k_radj3d(const float *__restrict__ l1, const float *__restrict__ l2, float *__restrict__ lk,
         const float *__restrict__ pc, const L *__restrict__ pcl,
         const float *__restrict__ pk, const L *__restrict__ pkl, float *__restrict__ po,
         L *__restrict__ pol, int wr, const float2 *__restrict__ sp2,
         const int *__restrict__ src_cell, const double *__restrict__ sincd,
         const int *__restrict__ rec_head, const int *__restrict__ rec_next,
         const int *__restrict__ row_rec, const float *__restrict__ res, float *__restrict__ g_hi,
         float *__restrict__ g_lo, int nz, int ny, int nx, int Bp, int RZ, int nxs, int nys,
         int nzc, int ng, int blk) {
    constexpr int W = kWave / LB;
    constexpr int CC = CX * CY;
    const int lane = threadIdx.x & 63;
    int sg, xs, ys, zc;
    const int wy = YSH ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    if constexpr (YSH) {
        int t = xcd_remap(blockIdx.x, gridDim.x);
        sg = t % ng;
        t /= ng;
        xs = t % nxs;
        t /= nxs;
        const int nyb = (nys + kYsh - 1) / kYsh;
        ys = (t % nyb) * kYsh + wy;
        zc = t / nyb;
        if (zc >= nzc) return;  // uniform over the work-group
    } else {
        const int wv = wave_id();
        sg = wv % ng;
        if (!strip3d(wv / ng, nxs, nys, nzc, blk, xs, ys, zc)) return;  // uniform over the wavefront
    }
    const int sb = lane % LB, st = lane / LB;
    const int b = sg * LB + sb;
    const int x0 = (xs * W + st) * CX;
    const int y0 = ys * CY;
    const int z0 = zc * RZ;
    const int z1 = min(z0 + RZ, nz);
    const size_t rowE = (size_t)nx * Bp;
    const size_t planeE = rowE * ny;
    const size_t planeC = (size_t)nx * ny;
    const size_t N = planeC * nz;
    const int my_src = src_cell[b];
    const double my_inc = sincd[b];
    const unsigned o4 = (unsigned)(x0 * Bp + b) * 4u;
    const unsigned o2 = (unsigned)(x0 * Bp + b) * (unsigned)sizeof(L);
    const unsigned oc = (unsigned)x0 * 8u;
    auto row = [&](int j) -> size_t { return (size_t)min(max(y0 + j, 0), ny - 1); };
    auto F = [&](const float *plane, int j, int i) -> float {
        return ldb(plane + row(j) * rowE + (ptrdiff_t)i * Bp, o4);
    };
    auto Q = [&](const float *plane, const float2 *splane, int j, int i) -> float {
        return smul(ldb(splane + row(j) * nx + i, oc), F(plane, j, i));
    };
    // full state hi + lo of level k-1 at (row j, column i) of a plane, in FP64
    auto P = [&](size_t plane, int j, int i) -> double {
        const size_t e = plane + row(j) * rowE + (size_t)((ptrdiff_t)i * Bp);
        return (double)ldb(pc + e, o4) + (double)lo_f(ldb(pcl + e, o2));
    };
    float qm[CY][CX], qc[CY + 2][CX + 2], qp[CY][CX], lc[CY][CX], lpn[CY][CX];
    double zm[CY][CX], zcu[CY + 2][CX + 2], zp[CY][CX];
    auto load_halo = [&](const float *lz, const float2 *sz, size_t pl) {
#pragma unroll
        for (int i = 0; i < CX; ++i) {
            qc[0][i + 1] = Q(lz, sz, -1, i);
            qc[CY + 1][i + 1] = Q(lz, sz, CY, i);
            zcu[0][i + 1] = P(pl, -1, i);
            zcu[CY + 1][i + 1] = P(pl, CY, i);
        }
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            qc[j + 1][0] = Q(lz, sz, j, -1);
            qc[j + 1][CX + 1] = Q(lz, sz, j, CX);
            zcu[j + 1][0] = P(pl, j, -1);
            zcu[j + 1][CX + 1] = P(pl, j, CX);
        }
    };
    // YSH: edge rows [plane parity][wavefront][top, bottom][cell][lane]
    constexpr int kEdge = YSH ? 2 * kYsh * 2 * CX * 64 : 1;
    __shared__ float eq[kEdge];
    __shared__ double ez[kEdge];
    int par = 0;
    auto halo = [&](const float *lz, const float2 *sz, size_t pl) {
        if constexpr (!YSH) {
            load_halo(lz, sz, pl);
        } else {
#pragma unroll
            for (int j = 0; j < CY; ++j) {
                qc[j + 1][0] = Q(lz, sz, j, -1);
                qc[j + 1][CX + 1] = Q(lz, sz, j, CX);
                zcu[j + 1][0] = P(pl, j, -1);
                zcu[j + 1][CX + 1] = P(pl, j, CX);
            }
            if (wy == 0)
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    qc[0][i + 1] = Q(lz, sz, -1, i);
                    zcu[0][i + 1] = P(pl, -1, i);
                }
            if (wy == kYsh - 1)
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    qc[CY + 1][i + 1] = Q(lz, sz, CY, i);
                    zcu[CY + 1][i + 1] = P(pl, CY, i);
                }
            const int e0 = par * (kYsh * 2 * CX * 64) + lane;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                eq[e0 + ((wy * 2 + 0) * CX + i) * 64] = qc[1][i + 1];
                eq[e0 + ((wy * 2 + 1) * CX + i) * 64] = qc[CY][i + 1];
                ez[e0 + ((wy * 2 + 0) * CX + i) * 64] = zcu[1][i + 1];
                ez[e0 + ((wy * 2 + 1) * CX + i) * 64] = zcu[CY][i + 1];
            }
            __syncthreads();
            if (wy > 0)
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    qc[0][i + 1] = eq[e0 + (((wy - 1) * 2 + 1) * CX + i) * 64];
                    zcu[0][i + 1] = ez[e0 + (((wy - 1) * 2 + 1) * CX + i) * 64];
                }
            if (wy < kYsh - 1)
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    qc[CY + 1][i + 1] = eq[e0 + (((wy + 1) * 2 + 0) * CX + i) * 64];
                    zcu[CY + 1][i + 1] = ez[e0 + (((wy + 1) * 2 + 0) * CX + i) * 64];
                }
            par ^= 1;
        }
    };
    // raw loads issued one plane ahead: own cells of lam^{k+1} and of level
    // k-1 on plane z+1 (an, nh/nl), of lam^{k+2} and of level k on plane z
    // (l2n, kh/kl)
    float an[CY][CX], nh[CY][CX], l2n[CY][CX], kh[CY][CX];
    L nl[CY][CX], kl[CY][CX];
    auto load_raw = [&](size_t pl, size_t pl2) {
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const size_t e = pl + row(j) * rowE + (size_t)((ptrdiff_t)i * Bp);
                const size_t e2 = pl2 + row(j) * rowE + (size_t)((ptrdiff_t)i * Bp);
                an[j][i] = ldb(l1 + e, o4);
                nh[j][i] = ldb(pc + e, o4);
                nl[j][i] = ldb(pcl + e, o2);
                l2n[j][i] = ldn<USCT_NTR3D>(l2 + e2, o4);
                kh[j][i] = ldn<USCT_NTR3D>(pk + e2, o4);
                kl[j][i] = ldn<USCT_NTR3D>(pkl + e2, o2);
            }
    };
    {
        const int zmi = max(z0 - 1, 0);
        const float *lm = l1 + (size_t)zmi * planeE, *l0 = l1 + (size_t)z0 * planeE;
        const float2 *sm = sp2 + (size_t)zmi * planeC, *s0 = sp2 + (size_t)z0 * planeC;
        const size_t pm = (size_t)zmi * planeE, p0 = (size_t)z0 * planeE;
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                qm[j][i] = Q(lm, sm, j, i);
                zm[j][i] = P(pm, j, i);
                const float v = F(l0, j, i);
                lc[j][i] = v;
                qc[j + 1][i + 1] = smul(ldb(s0 + row(j) * nx + i, oc), v);
                zcu[j + 1][i + 1] = P(p0, j, i);
            }
        halo(l0, s0, p0);
        load_raw((size_t)min(z0 + 1, nz - 1) * planeE, p0);
    }
    // the lane's reduced cell (the same on every plane) and its gradient
    // accumulators, loaded one plane ahead so the update does not wait on them
    const bool gown = (sb & (LB / CC - 1)) == 0;
    const int gci = sum_cell<LB, CC>(lane);
    const int gx = x0 + gci % CX, gy = y0 + gci / CX;
    const bool gval = gown && gx < nx && gy < ny;
    const size_t gbase = (size_t)sg * N + (size_t)gy * nx + gx;
    float ghn = 0.0f, gln = 0.0f;
    if (gval) {
        ghn = g_hi[gbase + (size_t)z0 * planeC];
        gln = g_lo[gbase + (size_t)z0 * planeC];
    }
    for (int z = z0; z < z1; ++z) {
        const int zpi = min(z + 1, nz - 1);
        const float *lnz = l1 + (size_t)zpi * planeE;
        const float2 *snz = sp2 + (size_t)zpi * planeC;
        const size_t pnz = (size_t)zpi * planeE;
        float l2c[CY][CX];
        double q[CY][CX];
        const float gh0 = ghn, gl0 = gln;
        if (gval && z + 1 < z1) {
            ghn = g_hi[gbase + (size_t)(z + 1) * planeC];
            gln = g_lo[gbase + (size_t)(z + 1) * planeC];
        }
#pragma unroll
        for (int j = 0; j < CY; ++j)
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const float v = an[j][i];
                lpn[j][i] = v;
                qp[j][i] = smul(ldb(snz + row(j) * nx + i, oc), v);
                zp[j][i] = hl(nh[j][i], nl[j][i]);
                l2c[j][i] = l2n[j][i];
                q[j][i] = hl(kh[j][i], kl[j][i]);
            }
        if (z + 1 < z1) load_raw((size_t)min(z + 2, nz - 1) * planeE, pnz);
        const bool zint = (z >= 1) && (z <= nz - 2);
        // The arithmetic of every cell of the patch first, with no branch between
        // cells, so their FP64 chains interleave; then receivers, stores and the
        // imaging products of the cells inside the grid.  Cells outside the grid
        // (strips rounded up) compute on clamped rows and are not stored.
        float vv[CY][CX], pvv[CY][CX];
        double incv[CY][CX];
        L lbv[CY][CX];
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            const int y = y0 + j;
            const bool yint = zint && (y >= 1) && (y <= ny - 2);
            const float2 *spr = sp2 + (size_t)z * planeC + row(j) * nx;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                const bool in = yint && x >= 1 && x <= nx - 2;
                const float lq = in ? lap7(qc[j + 1][i + 1], qp[j][i], qm[j][i], qc[j + 2][i + 1],
                                           qc[j][i + 1], qc[j + 1][i + 2], qc[j + 1][i])
                                    : 0.0f;
                vv[j][i] = leap(lc[j][i], l2c[j][i], lq);
                const double lu = in ? lap7d(zcu[j + 1][i + 1], zp[j][i], zm[j][i], zcu[j + 2][i + 1],
                                             zcu[j][i + 1], zcu[j + 1][i + 2], zcu[j + 1][i])
                                     : 0.0;
                double inc = s2d(ldb(spr + i, oc)) * lu;
                if ((z * ny + y) * nx + x == my_src) inc = inc + my_inc;
                incv[j][i] = inc;
                const double nv = (2.0 * zcu[j + 1][i + 1] - q[j][i]) + inc;
                pvv[j][i] = split_hl(nv, lbv[j][i]);
            }
        }
        float tv[CC];
#pragma unroll
        for (int j = 0; j < CY; ++j) {
            const int y = y0 + j;
            const bool has_rec = y < ny && row_rec[z * ny + y] > 0;
            const size_t ro = (size_t)z * planeE + (size_t)y * rowE;
#pragma unroll
            for (int i = 0; i < CX; ++i) {
                const int x = x0 + i;
                tv[j * CX + i] = 0.0f;
                if (x < nx && y < ny) {
                    const size_t ei = ro + (size_t)i * Bp;
                    float v = vv[j][i];
                    if (has_rec) {
                        const int c = (z * ny + y) * nx + x;
                        for (int r = rec_head[c]; r >= 0; r = rec_next[r])
                            v = v - res[(size_t)r * Bp + b];
                    }
                    stn<USCT_NTR3D>(lk + ei, o4, v);
                    if (wr) {
                        stn<USCT_NTR3D>(po + ei, o4, pvv[j][i]);
                        stn<USCT_NTR3D>(pol + ei, o2, lbv[j][i]);
                    }
                    tv[j * CX + i] = v * (float)incv[j][i];
                }
            }
        }
        const float s = reduce_scatter<LB, CC>(tv, lane);
        if (gval) {
            float hi = gh0, lo = gl0;
            neumaier(hi, lo, s);
            g_hi[gbase + (size_t)z * planeC] = hi;
            g_lo[gbase + (size_t)z * planeC] = lo;
        }
        if (z + 1 < z1) {
#pragma unroll
            for (int j = 0; j < CY; ++j)
#pragma unroll
                for (int i = 0; i < CX; ++i) {
                    qm[j][i] = qc[j + 1][i + 1];
                    qc[j + 1][i + 1] = qp[j][i];
                    lc[j][i] = lpn[j][i];
                    zm[j][i] = zcu[j + 1][i + 1];
                    zcu[j + 1][i + 1] = zp[j][i];
                }
            halo(lnz, snz, pnz);
        }
    }
}

// ---------------------------------------------------------------------------
// Plain (one shot) kernels: one thread per cell, dimension-general (2D: nz = 1).
// ---------------------------------------------------------------------------
template <bool SPONGE, bool REC, class L = uint16_t>
__global__ void __launch_bounds__(kBlock)
k_fwd_plain(const float *__restrict__ pc, const float *__restrict__ pp, float *__restrict__ pn,
            const L *__restrict__ pcl, const L *__restrict__ ppl,
            L *__restrict__ pnl, const float2 *__restrict__ sp2,
            const float *__restrict__ sponge, int src, const double *__restrict__ sincd,
            const int *__restrict__ rec_head, const int *__restrict__ rec_next,
            float *__restrict__ rec, float *__restrict__ rec_lo, float *__restrict__ srcd2, int nz,
            int ny, int nx, int ndim) {
    const int64_t N = (int64_t)nz * ny * nx;
    const int64_t c64 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c64 >= N) return;
    const int c = (int)c64;
    const int x = c % nx;
    const int y = (c / nx) % ny;
    const int z = c / (nx * ny);
    // FP64 update from the full hi + lo state (as the batched kernels)
    auto P = [&](int e) { return hl(pc[e], pcl[e]); };
    double lu = 0.0;
    if (ndim == 3) {
        const int sz = nx * ny;
        if (z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2)
            lu = lap7d(P(c), P(c + sz), P(c - sz), P(c + nx), P(c - nx), P(c + 1), P(c - 1));
    } else {
        if (y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2)
            lu = lap5d(P(c), P(c + nx), P(c - nx), P(c + 1), P(c - 1));
    }
    double inc = s2d(sp2[c]) * lu;
    if (c == src) inc = inc + *sincd;
    double nv = (2.0 * P(c) - hl(pp[c], ppl[c])) + inc;
    if (SPONGE) nv = nv * (double)sponge[c];
    if (c == src) *srcd2 = (float)inc;
    L lb;
    const float v = split_hl(nv, lb);
    pn[c] = v;
    pnl[c] = lb;
    if (REC) {
        for (int j = rec_head[c]; j >= 0; j = rec_next[j]) {
            rec[j] = v;
            rec_lo[j] = lo_f(lb);
        }
    }
}

template <bool SPONGE>
__global__ void __launch_bounds__(kBlock)
k_adj_plain(const float *__restrict__ l1, const float *__restrict__ l2, float *__restrict__ lk,
            const float *__restrict__ pa, const float *__restrict__ pb,
            const float *__restrict__ pc2, const float2 *__restrict__ sp2,
            const float *__restrict__ sponge, int src, const float *__restrict__ sinc,
            const float *__restrict__ srcd2, const int *__restrict__ rec_head,
            const int *__restrict__ rec_next, const float *__restrict__ res,
            float *__restrict__ g_hi, float *__restrict__ g_lo, int nz, int ny, int nx, int ndim) {
    const int64_t N = (int64_t)nz * ny * nx;
    const int64_t c64 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c64 >= N) return;
    const int c = (int)c64;
    const int x = c % nx;
    const int y = (c / nx) % ny;
    const int z = c / (nx * ny);
    bool in;
    float lq = 0.0f, lp = 0.0f;
    auto q = [&](int cc) { return smul(sp2[cc], l1[cc]); };
    if (ndim == 2) {
        in = (y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2);
        if (in) {
            lq = lap5(q(c), q(c + nx), q(c - nx), q(c + 1), q(c - 1));
            if (!SPONGE) lp = lap5(pa[c], pa[c + nx], pa[c - nx], pa[c + 1], pa[c - 1]);
        }
    } else {
        const int sz = nx * ny;
        in = (z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2);
        if (in) {
            lq = lap7(q(c), q(c + sz), q(c - sz), q(c + nx), q(c - nx), q(c + 1), q(c - 1));
            if (!SPONGE)
                lp = lap7(pa[c], pa[c + sz], pa[c - sz], pa[c + nx], pa[c - nx], pa[c + 1], pa[c - 1]);
        }
    }
    float v = leap(l1[c], l2[c], lq);
    if (SPONGE) v = v * sponge[c];
    for (int j = rec_head[c]; j >= 0; j = rec_next[j]) v = v - res[j];
    lk[c] = v;
    float d2;
    if (SPONGE) {
        d2 = (pa[c] - 2.0f * pb[c]) + pc2[c];  // stored levels: the source is in them
    } else {
        d2 = smul(sp2[c], lp);
        if (c == src) d2 = ndim == 3 ? *srcd2 : d2 + *sinc;
    }
    float hi = g_hi[c], lo = g_lo[c];
    neumaier(hi, lo, v * d2);
    g_hi[c] = hi;
    g_lo[c] = lo;
}

// ---------------------------------------------------------------------------
// Born (linearised) forward step, one thread per shot-cell of the [cell][shot]
// layout (Bp shots per cell; the plain path has Bp = 1), fwi.py:366-377:
//   dp^n = 2 dp^{n-1} - dp^{n-2} + S L(dp^{n-1}) - (v/m)(U^n - 2U^{n-1} + U^{n-2}),
// times the sponge.  Background levels U (u0 = U^n, u1 = U^{n-1}, u2 = U^{n-2})
// and the Born state are FP32 hi + bfloat16 lo, combined and updated in FP64
// as in the forward; the Laplacian is zero on border cells.  Each receiver j on
// the cell takes dp^n into res[j * Bp + b] (row n of the adjoint's source).
// Neighbour offsets are +-Bp (x), +-nx Bp (y), +-nx ny Bp (z).
// ---------------------------------------------------------------------------
template <bool SPONGE>
__global__ void __launch_bounds__(256)
k_born(const float *__restrict__ u0, const uint16_t *__restrict__ u0l,
       const float *__restrict__ u1, const uint16_t *__restrict__ u1l,
       const float *__restrict__ u2, const uint16_t *__restrict__ u2l,
       const float *__restrict__ d1, const uint16_t *__restrict__ d1l,
       const float *__restrict__ d2, const uint16_t *__restrict__ d2l,
       float *__restrict__ dn, uint16_t *__restrict__ dnl, const float2 *__restrict__ sp2,
       const float2 *__restrict__ vm2, const float *__restrict__ sponge,
       const int *__restrict__ rec_head, const int *__restrict__ rec_next,
       float *__restrict__ res, int nz, int ny, int nx, int ndim, int Bp) {
    const int64_t total = (int64_t)nz * ny * nx * Bp;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int c, b;
    if (total <= (int64_t)0xffffffffu) {
        const uint32_t u = (uint32_t)i;
        c = (int)(u / (uint32_t)Bp);
        b = (int)(u - (uint32_t)c * (uint32_t)Bp);
    } else {
        c = (int)(i / Bp);
        b = (int)(i - (int64_t)c * Bp);
    }
    const int x = c % nx;
    const int y = (c / nx) % ny;
    const int z = c / (nx * ny);
    auto D = [&](int64_t e) { return hl(d1[e], d1l[e]); };
    const int64_t sx = Bp, sy = (int64_t)nx * Bp;
    double lu = 0.0;
    if (ndim == 3) {
        const int64_t sz = sy * ny;
        if (z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2)
            lu = lap7d(D(i), D(i + sz), D(i - sz), D(i + sy), D(i - sy), D(i + sx), D(i - sx));
    } else if (y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2) {
        lu = lap5d(D(i), D(i + sy), D(i - sy), D(i + sx), D(i - sx));
    }
    // background levels and dp^{n-2} are read once: non-temporal (no L2 retention)
    const double d2u = (hl(ld_nt(u0 + i), ld_nt(u0l + i)) - 2.0 * hl(ld_nt(u1 + i), ld_nt(u1l + i))) +
                       hl(ld_nt(u2 + i), ld_nt(u2l + i));
    double nv = (2.0 * D(i) - hl(ld_nt(d2 + i), ld_nt(d2l + i))) + s2d(sp2[c]) * lu -
                s2d(vm2[c]) * d2u;
    if (SPONGE) nv = nv * (double)sponge[c];
    uint16_t lb;
    const float v = split_hl(nv, lb);
    st_nt(dn + i, v);
    st_nt(dnl + i, lb);
    for (int j = rec_head[c]; j >= 0; j = rec_next[j]) res[(int64_t)j * Bp + b] = v;
}

// ---------------------------------------------------------------------------
// Finite-aperture elements (footprints, fwi.py:74-89, 316-319): an element is a
// group of cells with weights.  Plans with such elements run the stencil
// kernels with no point source and no receivers; after each level these
// kernels add the weighted sources, record the receiver groups and correct the
// adjoint.  Lists of the current shot pass are padded to a fixed stride and
// selected by *pass (a device int set before each pass), so one captured
// graph serves every pass.
// ---------------------------------------------------------------------------

// coef[e] = S_c w_e = (dt^2 / m[cell_e]) w_e for every transmit entry (set up per
// call, outside the captured sequence, from the caller's model).
__global__ void k_fp_coef(const int *__restrict__ cell, const double *__restrict__ wt,
                          double *__restrict__ coef, int64_t count, const double *__restrict__ m,
                          double dt2) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    coef[e] = cell[e] >= 0 ? (dt2 / m[cell[e]]) * wt[e] : 0.0;
}

// Transmit entries of the pass (shot eb, cell ec, coefficient ecoef = S_c w_i;
// eb < 0 pads): level n gains ecoef wav[n-1], times the sponge, added in FP64 to
// the stored hi + lo (lap[cell] += wavelet[n-1] * w_i before p_new = 2p - p_prev
// + S lap, fwi.py:64-69).  Each (cell, shot) appears once per pass.
__global__ void k_fp_source(float *__restrict__ hi, uint16_t *__restrict__ lo,
                            const int *__restrict__ pass, int stride, const int *__restrict__ eb,
                            const int *__restrict__ ec, const double *__restrict__ ecoef,
                            const double *__restrict__ wav, int n,
                            const float *__restrict__ sponge, int Bp) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= stride) return;
    const int64_t q = (int64_t)(*pass) * stride + e;
    const int b = eb[q];
    if (b < 0) return;
    const int c = ec[q];
    const int64_t i = (int64_t)c * Bp + b;
    double add = ecoef[q] * wav[n - 1];
    if (sponge) add = add * (double)sponge[c];
    uint16_t l;
    const float h = split_hl(hl(hi[i], lo[i]) + add, l);
    hi[i] = h;
    lo[i] = l;
}

// Receiver group j of shot b at one level: sum_i w_i p[cell_i] in FP64, in
// footprint order (fwi.py:74-76), into out_hi[j * Bp + b] and, when out_lo is
// given, the FP32 remainder.  Thread per (j, b).
__global__ void k_fp_record(const float *__restrict__ hi, const uint16_t *__restrict__ lo,
                            const int *__restrict__ roff, const int *__restrict__ rcell,
                            const double *__restrict__ rw, float *__restrict__ out_hi,
                            float *__restrict__ out_lo, int n_rec, int Bp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_rec * Bp) return;
    const int j = (int)(t / Bp), b = (int)(t - (int64_t)j * Bp);
    double acc = 0.0;
    for (int k = roff[j]; k < roff[j + 1]; ++k) {
        const int64_t i = (int64_t)rcell[k] * Bp + b;
        acc = acc + rw[k] * hl(hi[i], lo[i]);
    }
    const float h = (float)acc;
    out_hi[t] = h;
    if (out_lo) out_lo[t] = (float)(acc - (double)h);
}

// After adjoint level k: lk holds lam'^k from the stencil kernel (sponge
// applied, no receiver injected) and the group accumulators hold lam'^k d2 with
// d2 the kernel's second difference (no sponge: S' lap(p^{k-1}) with no source
// term; sponge: p^k - 2p^{k-1} + p^{k-2} of the stored levels).  For each fix
// cell c of the pass and shot group g (one thread, shots in order):
//   lam^k[c] = lam'^k[c] - sum_j res[k][j] w_j   (receiver order, fwi.py:316-319)
//   g[g][c] += (lam^k - lam'^k) d2 + lam^k s     (s = S_c w_i wav[k-1] for a transmit
//                                                 cell of that shot; 0 with the sponge,
//                                                 whose d2 already holds the source)
// Lists per fix cell: receiver entries [frb, fre) of (rj, rw), transmit entries
// [fsb, fse) of (sb, scoef = S_c w_i).  fcell < 0 pads.
template <bool SPONGE>
__global__ void __launch_bounds__(256)
k_fp_adjoint(float *__restrict__ lk, const float *__restrict__ res, const float *__restrict__ pa,
             const float *__restrict__ pb, const float *__restrict__ pc2,
             const float2 *__restrict__ sp2, const int *__restrict__ pass, int stride,
             const int *__restrict__ fcell, const int *__restrict__ frb,
             const int *__restrict__ fre, const int *__restrict__ rj,
             const float *__restrict__ rw, const int *__restrict__ fsb,
             const int *__restrict__ fse, const int *__restrict__ sb,
             const double *__restrict__ scoef, const double *__restrict__ wav, int k,
             float *__restrict__ g_hi,
             float *__restrict__ g_lo, int LB, int ng, int nz, int ny, int nx, int ndim, int Bp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)stride * ng) return;
    const int f = (int)(t / ng), g = (int)(t - (int64_t)f * ng);
    const int64_t q = (int64_t)(*pass) * stride + f;
    const int c = fcell[q];
    if (c < 0) return;
    const int x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
    const bool in = ndim == 3 ? (z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2 && x >= 1 &&
                                 x <= nx - 2)
                              : (y >= 1 && y <= ny - 2 && x >= 1 && x <= nx - 2);
    const int64_t sx = Bp, sy = (int64_t)nx * Bp, sz = sy * ny;
    const int64_t gi = (int64_t)g * nz * ny * nx + c;
    float gh = g_hi[gi], gl = g_lo[gi];
    for (int bb = 0; bb < LB; ++bb) {
        const int b = g * LB + bb;
        const int64_t i = (int64_t)c * Bp + b;
        const float l0 = lk[i];
        float v = l0;
        for (int r = frb[q]; r < fre[q]; ++r) v = v - res[(int64_t)rj[r] * Bp + b] * rw[r];
        lk[i] = v;
        float d2, src = 0.0f;
        if (SPONGE) {
            d2 = (pa[i] - 2.0f * pb[i]) + pc2[i];
        } else {
            float lp = 0.0f;
            if (in)
                lp = ndim == 3 ? lap7(pa[i], pa[i + sz], pa[i - sz], pa[i + sy], pa[i - sy],
                                      pa[i + sx], pa[i - sx])
                               : lap5(pa[i], pa[i + sy], pa[i - sy], pa[i + sx], pa[i - sx]);
            d2 = smul(sp2[c], lp);
            for (int e = fsb[q]; e < fse[q]; ++e)
                if (sb[e] == b) src = src + (float)(scoef[e] * wav[k - 1]);
        }
        neumaier(gh, gl, (v - l0) * d2 + v * src);
    }
    g_hi[gi] = gh;
    g_lo[gi] = gl;
}

// ---------------------------------------------------------------------------
// Set-up and trace kernels
// ---------------------------------------------------------------------------
// vm2[c] = v[c] / m[c] as FP32 hi + lo (the Born source's scale).
__global__ void k_setup_vm(const double *__restrict__ v, const double *__restrict__ m,
                           float2 *__restrict__ vm2, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double s = v[i] / m[i];
    const float hi = (float)s;
    vm2[i] = make_float2(hi, (float)(s - (double)hi));
}

__global__ void k_setup_s(const double *__restrict__ m, float2 *__restrict__ sp2, int64_t N,
                          double dt2, double inv_h2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double s = (dt2 / m[i]) * inv_h2;
    const float hi = (float)s;
    sp2[i] = make_float2(hi, (float)(s - (double)hi));
}

// sinc[n][b] = S(src_b) * w[n-1] for n >= 1 (0 for padded shots and n = 0).
__global__ void k_setup_sinc(const double *__restrict__ m, const double *__restrict__ w,
                             const int *__restrict__ src_cell, float *__restrict__ sinc,
                             double *__restrict__ sincd, int nt, int Bp, double dt2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nt * Bp) return;
    const int n = (int)(i / Bp), b = (int)(i % Bp);
    const int s = src_cell[b];
    double v = 0.0;
    if (n >= 1 && s >= 0) v = (dt2 / m[s]) * w[n - 1];
    sinc[i] = (float)v;
    sincd[i] = v;
}

__global__ void k_d2f(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_fill(float *__restrict__ p, int64_t n, float v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_i32(int *__restrict__ p, int64_t n, int v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Receiver chains on the device (one thread, as the host builds them for a
// plan): head[c] = first receiver on cell c, next[j] = the following one, in
// ascending receiver order, so duplicate receivers accumulate in the order
// fwi.py:321-322 subtracts them.  head must hold -1 everywhere on entry.
// Also narrows the transmit cells to int (the source kernels' type).
__global__ void k_rec_chains(const int64_t *__restrict__ rec, int n_rec, int *__restrict__ head,
                             int *__restrict__ next, const int64_t *__restrict__ tx, int n_tx,
                             int *__restrict__ tx32) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int j = n_rec - 1; j >= 0; --j) {
        const int64_t c = rec[j];
        next[j] = head[c];
        head[c] = j;
    }
    for (int i = 0; i < n_tx; ++i) tx32[i] = (int)tx[i];
}

// Traces move between the caller's layout data[b][n][j] (FP64) and the
// batched layout [n][j][b] (FP32 hi + lo) through 32 x 32 LDS tiles, so both
// the reads and the writes are coalesced.  Grid: (ceil(nt*n_rec / 32),
// ceil(Bp / 32)), 256 threads.
constexpr int kTT = 32;

// dT[nj][b] + dT_lo[nj][b] = dobs[b0 + b][nj] (b < nb; padded shots get 0).
__global__ void __launch_bounds__(256)
k_dobs_to_batched(const double *__restrict__ dobs, float *__restrict__ dT,
                  float *__restrict__ dT_lo, int64_t NJ, int Bp, int b0, int nb) {
    __shared__ double tile[kTT][kTT + 1];
    const int64_t nj0 = (int64_t)blockIdx.x * kTT;
    const int bb0 = (int)blockIdx.y * kTT;
    const int tx = threadIdx.x & (kTT - 1), ty = threadIdx.x / kTT;
#pragma unroll
    for (int r = ty; r < kTT; r += 256 / kTT) {
        const int b = bb0 + r;
        const int64_t nj = nj0 + tx;
        tile[r][tx] = (b < nb && nj < NJ) ? dobs[(int64_t)(b0 + b) * NJ + nj] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = ty; r < kTT; r += 256 / kTT) {
        const int64_t nj = nj0 + r;
        const int b = bb0 + tx;
        if (nj < NJ && b < Bp) {
            const double v = tile[tx][r];
            const float h = (float)v;
            dT[nj * Bp + b] = h;
            dT_lo[nj * Bp + b] = (float)(v - (double)h);
        }
    }
}

// data[b0 + b][nj] = rec[nj][b] + rec_lo[nj][b] in FP64, for b < nb.
__global__ void __launch_bounds__(256)
k_batched_to_data(const float *__restrict__ rec, const float *__restrict__ rec_lo,
                  double *__restrict__ data, int64_t NJ, int Bp, int b0, int nb) {
    __shared__ double tile[kTT][kTT + 1];
    const int64_t nj0 = (int64_t)blockIdx.x * kTT;
    const int bb0 = (int)blockIdx.y * kTT;
    const int tx = threadIdx.x & (kTT - 1), ty = threadIdx.x / kTT;
#pragma unroll
    for (int r = ty; r < kTT; r += 256 / kTT) {
        const int64_t nj = nj0 + r;
        const int b = bb0 + tx;
        tile[r][tx] = (nj < NJ && b < Bp) ? (double)rec[nj * Bp + b] + (double)rec_lo[nj * Bp + b]
                                          : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = ty; r < kTT; r += 256 / kTT) {
        const int b = bb0 + r;
        const int64_t nj = nj0 + tx;
        if (b < nb && nj < NJ) data[(int64_t)(b0 + b) * NJ + nj] = tile[tx][r];
    }
}

// L2 residual r = dsyn - dobs over `count` entries (padded entries are 0 - 0),
// both as FP32 hi + lo pairs: r = (s_hi - o_hi) + (s_lo - o_lo); and
// per-block Neumaier partial sums of r^2 (fwi.py:155-157).
__global__ void __launch_bounds__(kBlock)
k_residual_l2(const float *__restrict__ dsyn, const float *__restrict__ dsyn_lo,
              const float *__restrict__ dobs, const float *__restrict__ dobs_lo,
              const float *__restrict__ W, float *__restrict__ res, int64_t count,
              float2 *__restrict__ part) {
    float s = 0.0f, c = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        float r = (dsyn[i] - dobs[i]) + (dsyn_lo[i] - dobs_lo[i]);
        // windowed (fwi.py:298-301): residual of W s against W o, adjoint source times W
        if (W) r = r * W[i];
        if (res) res[i] = W ? r * W[i] : r;
        neumaier(s, c, r * r);
    }
    __shared__ float sh[kBlock], ch[kBlock];
    sh[threadIdx.x] = s;
    ch[threadIdx.x] = c;
    __syncthreads();
    for (int off = kBlock / 2; off >= 1; off >>= 1) {
        if ((int)threadIdx.x < off) {
            float a = sh[threadIdx.x], ac = ch[threadIdx.x] + ch[threadIdx.x + off];
            neumaier(a, ac, sh[threadIdx.x + off]);
            sh[threadIdx.x] = a;
            ch[threadIdx.x] = ac;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_float2(sh[0], ch[0]);
}

// J = scale * sum of the partials (fixed order, FP64 at the end).
__global__ void k_reduce_partials(const float2 *__restrict__ part, int n, double scale,
                                  double *__restrict__ J, int accumulate) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.0f, c = 0.0f;
    for (int i = 0; i < n; ++i) {
        neumaier(s, c, part[i].x);
        c = c + part[i].y;
    }
    const double v = scale * ((double)s + (double)c);
    if (accumulate) *J += v;
    else *J = v;
}

// ---------------------------------------------------------------------------
// Graph-space optimal transport misfit (fwi.py:193-217), one work-group per
// trace (transmit t, receiver r).  Synthetic and observed samples, both divided
// by A = max|observed| + 1e-30, are points (tau_i, s_i) and (tau_j, o_j) with
// tau_i = i / nt; the assignment sigma minimising sum_i C(i, sigma(i)),
//   C(i, j) = (s_i - o_j)^2 + eta^2 (tau_i - tau_j)^2,
// gives J_trace = sum_i C(i, sigma(i)) and the adjoint source
// 2 (s_i - o_sigma(i)) / A.  sigma comes from the shortest-augmenting-path
// Hungarian method: rows enter one at a time, a Dijkstra search over columns
// on the reduced costs C - u - v finds the augmenting path, and the dual
// potentials u, v are updated along it.  Costs are evaluated on the fly; the
// column scan and the potential update run over the work-group's lanes, with
// ties broken on the smallest column so the result is deterministic.  Each
// search visits at most nt columns and each augmentation walks at most nt
// steps; non-finite costs, or a search that finds no column, set *status
// and end the trace's work (its J is NaN).  Indices below are 1-based with
// column 0 the search root, as the method is usually written.
// Workspace per trace: 5 (nt + 1) doubles (u, v, minv, s/A, o/A) and
// 3 (nt + 1) ints (p: row of each column, way, used).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double gsot_cost(const double *sh, const double *oh, int i, int j, int nt,
                                            double eta2) {
    const double ds = sh[i] - oh[j];
    const double dtau = (double)(i - 1) / nt - (double)(j - 1) / nt;
    return ds * ds + eta2 * (dtau * dtau);
}

__global__ void __launch_bounds__(256)
k_gsot(const double *__restrict__ dsyn, const double *__restrict__ dobs, int nt, int n_rx,
       int n_traces, double eta2, double *__restrict__ wsd, int *__restrict__ wsi,
       double *__restrict__ J_trace, double *__restrict__ adj, int *__restrict__ status) {
    const int tr = blockIdx.x;
    if (tr >= n_traces) return;  // uniform over the work-group
    const int t = tr / n_rx, r = tr % n_rx;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int n = nt;
    const size_t W = (size_t)n + 1;
    double *u = wsd + (size_t)tr * 5 * W, *v = u + W, *minv = v + W, *sh = minv + W, *oh = sh + W;
    int *p = wsi + (size_t)tr * 3 * W, *way = p + W, *used = way + W;
    const double *s = dsyn + (size_t)t * nt * n_rx + r;
    const double *o = dobs + (size_t)t * nt * n_rx + r;
    __shared__ double rv[256];
    __shared__ int ri[256];
    __shared__ int s_j0, s_bad;
    // A = max |o| + 1e-30
    double a = 0.0;
    for (int j = tid; j < n; j += nth) a = fmax(a, fabs(o[(size_t)j * n_rx]));
    rv[tid] = a;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k && tid + k < nth) rv[tid] = fmax(rv[tid], rv[tid + k]);
        __syncthreads();
    }
    const double A = rv[0] + 1e-30;
    for (int j = tid; j <= n; j += nth) {
        if (j >= 1) {
            sh[j] = s[(size_t)(j - 1) * n_rx] / A;
            oh[j] = o[(size_t)(j - 1) * n_rx] / A;
            if (!__builtin_isfinite(sh[j]) || !__builtin_isfinite(oh[j])) s_bad = 1;
        }
        u[j] = 0.0;
        v[j] = 0.0;
        p[j] = 0;
        way[j] = 0;
    }
    __syncthreads();
    for (int i = 1; i <= n && !s_bad; ++i) {
        for (int j = tid; j <= n; j += nth) {
            minv[j] = INFINITY;
            used[j] = 0;
        }
        if (tid == 0) {
            p[0] = i;
            s_j0 = 0;
        }
        __syncthreads();
        for (int step = 0; step <= n; ++step) {
            const int j0 = s_j0;
            if (tid == 0) used[j0] = 1;
            __syncthreads();
            const int i0 = p[j0];
            double best = INFINITY;
            int bj = 0;
            for (int j = 1 + tid; j <= n; j += nth) {
                if (used[j]) continue;
                const double cur = gsot_cost(sh, oh, i0, j, n, eta2) - u[i0] - v[j];
                if (cur < minv[j]) {
                    minv[j] = cur;
                    way[j] = j0;
                }
                if (minv[j] < best) {
                    best = minv[j];
                    bj = j;
                }
            }
            rv[tid] = best;
            ri[tid] = bj;
            __syncthreads();
            for (int k = 128; k > 0; k >>= 1) {
                if (tid < k && tid + k < nth) {
                    const double ov = rv[tid + k];
                    const int oi = ri[tid + k];
                    if (oi != 0 && (ri[tid] == 0 || ov < rv[tid] || (ov == rv[tid] && oi < ri[tid]))) {
                        rv[tid] = ov;
                        ri[tid] = oi;
                    }
                }
                __syncthreads();
            }
            const double delta = rv[0];
            const int j1 = ri[0];
            if (j1 == 0 || !__builtin_isfinite(delta)) {  // no reachable column: not a valid problem
                if (tid == 0) s_bad = 1;
                __syncthreads();
                break;
            }
            for (int j = tid; j <= n; j += nth) {
                if (used[j]) {
                    u[p[j]] += delta;
                    v[j] -= delta;
                } else {
                    minv[j] -= delta;
                }
            }
            __syncthreads();
            if (tid == 0) s_j0 = j1;
            __syncthreads();
            if (p[j1] == 0) break;  // free column reached
            if (step == n) {
                if (tid == 0) s_bad = 1;
                __syncthreads();
            }
        }
        __syncthreads();
        if (tid == 0 && !s_bad) {
            int j0 = s_j0;
            for (int k = 0; k <= n && j0 != 0; ++k) {
                const int j1 = way[j0];
                p[j0] = p[j1];
                j0 = j1;
            }
            if (j0 != 0) s_bad = 1;
        }
        __syncthreads();
    }
    if (s_bad) {
        if (tid == 0) {
            *status = 1;
            J_trace[tr] = NAN;
        }
        for (int i = tid; i < n; i += nth) adj[((size_t)t * nt + i) * n_rx + r] = 0.0;
        return;
    }
    // row -> column (way is free now), the trace's misfit and adjoint source
    for (int j = 1 + tid; j <= n; j += nth) way[p[j]] = j;
    __syncthreads();
    double jt = 0.0;
    for (int i = 1 + tid; i <= n; i += nth) {
        const int c = way[i];
        jt += gsot_cost(sh, oh, i, c, n, eta2);
        adj[((size_t)t * nt + (i - 1)) * n_rx + r] = 2.0 * (sh[i] - oh[c]) / A;
    }
    rv[tid] = jt;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k && tid + k < nth) rv[tid] += rv[tid + k];
        __syncthreads();
    }
    if (tid == 0) J_trace[tr] = rv[0];
}

// GCN per-trace statistics over time (fwi.py:160-164): for trace t = (b, j)
// stats[t] = (|s|^2, |o|^2, <s,o>) in FP64; layout of traces [n][j][b] with
// stride n_rec*Bp between time samples.
__global__ void k_gcn_stats(const float *__restrict__ dsyn, const float *__restrict__ dsyn_lo,
                            const float *__restrict__ dobs, const float *__restrict__ dobs_lo,
                            const float *__restrict__ W, double *__restrict__ stats, int nt,
                            int64_t ntr) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntr) return;
    double ss = 0.0, oo = 0.0, so = 0.0;
    for (int n = 0; n < nt; ++n) {
        const int64_t q = (int64_t)n * ntr + t;
        const double wq = W ? (double)W[q] : 1.0;
        const double s = ((double)dsyn[q] + (double)dsyn_lo[q]) * wq;
        const double o = ((double)dobs[q] + (double)dobs_lo[q]) * wq;
        ss += s * s;
        oo += o * o;
        so += s * o;
    }
    stats[3 * t] = ss;
    stats[3 * t + 1] = oo;
    stats[3 * t + 2] = so;
}

// GCN adjoint source res = -(o/no - c s/ns)/ns and per-trace misfit 1 - c
// (fwi.py:164-167); tmask[t] = 0 for padded traces.  With a window W the
// traces are W s and W o and the source is multiplied by W (fwi.py:298-301).
__global__ void k_gcn_residual(const float *__restrict__ dsyn, const float *__restrict__ dsyn_lo,
                               const float *__restrict__ dobs, const float *__restrict__ dobs_lo,
                               const float *__restrict__ W, const double *__restrict__ stats,
                               const unsigned char *__restrict__ tmask, float *__restrict__ res,
                               int nt, int64_t ntr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nt * ntr) return;
    const int64_t t = i % ntr;
    if (!tmask[t]) { res[i] = 0.0f; return; }
    const double eps = 1e-12;
    const double ns = sqrt(stats[3 * t]) + eps, no = sqrt(stats[3 * t + 1]) + eps;
    const double c = stats[3 * t + 2] / (ns * no);
    const double wq = W ? (double)W[i] : 1.0;
    const double s = ((double)dsyn[i] + (double)dsyn_lo[i]) * wq;
    const double o = ((double)dobs[i] + (double)dobs_lo[i]) * wq;
    res[i] = (float)((-(o / no - c * (s / ns)) / ns) * wq);
}

__global__ void k_gcn_misfit(const double *__restrict__ stats, const unsigned char *__restrict__ tmask,
                             int64_t ntr, double *__restrict__ J, int accumulate) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double eps = 1e-12;
    double s = 0.0, comp = 0.0;
    for (int64_t t = 0; t < ntr; ++t) {
        if (!tmask[t]) continue;
        const double ns = sqrt(stats[3 * t]) + eps, no = sqrt(stats[3 * t + 1]) + eps;
        const double x = 1.0 - stats[3 * t + 2] / (ns * no);
        const double y = s + x;
        comp += (fabs(s) >= fabs(x)) ? (s - y) + x : (x - y) + s;
        s = y;
    }
    if (accumulate) *J += s + comp;
    else *J = s + comp;
}

// grad[c] = (sum_g hi + lo) / m[c]   (fwi.py:331), FP64.
__global__ void k_finalize_grad(const float *__restrict__ g_hi, const float *__restrict__ g_lo,
                                int ng, int64_t N, const double *__restrict__ m,
                                double *__restrict__ grad, int accumulate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double s = 0.0;
    for (int g = 0; g < ng; ++g) s += (double)g_hi[(int64_t)g * N + i] + (double)g_lo[(int64_t)g * N + i];
    const double v = s / m[i];
    grad[i] = accumulate ? grad[i] + v : v;
}

}  // namespace usct

// usct.hip -- plans, orchestration and C ABI of libusct.so (gfx950).
//
// A plan fixes the grid, time axis, transmit and receiver cells and allocates
// the device workspace.  Two execution methods:
//
//  plain    one shot at a time, one kernel launch per time step, the full
//           forward history kept in HBM (what a direct port of the reference's
//           per-shot loop does: software/python/uap_gpu.py:120-140).
//  batched  every shot of a pass in each launch (field layout cell-major,
//           shot-minor).  Forward levels are checkpointed every K steps and
//           re-computed window by window during the adjoint sweep; when the
//           whole history fits the memory budget K covers the whole time axis
//           and nothing is re-computed.  The launch sequence of a pass is
//           captured once into a hipGraph and replayed on every call.
//
// Level bookkeeping of the batched gradient (levels n = -1..nt-1, p^-1 = p^0 = 0):
//   window s (s = 0..NS-1) serves adjoint steps k in [sK+1, min(sK+K, nt-1)]
//   and holds levels sK-1 .. sK+K in slots W[0..K+1] (W[i] = level sK-1+i).
//   Checkpoint pair of window s (1 <= s <= NS-2): levels (sK-1, sK).
//   The forward pass writes the checkpoint levels straight into their slots
//   and the last window's levels straight into W, so checkpointing adds no
//   copies; window s < NS-1 is re-computed from its pair before its adjoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/usct.h"
#include "usct_kernels.hpp"

using namespace usct;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPX(expr)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(USCT_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                             \
    } while (0)

#define TRY(expr)              \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// MODE_ADJSRC: the gradient for a caller-supplied adjoint source (staged into
// res per pass in place of the residual).
enum Mode { MODE_FWD = 0, MODE_MISFIT = 1, MODE_GRAD = 2, MODE_GNHV = 3, MODE_ADJSRC = 4 };

// Time levels per launch of the fused 2D kernels: two, or three for plans created
// with USCT_FLAG_LEVELS_3 (which need kLvMax + 2 rotation buffers).
constexpr int kLvMax = 3;
constexpr int kMaxRot = kLvMax + 2;

struct Tiling {
    int LB, CX, CY, RY;  // lanes per shot group, cells per lane (x, y), rows per wavefront
    int lds = 0;         // 3D: 1 = LDS-staged work-group tiles of CX x CY cells (patch per
                         // lane from lds_patch), rows = planes per work-group; 2 = the same
                         // tiles advancing two forward levels per launch; 3 = CX x CY lane
                         // patches advancing two forward levels per launch (k_fwd3d_x2);
                         // 4 (reversing adjoint, 3D forward) = lane patches in work-groups
                         // of kYsh strips stacked along y that share their edge rows
                         // (blk = 1 x kYsh)
    int sched = 0;       // 2D two-level kernels: bit 0 = odd row chunks march upwards,
                         // bit 1 = row chunk fastest in the wavefront order, bit 2 = shot
                         // group slowest
    int blk = kBlk3dDefault;  // 3D patch kernels: strip blocks of (blk >> 8) x (blk & 255)
    int lv = 2;               // 2D fused kernels: time levels per launch (2: k_fwd2d_x2 /
                              // k_adj2d_x2, 3: k_fwd2d_xn / k_adj2d_xn)
};

// Lane patch of an LDS tile: 256 lanes = LB shots x (TX/PX) x (TY/PY) patches.
inline void lds_patch(int LB, int TX, int TY, int &PX, int &PY) {
    const int cells = TX * TY * LB / 256;  // cells per lane
    PX = cells >= 4 ? 4 : 2;
    PY = cells / PX;
}

struct Profile {
    std::vector<hipEvent_t> fwd_ev, adj_ev;  // start/stop pairs
    std::vector<int> fwd_lv, adj_lv;         // time levels advanced by each timed launch
};

}  // namespace

struct usct_plan {
    // problem
    int ndim = 2, nz = 1, ny = 0, nx = 0;
    int64_t N = 0;
    double h = 0, dt = 0;
    int nt = 0, n_tx = 0, n_rec = 0;
    std::vector<int64_t> tx, rec;
    int method = USCT_METHOD_BATCHED, flags = 0;
    bool sponge = false;
    // kernel configuration
    // Tiles of the forward and of the adjoint kernels (chosen independently: the
    // field layout does not depend on them; the gradient accumulators use the
    // adjoint's lane groups).
    // Defaults (plans without the autotuner: serving, tests) are the shapes the
    // tuner picks on the 512^2 / 256^3 workloads: 2D forward 8 and adjoint 4
    // cells per lane, 3D 2 x 2 lane patches.
    Tiling tf{64, 8, 1, 32, 0}, ta{64, 4, 1, 32, 0};
    int Bp = 1;        // padded shots per pass
    int passes = 1;
    int K = 0, NS = 1;  // checkpoint interval and number of windows
    // 3D gradients without stored levels: the adjoint sweep runs the forward
    // backwards alongside it (k_radj3d); Wslot then holds three adjoint levels
    bool rev = false;
    int64_t budget = 0;
    // device memory: blocks (pointer, bytes held) and the bytes the plan asked for
    std::vector<std::pair<void *, int64_t>> allocs;
    int64_t bytes = 0, held = 0;
    int dev = 0;  // the device current at creation: the plan's memory and stream live there
    float2 *sp2 = nullptr;
    float *spg = nullptr;  // sponge (FP32)
    int *rec_head = nullptr, *rec_next = nullptr;
    int *row_rec = nullptr;  // receivers per grid row (2D: ny rows, 3D: nz*ny rows)
    int *src_all = nullptr;  // [passes][Bp] source cells (-1 = padding)
    int *src_cell = nullptr; // [Bp] cells of the current pass
    float *sinc = nullptr;   // [nt][Bp] S_src w[n-1], FP32 (2D kernels)
    double *sincd = nullptr; // [nt][Bp] the same in FP64 (3D kernels)
    float *srcd2 = nullptr;  // [nt][Bp] recorded increment at each shot's source cell
    float *Z = nullptr;      // zero field
    // rotation of levels that are not kept: a launch reads levels n-1, n-2 and writes
    // up to n+LV-1, so nrot >= LV + 2 buffers (4, or 5 where three-level launches run)
    int nrot = 4;
    float *F[kMaxRot] = {};
    std::vector<float *> Wslot, CK;
    // bfloat16 low parts of the forward state: rotation, zero level, checkpoints
    uint16_t *Lo[kMaxRot] = {};
    uint16_t *Zlo = nullptr;
    std::vector<uint16_t *> CKlo;
    // USCT_FLAG_FP32_LO: the low parts are FP32 (two uint16_t slots per element
    // in the buffers above, passed to the kernels as float)
    bool lo32 = false;
    int64_t lo_units() const { return lo32 ? 2 : 1; }
    float *recb = nullptr, *recb_lo = nullptr, *res = nullptr;
    float *dobsT = nullptr, *dobsT_lo = nullptr;  // observed traces, FP32 hi + lo
    float2 *part = nullptr;
    int npart = 1024;
    float *g_hi = nullptr, *g_lo = nullptr;
    double *J_acc = nullptr;
    double *gstats = nullptr;
    // Gauss-Newton: Born state over three levels (FP32 hi + bfloat16 lo) and
    // v/m, allocated by the first Hessian-vector product
    float *Dh[4] = {nullptr, nullptr, nullptr, nullptr};
    uint16_t *Dl[4] = {nullptr, nullptr, nullptr, nullptr};
    float2 *vm2 = nullptr;
    // Finite-aperture elements (groups of cells with weights).  Per-pass lists
    // are padded to fp_SE transmit entries and fp_FC fix cells and selected by
    // *fp_pass.
    bool fp = false;
    int fp_SE = 0, fp_FC = 0;
    int *fp_pass = nullptr, *fp_pass_ids = nullptr;
    int *fp_eb = nullptr, *fp_ec = nullptr;   // transmit entries: shot, cell
    double *fp_ew = nullptr, *fp_ecoef = nullptr;  // weight, S_c * weight
    int *fp_roff = nullptr, *fp_rcell = nullptr;  // receiver groups (CSR)
    double *fp_rw = nullptr;
    int *fp_fcell = nullptr, *fp_frb = nullptr, *fp_fre = nullptr, *fp_fsb = nullptr,
        *fp_fse = nullptr;                     // fix cells and their entry ranges
    int *fp_rj = nullptr, *fp_sb = nullptr, *fp_sc = nullptr;  // fix entries: receiver;
                                                                // shot and cell
    float *fp_rwf = nullptr;                   // receiver weight (FP32)
    double *fp_sw = nullptr, *fp_scoef = nullptr;  // transmit weight, S_c * weight
    int64_t fp_nse = 0;                        // fix transmit entries over all passes
    double *wav = nullptr;                     // the call's wavelet (nt)
    // time windows of the current pass, batched layout (allocated on first use)
    float *twT = nullptr, *twT_lo = nullptr;
    unsigned char *tmask = nullptr;  // [passes][n_rec*Bp]
    unsigned char *tmask_cur = nullptr;
    // execution
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::map<int, hipGraphExec_t> graphs;  // key: mode*4 + misfit

    int64_t field_elems() const { return N * (int64_t)Bp; }
    int ng_adj() const { return Bp / ta.LB; }  // shot groups of the gradient accumulators
};

namespace {

// Device memory of destroyed plans kept for later plans (usct_device_cache): a
// serving loop creates a plan per request, and hipFree waits for the whole
// device while a first hipMalloc of many GB maps its pages.  Kept blocks are
// 2 MiB multiples; a request takes the smallest kept block of its device that
// is at least its size and at most 1.25 times it.
constexpr int64_t kCacheGrain = (int64_t)2 << 20;
struct KeptBlock {
    void *p;
    int64_t bytes;
    int dev;
};
std::mutex g_cache_mu;
std::vector<KeptBlock> g_kept;
int64_t g_cache_cap = 0, g_kept_bytes = 0;

void free_block(const KeptBlock &b) {
    int cur = 0;
    const bool other = hipGetDevice(&cur) == hipSuccess && cur != b.dev;
    if (other) (void)hipSetDevice(b.dev);
    (void)hipFree(b.p);
    if (other) (void)hipSetDevice(cur);
}

// Free kept blocks (of device dev, or of every device when dev < 0) until at most
// `keep` bytes are kept; returns the bytes freed.  Caller holds g_cache_mu.
int64_t release_kept(int dev, int64_t keep) {
    int64_t freed = 0;
    for (size_t i = g_kept.size(); i-- > 0 && g_kept_bytes > keep;) {
        if (dev >= 0 && g_kept[i].dev != dev) continue;
        free_block(g_kept[i]);
        g_kept_bytes -= g_kept[i].bytes;
        freed += g_kept[i].bytes;
        g_kept.erase(g_kept.begin() + (ptrdiff_t)i);
    }
    return freed;
}

int64_t kept_on(int dev) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    int64_t b = 0;
    for (const KeptBlock &k : g_kept) b += k.dev == dev ? k.bytes : 0;
    return b;
}

int dalloc(usct_plan *P, void **p, int64_t bytes) {
    if (bytes <= 0) bytes = 16;
    const int dev = P->dev;
    int cur = dev;
    (void)hipGetDevice(&cur);
    if (cur != dev) (void)hipSetDevice(dev);
    struct Restore {
        int cur, dev;
        ~Restore() {
            if (cur != dev) (void)hipSetDevice(cur);
        }
    } restore{cur, dev};
    int64_t size = bytes;
    *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        if (g_cache_cap > 0) {
            size = cdiv(bytes, kCacheGrain) * kCacheGrain;
            int best = -1;
            for (size_t i = 0; i < g_kept.size(); ++i) {
                const KeptBlock &k = g_kept[i];
                if (k.dev == dev && k.bytes >= size && k.bytes <= size + size / 4 &&
                    (best < 0 || k.bytes < g_kept[best].bytes))
                    best = (int)i;
            }
            if (best >= 0) {
                *p = g_kept[best].p;
                size = g_kept[best].bytes;
                g_kept_bytes -= size;
                g_kept.erase(g_kept.begin() + best);
            }
        }
    }
    if (!*p) {
        hipError_t e = hipMalloc(p, (size_t)size);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            int64_t freed = 0;
            {
                std::lock_guard<std::mutex> lk(g_cache_mu);
                freed = release_kept(dev, 0);
            }
            if (freed > 0) e = hipMalloc(p, (size_t)size);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            return fail(USCT_ENOMEM, "hipMalloc(%lld bytes) failed: %s", (long long)size,
                        hipGetErrorString(e));
        }
    }
    P->allocs.emplace_back(*p, size);
    P->bytes += bytes;
    P->held += size;
    return 0;
}

template <class T>
int dalloc_t(usct_plan *P, T **p, int64_t count) {
    return dalloc(P, (void **)p, count * (int64_t)sizeof(T));
}

// Stencil operands of the batched kernels are read with unclamped columns: the
// x halo of a strip may fall up to kGuardCells cells (times the shot stride)
// before the first or after the last element.  Such buffers carry a zeroed
// guard zone on both sides; the returned pointer is the first element.
constexpr int64_t kGuardCells = 64;

template <class T>
int dalloc_guarded(usct_plan *P, T **p, int64_t count, int64_t stride) {
    const int64_t g = kGuardCells * std::max<int64_t>(stride, 1);
    T *base = nullptr;
    TRY(dalloc_t(P, &base, count + 2 * g));
    HIPX(hipMemset(base, 0, sizeof(T) * g));
    HIPX(hipMemset(base + g + count, 0, sizeof(T) * g));
    *p = base + g;
    return 0;
}

void free_plan_memory(usct_plan *P) {
    for (auto &kv : P->graphs) (void)hipGraphExecDestroy(kv.second);
    P->graphs.clear();
    // blocks go back to the cache only once no work of this plan can touch them
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        for (auto &a : P->allocs) {
            if (g_cache_cap > 0 && g_kept_bytes + a.second <= g_cache_cap) {
                g_kept.push_back({a.first, a.second, P->dev});
                g_kept_bytes += a.second;
            } else {
                free_block({a.first, a.second, P->dev});
            }
        }
    }
    P->allocs.clear();
    P->held = 0;
    P->Wslot.clear();
    P->CK.clear();
    P->CKlo.clear();
    for (int i = 0; i < 4; ++i) {
        P->Dh[i] = nullptr;
        P->Dl[i] = nullptr;
    }
    P->vm2 = nullptr;
    P->twT = P->twT_lo = nullptr;
    P->bytes = 0;
}

inline unsigned grid_for(int64_t n, int block = 256) {
    return (unsigned)std::max<int64_t>(1, cdiv(n, block));
}

// ---------------------------------------------------------------------------
// Kernel dispatch (template instantiation tables)
// ---------------------------------------------------------------------------
struct Tile {
    int nxs, nys, nyc;  // 2D: nxs, nyc; 3D: nxs, nys, nyc = z chunks
    int64_t waves;
};

Tile tile_of(const usct_plan *P, const Tiling &T) {
    Tile t{};
    const int W = kWave / T.LB;
    const int64_t ng = P->Bp / T.LB;
    t.nxs = (int)cdiv(P->nx, (int64_t)W * T.CX);
    if (P->ndim == 2) {
        t.nys = 1;
        t.nyc = (int)cdiv(P->ny, T.RY);
        t.waves = ng * t.nxs * t.nyc;
    } else {
        t.nys = (int)cdiv(P->ny, T.CY);
        t.nyc = (int)cdiv(P->nz, T.RY);
        t.waves = ng * tasks3d(t.nxs, t.nys, t.nyc, T.blk);
    }
    return t;
}

// A forward level: FP32 high part and bfloat16 low part.
struct Lvl {
    float *hi;
    uint16_t *lo;
};

template <int LB, int CX, bool SP, bool REC>
void fwd2d_t(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pn, const usct_plan *P,
             const float *sinc, float *rec, float *rec_lo, const Tile &t) {
    const double *sd = P->sincd + (sinc - P->sinc);
    auto go = [&](auto lo_type) {
        using L = decltype(lo_type);
        hipLaunchKernelGGL((k_fwd2d<LB, CX, SP, REC, L>), g, dim3(kBlock), 0, s, pc.hi, pp.hi,
                           pn.hi, (const L *)pc.lo, (const L *)pp.lo, (L *)pn.lo, P->sp2, P->spg,
                           P->src_cell, sd, P->rec_head, P->rec_next, P->row_rec, rec, rec_lo,
                           P->ny, P->nx, P->Bp, P->tf.RY, t.nxs, t.nyc, P->Bp / P->tf.LB);
    };
    if (P->lo32) go(float{});
    else go(uint16_t{});
}

template <int LB, bool SP, bool REC>
void fwd2d_lb(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pn, const usct_plan *P,
              const float *sinc, float *rec, float *rec_lo, const Tile &t) {
    if (P->tf.CX == 4) fwd2d_t<LB, 4, SP, REC>(g, s, pc, pp, pn, P, sinc, rec, rec_lo, t);
    else fwd2d_t<LB, 8, SP, REC>(g, s, pc, pp, pn, P, sinc, rec, rec_lo, t);
}

template <bool SP, bool REC>
void fwd2d_sel(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pn, const usct_plan *P,
               const float *sinc, float *rec, float *rl, const Tile &t) {
    switch (P->tf.LB) {
        case 16: fwd2d_lb<16, SP, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t); break;
        case 32: fwd2d_lb<32, SP, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t); break;
        default: fwd2d_lb<64, SP, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t); break;
    }
}

template <int LB, int CX, int CY, bool REC>
void fwd3d_t(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pn, const usct_plan *P,
             const float *sinc, float *rec, float *rl, const Tile &t) {
    float *d2 = P->srcd2 + (sinc - P->sinc);
    const double *sd = P->sincd + (sinc - P->sinc);
    auto go = [&](auto lo_type, auto ysh) {
        using L = decltype(lo_type);
        hipLaunchKernelGGL((k_fwd3d<LB, CX, CY, REC, L, decltype(ysh)::value>), g, dim3(kBlock), 0,
                           s, pc.hi, pp.hi, pn.hi, (const L *)pc.lo, (const L *)pp.lo, (L *)pn.lo,
                           P->sp2, P->src_cell, sd, P->rec_head, P->rec_next, P->row_rec, rec, rl,
                           d2, P->nz, P->ny, P->nx, P->Bp, P->tf.RY, t.nxs, t.nys, t.nyc,
                           P->Bp / P->tf.LB, P->tf.blk);
    };
    // tile kind 4: work-groups of y-stacked strips sharing their edge rows (built
    // for 64-lane groups with 2 x 2 patches)
    if constexpr (LB == 64 && CX == 2 && CY == 2) {
        if (P->tf.lds == 4) {
            if (P->lo32) go(float{}, std::true_type{});
            else go(uint16_t{}, std::true_type{});
            return;
        }
    }
    if (P->lo32) go(float{}, std::false_type{});
    else go(uint16_t{}, std::false_type{});
}

template <int LB, bool REC>
void fwd3d_lb(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pn, const usct_plan *P,
              const float *sinc, float *rec, float *rl, const Tile &t) {
    if (P->tf.CX == 4 && P->tf.CY == 4) fwd3d_t<LB, 4, 4, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
    else if (P->tf.CX == 8 && P->tf.CY == 2) fwd3d_t<LB, 8, 2, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
    else if (P->tf.CX == 2 && P->tf.CY == 2) fwd3d_t<LB, 2, 2, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
    else if (P->tf.CX == 2 && P->tf.CY == 4) fwd3d_t<LB, 2, 4, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
    else if (P->tf.CX == 2 && P->tf.CY == 1) fwd3d_t<LB, 2, 1, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
    else if (P->tf.CX == 1 && P->tf.CY == 2) fwd3d_t<LB, 1, 2, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
    else if (P->tf.CX == 1 && P->tf.CY == 1) fwd3d_t<LB, 1, 1, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
    else fwd3d_t<LB, 4, 2, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
}

template <int LB, int TX, int TY, int PX, int PY, bool REC>
void fwd3d_lds_t(hipStream_t s, Lvl pc, Lvl pp, Lvl pn, const usct_plan *P, const float *sinc,
                 float *rec, float *rl) {
    const Tiling &T = P->tf;
    const int ntx = (int)cdiv(P->nx, TX), nty = (int)cdiv(P->ny, TY);
    const int nzc = (int)cdiv(P->nz, T.RY), ng = P->Bp / LB;
    const dim3 g((unsigned)((int64_t)ng * ntx * nty * nzc));
    float *d2 = P->srcd2 + (sinc - P->sinc);
    const double *sd = P->sincd + (sinc - P->sinc);
    hipLaunchKernelGGL((k_fwd3d_lds<LB, TX, TY, PX, PY, REC>), g, dim3(kBlock), 0, s, pc.hi, pp.hi,
                       pn.hi, pc.lo, pp.lo, pn.lo, P->sp2, P->src_cell, sd, P->rec_head,
                       P->rec_next, P->row_rec, rec, rl, d2, P->nz, P->ny, P->nx, P->Bp, T.RY, ntx,
                       nty, nzc, ng);
}

// LDS tile shapes (the autotuner's candidates): LB 16 -> 8x8 or 16x8, LB 32 -> 8x8,
// LB 64 -> 8x4 cells per work-group.
template <bool REC>
void fwd3d_lds_sel(hipStream_t s, Lvl pc, Lvl pp, Lvl pn, const usct_plan *P, const float *sinc,
                   float *rec, float *rl) {
    const Tiling &T = P->tf;
    if (T.LB == 16 && T.CX == 16) fwd3d_lds_t<16, 16, 8, 4, 2, REC>(s, pc, pp, pn, P, sinc, rec, rl);
    else if (T.LB == 16) fwd3d_lds_t<16, 8, 8, 2, 2, REC>(s, pc, pp, pn, P, sinc, rec, rl);
    else if (T.LB == 32) fwd3d_lds_t<32, 8, 8, 4, 2, REC>(s, pc, pp, pn, P, sinc, rec, rl);
    else fwd3d_lds_t<64, 8, 4, 4, 2, REC>(s, pc, pp, pn, P, sinc, rec, rl);
}

template <bool REC>
void fwd3d_sel(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pn, const usct_plan *P,
               const float *sinc, float *rec, float *rl, const Tile &t) {
    switch (P->tf.LB) {
        case 16: fwd3d_lb<16, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t); break;
        case 32: fwd3d_lb<32, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t); break;
        default: fwd3d_lb<64, REC>(g, s, pc, pp, pn, P, sinc, rec, rl, t); break;
    }
}

void launch_fwd_batched(const usct_plan *P, hipStream_t s, Lvl pc, Lvl pp, Lvl pn, int n,
                        bool record) {
    const Tile t = tile_of(P, P->tf);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
    const float *sinc = P->sinc + (int64_t)n * P->Bp;
    float *rec = P->recb + (int64_t)n * P->n_rec * P->Bp;
    float *rl = P->recb_lo + (int64_t)n * P->n_rec * P->Bp;
    if (P->ndim == 2) {
        if (P->sponge) {
            if (record) fwd2d_sel<true, true>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
            else fwd2d_sel<true, false>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
        } else {
            if (record) fwd2d_sel<false, true>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
            else fwd2d_sel<false, false>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
        }
    } else {
        if (P->tf.lds == 1 || P->tf.lds == 2) {
            if (record) fwd3d_lds_sel<true>(s, pc, pp, pn, P, sinc, rec, rl);
            else fwd3d_lds_sel<false>(s, pc, pp, pn, P, sinc, rec, rl);
        } else {
            if (record) fwd3d_sel<true>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
            else fwd3d_sel<false>(g, s, pc, pp, pn, P, sinc, rec, rl, t);
        }
    }
}

// Two forward levels per launch (2D, no sponge): levels n and n+1.
template <int LB, int CX, bool REC>
void fwd2d_x2_t(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pa, Lvl pb, const usct_plan *P,
                int n, const Tile &t) {
    const double *sa = P->sincd + (int64_t)n * P->Bp;
    const int64_t tr = (int64_t)P->n_rec * P->Bp;
    float *ra = P->recb + (int64_t)n * tr, *rla = P->recb_lo + (int64_t)n * tr;
    auto go = [&](auto lo_type) {
        using L = decltype(lo_type);
        hipLaunchKernelGGL((k_fwd2d_x2<LB, CX, REC, L>), g, dim3(kBlock), 0, s, pc.hi, pp.hi,
                           pa.hi, pb.hi, (const L *)pc.lo, (const L *)pp.lo, (L *)pa.lo,
                           (L *)pb.lo, P->sp2, P->src_cell, sa, sa + P->Bp, P->rec_head,
                           P->rec_next, P->row_rec, ra, rla, ra + tr, rla + tr, P->ny, P->nx,
                           P->Bp, P->tf.RY, t.nxs, t.nyc, P->Bp / P->tf.LB, P->tf.sched);
    };
    if (P->lo32) go(float{});
    else go(uint16_t{});
}

template <int LB, bool REC>
void fwd2d_x2_lb(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pa, Lvl pb, const usct_plan *P,
                 int n, const Tile &t) {
    if (P->tf.CX == 4) fwd2d_x2_t<LB, 4, REC>(g, s, pc, pp, pa, pb, P, n, t);
    else fwd2d_x2_t<LB, 8, REC>(g, s, pc, pp, pa, pb, P, n, t);
}

// Two 3D forward levels per launch on 8 x 8 LDS tiles of 16 shots.
template <bool REC>
void fwd3d_x2_lds_t(hipStream_t s, Lvl pc, Lvl pp, Lvl pa, Lvl pb, const usct_plan *P, int n) {
    constexpr int LB = 16, TX = 8, TY = 8;
    const Tiling &T = P->tf;
    const int ntx = (int)cdiv(P->nx, TX), nty = (int)cdiv(P->ny, TY);
    const int nzc = (int)cdiv(P->nz, T.RY), ng = P->Bp / LB;
    const dim3 g((unsigned)((int64_t)ng * ntx * nty * nzc));
    const double *sa = P->sincd + (int64_t)n * P->Bp;
    float *da = P->srcd2 + (int64_t)n * P->Bp;
    const int64_t tr = (int64_t)P->n_rec * P->Bp;
    float *ra = P->recb + (int64_t)n * tr, *rla = P->recb_lo + (int64_t)n * tr;
    hipLaunchKernelGGL((k_fwd3d_x2_lds<LB, TX, TY, 2, 2, REC>), g, dim3(kBlock), 0, s, pc.hi,
                       pp.hi, pa.hi, pb.hi, pc.lo, pp.lo, pa.lo, pb.lo, P->sp2, P->src_cell, sa,
                       sa + P->Bp, P->rec_head, P->rec_next, P->row_rec, ra, rla, ra + tr,
                       rla + tr, da, da + P->Bp, P->nz, P->ny, P->nx, P->Bp, T.RY, ntx, nty, nzc,
                       ng);
}

// Two 3D forward levels per launch on CX x CY lane patches (register z-march).
template <int LB, int CX, int CY, bool REC>
void fwd3d_x2_t(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pa, Lvl pb, const usct_plan *P, int n,
                const Tile &t) {
    const double *sa = P->sincd + (int64_t)n * P->Bp;
    float *da = P->srcd2 + (int64_t)n * P->Bp;
    const int64_t tr = (int64_t)P->n_rec * P->Bp;
    float *ra = P->recb + (int64_t)n * tr, *rla = P->recb_lo + (int64_t)n * tr;
    hipLaunchKernelGGL((k_fwd3d_x2<LB, CX, CY, REC>), g, dim3(kBlock), 0, s, pc.hi, pp.hi, pa.hi,
                       pb.hi, pc.lo, pp.lo, pa.lo, pb.lo, P->sp2, P->src_cell, sa, sa + P->Bp,
                       P->rec_head, P->rec_next, P->row_rec, ra, rla, ra + tr, rla + tr, da,
                       da + P->Bp, P->nz, P->ny, P->nx, P->Bp, P->tf.RY, t.nxs, t.nys, t.nyc,
                       P->Bp / P->tf.LB, P->tf.blk);
}

template <int LB, bool REC>
void fwd3d_x2_lb(dim3 g, hipStream_t s, Lvl pc, Lvl pp, Lvl pa, Lvl pb, const usct_plan *P, int n,
                 const Tile &t) {
    if (P->tf.CX == 2 && P->tf.CY == 1) fwd3d_x2_t<LB, 2, 1, REC>(g, s, pc, pp, pa, pb, P, n, t);
    else if (P->tf.CX == 1 && P->tf.CY == 2) fwd3d_x2_t<LB, 1, 2, REC>(g, s, pc, pp, pa, pb, P, n, t);
    else fwd3d_x2_t<LB, 1, 1, REC>(g, s, pc, pp, pa, pb, P, n, t);
}

void launch_fwd2_batched(const usct_plan *P, hipStream_t s, Lvl pc, Lvl pp, Lvl pa, Lvl pb,
                         int n, bool record) {
    if (P->ndim == 3 && P->tf.lds == 3) {
        const Tile t = tile_of(P, P->tf);
        const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
#define USCT_X23(LB_)                                                         \
    if (record) fwd3d_x2_lb<LB_, true>(g, s, pc, pp, pa, pb, P, n, t);        \
    else fwd3d_x2_lb<LB_, false>(g, s, pc, pp, pa, pb, P, n, t);
        switch (P->tf.LB) {
            case 16: USCT_X23(16) break;
            case 32: USCT_X23(32) break;
            default: USCT_X23(64) break;
        }
#undef USCT_X23
        return;
    }
    if (P->ndim == 3) {
        if (record) fwd3d_x2_lds_t<true>(s, pc, pp, pa, pb, P, n);
        else fwd3d_x2_lds_t<false>(s, pc, pp, pa, pb, P, n);
        return;
    }
    const Tile t = tile_of(P, P->tf);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
#define USCT_X2(LB_)                                                          \
    if (record) fwd2d_x2_lb<LB_, true>(g, s, pc, pp, pa, pb, P, n, t);        \
    else fwd2d_x2_lb<LB_, false>(g, s, pc, pp, pa, pb, P, n, t);
    switch (P->tf.LB) {
        case 16: USCT_X2(16) break;
        case 32: USCT_X2(32) break;
        default: USCT_X2(64) break;
    }
#undef USCT_X2
}

// LV forward levels per launch (2D, no sponge): levels n .. n+LV-1 go to
// out[0 .. LV-1]; a null hi or lo is not stored.
template <int LB, int CX, int LV, bool REC>
void fwd2d_xn_t(dim3 g, hipStream_t s, Lvl pc, Lvl pp, const Lvl *out, const usct_plan *P, int n,
                const Tile &t) {
    const int64_t tr = (int64_t)P->n_rec * P->Bp;
    auto go = [&](auto lo_type) {
        using L = decltype(lo_type);
        FwdLevels<LV, L> o{};
        for (int i = 0; i < LV; ++i) {
            o.hi[i] = out[i].hi;
            o.lo[i] = (L *)out[i].lo;
            o.sincd[i] = P->sincd + (int64_t)(n + i) * P->Bp;
            o.rec[i] = P->recb + (int64_t)(n + i) * tr;
            o.rec_lo[i] = P->recb_lo + (int64_t)(n + i) * tr;
        }
        hipLaunchKernelGGL((k_fwd2d_xn<LB, CX, LV, REC, L>), g, dim3(kBlock), 0, s, pc.hi, pp.hi,
                           (const L *)pc.lo, (const L *)pp.lo, o, P->sp2, P->src_cell,
                           P->rec_head, P->rec_next, P->row_rec, P->ny, P->nx, P->Bp, P->tf.RY,
                           t.nxs, t.nyc, P->Bp / P->tf.LB, P->tf.sched);
    };
    if (P->lo32) go(float{});
    else go(uint16_t{});
}

// Three forward levels per launch: k_fwd2d_xn with four cells per lane (eight
// spill at two wavefronts per SIMD).
void launch_fwd3_batched(const usct_plan *P, hipStream_t s, Lvl pc, Lvl pp, const Lvl *out, int n,
                         bool record) {
    const Tile t = tile_of(P, P->tf);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
#define USCT_X3(LB_)                                                           \
    if (record) fwd2d_xn_t<LB_, 4, 3, true>(g, s, pc, pp, out, P, n, t);       \
    else fwd2d_xn_t<LB_, 4, 3, false>(g, s, pc, pp, out, P, n, t);
    switch (P->tf.LB) {
        case 16: USCT_X3(16) break;
        case 32: USCT_X3(32) break;
        default: USCT_X3(64) break;
    }
#undef USCT_X3
}

// Whether the plan's forward / adjoint launches advance two time levels: 2D both
// (no sponge), 3D the forward when its tile is the two-level LDS kind.  Footprint
// plans correct every level after its launch, so one level per launch.
bool fusable(const usct_plan *P) {
    return P->method == USCT_METHOD_BATCHED && !P->sponge && !P->fp &&
           !(P->flags & USCT_FLAG_NO_FUSE);
}
bool fused_fwd(const usct_plan *P) {
    return fusable(P) && (P->ndim == 2 || P->tf.lds == 2 || P->tf.lds == 3);
}
bool fused_adj(const usct_plan *P) { return fusable(P) && P->ndim == 2; }

template <int LB, int CX, bool SP>
void adj2d_t(dim3 g, hipStream_t s, const float *l1, const float *l2, float *lk,
             const float *pa, const float *pb, const float *pc2, const usct_plan *P,
             const float *sinc, const float *res, const Tile &t) {
    const float *d2 = P->srcd2 + (sinc - P->sinc);
    hipLaunchKernelGGL((k_adj2d<LB, CX, SP>), g, dim3(kBlock), 0, s, l1, l2, lk, pa, pb, pc2,
                       P->sp2, P->spg, P->src_cell, sinc, P->rec_head, P->rec_next, P->row_rec,
                       res, d2, P->g_hi, P->g_lo, P->ny, P->nx, P->Bp, P->ta.RY, t.nxs, t.nyc,
                       P->ng_adj());
}

template <int LB, bool SP>
void adj2d_lb(dim3 g, hipStream_t s, const float *l1, const float *l2, float *lk,
              const float *pa, const float *pb, const float *pc2, const usct_plan *P,
              const float *sinc, const float *res, const Tile &t) {
    if (P->ta.CX == 4) adj2d_t<LB, 4, SP>(g, s, l1, l2, lk, pa, pb, pc2, P, sinc, res, t);
    else adj2d_t<LB, 8, SP>(g, s, l1, l2, lk, pa, pb, pc2, P, sinc, res, t);
}

template <bool SP>
void adj2d_sel(dim3 g, hipStream_t s, const float *l1, const float *l2, float *lk,
               const float *pa, const float *pb, const float *pc2, const usct_plan *P,
               const float *sinc, const float *res, const Tile &t) {
    switch (P->ta.LB) {
        case 16: adj2d_lb<16, SP>(g, s, l1, l2, lk, pa, pb, pc2, P, sinc, res, t); break;
        case 32: adj2d_lb<32, SP>(g, s, l1, l2, lk, pa, pb, pc2, P, sinc, res, t); break;
        default: adj2d_lb<64, SP>(g, s, l1, l2, lk, pa, pb, pc2, P, sinc, res, t); break;
    }
}

// Two adjoint levels per launch (2D, no sponge): lam^k and lam^{k-1}.
template <int LB, int CX>
void adj2d_x2_t(dim3 g, hipStream_t s, const float *l1, const float *l2, float *lu, float *lv,
                const float *pa1, const float *pa2, const usct_plan *P, int k, const Tile &t) {
    const float *su = P->sinc + (int64_t)k * P->Bp;
    const int64_t tr = (int64_t)P->n_rec * P->Bp;
    const float *ru = P->res + (int64_t)k * tr;
    hipLaunchKernelGGL((k_adj2d_x2<LB, CX>), g, dim3(kBlock), 0, s, l1, l2, lu, lv, pa1, pa2,
                       P->sp2, P->src_cell, su, su - P->Bp, P->rec_head, P->rec_next, P->row_rec,
                       ru, ru - tr, P->g_hi, P->g_lo, P->ny, P->nx, P->Bp, P->ta.RY, t.nxs, t.nyc,
                       P->ng_adj(), P->ta.sched);
}

void launch_adj2_batched(const usct_plan *P, hipStream_t s, const float *l1, const float *l2,
                         float *lu, float *lv, const float *pa1, const float *pa2, int k) {
    const Tile t = tile_of(P, P->ta);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
#define USCT_A2(LB_)                                                                        \
    if (P->ta.CX == 4) adj2d_x2_t<LB_, 4>(g, s, l1, l2, lu, lv, pa1, pa2, P, k, t);            \
    else adj2d_x2_t<LB_, 8>(g, s, l1, l2, lu, lv, pa1, pa2, P, k, t);
    switch (P->ta.LB) {
        case 16: USCT_A2(16) break;
        case 32: USCT_A2(32) break;
        default: USCT_A2(64) break;
    }
#undef USCT_A2
}

// Three adjoint levels per launch (2D, no sponge): lam^k, lam^{k-1}, lam^{k-2}
// to out[0 .. 2] (null: not stored), imaged against p^{k-1}, p^{k-2}, p^{k-3} (pf).
template <int LB>
void adj2d_x3_t(dim3 g, hipStream_t s, const float *l1, const float *l2, float *const *out,
                const float *const *pf, const usct_plan *P, int k, const Tile &t) {
    constexpr int LV = 3;
    const int64_t tr = (int64_t)P->n_rec * P->Bp;
    AdjLevels<LV> o{};
    for (int i = 0; i < LV; ++i) {
        o.lam[i] = out[i];
        o.pf[i] = pf[i];
        o.sinc[i] = P->sinc + (int64_t)(k - i) * P->Bp;
        o.res[i] = P->res + (int64_t)(k - i) * tr;
    }
    hipLaunchKernelGGL((k_adj2d_xn<LB, 4, LV>), g, dim3(kBlock), 0, s, l1, l2, o, P->sp2,
                       P->src_cell, P->rec_head, P->rec_next, P->row_rec, P->g_hi, P->g_lo,
                       P->ny, P->nx, P->Bp, P->ta.RY, t.nxs, t.nyc, P->ng_adj(), P->ta.sched);
}

void launch_adj3_batched(const usct_plan *P, hipStream_t s, const float *l1, const float *l2,
                         float *const *out, const float *const *pf, int k) {
    const Tile t = tile_of(P, P->ta);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
    switch (P->ta.LB) {
        case 16: adj2d_x3_t<16>(g, s, l1, l2, out, pf, P, k, t); break;
        case 32: adj2d_x3_t<32>(g, s, l1, l2, out, pf, P, k, t); break;
        default: adj2d_x3_t<64>(g, s, l1, l2, out, pf, P, k, t); break;
    }
}

template <int LB, int CX, int CY>
void adj3d_t(dim3 g, hipStream_t s, const float *l1, const float *l2, float *lk,
             const float *pa, const usct_plan *P, const float *sinc, const float *res,
             const Tile &t) {
    const float *d2 = P->srcd2 + (sinc - P->sinc);
    hipLaunchKernelGGL((k_adj3d<LB, CX, CY>), g, dim3(kBlock), 0, s, l1, l2, lk, pa, P->sp2,
                       P->src_cell, sinc, P->rec_head, P->rec_next, P->row_rec, res, d2, P->g_hi,
                       P->g_lo, P->nz, P->ny, P->nx, P->Bp, P->ta.RY, t.nxs, t.nys, t.nyc, P->ng_adj(),
                       P->ta.blk);
}

template <int LB>
void adj3d_lb(dim3 g, hipStream_t s, const float *l1, const float *l2, float *lk,
              const float *pa, const usct_plan *P, const float *sinc, const float *res,
              const Tile &t) {
    if (P->ta.CX == 4 && P->ta.CY == 4) adj3d_t<LB, 4, 4>(g, s, l1, l2, lk, pa, P, sinc, res, t);
    else if (P->ta.CX == 8 && P->ta.CY == 2) adj3d_t<LB, 8, 2>(g, s, l1, l2, lk, pa, P, sinc, res, t);
    else if (P->ta.CX == 2 && P->ta.CY == 2) adj3d_t<LB, 2, 2>(g, s, l1, l2, lk, pa, P, sinc, res, t);
    else if (P->ta.CX == 2 && P->ta.CY == 4) adj3d_t<LB, 2, 4>(g, s, l1, l2, lk, pa, P, sinc, res, t);
    else if (P->ta.CX == 2 && P->ta.CY == 1) adj3d_t<LB, 2, 1>(g, s, l1, l2, lk, pa, P, sinc, res, t);
    else if (P->ta.CX == 1 && P->ta.CY == 2) adj3d_t<LB, 1, 2>(g, s, l1, l2, lk, pa, P, sinc, res, t);
    else if (P->ta.CX == 1 && P->ta.CY == 1) adj3d_t<LB, 1, 1>(g, s, l1, l2, lk, pa, P, sinc, res, t);
    else adj3d_t<LB, 4, 2>(g, s, l1, l2, lk, pa, P, sinc, res, t);
}

template <int LB, int TX, int TY, int PX, int PY>
void adj3d_lds_t(hipStream_t s, const float *l1, const float *l2, float *lk, const float *pa,
                 const usct_plan *P, const float *sinc, const float *res) {
    const Tiling &T = P->ta;
    const int ntx = (int)cdiv(P->nx, TX), nty = (int)cdiv(P->ny, TY);
    const int nzc = (int)cdiv(P->nz, T.RY), ng = P->Bp / LB;
    const dim3 g((unsigned)((int64_t)ng * ntx * nty * nzc));
    const float *d2 = P->srcd2 + (sinc - P->sinc);
    hipLaunchKernelGGL((k_adj3d_lds<LB, TX, TY, PX, PY>), g, dim3(kBlock), 0, s, l1, l2, lk, pa,
                       P->sp2, P->src_cell, P->rec_head, P->rec_next, P->row_rec, res, d2,
                       P->g_hi, P->g_lo, P->nz, P->ny, P->nx, P->Bp, T.RY, ntx, nty, nzc, ng);
}

void launch_adj_batched(const usct_plan *P, hipStream_t s, const float *l1, const float *l2,
                        float *lk, const float *pa, const float *pb, const float *pc2, int k) {
    const Tile t = tile_of(P, P->ta);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
    const float *sinc = P->sinc + (int64_t)k * P->Bp;
    const float *res = P->res + (int64_t)k * P->n_rec * P->Bp;
    if (P->ndim == 2) {
        if (P->sponge) adj2d_sel<true>(g, s, l1, l2, lk, pa, pb, pc2, P, sinc, res, t);
        else adj2d_sel<false>(g, s, l1, l2, lk, pa, pb, pc2, P, sinc, res, t);
    } else if (P->ta.lds) {
        const Tiling &T = P->ta;
        if (T.LB == 16 && T.CX == 16) adj3d_lds_t<16, 16, 8, 4, 2>(s, l1, l2, lk, pa, P, sinc, res);
        else if (T.LB == 16) adj3d_lds_t<16, 8, 8, 2, 2>(s, l1, l2, lk, pa, P, sinc, res);
        else if (T.LB == 32) adj3d_lds_t<32, 8, 8, 4, 2>(s, l1, l2, lk, pa, P, sinc, res);
        else adj3d_lds_t<64, 8, 4, 4, 2>(s, l1, l2, lk, pa, P, sinc, res);
    } else {
        switch (P->ta.LB) {
            case 16: adj3d_lb<16>(g, s, l1, l2, lk, pa, P, sinc, res, t); break;
            case 32: adj3d_lb<32>(g, s, l1, l2, lk, pa, P, sinc, res, t); break;
            default: adj3d_lb<64>(g, s, l1, l2, lk, pa, P, sinc, res, t); break;
        }
    }
}

// Adjoint level k with the forward run backwards (3D plans that keep no
// levels): lam^{k+1}, lam^{k+2} -> lam^k; forward levels k-1 (stencil) and k
// -> k-2, written when k >= 3.  Lane patches 2x2, 2x1, 1x2 or 1x1: the FP64
// halo of the forward level makes larger patches spill.
template <int LB, int CX, int CY>
void radj3d_t(dim3 g, hipStream_t s, const float *l1, const float *l2, float *lk, Lvl pc,
              Lvl pk, Lvl po, int wr, const usct_plan *P, int k, const Tile &t) {
    const double *sd = P->sincd + (int64_t)k * P->Bp;
    const float *res = P->res + (int64_t)k * P->n_rec * P->Bp;
    auto go = [&](auto lo_type, auto ysh) {
        using L = decltype(lo_type);
        hipLaunchKernelGGL((k_radj3d<LB, CX, CY, L, decltype(ysh)::value>), g, dim3(kBlock), 0, s,
                           l1, l2, lk, pc.hi, (const L *)pc.lo, pk.hi, (const L *)pk.lo, po.hi,
                           (L *)po.lo, wr, P->sp2, P->src_cell, sd, P->rec_head, P->rec_next,
                           P->row_rec, res, P->g_hi, P->g_lo, P->nz, P->ny, P->nx, P->Bp, P->ta.RY,
                           t.nxs, t.nys, t.nyc, P->ng_adj(), P->ta.blk);
    };
    // work-groups of kYsh strips stacked along y sharing their edge rows
    // (tile kind 4; built for the 2 x 2 patch)
    constexpr bool kYshBuilt = CX == 2 && CY == 2;
    if constexpr (kYshBuilt) {
        if (P->ta.lds == 4) {
            if (P->lo32) go(float{}, std::true_type{});
            else go(uint16_t{}, std::true_type{});
            return;
        }
    }
    if (P->lo32) go(float{}, std::false_type{});
    else go(uint16_t{}, std::false_type{});
}

template <int LB>
void radj3d_lb(dim3 g, hipStream_t s, const float *l1, const float *l2, float *lk, Lvl pc,
               Lvl pk, Lvl po, int wr, const usct_plan *P, int k, const Tile &t) {
    if (P->ta.CX == 2 && P->ta.CY == 2) radj3d_t<LB, 2, 2>(g, s, l1, l2, lk, pc, pk, po, wr, P, k, t);
    else if (P->ta.CX == 2 && P->ta.CY == 1) radj3d_t<LB, 2, 1>(g, s, l1, l2, lk, pc, pk, po, wr, P, k, t);
    else if (P->ta.CX == 1 && P->ta.CY == 2) radj3d_t<LB, 1, 2>(g, s, l1, l2, lk, pc, pk, po, wr, P, k, t);
    else radj3d_t<LB, 1, 1>(g, s, l1, l2, lk, pc, pk, po, wr, P, k, t);
}

bool radj_patch(int cx, int cy) {
    return (cx == 2 && cy == 2) || (cx == 2 && cy == 1) || (cx == 1 && cy == 2) ||
           (cx == 1 && cy == 1);
}

void launch_radj_batched(const usct_plan *P, hipStream_t s, const float *l1, const float *l2,
                         float *lk, Lvl pc, Lvl pk, Lvl po, bool wr, int k) {
    const Tile t = tile_of(P, P->ta);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
    switch (P->ta.LB) {
        case 16: radj3d_lb<16>(g, s, l1, l2, lk, pc, pk, po, wr ? 1 : 0, P, k, t); break;
        case 32: radj3d_lb<32>(g, s, l1, l2, lk, pc, pk, po, wr ? 1 : 0, P, k, t); break;
        default: radj3d_lb<64>(g, s, l1, l2, lk, pc, pk, po, wr ? 1 : 0, P, k, t); break;
    }
}

// ---------------------------------------------------------------------------
// Level pointers
// ---------------------------------------------------------------------------
// Checkpoint slot (0..) holding level n, or -1.
int ck_slot(const usct_plan *P, int n) {
    if (P->K <= 0 || n < 1) return -1;
    const int s = (n + 1) / P->K;
    if (s >= 1 && s <= P->NS - 2) {
        if (n == s * P->K - 1) return 2 * (s - 1);
        if (n == s * P->K) return 2 * (s - 1) + 1;
    }
    return -1;
}

// Where level n lives during the forward pass.  Low parts rotate over four
// buffers (a two-level launch touches levels n-2 .. n+1) except at checkpoint
// levels, whose low parts are kept.
Lvl fwd_level(const usct_plan *P, int n, bool store) {
    if (n <= 0) return {P->Z, P->Zlo};
    if (store && !P->rev) {
        const int c = ck_slot(P, n);
        if (c >= 0) return {P->CK[c], P->CKlo[c]};
        const int base_last = (P->NS - 1) * P->K;
        if (n >= base_last - 1) return {P->Wslot[n - base_last + 1], P->Lo[n % P->nrot]};
    }
    return {P->F[n % P->nrot], P->Lo[n % P->nrot]};
}

// Where level `level` lives while window s is re-computed / consumed.
Lvl win_level(const usct_plan *P, int s, int level) {
    if (level <= 0) return {P->Z, P->Zlo};
    const int base = s * P->K;
    const int i = level - base + 1;
    if (s == P->NS - 1 || i >= 2) return {P->Wslot[i], P->Lo[level % P->nrot]};
    const int c = 2 * (s - 1) + i;
    return {P->CK[c], P->CKlo[c]};
}

float *lam_level(const usct_plan *P, int k) { return k >= P->nt ? P->Z : P->F[k % P->nrot]; }
// Adjoint levels of a reversing plan: three slots of their own (the forward
// rotation F holds the levels being run backwards).
float *lam_rev(const usct_plan *P, int k) { return k >= P->nt ? P->Z : P->Wslot[k % 3]; }

// Forward levels a gradient keeps beyond the launch after next: checkpoint
// pairs (hi and lo) and the last window's levels (hi).  The others live in the
// rotation, where only the last two levels of a launch are read again.
bool level_kept(const usct_plan *P, int n, bool store) {
    if (!store || P->rev || n <= 0) return false;
    return ck_slot(P, n) >= 0 || n >= (P->NS - 1) * P->K - 1;
}

// Time levels per launch of a plan's fused forward / adjoint kernels (1: none).
int fwd_levels(const usct_plan *P) {
    if (!fused_fwd(P)) return 1;
    return P->ndim == 2 ? P->tf.lv : 2;
}
int adj_levels(const usct_plan *P) { return fused_adj(P) ? P->ta.lv : 1; }

void ev_start(Profile *pr, std::vector<hipEvent_t> &v, hipStream_t s) {
    if (!pr) return;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, s);
    v.push_back(a);
    v.push_back(b);
}
void ev_stop(Profile *pr, std::vector<hipEvent_t> &v, hipStream_t s) {
    if (!pr) return;
    (void)hipEventRecord(v.back(), s);
}

// Forward levels n, n+1 (two-level launch) or n alone; returns levels advanced.
int fwd_step(usct_plan *P, hipStream_t s, Lvl l2, Lvl l1, Lvl a, Lvl b, int n, bool two,
             bool record, Profile *pr) {
    if (pr) ev_start(pr, pr->fwd_ev, s);
    if (two) launch_fwd2_batched(P, s, l1, l2, a, b, n, record);
    else launch_fwd_batched(P, s, l1, l2, a, n, record);
    if (pr) {
        ev_stop(pr, pr->fwd_ev, s);
        pr->fwd_lv.push_back(two ? 2 : 1);
    }
    return two ? 2 : 1;
}

// Residual and misfit of the traces in recb against dobsT for the current pass.
void launch_residual(usct_plan *P, hipStream_t s, int misfit, bool want_res, bool windowed) {
    const int64_t ntr = (int64_t)P->n_rec * P->Bp;
    const int64_t count = (int64_t)P->nt * ntr;
    const float *W = windowed ? P->twT : nullptr;
    if (misfit == USCT_MISFIT_L2) {
        const int nb = (int)std::min<int64_t>(P->npart, grid_for(count));
        hipLaunchKernelGGL(k_residual_l2, dim3(nb), dim3(kBlock), 0, s, P->recb, P->recb_lo,
                           P->dobsT, P->dobsT_lo, W, want_res ? P->res : nullptr, count, P->part);
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, s, P->part, nb, 0.5,
                           P->J_acc, 1);
    } else {
        hipLaunchKernelGGL(k_gcn_stats, dim3(grid_for(ntr)), dim3(256), 0, s, P->recb, P->recb_lo,
                           P->dobsT, P->dobsT_lo, W, P->gstats, P->nt, ntr);
        hipLaunchKernelGGL(k_gcn_misfit, dim3(1), dim3(64), 0, s, P->gstats, P->tmask_cur, ntr,
                           P->J_acc, 1);
        if (want_res)
            hipLaunchKernelGGL(k_gcn_residual, dim3(grid_for(count)), dim3(256), 0, s, P->recb,
                               P->recb_lo, P->dobsT, P->dobsT_lo, W, P->gstats, P->tmask_cur,
                               P->res, P->nt, ntr);
    }
}

// Born level l lives in Dh/Dl[born_slot(l)]: four buffers, so a two-level
// launch never writes level n+1 over level n-2.
int born_slot(int l) { return ((l % 4) + 4) % 4; }

// Born state of levels -1 and 0 (zero), before the first Born step of a pass.
void born_clear(usct_plan *P, hipStream_t s) {
    const int64_t FE = P->field_elems();
    for (int i : {born_slot(0), born_slot(-1)}) {
        hipLaunchKernelGGL(k_fill, dim3(grid_for(FE)), dim3(256), 0, s, P->Dh[i], FE, 0.0f);
        hipLaunchKernelGGL(k_fill, dim3(grid_for(cdiv(FE, 2))), dim3(256), 0, s,
                           reinterpret_cast<float *>(P->Dl[i]), cdiv(FE, 2), 0.0f);
    }
}

// Born level n from the background levels n, n-1, n-2; its receiver samples
// become row n of the adjoint source res.
void launch_born(usct_plan *P, hipStream_t s, int n, Lvl u0, Lvl u1, Lvl u2) {
    const int64_t FE = P->field_elems();
    const int a = born_slot(n), b = born_slot(n - 1), c = born_slot(n - 2);
    float *res = P->res + (int64_t)n * P->n_rec * P->Bp;
    const dim3 g((unsigned)cdiv(FE, 256));
    if (P->sponge)
        hipLaunchKernelGGL(k_born<true>, g, dim3(256), 0, s, u0.hi, u0.lo, u1.hi, u1.lo, u2.hi,
                           u2.lo, P->Dh[b], P->Dl[b], P->Dh[c], P->Dl[c], P->Dh[a], P->Dl[a],
                           P->sp2, P->vm2, P->spg, P->rec_head, P->rec_next, res, P->nz, P->ny,
                           P->nx, P->ndim, P->Bp);
    else
        hipLaunchKernelGGL(k_born<false>, g, dim3(256), 0, s, u0.hi, u0.lo, u1.hi, u1.lo, u2.hi,
                           u2.lo, P->Dh[b], P->Dl[b], P->Dh[c], P->Dl[c], P->Dh[a], P->Dl[a],
                           P->sp2, P->vm2, P->spg, P->rec_head, P->rec_next, res, P->nz, P->ny,
                           P->nx, P->ndim, P->Bp);
}

// Footprint plans: weighted transmit cells of level n, then (record) the
// receiver groups into trace row n.
void fp_forward_fix(usct_plan *P, hipStream_t s, Lvl lv, int n, bool record) {
    hipLaunchKernelGGL(k_fp_source, dim3(grid_for(P->fp_SE)), dim3(256), 0, s, lv.hi, lv.lo,
                       P->fp_pass, P->fp_SE, P->fp_eb, P->fp_ec, P->fp_ecoef, P->wav, n,
                       P->sponge ? P->spg : nullptr, P->Bp);
    if (!record) return;
    const int64_t ntr = (int64_t)P->n_rec * P->Bp;
    hipLaunchKernelGGL(k_fp_record, dim3(grid_for(ntr)), dim3(256), 0, s, lv.hi, lv.lo, P->fp_roff,
                       P->fp_rcell, P->fp_rw, P->recb + (int64_t)n * ntr,
                       P->recb_lo + (int64_t)n * ntr, P->n_rec, P->Bp);
}

// Background levels n, n+1 and Born levels n, n+1 in one launch (2D, no
// sponge, no footprints): k_born2d_x2 on the forward's lane groups, USCT_GNCX
// cells per lane and USCT_GNRY rows per wavefront (on the 512^2, 128-shot
// product: 4 cells and 32 rows at two wavefronts per SIMD 308 ms; 2 cells
// 319-325 ms; 8 or 16 rows 313-338 ms; one wavefront per SIMD 381 ms).
#ifndef USCT_GNCX
#define USCT_GNCX 4
#endif
#ifndef USCT_GNRY
#define USCT_GNRY 32
#endif
template <int LB>
void born2_lb(hipStream_t s, const usct_plan *P, int n, Lvl u2, Lvl u1, Lvl a, Lvl b) {
    constexpr int CX = USCT_GNCX;
    Tiling tg = P->tf;
    tg.CX = CX;
    tg.RY = USCT_GNRY;
    const Tile t = tile_of(P, tg);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
    const double *sa = P->sincd + (int64_t)n * P->Bp;
    const int64_t tr = (int64_t)P->n_rec * P->Bp;
    float *ra = P->res + (int64_t)n * tr;
    const int dc = born_slot(n - 1), dp = born_slot(n - 2), da = born_slot(n), db = born_slot(n + 1);
    hipLaunchKernelGGL((k_born2d_x2<LB, CX>), g, dim3(kBlock), 0, s, u1.hi, u2.hi, a.hi, b.hi,
                       u1.lo, u2.lo, a.lo, b.lo, P->Dh[dc], P->Dh[dp], P->Dh[da], P->Dh[db],
                       P->Dl[dc], P->Dl[dp], P->Dl[da], P->Dl[db], P->sp2, P->vm2, P->src_cell,
                       sa, sa + P->Bp, P->rec_head, P->rec_next, P->row_rec, ra, ra + tr, P->ny,
                       P->nx, P->Bp, tg.RY, t.nxs, t.nyc, P->Bp / tg.LB);
}

void launch_born2_batched(const usct_plan *P, hipStream_t s, int n, Lvl u2, Lvl u1, Lvl a, Lvl b) {
    switch (P->tf.LB) {
        case 16: born2_lb<16>(s, P, n, u2, u1, a, b); break;
        case 32: born2_lb<32>(s, P, n, u2, u1, a, b); break;
        default: born2_lb<64>(s, P, n, u2, u1, a, b); break;
    }
}

// Background and Born level n in one launch (3D, no footprints): k_born3d on
// the forward's lane groups and planes per wavefront, USCT_GN3CX x USCT_GN3CY
// cells per lane (two FP64 halos per lane: smaller patches than the forward's;
// on the 256^3, 64-shot product 2x1 takes 33.5 s, 1x2 33.9, 1x1 34.6, 2x2 36.7,
// against 39.5 for separate k_fwd3d and k_born launches).
#ifndef USCT_GN3CX
#define USCT_GN3CX 2
#endif
#ifndef USCT_GN3CY
#define USCT_GN3CY 1
#endif
template <int LB>
void born3_lb(hipStream_t s, const usct_plan *P, int n, Lvl u2, Lvl u1, Lvl u0) {
    Tiling tg = P->tf;
    tg.CX = USCT_GN3CX;
    tg.CY = USCT_GN3CY;
    tg.lds = 0;
    const Tile t = tile_of(P, tg);
    const dim3 g((unsigned)cdiv(t.waves, kBlock / kWave));
    const double *sd = P->sincd + (int64_t)n * P->Bp;
    float *d2 = P->srcd2 + (int64_t)n * P->Bp;
    float *res = P->res + (int64_t)n * P->n_rec * P->Bp;
    const int a = born_slot(n), b1 = born_slot(n - 1), c2 = born_slot(n - 2);
    hipLaunchKernelGGL((k_born3d<LB, USCT_GN3CX, USCT_GN3CY>), g, dim3(kBlock), 0, s, u1.hi, u2.hi,
                       u0.hi, u1.lo, u2.lo, u0.lo, P->Dh[b1], P->Dh[c2], P->Dh[a], P->Dl[b1],
                       P->Dl[c2], P->Dl[a], P->sp2, P->vm2, P->src_cell, sd, P->rec_head,
                       P->rec_next, P->row_rec, res, d2, P->nz, P->ny, P->nx, P->Bp, tg.RY, t.nxs,
                       t.nys, t.nyc, P->Bp / tg.LB, tg.blk);
}

void launch_born3_batched(const usct_plan *P, hipStream_t s, int n, Lvl u2, Lvl u1, Lvl u0) {
    switch (P->tf.LB) {
        case 16: born3_lb<16>(s, P, n, u2, u1, u0); break;
        case 32: born3_lb<32>(s, P, n, u2, u1, u0); break;
        default: born3_lb<64>(s, P, n, u2, u1, u0); break;
    }
}

// Footprint plans: receiver groups of Born level n into row n of the adjoint source.
void fp_born_record(usct_plan *P, hipStream_t s, int n) {
    const int64_t ntr = (int64_t)P->n_rec * P->Bp;
    hipLaunchKernelGGL(k_fp_record, dim3(grid_for(ntr)), dim3(256), 0, s, P->Dh[born_slot(n)],
                       P->Dl[born_slot(n)], P->fp_roff, P->fp_rcell, P->fp_rw,
                       P->res + (int64_t)n * ntr, nullptr, P->n_rec, P->Bp);
}

// Footprint plans: receiver injection and imaging corrections after adjoint level k.
void fp_adjoint_fix(usct_plan *P, hipStream_t s, float *lk, const float *pa, const float *pb,
                    const float *pc2, int k) {
    const int ng = P->ng_adj();
    const float *res = P->res + (int64_t)k * P->n_rec * P->Bp;
    const dim3 g(grid_for((int64_t)P->fp_FC * ng));
    if (P->sponge)
        hipLaunchKernelGGL(k_fp_adjoint<true>, g, dim3(256), 0, s, lk, res, pa, pb, pc2, P->sp2,
                           P->fp_pass, P->fp_FC, P->fp_fcell, P->fp_frb, P->fp_fre, P->fp_rj,
                           P->fp_rwf, P->fp_fsb, P->fp_fse, P->fp_sb, P->fp_scoef, P->wav, k,
                           P->g_hi, P->g_lo, P->ta.LB, ng, P->nz, P->ny, P->nx, P->ndim, P->Bp);
    else
        hipLaunchKernelGGL(k_fp_adjoint<false>, g, dim3(256), 0, s, lk, res, pa, pb, pc2, P->sp2,
                           P->fp_pass, P->fp_FC, P->fp_fcell, P->fp_frb, P->fp_fre, P->fp_rj,
                           P->fp_rwf, P->fp_fsb, P->fp_fse, P->fp_sb, P->fp_scoef, P->wav, k,
                           P->g_hi, P->g_lo, P->ta.LB, ng, P->nz, P->ny, P->nx, P->ndim, P->Bp);
}

// The launch sequence of one batched pass (captured into a hipGraph).  The
// Gauss-Newton product runs the Born level of each background level right
// after it (levels n-2 .. n+1 are resident then) and takes the Born traces,
// instead of a residual, as the adjoint's source.
void batched_sequence(usct_plan *P, hipStream_t s, int mode, int misfit, bool windowed,
                      Profile *pr) {
    const bool born = mode == MODE_GNHV;
    const bool store = mode == MODE_GRAD || born || mode == MODE_ADJSRC;
    const int nt = P->nt;
    // rec row 0 = 0 (fwi.py:58: rec starts as zeros, level 0 is never written)
    hipLaunchKernelGGL(k_fill, dim3(grid_for((int64_t)P->n_rec * P->Bp)), dim3(256), 0, s,
                       P->recb, (int64_t)P->n_rec * P->Bp, 0.0f);
    hipLaunchKernelGGL(k_fill, dim3(grid_for((int64_t)P->n_rec * P->Bp)), dim3(256), 0, s,
                       P->recb_lo, (int64_t)P->n_rec * P->Bp, 0.0f);
    const bool x2 = fused_fwd(P), x2a = fused_adj(P);
    const int lvf = fwd_levels(P), lva = adj_levels(P);
    if (born) born_clear(P, s);
    for (int n = 1; n < nt;) {
        const bool two = x2 && n + 1 < nt;
        if (!born && lvf >= 3 && n + 2 < nt) {
            // three levels: the first is stored only where the gradient keeps it
            Lvl out[3];
            for (int i = 0; i < 3; ++i) out[i] = fwd_level(P, n + i, store);
            if (!level_kept(P, n, store)) out[0].hi = nullptr;
            if (ck_slot(P, n) < 0 || !level_kept(P, n, store)) out[0].lo = nullptr;
            if (pr) ev_start(pr, pr->fwd_ev, s);
            launch_fwd3_batched(P, s, fwd_level(P, n - 1, store), fwd_level(P, n - 2, store), out,
                                n, true);
            if (pr) {
                ev_stop(pr, pr->fwd_ev, s);
                pr->fwd_lv.push_back(3);
            }
            n += 3;
            continue;
        }
        if (born && two && P->ndim == 2) {
            // background and Born levels n, n+1 together (fusable: no sponge, no footprints)
            if (pr) ev_start(pr, pr->fwd_ev, s);
            launch_born2_batched(P, s, n, fwd_level(P, n - 2, true), fwd_level(P, n - 1, true),
                                 fwd_level(P, n, true), fwd_level(P, n + 1, true));
            if (pr) {
                ev_stop(pr, pr->fwd_ev, s);
                pr->fwd_lv.push_back(2);
            }
            n += 2;
            continue;
        }
        if (born && P->ndim == 3 && fusable(P)) {
            // background and Born level n together
            if (pr) ev_start(pr, pr->fwd_ev, s);
            launch_born3_batched(P, s, n, fwd_level(P, n - 2, true), fwd_level(P, n - 1, true),
                                 fwd_level(P, n, true));
            if (pr) {
                ev_stop(pr, pr->fwd_ev, s);
                pr->fwd_lv.push_back(1);
            }
            n += 1;
            continue;
        }
        const int adv = fwd_step(P, s, fwd_level(P, n - 2, store), fwd_level(P, n - 1, store),
                                 fwd_level(P, n, store), two ? fwd_level(P, n + 1, store) : Lvl{},
                                 n, two, true, pr);
        if (P->fp)
            for (int l = n; l < n + adv; ++l) fp_forward_fix(P, s, fwd_level(P, l, store), l, true);
        if (born)
            for (int l = n; l < n + adv; ++l) {
                launch_born(P, s, l, fwd_level(P, l, true), fwd_level(P, l - 1, true),
                            fwd_level(P, l - 2, true));
                if (P->fp) fp_born_record(P, s, l);
            }
        n += adv;
    }
    if (mode == MODE_FWD) return;
    if (!born && mode != MODE_ADJSRC) launch_residual(P, s, misfit, store, windowed);
    if (!store) return;
    if (P->rev) {
        // the forward's last two levels are in the rotation; each launch takes
        // them one level back while the adjoint advances
        for (int k = nt - 1; k >= 1; --k) {
            if (pr) ev_start(pr, pr->adj_ev, s);
            launch_radj_batched(P, s, lam_rev(P, k + 1), lam_rev(P, k + 2), lam_rev(P, k),
                                fwd_level(P, k - 1, false), fwd_level(P, k, false),
                                fwd_level(P, k - 2, false), k >= 3, k);
            if (pr) {
                ev_stop(pr, pr->adj_ev, s);
                pr->adj_lv.push_back(1);
            }
        }
        return;
    }
    for (int w = P->NS - 1; w >= 0; --w) {
        const int base = w * P->K;
        const int khi = std::min(base + P->K, nt - 1);
        if (w < P->NS - 1) {
            for (int n = base + 1; n <= khi;) {
                if (lvf >= 3 && n + 2 <= khi) {
                    Lvl out[3];
                    for (int i = 0; i < 3; ++i) out[i] = win_level(P, w, n + i);
                    out[0].lo = nullptr;  // levels of a window keep their hi only
                    if (pr) ev_start(pr, pr->fwd_ev, s);
                    launch_fwd3_batched(P, s, win_level(P, w, n - 1), win_level(P, w, n - 2), out,
                                        n, false);
                    if (pr) {
                        ev_stop(pr, pr->fwd_ev, s);
                        pr->fwd_lv.push_back(3);
                    }
                    n += 3;
                    continue;
                }
                const bool two = x2 && n + 1 <= khi;
                const int adv = fwd_step(P, s, win_level(P, w, n - 2), win_level(P, w, n - 1),
                                         win_level(P, w, n), two ? win_level(P, w, n + 1) : Lvl{},
                                         n, two, false, pr);
                if (P->fp)
                    for (int l = n; l < n + adv; ++l) fp_forward_fix(P, s, win_level(P, w, l), l, false);
                n += adv;
            }
        }
        for (int k = khi; k >= base + 1;) {
            if (lva >= 3 && k - 2 >= base + 1) {
                // lam^k is read by no later launch: not stored
                float *out[3] = {nullptr, lam_level(P, k - 1), lam_level(P, k - 2)};
                const float *pf[3] = {win_level(P, w, k - 1).hi, win_level(P, w, k - 2).hi,
                                      win_level(P, w, k - 3).hi};
                if (pr) ev_start(pr, pr->adj_ev, s);
                launch_adj3_batched(P, s, lam_level(P, k + 1), lam_level(P, k + 2), out, pf, k);
                if (pr) {
                    ev_stop(pr, pr->adj_ev, s);
                    pr->adj_lv.push_back(3);
                }
                k -= 3;
                continue;
            }
            const bool two = x2a && k - 1 >= base + 1;
            if (pr) ev_start(pr, pr->adj_ev, s);
            if (two) {
                launch_adj2_batched(P, s, lam_level(P, k + 1), lam_level(P, k + 2),
                                    lam_level(P, k), lam_level(P, k - 1),
                                    win_level(P, w, k - 1).hi, win_level(P, w, k - 2).hi, k);
            } else {
                const float *pa, *pb = nullptr, *pc2 = nullptr;
                if (P->sponge) {
                    pa = win_level(P, w, k).hi;
                    pb = win_level(P, w, k - 1).hi;
                    pc2 = win_level(P, w, k - 2).hi;
                } else {
                    pa = win_level(P, w, k - 1).hi;
                }
                launch_adj_batched(P, s, lam_level(P, k + 1), lam_level(P, k + 2),
                                   lam_level(P, k), pa, pb, pc2, k);
                if (P->fp) fp_adjoint_fix(P, s, lam_level(P, k), pa, pb, pc2, k);
            }
            if (pr) {
                ev_stop(pr, pr->adj_ev, s);
                pr->adj_lv.push_back(two ? 2 : 1);
            }
            k -= two ? 2 : 1;
        }
    }
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(USCT_EHIP, "kernel launch failed: %s", hipGetErrorString(e));
    return 0;
}

int run_pass_sequence(usct_plan *P, int mode, int misfit, bool windowed, Profile *pr) {
    const bool use_graph = !(P->flags & USCT_FLAG_NO_GRAPH) && pr == nullptr;
    if (!use_graph) {
        batched_sequence(P, P->stream, mode, misfit, windowed, pr);
        return check_launch();
    }
    const int key = (mode * 4 + misfit) * 2 + (windowed ? 1 : 0);
    auto it = P->graphs.find(key);
    if (it == P->graphs.end()) {
        hipGraph_t graph;
        HIPX(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
        batched_sequence(P, P->stream, mode, misfit, windowed, nullptr);
        hipError_t le = hipGetLastError();
        HIPX(hipStreamEndCapture(P->stream, &graph));
        if (le != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(USCT_EHIP, "launch during capture failed: %s", hipGetErrorString(le));
        }
        hipGraphExec_t exec;
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail(USCT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
        it = P->graphs.emplace(key, exec).first;
    }
    HIPX(hipGraphLaunch(it->second, P->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Batched driver
// ---------------------------------------------------------------------------
int run_batched(usct_plan *P, int mode, int misfit, const double *m, const double *w,
                const double *sponge, const double *dobs, double *data, double *J, double *grad,
                Profile *pr, const double *tw) {
    hipStream_t s = P->stream;
    const int64_t N = P->N;
    hipLaunchKernelGGL(k_setup_s, dim3(grid_for(N)), dim3(256), 0, s, m, P->sp2, N,
                       P->dt * P->dt, 1.0 / (P->h * P->h));
    if (P->sponge)  // FP64 mask -> FP32 working copy
        hipLaunchKernelGGL(k_d2f, dim3(grid_for(N)), dim3(256), 0, s, sponge, P->spg, N);
    if (J) HIPX(hipMemsetAsync(P->J_acc, 0, sizeof(double), s));
    if (grad) {
        HIPX(hipMemsetAsync(P->g_hi, 0, sizeof(float) * N * P->ng_adj(), s));
        HIPX(hipMemsetAsync(P->g_lo, 0, sizeof(float) * N * P->ng_adj(), s));
    }
    const int64_t ntr = (int64_t)P->n_rec * P->Bp;
    for (int p = 0; p < P->passes; ++p) {
        const int b0 = p * P->Bp;
        const int nb = std::min(P->Bp, P->n_tx - b0);
        HIPX(hipMemcpyAsync(P->src_cell, P->src_all + (int64_t)p * P->Bp, sizeof(int) * P->Bp,
                            hipMemcpyDeviceToDevice, s));
        if (misfit == USCT_MISFIT_GCN)
            HIPX(hipMemcpyAsync(P->tmask_cur, P->tmask + p * ntr, ntr, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_setup_sinc, dim3(grid_for((int64_t)P->nt * P->Bp)), dim3(256), 0, s,
                           m, w, P->src_cell, P->sinc, P->sincd, P->nt, P->Bp, P->dt * P->dt);
        const int64_t NJ = (int64_t)P->nt * P->n_rec;
        const dim3 gt((unsigned)cdiv(NJ, kTT), (unsigned)cdiv(P->Bp, kTT));
        if (dobs)  // observed data, or (MODE_ADJSRC) the adjoint source straight into res
            hipLaunchKernelGGL(k_dobs_to_batched, gt, dim3(256), 0, s, dobs,
                               mode == MODE_ADJSRC ? P->res : P->dobsT, P->dobsT_lo, NJ, P->Bp,
                               b0, nb);
        if (tw)
            hipLaunchKernelGGL(k_dobs_to_batched, gt, dim3(256), 0, s, tw, P->twT, P->twT_lo, NJ,
                               P->Bp, b0, nb);
        if (P->fp)
            HIPX(hipMemcpyAsync(P->fp_pass, P->fp_pass_ids + p, sizeof(int),
                                hipMemcpyDeviceToDevice, s));
        TRY(check_launch());
        TRY(run_pass_sequence(P, mode, misfit, tw != nullptr, pr));
        if (data)
            hipLaunchKernelGGL(k_batched_to_data, gt, dim3(256), 0, s, P->recb, P->recb_lo, data,
                               NJ, P->Bp, b0, nb);
    }
    if (grad)
        hipLaunchKernelGGL(k_finalize_grad, dim3(grid_for(N)), dim3(256), 0, s, P->g_hi, P->g_lo,
                           P->ng_adj(), N, m, grad, 0);
    if (J) HIPX(hipMemcpyAsync(J, P->J_acc, sizeof(double), hipMemcpyDeviceToDevice, s));
    return check_launch();
}

// ---------------------------------------------------------------------------
// Plain driver: one shot at a time, one launch per time step, full history.
// ---------------------------------------------------------------------------
int run_plain(usct_plan *P, int mode, int misfit, const double *m, const double *w,
              const double *sponge, const double *dobs, double *data, double *J, double *grad,
              Profile *pr, const double *tw) {
    hipStream_t s = P->stream;
    const int64_t N = P->N;
    const unsigned gN = grid_for(N);
    hipLaunchKernelGGL(k_setup_s, dim3(gN), dim3(256), 0, s, m, P->sp2, N, P->dt * P->dt,
                       1.0 / (P->h * P->h));
    if (P->sponge)
        hipLaunchKernelGGL(k_d2f, dim3(gN), dim3(256), 0, s, sponge, P->spg, N);
    if (J) HIPX(hipMemsetAsync(P->J_acc, 0, sizeof(double), s));
    if (grad) {
        HIPX(hipMemsetAsync(P->g_hi, 0, sizeof(float) * N, s));
        HIPX(hipMemsetAsync(P->g_lo, 0, sizeof(float) * N, s));
    }
    const bool born = mode == MODE_GNHV;
    const bool store = mode == MODE_GRAD || born || mode == MODE_ADJSRC;
    const int nt = P->nt, nr = P->n_rec;
    const int nd = P->ndim;
    // History slots: level n in Wslot[n] (n >= 1), levels <= 0 are Z.
    auto lvl = [&](int n) -> float * {
        if (n <= 0) return P->Z;
        return store ? P->Wslot[n] : P->F[n % 4];
    };
    auto llo = [&](int n) -> uint16_t * { return n <= 0 ? P->Zlo : P->Lo[n % 4]; };
    for (int b = 0; b < P->n_tx; ++b) {
        // footprint plans add their weighted sources after each level instead
        const int src = P->fp ? -1 : (int)P->tx[b];
        if (P->fp)
            HIPX(hipMemcpyAsync(P->fp_pass, P->fp_pass_ids + b, sizeof(int),
                                hipMemcpyDeviceToDevice, s));
        // S_src * w[n-1] per level, same FP64 product as the batched path.
        hipLaunchKernelGGL(k_setup_sinc, dim3(grid_for(nt)), dim3(256), 0, s, m, w,
                           P->src_all + b, P->sinc, P->sincd, nt, 1, P->dt * P->dt);
        HIPX(hipMemsetAsync(P->recb, 0, sizeof(float) * nr, s));
        HIPX(hipMemsetAsync(P->recb_lo, 0, sizeof(float) * nr, s));
        if (born) born_clear(P, s);
        for (int n = 1; n < nt; ++n) {
            if (pr) ev_start(pr, pr->fwd_ev, s);
            auto go = [&](auto lo_type, auto sponge_tag) {
                using L = decltype(lo_type);
                constexpr bool SP = decltype(sponge_tag)::value;
                hipLaunchKernelGGL((k_fwd_plain<SP, true, L>), dim3(gN), dim3(256), 0, s,
                                   lvl(n - 1), lvl(n - 2), lvl(n), (const L *)llo(n - 1),
                                   (const L *)llo(n - 2), (L *)llo(n), P->sp2, P->spg, src,
                                   P->sincd + n, P->rec_head, P->rec_next,
                                   P->recb + (int64_t)n * nr, P->recb_lo + (int64_t)n * nr,
                                   P->srcd2 + n, P->nz, P->ny, P->nx, nd);
            };
            if (P->lo32) {
                if (P->sponge) go(float{}, std::true_type{});
                else go(float{}, std::false_type{});
            } else {
                if (P->sponge) go(uint16_t{}, std::true_type{});
                else go(uint16_t{}, std::false_type{});
            }
            if (pr) {
                ev_stop(pr, pr->fwd_ev, s);
                pr->fwd_lv.push_back(1);
            }
            if (P->fp) fp_forward_fix(P, s, {lvl(n), llo(n)}, n, true);
            if (born) {
                launch_born(P, s, n, {lvl(n), llo(n)}, {lvl(n - 1), llo(n - 1)},
                            {lvl(n - 2), llo(n - 2)});
                if (P->fp) fp_born_record(P, s, n);
            }
        }
        const int64_t NJ = (int64_t)nt * nr;
        const dim3 gt((unsigned)cdiv(NJ, kTT), 1);
        if (data)
            hipLaunchKernelGGL(k_batched_to_data, gt, dim3(256), 0, s, P->recb, P->recb_lo, data,
                               NJ, 1, b, 1);
        if (mode == MODE_FWD) continue;
        if (mode == MODE_ADJSRC) {
            hipLaunchKernelGGL(k_dobs_to_batched, gt, dim3(256), 0, s, dobs, P->res, P->dobsT_lo,
                               NJ, 1, b, 1);
        } else if (!born) {
            // residual of this shot against dobs[b] (FP64 -> FP32 hi + lo)
            hipLaunchKernelGGL(k_dobs_to_batched, gt, dim3(256), 0, s, dobs, P->dobsT,
                               P->dobsT_lo, NJ, 1, b, 1);
            if (tw)
                hipLaunchKernelGGL(k_dobs_to_batched, gt, dim3(256), 0, s, tw, P->twT, P->twT_lo,
                                   NJ, 1, b, 1);
            if (misfit == USCT_MISFIT_GCN)
                HIPX(hipMemcpyAsync(P->tmask_cur, P->tmask, nr, hipMemcpyDeviceToDevice, s));
            launch_residual(P, s, misfit, store, tw != nullptr);
        }
        if (!store) continue;
        for (int k = nt - 1; k >= 1; --k) {
            const float *pa, *pb = nullptr, *pc2 = nullptr;
            if (P->sponge) { pa = lvl(k); pb = lvl(k - 1); pc2 = lvl(k - 2); }
            else pa = lvl(k - 1);
            float *lk = lam_level(P, k);
            if (pr) ev_start(pr, pr->adj_ev, s);
            if (P->sponge)
                hipLaunchKernelGGL((k_adj_plain<true>), dim3(gN), dim3(256), 0, s, lam_level(P, k + 1),
                                   lam_level(P, k + 2), lk, pa, pb, pc2, P->sp2, P->spg, src, P->sinc + k, P->srcd2 + k,
                                   P->rec_head, P->rec_next, P->res + (int64_t)k * nr, P->g_hi,
                                   P->g_lo, P->nz, P->ny, P->nx, nd);
            else
                hipLaunchKernelGGL((k_adj_plain<false>), dim3(gN), dim3(256), 0, s,
                                   lam_level(P, k + 1), lam_level(P, k + 2), lk, pa, pb, pc2, P->sp2,
                                   P->spg, src, P->sinc + k, P->srcd2 + k, P->rec_head, P->rec_next,
                                   P->res + (int64_t)k * nr, P->g_hi, P->g_lo, P->nz, P->ny, P->nx,
                                   nd);
            if (P->fp) fp_adjoint_fix(P, s, lk, pa, pb, pc2, k);
            if (pr) {
                ev_stop(pr, pr->adj_ev, s);
                pr->adj_lv.push_back(1);
            }
        }
        TRY(check_launch());
    }
    if (grad)
        hipLaunchKernelGGL(k_finalize_grad, dim3(gN), dim3(256), 0, s, P->g_hi, P->g_lo, 1, N, m,
                           grad, 0);
    if (J) HIPX(hipMemcpyAsync(J, P->J_acc, sizeof(double), hipMemcpyDeviceToDevice, s));
    return check_launch();
}

// ---------------------------------------------------------------------------
// Planning
// ---------------------------------------------------------------------------
struct MemPlan {
    int Bp, passes, K, NS;
    int64_t bytes;
    double cost;  // relative time estimate
};

int64_t batched_bytes(const usct_plan *P, int Bp, int K, int NS, bool grad) {
    const int64_t field = P->N * (int64_t)Bp * 4;
    const int64_t traces = (int64_t)P->nt * P->n_rec * Bp * 4;
    int64_t fields = 1 + P->nrot;  // Z + F
    int64_t halves = 1 + P->nrot;  // Zlo + Lo
    if (grad) {
        fields += (K + 2) + 2 * std::max(NS - 2, 0);
        halves += 2 * std::max(NS - 2, 0);
    }
    const int64_t lo_b = P->lo32 ? 4 : 2;  // bytes of a low part
    int64_t b = fields * field + halves * (field / 4 * lo_b) + 5 * traces;  // rec hi/lo, dobs hi/lo, res
    b += (fields * 4 + halves * lo_b) * 2 * kGuardCells * Bp;
    b += P->N * 8 /*sp2*/ + P->N * 4 /*sponge*/ + 2 * P->N * 4 * (Bp / 16) /*g*/;
    b += (int64_t)P->nt * Bp * 4 + (int64_t)P->n_rec * Bp * 24 + (int64_t)P->N * 4 + (1 << 20);
    return b;
}

// A 3D gradient plan may run the forward backwards instead of keeping levels
// (no sponge in 3D, point elements only: footprint plans keep levels).
bool reversible(const usct_plan *P) {
    return P->ndim == 3 && !P->sponge && P->method == USCT_METHOD_BATCHED &&
           !(P->flags & (USCT_FLAG_FORWARD_ONLY | USCT_FLAG_NO_REVERSE));
}

int choose_batched(usct_plan *P, int ckpt_every) {
    const int B = P->n_tx;
    const int nt = P->nt;
    // HBM bytes streamed per shot-cell and time level (forward, adjoint,
    // re-computed forward): the two-level 2D kernels move about 13-14 bytes per
    // level, the single-level ones 18-20; the reversing adjoint takes 1.7x the
    // forward's time (7.4 against 4.3 ms per level on 256^3 with 64 shots).
    const bool x2 = P->ndim == 2 && !P->sponge && !(P->flags & USCT_FLAG_NO_FUSE);
    const double b_fwd = x2 ? 13.0 : 19.0, b_adj = x2 ? 14.0 : 17.0, b_radj = 33.0;
    // plans that launch three levels at a time rotate over five buffers
    P->nrot = x2 && (P->flags & USCT_FLAG_LEVELS_3) ? kMaxRot : 4;
    const bool force_rev = (P->flags & USCT_FLAG_REVERSE) != 0;
    if (force_rev && !reversible(P))
        return fail(USCT_ENOTSUP, "USCT_FLAG_REVERSE needs a batched 3D gradient plan without "
                                  "footprints or USCT_FLAG_NO_REVERSE");
    const bool try_rev = reversible(P) && (force_rev || ckpt_every <= 0);
    MemPlan best{0, 0, 0, 0, 0, 1e300};
    bool best_rev = false;
    int best_lb = 64;
    // Every lane-group width, every number of groups per pass (padded shots
    // count in the cost), whole history or checkpoints: the cheapest plan that
    // fits the budget.  Narrow groups let a pass hold few shots when a wide one
    // does not fit.
    for (int lb : {64, 32, 16}) {
        const int64_t total_groups = cdiv(B, lb);
        for (int64_t gpp = total_groups; gpp >= 1; --gpp) {
            const int Bp = (int)(gpp * lb);
            const int passes = (int)cdiv(total_groups, gpp);
            // candidate K values: whole history, requested, and the memory minimiser
            std::vector<int> Ks;
            if (ckpt_every > 0) Ks.push_back(std::min(std::max(ckpt_every, 2), std::max(nt - 1, 2)));
            else {
                Ks.push_back(std::max(nt - 1, 2));
                int kmin = std::max(2, (int)std::lround(std::sqrt(2.0 * std::max(nt - 1, 1))));
                Ks.push_back(kmin);
            }
            // narrower lane groups stream a little less efficiently
            const double eff = lb == 64 ? 1.0 : (lb == 32 ? 0.95 : 0.88);
            if (!force_rev)
                for (int K : Ks) {
                    const int NS = (int)std::max<int64_t>(1, cdiv(nt - 1, K));
                    const int64_t bytes =
                        batched_bytes(P, Bp, K, NS, !(P->flags & USCT_FLAG_FORWARD_ONLY));
                    if (bytes > P->budget) continue;
                    const double per = b_fwd + b_adj + (NS > 1 ? b_fwd * (1.0 - 1.0 / NS) : 0.0);
                    // launch overhead of ~3 us per step launch, in streamed-byte units at 5 TB/s
                    const double launches = (double)passes * nt * (NS > 1 ? 3.0 : 2.0);
                    const double cost = (double)passes * Bp * per * (double)P->N * nt / eff +
                                        launches * 15e6;
                    if (cost < best.cost) {
                        best = MemPlan{Bp, passes, K, NS, bytes, cost};
                        best_rev = false;
                        best_lb = lb;
                    }
                }
            if (try_rev) {  // three forward and three adjoint levels, nothing kept
                const int64_t bytes = batched_bytes(P, Bp, 1, 1, true);
                if (bytes <= P->budget) {
                    const double cost = (double)passes * Bp * (b_fwd + b_radj) * (double)P->N * nt /
                                            eff + (double)passes * nt * 2.0 * 15e6;
                    if (cost < best.cost) {
                        best = MemPlan{Bp, passes, 1, 1, bytes, cost};
                        best_rev = true;
                        best_lb = lb;
                    }
                }
            }
        }
    }
    if (best.Bp == 0)
        return fail(USCT_ENOMEM, "batched plan does not fit the memory budget of %lld bytes",
                    (long long)P->budget);
    P->tf.LB = P->ta.LB = best_lb;
    P->Bp = best.Bp;
    P->passes = best.passes;
    P->K = best.K;
    P->NS = best.NS;
    P->rev = best_rev;
    // the reversing adjoint has its own (smaller) set of lane patches
    if (P->rev && !radj_patch(P->ta.CX, P->ta.CY)) P->ta.CX = P->ta.CY = 2;
    return 0;
}

// Rows (2D) or planes (3D) per wavefront: the largest power of two in [4, 64]
// that still gives about 8 wavefronts per SIMD (256 CUs x 4 SIMDs), so the
// streaming loads of enough wavefronts are in flight to cover HBM latency.
void default_rows(const usct_plan *P, Tiling &T) {
    const int W = kWave / T.LB;
    const int64_t ng = P->Bp / T.LB;
    const int64_t nxs = cdiv(P->nx, (int64_t)W * T.CX);
    const int64_t cols = P->ndim == 2 ? ng * nxs : ng * nxs * cdiv(P->ny, T.CY);
    const int64_t depth = P->ndim == 2 ? P->ny : P->nz;
    const int64_t target = 256 * 4 * 8;
    // 3D: at most 16 planes per wavefront (the autotuner's picks on 256^3 are 8
    // and 16; 64-plane chunks measured 9-13 % slower launches)
    int r = P->ndim == 3 ? 16 : 64;
    while (r > 4 && cols * cdiv(depth, r) < target) r >>= 1;
    T.RY = r;
}

int allocate_batched(usct_plan *P) {
    const int64_t FE = P->field_elems();
    TRY(dalloc_guarded(P, &P->Z, FE, P->Bp));
    HIPX(hipMemset(P->Z, 0, sizeof(float) * FE));
    for (int i = 0; i < P->nrot; ++i) TRY(dalloc_guarded(P, &P->F[i], FE, P->Bp));
    const int64_t u = P->lo_units();
    for (int i = 0; i < P->nrot; ++i) TRY(dalloc_guarded(P, &P->Lo[i], FE * u, P->Bp * u));
    TRY(dalloc_guarded(P, &P->Zlo, FE * u, P->Bp * u));
    HIPX(hipMemset(P->Zlo, 0, sizeof(uint16_t) * FE * u));
    const bool fwd_only = (P->flags & USCT_FLAG_FORWARD_ONLY) != 0;
    const int nW = fwd_only ? 0 : P->K + 2;
    P->Wslot.resize(nW);
    for (int i = 0; i < nW; ++i) TRY(dalloc_guarded(P, &P->Wslot[i], FE, P->Bp));
    const int nC = fwd_only ? 0 : 2 * std::max(P->NS - 2, 0);
    P->CK.resize(nC);
    for (int i = 0; i < nC; ++i) TRY(dalloc_guarded(P, &P->CK[i], FE, P->Bp));
    P->CKlo.resize(nC);
    for (int i = 0; i < nC; ++i) TRY(dalloc_guarded(P, &P->CKlo[i], FE * u, P->Bp * u));
    const int64_t ntr = (int64_t)P->n_rec * P->Bp;
    TRY(dalloc_t(P, &P->recb, (int64_t)P->nt * ntr));
    TRY(dalloc_t(P, &P->recb_lo, (int64_t)P->nt * ntr));
    TRY(dalloc_t(P, &P->dobsT, (int64_t)P->nt * ntr));
    TRY(dalloc_t(P, &P->dobsT_lo, (int64_t)P->nt * ntr));
    TRY(dalloc_t(P, &P->res, (int64_t)P->nt * ntr));
    // sized for the narrowest lane group the autotuner may pick (16 lanes)
    TRY(dalloc_t(P, &P->g_hi, P->N * std::max(P->Bp / 16, 1)));
    TRY(dalloc_t(P, &P->g_lo, P->N * std::max(P->Bp / 16, 1)));
    TRY(dalloc_t(P, &P->sinc, (int64_t)P->nt * P->Bp));
    TRY(dalloc_t(P, &P->sincd, (int64_t)P->nt * P->Bp));
    TRY(dalloc_t(P, &P->srcd2, (int64_t)P->nt * P->Bp));
    TRY(dalloc_t(P, &P->src_cell, P->Bp));
    TRY(dalloc_t(P, &P->gstats, 3 * ntr));
    TRY(dalloc_t(P, &P->tmask_cur, ntr));
    // per-pass source cells and trace masks
    std::vector<int> src((size_t)P->passes * P->Bp, -1);
    std::vector<unsigned char> tm((size_t)P->passes * ntr, 0);
    for (int i = 0; i < P->n_tx; ++i) {
        src[i] = (int)P->tx[i];
        const int p = i / P->Bp, b = i % P->Bp;
        for (int j = 0; j < P->n_rec; ++j) tm[(size_t)p * ntr + (size_t)j * P->Bp + b] = 1;
    }
    TRY(dalloc_t(P, &P->src_all, (int64_t)src.size()));
    HIPX(hipMemcpy(P->src_all, src.data(), sizeof(int) * src.size(), hipMemcpyHostToDevice));
    TRY(dalloc_t(P, &P->tmask, (int64_t)tm.size()));
    HIPX(hipMemcpy(P->tmask, tm.data(), tm.size(), hipMemcpyHostToDevice));
    return 0;
}

int allocate_plain(usct_plan *P) {
    const int64_t N = P->N;
    TRY(dalloc_t(P, &P->Z, N));
    HIPX(hipMemset(P->Z, 0, sizeof(float) * N));
    for (int i = 0; i < 4; ++i) TRY(dalloc_t(P, &P->F[i], N));
    for (int i = 0; i < 4; ++i) TRY(dalloc_t(P, &P->Lo[i], N * P->lo_units()));
    TRY(dalloc_t(P, &P->Zlo, N * P->lo_units()));
    HIPX(hipMemset(P->Zlo, 0, sizeof(uint16_t) * N * P->lo_units()));
    // full history, levels 1..nt-1 (slot 0 unused)
    const int64_t hist = (int64_t)P->nt * N * 4;
    if (hist + P->bytes > P->budget)
        return fail(USCT_ENOMEM, "plain plan: full history (%lld bytes) exceeds the budget",
                    (long long)hist);
    float *H = nullptr;
    TRY(dalloc_t(P, &H, (int64_t)P->nt * N));
    P->Wslot.resize(P->nt);
    for (int n = 0; n < P->nt; ++n) P->Wslot[n] = H + (int64_t)n * N;
    const int nr = P->n_rec;
    TRY(dalloc_t(P, &P->recb, (int64_t)P->nt * nr));
    TRY(dalloc_t(P, &P->recb_lo, (int64_t)P->nt * nr));
    TRY(dalloc_t(P, &P->dobsT, (int64_t)P->nt * nr));
    TRY(dalloc_t(P, &P->dobsT_lo, (int64_t)P->nt * nr));
    TRY(dalloc_t(P, &P->res, (int64_t)P->nt * nr));
    TRY(dalloc_t(P, &P->g_hi, N));
    TRY(dalloc_t(P, &P->g_lo, N));
    TRY(dalloc_t(P, &P->sinc, P->nt));
    TRY(dalloc_t(P, &P->sincd, P->nt));
    TRY(dalloc_t(P, &P->srcd2, P->nt));
    TRY(dalloc_t(P, &P->gstats, 3 * (int64_t)nr));
    TRY(dalloc_t(P, &P->tmask, nr));
    TRY(dalloc_t(P, &P->tmask_cur, nr));
    std::vector<unsigned char> tm(nr, 1);
    HIPX(hipMemcpy(P->tmask, tm.data(), nr, hipMemcpyHostToDevice));
    std::vector<int> src(P->n_tx);
    for (int i = 0; i < P->n_tx; ++i) src[i] = (int)P->tx[i];
    TRY(dalloc_t(P, &P->src_all, std::max(P->n_tx, 1)));
    if (P->n_tx)
        HIPX(hipMemcpy(P->src_all, src.data(), sizeof(int) * P->n_tx, hipMemcpyHostToDevice));
    return 0;
}

// Time a few forward (adjoint = false) or adjoint launches of the current
// tile choice of that kernel, rotating over the plan's buffers.
double time_tile(usct_plan *P, bool adjoint, int reps = 4, int trials = 1) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    const int lv = adjoint ? adj_levels(P) : fwd_levels(P);
    // the launches write time levels 1 .. lv (traces, source rows): they must exist
    if (lv >= 3 && (P->nt < 4 || P->nrot < 5)) return 1e30;
    const bool x2 = lv == 2 && P->nt >= 3;
    // The forward pass of a gradient records traces as it goes: time the same
    // kernel instances (they differ in registers and schedule from the ones that
    // do not record); the launches write trace row 1, which every forward pass
    // writes again.
    auto step = [&](int r) {
        Lvl L[kMaxRot];
        for (int i = 0; i < P->nrot; ++i) L[i] = {P->F[(r + i) % P->nrot], P->Lo[(r + i) % P->nrot]};
        if (!adjoint && lv >= 3) {
            launch_fwd3_batched(P, P->stream, L[1], L[0], &L[2], 1, true);
        } else if (adjoint && lv >= 3) {
            float *out[3] = {L[2].hi, L[3].hi, L[4].hi};
            const float *pf[3] = {P->Wslot[0], P->Wslot[0], P->Wslot[0]};
            launch_adj3_batched(P, P->stream, L[0].hi, L[1].hi, out, pf, 3);
        } else if (!adjoint) {
            if (x2) launch_fwd2_batched(P, P->stream, L[1], L[0], L[2], L[3], 1, true);
            else launch_fwd_batched(P, P->stream, L[1], L[0], L[2], 1, true);
        } else if (P->rev) {
            launch_radj_batched(P, P->stream, P->Wslot[r % 3], P->Wslot[(r + 1) % 3],
                                P->Wslot[(r + 2) % 3], L[0], L[1], L[2], true, 1);
        } else if (x2) {
            launch_adj2_batched(P, P->stream, L[0].hi, L[1].hi, L[2].hi, L[3].hi, P->Wslot[0],
                                P->Wslot[0], 2);
        } else {
            launch_adj_batched(P, P->stream, L[0].hi, L[1].hi, L[2].hi, P->Wslot[0],
                               P->Wslot[0], P->Wslot[0], 1);
        }
    };
    step(0);  // warm-up
    float ms = 1e30f;
    for (int tr = 0; tr < trials; ++tr) {
        (void)hipEventRecord(a, P->stream);
        for (int r = 0; r < reps; ++r) step(r);
        (void)hipEventRecord(b, P->stream);
        (void)hipEventSynchronize(b);
        float t = 1e30f;
        if (hipGetLastError() != hipSuccess) {
            ms = 1e30f;
            break;
        }
        (void)hipEventElapsedTime(&t, a, b);
        ms = std::min(ms, t * 4.0f / (float)reps);  // per four launches, as before
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    // per time level: candidates advancing different numbers of levels compare
    return lv >= 3 ? ms / lv : (x2 ? ms / 2.0 : ms);
}

// Autotuner: lane-group width (among those that divide the padded shot count),
// cells per lane and rows per wavefront, for the forward and the adjoint kernel
// separately.  Results are kept per (ndim, dims, padded shots, sponge) for the
// life of the process.
// Plans share tuned tiles when the candidates and kernels they time are the same:
// grid, shots per pass, sponge, reversal, fusion (whether two-level kernels are
// candidates) and the low part's type (which kernel kinds can take it).
struct TuneKey {
    int ndim, nz, ny, nx, Bp, sponge, rev, fuse, lo32, levels;
    bool operator<(const TuneKey &o) const {
        return std::tie(ndim, nz, ny, nx, Bp, sponge, rev, fuse, lo32, levels) <
               std::tie(o.ndim, o.nz, o.ny, o.nx, o.Bp, o.sponge, o.rev, o.fuse, o.lo32, o.levels);
    }
};
struct TuneVal { Tiling tf, ta; };
std::map<TuneKey, TuneVal> g_tuned;
std::mutex g_tuned_mu;  // plans may be created on several host threads

int autotune(usct_plan *P) {
    // The timing launches advance level 1 (two-level candidates also level 2) and
    // read that row of the source and residual arrays; with nt = 1 no level is
    // ever advanced, so there is nothing to time and row 1 does not exist.
    if (P->nt < 2) return 0;
    const TuneKey key{P->ndim, P->nz, P->ny, P->nx, P->Bp, P->sponge ? 1 : 0, P->rev ? 1 : 0,
                      fusable(P) ? 1 : 0, P->lo32 ? 1 : 0,
                      P->flags & (USCT_FLAG_LEVELS_2 | USCT_FLAG_LEVELS_3)};
    // USCT_TUNE_KIND=k keeps only the candidates of tile kind k, where there are
    // any (tests use it to run a kernel that the timing would not pick); such a
    // choice is neither taken from nor kept in the cache
    const char *only_kind = std::getenv("USCT_TUNE_KIND");
    const bool restricted = only_kind && *only_kind;
    if (!restricted) {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        auto hit = g_tuned.find(key);
        if (hit != g_tuned.end()) {
            P->tf = hit->second.tf;
            P->ta = hit->second.ta;
            return 0;
        }
    }
    // zero the fields used by the timing launches
    for (int i = 0; i < P->nrot; ++i) {
        HIPX(hipMemset(P->F[i], 0, sizeof(float) * P->field_elems()));
        HIPX(hipMemset(P->Lo[i], 0, sizeof(uint16_t) * P->field_elems() * P->lo_units()));
    }
    for (size_t i = 0; i < std::min<size_t>(P->Wslot.size(), 3); ++i)
        HIPX(hipMemset(P->Wslot[i], 0, sizeof(float) * P->field_elems()));
    HIPX(hipMemset(P->sinc, 0, sizeof(float) * P->nt * P->Bp));
    HIPX(hipMemset(P->sincd, 0, sizeof(double) * P->nt * P->Bp));
    HIPX(hipMemset(P->sp2, 0, sizeof(float2) * P->N));
    HIPX(hipMemset(P->g_hi, 0, sizeof(float) * P->N * std::max(P->Bp / 16, 1)));
    HIPX(hipMemset(P->g_lo, 0, sizeof(float) * P->N * std::max(P->Bp / 16, 1)));
    HIPX(hipMemset(P->res, 0, sizeof(float) * P->nt * P->n_rec * P->Bp));
    HIPX(hipMemcpy(P->src_cell, P->src_all, sizeof(int) * P->Bp, hipMemcpyDeviceToDevice));
    std::vector<int> lbs;
    for (int c : {64, 32, 16})
        if (P->Bp % c == 0) lbs.push_back(c);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const char *log = std::getenv("USCT_TUNE_LOG");  // print every candidate's time
    for (int adjoint = 0; adjoint < 2; ++adjoint) {
        if (adjoint && P->Wslot.empty()) break;  // forward-only plan
        std::vector<Tiling> cands;
        if (P->ndim == 2) {
            // the two-level kernels also try the chunk orders and march directions
            const bool x2 = adjoint ? fused_adj(P) : fused_fwd(P);
            const bool only3 = x2 && P->nrot >= kMaxRot;
            for (int lb : lbs)
                for (int cx : {4, 8}) {
                    if (only3) break;
                    // rows per wavefront: powers of two, and the counts that make the
                    // grid one, two or three whole rounds of resident wavefronts
                    // (3 per SIMD), so that no round runs part-empty
                    std::vector<int> rys{4, 8, 16, 32, 64};
                    const int64_t per_chunk =
                        (P->Bp / lb) * cdiv(P->nx, (int64_t)(kWave / lb) * cx);
                    for (int k = 1; k <= 3; ++k) {
                        const int64_t nyc =
                            std::max<int64_t>(1, cdiv((int64_t)k * cus * 4 * 3, per_chunk));
                        const int ry = (int)cdiv(P->ny, nyc);
                        if (ry >= 2 && std::find(rys.begin(), rys.end(), ry) == rys.end())
                            rys.push_back(ry);
                    }
                    for (int ry : rys)
                        for (int sc = 0; sc < (x2 ? 8 : 1); ++sc)
                            cands.push_back({lb, cx, 1, ry, 0, sc});
                }
            // three levels per launch (four cells per lane, two wavefronts per SIMD)
            if (only3 && P->nt >= 4)
                for (int lb : lbs) {
                    std::vector<int> rys{8, 16, 32, 64};
                    const int64_t per_chunk = (P->Bp / lb) * cdiv(P->nx, (int64_t)(kWave / lb) * 4);
                    for (int k = 1; k <= 3; ++k) {
                        const int64_t nyc =
                            std::max<int64_t>(1, cdiv((int64_t)k * cus * 4 * 2, per_chunk));
                        const int ry = (int)cdiv(P->ny, nyc);
                        if (ry >= 3 && std::find(rys.begin(), rys.end(), ry) == rys.end())
                            rys.push_back(ry);
                    }
                    for (int ry : rys)
                        for (int sc = 0; sc < 8; ++sc)
                            cands.push_back({lb, 4, 1, ry, 0, sc, kBlk3dDefault, 3});
                }
        } else if (adjoint && P->rev) {
            for (int lb : lbs)
                for (auto pc : {std::pair<int, int>{2, 2}, {2, 1}, {1, 2}, {1, 1}})
                    for (int rz : {4, 8, 16, 32, 64}) cands.push_back({lb, pc.first, pc.second, rz, 0});
            // work-groups of four strips stacked along y that pass their edge rows
            // through LDS (tile kind 4; the grid covers the y strips in blocks of 4).
            // On 256^3 with 64 shots it reads 27.7 instead of 31.3 B per shot-cell
            // and times level with kind 0 (2 x 2 patches; 2 x 1 patches 1.7x slower)
            for (int lb : lbs)
                for (int rz : {8, 16}) cands.push_back({lb, 2, 2, rz, 4, 0, (1 << 8) | kYsh});
        } else {
            for (int lb : lbs)
                for (auto pc : {std::pair<int, int>{4, 4}, {4, 2}, {8, 2}, {2, 2}, {2, 4}, {2, 1},
                                {1, 2}, {1, 1}})
                    for (int rz : {4, 8, 16, 32, 64}) {
                        // the adjoint reduces a lane's CX*CY cells over its lane group
                        if (adjoint && pc.first * pc.second > lb) continue;
                        cands.push_back({lb, pc.first, pc.second, rz, 0});
                    }
            // LDS-staged work-group tiles (TX x TY cells)
            for (int lb : lbs) {
                std::vector<std::pair<int, int>> tiles;
                if (lb == 16) tiles = {{8, 8}, {16, 8}};
                else if (lb == 32) tiles = {{8, 8}};
                else tiles = {{8, 4}};
                for (auto tl : tiles)
                    for (int rz : {8, 16, 32, 64}) cands.push_back({lb, tl.first, tl.second, rz, 1});
            }
            // forward: work-groups of four y-stacked strips sharing their edge rows
            // (tile kind 4, 64-lane groups, 2 x 2 patches)
            if (!adjoint && std::find(lbs.begin(), lbs.end(), 64) != lbs.end())
                for (int rz : {8, 16, 32}) cands.push_back({64, 2, 2, rz, 4, 0, (1 << 8) | kYshF});
            // two forward levels per launch on 8 x 8 tiles of 16 shots
            if (!adjoint && P->Bp % 16 == 0 && fusable(P))
                for (int rz : {8, 16, 32, 64}) cands.push_back({16, 8, 8, rz, 2});
            // two forward levels per launch on lane patches, marching z in registers
            // (long z chunks: the planes above and below a chunk cost it the most;
            // 1 x 2 measured level with 2 x 1)
            if (!adjoint && fusable(P))
                for (int lb : lbs)
                    for (auto pc : {std::pair<int, int>{1, 1}, {2, 1}})
                        for (int rz : {32, 64}) cands.push_back({lb, pc.first, pc.second, rz, 3});
        }
        if (P->lo32 && !adjoint)  // the LDS and two-level 3D forwards take bfloat16 low parts
            cands.erase(std::remove_if(cands.begin(), cands.end(),
                                       [](const Tiling &c) { return c.lds != 0 && c.lds != 4; }),
                        cands.end());
        if (restricted) {
            const int k = std::atoi(only_kind);
            if (std::any_of(cands.begin(), cands.end(), [k](const Tiling &c) { return c.lds == k; }))
                cands.erase(std::remove_if(cands.begin(), cands.end(),
                                           [k](const Tiling &c) { return c.lds != k; }),
                            cands.end());
        }
        Tiling &T = adjoint ? P->ta : P->tf;
        double best = 1e300;
        Tiling bt = T;
        std::vector<std::pair<double, Tiling>> timed;
        for (const Tiling &c : cands) {
            T = c;
            const double t = time_tile(P, adjoint != 0);
            if (log && *log == '1')
                std::fprintf(stderr,
                             "[usct tune] %s lanes %d %s %dx%d rows %d sched %d levels %d: %.3f ms\n",
                             adjoint ? "adjoint" : "forward", c.LB,
                             c.lds == 4 ? "patch-ysh"
                             : c.lds == 3 ? "patch-x2"
                                          : (c.lds == 2 ? "lds-tile-x2" : (c.lds ? "lds-tile" : "patch")),
                             c.CX, c.CY, c.RY, c.sched, c.lv, t);
            timed.emplace_back(t, c);
            if (t < best) { best = t; bt = c; }
        }
        // 2D launches are short and one trial of four ranks candidates only within
        // a few per cent: the six fastest are timed again in three trials of eight
        // launches and the fastest trial decides (one 4-launch trial sometimes
        // ranked 32 rows first on the 512^2, 128-shot pass, a grid whose second
        // round runs a third full and whose adjoint launches are 14 % slower than
        // 43 rows).  3D launches take milliseconds and keep the single trial.
        if (P->ndim == 2 && timed.size() > 1) {
            std::stable_sort(timed.begin(), timed.end(),
                             [](const std::pair<double, Tiling> &x,
                                const std::pair<double, Tiling> &y) { return x.first < y.first; });
            best = 1e300;
            for (size_t i = 0; i < std::min<size_t>(timed.size(), 6); ++i) {
                T = timed[i].second;
                const double t = time_tile(P, adjoint != 0, 8, 3);
                if (log && *log == '1')
                    std::fprintf(stderr, "[usct tune] %s re-timed rows %d sched %d lanes %d cells %d: %.3f ms\n",
                                 adjoint ? "adjoint" : "forward", T.RY, T.sched, T.LB, T.CX, t);
                if (t < best) { best = t; bt = T; }
            }
        }
        T = bt;
        // 3D patch kernels: the order of the strips (blocks of bx x by strips) for
        // the chosen tile, so that strips sharing halo cells run together
        if (P->ndim == 3 && (T.lds == 0 || T.lds == 3)) {
            for (int blk : {(2 << 8) | 2, (4 << 8) | 4}) {
                T.blk = blk;
                const double t = time_tile(P, adjoint != 0);
                if (log && *log == '1')
                    std::fprintf(stderr, "[usct tune] %s strip blocks %dx%d: %.3f ms\n",
                                 adjoint ? "adjoint" : "forward", blk >> 8, blk & 255, t);
                if (t < best) { best = t; bt = T; }
            }
            T = bt;
        }
    }
    if (!restricted) {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        g_tuned[key] = TuneVal{P->tf, P->ta};
    }
    return 0;
}

int build_receivers(usct_plan *P) {
    std::vector<int> head(P->N, -1), next(std::max(P->n_rec, 1), -1);
    // chains in ascending receiver order (fwi.py:321-322 subtracts j = 0, 1, ...)
    for (int j = P->n_rec - 1; j >= 0; --j) {
        const int64_t c = P->rec[j];
        next[j] = head[c];
        head[c] = j;
    }
    TRY(dalloc_t(P, &P->rec_head, P->N));
    TRY(dalloc_t(P, &P->rec_next, (int64_t)next.size()));
    HIPX(hipMemcpy(P->rec_head, head.data(), sizeof(int) * P->N, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(P->rec_next, next.data(), sizeof(int) * next.size(), hipMemcpyHostToDevice));
    const int64_t rows = P->N / P->nx;
    std::vector<int> rr(rows, 0);
    for (int j = 0; j < P->n_rec; ++j) rr[P->rec[j] / P->nx]++;
    TRY(dalloc_t(P, &P->row_rec, rows));
    HIPX(hipMemcpy(P->row_rec, rr.data(), sizeof(int) * rows, hipMemcpyHostToDevice));
    return 0;
}

int validate(int ndim, const int *dims, double h, double dt, int nt, const int64_t *tx, int n_tx,
             const int64_t *rec, int n_rec, int64_t *N) {
    if (ndim != 2 && ndim != 3) return fail(USCT_EINVAL, "ndim must be 2 or 3 (got %d)", ndim);
    int64_t n = 1;
    for (int d = 0; d < ndim; ++d) {
        if (dims[d] < 1) return fail(USCT_EINVAL, "dims[%d] = %d must be positive", d, dims[d]);
        n *= dims[d];
    }
    if (n >= (int64_t)1 << 31) return fail(USCT_EINVAL, "grid of %lld cells exceeds 2^31", (long long)n);
    // kernels address a row (2D) or plane (3D) of all shots with 32-bit byte offsets
    const int64_t plane = ndim == 2 ? dims[1] : (int64_t)dims[1] * dims[2];
    if (plane * 4 * (int64_t)((n_tx + 63) / 64 * 64) >= ((int64_t)1 << 32))
        return fail(USCT_EINVAL, "a grid %s of %lld cells times %d shots exceeds 32-bit offsets",
                    ndim == 2 ? "row" : "plane", (long long)plane, n_tx);
    if (!(h > 0) || !(dt > 0)) return fail(USCT_EINVAL, "h and dt must be positive");
    if (nt < 1) return fail(USCT_EINVAL, "nt must be >= 1 (got %d)", nt);
    if (n_tx < 0 || n_rec < 0) return fail(USCT_EINVAL, "negative shot or receiver count");
    if (n_rec < 1) return fail(USCT_EINVAL, "at least one receiver is required");
    for (int i = 0; i < n_tx; ++i)
        if (tx[i] < 0 || tx[i] >= n)
            return fail(USCT_EINVAL, "tx_lin[%d] = %lld outside the grid of %lld cells", i,
                        (long long)tx[i], (long long)n);
    for (int j = 0; j < n_rec; ++j)
        if (rec[j] < 0 || rec[j] >= n)
            return fail(USCT_EINVAL, "rec_lin[%d] = %lld outside the grid of %lld cells", j,
                        (long long)rec[j], (long long)n);
    *N = n;
    return 0;
}

int enter(usct_plan *P, void *stream) {
    hipStream_t user = (hipStream_t)stream;
    HIPX(hipEventRecord(P->ev_in, user));
    HIPX(hipStreamWaitEvent(P->stream, P->ev_in, 0));
    return 0;
}

int leave(usct_plan *P, void *stream) {
    hipStream_t user = (hipStream_t)stream;
    HIPX(hipEventRecord(P->ev_out, P->stream));
    HIPX(hipStreamWaitEvent(user, P->ev_out, 0));
    return 0;
}

// Born state and v/m buffers, allocated once per plan (Gauss-Newton only).
// The Born fields and v/m carry the guard zones of the forward fields: the
// two-level Born kernel reads halo columns past the first and last cell.
// Both are held against the plan's memory budget: a call whose buffers do not
// fit fails with USCT_ENOMEM before anything is allocated.
int fits_budget(const usct_plan *P, int64_t need, const char *what) {
    if (P->bytes + need > P->budget)
        return fail(USCT_ENOMEM,
                    "%s need %lld bytes; the plan holds %lld of its %lld-byte memory budget",
                    what, (long long)need, (long long)P->bytes, (long long)P->budget);
    return 0;
}

int ensure_gn(usct_plan *P) {
    if (P->vm2) return 0;
    const int64_t FE = P->field_elems();
    const int64_t g = kGuardCells * std::max<int64_t>(P->Bp, 1);
    const int64_t need = 4 * ((FE + 2 * g) * 4 + (2 * cdiv(FE, 2) + 2 * g) * 2) +
                         (P->N + 2 * kGuardCells) * 8;
    TRY(fits_budget(P, need, "the Gauss-Newton Born state and v/m"));
    for (int i = 0; i < 4; ++i) {
        TRY(dalloc_guarded(P, &P->Dh[i], FE, P->Bp));
        TRY(dalloc_guarded(P, &P->Dl[i], 2 * cdiv(FE, 2), P->Bp));
    }
    return dalloc_guarded(P, &P->vm2, P->N, 1);
}

// Time-window traces of a pass (batched layout), allocated on first use.
int ensure_windows(usct_plan *P) {
    if (P->twT) return 0;
    const int64_t count = (int64_t)P->nt * P->n_rec * P->Bp;
    TRY(fits_budget(P, 2 * count * (int64_t)sizeof(float), "the time-window traces"));
    TRY(dalloc_t(P, &P->twT_lo, count));
    return dalloc_t(P, &P->twT, count);
}

int run(usct_plan *P, int mode, int misfit, const double *m, const double *w,
        const double *sponge, const double *dobs, double *data, double *J, double *grad,
        void *stream, Profile *pr, const double *v = nullptr, const double *tw = nullptr) {
    if (!P) return fail(USCT_EINVAL, "null plan");
    if (!m || !w) return fail(USCT_EINVAL, "m and wavelet are required");
    if (misfit != USCT_MISFIT_L2 && misfit != USCT_MISFIT_GCN)
        return fail(USCT_EINVAL, "unknown misfit type %d", misfit);
    if ((sponge != nullptr) != P->sponge)
        return fail(USCT_EINVAL, sponge ? "sponge passed to a plan created without USCT_FLAG_SPONGE"
                                        : "plan created with USCT_FLAG_SPONGE needs a sponge");
    if ((mode == MODE_MISFIT || mode == MODE_GRAD) && !dobs)
        return fail(USCT_EINVAL, "dobs is required");
    if (mode == MODE_ADJSRC && !dobs) return fail(USCT_EINVAL, "the adjoint source is required");
    if ((mode == MODE_GRAD || mode == MODE_GNHV || mode == MODE_ADJSRC) &&
        (P->flags & USCT_FLAG_FORWARD_ONLY))
        return fail(USCT_ENOTSUP, "plan was created with USCT_FLAG_FORWARD_ONLY");
    if (mode == MODE_GNHV && !v) return fail(USCT_EINVAL, "v is required");
    if (mode == MODE_GNHV && P->lo32)
        return fail(USCT_ENOTSUP, "Gauss-Newton products are not implemented for plans with "
                                  "USCT_FLAG_FP32_LO");
    if (P->n_tx == 0) {
        TRY(enter(P, stream));
        if (J) HIPX(hipMemsetAsync(J, 0, sizeof(double), P->stream));
        if (grad) HIPX(hipMemsetAsync(grad, 0, sizeof(double) * P->N, P->stream));
        return leave(P, stream);
    }
    if (mode == MODE_GNHV) TRY(ensure_gn(P));
    if (tw) TRY(ensure_windows(P));
    TRY(enter(P, stream));
    if (mode == MODE_GNHV)
        hipLaunchKernelGGL(k_setup_vm, dim3(grid_for(P->N)), dim3(256), 0, P->stream, v, m,
                           P->vm2, P->N);
    if (P->fp) {
        // the call's wavelet and S_c w_i of every transmit cell, read by the
        // (captured) footprint kernels from plan-owned buffers
        HIPX(hipMemcpyAsync(P->wav, w, sizeof(double) * P->nt, hipMemcpyDeviceToDevice,
                            P->stream));
        const double dt2 = P->dt * P->dt;
        const int64_t ne = (int64_t)P->fp_SE * P->passes;
        hipLaunchKernelGGL(k_fp_coef, dim3(grid_for(ne)), dim3(256), 0, P->stream, P->fp_ec,
                           P->fp_ew, P->fp_ecoef, ne, m, dt2);
        if (P->fp_nse)
            hipLaunchKernelGGL(k_fp_coef, dim3(grid_for(P->fp_nse)), dim3(256), 0, P->stream,
                               P->fp_sc, P->fp_sw, P->fp_scoef, P->fp_nse, m, dt2);
    }
    int rc = P->method == USCT_METHOD_PLAIN
                 ? run_plain(P, mode, misfit, m, w, sponge, dobs, data, J, grad, pr, tw)
                 : run_batched(P, mode, misfit, m, w, sponge, dobs, data, J, grad, pr, tw);
    int rc2 = leave(P, stream);
    return rc ? rc : rc2;
}

// Element groups in CSR form (group g: cells[off[g] .. off[g+1]) with weights w).
int check_groups_arg(const char *what, int n, const int64_t *off, const int64_t *cells,
                     const double *w, int64_t N) {
    if (n > 0 && (!off || !cells || !w)) return fail(USCT_EINVAL, "%s groups: null argument", what);
    if (n > 0 && off[0] != 0) return fail(USCT_EINVAL, "%s_off[0] must be 0", what);
    for (int g = 0; g < n; ++g) {
        if (off[g + 1] <= off[g]) return fail(USCT_EINVAL, "%s group %d is empty", what, g);
        for (int64_t i = off[g]; i < off[g + 1]; ++i) {
            if (cells[i] < 0 || cells[i] >= N)
                return fail(USCT_EINVAL, "%s group %d: cell %lld outside the grid of %lld cells",
                            what, g, (long long)cells[i], (long long)N);
            if (!std::isfinite(w[i]))
                return fail(USCT_EINVAL, "%s group %d: weight %lld is not finite", what, g,
                            (long long)(i - off[g]));
        }
    }
    return 0;
}

template <class T>
int upload(usct_plan *P, T **dst, const std::vector<T> &v) {
    TRY(dalloc_t(P, dst, std::max<int64_t>((int64_t)v.size(), 1)));
    if (!v.empty()) HIPX(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return 0;
}

// Device lists of a plan whose elements are groups of cells (its passes fixed):
// the stencil kernels get no point source and no receivers, and the footprint
// kernels take the transmit entries per pass, the receiver groups and, per pass,
// the fix cells (every receiver cell and the pass's transmit cells) with their
// receiver entries in receiver order and their transmit entries.
int attach_groups(usct_plan *P, const int64_t *tx_off, const int64_t *tx_cells,
                  const double *tx_w, const int64_t *rec_off, const int64_t *rec_cells,
                  const double *rec_w) {
    const int Bp = P->Bp, passes = P->passes, n_rec = P->n_rec;
    if ((int64_t)passes * Bp > 0)
        HIPX(hipMemset(P->src_all, 0xff, sizeof(int) * (size_t)passes * Bp));
    HIPX(hipMemset(P->rec_head, 0xff, sizeof(int) * (size_t)P->N));
    HIPX(hipMemset(P->row_rec, 0, sizeof(int) * (size_t)(P->N / P->nx)));
    std::vector<int> roff(n_rec + 1, 0), rcell;
    std::vector<double> rw;
    std::map<int, std::vector<std::pair<int, float>>> rec_at;
    for (int j = 0; j < n_rec; ++j) {
        for (int64_t i = rec_off[j]; i < rec_off[j + 1]; ++i) {
            rcell.push_back((int)rec_cells[i]);
            rw.push_back(rec_w[i]);
            rec_at[(int)rec_cells[i]].push_back({j, (float)rec_w[i]});
        }
        roff[j + 1] = (int)rcell.size();
    }
    // transmit entries per pass; a cell repeated within one group is driven once
    // with the summed weight
    std::vector<std::vector<std::tuple<int, int, double>>> ent(std::max(passes, 1));
    std::vector<std::map<int, std::vector<std::pair<int, double>>>> src_at(std::max(passes, 1));
    for (int t = 0; t < P->n_tx; ++t) {
        const int p = t / Bp, b = t % Bp;
        std::vector<std::pair<int, double>> cw;
        for (int64_t i = tx_off[t]; i < tx_off[t + 1]; ++i) {
            const int c = (int)tx_cells[i];
            auto it = std::find_if(cw.begin(), cw.end(), [c](const auto &e) { return e.first == c; });
            if (it == cw.end()) cw.push_back({c, tx_w[i]});
            else it->second += tx_w[i];
        }
        for (auto &e : cw) {
            ent[p].push_back({b, e.first, e.second});
            src_at[p][e.first].push_back({b, e.second});
        }
    }
    int SE = 0, FC = 0;
    std::vector<std::vector<int>> fix(std::max(passes, 1));
    for (int p = 0; p < passes; ++p) {
        SE = std::max(SE, (int)ent[p].size());
        std::vector<int> &f = fix[p];
        for (auto &kv : rec_at) f.push_back(kv.first);
        for (auto &kv : src_at[p])
            if (!rec_at.count(kv.first)) f.push_back(kv.first);
        std::sort(f.begin(), f.end());
        FC = std::max(FC, (int)f.size());
    }
    std::vector<int> eb((size_t)passes * SE, -1), ec((size_t)passes * SE, -1);
    std::vector<double> ew((size_t)passes * SE, 0.0);
    std::vector<int> fcell((size_t)passes * FC, -1), frb((size_t)passes * FC, 0),
        fre((size_t)passes * FC, 0), fsb((size_t)passes * FC, 0), fse((size_t)passes * FC, 0);
    std::vector<int> rj, sb, sc;
    std::vector<float> rwf;
    std::vector<double> sw;
    for (int p = 0; p < passes; ++p) {
        for (size_t e = 0; e < ent[p].size(); ++e) {
            const size_t q = (size_t)p * SE + e;
            eb[q] = std::get<0>(ent[p][e]);
            ec[q] = std::get<1>(ent[p][e]);
            ew[q] = std::get<2>(ent[p][e]);
        }
        for (size_t f = 0; f < fix[p].size(); ++f) {
            const size_t q = (size_t)p * FC + f;
            const int c = fix[p][f];
            fcell[q] = c;
            frb[q] = (int)rj.size();
            auto r = rec_at.find(c);
            if (r != rec_at.end())
                for (auto &e : r->second) {
                    rj.push_back(e.first);
                    rwf.push_back(e.second);
                }
            fre[q] = (int)rj.size();
            fsb[q] = (int)sb.size();
            auto t = src_at[p].find(c);
            if (t != src_at[p].end())
                for (auto &e : t->second) {
                    sb.push_back(e.first);
                    sc.push_back(c);
                    sw.push_back(e.second);
                }
            fse[q] = (int)sb.size();
        }
    }
    if (rj.size() > (size_t)INT32_MAX || sb.size() > (size_t)INT32_MAX)
        return fail(USCT_EINVAL, "footprint lists exceed 2^31 entries");
    std::vector<int> ids(std::max(passes, 1));
    for (int p = 0; p < (int)ids.size(); ++p) ids[p] = p;
    TRY(upload(P, &P->fp_roff, roff));
    TRY(upload(P, &P->fp_rcell, rcell));
    TRY(upload(P, &P->fp_rw, rw));
    TRY(upload(P, &P->fp_eb, eb));
    TRY(upload(P, &P->fp_ec, ec));
    TRY(upload(P, &P->fp_ew, ew));
    TRY(upload(P, &P->fp_fcell, fcell));
    TRY(upload(P, &P->fp_frb, frb));
    TRY(upload(P, &P->fp_fre, fre));
    TRY(upload(P, &P->fp_fsb, fsb));
    TRY(upload(P, &P->fp_fse, fse));
    TRY(upload(P, &P->fp_rj, rj));
    TRY(upload(P, &P->fp_rwf, rwf));
    TRY(upload(P, &P->fp_sb, sb));
    TRY(upload(P, &P->fp_sc, sc));
    TRY(upload(P, &P->fp_sw, sw));
    TRY(upload(P, &P->fp_pass_ids, ids));
    TRY(dalloc_t(P, &P->fp_pass, 1));
    HIPX(hipMemcpy(P->fp_pass, ids.data(), sizeof(int), hipMemcpyHostToDevice));
    TRY(dalloc_t(P, &P->fp_ecoef, std::max<int64_t>((int64_t)eb.size(), 1)));
    TRY(dalloc_t(P, &P->fp_scoef, std::max<int64_t>((int64_t)sb.size(), 1)));
    TRY(dalloc_t(P, &P->wav, P->nt));
    P->fp_SE = SE;
    P->fp_FC = FC;
    P->fp_nse = (int64_t)sb.size();
    P->fp = true;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Forward without device allocation (usct_forward_fmc_noalloc)
//
// The plain one-shot-at-a-time forward with every buffer it needs (S' pairs,
// the zero level and three rotating levels with their low parts, receiver
// chains, source terms, one shot's traces) carved from one file-scope device
// array.  Each device the library runs on holds its own copy of the array
// until the process ends; host code reaches a device's copy only through the
// address hipGetSymbolAddress returns for it on that device, kept per device.
// Calls on one device share the array: calls that overlap in time on one
// device (different streams or host threads) race; calls ordered on one
// stream do not.  Everything a call reads before writing (the zero level,
// receiver chains, trace row 0) is cleared by kernels at the start of the call.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kScratchBytes = (int64_t)64 << 20;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch` (looked up once per
// device, with that device current).
int scratch_ptr(void **out) {
    int dev = 0;
    HIPX(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(USCT_ENOTSUP, "device %d: the static workspace supports devices 0..%d", dev,
                    kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPX(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(USCT_EHIP, "hipGetSymbolAddress returned a null workspace address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Consecutive 256-byte aligned pieces of a workspace (base nullptr: sizing only).
struct Carve {
    char *base;
    int64_t off = 0;
    template <class T>
    T *take(int64_t count) {
        T *p = reinterpret_cast<T *>(base + off);
        off += cdiv(std::max<int64_t>(count, 1) * (int64_t)sizeof(T), 256) * 256;
        return p;
    }
};

struct NoallocBufs {
    float2 *sp2;
    float *Z, *F[3];
    uint16_t *Zlo, *Lo[3];
    int *head, *next, *tx32;
    int64_t *rec64, *tx64;
    float *sinc, *srcd2, *recb, *recb_lo;
    double *sincd;
};

int64_t carve_noalloc(char *base, int64_t N, int nt, int n_tx, int n_rec, NoallocBufs &B) {
    Carve c{base};
    B.sp2 = c.take<float2>(N);
    B.Z = c.take<float>(N);
    for (int i = 0; i < 3; ++i) B.F[i] = c.take<float>(N);
    // low parts padded to whole 4-byte words (cleared as floats)
    B.Zlo = c.take<uint16_t>(2 * cdiv(N, 2));
    for (int i = 0; i < 3; ++i) B.Lo[i] = c.take<uint16_t>(N);
    B.head = c.take<int>(N);
    B.next = c.take<int>(n_rec);
    B.rec64 = c.take<int64_t>(n_rec);
    B.tx64 = c.take<int64_t>(n_tx);
    B.tx32 = c.take<int>(n_tx);
    B.sinc = c.take<float>(nt);
    B.sincd = c.take<double>(nt);
    B.srcd2 = c.take<float>(nt);
    B.recb = c.take<float>((int64_t)nt * n_rec);
    B.recb_lo = c.take<float>((int64_t)nt * n_rec);
    return c.off;
}

// Host arrays of the libuap interface (int cells, host memory) staged into HBM
// for a plan: the device buffers of one call, freed when it returns.
struct HostStage {
    std::vector<void *> bufs;
    ~HostStage() {
        for (void *p : bufs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, int64_t count) {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(count, 1)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(USCT_ENOMEM, "staging buffer of %lld bytes",
                        (long long)(sizeof(T) * count));
        }
        bufs.push_back(q);
        *p = (T *)q;
        return 0;
    }
};

int widen_cells(const int *a, int n, const char *what, std::vector<int64_t> &out) {
    if (n > 0 && !a) return fail(USCT_EINVAL, "%s is null", what);
    out.assign(std::max(n, 0), 0);
    for (int i = 0; i < n; ++i) out[i] = a[i];
    return 0;
}

int64_t cells_of(const int *dims, int ndim) {
    int64_t n = 1;
    for (int d = 0; d < ndim; ++d) n *= std::max(dims[d], 0);
    return n;
}

}  // namespace

extern "C" {

int usct_plan_create(usct_plan **out, int ndim, const int *dims, double h, double dt, int nt,
                     const int64_t *tx_lin, int n_tx, const int64_t *rec_lin, int n_rec,
                     int method, int ckpt_every, int64_t mem_budget, int flags) {
    if (!out || !dims) return fail(USCT_EINVAL, "null argument");
    *out = nullptr;
    int64_t N = 0;
    TRY(validate(ndim, dims, h, dt, nt, tx_lin, n_tx, rec_lin, n_rec, &N));
    if ((flags & USCT_FLAG_SPONGE) && ndim != 2)
        return fail(USCT_ENOTSUP, "the sponge is defined for 2D grids only (solver.py:30)");
    if (method != USCT_METHOD_AUTO && method != USCT_METHOD_PLAIN && method != USCT_METHOD_BATCHED)
        return fail(USCT_EINVAL, "unknown method %d", method);
    if ((flags & USCT_FLAG_LEVELS_2) && (flags & USCT_FLAG_LEVELS_3))
        return fail(USCT_EINVAL, "USCT_FLAG_LEVELS_2 and USCT_FLAG_LEVELS_3 exclude each other");
    if ((flags & USCT_FLAG_LEVELS_3) &&
        (ndim != 2 || method == USCT_METHOD_PLAIN || (flags & (USCT_FLAG_SPONGE | USCT_FLAG_NO_FUSE))))
        return fail(USCT_ENOTSUP, "three-level launches need a batched 2D plan without the sponge "
                                  "or USCT_FLAG_NO_FUSE");
    usct_plan *P = new usct_plan();
    P->ndim = ndim;
    P->nz = ndim == 3 ? dims[0] : 1;
    P->ny = dims[ndim - 2];
    P->nx = dims[ndim - 1];
    P->N = N;
    P->h = h;
    P->dt = dt;
    P->nt = nt;
    P->n_tx = n_tx;
    P->n_rec = n_rec;
    P->tx.assign(tx_lin, tx_lin + n_tx);
    P->rec.assign(rec_lin, rec_lin + n_rec);
    P->method = method == USCT_METHOD_PLAIN ? USCT_METHOD_PLAIN : USCT_METHOD_BATCHED;
    P->flags = flags;
    P->sponge = (flags & USCT_FLAG_SPONGE) != 0;
    P->lo32 = (flags & USCT_FLAG_FP32_LO) != 0;
    if (ndim == 3) P->tf = P->ta = Tiling{64, 2, 2, 32, 0};
    int rc = 0;
    auto bail = [&](int code) {
        free_plan_memory(P);
        if (P->stream) (void)hipStreamDestroy(P->stream);
        if (P->ev_in) (void)hipEventDestroy(P->ev_in);
        if (P->ev_out) (void)hipEventDestroy(P->ev_out);
        delete P;
        return code;
    };
    if (hipGetDevice(&P->dev) != hipSuccess) return bail(fail(USCT_EHIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&P->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&P->ev_out, hipEventDisableTiming) != hipSuccess)
        return bail(fail(USCT_EHIP, "stream/event creation failed"));
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
        return bail(fail(USCT_EHIP, "hipMemGetInfo failed"));
    // memory kept from destroyed plans (usct_device_cache) is free for this one
    P->budget = mem_budget > 0 ? mem_budget
                               : (int64_t)(0.85 * (double)((int64_t)free_b + kept_on(P->dev)));
    if ((rc = build_receivers(P))) return bail(rc);
    if ((rc = dalloc_guarded(P, &P->sp2, N, 1))) return bail(rc);
    if (P->sponge && (rc = dalloc_t(P, &P->spg, N))) return bail(rc);
    if ((rc = dalloc_t(P, &P->J_acc, 1))) return bail(rc);
    if ((rc = dalloc_t(P, &P->part, P->npart))) return bail(rc);
    if (P->method == USCT_METHOD_PLAIN) {
        P->tf = P->ta = Tiling{1, 1, 1, 1, 0};
        P->Bp = 1;
        P->passes = n_tx;
        P->K = nt;
        P->NS = 1;
        if ((rc = allocate_plain(P))) return bail(rc);
    } else {
        if (n_tx == 0) {
            P->Bp = 64; P->passes = 0; P->K = std::max(nt - 1, 2); P->NS = 1;
        } else {
            P->budget -= P->bytes;
            if ((rc = choose_batched(P, ckpt_every))) return bail(rc);
            P->budget += P->bytes;
            if ((rc = allocate_batched(P))) return bail(rc);
            if (P->nrot >= kMaxRot) {  // three time levels per fused 2D launch
                P->tf.lv = P->ta.lv = 3;
                P->tf.CX = P->ta.CX = 4;
            }
            default_rows(P, P->tf);
            default_rows(P, P->ta);
            if ((flags & USCT_FLAG_AUTOTUNE) && (rc = autotune(P))) return bail(rc);
        }
    }
    *out = P;
    return 0;
}

int usct_plan_create_groups(usct_plan **out, int ndim, const int *dims, double h, double dt,
                            int nt, int n_tx, const int64_t *tx_off, const int64_t *tx_cells,
                            const double *tx_w, int n_rec, const int64_t *rec_off,
                            const int64_t *rec_cells, const double *rec_w, int method,
                            int ckpt_every, int64_t mem_budget, int flags) {
    if (!out || !dims) return fail(USCT_EINVAL, "null argument");
    *out = nullptr;
    if (ndim != 2 && ndim != 3) return fail(USCT_EINVAL, "ndim must be 2 or 3 (got %d)", ndim);
    int64_t N = 1;
    for (int d = 0; d < ndim; ++d) {
        if (dims[d] < 1) return fail(USCT_EINVAL, "dims[%d] = %d must be positive", d, dims[d]);
        N *= dims[d];
    }
    if (n_tx < 0 || n_rec < 1) return fail(USCT_EINVAL, "need n_tx >= 0 and n_rec >= 1");
    TRY(check_groups_arg("tx", n_tx, tx_off, tx_cells, tx_w, N));
    TRY(check_groups_arg("rec", n_rec, rec_off, rec_cells, rec_w, N));
    std::vector<int64_t> tx_first(std::max(n_tx, 1)), rec_first(n_rec);
    for (int t = 0; t < n_tx; ++t) tx_first[t] = tx_cells[tx_off[t]];
    for (int j = 0; j < n_rec; ++j) rec_first[j] = rec_cells[rec_off[j]];
    if (flags & USCT_FLAG_REVERSE)
        return fail(USCT_ENOTSUP, "footprint plans keep forward levels (USCT_FLAG_REVERSE)");
    if (flags & USCT_FLAG_FP32_LO)
        return fail(USCT_ENOTSUP, "USCT_FLAG_FP32_LO is not implemented for footprint plans");
    usct_plan *P = nullptr;
    // the footprint corrections read stored forward levels
    TRY(usct_plan_create(&P, ndim, dims, h, dt, nt, tx_first.data(), n_tx, rec_first.data(), n_rec,
                         method, ckpt_every, mem_budget, flags | USCT_FLAG_NO_REVERSE));
    const int rc = attach_groups(P, tx_off, tx_cells, tx_w, rec_off, rec_cells, rec_w);
    if (rc) {
        usct_plan_destroy(P);
        return rc;
    }
    *out = P;
    return 0;
}

int usct_device_cache(int64_t max_bytes, int64_t *held) {
    if (max_bytes < 0) return fail(USCT_EINVAL, "max_bytes must be >= 0");
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache_cap = max_bytes;
    release_kept(-1, max_bytes);
    if (held) *held = g_kept_bytes;
    return 0;
}

int usct_plan_destroy(usct_plan *P) {
    if (!P) return 0;
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    free_plan_memory(P);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    if (P->ev_in) (void)hipEventDestroy(P->ev_in);
    if (P->ev_out) (void)hipEventDestroy(P->ev_out);
    delete P;
    return 0;
}

int usct_plan_get_info(const usct_plan *P, usct_plan_info *info) {
    if (!P || !info) return fail(USCT_EINVAL, "null argument");
    info->method = P->method;
    info->lanes_per_group = P->ta.LB;
    info->tile_x = P->ta.CX;
    info->tile_y = P->ta.CY;
    info->rows_per_wave = P->ta.RY;
    info->fwd_lanes_per_group = P->tf.LB;
    info->fwd_tile_x = P->tf.CX;
    info->fwd_tile_y = P->tf.CY;
    info->fwd_rows_per_wave = P->tf.RY;
    info->tile_kind = P->ta.lds;
    info->fwd_tile_kind = P->tf.lds;
    info->shots_padded = P->method == USCT_METHOD_PLAIN ? P->n_tx : P->Bp * P->passes;
    info->shot_passes = P->passes;
    info->shots_per_pass = P->Bp;
    info->ckpt_every = P->rev ? 0 : P->K;
    info->n_windows = P->rev ? 0 : P->NS;
    info->workspace_bytes = P->held;
    info->reverse = P->rev ? 1 : 0;
    info->sched = P->ta.sched;
    info->fwd_sched = P->tf.sched;
    info->strip_block = P->ta.blk;
    info->fwd_strip_block = P->tf.blk;
    info->levels_per_launch = adj_levels(P);
    info->fwd_levels_per_launch = fwd_levels(P);
    return 0;
}

int usct_forward_fmc(usct_plan *P, const double *m, const double *wavelet, const double *sponge,
                     double *data, void *stream) {
    if (!data) return fail(USCT_EINVAL, "data is required");
    return run(P, MODE_FWD, USCT_MISFIT_L2, m, wavelet, sponge, nullptr, data, nullptr, nullptr,
               stream, nullptr);
}

int usct_misfit(usct_plan *P, const double *m, const double *wavelet, const double *sponge,
                const double *dobs, int misfit_type, double *J, void *stream) {
    if (!J) return fail(USCT_EINVAL, "J is required");
    return run(P, MODE_MISFIT, misfit_type, m, wavelet, sponge, dobs, nullptr, J, nullptr, stream,
               nullptr);
}

int usct_misfit_and_gradient(usct_plan *P, const double *m, const double *wavelet,
                             const double *sponge, const double *dobs, int misfit_type, double *J,
                             double *grad, void *stream) {
    if (!J || !grad) return fail(USCT_EINVAL, "J and grad are required");
    return run(P, MODE_GRAD, misfit_type, m, wavelet, sponge, dobs, nullptr, J, grad, stream,
               nullptr);
}

int usct_gradient_from_source(usct_plan *P, const double *m, const double *wavelet,
                              const double *sponge, const double *adj, double *grad,
                              void *stream) {
    if (!grad) return fail(USCT_EINVAL, "grad is required");
    return run(P, MODE_ADJSRC, USCT_MISFIT_L2, m, wavelet, sponge, adj, nullptr, nullptr, grad,
               stream, nullptr);
}

int64_t usct_gsot_workspace_bytes(int n_tx, int nt, int n_rec) {
    if (n_tx < 0 || nt < 1 || n_rec < 0) return -1;
    const int64_t W = (int64_t)nt + 1, ntr = (int64_t)n_tx * n_rec;
    return ntr * W * (5 * (int64_t)sizeof(double) + 3 * (int64_t)sizeof(int));
}

int usct_gsot(const double *dsyn, const double *dobs, int n_tx, int nt, int n_rec, double eta,
              void *workspace, int64_t workspace_bytes, double *J_trace, double *adj, int *status,
              void *stream) {
    if (n_tx < 0 || n_rec < 0 || nt < 1 || nt > (1 << 20))
        return fail(USCT_EINVAL, "gsot: bad sizes n_tx %d nt %d n_rec %d", n_tx, nt, n_rec);
    const int64_t ntr = (int64_t)n_tx * n_rec;
    if (ntr > INT32_MAX) return fail(USCT_EINVAL, "gsot: too many traces");
    const int64_t need = usct_gsot_workspace_bytes(n_tx, nt, n_rec);
    if (workspace_bytes < need)
        return fail(USCT_EINVAL, "gsot: workspace of %lld bytes, %lld needed",
                    (long long)workspace_bytes, (long long)need);
    if (!status) return fail(USCT_EINVAL, "gsot: status is required");
    if (ntr > 0 && (!dsyn || !dobs || !workspace || !J_trace || !adj))
        return fail(USCT_EINVAL, "gsot: null argument");
    if (!std::isfinite(eta)) return fail(USCT_EINVAL, "gsot: eta is not finite");
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPX(hipMemsetAsync(status, 0, sizeof(int), s));
    if (ntr == 0) return 0;
    double *wsd = static_cast<double *>(workspace);
    int *wsi = reinterpret_cast<int *>(wsd + ntr * 5 * ((int64_t)nt + 1));
    hipLaunchKernelGGL(k_gsot, dim3((unsigned)ntr), dim3(256), 0, s, dsyn, dobs, nt, n_rec,
                       (int)ntr, eta * eta, wsd, wsi, J_trace, adj, status);
    return check_launch();
}

int usct_misfit_tw(usct_plan *P, const double *m, const double *wavelet, const double *sponge,
                   const double *dobs, const double *time_weights, int misfit_type, double *J,
                   void *stream) {
    if (!J) return fail(USCT_EINVAL, "J is required");
    return run(P, MODE_MISFIT, misfit_type, m, wavelet, sponge, dobs, nullptr, J, nullptr, stream,
               nullptr, nullptr, time_weights);
}

int usct_misfit_and_gradient_tw(usct_plan *P, const double *m, const double *wavelet,
                                const double *sponge, const double *dobs,
                                const double *time_weights, int misfit_type, double *J,
                                double *grad, void *stream) {
    if (!J || !grad) return fail(USCT_EINVAL, "J and grad are required");
    return run(P, MODE_GRAD, misfit_type, m, wavelet, sponge, dobs, nullptr, J, grad, stream,
               nullptr, nullptr, time_weights);
}

int usct_gn_hessian_vector(usct_plan *P, const double *m, const double *wavelet,
                           const double *sponge, const double *v, double *hv, void *stream) {
    if (!hv) return fail(USCT_EINVAL, "hv is required");
    return run(P, MODE_GNHV, USCT_MISFIT_L2, m, wavelet, sponge, nullptr, nullptr, nullptr, hv,
               stream, nullptr, v);
}

int usct_profile_gradient(usct_plan *P, const double *m, const double *wavelet,
                          const double *sponge, const double *dobs, double *fwd_ms, double *adj_ms,
                          int *n_fwd, int *n_adj, int *levels_per_launch, void *stream) {
    if (!P) return fail(USCT_EINVAL, "null plan");
    double *Jd = nullptr, *gd = nullptr;
    HIPX(hipMalloc(&Jd, sizeof(double)));
    if (hipMalloc(&gd, sizeof(double) * P->N) != hipSuccess) {
        (void)hipFree(Jd);
        return fail(USCT_ENOMEM, "profile: gradient buffer");
    }
    Profile pr;
    int rc = run(P, MODE_GRAD, USCT_MISFIT_L2, m, wavelet, sponge, dobs, nullptr, Jd, gd, stream, &pr);
    if (!rc) rc = hipStreamSynchronize(P->stream) == hipSuccess ? 0 : fail(USCT_EHIP, "sync");
    // per kind: mean over the launches that advance the most levels; levels
    // counted over all launches
    int lvf = 1, lva = 1;
    for (int l : pr.fwd_lv) lvf = std::max(lvf, l);
    for (int l : pr.adj_lv) lva = std::max(lva, l);
    auto sum = [](std::vector<hipEvent_t> &v, const std::vector<int> &lvs, int lv, double *ms,
                  int *levels) {
        double s = 0, s_any = 0;
        int c = 0, c_any = 0, total = 0;
        for (size_t i = 0; i + 1 < v.size(); i += 2) {
            float t = 0;
            const int l = i / 2 < lvs.size() ? lvs[i / 2] : 1;
            total += l;
            if (hipEventElapsedTime(&t, v[i], v[i + 1]) != hipSuccess) continue;
            s_any += t;
            ++c_any;
            if (l == lv) { s += t; ++c; }
        }
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
        if (ms) *ms = c ? s / c : (c_any ? s_any / c_any : 0.0);
        if (levels) *levels = total;
    };
    sum(pr.fwd_ev, pr.fwd_lv, lvf, fwd_ms, n_fwd);
    sum(pr.adj_ev, pr.adj_lv, lva, adj_ms, n_adj);
    if (levels_per_launch) {
        levels_per_launch[0] = lvf;
        levels_per_launch[1] = lva;
    }
    (void)hipFree(Jd);
    (void)hipFree(gd);
    return rc;
}

int usct_forward_fmc_direct(const double *m, const int *dims, int ndim, double h, double dt, int nt,
                            const int64_t *tx_lin, int n_tx, const int64_t *rec_lin, int n_rec,
                            const double *wavelet, double *data, int method, void *stream) {
    usct_plan *P = nullptr;
    TRY(usct_plan_create(&P, ndim, dims, h, dt, nt, tx_lin, n_tx, rec_lin, n_rec, method, 0, 0,
                         USCT_FLAG_NO_GRAPH));
    int rc = usct_forward_fmc(P, m, wavelet, nullptr, data, stream);
    if (!rc && hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
        rc = fail(USCT_EHIP, "stream synchronize failed");
    usct_plan_destroy(P);
    return rc;
}

int usct_misfit_and_gradient_direct(const double *m, const int *dims, int ndim, double h,
                                    double dt, int nt, const int64_t *tx_lin, int n_tx,
                                    const int64_t *rec_lin, int n_rec, const double *wavelet,
                                    const double *dobs, double *grad, double *J_host,
                                    int misfit_type, int method, void *stream) {
    if (!J_host) return fail(USCT_EINVAL, "J_host is required");
    usct_plan *P = nullptr;
    TRY(usct_plan_create(&P, ndim, dims, h, dt, nt, tx_lin, n_tx, rec_lin, n_rec, method, 0, 0,
                         USCT_FLAG_NO_GRAPH));
    double *Jd = nullptr;
    if (hipMalloc(&Jd, sizeof(double)) != hipSuccess) {
        usct_plan_destroy(P);
        return fail(USCT_ENOMEM, "J buffer");
    }
    int rc = usct_misfit_and_gradient(P, m, wavelet, nullptr, dobs, misfit_type, Jd, grad, stream);
    if (!rc && hipMemcpyAsync(J_host, Jd, sizeof(double), hipMemcpyDeviceToHost,
                              (hipStream_t)stream) != hipSuccess)
        rc = fail(USCT_EHIP, "J copy failed");
    if (!rc && hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
        rc = fail(USCT_EHIP, "stream synchronize failed");
    (void)hipFree(Jd);
    usct_plan_destroy(P);
    return rc;
}

int usct_forward_fmc_noalloc(const double *m, const int *dims, int ndim, double h, double dt,
                             int nt, const int64_t *tx_lin, int n_tx, const int64_t *rec_lin,
                             int n_rec, const double *wavelet, double *data, void *stream) {
    if (!dims) return fail(USCT_EINVAL, "null argument");
    int64_t N = 0;
    TRY(validate(ndim, dims, h, dt, nt, tx_lin, n_tx, rec_lin, n_rec, &N));
    if (!m || !wavelet || (n_tx > 0 && !data))
        return fail(USCT_EINVAL, "m, wavelet and data are required");
    NoallocBufs B;
    const int64_t need = carve_noalloc(nullptr, N, nt, n_tx, n_rec, B);
    if (need > kScratchBytes)
        return fail(USCT_ENOMEM,
                    "the forward of %lld cells, %d levels, %d transmits and %d receivers needs "
                    "%lld bytes of workspace; the static workspace holds %lld "
                    "(use usct_forward_fmc_direct or a plan)",
                    (long long)N, nt, n_tx, n_rec, (long long)need, (long long)kScratchBytes);
    if (n_tx == 0) return 0;
    void *base = nullptr;
    TRY(scratch_ptr(&base));
    carve_noalloc((char *)base, N, nt, n_tx, n_rec, B);
    hipStream_t s = (hipStream_t)stream;
    const int nz = ndim == 3 ? dims[0] : 1, ny = dims[ndim - 2], nx = dims[ndim - 1];
    const unsigned gN = grid_for(N);
    // levels <= 0 are the zero level; level n >= 1 rotates over three buffers
    auto lvl = [&](int n) { return n <= 0 ? B.Z : B.F[n % 3]; };
    auto llo = [&](int n) { return n <= 0 ? B.Zlo : B.Lo[n % 3]; };
    hipLaunchKernelGGL(k_fill, dim3(gN), dim3(256), 0, s, B.Z, N, 0.0f);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(cdiv(N, 2))), dim3(256), 0, s, (float *)B.Zlo,
                       cdiv(N, 2), 0.0f);
    hipLaunchKernelGGL(k_fill_i32, dim3(gN), dim3(256), 0, s, B.head, N, -1);
    HIPX(hipMemcpyAsync(B.rec64, rec_lin, sizeof(int64_t) * n_rec, hipMemcpyHostToDevice, s));
    HIPX(hipMemcpyAsync(B.tx64, tx_lin, sizeof(int64_t) * n_tx, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_rec_chains, dim3(1), dim3(64), 0, s, B.rec64, n_rec, B.head, B.next,
                       B.tx64, n_tx, B.tx32);
    hipLaunchKernelGGL(k_setup_s, dim3(gN), dim3(256), 0, s, m, B.sp2, N, dt * dt, 1.0 / (h * h));
    const int64_t NJ = (int64_t)nt * n_rec;
    const dim3 gt((unsigned)cdiv(NJ, kTT), 1);
    for (int b = 0; b < n_tx; ++b) {
        hipLaunchKernelGGL(k_setup_sinc, dim3(grid_for(nt)), dim3(256), 0, s, m, wavelet,
                           B.tx32 + b, B.sinc, B.sincd, nt, 1, dt * dt);
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n_rec)), dim3(256), 0, s, B.recb,
                           (int64_t)n_rec, 0.0f);
        hipLaunchKernelGGL(k_fill, dim3(grid_for(n_rec)), dim3(256), 0, s, B.recb_lo,
                           (int64_t)n_rec, 0.0f);
        for (int n = 1; n < nt; ++n)
            hipLaunchKernelGGL((k_fwd_plain<false, true>), dim3(gN), dim3(256), 0, s, lvl(n - 1),
                               lvl(n - 2), lvl(n), llo(n - 1), llo(n - 2), llo(n), B.sp2,
                               (const float *)nullptr, (int)tx_lin[b], B.sincd + n, B.head, B.next,
                               B.recb + (int64_t)n * n_rec, B.recb_lo + (int64_t)n * n_rec,
                               B.srcd2 + n, nz, ny, nx, ndim);
        hipLaunchKernelGGL(k_batched_to_data, gt, dim3(256), 0, s, B.recb, B.recb_lo, data, NJ,
                           1, b, 1);
        TRY(check_launch());
    }
    return check_launch();
}

int usct_uap_forward_fmc(const double *m, const int *dims, int ndim, double h, double dt, int nt,
                         const int *tx_lin, int n_tx, const int *rec_lin, int n_rec,
                         const double *wavelet, double *data) {
    if (!dims || !m || !wavelet || (n_tx > 0 && !data)) return fail(USCT_EINVAL, "null argument");
    if (ndim != 2 && ndim != 3) return fail(USCT_EINVAL, "ndim must be 2 or 3 (got %d)", ndim);
    std::vector<int64_t> tx, rec;
    TRY(widen_cells(tx_lin, n_tx, "tx_lin", tx));
    TRY(widen_cells(rec_lin, n_rec, "rec_lin", rec));
    const int64_t N = cells_of(dims, ndim);
    const int64_t nd = (int64_t)std::max(n_tx, 0) * std::max(nt, 0) * std::max(n_rec, 0);
    usct_plan *P = nullptr;
    TRY(usct_plan_create(&P, ndim, dims, h, dt, nt, tx.data(), n_tx, rec.data(), n_rec,
                         USCT_METHOD_BATCHED, 0, 0, USCT_FLAG_NO_GRAPH | USCT_FLAG_FORWARD_ONLY));
    HostStage st;
    double *md = nullptr, *wd = nullptr, *dd = nullptr;
    int rc = st.alloc(&md, N);
    if (!rc) rc = st.alloc(&wd, nt);
    if (!rc) rc = st.alloc(&dd, nd);
    if (!rc && hipMemcpy(md, m, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(USCT_EHIP, "copy of m failed");
    if (!rc && hipMemcpy(wd, wavelet, sizeof(double) * nt, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(USCT_EHIP, "copy of the wavelet failed");
    if (!rc) rc = usct_forward_fmc(P, md, wd, nullptr, dd, nullptr);
    if (!rc && nd && hipMemcpy(data, dd, sizeof(double) * nd, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(USCT_EHIP, "copy of the traces failed");
    usct_plan_destroy(P);
    return rc;
}

double usct_uap_misfit_and_gradient(const double *m, const int *dims, int ndim, double h,
                                    double dt, int nt, const int *tx_lin, int n_tx,
                                    const int *rec_lin, int n_rec, const double *wavelet,
                                    const double *dobs, double *grad, int misfit_type) {
    const double bad = std::nan("");
    if (!dims || !m || !wavelet || !grad || (n_tx > 0 && !dobs)) {
        fail(USCT_EINVAL, "null argument");
        return bad;
    }
    if (ndim != 2 && ndim != 3) {
        fail(USCT_EINVAL, "ndim must be 2 or 3 (got %d)", ndim);
        return bad;
    }
    std::vector<int64_t> tx, rec;
    if (widen_cells(tx_lin, n_tx, "tx_lin", tx) || widen_cells(rec_lin, n_rec, "rec_lin", rec))
        return bad;
    const int64_t N = cells_of(dims, ndim);
    const int64_t nd = (int64_t)std::max(n_tx, 0) * std::max(nt, 0) * std::max(n_rec, 0);
    usct_plan *P = nullptr;
    if (usct_plan_create(&P, ndim, dims, h, dt, nt, tx.data(), n_tx, rec.data(), n_rec,
                         USCT_METHOD_BATCHED, 0, 0, USCT_FLAG_NO_GRAPH))
        return bad;
    double J = bad;
    {
        HostStage st;
        double *md = nullptr, *wd = nullptr, *od = nullptr, *gd = nullptr, *Jd = nullptr;
        int rc = st.alloc(&md, N);
        if (!rc) rc = st.alloc(&wd, nt);
        if (!rc) rc = st.alloc(&od, nd);
        if (!rc) rc = st.alloc(&gd, N);
        if (!rc) rc = st.alloc(&Jd, 1);
        if (!rc && hipMemcpy(md, m, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(USCT_EHIP, "copy of m failed");
        if (!rc && hipMemcpy(wd, wavelet, sizeof(double) * nt, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(USCT_EHIP, "copy of the wavelet failed");
        if (!rc && nd && hipMemcpy(od, dobs, sizeof(double) * nd, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(USCT_EHIP, "copy of dobs failed");
        if (!rc) rc = usct_misfit_and_gradient(P, md, wd, nullptr, od, misfit_type, Jd, gd, nullptr);
        double Jh = bad;
        if (!rc && hipMemcpy(grad, gd, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(USCT_EHIP, "copy of the gradient failed");
        if (!rc && hipMemcpy(&Jh, Jd, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(USCT_EHIP, "copy of J failed");
        if (!rc) J = Jh;
    }
    usct_plan_destroy(P);
    return J;
}

const char *usct_last_error(void) { return g_err.c_str(); }

int usct_version(void) { return 1; }

}  // extern "C"

/*
 * usct.h -- C ABI of the MI355X acoustic FD solver and its exact discrete-adjoint
 * FWI gradient (libusct.so, built for gfx950).
 *
 * All array arguments marked "device" are device pointers in HBM (any
 * allocator: torch, hipMalloc, ...).  Index lists and dims are host arrays.
 * Fields are FP32 on the device; models, wavelets, traces (synthetic and
 * observed), gradients and misfits are FP64 at the boundary, like the
 * reference's core (software/core/libuap.hpp:109-116, 169-173).  Every entry point returns 0 on success or a
 * usct_status code, with a message in usct_last_error() (thread-local).
 * Work is ordered after prior work on `stream` (a hipStream_t, NULL = the null
 * stream) and later work on `stream` sees the results.
 *
 * Reference interfaces these replace (paths relative to the reference root):
 *   usct_forward_fmc_direct         software/core/libuap.hpp:109-116 (uap::forward_fmc)
 *                                   software/python/_uap.cpp:25-37 (_uap.forward_fmc)
 *   usct_misfit_and_gradient_direct software/core/libuap.hpp:169-224
 *                                   (uap::misfit_and_gradient),
 *                                   software/python/_uap.cpp:41-55
 *   usct_plan_create_groups / usct_misfit_tw / usct_misfit_and_gradient_tw
 *                                   finite-aperture elements and time windows
 *                                   (simulation/ringfwi/fwi.py:74-89, 246-332),
 *                                   which the reference runs in Python only
 *                                   (fwi.py:594-615)
 *   usct_gn_hessian_vector          simulation/ringfwi/fwi.py:410-431
 *   usct_uap_forward_fmc            software/core/libuap.hpp:109-116, same argument types
 *   usct_uap_misfit_and_gradient    software/core/libuap.hpp:169-224, same argument types
 *   usct_plan_* / usct_forward_fmc / usct_misfit / usct_misfit_and_gradient
 *                                   the same three operations (plus misfit-only,
 *                                   software/python/uap.py:56-67) split into a
 *                                   reusable plan, as the reference's CuPy backend
 *                                   re-uses its kernels across calls
 *                                   (software/python/uap_gpu.py:166-233).
 */
#ifndef USCT_H
#define USCT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    USCT_OK = 0,
    USCT_EINVAL = 1,   /* bad argument: shape, index out of range, unknown misfit */
    USCT_ENOMEM = 2,   /* workspace does not fit the memory budget / HBM */
    USCT_EHIP = 3,     /* HIP runtime error */
    USCT_ENOTSUP = 4   /* valid request this build does not implement */
} usct_status;

/* Execution method of a plan. */
enum {
    USCT_METHOD_AUTO = 0,    /* batched */
    USCT_METHOD_PLAIN = 1,   /* one launch per shot per time step, full history */
    USCT_METHOD_BATCHED = 2  /* all shots per launch, checkpointed, hipGraph-replayed */
};

/* Plan flags. */
enum {
    USCT_FLAG_SPONGE = 1,    /* calls pass a multiplicative sponge mask (2D only) */
    USCT_FLAG_AUTOTUNE = 2,  /* time candidate tile shapes at creation, keep the best */
    USCT_FLAG_NO_GRAPH = 4,  /* batched: launch eagerly instead of replaying a hipGraph */
    USCT_FLAG_FORWARD_ONLY = 8, /* no gradient workspace (forward_fmc / misfit only) */
    USCT_FLAG_NO_FUSE = 16,     /* batched 2D: one time level per launch instead of two */
    USCT_FLAG_NO_REVERSE = 32,  /* batched 3D gradients: keep forward levels (whole history or
                                   checkpoints) instead of running the forward backwards */
    USCT_FLAG_REVERSE = 64,     /* batched 3D gradients: run the forward backwards alongside the
                                   adjoint even where the whole history would fit */
    USCT_FLAG_FP32_LO = 128,    /* carry the forward state as FP32 hi + FP32 lo (8 bytes per
                                   shot-cell instead of 6; about 48 significant bits instead of 32):
                                   for time axes that end soon after the scattered arrivals, where
                                   the traces' rounding is a visible share of a small residual;
                                   point-element plans; Gauss-Newton products return USCT_ENOTSUP */
    USCT_FLAG_LEVELS_2 = 256,   /* batched 2D without sponge: fused launches advance two time
                                   levels (the default) */
    USCT_FLAG_LEVELS_3 = 512    /* batched 2D without sponge: fused launches advance three time
                                   levels (fewer bytes per level, more registers and arithmetic:
                                   slower than two on MI355X at 512^2 x 128 shots); one more
                                   rotation buffer; USCT_ENOTSUP where launches cannot be fused */
};

/* Misfit codes (software/python/uap.py:41). */
enum { USCT_MISFIT_L2 = 0, USCT_MISFIT_GCN = 1 };

typedef struct usct_plan usct_plan;

typedef struct {
    int method;            /* USCT_METHOD_PLAIN or USCT_METHOD_BATCHED */
    int lanes_per_group;   /* adjoint kernel: shots per wavefront lane group (16, 32, 64);
                              1 for plain */
    int tile_x, tile_y;    /* adjoint kernel: cells per lane along x (and y in 3D) */
    int rows_per_wave;     /* adjoint kernel: y rows (2D) or z planes (3D) per wavefront */
    int shots_padded;      /* shot count rounded up to a lane-group multiple */
    int shot_passes;       /* batched: passes over shot subsets that fit the budget */
    int shots_per_pass;    /* batched: padded shots per pass */
    int ckpt_every;        /* checkpoint interval K (>= nt-1: whole history stored) */
    int n_windows;         /* recompute windows per pass */
    int64_t workspace_bytes;  /* device memory held by the plan (with usct_device_cache: the
                                 blocks it holds, which may exceed what it asked for) */
    int fwd_lanes_per_group;  /* the same four for the forward kernel */
    int fwd_tile_x, fwd_tile_y;
    int fwd_rows_per_wave;
    int tile_kind;            /* adjoint, 3D: 0 = cells per lane (tile_x x tile_y),
                                 1 = LDS-staged work-group tile of tile_x x tile_y cells,
                                 4 = cells per lane in work-groups of four strips stacked
                                 along y that pass their edge rows through LDS (plans with
                                 reverse = 1) */
    int fwd_tile_kind;        /* the same for the forward kernel; 2 = LDS-staged tiles
                                 advancing two time levels per launch; 3 = cells per lane
                                 advancing two time levels per launch (z marched in
                                 registers); 4 = as for the adjoint (64-lane groups) */
    int reverse;              /* 1: the gradient keeps no forward levels; the adjoint sweep
                                 runs the forward backwards (3D, no sponge, point elements) */
    int sched;                /* 2D two-level adjoint kernel: row-chunk order and march
                                 (bit 0: odd chunks march upwards; bit 1: chunk index fastest;
                                 bit 2: shot group slowest) */
    int fwd_sched;            /* the same for the two-level forward kernel */
    int strip_block;          /* 3D patch adjoint kernel: wavefront order over blocks of
                                 (strip_block >> 8) x (strip_block & 255) x-by-y strips */
    int fwd_strip_block;      /* the same for the forward kernel */
    int levels_per_launch;    /* time levels one adjoint launch advances (1, 2 or 3) */
    int fwd_levels_per_launch;  /* the same for the forward */
} usct_plan_info;

/*
 * Create a plan for a grid, a time axis, transmit cells and receiver cells.
 *   dims      ndim entries, row-major (2D: ny, nx; 3D: nz, ny, nx)
 *   tx_lin    n_tx linear source cells (one shot each; uap.py:25-30)
 *   rec_lin   n_rec linear receiver cells (duplicates allowed and accumulate)
 *   ckpt_every  0 = choose from the memory budget (3D: whole history, checkpoints or
 *               running the forward backwards, whichever streams the fewest bytes)
 *   mem_budget  bytes of HBM the plan may hold; 0 = 85% of free memory
 */
int usct_plan_create(usct_plan **plan, int ndim, const int *dims, double h, double dt,
                     int nt, const int64_t *tx_lin, int n_tx, const int64_t *rec_lin,
                     int n_rec, int method, int ckpt_every, int64_t mem_budget, int flags);

/*
 * Create a plan whose elements are finite apertures: each transmit and each
 * receiver is a group of cells with weights (simulation/ringfwi/geometry.py:212-278
 * build_footprints; fwi.py:74-89 drive every cell of the transmit group with
 * wavelet * w_i and record sum_i w_i p[cell_i] per receiver group; fwi.py:316-319
 * spread the adjoint source the same way).  Groups are in CSR form: group g
 * holds cells[off[g] .. off[g+1]) with weights w[...], off[0] = 0, no empty
 * group.  Other arguments as usct_plan_create.  Such plans advance one time
 * level per launch; every entry point taking a plan accepts them.
 */
int usct_plan_create_groups(usct_plan **plan, int ndim, const int *dims, double h, double dt,
                            int nt, int n_tx, const int64_t *tx_off, const int64_t *tx_cells,
                            const double *tx_w, int n_rec, const int64_t *rec_off,
                            const int64_t *rec_cells, const double *rec_w, int method,
                            int ckpt_every, int64_t mem_budget, int flags);
int usct_plan_destroy(usct_plan *plan);

/*
 * Device memory kept from destroyed plans for later plans, per process: at most
 * max_bytes (0, the default: nothing is kept, and what is kept is freed).  While a
 * cap is set, a plan's buffers are 2 MiB multiples taken from a kept block of its
 * device within 1.25x of their size where one exists (hipMalloc otherwise), and
 * usct_plan_destroy keeps them instead of calling hipFree, which waits for the
 * whole device.  A hipMalloc that fails frees the device's kept blocks and is tried
 * once more; kept memory counts as free for a plan's default budget.  *held (may
 * be NULL): bytes kept after the call.  No counterpart in the reference, whose
 * calls hold no device memory between them; for loops that create a plan per
 * request (openusct_amd.serve).
 */
int usct_device_cache(int64_t max_bytes, int64_t *held);
int usct_plan_get_info(const usct_plan *plan, usct_plan_info *info);

/*
 * Full-matrix-capture forward (fwi.py:102-119).
 *   m        device, N doubles, squared slowness
 *   wavelet  device, nt doubles
 *   sponge   device, N doubles or NULL (plan must have USCT_FLAG_SPONGE if not NULL)
 *   data     device, n_tx * nt * n_rec doubles (row 0 of each shot is zero)
 */
int usct_forward_fmc(usct_plan *plan, const double *m, const double *wavelet,
                     const double *sponge, double *data, void *stream);

/* Misfit only (fwi.py:246-263).  dobs: device n_tx*nt*n_rec doubles; J: device double. */
int usct_misfit(usct_plan *plan, const double *m, const double *wavelet,
                const double *sponge, const double *dobs, int misfit_type, double *J,
                void *stream);

/* Misfit and exact discrete-adjoint gradient w.r.t. m (fwi.py:266-332).
 * grad: device N doubles; J: device double. */
int usct_misfit_and_gradient(usct_plan *plan, const double *m, const double *wavelet,
                             const double *sponge, const double *dobs, int misfit_type,
                             double *J, double *grad, void *stream);

/* Windowed misfit (fwi.py:246-263 with time_weights, fwi.py:255-262): the misfit
 * of W*dsyn against W*dobs per transmit, and with it the gradient, whose adjoint
 * source is multiplied by W again (fwi.py:298-301).  time_weights: device
 * n_tx*nt*n_rec doubles in the layout of dobs, or NULL (no window).  The first
 * windowed call allocates the plan's window traces (8 bytes per trace sample of
 * a pass) and fails with USCT_ENOMEM when they would exceed its mem_budget. */
int usct_misfit_tw(usct_plan *plan, const double *m, const double *wavelet,
                   const double *sponge, const double *dobs, const double *time_weights,
                   int misfit_type, double *J, void *stream);
int usct_misfit_and_gradient_tw(usct_plan *plan, const double *m, const double *wavelet,
                                const double *sponge, const double *dobs,
                                const double *time_weights, int misfit_type, double *J,
                                double *grad, void *stream);

/* Exact discrete-adjoint gradient for a caller-supplied adjoint source: the
 * adjoint half of fwi.py:266-332 with `res` given (fwi.py:295-301 computes it
 * there from the chosen misfit).  adj = dJ/d(traces), device n_tx*nt*n_rec
 * doubles in the layout of dobs (row 0 of each transmit is never injected, as
 * the reference's loop stops at k = 1); it is injected as -R^T adj
 * (fwi.py:316-322, footprints included).  grad: device N doubles, dJ/dm.
 * With it any misfit of the forward traces (envelope, normalised envelope,
 * traveltime, optimal transport; fwi.py:134-243) is differentiated through the
 * device adjoint. */
int usct_gradient_from_source(usct_plan *plan, const double *m, const double *wavelet,
                              const double *sponge, const double *adj, double *grad,
                              void *stream);

/* Graph-space optimal transport misfit (fwi.py:193-217) of device traces
 * dsyn, dobs (n_tx*nt*n_rec doubles, layout of dobs): per trace the optimal
 * assignment between the samples (tau_i = i/nt, s_i/A) and (tau_j, o_j/A),
 * A = max|o| + 1e-30, for the cost (s_i - o_j)^2 + eta^2 (tau_i - tau_j)^2
 * (the reference solves it with scipy's linear_sum_assignment; here a
 * Hungarian method runs in one work-group per trace).  J_trace: device n_tx*n_rec
 * doubles, the assignment's cost per trace (NaN for a trace with non-finite
 * samples); adj: device, layout of dsyn, the adjoint source 2 (s_i -
 * o_sigma(i)) / A.  workspace: device, usct_gsot_workspace_bytes() bytes,
 * brought by the caller (the call allocates nothing); status: device int, set
 * to 0, then 1 if any trace had non-finite samples.  The reference's eta is 1. */
int64_t usct_gsot_workspace_bytes(int n_tx, int nt, int n_rec);
int usct_gsot(const double *dsyn, const double *dobs, int n_tx, int nt, int n_rec, double eta,
              void *workspace, int64_t workspace_bytes, double *J_trace, double *adj,
              int *status, void *stream);

/* Gauss-Newton Hessian-vector product hv = (J^T J) v summed over the plan's
 * transmits (simulation/ringfwi/fwi.py:410-431, gn_hessian_vector): the exact
 * discrete Born forward of v (fwi.py:350-378) on each transmit's background
 * wavefield, then the exact adjoint of those Born traces (fwi.py:382-407),
 * divided by m.  v, hv: device N doubles.  The background wavefield is
 * recomputed each call (checkpointed, as for the gradient); the plan needs
 * gradient workspace (no USCT_FLAG_FORWARD_ONLY) and allocates the Born state
 * (four levels of the plan's field size, 6 bytes per element) on first use;
 * that first call fails with USCT_ENOMEM when the Born state would take the
 * plan past its mem_budget.
 * Plans that advance two levels per launch (2D, no sponge, no footprints)
 * advance the background and the Born field together. */
int usct_gn_hessian_vector(usct_plan *plan, const double *m, const double *wavelet,
                           const double *sponge, const double *v, double *hv, void *stream);

/*
 * Time one gradient launch by launch with HIP events on the plan's stream.
 * levels_per_launch: int[2], time levels one forward / one adjoint launch
 *   advances (2 for two-level kernels, else 1);
 * fwd_ms / adj_ms: mean duration of the forward / adjoint launches that advance
 *   that many levels (a trailing single-level launch is not averaged in);
 * n_fwd / n_adj: time levels advanced in all (re-computed levels count as forward).
 */
int usct_profile_gradient(usct_plan *plan, const double *m, const double *wavelet,
                          const double *sponge, const double *dobs, double *fwd_ms,
                          double *adj_ms, int *n_fwd, int *n_adj, int *levels_per_launch,
                          void *stream);

/* Stateless forms with the argument order of uap::forward_fmc and
 * uap::misfit_and_gradient (device pointers as above, J returned on the host). */
int usct_forward_fmc_direct(const double *m, const int *dims, int ndim, double h, double dt,
                            int nt, const int64_t *tx_lin, int n_tx, const int64_t *rec_lin,
                            int n_rec, const double *wavelet, double *data, int method,
                            void *stream);
int usct_misfit_and_gradient_direct(const double *m, const int *dims, int ndim, double h,
                                    double dt, int nt, const int64_t *tx_lin, int n_tx,
                                    const int64_t *rec_lin, int n_rec, const double *wavelet,
                                    const double *dobs, double *grad, double *J_host,
                                    int misfit_type, int method, void *stream);

/* The libuap interface itself, host memory in and out: the exact argument types
 * of uap::forward_fmc (software/core/libuap.hpp:109-112) and
 * uap::misfit_and_gradient (libuap.hpp:169-173, returning J), so a C or MEX
 * caller written against libuap binds unchanged.  m, wavelet, dobs, data and
 * grad are host arrays; tx_lin / rec_lin are int linear cells.  Each call
 * stages its arrays into HBM, runs a batched plan (created and destroyed
 * inside the call) on the null stream and copies the results back before it
 * returns.  usct_uap_forward_fmc returns a usct_status;
 * usct_uap_misfit_and_gradient returns J, or NaN on failure with the reason in
 * usct_last_error(). */
int usct_uap_forward_fmc(const double *m, const int *dims, int ndim, double h, double dt, int nt,
                         const int *tx_lin, int n_tx, const int *rec_lin, int n_rec,
                         const double *wavelet, double *data);
double usct_uap_misfit_and_gradient(const double *m, const int *dims, int ndim, double h,
                                    double dt, int nt, const int *tx_lin, int n_tx,
                                    const int *rec_lin, int n_rec, const double *wavelet,
                                    const double *dobs, double *grad, int misfit_type);

/* usct_forward_fmc_direct with method USCT_METHOD_PLAIN, allocating no device
 * memory.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture elsewhere in their program, tight
 * serving loops, harnesses that forbid hipMalloc).  Arguments and results as
 * usct_forward_fmc_direct (m, wavelet, data: device; tx_lin, rec_lin: host);
 * the traces equal its plain-method traces bit for bit.  Every buffer the
 * computation needs lives in one static device array of 64 MiB that each
 * device the library runs on holds until the process ends; a problem needs
 * about 36 bytes per grid cell plus 8 per trace sample (nt * n_rec), and
 * larger ones fail with USCT_ENOMEM before any work is queued.  Calls on one
 * device share that array: calls that overlap in time on one device (other
 * streams or host threads) race; calls ordered on one stream do not.  The call
 * queues its work on `stream` and returns; host index arrays may be reused
 * when it returns. */
int usct_forward_fmc_noalloc(const double *m, const int *dims, int ndim, double h, double dt,
                             int nt, const int64_t *tx_lin, int n_tx, const int64_t *rec_lin,
                             int n_rec, const double *wavelet, double *data, void *stream);

const char *usct_last_error(void);
int usct_version(void);

#ifdef __cplusplus
}
#endif
#endif /* USCT_H */

/*
 * usct_bw.h -- C ABI of libusct_bw.so: streaming kernels that measure the HBM
 * bandwidth one gfx950 device reaches (no counterpart in the reference, which
 * publishes no bandwidth figures; measurement infrastructure for the solver's
 * roofline and for calibrating the FETCH_SIZE / WRITE_SIZE counters).
 */
#ifndef USCT_BW_H
#define USCT_BW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { USCT_BW_READ = 0, USCT_BW_WRITE = 1, USCT_BW_COPY = 2 };

/*
 * Stream `bytes` (a multiple of 16) with 16-byte accesses: READ reads src,
 * WRITE fills dst, COPY reads src and writes dst.  The grid has blocks_per_cu
 * work-groups of 256 lanes per CU (grid-stride loops, four loads in flight per
 * lane); nontemporal != 0 uses non-temporal loads and stores.  partials: device
 * buffer of (CUs x blocks_per_cu) floats for READ's per-work-group sums.  One
 * untimed launch, then `reps` launches timed with HIP events on `stream`;
 * *ms_out = mean milliseconds per launch.  Returns 0, 1 (bad argument) or
 * 3 (HIP error).  Buffers are device pointers.
 */
int usct_bw_run(int kind, const void *src, void *dst, float *partials, int64_t bytes,
                int blocks_per_cu, int nontemporal, int reps, float *ms_out, void *stream);

/*
 * The same in the single-pass layout: a grid of usct_bw_pass_groups(bytes,
 * per_lane) work-groups of 256 lanes covers the buffer once, each lane making
 * per_lane (1, 2, 4 or 8) 16-byte accesses 4 KiB apart and retiring.
 * partials: that many floats for READ.  Returns as usct_bw_run.
 */
int usct_bw_run_pass(int kind, const void *src, void *dst, float *partials, int64_t bytes,
                     int per_lane, int nontemporal, int reps, float *ms_out, void *stream);

/* Work-groups of the single-pass grid (READ's partials), or -1. */
int64_t usct_bw_pass_groups(int64_t bytes, int per_lane);

/* Compute units of the current device, or -1. */
int usct_bw_device_cus(void);

#ifdef __cplusplus
}
#endif
#endif /* USCT_BW_H */

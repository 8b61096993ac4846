// usct_bw.hip -- streaming kernels that measure what HBM bandwidth one gfx950
// device reaches (libusct_bw.so): read-only, write-only and copy over a buffer
// much larger than the L2s and the Infinity Cache, 16-byte accesses, in two
// layouts: grid-stride loops over a grid of a chosen number of work-groups per
// CU, and single-pass grids that cover the buffer once.  They give the stencil kernels' bandwidth its ceiling on
// the same device, and the counter runs a copy of known size to calibrate the
// FETCH_SIZE / WRITE_SIZE counters against.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/usct_bw.h"

namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 4;  // 16-byte loads or stores in flight per lane per iteration
// a native vector type: the non-temporal builtins take it (HIP's float4 is a class)
typedef float v4f __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ v4f ld4(const v4f *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool NT>
__device__ __forceinline__ void st4(v4f *p, v4f v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Read every 16-byte element of src once; one float per work-group goes to out so that
// the loads cannot be dropped.
template <bool NT>
__global__ void __launch_bounds__(kBlock) k_bw_read(const v4f *__restrict__ src, int64_t n,
                                                     float *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    float acc = 0.0f;
    for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
        v4f v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) v[u] = ld4<NT>(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    for (; i < n; i += stride) {
        const v4f v = ld4<NT>(src + i);
        acc += (v.x + v.y) + (v.z + v.w);
    }
    __shared__ float red[kBlock];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <bool NT>
__global__ void __launch_bounds__(kBlock) k_bw_write(v4f *__restrict__ dst, int64_t n, float v) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const v4f w = v4f{v, v, v, v};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        st4<NT>(dst + i, w);
}

template <bool NT>
__global__ void __launch_bounds__(kBlock) k_bw_copy(const v4f *__restrict__ src,
                                                     v4f *__restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
        v4f v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) v[u] = ld4<NT>(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) st4<NT>(dst + i + u * stride, v[u]);
    }
    for (; i < n; i += stride) st4<NT>(dst + i, ld4<NT>(src + i));
}

// Single-pass forms: a grid that covers the buffer once.  Work-group g owns
// the U * 256 consecutive elements from g * U * 256; lane t touches t, t + 256,
// ...  Each wavefront issues its U accesses and retires, so the device keeps
// every wavefront slot busy with fresh requests (on gfx950 this streams faster
// than the grid-stride loops, whose few resident wavefronts each wait on their
// own loads).
template <int U, bool NT>
__global__ void __launch_bounds__(kBlock) k_bw_read_pass(const v4f *__restrict__ src, int64_t n,
                                                          float *__restrict__ out) {
    const int64_t base = (int64_t)blockIdx.x * kBlock * U + threadIdx.x;
    float acc = 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = base + (int64_t)u * kBlock;
        if (i < n) {
            const v4f v = ld4<NT>(src + i);
            acc += (v.x + v.y) + (v.z + v.w);
        }
    }
    // the work-group's sum (reduced through LDS) to out keeps the loads live
    __shared__ float red[kBlock];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <int U, bool NT>
__global__ void __launch_bounds__(kBlock) k_bw_write_pass(v4f *__restrict__ dst, int64_t n, float v) {
    const int64_t base = (int64_t)blockIdx.x * kBlock * U + threadIdx.x;
    const v4f w = v4f{v, v, v, v};
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = base + (int64_t)u * kBlock;
        if (i < n) st4<NT>(dst + i, w);
    }
}

template <int U, bool NT>
__global__ void __launch_bounds__(kBlock) k_bw_copy_pass(const v4f *__restrict__ src,
                                                          v4f *__restrict__ dst, int64_t n) {
    const int64_t base = (int64_t)blockIdx.x * kBlock * U + threadIdx.x;
    v4f v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = base + (int64_t)u * kBlock;
        v[u] = i < n ? ld4<NT>(src + i) : v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = base + (int64_t)u * kBlock;
        if (i < n) st4<NT>(dst + i, v[u]);
    }
}

template <int U, bool NT>
void launch_pass_u(int kind, const v4f *src, v4f *dst, int64_t n, float *red, hipStream_t s) {
    const dim3 g((unsigned)((n + (int64_t)kBlock * U - 1) / ((int64_t)kBlock * U)));
    if (kind == USCT_BW_READ) hipLaunchKernelGGL((k_bw_read_pass<U, NT>), g, dim3(kBlock), 0, s, src, n, red);
    else if (kind == USCT_BW_WRITE)
        hipLaunchKernelGGL((k_bw_write_pass<U, NT>), g, dim3(kBlock), 0, s, dst, n, 1.0f);
    else hipLaunchKernelGGL((k_bw_copy_pass<U, NT>), g, dim3(kBlock), 0, s, src, dst, n);
}

template <bool NT>
void launch_pass_nt(int kind, int per_lane, const v4f *src, v4f *dst, int64_t n, float *red,
                    hipStream_t s) {
    switch (per_lane) {
        case 1: launch_pass_u<1, NT>(kind, src, dst, n, red, s); break;
        case 2: launch_pass_u<2, NT>(kind, src, dst, n, red, s); break;
        case 4: launch_pass_u<4, NT>(kind, src, dst, n, red, s); break;
        default: launch_pass_u<8, NT>(kind, src, dst, n, red, s); break;
    }
}

void launch(int kind, bool nt, const v4f *src, v4f *dst, int64_t n, int grid,
            float *red, hipStream_t s) {
    if (kind == USCT_BW_READ) {
        if (nt) hipLaunchKernelGGL(k_bw_read<true>, dim3(grid), dim3(kBlock), 0, s, src, n, red);
        else hipLaunchKernelGGL(k_bw_read<false>, dim3(grid), dim3(kBlock), 0, s, src, n, red);
    } else if (kind == USCT_BW_WRITE) {
        if (nt) hipLaunchKernelGGL(k_bw_write<true>, dim3(grid), dim3(kBlock), 0, s, dst, n, 1.0f);
        else hipLaunchKernelGGL(k_bw_write<false>, dim3(grid), dim3(kBlock), 0, s, dst, n, 1.0f);
    } else {
        if (nt) hipLaunchKernelGGL(k_bw_copy<true>, dim3(grid), dim3(kBlock), 0, s, src, dst, n);
        else hipLaunchKernelGGL(k_bw_copy<false>, dim3(grid), dim3(kBlock), 0, s, src, dst, n);
    }
}

}  // namespace

extern "C" {

int usct_bw_run(int kind, const void *src, void *dst, float *partials, int64_t bytes,
                int blocks_per_cu, int nontemporal, int reps, float *ms_out, void *stream) {
    if (kind < USCT_BW_READ || kind > USCT_BW_COPY || bytes < 16 || bytes % 16 || reps < 1 ||
        blocks_per_cu < 1 || blocks_per_cu > 64 || !ms_out)
        return 1;
    if ((kind != USCT_BW_WRITE && !src) || (kind != USCT_BW_READ && !dst) ||
        (kind == USCT_BW_READ && !partials))
        return 1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 3;
    const int grid = cus * blocks_per_cu;
    const int64_t n = bytes / 16;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return 3;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return 3;
    }
    const v4f *s4 = (const v4f *)src;
    v4f *d4 = (v4f *)dst;
    launch(kind, nontemporal != 0, s4, d4, n, grid, partials, s);  // warm-up
    (void)hipEventRecord(a, s);
    for (int r = 0; r < reps; ++r) launch(kind, nontemporal != 0, s4, d4, n, grid, partials, s);
    (void)hipEventRecord(b, s);
    int rc = hipEventSynchronize(b) == hipSuccess && hipGetLastError() == hipSuccess ? 0 : 3;
    float ms = 0.0f;
    if (!rc && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = 3;
    *ms_out = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

int64_t usct_bw_pass_groups(int64_t bytes, int per_lane) {
    if (bytes < 16 || per_lane < 1) return -1;
    const int64_t n = bytes / 16;
    return (n + (int64_t)kBlock * per_lane - 1) / ((int64_t)kBlock * per_lane);
}

int usct_bw_run_pass(int kind, const void *src, void *dst, float *partials, int64_t bytes,
                     int per_lane, int nontemporal, int reps, float *ms_out, void *stream) {
    if (kind < USCT_BW_READ || kind > USCT_BW_COPY || bytes < 16 || bytes % 16 || reps < 1 ||
        !ms_out || (per_lane != 1 && per_lane != 2 && per_lane != 4 && per_lane != 8))
        return 1;
    if ((kind != USCT_BW_WRITE && !src) || (kind != USCT_BW_READ && !dst) ||
        (kind == USCT_BW_READ && !partials))
        return 1;
    // the grid must fit one launch dimension
    if (usct_bw_pass_groups(bytes, per_lane) > 0x7fffffffLL) return 1;
    const int64_t n = bytes / 16;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return 3;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return 3;
    }
    const v4f *s4 = (const v4f *)src;
    v4f *d4 = (v4f *)dst;
    auto go = [&] {
        if (nontemporal) launch_pass_nt<true>(kind, per_lane, s4, d4, n, partials, s);
        else launch_pass_nt<false>(kind, per_lane, s4, d4, n, partials, s);
    };
    go();  // warm-up
    (void)hipEventRecord(a, s);
    for (int r = 0; r < reps; ++r) go();
    (void)hipEventRecord(b, s);
    int rc = hipEventSynchronize(b) == hipSuccess && hipGetLastError() == hipSuccess ? 0 : 3;
    float ms = 0.0f;
    if (!rc && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = 3;
    *ms_out = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

int usct_bw_device_cus(void) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    return cus;
}

}  // extern "C"

// Practical peaks of the device, measured by streaming and FMA kernels of
// this file.  The rooflines of the update kernels are quoted against the
// MI355X datasheet (8 TB/s HBM3E, 78.6 TFLOP/s FP64 vector); these kernels
// give what a plain kernel reaches on the same device, so the fractions can
// be read against both.
//
//   HBM: read (one sum per block written), write (a fill) and copy of a
//        buffer much larger than the caches, 16-byte accesses per lane
//        (four loads in flight per lane in read and copy), grid-stride.
//   FMA: eight independent FMA chains per lane in FP64 or FP32.  The library
//        is built without packed FP32 instructions, so the FP32 rate is that
//        of one FP32 FMA per lane and cycle, the rate its mixed-precision
//        kernels can reach.
//
// Not part of the solver's C ABI (include/tenjin.h): a measurement tool,
// loaded by bench.py.
#include <hip/hip_runtime.h>

namespace {

constexpr int BLOCK = 256;
constexpr int UNROLL = 4;

__global__ __launch_bounds__(BLOCK) void k_read(const double2* __restrict__ a, long long n, double* __restrict__ out) {
    __shared__ double part[BLOCK / 64];
    const long long stride = (long long)gridDim.x * BLOCK;
    double s = 0.0;
    long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    for (; i + (UNROLL - 1) * stride < n; i += UNROLL * stride) {
        double2 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) v[u] = a[i + u * stride];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) s += v[u].x + v[u].y;
    }
    for (; i < n; i += stride) s += a[i].x + a[i].y;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) t += part[w];
        out[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(BLOCK) void k_write(double2* __restrict__ a, long long n, double x) {
    const long long stride = (long long)gridDim.x * BLOCK;
    for (long long i = (long long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) a[i] = make_double2(x, x);
}

__global__ __launch_bounds__(BLOCK) void k_copy(const double2* __restrict__ a, double2* __restrict__ b, long long n) {
    const long long stride = (long long)gridDim.x * BLOCK;
    long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    for (; i + (UNROLL - 1) * stride < n; i += UNROLL * stride) {
        double2 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) v[u] = a[i + u * stride];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) b[i + u * stride] = v[u];
    }
    for (; i < n; i += stride) b[i] = a[i];
}

constexpr int CHAINS = 8;

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_fma(T* __restrict__ out, T x, T y, int iters) {
    T c[CHAINS];
#pragma unroll
    for (int q = 0; q < CHAINS; ++q) c[q] = T(threadIdx.x + q);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < CHAINS; ++q) c[q] = fma_t(c[q], x, y);
    }
    T s = T(0);
#pragma unroll
    for (int q = 0; q < CHAINS; ++q) s += c[q];
    out[(long long)blockIdx.x * BLOCK + threadIdx.x] = s;  // keeps the chains live
}

// blocks of a streaming launch: eight per CU of the 256 CUs
constexpr int STREAM_BLOCKS = 256 * 8;
// FMA launch: eight waves per SIMD on every CU (4 SIMDs, 256 CUs)
constexpr int FMA_BLOCKS = 256 * 4 * 8 / (BLOCK / 64);

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" {

// kind 0 read, 1 write, 2 copy (reads one buffer of `bytes`, writes another).
// Average over `reps` launches after one untimed launch.  *gbs: bytes moved
// per launch (bytes for read or write, 2 x bytes for copy) / average time.
// Returns 0, or the HIP error code (1 for invalid arguments).
int tjp_hbm(int kind, long long bytes, int reps, double* gbs, double* ms) {
    if (kind < 0 || kind > 2 || bytes < 16 * (long long)BLOCK || reps < 1 || !gbs || !ms) return 1;
    const long long n = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    double* part = nullptr;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess && kind == 2) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess) e = hipMalloc(&part, STREAM_BLOCKS * sizeof(double));
    Events ev;
    if (e == hipSuccess) e = hipEventCreate(&ev.a);
    if (e == hipSuccess) e = hipEventCreate(&ev.b);
    auto launch = [&]() {
        if (kind == 0) hipLaunchKernelGGL(k_read, dim3(STREAM_BLOCKS), dim3(BLOCK), 0, 0, a, n, part);
        else if (kind == 1) hipLaunchKernelGGL(k_write, dim3(STREAM_BLOCKS), dim3(BLOCK), 0, 0, a, n, 1.0);
        else hipLaunchKernelGGL(k_copy, dim3(STREAM_BLOCKS), dim3(BLOCK), 0, 0, a, b, n);
    };
    if (e == hipSuccess) {
        // defined contents for the read and copy kernels
        hipLaunchKernelGGL(k_write, dim3(STREAM_BLOCKS), dim3(BLOCK), 0, 0, a, n, 0.5);
        launch();
        e = hipEventRecord(ev.a, 0);
    }
    if (e == hipSuccess) {
        for (int r = 0; r < reps; ++r) launch();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev.b, 0);
    if (e == hipSuccess) e = hipEventSynchronize(ev.b);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ev.a, ev.b);
    if (e == hipSuccess) {
        *ms = (double)t / reps;
        *gbs = (double)(kind == 2 ? 2 : 1) * (double)(n * 16) / (*ms * 1e6);
    }
    if (part) (void)hipFree(part);
    if (b) (void)hipFree(b);
    if (a) (void)hipFree(a);
    return e == hipSuccess ? 0 : (int)e;
}

// FMA throughput: fp64 != 0 for FP64, else FP32 (one FMA per lane, no packed
// form).  *tflops counts an FMA as two operations.  Returns 0 or the HIP
// error code (1 for invalid arguments).
int tjp_fma(int fp64, int iters, int reps, double* tflops, double* ms) {
    if (iters < 1 || reps < 1 || !tflops || !ms) return 1;
    const long long threads = (long long)FMA_BLOCKS * BLOCK;
    void* out = nullptr;
    hipError_t e = hipMalloc(&out, threads * 8);
    Events ev;
    if (e == hipSuccess) e = hipEventCreate(&ev.a);
    if (e == hipSuccess) e = hipEventCreate(&ev.b);
    auto launch = [&]() {
        if (fp64)
            hipLaunchKernelGGL(k_fma<double>, dim3(FMA_BLOCKS), dim3(BLOCK), 0, 0, (double*)out, 0.999999, 1e-7, iters);
        else
            hipLaunchKernelGGL(k_fma<float>, dim3(FMA_BLOCKS), dim3(BLOCK), 0, 0, (float*)out, 0.999f, 1e-4f, iters);
    };
    if (e == hipSuccess) {
        launch();
        e = hipEventRecord(ev.a, 0);
    }
    if (e == hipSuccess) {
        for (int r = 0; r < reps; ++r) launch();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev.b, 0);
    if (e == hipSuccess) e = hipEventSynchronize(ev.b);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ev.a, ev.b);
    if (e == hipSuccess) {
        *ms = (double)t / reps;
        *tflops = 2.0 * CHAINS * (double)iters * (double)threads / (*ms * 1e9);
    }
    if (out) (void)hipFree(out);
    return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"

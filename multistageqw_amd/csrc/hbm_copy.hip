// Calibration program for HBM traffic counters and the streaming ceiling: a
// plain copy of a known number of bytes in a few launch shapes.
// rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this program gives the counter
// reading that corresponds to exactly `bytes` read and `bytes` written per
// dispatch, the scale for the term kernel's counter readings.  A last
// kernel reads four arrays and writes a fifth, the sweep's stream mix.
// Usage: hbm_copy [MiB] [reps]
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>

typedef float f4 __attribute__((ext_vector_type(4)));

// grid-stride, 16 workgroups per CU
__global__ void __launch_bounds__(256) copy_stride(const f4* __restrict__ src, f4* __restrict__ dst, size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

// U float4 per thread, one pass, coalesced per load
template <int U, bool NT>
__global__ void __launch_bounds__(256) copy_flat(const f4* __restrict__ src, f4* __restrict__ dst, size_t n4)
{
    const size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x;
    f4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < n4) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < n4) {
            if (NT)
                __builtin_nontemporal_store(v[u], dst + i);
            else
                dst[i] = v[u];
        }
    }
}

// The sweep kernel's stream mix without its partner loads: four streams read
// and one written per element (b_{k+1}, b_{k+2}, psi, E -> b_k), one float4
// per thread per stream, non-temporal like the sweep's streams.
__global__ void __launch_bounds__(256)
mix4r1w(const f4* __restrict__ a, const f4* __restrict__ b, const f4* __restrict__ c, const f4* __restrict__ d,
        f4* __restrict__ o, size_t n4)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n4) {
        const f4 v = a[i] + __builtin_nontemporal_load(b + i) + __builtin_nontemporal_load(c + i) +
                     __builtin_nontemporal_load(d + i);
        __builtin_nontemporal_store(v, o + i);
    }
}

static double timed(void (*launch)(const f4*, f4*, size_t, int), const f4* a, f4* b, size_t n4, int cus)
{
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, 0);
    launch(a, b, n4, cus);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return ms;
}

static void l_stride(const f4* a, f4* b, size_t n4, int cus)
{
    hipLaunchKernelGGL(copy_stride, dim3((unsigned)cus * 16), dim3(256), 0, 0, a, b, n4);
}
template <int U, bool NT>
static void l_flat(const f4* a, f4* b, size_t n4, int)
{
    const size_t per = 256 * (size_t)U;
    hipLaunchKernelGGL((copy_flat<U, NT>), dim3((unsigned)((n4 + per - 1) / per)), dim3(256), 0, 0, a, b, n4);
}

int main(int argc, char** argv)
{
    const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const size_t bytes = mib << 20, n4 = bytes / 16;
    f4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    hipMemset(a, 0, bytes);
    hipMemset(b, 0, bytes);
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
    struct V {
        const char* name;
        void (*fn)(const f4*, f4*, size_t, int);
    } vs[] = {{"stride16", l_stride}, {"flat1", l_flat<1, false>}, {"flat4", l_flat<4, false>},
              {"flat4_nt", l_flat<4, true>}, {"flat8_nt", l_flat<8, true>}};
    for (const V& v : vs) {
        double best = 1e30;
        for (int r = 0; r < reps; r++) {
            const double ms = timed(v.fn, a, b, n4, cus);
            if (ms < best) best = ms;
        }
        printf("copy %-9s %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", v.name, mib, best,
               2.0 * bytes / (best * 1e-3) / 1e9);
    }
    // 4-read / 1-write mix over five separate arrays of `bytes` each
    f4 *c = nullptr, *d = nullptr, *o = nullptr;
    if (hipMalloc(&c, bytes) == hipSuccess && hipMalloc(&d, bytes) == hipSuccess && hipMalloc(&o, bytes) == hipSuccess) {
        hipMemset(c, 0, bytes);
        hipMemset(d, 0, bytes);
        hipMemset(o, 0, bytes);
        double best = 1e30;
        for (int r = 0; r < reps; r++) {
            hipEvent_t e0, e1;
            hipEventCreate(&e0);
            hipEventCreate(&e1);
            hipEventRecord(e0, 0);
            hipLaunchKernelGGL(mix4r1w, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, 0, a, b, c, d, o, n4);
            hipEventRecord(e1, 0);
            hipEventSynchronize(e1);
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            hipEventDestroy(e0);
            hipEventDestroy(e1);
            if (ms < best) best = ms;
        }
        printf("mix4r1w   %zu MiB x5: %.3f ms, %.1f GB/s (4 read + 1 write)\n", mib, best,
               5.0 * bytes / (best * 1e-3) / 1e9);
    }
    hipFree(c);
    hipFree(d);
    hipFree(o);
    hipFree(a);
    hipFree(b);
    return 0;
}

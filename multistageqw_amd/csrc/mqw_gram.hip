// Gram matrices of the shared first-stage bases (the compressed second stage,
// see GramJob) and of the last-stage adjoint bases X: per instance, the sums
// over all amplitudes of V_k1 V_k2 for k1 <= k2 of equal parity, with
// V_0 = psi0 * ones, (X_0) a unit vector or a stored vector.  Products of FP32 values
// are exact in FP64 and the sums are FP64: per workgroup over a chunk of
// amplitudes, then over the chunks in a fixed order.
#include <hip/hip_runtime.h>

#include "mqw_internal.h"

using namespace mqw;

namespace {

constexpr int kSlab = 64; // amplitudes per LDS slab

// gram_pair_start / gram_pairs of mqw_internal.h, on the device
__device__ __forceinline__ int pair_start(int k2) { return k2 == 0 ? 0 : k2 + ((k2 - 1) * (k2 - 1)) / 4; }

// One workgroup: one job, one chunk of N / nchunk amplitudes.  Thread t keeps
// the FP64 sums of pairs t, t + 256, ... (at most PMAX / 256 of them).
template <int KC>
__global__ void __launch_bounds__(256)
gram_kernel(const GramJob* __restrict__ jobs, uint32_t njobs, int n, uint32_t nchunk, double* __restrict__ partial)
{
    constexpr int PMAX = KC + ((KC - 1) * (KC - 1)) / 4;
    constexpr int PT = (PMAX + 255) / 256;
    __shared__ float s[KC][kSlab + 1];
    __shared__ uint16_t pr[PMAX];
    const uint32_t job = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    if (job >= njobs) return;
    const GramJob J = jobs[job];
    const int kc = (int)J.kc;
    const int t = threadIdx.x;
    const int P = pair_start(kc);
    if (t < kc) { // the pairs (k1, t) of equal parity, k1 <= t
        const int b = pair_start(t);
        for (int k1 = t & 1, q = 0; k1 <= t; k1 += 2, q++) pr[b + q] = (uint16_t)(k1 | (t << 8));
    }
    const uint64_t N = 1ull << n;
    const uint64_t C = N / nchunk;
    const uint64_t base = (uint64_t)chunk * C;
    double acc[PT];
#pragma unroll
    for (int i = 0; i < PT; i++) acc[i] = 0.0;
    for (uint64_t a0 = 0; a0 < C; a0 += kSlab) {
        __syncthreads();
        for (int i = t; i < kc * kSlab; i += 256) {
            const int k = i / kSlab, x = i % kSlab;
            const uint64_t a = base + a0 + (uint64_t)x;
            s[k][x] = k > 0     ? J.V[(uint64_t)(k - 1) * N + a]
                      : J.V0    ? J.V0[a]
                      : J.unit1 ? (a + 1 == J.unit1 ? 1.0f : 0.0f)
                                : J.psi0;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int p = t + 256 * i;
            if (p < P) {
                const int k1 = pr[p] & 0xff, k2 = pr[p] >> 8;
                double a = acc[i];
                for (int x = 0; x < kSlab; x++) a = fma((double)s[k1][x], (double)s[k2][x], a);
                acc[i] = a;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < PT; i++) {
        const int p = t + 256 * i;
        if (p < P) partial[((uint64_t)job * nchunk + chunk) * PMAX + p] = acc[i];
    }
}

// One thread per pair (workgroups: jobs x pair blocks of 256), the chunks in
// order: consecutive threads read consecutive partial sums.
template <int KC>
__global__ void __launch_bounds__(256)
gram_reduce_kernel(const GramJob* __restrict__ jobs, uint32_t nchunk, const double* __restrict__ partial)
{
    constexpr int PMAX = KC + ((KC - 1) * (KC - 1)) / 4;
    const uint32_t job = blockIdx.x;
    const GramJob J = jobs[job];
    const int P = pair_start((int)J.kc);
    const int p = (int)(blockIdx.y * 256 + threadIdx.x);
    if (p >= P) return;
    double sum = 0.0;
    for (uint32_t c = 0; c < nchunk; c++) sum += partial[((uint64_t)job * nchunk + c) * PMAX + p];
    J.out[p] = sum;
}

// Wide bases: the same pair sums in register blocks.  Per parity class cls,
// class index i is basis index 2 i + cls, and an item is a block of 4 x 4
// class indices [4 bi, 4 bi + 4) x [4 bj, 4 bj + 4) with bi <= bj; a thread
// keeps the FP64 sums of up to PT items (16 each) and per amplitude of an LDS
// slab reads 8 values for 16 fused multiply-adds (against 2 per pair above).
// The partial sums have gram_kernel<KC>'s layout, so gram_reduce_kernel<KC>
// sums them; each pair's sum is in a fixed order, the same on every run.
template <int KC>
__global__ void __launch_bounds__(512)
gram_blocked_kernel(const GramJob* __restrict__ jobs, uint32_t njobs, int n, uint32_t nchunk,
                    double* __restrict__ partial)
{
    constexpr int PMAX = KC + ((KC - 1) * (KC - 1)) / 4;
    constexpr int NBM = (KC / 2 + 3) / 4;     // blocks per side of a class, at most
    constexpr int ITEMS = NBM * (NBM + 1);    // upper-triangle blocks of both classes, at most
    constexpr int PT = (ITEMS + 511) / 512;
    static_assert(NBM <= 32, "block indices fit 5 bits");
    __shared__ float s[KC][kSlab + 1];
    __shared__ uint16_t code[ITEMS];
    const uint32_t job = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    if (job >= njobs) return;
    const GramJob J = jobs[job];
    const int kc = (int)J.kc;
    const int t = threadIdx.x;
    int nc[2], nb[2];
    for (int c = 0; c < 2; c++) {
        nc[c] = (kc - c + 1) / 2;
        nb[c] = (nc[c] + 3) / 4;
    }
    const int T0 = nb[0] * (nb[0] + 1) / 2, nitems = T0 + nb[1] * (nb[1] + 1) / 2;
    if (t == 0) { // item codes: class << 10 | bi << 5 | bj, class 0 first, bi <= bj
        int q = 0;
        for (int c = 0; c < 2; c++)
            for (int bi = 0; bi < nb[c]; bi++)
                for (int bj = bi; bj < nb[c]; bj++) code[q++] = (uint16_t)(c << 10 | bi << 5 | bj);
    }
    // rows past the basis stay zero (blocks at the class edge read them)
    for (int i = t; i < (KC - kc) * kSlab; i += 512) s[kc + i / kSlab][i % kSlab] = 0.0f;
    const uint64_t N = 1ull << n;
    const uint64_t C = N / nchunk;
    const uint64_t base = (uint64_t)chunk * C;
    double acc[PT][16];
#pragma unroll
    for (int p = 0; p < PT; p++)
#pragma unroll
        for (int e = 0; e < 16; e++) acc[p][e] = 0.0;
    for (uint64_t a0 = 0; a0 < C; a0 += kSlab) {
        __syncthreads();
        for (int i = t; i < kc * kSlab; i += 512) {
            const int k = i / kSlab, x = i % kSlab;
            const uint64_t a = base + a0 + (uint64_t)x;
            s[k][x] = k > 0     ? J.V[(uint64_t)(k - 1) * N + a]
                      : J.V0    ? J.V0[a]
                      : J.unit1 ? (a + 1 == J.unit1 ? 1.0f : 0.0f)
                                : J.psi0;
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PT; p++) {
            const int it = t + 512 * p;
            if (it < nitems) {
                const int cd = code[it], c = cd >> 10, bi = (cd >> 5) & 31, bj = cd & 31;
                int ri[4], rj[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    ri[q] = min(2 * (4 * bi + q) + c, KC - 1);
                    rj[q] = min(2 * (4 * bj + q) + c, KC - 1);
                }
                for (int x = 0; x < kSlab; x++) {
                    double u[4], v[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        u[q] = (double)s[ri[q]][x];
                        v[q] = (double)s[rj[q]][x];
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++)
#pragma unroll
                        for (int r = 0; r < 4; r++) acc[p][4 * q + r] = fma(u[q], v[r], acc[p][4 * q + r]);
                }
            }
        }
    }
#pragma unroll
    for (int p = 0; p < PT; p++) {
        const int it = t + 512 * p;
        if (it >= nitems) continue;
        const int cd = code[it], c = cd >> 10, bi = (cd >> 5) & 31, bj = cd & 31;
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 4 * bi + q, j = 4 * bj + r; // class indices, pair (i, j) with i <= j
                if (i <= j && j < nc[c])
                    partial[((uint64_t)job * nchunk + chunk) * PMAX + pair_start(2 * j + c) + i] = acc[p][4 * q + r];
            }
    }
}

template <int KC>
int launch_gram_kc(const GramJob* jobs, uint32_t njobs, int n, uint32_t nchunk, double* partial, hipStream_t st)
{
    constexpr int PMAX = KC + ((KC - 1) * (KC - 1)) / 4;
    const dim3 rgrid(njobs, (PMAX + 255) / 256);
    if constexpr (KC > 64) {
        hipLaunchKernelGGL(gram_blocked_kernel<KC>, dim3(njobs * nchunk), dim3(512), 0, st, jobs, njobs, n, nchunk,
                           partial);
        if (check_launch("gram")) return -1;
        hipLaunchKernelGGL(gram_reduce_kernel<KC>, rgrid, dim3(256), 0, st, jobs, nchunk, partial);
        return check_launch("gram_reduce");
    }
    hipLaunchKernelGGL(gram_kernel<KC>, dim3(njobs * nchunk), dim3(256), 0, st, jobs, njobs, n, nchunk, partial);
    if (check_launch("gram")) return -1;
    hipLaunchKernelGGL(gram_reduce_kernel<KC>, rgrid, dim3(256), 0, st, jobs, nchunk, partial);
    return check_launch("gram_reduce");
}

} // namespace

namespace mqw {

int launch_gram(const GramJob* jobs_dev, uint32_t njobs, uint32_t kc_max, int n, uint32_t nchunk, double* partial,
                void* stream)
{
    if (njobs == 0) return 0;
    const uint64_t N = 1ull << n;
    if (nchunk == 0 || N % nchunk != 0 || (N / nchunk) % kSlab != 0) return -2;
    if ((uint64_t)njobs * nchunk > 0x7fffffffull || njobs > 0x7fffffffu) return -3;
    hipStream_t st = (hipStream_t)stream;
    if (kc_max <= 32) return launch_gram_kc<32>(jobs_dev, njobs, n, nchunk, partial, st);
    if (kc_max <= 64) return launch_gram_kc<64>(jobs_dev, njobs, n, nchunk, partial, st);
    if (kc_max <= 128) return launch_gram_kc<128>(jobs_dev, njobs, n, nchunk, partial, st);
    if (kc_max <= (uint32_t)kGramMaxK) return launch_gram_kc<kGramMaxK>(jobs_dev, njobs, n, nchunk, partial, st);
    return -2;
}

} // namespace mqw

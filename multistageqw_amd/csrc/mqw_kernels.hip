// gfx950 kernels of the multi-stage quantum-walk path.
//
//   energy_*    : FP64 Gray-code energy build + ground-state search
//                 (MultiQW.cpp:325-362), one thread per 16 Gray steps,
//                 exact per-step deltas recomputed in two passes around a
//                 block-sum scan instead of the reference's serial walk.
//   term_kernel : one fused complex Clenshaw term (MultiQW.cpp:88-173,177-236)
//                 for a list of trajectories: the state tile of 2^TB amplitudes
//                 is flipped along its low TB bits through LDS / registers and
//                 along the remaining bits by coalesced partner-tile loads, in
//                 one read of b_{k+1}, b_{k+2}, psi, E and one write of b_k.
//   reduce_kernel: FP64 norm of a finished stage, 1/||psi|| for the next stage
//                 and the ground-state probability readout (MultiQW.cpp:443-446).
//   lds_kernel  : small n: a whole trajectory (all stages, all terms) inside one
//                 workgroup with the state resident in LDS and registers.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "mqw_internal.h"

namespace mqw {

static thread_local const char* g_kernel_error = "";
const char* last_kernel_error() { return g_kernel_error; }

int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_kernel_error = hipGetErrorString(e);
        (void)what;
        return -1;
    }
    return 0;
}

// ===========================================================================
// Energies
// ===========================================================================
constexpr int kEnergyThreads = 256;
constexpr int kEnergyPer = 16;
constexpr uint64_t kEnergyChunk = (uint64_t)kEnergyThreads * kEnergyPer;

size_t energy_blocks(int n)
{
    const uint64_t N = 1ull << n;
    return (size_t)((N + kEnergyChunk - 1) / kEnergyChunk);
}

// Change of energy at Gray step i (i >= 1): spin f = ctz(i) flips into the
// state g = grey(i).  The reference (MultiQW.cpp:340-346) forms
//   delta = float(4 * S - 2 h_f st_ff),  S = float sum_j J_fj st_fj in j order,
// with st_fj = s_f s_j and st_ff = -s_f.  Since negation is exact and rounding
// symmetric, delta = s_f * float(4 T + 2 h_f) with T = float sum_{j != f} J_fj s_j.
template <bool EXACT>
__device__ __forceinline__ double gray_delta(int n, const float* __restrict__ sJ, const float* __restrict__ sh,
                                            uint64_t i)
{
    if (i == 0) return 0.0;
    const int f = __builtin_ctzll(i);
    const uint64_t g = i ^ (i >> 1);
    const float* row = sJ + f * n;
    if (EXACT) {
        double T = 0.0;
        for (int j = 0; j < n; j++) {
            if (j == f) continue;
            const double v = (double)row[j];
            T += ((g >> j) & 1ull) ? v : -v;
        }
        const double d = 4.0 * T + 2.0 * (double)sh[f];
        return ((g >> f) & 1ull) ? d : -d;
    } else {
        float T = 0.0f;
        for (int j = 0; j < n; j++) {
            if (j == f) continue;
            const float v = row[j];
            T = T + (((g >> j) & 1ull) ? v : -v);
        }
        const float d = 4.0f * T + 2.0f * sh[f];
        return (double)(((g >> f) & 1ull) ? d : -d);
    }
}

__device__ __forceinline__ void load_couplings(int n, const float* J, const float* h, float* sJ, float* sh)
{
    for (int k = threadIdx.x; k < n * n; k += blockDim.x) sJ[k] = J[k];
    for (int k = threadIdx.x; k < n; k += blockDim.x) sh[k] = h[k];
    __syncthreads();
}

// Fixed-order block sum of one double per thread (kEnergyThreads threads).
__device__ __forceinline__ double block_sum_fixed(double v, double* red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = kEnergyThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Problem blockIdx.y of a batch: its couplings at J + y sJ (fields h at + n^2),
// its block tables at + y sB (EnergyBatch below; one problem: gridDim.y = 1).
template <bool EXACT>
__global__ void __launch_bounds__(kEnergyThreads)
energy_blocksum_kernel(int n, const float* __restrict__ J, const float* __restrict__ h, double* __restrict__ block_sum,
                       uint32_t sJ_stride, uint32_t sB_stride)
{
    __shared__ float sJ[kMaxN * kMaxN];
    __shared__ float sh[kMaxN];
    __shared__ double red[kEnergyThreads];
    J += (size_t)blockIdx.y * sJ_stride;
    h += (size_t)blockIdx.y * sJ_stride;
    block_sum += (size_t)blockIdx.y * sB_stride;
    load_couplings(n, J, h, sJ, sh);
    const uint64_t N = 1ull << n;
    const uint64_t i0 = ((uint64_t)blockIdx.x * kEnergyThreads + threadIdx.x) * kEnergyPer;
    double s = 0.0;
    for (int k = 0; k < kEnergyPer; k++) {
        const uint64_t i = i0 + k;
        if (i < N) s += gray_delta<EXACT>(n, sJ, sh, i);
    }
    const double tot = block_sum_fixed(s, red);
    if (threadIdx.x == 0) block_sum[blockIdx.x] = tot;
}

// Exclusive scan of the block sums in place, one workgroup, fixed order.
__global__ void __launch_bounds__(1024) energy_scan_kernel(double* __restrict__ v, uint32_t nb, uint32_t sB_stride)
{
    __shared__ double part[1024];
    v += (size_t)blockIdx.y * sB_stride;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t lo = t * per, hi = min(nb, lo + per);
    double s = 0.0;
    for (uint32_t k = lo; k < hi; k++) s += v[k];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        double run = 0.0;
        for (int k = 0; k < 1024; k++) {
            const double x = part[k];
            part[k] = run;
            run += x;
        }
    }
    __syncthreads();
    double run = part[t];
    for (uint32_t k = lo; k < hi; k++) {
        const double x = v[k];
        v[k] = run;
        run += x;
    }
}

template <bool EXACT>
__global__ void __launch_bounds__(kEnergyThreads)
energy_write_kernel(int n, const float* __restrict__ J, const float* __restrict__ h, double e_init,
                    const double* __restrict__ e_inits, const double* __restrict__ block_prefix,
                    float* __restrict__ hp, double* __restrict__ e64, double* __restrict__ block_min,
                    uint32_t* __restrict__ block_arg, double* __restrict__ block_max, uint32_t sJ_stride,
                    uint32_t sB_stride)
{
    __shared__ float sJ[kMaxN * kMaxN];
    __shared__ float sh[kMaxN];
    __shared__ double scan[kEnergyThreads];
    __shared__ double mn[kEnergyThreads];
    __shared__ double mx[kEnergyThreads];
    __shared__ uint32_t ag[kEnergyThreads];
    const uint64_t N = 1ull << n;
    {
        const size_t y = blockIdx.y;
        J += y * sJ_stride;
        h += y * sJ_stride;
        if (e_inits) e_init = e_inits[y];
        block_prefix += y * sB_stride;
        block_min += y * sB_stride;
        block_arg += y * sB_stride;
        block_max += y * sB_stride;
        if (gridDim.y > 1) {
            hp += y * N;
            if (e64) e64 += y * N;
        }
    }
    load_couplings(n, J, h, sJ, sh);
    const int t = threadIdx.x;
    const uint64_t i0 = ((uint64_t)blockIdx.x * kEnergyThreads + t) * kEnergyPer;
    double d[kEnergyPer];
    double s = 0.0;
    for (int k = 0; k < kEnergyPer; k++) {
        const uint64_t i = i0 + k;
        d[k] = (i < N) ? gray_delta<EXACT>(n, sJ, sh, i) : 0.0;
        s += d[k];
    }
    // exclusive scan of the thread sums (Hillis-Steele on a fixed layout)
    scan[t] = s;
    __syncthreads();
    for (int w = 1; w < kEnergyThreads; w <<= 1) {
        const double add = (t >= w) ? scan[t - w] : 0.0;
        __syncthreads();
        scan[t] += add;
        __syncthreads();
    }
    // Every partial sum here is a walk energy: float deltas of magnitude
    // >= 2^-22 added to values below 2^8 are exact in FP64, so this tree order
    // reproduces the reference's serial accumulation (MultiQW.cpp:345).
    double run = (e_init + block_prefix[blockIdx.x]) + (t > 0 ? scan[t - 1] : 0.0);
    double lo = __builtin_inf(), hi = -__builtin_inf();
    uint32_t arg = 0xffffffffu;
    for (int k = 0; k < kEnergyPer; k++) {
        const uint64_t i = i0 + k;
        if (i >= N) break;
        run += d[k];
        const uint64_t g = i ^ (i >> 1);
        hp[g] = (float)run;
        if (e64) e64[g] = run;
        if (run < lo) {
            lo = run;
            arg = (uint32_t)i;
        }
        if (run > hi) hi = run;
    }
    mn[t] = lo;
    mx[t] = hi;
    ag[t] = arg;
    __syncthreads();
    for (int w = kEnergyThreads / 2; w > 0; w >>= 1) {
        if (t < w) {
            // strict minimum, earliest Gray step on ties (MultiQW.cpp:349)
            const double a = mn[t], b = mn[t + w];
            if (b < a || (b == a && ag[t + w] < ag[t])) {
                mn[t] = b;
                ag[t] = ag[t + w];
            }
            if (mx[t + w] > mx[t]) mx[t] = mx[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        block_min[blockIdx.x] = mn[0];
        block_arg[blockIdx.x] = ag[0];
        block_max[blockIdx.x] = mx[0];
    }
}

__global__ void __launch_bounds__(256)
energy_finalize_kernel(uint32_t nb, const double* __restrict__ block_min, const uint32_t* __restrict__ block_arg,
                       const double* __restrict__ block_max, EnergyDev* __restrict__ info, uint32_t sB_stride)
{
    __shared__ double mn[256], mx[256];
    __shared__ uint32_t ag[256];
    const int t = threadIdx.x;
    block_min += (size_t)blockIdx.y * sB_stride;
    block_arg += (size_t)blockIdx.y * sB_stride;
    block_max += (size_t)blockIdx.y * sB_stride;
    info += blockIdx.y;
    double lo = __builtin_inf(), hi = -__builtin_inf();
    uint32_t arg = 0xffffffffu;
    for (uint32_t k = t; k < nb; k += 256) {
        const double a = block_min[k];
        if (a < lo || (a == lo && block_arg[k] < arg)) {
            lo = a;
            arg = block_arg[k];
        }
        if (block_max[k] > hi) hi = block_max[k];
    }
    mn[t] = lo;
    mx[t] = hi;
    ag[t] = arg;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            const double a = mn[t], b = mn[t + w];
            if (b < a || (b == a && ag[t + w] < ag[t])) {
                mn[t] = b;
                ag[t] = ag[t + w];
            }
            if (mx[t + w] > mx[t]) mx[t] = mx[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double e0 = mn[0], emax = mx[0];
        const uint32_t i = ag[0];
        info->e0 = e0;
        info->emax = emax;
        info->eloc = i ^ (i >> 1);
        info->shift = (float)((emax + e0) / 2);
        info->e_abs = (float)((emax - e0) / 2);
        info->center = (emax + e0) / 2;
    }
}

__global__ void __launch_bounds__(256)
energy_shift_kernel(uint64_t N, float* __restrict__ hp, double* __restrict__ e64, int e64_shift,
                    const EnergyDev* __restrict__ info)
{
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= N) return;
    if (gridDim.y > 1) {
        hp += (size_t)blockIdx.y * N;
        if (e64) e64 += (size_t)blockIdx.y * N;
        info += blockIdx.y;
    }
    const float s = info->shift;
    hp[k] = hp[k] - s;
    if (e64 && e64_shift) e64[k] = e64[k] - info->center;
}

int launch_energy(int n, const float* J_dev, const float* h_dev, double e_init, int exact, float* hp, double* e64,
                  int e64_shift, const EnergyScratch& scr, EnergyDev* info_dev, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nb = (uint32_t)energy_blocks(n);
    const uint64_t N = 1ull << n;
    if (exact)
        hipLaunchKernelGGL(energy_blocksum_kernel<true>, dim3(nb), dim3(kEnergyThreads), 0, st, n, J_dev, h_dev,
                           scr.block_sum, 0u, 0u);
    else
        hipLaunchKernelGGL(energy_blocksum_kernel<false>, dim3(nb), dim3(kEnergyThreads), 0, st, n, J_dev, h_dev,
                           scr.block_sum, 0u, 0u);
    hipLaunchKernelGGL(energy_scan_kernel, dim3(1), dim3(1024), 0, st, scr.block_sum, nb, 0u);
    if (exact)
        hipLaunchKernelGGL(energy_write_kernel<true>, dim3(nb), dim3(kEnergyThreads), 0, st, n, J_dev, h_dev, e_init,
                           nullptr, scr.block_sum, hp, e64, scr.block_min, scr.block_arg, scr.block_max, 0u, 0u);
    else
        hipLaunchKernelGGL(energy_write_kernel<false>, dim3(nb), dim3(kEnergyThreads), 0, st, n, J_dev, h_dev,
                           e_init, nullptr, scr.block_sum, hp, e64, scr.block_min, scr.block_arg, scr.block_max, 0u,
                           0u);
    hipLaunchKernelGGL(energy_finalize_kernel, dim3(1), dim3(256), 0, st, nb, scr.block_min, scr.block_arg,
                       scr.block_max, info_dev, 0u);
    hipLaunchKernelGGL(energy_shift_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, hp, e64,
                       e64_shift, info_dev);
    return check_launch("energy");
}

int launch_energies(int n, uint32_t cnt, const float* Jh_dev, const double* e_inits_dev, int exact, float* hp,
                    double* e64, int e64_shift, const EnergyScratch& scr, EnergyDev* info_dev, void* stream)
{
    if (cnt == 0) return 0;
    if (cnt > 65535u || n < 1 || n > kMaxN) return -2;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nb = (uint32_t)energy_blocks(n);
    const uint64_t N = 1ull << n;
    const uint32_t sj = (uint32_t)(n * n + n); // couplings then fields, per problem
    const float* J = Jh_dev;
    const float* h = Jh_dev + n * n;
    if (exact)
        hipLaunchKernelGGL(energy_blocksum_kernel<true>, dim3(nb, cnt), dim3(kEnergyThreads), 0, st, n, J, h,
                           scr.block_sum, sj, nb);
    else
        hipLaunchKernelGGL(energy_blocksum_kernel<false>, dim3(nb, cnt), dim3(kEnergyThreads), 0, st, n, J, h,
                           scr.block_sum, sj, nb);
    hipLaunchKernelGGL(energy_scan_kernel, dim3(1, cnt), dim3(1024), 0, st, scr.block_sum, nb, nb);
    if (exact)
        hipLaunchKernelGGL(energy_write_kernel<true>, dim3(nb, cnt), dim3(kEnergyThreads), 0, st, n, J, h, 0.0,
                           e_inits_dev, scr.block_sum, hp, e64, scr.block_min, scr.block_arg, scr.block_max, sj, nb);
    else
        hipLaunchKernelGGL(energy_write_kernel<false>, dim3(nb, cnt), dim3(kEnergyThreads), 0, st, n, J, h, 0.0,
                           e_inits_dev, scr.block_sum, hp, e64, scr.block_min, scr.block_arg, scr.block_max, sj, nb);
    hipLaunchKernelGGL(energy_finalize_kernel, dim3(1, cnt), dim3(256), 0, st, nb, scr.block_min, scr.block_arg,
                       scr.block_max, info_dev, nb);
    hipLaunchKernelGGL(energy_shift_kernel, dim3((unsigned)((N + 255) / 256), cnt), dim3(256), 0, st, N, hp, e64,
                       e64_shift, info_dev);
    return check_launch("energies");
}

// ===========================================================================
// Fused Clenshaw term
// ===========================================================================
// Pointers fetched from a job descriptor are generic; casting them to the
// global address space makes the compiler emit global_load/store (counted by
// vmcnt only) instead of flat accesses that would also stall LDS waits.
#define MQW_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const MQW_GLOBAL T* gview(const T* p)
{
    return (const MQW_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ MQW_GLOBAL T* gview(T* p)
{
    return (MQW_GLOBAL T*)p;
}

template <typename real>
struct NVec16;
template <>
struct NVec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int W = 4;
};
template <>
struct NVec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int W = 2;
};

template <typename real, int R>
__device__ __forceinline__ void vload(const real* __restrict__ p, real (&v)[R], int stride = NVec16<real>::W)
{
    using vt = typename NVec16<real>::type;
    constexpr int W = NVec16<real>::W;
#pragma unroll
    for (int q = 0; q < R / W; q++) {
        const vt x = *gview(reinterpret_cast<const vt*>(p + q * stride));
#pragma unroll
        for (int w = 0; w < W; w++) v[q * W + w] = x[w];
    }
}

template <typename real, int R>
__device__ __forceinline__ void vstore(real* p, const real (&v)[R], int stride = NVec16<real>::W)
{
    using vt = typename NVec16<real>::type;
    constexpr int W = NVec16<real>::W;
#pragma unroll
    for (int q = 0; q < R / W; q++) {
        vt x;
#pragma unroll
        for (int w = 0; w < W; w++) x[w] = v[q * W + w];
        *gview(reinterpret_cast<vt*>(p + q * stride)) = x;
    }
}

// Streaming accesses: data read once per sweep (b_{k+2}, psi, E) carry the
// non-temporal hint so L2 keeps the b_{k+1} tiles that neighbouring
// workgroups re-read as partner tiles.
template <typename real, int R>
__device__ __forceinline__ void vload_nt(const real* __restrict__ p, real (&v)[R], int stride = NVec16<real>::W)
{
    using vt = typename NVec16<real>::type;
    constexpr int W = NVec16<real>::W;
#pragma unroll
    for (int q = 0; q < R / W; q++) {
        const vt x = __builtin_nontemporal_load(gview(reinterpret_cast<const vt*>(p + q * stride)));
#pragma unroll
        for (int w = 0; w < W; w++) v[q * W + w] = x[w];
    }
}

template <typename real, int R>
__device__ __forceinline__ void vstore_nt(real* p, const real (&v)[R], int stride = NVec16<real>::W)
{
    using vt = typename NVec16<real>::type;
    constexpr int W = NVec16<real>::W;
#pragma unroll
    for (int q = 0; q < R / W; q++) {
        vt x;
#pragma unroll
        for (int w = 0; w < W; w++) x[w] = v[q * W + w];
        __builtin_nontemporal_store(x, gview(reinterpret_cast<vt*>(p + q * stride)));
    }
}

// LDS accesses of the tile exchange buffer (generic pointers into __shared__).
template <typename real, int R>
__device__ __forceinline__ void lds_read(const real* p, real (&v)[R], int stride = NVec16<real>::W)
{
    using vt = typename NVec16<real>::type;
    constexpr int W = NVec16<real>::W;
#pragma unroll
    for (int q = 0; q < R / W; q++) {
        const vt x = *reinterpret_cast<const vt*>(p + q * stride);
#pragma unroll
        for (int w = 0; w < W; w++) v[q * W + w] = x[w];
    }
}

template <typename real, int R>
__device__ __forceinline__ void lds_write(real* p, const real (&v)[R], int stride = NVec16<real>::W)
{
    using vt = typename NVec16<real>::type;
    constexpr int W = NVec16<real>::W;
#pragma unroll
    for (int q = 0; q < R / W; q++) {
        vt x;
#pragma unroll
        for (int w = 0; w < W; w++) x[w] = v[q * W + w];
        *reinterpret_cast<vt*>(p + q * stride) = x;
    }
}

// Deterministic workgroup sum of one double per thread, result valid in thread 0.
// Full waves reduce by cross-lane shuffles; a workgroup narrower than one
// wave goes through LDS so no inactive lane is ever read.
template <int NT>
__device__ __forceinline__ double wg_sum_t0(double v, double* red)
{
    const int t = threadIdx.x;
    if constexpr (NT >= 64) {
#pragma unroll
        for (int w = 32; w > 0; w >>= 1) v += __shfl_xor(v, w, 64);
        if ((t & 63) == 0) red[t >> 6] = v;
        __syncthreads();
        double s = 0.0;
        if (t == 0)
            for (int w = 0; w < NT / 64; w++) s += red[w];
        return s;
    } else {
        red[t] = v;
        __syncthreads();
        double s = 0.0;
        if (t == 0)
            for (int w = 0; w < NT; w++) s += red[w];
        return s;
    }
}

// blockIdx -> (job, tile).  Workgroups are dispatched round-robin over the 8
// XCDs; consecutive runs of 2^lc tiles (a sub-cube of the outer bits) are kept
// on one XCD so their partner-tile loads meet in that XCD's L2.
__device__ __forceinline__ uint64_t xcd_unit(uint32_t bid, int lc)
{
    const uint32_t xcd = bid & 7u, slot = bid >> 3;
    const uint64_t chunk = ((uint64_t)(slot >> lc) << 3) | xcd;
    return (chunk << lc) | (uint64_t)(slot & ((1u << lc) - 1u));
}

template <typename real, int TB, int RB, bool NTS>
__global__ void __launch_bounds__(1 << (TB - RB), 8)
term_kernel(const TermJob<real>* __restrict__ jobs, uint32_t njobs, int n, int lc)
{
    constexpr int NT = 1 << (TB - RB);
    constexpr int R = 1 << RB;
    constexpr int TS = 1 << TB;
    constexpr int NRED = NT >= 64 ? NT / 64 : NT;
    __shared__ real s_re[TS];
    __shared__ real s_im[TS];
    __shared__ double s_red[NRED];
    __shared__ double s_bcast;

    const int ob = n - TB;
    const uint64_t unit = xcd_unit(blockIdx.x, lc);
    const uint64_t job = unit >> ob;
    if (job >= njobs) return;
    const uint64_t tile = unit & ((1ull << ob) - 1ull);
    const TermJob<real>& J = jobs[job];
    const uint32_t flags = J.flags;
    const uint32_t src1 = flags & 3u, src2 = flags & 12u;
    const bool psi_const = (flags & F_PSI_CONST) != 0;
    const uint64_t N = 1ull << n;
    const int t = threadIdx.x;
    // plane layout: chunk c of thread t holds elements c*(NT*W) + t*W + [0, W)
    constexpr int W = NVec16<real>::W;
    constexpr int PS = NT * W;
    const uint64_t off = (tile << TB) + (uint64_t)t * W;
    real ps = real(1);
    if (flags & F_NORM_IN) {
        // 1/||Psi|| from the previous stage's partial sums, same order in every workgroup
        const uint32_t nt = 1u << ob;
        double s = 0.0;
        for (uint32_t k = t; k < nt; k += NT) s += J.norm_in[k];
        const double tot0 = wg_sum_t0<NT>(s, s_red);
        if (t == 0) s_bcast = tot0;
        __syncthreads();
        const double tot = s_bcast;
        ps = (real)(1.0 / sqrt(tot));
        if ((flags & F_READOUT) && tile == 0 && t == 0) {
            const double re = (double)J.psi[J.eloc], im = (double)J.psi[N + J.eloc];
            *J.prob = (re * re + im * im) / tot;
        }
    }

    real pr[R], pi[R], e[R];
    if (!psi_const) {
        vload_nt<real, R>(J.psi + off, pr, PS);
        vload_nt<real, R>(J.psi + N + off, pi, PS);
    } else {
#pragma unroll
        for (int q = 0; q < R; q++) {
            pr[q] = J.psi0;
            pi[q] = real(0);
        }
    }
    vload_nt<real, R>(J.E + off, e, PS);
    real yr[R], yi[R];
    if (src2 == SRC2_ARRAY) {
        vload_nt<real, R>(J.b2 + off, yr, PS);
        vload_nt<real, R>(J.b2 + N + off, yi, PS);
    }

    real ur[R], ui[R];
    if (src1 == SRC1_ZERO) {
#pragma unroll
        for (int q = 0; q < R; q++) ur[q] = ui[q] = real(0);
    } else if (src1 == SRC1_PSI && psi_const) {
        // |+> is an eigenvector of H_G with eigenvalue n
#pragma unroll
        for (int q = 0; q < R; q++) {
            ur[q] = J.alpha * e[q] * J.psi0 - J.beta * (real)n * J.psi0;
            ui[q] = real(0);
        }
    } else {
        const real* src = (src1 == SRC1_PSI) ? J.psi : J.b1;
        real xr[R], xi[R];
        if (src1 == SRC1_PSI) {
#pragma unroll
            for (int q = 0; q < R; q++) {
                xr[q] = pr[q];
                xi[q] = pi[q];
            }
        } else {
            vload<real, R>(src + off, xr, PS);
            vload<real, R>(src + N + off, xi, PS);
        }
        lds_write<real, R>(s_re + t * W, xr, PS);
        lds_write<real, R>(s_im + t * W, xi, PS);
        real ar[R], ai[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            ar[q] = real(0);
            ai[q] = real(0);
        }
        // flips of the register bits
#pragma unroll
        for (int b = 0; b < RB; b++)
#pragma unroll
            for (int q = 0; q < R; q++) {
                ar[q] += xr[q ^ (1 << b)];
                ai[q] += xi[q ^ (1 << b)];
            }
        // flips of the outer bits: partner tiles, coalesced loads
#pragma unroll 4
        for (int j = 0; j < ob; j++) {
            const uint64_t poff = ((tile ^ (1ull << j)) << TB) + (uint64_t)t * W;
            real yr[R], yi[R];
            vload<real, R>(src + poff, yr, PS);
            vload<real, R>(src + N + poff, yi, PS);
#pragma unroll
            for (int q = 0; q < R; q++) {
                ar[q] += yr[q];
                ai[q] += yi[q];
            }
        }
        __syncthreads();
        // flips of the thread bits through LDS
#pragma unroll
        for (int b = 0; b < TB - RB; b++) {
            const int pt = t ^ (1 << b);
            real yr[R], yi[R];
            lds_read<real, R>(s_re + pt * W, yr, PS);
            lds_read<real, R>(s_im + pt * W, yi, PS);
#pragma unroll
            for (int q = 0; q < R; q++) {
                ar[q] += yr[q];
                ai[q] += yi[q];
            }
        }
#pragma unroll
        for (int q = 0; q < R; q++) {
            ur[q] = J.alpha * e[q] * xr[q] - J.beta * ar[q];
            ui[q] = J.alpha * e[q] * xi[q] - J.beta * ai[q];
        }
    }

    real w1r = J.w1r, w1i = J.w1i, w2r = J.w2r, w2i = J.w2i;
    if (src1 == SRC1_PSI) {
        w1r *= ps;
        w1i *= ps;
    }
    if (src2 == SRC2_ZERO) {
#pragma unroll
        for (int q = 0; q < R; q++) yr[q] = yi[q] = real(0);
    } else if (src2 == SRC2_PSI) {
        w2r *= ps;
        w2i *= ps;
#pragma unroll
        for (int q = 0; q < R; q++) {
            yr[q] = pr[q];
            yi[q] = pi[q];
        }
    }
    const real cr = J.cr * ps, ci = J.ci * ps;
    real orr[R], oi[R];
    double nrm = 0.0;
#pragma unroll
    for (int q = 0; q < R; q++) {
        orr[q] = (w1r * ur[q] - w1i * ui[q]) - (w2r * yr[q] - w2i * yi[q]) + (cr * pr[q] - ci * pi[q]);
        oi[q] = (w1r * ui[q] + w1i * ur[q]) - (w2r * yi[q] + w2i * yr[q]) + (cr * pi[q] + ci * pr[q]);
        nrm += (double)orr[q] * (double)orr[q] + (double)oi[q] * (double)oi[q];
    }
    if (NTS) {
        vstore_nt<real, R>(J.out + off, orr, PS);
        vstore_nt<real, R>(J.out + N + off, oi, PS);
    } else {
        vstore<real, R>(J.out + off, orr, PS);
        vstore<real, R>(J.out + N + off, oi, PS);
    }

    if (flags & F_FINAL) {
        const double s = wg_sum_t0<NT>(nrm, s_red);
        if (t == 0) J.norm_part[tile] = s;
    }
}

template <typename real, int TB>
static int launch_term_tb(const void* jobs, uint32_t njobs, int n, int lc, int variant, hipStream_t st)
{
    // 16-byte vectors per thread at least; at most 1024 threads per workgroup
    constexpr int RB0 = (sizeof(real) == 4) ? 2 : 1;
    constexpr int RB = (TB - RB0 > 10) ? TB - 10 : RB0;
    constexpr int NT = 1 << (TB - RB);
    static_assert(NT <= 1024, "workgroup too large");
    const int ob = n - TB;
    if (ob < 0) return -2;
    if (lc > ob + 16) lc = ob + 16;
    const uint64_t units = (uint64_t)njobs << ob;
    const uint64_t per = 8ull << lc;
    const uint64_t grid = (units + per - 1) / per * per;
    if (grid > 0x7fffffffull) return -3;
    if (variant == 1)
        hipLaunchKernelGGL((term_kernel<real, TB, RB, true>), dim3((uint32_t)grid), dim3(NT), 0, st,
                           (const TermJob<real>*)jobs, njobs, n, lc);
    else
        hipLaunchKernelGGL((term_kernel<real, TB, RB, false>), dim3((uint32_t)grid), dim3(NT), 0, st,
                           (const TermJob<real>*)jobs, njobs, n, lc);
    return check_launch("term");
}

int term_tile_bits_supported(int precision, int tb)
{
    (void)precision;
    return tb >= 4 && tb <= 12;
}

int launch_term(int precision, int tb, int variant, const void* jobs_dev, uint32_t njobs, int n, int lc,
                void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (precision == MQW_FP32) {
        switch (tb) {
        case 4: return launch_term_tb<float, 4>(jobs_dev, njobs, n, lc, variant, st);
        case 5: return launch_term_tb<float, 5>(jobs_dev, njobs, n, lc, variant, st);
        case 6: return launch_term_tb<float, 6>(jobs_dev, njobs, n, lc, variant, st);
        case 7: return launch_term_tb<float, 7>(jobs_dev, njobs, n, lc, variant, st);
        case 8: return launch_term_tb<float, 8>(jobs_dev, njobs, n, lc, variant, st);
        case 9: return launch_term_tb<float, 9>(jobs_dev, njobs, n, lc, variant, st);
        case 10: return launch_term_tb<float, 10>(jobs_dev, njobs, n, lc, variant, st);
        case 11: return launch_term_tb<float, 11>(jobs_dev, njobs, n, lc, variant, st);
        case 12: return launch_term_tb<float, 12>(jobs_dev, njobs, n, lc, variant, st);
        }
    } else {
        switch (tb) {
        case 4: return launch_term_tb<double, 4>(jobs_dev, njobs, n, lc, variant, st);
        case 5: return launch_term_tb<double, 5>(jobs_dev, njobs, n, lc, variant, st);
        case 6: return launch_term_tb<double, 6>(jobs_dev, njobs, n, lc, variant, st);
        case 7: return launch_term_tb<double, 7>(jobs_dev, njobs, n, lc, variant, st);
        case 8: return launch_term_tb<double, 8>(jobs_dev, njobs, n, lc, variant, st);
        case 9: return launch_term_tb<double, 9>(jobs_dev, njobs, n, lc, variant, st);
        case 10: return launch_term_tb<double, 10>(jobs_dev, njobs, n, lc, variant, st);
        case 11: return launch_term_tb<double, 11>(jobs_dev, njobs, n, lc, variant, st);
        case 12: return launch_term_tb<double, 12>(jobs_dev, njobs, n, lc, variant, st);
        }
    }
    return -1;
}

// ===========================================================================
// Real-half sweep (FP32).  The real and imaginary parts of b obey independent
// real recurrences, coupled only through the injected c_r i^(r&1) psi and the
// complex scalars of a stage's first two terms, each of which is purely real
// or purely imaginary.  So one workgroup produces one real half of a tile of
// 2^HB amplitudes from one real half of each source: with the same LDS as a
// complex tile of 2^(HB-1), two more tile bits are exchanged on chip and
// fewer flips need partner-tile loads.  The diagonal is read by both halves.
// ===========================================================================
template <int HB, int UNR, int OCC, bool NTS, int MODE>
__global__ void __launch_bounds__(1024, OCC)
half_term_kernel(const TermJob<float>* __restrict__ jobs, uint32_t njobs, int n, int lc, int hi)
{
    constexpr int NT = 1024;
    constexpr int RB = HB - 10;
    constexpr int R = 1 << RB;
    constexpr int TS = 1 << HB;
    // MODE 0: one workgroup per (half, tile).
    // MODE 1: per (half, tile pair): the two tiles that differ in the top tile
    //         bit, so that flip is an exchange of registers, not a load.
    // MODE 2: per tile: both halves, one read of the diagonal.
    // MODE 3: per tile, the real half only (real recurrences: the imaginary
    //         halves of the arrays are never touched and may be absent).
    constexpr int NS = (MODE == 0 || MODE == 3) ? 1 : 2;
    __shared__ float s_x[TS];
    __shared__ double s_red[NT / 64];
    __shared__ double s_bcast;

    const int ob = n - HB;
    const int pb = ob - (MODE == 1 ? 1 : 0); // tile bits that are partner loads
    const uint64_t unit = xcd_unit(blockIdx.x, lc);
    // unit = (job, half, tile) in modes 0 and 1 (mode 1 drops the top tile
    // bit), (job, tile) in modes 2 and 3
    const int ub = (MODE == 2 || MODE == 3) ? pb : pb + 1;
    const uint64_t job = unit >> ub;
    if (job >= njobs) return;
    const int h0 = (MODE == 2 || MODE == 3) ? 0 : (int)((unit >> pb) & 1u);
    const uint64_t tile0 = unit & ((1ull << pb) - 1ull);
    const TermJob<float>& J = jobs[job];
    const uint32_t flags = J.flags;
    const uint32_t src1 = flags & 3u, src2 = flags & 12u;
    const bool psi_const = (flags & F_PSI_CONST) != 0;
    const uint64_t N = 1ull << n;
    const int t = threadIdx.x;
    // plane layout: chunk c of thread t holds elements c*4096 + 4t + [0, 4)
    constexpr int PS = NT * 4;

    float ps = 1.0f;
    if (flags & F_NORM_IN) {
        const uint32_t nt = 2u << ob;
        double sacc = 0.0;
        for (uint32_t k = t; k < nt; k += NT) sacc += J.norm_in[k];
        const double tot0 = wg_sum_t0<NT>(sacc, s_red);
        if (t == 0) s_bcast = tot0;
        __syncthreads();
        const double tot = s_bcast;
        ps = (float)(1.0 / sqrt(tot));
        if ((flags & F_READOUT) && tile0 == 0 && h0 == 0 && t == 0) {
            const double re = (double)J.psi[J.eloc], im = (double)J.psi[N + J.eloc];
            *J.prob = (re * re + im * im) / tot;
        }
    }

    const bool src1_arr = src1 != SRC1_ZERO && !(src1 == SRC1_PSI && psi_const);
    const float* src_base = (src1 == SRC1_PSI) ? J.psi : J.b1;
    const bool swap1 = J.w1i != 0.0f; // the first scalar is imaginary: halves swap

    // own b_{k+1} tiles of every side, loaded up front
    float xs[NS][R];
    if (src1_arr) {
#pragma unroll
        for (int sd = 0; sd < NS; sd++) {
            const int hs = MODE == 2 ? sd : h0;
            const uint64_t ts = MODE == 1 ? (tile0 | ((uint64_t)sd << pb)) : tile0;
            vload<float, R>(src_base + (uint64_t)(swap1 ? 1 - hs : hs) * N + (ts << HB) + (uint64_t)t * 4, xs[sd], PS);
        }
    }
    // mode 2: the diagonal of the tile, shared by both halves
    float e2[R];
    if (MODE == 2) vload_nt<float, R>(J.E + (tile0 << HB) + (uint64_t)t * 4, e2, PS);

#pragma unroll
    for (int sd = 0; sd < NS; sd++) {
        const int h = MODE == 2 ? sd : h0;
        const uint64_t tile = MODE == 1 ? (tile0 | ((uint64_t)sd << pb)) : tile0;
        const uint64_t off = (tile << HB) + (uint64_t)t * 4;
        // out_h = a1 u_{h1} - a2 s2_{h2} + ac Psi_{hc}: each complex scalar is
        // real or imaginary, so it maps one source half onto this output half.
        float a1 = J.w1r, a2 = J.w2r, ac = J.cr;
        int h1 = h, h2 = h, hc = h;
        if (swap1) {
            a1 = h == 0 ? -J.w1i : J.w1i;
            h1 = 1 - h;
        }
        if (J.w2i != 0.0f) {
            a2 = h == 0 ? -J.w2i : J.w2i;
            h2 = 1 - h;
        }
        if (J.ci != 0.0f) {
            ac = h == 0 ? -J.ci : J.ci;
            hc = 1 - h;
        }
        if (src1 == SRC1_PSI) a1 *= ps;
        if (src2 == SRC2_PSI) a2 *= ps;
        ac *= ps;
        const float* src = src_base + (uint64_t)h1 * N;

        // the previous side's LDS reads and norm sum are complete
        if (sd > 0) __syncthreads();
        float o[R];
#pragma unroll
        for (int q = 0; q < R; q++) o[q] = 0.0f;
        if (src1 == SRC1_PSI && psi_const) {
            if (h1 == 0) {
                float e[R];
                if (MODE == 2) {
#pragma unroll
                    for (int q = 0; q < R; q++) e[q] = e2[q];
                } else {
                    vload_nt<float, R>(J.E + off, e, PS);
                }
                const float base = J.beta * (float)n * J.psi0;
#pragma unroll
                for (int q = 0; q < R; q++) o[q] = a1 * (J.alpha * e[q] * J.psi0 - base);
            }
        } else if (src1_arr) {
            const float(&x)[R] = xs[sd];
            float acc[R];
            lds_write<float, R>(s_x + t * 4, x, PS);
#pragma unroll
            for (int q = 0; q < R; q++) acc[q] = 0.0f;
#pragma unroll
            for (int b = 0; b < RB; b++)
#pragma unroll
                for (int q = 0; q < R; q++) acc[q] += x[q ^ (1 << b)];
            // partner tiles in bit order, UNR loads in flight per thread; with
            // F_HSUM the top hi bits come from the high-bit pass instead.  Mode 3
            // jobs flagged F_LATE (the basis recurrences where the on-chip ones
            // apply) add them after the tile's own flips, in their order
            // (mqw_chain.hip), so the two give bit-identical outputs.
            const int pl = (flags & F_HSUM) ? pb - hi : pb;
            auto partners = [&]() {
                int j = 0;
                for (; j + UNR <= pl; j += UNR) {
                    float y[UNR][R];
#pragma unroll
                    for (int u = 0; u < UNR; u++)
                        vload<float, R>(src + ((tile ^ (1ull << (j + u))) << HB) + (uint64_t)t * 4, y[u], PS);
#pragma unroll
                    for (int u = 0; u < UNR; u++)
#pragma unroll
                        for (int q = 0; q < R; q++) acc[q] += y[u][q];
                }
                for (; j < pl; j++) {
                    float y[R];
                    vload<float, R>(src + ((tile ^ (1ull << j)) << HB) + (uint64_t)t * 4, y, PS);
#pragma unroll
                    for (int q = 0; q < R; q++) acc[q] += y[q];
                }
            };
            const bool late = MODE == 3 && (flags & F_LATE) != 0;
            if (!late) partners();
            if (flags & F_HSUM) {
                float y[R];
                vload_nt<float, R>(J.hsum + (uint64_t)h1 * N + off, y, PS);
#pragma unroll
                for (int q = 0; q < R; q++) acc[q] += y[q];
            }
            if (MODE == 1) {
                // the top tile bit, last in the same order
#pragma unroll
                for (int q = 0; q < R; q++) acc[q] += xs[NS - 1 - sd][q];
            }
            float e[R];
            if (MODE == 2) {
#pragma unroll
                for (int q = 0; q < R; q++) e[q] = e2[q];
            } else if (MODE == 3) {
                // the recurrences of one instance share its diagonal: cached loads
                vload<float, R>(J.E + off, e, PS);
            } else {
                vload_nt<float, R>(J.E + off, e, PS);
            }
            __syncthreads();
#pragma unroll
            for (int b = 0; b < HB - RB; b++) {
                float y[R];
                lds_read<float, R>(s_x + (t ^ (1 << b)) * 4, y, PS);
#pragma unroll
                for (int q = 0; q < R; q++) acc[q] += y[q];
            }
            if (late) partners();
#pragma unroll
            for (int q = 0; q < R; q++) o[q] = a1 * (J.alpha * e[q] * x[q] - J.beta * acc[q]);
        }
        if (src2 != SRC2_ZERO) {
            float y[R];
            if (src2 == SRC2_PSI && psi_const) {
#pragma unroll
                for (int q = 0; q < R; q++) y[q] = h2 == 0 ? J.psi0 : 0.0f;
            } else {
                const float* s2 = ((src2 == SRC2_PSI) ? J.psi : J.b2) + (uint64_t)h2 * N;
                vload_nt<float, R>(s2 + off, y, PS);
            }
#pragma unroll
            for (int q = 0; q < R; q++) o[q] -= a2 * y[q];
        }
        if (ac != 0.0f) {
            float y[R];
            if (psi_const) {
#pragma unroll
                for (int q = 0; q < R; q++) y[q] = hc == 0 ? J.psi0 : 0.0f;
            } else {
                vload_nt<float, R>(J.psi + (uint64_t)hc * N + off, y, PS);
            }
#pragma unroll
            for (int q = 0; q < R; q++) o[q] += ac * y[q];
        }
        if (NTS)
            vstore_nt<float, R>(J.out + (uint64_t)h * N + off, o, PS);
        else
            vstore<float, R>(J.out + (uint64_t)h * N + off, o, PS);
        if (flags & F_FINAL) {
            double nrm = 0.0;
#pragma unroll
            for (int q = 0; q < R; q++) nrm += (double)o[q] * (double)o[q];
            const double sum = wg_sum_t0<NT>(nrm, s_red);
            if (t == 0) J.norm_part[((uint64_t)h << ob) + tile] = sum;
        }
    }
}

int half_tile_bits_supported(int hb) { return hb == 13 || hb == 14; }

// variant = v + 3 nts + 6 mode: v picks the loads in flight per thread and
// the occupancy, nts = non-temporal output stores, mode = one tile per
// workgroup (0), tile pairs across the top tile bit (1), both halves of a
// tile (2).  Mode 3 (real halves only) is launch_half_real's.
int half_variants(int hb) { return 18; }
int term_variants() { return 2; }

template <int HB, int UNR, int OCC, int MODE>
static void launch_half_m(uint32_t grid, const TermJob<float>* j, uint32_t njobs, int n, int lc, bool nts, int hi,
                          hipStream_t st)
{
    if (nts)
        hipLaunchKernelGGL((half_term_kernel<HB, UNR, OCC, true, MODE>), dim3(grid), dim3(1024), 0, st, j, njobs, n,
                           lc, hi);
    else
        hipLaunchKernelGGL((half_term_kernel<HB, UNR, OCC, false, MODE>), dim3(grid), dim3(1024), 0, st, j, njobs, n,
                           lc, hi);
}

template <int HB, int UNR, int OCC>
static void launch_half_v(uint32_t grid, const void* jobs, uint32_t njobs, int n, int lc, bool nts, int mode,
                          int hi, hipStream_t st)
{
    const TermJob<float>* j = (const TermJob<float>*)jobs;
    switch (mode) {
    case 1: launch_half_m<HB, UNR, OCC, 1>(grid, j, njobs, n, lc, nts, hi, st); break;
    case 2: launch_half_m<HB, UNR, OCC, 2>(grid, j, njobs, n, lc, nts, hi, st); break;
    case 3: launch_half_m<HB, UNR, OCC, 3>(grid, j, njobs, n, lc, nts, hi, st); break;
    default: launch_half_m<HB, UNR, OCC, 0>(grid, j, njobs, n, lc, nts, hi, st); break;
    }
}

static int launch_half_mode(int hb, int variant, int mode, const void* jobs_dev, uint32_t njobs, int n, int lc, int hi,
                            hipStream_t st);

int launch_half_term(int hb, int variant, const void* jobs_dev, uint32_t njobs, int n, int lc, int hi, void* stream)
{
    if (variant < 0 || variant >= half_variants(hb)) variant = 0;
    return launch_half_mode(hb, variant, variant / 6, jobs_dev, njobs, n, lc, hi, (hipStream_t)stream);
}

int launch_half_real(int hb, int variant, const void* jobs_dev, uint32_t njobs, int n, int lc, void* stream)
{
    if (variant < 0 || variant >= half_variants(hb)) variant = 0;
    return launch_half_mode(hb, variant, 3, jobs_dev, njobs, n, lc, 0, (hipStream_t)stream);
}

static int launch_half_mode(int hb, int variant, int mode, const void* jobs_dev, uint32_t njobs, int n, int lc, int hi,
                            hipStream_t st)
{
    const int ob = n - hb;
    if (ob < 0 || !half_tile_bits_supported(hb)) return -2;
    if (hi < 0 || hi > ob) return -2;
    // a pair needs a top tile bit to pair across, and one the high-bit pass leaves
    if (mode == 1 && (ob < 1 || hi > 0)) mode = 0;
    const bool nts = (variant % 6) >= 3;
    const int v = variant % 3;
    const int pb = ob - (mode == 1 ? 1 : 0);
    const int ub = (mode == 2 || mode == 3) ? pb : pb + 1; // unit bits per job
    if (lc > ub + 16) lc = ub + 16;
    const uint64_t units = (uint64_t)njobs << ub;
    const uint64_t per = 8ull << lc;
    const uint64_t grid64 = (units + per - 1) / per * per;
    if (grid64 > 0x7fffffffull) return -3;
    const uint32_t grid = (uint32_t)grid64;
    if (hb == 14) {
        switch (v) {
        case 1: launch_half_v<14, 2, 4>(grid, jobs_dev, njobs, n, lc, nts, mode, hi, st); break;
        case 2: launch_half_v<14, 4, 4>(grid, jobs_dev, njobs, n, lc, nts, mode, hi, st); break;
        default: launch_half_v<14, 1, 8>(grid, jobs_dev, njobs, n, lc, nts, mode, hi, st); break;
        }
    } else {
        switch (v) {
        case 1: launch_half_v<13, 1, 8>(grid, jobs_dev, njobs, n, lc, nts, mode, hi, st); break;
        case 2: launch_half_v<13, 2, 8>(grid, jobs_dev, njobs, n, lc, nts, mode, hi, st); break;
        default: launch_half_v<13, 4, 8>(grid, jobs_dev, njobs, n, lc, nts, mode, hi, st); break;
        }
    }
    return check_launch("half_term");
}

// ===========================================================================
// X-string mixers (the reference's README.md:41 lists multi-X mixers as a
// to-do): the term of term_kernel with H_G = sum_i X_i replaced by
// H_M = sum_k X^{m_k}, X^m flipping every bit of the mask m.  Four consecutive
// amplitudes per thread; each mask's partners are one aligned 16-element-byte
// group whose lanes are permuted by the mask's two low bits, so every load of
// a wave stays inside the same cache lines.  FP32 and FP64.
// ===========================================================================
constexpr int kMaskTileBits = 10;

// Deterministic sum over the nt threads of a workgroup (tree in LDS), valid in every thread.
__device__ __forceinline__ double block_sum(double v, double* red, int nt)
{
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = nt >> 1; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// XY: the masks are pairs; the flip of pair m reaches amplitude x with weight
// 2 where exactly one of its two bits is set in x (X_a X_b + Y_a Y_b moves
// |01> <-> |10> with amplitude 2 and annihilates |00>, |11>).
template <typename real, bool XY>
__device__ __forceinline__ real mask_weight(uint64_t x, uint64_t m)
{
    if (!XY) return real(1);
    return __popcll(x & m) == 1 ? real(2) : real(0);
}

template <typename real, bool XY>
__global__ void __launch_bounds__(256)
mask_term_kernel(const TermJob<real>* __restrict__ jobs, uint32_t njobs, int n, const uint32_t* __restrict__ masks,
                 int nmask)
{
    __shared__ double red[256];
    const int tmb = n < kMaskTileBits ? n : kMaskTileBits;
    const int nt = 1 << (tmb - 2);
    const uint64_t tiles = 1ull << (n - tmb);
    const uint64_t job = blockIdx.x / tiles;
    if (job >= njobs) return;
    const uint64_t tile = blockIdx.x - job * tiles;
    const TermJob<real>& J = jobs[job];
    const uint32_t flags = J.flags;
    const uint32_t src1 = flags & 3u, src2 = flags & 12u;
    const bool psi_const = (flags & F_PSI_CONST) != 0;
    const uint64_t N = 1ull << n;
    const int t = threadIdx.x;
    const uint64_t x0 = (tile << tmb) + 4 * (uint64_t)t;

    real ps = real(1);
    if (flags & F_NORM_IN) {
        double s = 0.0;
        for (uint64_t k = t; k < tiles; k += nt) s += J.norm_in[k];
        const double tot = block_sum(s, red, nt);
        ps = (real)(1.0 / sqrt(tot));
        if ((flags & F_READOUT) && tile == 0 && t == 0) {
            const double re = (double)J.psi[J.eloc], im = (double)J.psi[N + J.eloc];
            *J.prob = (re * re + im * im) / tot;
        }
    }

    real pr[4], pi[4], e[4], yr[4], yi[4], ur[4], ui[4];
    if (!psi_const) {
        vload<real, 4>(J.psi + x0, pr);
        vload<real, 4>(J.psi + N + x0, pi);
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            pr[q] = J.psi0;
            pi[q] = real(0);
        }
    }
    vload<real, 4>(J.E + x0, e);
    if (src2 == SRC2_ARRAY) {
        vload<real, 4>(J.b2 + x0, yr);
        vload<real, 4>(J.b2 + N + x0, yi);
    }
    if (src1 == SRC1_ZERO) {
#pragma unroll
        for (int q = 0; q < 4; q++) ur[q] = ui[q] = real(0);
    } else if (src1 == SRC1_PSI && psi_const && !XY) {
        // |+> is an eigenvector of every X-string, eigenvalue 1 each
#pragma unroll
        for (int q = 0; q < 4; q++) {
            ur[q] = J.alpha * e[q] * J.psi0 - J.beta * (real)nmask * J.psi0;
            ui[q] = real(0);
        }
    } else if (src1 == SRC1_PSI && psi_const) {
        // |+> under XY pairs: psi0 times the sum of the weights at x
#pragma unroll
        for (int q = 0; q < 4; q++) {
            real w = real(0);
            for (int k = 0; k < nmask; k++) w += mask_weight<real, true>(x0 + q, masks[k]);
            ur[q] = J.alpha * e[q] * J.psi0 - J.beta * w * J.psi0;
            ui[q] = real(0);
        }
    } else {
        const real* src = (src1 == SRC1_PSI) ? J.psi : J.b1;
        real xr[4], xi[4], ar[4], ai[4];
        if (src1 == SRC1_PSI) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                xr[q] = pr[q];
                xi[q] = pi[q];
            }
        } else {
            vload<real, 4>(src + x0, xr);
            vload<real, 4>(src + N + x0, xi);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) ar[q] = ai[q] = real(0);
        for (int k = 0; k < nmask; k++) {
            const uint64_t m = masks[k];
            const uint64_t base = (x0 ^ m) & ~(uint64_t)3;
            const int lo = (int)(m & 3u);
            real tr[4], ti[4];
            vload<real, 4>(src + base, tr);
            vload<real, 4>(src + N + base, ti);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const real w = mask_weight<real, XY>(x0 + q, m);
                ar[q] += w * tr[q ^ lo];
                ai[q] += w * ti[q ^ lo];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            ur[q] = J.alpha * e[q] * xr[q] - J.beta * ar[q];
            ui[q] = J.alpha * e[q] * xi[q] - J.beta * ai[q];
        }
    }
    real w1r = J.w1r, w1i = J.w1i, w2r = J.w2r, w2i = J.w2i;
    if (src1 == SRC1_PSI) {
        w1r *= ps;
        w1i *= ps;
    }
    if (src2 == SRC2_ZERO) {
#pragma unroll
        for (int q = 0; q < 4; q++) yr[q] = yi[q] = real(0);
    } else if (src2 == SRC2_PSI) {
        w2r *= ps;
        w2i *= ps;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            yr[q] = pr[q];
            yi[q] = pi[q];
        }
    }
    const real cr = J.cr * ps, ci = J.ci * ps;
    real orr[4], oi[4];
    double nrm = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        orr[q] = (w1r * ur[q] - w1i * ui[q]) - (w2r * yr[q] - w2i * yi[q]) + (cr * pr[q] - ci * pi[q]);
        oi[q] = (w1r * ui[q] + w1i * ur[q]) - (w2r * yi[q] + w2i * yr[q]) + (cr * pi[q] + ci * pr[q]);
        nrm += (double)orr[q] * (double)orr[q] + (double)oi[q] * (double)oi[q];
    }
    vstore<real, 4>(J.out + x0, orr);
    vstore<real, 4>(J.out + N + x0, oi);
    if (flags & F_FINAL) {
        const double s = block_sum(nrm, red, nt);
        if (t == 0) J.norm_part[tile] = s;
    }
}

int mask_tile_bits(int n) { return n < kMaskTileBits ? n : kMaskTileBits; }

int launch_mask_term(int precision, const void* jobs_dev, uint32_t njobs, int n, const uint32_t* masks_dev, int nmask,
                     bool xy, void* stream)
{
    if (n < 4 || nmask < 1) return -2;
    const int tmb = mask_tile_bits(n);
    const uint64_t grid = (uint64_t)njobs << (n - tmb);
    if (grid == 0) return 0;
    if (grid > 0x7fffffffull) return -3;
    hipStream_t st = (hipStream_t)stream;
    const dim3 block(1u << (tmb - 2));
    if (precision == MQW_FP32 && xy)
        hipLaunchKernelGGL((mask_term_kernel<float, true>), dim3((uint32_t)grid), block, 0, st,
                           (const TermJob<float>*)jobs_dev, njobs, n, masks_dev, nmask);
    else if (precision == MQW_FP32)
        hipLaunchKernelGGL((mask_term_kernel<float, false>), dim3((uint32_t)grid), block, 0, st,
                           (const TermJob<float>*)jobs_dev, njobs, n, masks_dev, nmask);
    else if (xy)
        hipLaunchKernelGGL((mask_term_kernel<double, true>), dim3((uint32_t)grid), block, 0, st,
                           (const TermJob<double>*)jobs_dev, njobs, n, masks_dev, nmask);
    else
        hipLaunchKernelGGL((mask_term_kernel<double, false>), dim3((uint32_t)grid), block, 0, st,
                           (const TermJob<double>*)jobs_dev, njobs, n, masks_dev, nmask);
    return check_launch("mask_term");
}

// ===========================================================================
// High-bit pass (FP32).  For every job flagged F_HSUM, hsum[k] = sum over the
// top H bits b of s1[k ^ 2^b], both real halves.  One workgroup takes a pencil
// of 16 * 2^WB consecutive amplitudes (a row of 64 * 2^WB bytes) by all 2^H
// values of the top bits: the 2^H rows are staged in 64 KiB of LDS, two row
// bits are flipped in registers and the other H - 2 through LDS.  The sweep
// that follows reads hsum instead of loading the partner tiles of those bits,
// which at large n miss L2.
// ===========================================================================
template <int H, int WB, int PPW>
__global__ void __launch_bounds__(1 << (H + WB))
high_pass_kernel(const TermJob<float>* __restrict__ jobs, uint32_t njobs, int n)
{
    using f4 = NVec16<float>::type;
    constexpr int ROWS = 1 << H;
    constexpr int RG = ROWS / 4;      // row groups; thread t holds rows q*RG + rg, q < 4
    constexpr int CW = 4 << WB;       // float4 columns per row
    constexpr int CB = 4 + WB;        // amplitude bits inside a row
    __shared__ f4 s_rows[ROWS * CW];

    const int p0 = n - H; // the first of the top bits
    const uint64_t npen = 1ull << (p0 - CB);
    const uint64_t N = 1ull << n;
    const int t = threadIdx.x;
    const int col = t & (CW - 1), rg = t >> (2 + WB);
    // PPW consecutive (job, half, pencil) units per workgroup; runs of 8
    // workgroups on one XCD, so rows that share a 128 B line meet in one L2.
    // The next unit's rows are loaded while the current one is summed.
    const uint64_t u0 = xcd_unit(blockIdx.x, 3) * PPW;
    auto source = [&](uint64_t unit, const float*& src, float*& dst) -> bool {
        const uint64_t job = unit / (2 * npen);
        if (job >= njobs) return false;
        const TermJob<float>& J = jobs[job];
        if (!(J.flags & F_HSUM)) return false;
        const uint64_t rem = unit - job * 2 * npen;
        const uint64_t h = rem / npen, pen = rem - h * npen;
        src = (((J.flags & 3u) == SRC1_PSI) ? J.psi : J.b1) + h * N + (pen << CB);
        dst = J.hsum + h * N + (pen << CB);
        return true;
    };
    auto load = [&](const float* src, f4 (&x)[4]) {
#pragma unroll
        for (int q = 0; q < 4; q++)
            x[q] = *gview(reinterpret_cast<const f4*>(src + ((uint64_t)(q * RG + rg) << p0)) + col);
    };

    const float* src;
    float* dst;
    bool have = source(u0, src, dst);
    f4 x[4];
    if (have) load(src, x);
#pragma unroll
    for (int i = 0; i < PPW; i++) {
        const float* nsrc = nullptr;
        float* ndst = nullptr;
        f4 nx[4];
        const bool nhave = i + 1 < PPW && source(u0 + i + 1, nsrc, ndst);
        if (nhave) load(nsrc, nx);
        if (have) {
            if (i > 0) __syncthreads(); // the previous unit's LDS reads are done
#pragma unroll
            for (int q = 0; q < 4; q++) s_rows[(q * RG + rg) * CW + col] = x[q];
            // row bits H-2 and H-1 (the q bits) in registers
            f4 acc[4];
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = x[q ^ 1] + x[q ^ 2];
            __syncthreads();
#pragma unroll
            for (int b = 0; b < H - 2; b++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q] += s_rows[(q * RG + (rg ^ (1 << b))) * CW + col];
#pragma unroll
            for (int q = 0; q < 4; q++)
                __builtin_nontemporal_store(acc[q],
                                            gview(reinterpret_cast<f4*>(dst + ((uint64_t)(q * RG + rg) << p0)) + col));
        }
        if (nhave) {
#pragma unroll
            for (int q = 0; q < 4; q++) x[q] = nx[q];
            src = nsrc;
            dst = ndst;
        }
        have = nhave;
    }
}

// Register variant of the high-bit pass for up to 7 top bits, without LDS or
// barriers: one wave per pencil of 32 amplitudes (128 B rows) by 2^H rows,
// H = 3 + RB.  Lane l holds columns (l & 7) of rows (q << 3) | (l >> 3), so
// each load instruction reads 8 whole 128 B lines; the three lane row bits
// are exchanged by cross-lane shuffles, the RB others in registers.
template <int RB>
__global__ void __launch_bounds__(256)
high_pass_reg_kernel(const TermJob<float>* __restrict__ jobs, uint32_t njobs, int n)
{
    using f4 = NVec16<float>::type;
    constexpr int H = 3 + RB;
    constexpr int Q = 1 << RB;
    const int p0 = n - H;
    const uint64_t npen = 1ull << (p0 - 5);
    const int lane = threadIdx.x & 63;
    const uint64_t unit = ((uint64_t)blockIdx.x << 2) + (threadIdx.x >> 6); // one pencil per wave
    const uint64_t job = unit / (2 * npen);
    if (job >= njobs) return; // uniform per wave; the kernel has no barriers
    const TermJob<float>& J = jobs[job];
    if (!(J.flags & F_HSUM)) return;
    const uint64_t rem = unit - job * 2 * npen;
    const uint64_t h = rem / npen, pen = rem - h * npen;
    const uint64_t N = 1ull << n;
    const float* src = (((J.flags & 3u) == SRC1_PSI) ? J.psi : J.b1) + h * N + (pen << 5);
    float* dst = J.hsum + h * N + (pen << 5);
    const int col = lane & 7, rl = lane >> 3;

    f4 x[Q];
#pragma unroll
    for (int q = 0; q < Q; q++)
        x[q] = *gview(reinterpret_cast<const f4*>(src + ((uint64_t)((q << 3) | rl) << p0)) + col);
    f4 acc[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        acc[q] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < RB; b++) acc[q] += x[q ^ (1 << b)];
    }
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int q = 0; q < Q; q++) {
            f4 y;
#pragma unroll
            for (int w = 0; w < 4; w++) y[w] = __shfl_xor(x[q][w], 8 << b, 64);
            acc[q] += y;
        }
#pragma unroll
    for (int q = 0; q < Q; q++)
        __builtin_nontemporal_store(acc[q],
                                    gview(reinterpret_cast<f4*>(dst + ((uint64_t)((q << 3) | rl) << p0)) + col));
}

// wb: row of 64 * 2^wb bytes; hi + wb <= 10 keeps the rows in 64 KiB of LDS
int high_pass_supported(int n, int hi, int wb)
{
    if (wb < 0) return hi >= 4 && hi <= 7 && n - hi >= 5; // register variant, 128 B rows
    return wb <= 2 && hi >= 4 && hi + wb <= 10 && n - hi >= 4 + wb;
}

// pencils per workgroup (1 measured best: a second, prefetched pencil was
// 8 % slower at n = 24..26)
constexpr int kHighPassPPW = 1;

template <int H, int WB>
static void launch_hp(uint64_t grid, const void* jobs, uint32_t njobs, int n, hipStream_t st)
{
    hipLaunchKernelGGL((high_pass_kernel<H, WB, kHighPassPPW>), dim3((uint32_t)grid), dim3(1 << (H + WB)), 0, st,
                       (const TermJob<float>*)jobs, njobs, n);
}

template <int WB>
static void launch_hp_w(int hi, uint64_t grid, const void* jobs, uint32_t njobs, int n, hipStream_t st)
{
    switch (hi) {
    case 4: launch_hp<4, WB>(grid, jobs, njobs, n, st); break;
    case 5: launch_hp<5, WB>(grid, jobs, njobs, n, st); break;
    case 6: launch_hp<6, WB>(grid, jobs, njobs, n, st); break;
    case 7: launch_hp<7, WB>(grid, jobs, njobs, n, st); break;
    case 8: launch_hp<8, WB>(grid, jobs, njobs, n, st); break;
    case 9:
        if constexpr (WB <= 1) launch_hp<9, WB>(grid, jobs, njobs, n, st);
        break;
    default:
        if constexpr (WB == 0) launch_hp<10, WB>(grid, jobs, njobs, n, st);
        break;
    }
}

int launch_high_pass(int hi, int wb, const void* jobs_dev, uint32_t njobs, int n, void* stream)
{
    if (!high_pass_supported(n, hi, wb)) return -2;
    hipStream_t st = (hipStream_t)stream;
    if (wb < 0) {
        const uint64_t units = (uint64_t)njobs * 2ull * (1ull << (n - hi - 5));
        const uint64_t grid = (units + 3) / 4; // 4 waves, one pencil each
        if (grid == 0) return 0;
        if (grid > 0x7fffffffull) return -3;
        const TermJob<float>* j = (const TermJob<float>*)jobs_dev;
        switch (hi) {
        case 4: hipLaunchKernelGGL(high_pass_reg_kernel<1>, dim3((uint32_t)grid), dim3(256), 0, st, j, njobs, n); break;
        case 5: hipLaunchKernelGGL(high_pass_reg_kernel<2>, dim3((uint32_t)grid), dim3(256), 0, st, j, njobs, n); break;
        case 6: hipLaunchKernelGGL(high_pass_reg_kernel<3>, dim3((uint32_t)grid), dim3(256), 0, st, j, njobs, n); break;
        default: hipLaunchKernelGGL(high_pass_reg_kernel<4>, dim3((uint32_t)grid), dim3(256), 0, st, j, njobs, n); break;
        }
        return check_launch("high_pass_reg");
    }
    const uint64_t units = (uint64_t)njobs * 2ull * (1ull << (n - hi - 4 - wb));
    if (units == 0) return 0;
    const uint64_t wgs = (units + kHighPassPPW - 1) / kHighPassPPW;
    const uint64_t grid = (wgs + 63) / 64 * 64; // whole XCD runs of 8 workgroups
    if (grid > 0x7fffffffull) return -3;
    if (wb == 2)
        launch_hp_w<2>(hi, grid, jobs_dev, njobs, n, st);
    else if (wb == 1)
        launch_hp_w<1>(hi, grid, jobs_dev, njobs, n, st);
    else
        launch_hp_w<0>(hi, grid, jobs_dev, njobs, n, st);
    return check_launch("high_pass");
}

// ===========================================================================
// Shared first stage: each sample's output psi_j = sum_k c_j[k] i^(k&1) V_k
// from one instance's real Chebyshev vectors.  A workgroup takes one
// contiguous chunk of N / nparts amplitudes for up to kCombineBatch samples:
// every thread keeps the KR basis values of its amplitude in registers, reads
// each sample's coefficients from LDS (broadcast) and writes both real halves
// of that sample's output.  The |psi_j|^2 of the chunk goes to
// norm_part[chunk] of the sample: a fixed-order sum (cross-lane tree, then
// the 4 waves in order), so the norm is the same on every run.
// ===========================================================================
template <int KR>
__global__ void __launch_bounds__(256)
combine_kernel(const CombineJob* __restrict__ jobs, uint32_t njobs, const CombineItem* __restrict__ items,
               const float* __restrict__ coefs, int n, uint32_t nparts)
{
    __shared__ float s_c[kCombineBatch * KR];
    __shared__ double s_w[4][kCombineBatch];
    const uint32_t job = blockIdx.x / nparts, part = blockIdx.x % nparts;
    if (job >= njobs) return;
    const CombineJob& J = jobs[job];
    const int kc = (int)J.kc, ns = (int)J.ns;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = t; i < ns * KR; i += 256) {
        const int j = i / KR, k = i % KR;
        s_c[i] = k < kc ? coefs[items[J.item_off + j].coef_off + k] : 0.0f;
    }
    for (int i = t; i < 4 * kCombineBatch; i += 256) (&s_w[0][0])[i] = 0.0;
    __syncthreads();
    const uint64_t N = 1ull << n;
    const uint64_t C = N / nparts;
    const uint64_t base = (uint64_t)part * C;
    for (uint64_t a0 = 0; a0 < C; a0 += 256) {
        const uint64_t a = base + a0 + (uint64_t)t;
        const MQW_GLOBAL float* V = gview(J.V);
        float v[KR];
        v[0] = J.V0 ? gview(J.V0)[a] : J.unit1 ? (a + 1 == J.unit1 ? 1.0f : 0.0f) : J.psi0;
#pragma unroll
        for (int k = 1; k < KR; k++) v[k] = k < kc ? V[(uint64_t)(k - 1) * N + a] : 0.0f;
        // keep the column in registers for the whole sample loop (the loads
        // are otherwise re-issued per sample)
#pragma unroll
        for (int k = 0; k < KR; k++) asm volatile("" : "+v"(v[k]));
        for (int j = 0; j < ns; j++) {
            const float* c = s_c + j * KR;
            float re = 0.0f, im = 0.0f;
#pragma unroll
            for (int k = 0; k < KR; k += 2) {
                re += c[k] * v[k];
                im += c[k + 1] * v[k + 1];
            }
            MQW_GLOBAL float* o = gview(items[J.item_off + j].out);
            if (J.accum) {
                re += o[a];
                im += o[N + a];
            }
            o[a] = re;
            o[N + a] = im;
            double s = (double)re * (double)re + (double)im * (double)im;
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
            if (lane == 0) s_w[w][j] += s;
        }
    }
    __syncthreads();
    if (t < ns && !J.accum)
        items[J.item_off + t].norm_part[part] = ((s_w[0][t] + s_w[1][t]) + s_w[2][t]) + s_w[3][t];
}

int launch_combine(const CombineJob* jobs_dev, uint32_t njobs, const CombineItem* items_dev, const float* coefs_dev,
                   int n, uint32_t nparts, uint32_t kc_max, void* stream)
{
    if (njobs == 0) return 0;
    const uint64_t N = 1ull << n;
    // whole 256-amplitude rows per workgroup
    if (nparts == 0 || N % nparts != 0 || (N / nparts) % 256 != 0) return -2;
    if (kc_max < 1 || kc_max > (uint32_t)kCombineMaxKX) return -2;
    const uint64_t grid = (uint64_t)njobs * nparts;
    if (grid > 0x7fffffffull) return -3;
    hipStream_t st = (hipStream_t)stream;
    if (kc_max <= 16)
        hipLaunchKernelGGL(combine_kernel<16>, dim3((uint32_t)grid), dim3(256), 0, st, jobs_dev, njobs, items_dev,
                           coefs_dev, n, nparts);
    else if (kc_max <= 32)
        hipLaunchKernelGGL(combine_kernel<32>, dim3((uint32_t)grid), dim3(256), 0, st, jobs_dev, njobs, items_dev,
                           coefs_dev, n, nparts);
    else if (kc_max <= 48)
        hipLaunchKernelGGL(combine_kernel<48>, dim3((uint32_t)grid), dim3(256), 0, st, jobs_dev, njobs, items_dev,
                           coefs_dev, n, nparts);
    else if (kc_max <= 64)
        hipLaunchKernelGGL(combine_kernel<64>, dim3((uint32_t)grid), dim3(256), 0, st, jobs_dev, njobs, items_dev,
                           coefs_dev, n, nparts);
    else if (kc_max <= 96)
        hipLaunchKernelGGL(combine_kernel<96>, dim3((uint32_t)grid), dim3(256), 0, st, jobs_dev, njobs, items_dev,
                           coefs_dev, n, nparts);
    else
        hipLaunchKernelGGL(combine_kernel<128>, dim3((uint32_t)grid), dim3(256), 0, st, jobs_dev, njobs, items_dev,
                           coefs_dev, n, nparts);
    return check_launch("combine");
}

// ===========================================================================
// Shared second stage (see Combine2Tile): out_j = E_j W on the matrix cores.
// A workgroup of 8 waves takes one chunk of N / nparts amplitudes for up to
// 32 MB samples; per 256 amplitudes, wave w owns 32 of them and MB row blocks
// of 32 samples, both classes: 2 MB accumulator tiles of
// v_mfma_f32_32x32x2f32 (A = coefficients, 32 samples x 2 rows; B = basis,
// 2 rows x 32 amplitudes).  A = d_jl c_jk s is built by the workgroup, 32 rows
// at a time, into a double-buffered LDS chunk (one barrier per chunk); each
// lane's B entry is one load of row (l, k) of the basis, UNR row pairs ahead
// of the matrix ops.  The squares of each output go through a 32-lane tree
// into FP64 per wave, summed in wave order.
// ===========================================================================
typedef float mqw_f32x16 __attribute__((ext_vector_type(16)));

// One step of a reduce-scatter over the 32 lanes of each half of a wave, NV
// values per lane: the lane keeps the half of its values whose index bit
// (log2 NV - 1) equals its lane bit log2 D, plus its partner's at distance D;
// once one value is left, the remaining steps add the partner's (all-reduce).
template <int NV, int D>
__device__ __forceinline__ void norm_reduce_step(float* v, int col)
{
    if constexpr (D > 0) {
        if constexpr (NV >= 2) {
            constexpr int H = NV / 2;
            const bool up = (col & D) != 0;
#pragma unroll
            for (int i = 0; i < H; i++) {
                const float send = up ? v[i] : v[i + H];
                const float keep = up ? v[i + H] : v[i];
                v[i] = keep + __shfl_xor(send, D, 64);
            }
            norm_reduce_step<H, D / 2>(v, col);
        } else {
            v[0] += __shfl_xor(v[0], D, 64);
            norm_reduce_step<1, D / 2>(v, col);
        }
    }
}

template <int MB>
__global__ void __launch_bounds__(512, 4)
combine2_kernel(const Combine2Tile* __restrict__ tiles, uint32_t ntiles, const Combine2Item* __restrict__ items,
                const float* __restrict__ coefs, int n, uint32_t nparts)
{
    constexpr int NS = MB * 32;
    constexpr int CP = kC2MaxK + 1;   // padded coefficient rows: consecutive samples hit consecutive banks
    constexpr int UNR = 8;            // row pairs per block, loaded one block ahead
    constexpr int QC = 32;            // A rows per LDS chunk (a multiple of 2 UNR)
    constexpr int BPC = QC / (2 * UNR);
    __shared__ float s_c1[NS * CP];
    __shared__ float s_c2[NS * CP];
    __shared__ uint32_t s_q[kC2MaxK * kC2MaxK + 2];
    __shared__ float s_a[2][QC][NS];
    __shared__ double s_n[8][NS];
    const uint32_t tile = blockIdx.x / nparts, part = blockIdx.x % nparts;
    if (tile >= ntiles) return;
    const Combine2Tile& T = tiles[tile];
    const int ns = (int)T.ns, kc1 = (int)T.kc1, kc2 = (int)T.kc2;
    const int q0 = (int)T.q0, nq = (int)(T.q0 + T.q1);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int col = lane & 31, half = lane >> 5;
    for (int i = t; i < NS * CP; i += 512) {
        const int j = i / CP, k = i % CP;
        float a = 0.0f, b = 0.0f;
        if (j < ns) {
            const Combine2Item& I = items[T.item_off + j];
            if (k < kc1) a = coefs[I.c1_off + k];
            if (k < kc2) b = coefs[I.c2_off + k];
        }
        s_c1[i] = a;
        s_c2[i] = b;
    }
    for (int i = t; i < nq; i += 512) s_q[i] = T.qtab[i];
    for (int i = t; i < 8 * NS; i += 512) (&s_n[0][0])[i] = 0.0;
    __syncthreads();
    const uint64_t N = 1ull << n;
    const uint64_t C = N / nparts;
    const uint64_t base = (uint64_t)part * C;
    const MQW_GLOBAL float* plus = gview(T.plus);
    const MQW_GLOBAL float* V = gview(T.V);
    const MQW_GLOBAL float* W = gview(T.W);
    // A rows [r0, r1) into buffer b: s_a[b][r - r0][j] = s d_jl c_jk
    auto fill_a = [&](int b, int r0, int r1) {
        for (int i = t; i < QC * NS; i += 512) {
            const int r = i / NS, j = i % NS;
            float v = 0.0f;
            if (r0 + r < r1) {
                const uint32_t e = s_q[r0 + r];
                const int l = (int)(e & 255u), k = (int)((e >> 8) & 255u);
                const uint32_t sg = e >> 16;
                const float f = s_c2[j * CP + l] * s_c1[j * CP + k];
                v = sg == 1u ? f : (sg == 2u ? -f : 0.0f);
            }
            s_a[b][r][j] = v;
        }
    };
    // basis row q of this lane's half, at amplitude a
    auto load_b = [&](int q, uint64_t a) -> float {
        const uint32_t e = s_q[q];
        const int l = (int)(e & 255u), k = (int)((e >> 8) & 255u);
        const MQW_GLOBAL float* row =
            l == 0 ? (k == 0 ? plus : V + (uint64_t)(k - 1) * N) : W + ((uint64_t)k * (kc2 - 1) + (l - 1)) * N;
        return row[a];
    };
    for (uint64_t a0 = 0; a0 < C; a0 += 256) {
        const uint64_t a = base + a0 + (uint64_t)(w * 32 + col);
        mqw_f32x16 acc[2][MB];
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[c][mb][r] = 0.0f;
#pragma unroll
        for (int cls = 0; cls < 2; cls++) {
            const int qb = cls ? q0 : 0, qe = cls ? nq : q0;
            const int nblk = (qe - qb + 2 * UNR - 1) / (2 * UNR);
            if (nblk == 0) continue;
            const int nch = (qe - qb + QC - 1) / QC;
            float bv[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const int q = qb + 2 * u;
                bv[u] = q < qe ? load_b(q + half, a) : 0.0f;
            }
            fill_a(0, qb, qe < qb + QC ? qe : qb + QC);
            __syncthreads();
            for (int blk = 0; blk < nblk; blk++) {
                const int ch = blk / BPC;
                if (blk % BPC == 0 && ch + 1 < nch) {
                    const int r0 = qb + (ch + 1) * QC;
                    fill_a((ch + 1) & 1, r0, qe < r0 + QC ? qe : r0 + QC);
                }
                const int rs = qb + blk * 2 * UNR;
                float bn[UNR];
#pragma unroll
                for (int u = 0; u < UNR; u++) {
                    const int q = rs + 2 * UNR + 2 * u;
                    bn[u] = q < qe ? load_b(q + half, a) : 0.0f;
                }
                const float(*A)[NS] = s_a[ch & 1];
                const int rc = rs - (qb + ch * QC); // first row of the block inside the chunk
#pragma unroll
                for (int u = 0; u < UNR; u++) {
                    if (rs + 2 * u >= qe) break; // uniform
#pragma unroll
                    for (int mb = 0; mb < MB; mb++) {
                        const float av = A[rc + 2 * u + half][mb * 32 + col];
                        acc[cls][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv[u], acc[cls][mb], 0, 0, 0);
                    }
                }
#pragma unroll
                for (int u = 0; u < UNR; u++) bv[u] = bn[u];
                // end of a chunk: the next chunk is written, this one's reads are done
                if (blk % BPC == BPC - 1 || blk == nblk - 1) __syncthreads();
            }
        }
        // lane (col, half) holds rows 8 (r / 4) + 4 half + r % 4 of column col
        float sq[16 * MB];
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int j = mb * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
                const float re = acc[0][mb][r], im = acc[1][mb][r];
                if (j < ns) {
                    MQW_GLOBAL float* o = gview(items[T.item_off + j].out);
                    o[a] = re;
                    o[N + a] = im;
                }
                sq[16 * mb + r] = re * re + im * im;
            }
        // summed over the half's 32 lanes (see combine2h_kernel)
        norm_reduce_step<16 * MB, 16>(sq, col);
        constexpr int KQ = MB >= 2 ? MB / 2 : 1;
#pragma unroll
        for (int k = 0; k < KQ; k++) {
            const int q = MB >= 2 ? KQ * col + k : col >> 1;
            const int mb = q >> 4, r = q & 15;
            const int j = mb * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
            if (MB >= 2 || (col & 1) == 0) s_n[w][j] += (double)sq[k];
        }
    }
    __syncthreads();
    if (t < ns) {
        double sum = 0.0;
        for (int q = 0; q < 8; q++) sum += s_n[q][t];
        items[T.item_off + t].norm_part[part] = sum;
    }
}

int launch_combine2(const Combine2Tile* tiles_dev, uint32_t ntiles, const Combine2Item* items_dev,
                    const float* coefs_dev, int n, uint32_t nparts, uint32_t ns_max, void* stream)
{
    if (ntiles == 0) return 0;
    const uint64_t N = 1ull << n;
    if (nparts == 0 || N % nparts != 0 || (N / nparts) % 256 != 0) return -2;
    if (ns_max < 1 || ns_max > (uint32_t)kC2Tile) return -2;
    const uint64_t grid = (uint64_t)ntiles * nparts;
    if (grid > 0x7fffffffull) return -3;
    hipStream_t st = (hipStream_t)stream;
    if (ns_max <= 32)
        hipLaunchKernelGGL(combine2_kernel<1>, dim3((uint32_t)grid), dim3(512), 0, st, tiles_dev, ntiles, items_dev,
                           coefs_dev, n, nparts);
    else
        hipLaunchKernelGGL(combine2_kernel<2>, dim3((uint32_t)grid), dim3(512), 0, st, tiles_dev, ntiles, items_dev,
                           coefs_dev, n, nparts);
    return check_launch("combine2");
}

// ===========================================================================
// Shared second stage on the FP16 matrix cores with FP32 accuracy.  Each
// operand x is split as hi = fp16(x), lo = fp16(x - hi) (22 significant bits
// together; the basis scaled by 2^12 and the coefficients by the tile's power
// of two, so both parts stay in FP16's normal range for every entry that
// matters) and out = Ahi Bhi + Ahi Blo + Alo Bhi, accumulated in FP32 by
// v_mfma_f32_32x32x16_f16; the dropped Alo Blo is below 2^-22 of each product.
//
// combine2_prep_kernel writes each tile's A = s d_jl c_jk (scaled) once, as
// 16-row blocks [block][hi, lo][rows 0-7, rows 8-15][NSP samples][8 rows] of
// FP16 (so the lanes of a wave read consecutive 16-byte pieces), the layout the
// main kernel copies into LDS.  combine2h_kernel: a workgroup of NW = 16 waves
// takes one chunk of N / nparts amplitudes for up to NSP = 32 MB samples; per
// 512 amplitudes, wave w owns 32 of them and MB accumulator tiles of 32
// samples, one class (real rows, then imaginary rows) at a time.  Everything
// reaches LDS by direct global-to-LDS loads (16 B per lane): each wave copies
// its own 32 amplitudes of every 16-row basis block (128 B per row) into its
// own three-slot ring, three blocks ahead, and the waves copy each block of A
// into a three-buffer ring shared by the workgroup, two blocks ahead.  Lane
// (col, half) then reads rows 8 half + i (i < 8) of the block at its
// amplitude.  The copies' order makes a count of loads in flight enough to
// know a block has arrived (see the loop), so no wait drains the pipeline.
// ===========================================================================
typedef _Float16 mqw_h8 __attribute__((ext_vector_type(8)));

__global__ void __launch_bounds__(256)
combine2_prep_kernel(const Combine2Tile* __restrict__ tiles, const Combine2Item* __restrict__ items,
                     const float* __restrict__ coefs, _Float16* __restrict__ A, uint32_t nsp)
{
    const Combine2Tile& T = tiles[blockIdx.y];
    const uint32_t nq = T.q0 + T.q1;
    const uint32_t total = nq * nsp;
    _Float16* At = A + T.a_off;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < total; i += gridDim.x * 256u) {
        const uint32_t r = i / nsp, j = i % nsp;
        const uint32_t e = T.qtab[r];
        const uint32_t l = e & 255u, k = (e >> 8) & 255u, sg = e >> 16;
        float v = 0.0f;
        if (j < T.ns && sg != 0u) {
            const Combine2Item& I = items[T.item_off + j];
            const float f = coefs[I.c2_off + l] * coefs[I.c1_off + k];
            v = (sg == 1u ? f : -f) * T.ascale;
        }
        const _Float16 hi = (_Float16)v;
        const _Float16 lo = (_Float16)(v - (float)hi);
        const uint32_t b = r >> 4, rr = r & 15u;
        const size_t pos = ((size_t)(rr >> 3) * nsp + j) * 8u + (rr & 7u); // within one part of the block
        At[(size_t)(2u * b) * nsp * 16u + pos] = hi;
        At[(size_t)(2u * b + 1u) * nsp * 16u + pos] = lo;
    }
}

// s_waitcnt as inline assembly with a memory clobber, so the compiler moves no
// memory access across it: at most VM vector memory operations of this wave
// left in flight, and no LDS (or scalar memory) operation.
template <int VM>
__device__ __forceinline__ void wait_vm_lgkm0()
{
    static_assert(VM >= 0 && VM < 16, "vmcnt out of range");
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)" : : "n"(VM) : "memory");
}
// the same with a run-time count (wave-uniform), one of 0, 2, 3, 4, 5; any
// other value waits for everything
__device__ __forceinline__ void wait_vm_lgkm0(int vm)
{
    switch (vm) {
    case 5: wait_vm_lgkm0<5>(); break;
    case 4: wait_vm_lgkm0<4>(); break;
    case 3: wait_vm_lgkm0<3>(); break;
    case 2: wait_vm_lgkm0<2>(); break;
    default: wait_vm_lgkm0<0>(); break;
    }
}

// 16 bytes per lane from global memory straight into LDS: lane l's bytes land
// at LDS byte address lds + 16 l (lds is wave-uniform).  Inline assembly, so
// the compiler neither counts these loads nor makes its LDS reads wait for
// them (its wait insertion cannot tell LDS buffers apart): the caller waits
// explicitly.  This wave's earlier LDS reads complete first, since the copy
// may overwrite what they read.  The compiler's own vmcnt waits stay correct:
// loads it does not count only make them stricter (vector loads complete in
// issue order).
__device__ __forceinline__ void lds_copy16(const void* src, uint32_t lds)
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                 :
                 : "s"(lds), "v"(src)
                 : "memory", "m0");
}

template <int MB, int NW>
__global__ void __launch_bounds__(NW * 64, 4)
combine2h_kernel(const Combine2Tile* __restrict__ tiles, uint32_t ntiles, const Combine2Item* __restrict__ items,
                 const _Float16* __restrict__ A, int n, uint32_t nparts)
{
    constexpr int NT = NW * 64;
    constexpr int NSP = MB * 32;
    constexpr int BLK = 2 * NSP * 16;  // halves of one 16-row block (hi and lo)
    constexpr float kWScale = 4096.0f; // 2^12: basis entries are at most 1 in magnitude
    constexpr int SPB = BLK * 2 / 1024; // 1 KiB copy slots per block of A
    static_assert(SPB * 1024 == BLK * 2 && SPB <= NW, "a block is a whole number of slots, at most one per wave");
    // A: three buffers of one 16-row block, shared by the waves, two blocks
    // ahead.  Basis: per wave, three slots of one block (16 rows x the wave's
    // 32 amplitudes), three blocks ahead.
    __shared__ __attribute__((aligned(16))) _Float16 s_a[3 * BLK];
    __shared__ __attribute__((aligned(16))) float s_b[NW][3][16 * 32];
    __shared__ uint16_t s_rc[kC2MaxRows];
    __shared__ float* s_out[NSP];
    __shared__ double s_n[NW][NSP];
    const uint32_t tile = blockIdx.x / nparts, part = blockIdx.x % nparts;
    if (tile >= ntiles) return;
    const Combine2Tile& T = tiles[tile];
    const int ns = (int)T.ns, kc2 = (int)T.kc2;
    const int q0 = (int)T.q0, nq = (int)(T.q0 + T.q1);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int col = lane & 31, half = lane >> 5;
    const uint64_t N = 1ull << n;
    // row codes: 0 for |+>, k for V_k (k < 128), 128 + k (kc2 - 1) + l - 1
    // for W_lk (below 16384 with kc1, kc2 <= 128)
    for (int i = t; i < nq; i += NT) {
        const uint32_t e = T.qtab[i];
        const int l = (int)(e & 255u), k = (int)((e >> 8) & 255u);
        s_rc[i] = (uint16_t)(l == 0 ? k : 128 + k * (kc2 - 1) + (l - 1));
    }
    for (int i = t; i < NSP; i += NT) s_out[i] = i < ns ? items[T.item_off + i].out : nullptr;
    for (int i = t; i < NW * NSP; i += NT) (&s_n[0][0])[i] = 0.0;
    __syncthreads();
    const _Float16* At = A + T.a_off;
    const uint64_t C = N / nparts;
    const uint64_t base = (uint64_t)part * C;
    const float osc = T.oscale;
    const uint32_t s_a_lds = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) _Float16*)s_a;
    const uint32_t s_b_lds = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) float*)&s_b[w][0][0];
    // the tile's row bases in registers (the copies' memory clobbers would
    // otherwise reload them)
    const float* const rW = T.W;
    const float* const rV = T.V;
    const float* const rP = T.plus;
    auto row_ptr = [&](uint32_t code) -> const float* {
        const bool isw = code >= 128u;
        const float* const rb = isw ? rW : (code ? rV : rP);
        const uint32_t idx = isw ? code - 128u : (code ? code - 1u : 0u);
        return rb + (uint64_t)idx * N;
    };
    // block b of A into buffer bf: slot s (1 KiB) by wave s, so this wave
    // issues ca (0 or 1) instructions per block
    const int ca = w < SPB ? 1 : 0;
    auto copy_a = [&](int bf, int b) {
        if (ca) { // uniform per wave
            const int4* src = reinterpret_cast<const int4*>(At + (size_t)b * BLK) + w * 64 + lane;
            lds_copy16(src, __builtin_amdgcn_readfirstlane(s_a_lds + (uint32_t)(bf * BLK * 2 + w * 1024)));
        }
    };
    // rows [q, q + 16) of this wave's 32 amplitudes from aw into its slot sl:
    // two instructions of 8 rows, 8 lanes (128 B) per row
    auto copy_b = [&](int sl, int q, uint64_t aw) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const float* src = row_ptr(s_rc[q + 8 * h + (lane >> 3)]) + aw + (uint64_t)((lane & 7) * 4);
            lds_copy16(src, __builtin_amdgcn_readfirstlane(s_b_lds + (uint32_t)(sl * 2048 + h * 1024)));
        }
    };
    for (uint64_t a0 = 0; a0 < C; a0 += NW * 32) {
        const uint64_t aw = base + a0 + (uint64_t)(w * 32);
        const uint64_t a = aw + (uint64_t)col;
#pragma unroll 1
        for (int cls = 0; cls < 2; cls++) {
            const int qb = cls ? q0 : 0, qe = cls ? nq : q0;
            const int nb = (qe - qb) / kC2RowPad;
            mqw_f32x16 acc[MB];
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[mb][r] = 0.0f;
            if (nb > 0) {
                const int bb = qb / kC2RowPad;
                // This wave's copies, in issue order: A(0), A(1), B(0), B(1),
                // B(2), then per block b: A(b + 2), B(b + 3), each only if
                // that block exists (A: ca instructions, B: two).  So A(b) is
                // issued after B(b) for b >= 2, and at block b the copies
                // issued after both are B(b + 1), A(b + 1), B(b + 2) (only
                // B(1), B(2) at block 0): that many left in flight means A(b)
                // and B(b) are in LDS (vector loads complete in order).
                // First, every store of the previous class is complete and
                // every wave is done with the buffers.
                __builtin_amdgcn_sched_barrier(0);
                wait_vm_lgkm0<0>();
                __builtin_amdgcn_s_barrier();
                asm volatile("" ::: "memory");
                __builtin_amdgcn_sched_barrier(0);
                copy_a(0, bb);
                if (nb > 1) copy_a(1, bb + 1);
                copy_b(0, qb, aw);
                if (nb > 1) copy_b(1, qb + 16, aw);
                if (nb > 2) copy_b(2, qb + 32, aw);
#pragma unroll 1
                for (int b = 0; b < nb; b++) {
                    // after the barrier A(b) of every wave is in LDS, and
                    // every wave's reads of A(b - 1) are complete, so its
                    // buffer takes A(b + 2)
                    const int left = b == 0 ? 2 * (nb > 1) + 2 * (nb > 2)
                                            : (2 + ca) * (b + 1 < nb) + 2 * (b + 2 < nb);
                    __builtin_amdgcn_sched_barrier(0);
                    wait_vm_lgkm0(left);
                    __builtin_amdgcn_s_barrier();
                    asm volatile("" ::: "memory");
                    __builtin_amdgcn_sched_barrier(0);
                    const int sl = b % 3;
                    if (b + 2 < nb) copy_a((b + 2) % 3, bb + b + 2);
                    const float* sb = &s_b[w][sl][0];
                    mqw_h8 bh, bl;
#pragma unroll
                    for (int i = 0; i < 8; i++) {
                        const float xs = sb[(8 * half + i) * 32 + col] * kWScale;
                        const _Float16 h = (_Float16)xs;
                        bh[i] = h;
                        bl[i] = (_Float16)(xs - (float)h);
                    }
                    // the slot's reads above complete before the copy
                    if (b + 3 < nb) copy_b(sl, qb + (b + 3) * 16, aw);
                    const _Float16* Ab = s_a + sl * BLK;
                    // two sample tiles at a time, their products alternating, so
                    // no product waits for the one before it on the same tile
                    // (each tile's sums keep their order)
                    constexpr int MS = MB >= 2 ? 2 : 1;
#pragma unroll
                    for (int m0 = 0; m0 < MB; m0 += MS) {
                        mqw_h8 ah[MS], al[MS];
#pragma unroll
                        for (int u = 0; u < MS; u++) {
                            const int j = (m0 + u) * 32 + col;
                            ah[u] = *reinterpret_cast<const mqw_h8*>(Ab + (half * NSP + j) * 8);
                            al[u] = *reinterpret_cast<const mqw_h8*>(Ab + NSP * 16 + (half * NSP + j) * 8);
                        }
#pragma unroll
                        for (int u = 0; u < MS; u++)
                            acc[m0 + u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[u], bh, acc[m0 + u], 0, 0, 0);
#pragma unroll
                        for (int u = 0; u < MS; u++)
                            acc[m0 + u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[u], bl, acc[m0 + u], 0, 0, 0);
#pragma unroll
                        for (int u = 0; u < MS; u++)
                            acc[m0 + u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[u], bh, acc[m0 + u], 0, 0, 0);
                    }
                }
            }
            // lane (col, half) holds rows 8 (r / 4) + 4 half + r % 4 of column col;
            // the squares q = 16 mb + r summed over the half's 32 lanes by a
            // reduce-scatter (lane col ends with q = K col + k, K = MB / 2; for
            // MB = 1 with q = col / 2 on lanes col and col ^ 1)
            float sq[16 * MB];
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int j = mb * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
                    const float v = acc[mb][r] * osc;
                    if (j < ns) {
                        MQW_GLOBAL float* o = gview(s_out[j]);
                        o[(uint64_t)cls * N + a] = v;
                    }
                    sq[16 * mb + r] = v * v;
                }
            norm_reduce_step<16 * MB, 16>(sq, col);
            constexpr int KQ = MB >= 2 ? MB / 2 : 1;
#pragma unroll
            for (int k = 0; k < KQ; k++) {
                const int q = MB >= 2 ? KQ * col + k : col >> 1;
                const int mb = q >> 4, r = q & 15;
                const int j = mb * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
                if (MB >= 2 || (col & 1) == 0) s_n[w][j] += (double)sq[k];
            }
        }
    }
    __syncthreads();
    if (t < ns) {
        double sum = 0.0;
        for (int q = 0; q < NW; q++) sum += s_n[q][t];
        items[T.item_off + t].norm_part[part] = sum;
    }
}

int launch_combine2h(const Combine2Tile* tiles_dev, uint32_t ntiles, const Combine2Item* items_dev,
                     const float* coefs_dev, uint16_t* A_dev, int n, uint32_t nparts, uint32_t ns_max, void* stream)
{
    if (ntiles == 0) return 0;
    const uint64_t N = 1ull << n;
    if (nparts == 0 || N % nparts != 0 || (N / nparts) % 512 != 0) return -2;
    if (ns_max < 1 || ns_max > (uint32_t)kC2TileH || !A_dev) return -2;
    const uint64_t grid = (uint64_t)ntiles * nparts;
    if (grid > 0x7fffffffull || ntiles > 65535u) return -3;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nsp = combine2h_nsp(ns_max);
    _Float16* A = reinterpret_cast<_Float16*>(A_dev);
    hipLaunchKernelGGL(combine2_prep_kernel, dim3(64, ntiles), dim3(256), 0, st, tiles_dev, items_dev, coefs_dev, A,
                       nsp);
    // 16 waves, 512 amplitudes per pass (one workgroup per CU: 145 KB of LDS)
    const dim3 g((uint32_t)grid), b(1024);
    const _Float16* Ac = A;
    if (nsp == 32)
        hipLaunchKernelGGL((combine2h_kernel<1, 16>), g, b, 0, st, tiles_dev, ntiles, items_dev, Ac, n, nparts);
    else if (nsp == 64)
        hipLaunchKernelGGL((combine2h_kernel<2, 16>), g, b, 0, st, tiles_dev, ntiles, items_dev, Ac, n, nparts);
    else
        hipLaunchKernelGGL((combine2h_kernel<4, 16>), g, b, 0, st, tiles_dev, ntiles, items_dev, Ac, n, nparts);
    return check_launch("combine2h");
}

// ===========================================================================
// Adjoint readout of the last stage (see ReadItem).  A workgroup takes one
// chunk of N / nchunk amplitudes for a tile of up to kReadTile samples of one
// instance.  Every thread keeps the KR basis values X_k[a] of its amplitude
// in registers, forms phi_j[a] = sum_k a_jk X_k[a] with the sample's
// coefficients from LDS (broadcast), and adds phi_j[a] psi_j[a]; the sum over
// a wave is a cross-lane tree, and the waves' sums go to FP64 in LDS, in a
// fixed order.  So the basis is read once per tile and each state once.
// ===========================================================================
__global__ void __launch_bounds__(256) fill_kernel(float* __restrict__ x, uint64_t N, float v, uint64_t k, float vk)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (uint64_t)gridDim.x * 256)
        x[i] = i == k ? vk : v;
}

int launch_fill(float* x, uint64_t N, float v, uint64_t k, float vk, void* stream)
{
    if (!x) return -2;
    const uint64_t blocks = (N + 255) / 256 < 4096 ? (N + 255) / 256 : 4096;
    hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, x, N, v, k, vk);
    return check_launch("fill");
}

template <int KR>
__global__ void __launch_bounds__(256)
readout_dot_kernel(const ReadTile* __restrict__ tiles, uint32_t ntiles, const ReadItem* __restrict__ items,
                   const float* __restrict__ coefs, int n, uint32_t nchunk, double* __restrict__ dots)
{
    __shared__ float s_c[kReadTile * KR];
    __shared__ double s_w[4][kReadTile][2];
    const uint32_t tile = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    if (tile >= ntiles) return;
    const ReadTile& T = tiles[tile];
    const int kx = (int)T.kx, ns = (int)T.ns;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = t; i < ns * KR; i += 256) {
        const int j = i / KR, k = i % KR;
        s_c[i] = k < kx ? coefs[items[T.item_off + j].coef_off + k] : 0.0f;
    }
    for (int i = t; i < 4 * kReadTile * 2; i += 256) (&s_w[0][0][0])[i] = 0.0;
    __syncthreads();
    const uint64_t N = 1ull << n;
    const uint64_t C = N / nchunk;
    const uint64_t base = (uint64_t)chunk * C;
    for (uint64_t a0 = 0; a0 < C; a0 += 256) {
        const uint64_t a = base + a0 + (uint64_t)t;
        const MQW_GLOBAL float* X = gview(T.X);
        float x[KR];
#pragma unroll
        for (int k = 0; k < KR; k++) x[k] = k < kx ? X[(uint64_t)k * N + a] : 0.0f;
        // keep the column in registers for the whole sample loop
#pragma unroll
        for (int k = 0; k < KR; k++) asm volatile("" : "+v"(x[k]));
        for (int j = 0; j < ns; j++) {
            const float* c = s_c + j * KR;
            // phi_j = sum_k a_jk X_k: even k real, odd k imaginary
            float pr = 0.0f, pi = 0.0f;
#pragma unroll
            for (int k = 0; k < KR; k += 2) {
                pr += c[k] * x[k];
                pi += c[k + 1] * x[k + 1];
            }
            const MQW_GLOBAL float* psi = gview(items[T.item_off + j].psi);
            const float yr = psi[a], yi = psi[N + a];
            float re = pr * yr - pi * yi;
            float im = pr * yi + pi * yr;
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) {
                re += __shfl_xor(re, d, 64);
                im += __shfl_xor(im, d, 64);
            }
            if (lane == 0) {
                s_w[w][j][0] += (double)re;
                s_w[w][j][1] += (double)im;
            }
        }
    }
    __syncthreads();
    if (t < 2 * ns) {
        const int j = t >> 1, q = t & 1;
        dots[((uint64_t)tile * nchunk + chunk) * kReadTile * 2 + 2 * j + q] =
            ((s_w[0][j][q] + s_w[1][j][q]) + s_w[2][j][q]) + s_w[3][j][q];
    }
}

int launch_readout_dots(const ReadTile* tiles_dev, uint32_t ntiles, const ReadItem* items_dev, const float* coefs_dev,
                        int n, uint32_t nchunk, uint32_t kx_max, double* dots, void* stream)
{
    if (ntiles == 0) return 0;
    const uint64_t N = 1ull << n;
    if (nchunk == 0 || N % nchunk != 0 || (N / nchunk) % 256 != 0) return -2;
    if (kx_max < 1 || kx_max > (uint32_t)kReadMaxKF32) return -2;
    const uint64_t grid = (uint64_t)ntiles * nchunk;
    if (grid > 0x7fffffffull) return -3;
    hipStream_t st = (hipStream_t)stream;
    if (kx_max <= 32)
        hipLaunchKernelGGL(readout_dot_kernel<32>, dim3((uint32_t)grid), dim3(256), 0, st, tiles_dev, ntiles, items_dev,
                           coefs_dev, n, nchunk, dots);
    else if (kx_max <= 64)
        hipLaunchKernelGGL(readout_dot_kernel<64>, dim3((uint32_t)grid), dim3(256), 0, st, tiles_dev, ntiles, items_dev,
                           coefs_dev, n, nchunk, dots);
    else if (kx_max <= 96)
        hipLaunchKernelGGL(readout_dot_kernel<96>, dim3((uint32_t)grid), dim3(256), 0, st, tiles_dev, ntiles, items_dev,
                           coefs_dev, n, nchunk, dots);
    else
        hipLaunchKernelGGL(readout_dot_kernel<128>, dim3((uint32_t)grid), dim3(256), 0, st, tiles_dev, ntiles,
                           items_dev, coefs_dev, n, nchunk, dots);
    return check_launch("readout_dot");
}

// The same dots on the FP16 matrix cores (operands split as in combine2h):
// phi_j[a] = sum_k a_jk X_k[a] is D[j][a] = sum_k a_jk X_k[a] with the basis
// index as the inner dimension, even k (real part of a_jk X_k) and odd k
// (imaginary part) as two classes of rows k = 2 q + class.  A workgroup takes
// 32 samples of a tile and one amplitude chunk; per pass of 256 amplitudes,
// wave w takes 32 of them.  Lane (col, half) gives the coefficients of sample
// col (A, from LDS) and rows 8 half + i of each 16-row block at amplitude col
// (B), and afterwards holds phi of samples 8 (r / 4) + 4 half + r % 4 (r < 16)
// at amplitude col, so it reads those samples' psi there: 32 lanes, one
// 128-byte row.
//
// Each lane's loads come in "chunks" of 8: the X blocks of a pass, then psi
// of 4 samples (real and imaginary parts) four times.  They stream through a
// ring of four chunks, issued three chunks before their use and in the order
// of use, so no wait for one chunk waits for a later one (vector loads
// complete in order) and 24 loads stay in flight.  The blocks are padded to a
// multiple of 4, so a pass has a multiple of 4 chunks and every pass uses the
// ring the same way.  Every load is issued (no masked lanes), so the
// compiler's waits count exactly, and every address lies in its buffer: a row
// past the basis reads the basis's row rb (below), a sample past the tile
// reads sample 0, and the last pass's look-ahead reads its own amplitudes
// again; their coefficients are zero, so they add exact zeros (every value
// read is finite).  The products phi psi of a pass are summed over
// the wave's 32 amplitudes in FP32 by a reduce-scatter across the 32 lanes of
// each half (lane col ends with the real (col even) or imaginary part of
// sample r = col / 2), added to FP64 per lane, and the 8 waves' sums are
// added in order.
// One step of a reduce-scatter over the 32 lanes of each half of a wave: the
// lane keeps the S of its 2 S values whose index has bit log2(S) equal to
// that bit of its lane (the bits above already fixed by the earlier steps),
// plus those of the lane at distance S.  After the steps 16, 8, 4, 2, 1, v[0]
// of lane col is the half's sum of the values of index col.
template <int S>
__device__ __forceinline__ void reduce_scatter_step(float (&v)[32], int col)
{
    const bool up = (col & S) != 0;
#pragma unroll
    for (int i = 0; i < S; i++) {
        const float send = up ? v[i] : v[i + S];
        const float keep = up ? v[i + S] : v[i];
        // bit-mask swizzle within 32 lanes: lane ^ S
        const int got = __builtin_amdgcn_ds_swizzle(__float_as_int(send), 0x1F | (S << 10));
        v[i] = keep + __int_as_float(got);
    }
}

template <int OCC>
__global__ void __launch_bounds__(512, OCC)
readout_mfma_kernel(const ReadTile* __restrict__ tiles, uint32_t ntiles, const ReadItem* __restrict__ items,
                    const float* __restrict__ coefs, int n, uint32_t nchunk, double* __restrict__ dots)
{
    constexpr int NSP = 32;
    constexpr int SUB = kReadTile / NSP; // workgroups per tile and chunk, NSP samples each
    constexpr int KB = kReadMaxK / 32;   // 16-row blocks per class
    constexpr int BLK = 2 * NSP * 16;    // halves of one block (hi and lo)
    constexpr int RING = 4, AHEAD = 3; // chunks of 8 loads per lane in the ring, issued ahead of use
    constexpr float kXScale = 4096.0f; // 2^12: basis entries are at most 1 in magnitude
    __shared__ __attribute__((aligned(16))) _Float16 s_b[2 * KB * BLK];
    __shared__ const float* s_psi[NSP];
    __shared__ double s_w[8][NSP][2];
    // XCD-aware order: workgroup b runs on XCD b % 8 and each XCD takes a run
    // of consecutive units, chunk-major, so the tiles of one amplitude chunk
    // (the workgroups of an instance share its basis) meet in one L2
    const uint32_t u = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const uint32_t chunk = u / (ntiles * SUB), tile = (u / SUB) % ntiles;
    if (chunk >= nchunk) return;
    const ReadTile& T = tiles[tile];
    const int j0 = (int)(u % SUB) * NSP; // this workgroup's first sample of the tile
    const int kx = (int)T.kx, ns = min((int)T.ns - j0, NSP);
    if (ns <= 0) return; // a sub-tile past the tile's samples
    const ReadItem* const TI = items + T.item_off + j0;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int col = lane & 31, half = lane >> 5;
    for (int i = t; i < 2 * KB * NSP * 16; i += 512) {
        const int rr = i & 15, j = (i >> 4) % NSP, cb = (i >> 4) / NSP; // cb = class * KB + block
        const int k = 32 * (cb % KB) + 4 * (rr & 7) + 2 * (rr >> 3) + cb / KB;
        const float v = (j < ns && k < kx) ? coefs[TI[j].coef_off + k] * T.ascale : 0.0f;
        const _Float16 hi = (_Float16)v;
        // [block][hi, lo][rows 0-7, rows 8-15][NSP samples][8 rows]: a wave's
        // lanes read consecutive 16-byte pieces (no LDS bank conflicts)
        const int pos = ((rr >> 3) * NSP + j) * 8 + (rr & 7);
        s_b[cb * BLK + pos] = hi;
        s_b[cb * BLK + NSP * 16 + pos] = (_Float16)(v - (float)hi);
    }
    for (int i = t; i < NSP; i += 512) s_psi[i] = TI[i < ns ? i : 0].psi;
    for (int i = t; i < 8 * NSP * 2; i += 512) (&s_w[0][0][0])[i] = 0.0;
    __syncthreads();
    // byte offsets below 2^32: 124 N bytes for n <= 25 (the launcher checks)
    const uint32_t N = 1u << n;
    const uint32_t C = N / nchunk;
    const uint32_t base = chunk * C + (uint32_t)(w * 32 + col); // this lane's amplitude in pass 0
    const uint32_t oh = 4u * base + (half ? 8u * N : 0u);       // and its row pair's second row
    const float* const X = T.X;
    const float osc = T.oscale;
    // 16-row blocks, the classes alternating: block c has class c & 1 and
    // index c >> 1 within it, both classes padded to the same even count (a
    // padding row reads row kx - 1), so a pass has a multiple of 4 blocks
    const int nbc = (((kx + 1) / 2 + 15) / 16 + 1) & ~1;
    const int nbt = 2 * nbc;
    // X block c (< nbt) of the pass at offset p0 into dst: load i of half h
    // reads row 32 b + 4 i + 2 h + cls of block b of class cls, at a uniform
    // base row rb (the block's first row, or the basis's last) plus the lane's
    // byte offset, so the rows need no address arithmetic per load.  Every
    // address lies in the basis: a row past it reads row rb instead.
    auto issue_x = [&](float (&dst)[8], int c, uint32_t p0) {
        const int cls = c & 1, r0 = 32 * (c >> 1) + cls;
        const int rb = r0 < kx ? r0 : kx - 1;
        const int lim = r0 < kx ? kx - r0 - 2 * half : 0; // row rb + 4 i + 2 half in the basis iff 4 i < lim
        // (a uniform base, so the loads take it from scalar registers; the
        // builtin returns int, so each half is zero-extended through uint32_t)
        const uint64_t bp = (uint64_t)(uintptr_t)(X + (uint64_t)rb * N);
        const uint64_t bu = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)bp) |
                            ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(bp >> 32)) << 32);
        const MQW_GLOBAL char* const xb = reinterpret_cast<const MQW_GLOBAL char*>(bu);
        const uint32_t oa = 4u * (base + p0);
        const uint32_t o = oh + 4u * p0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t off = 4 * i < lim ? o + 16u * N * (uint32_t)i : oa;
            dst[i] = *reinterpret_cast<const MQW_GLOBAL float*>(xb + off);
        }
    };
    // psi of samples 8 q + 4 h + e (e < 4) of the pass at offset p0: real
    // part into dst[2 e], imaginary part into dst[2 e + 1]
    auto issue_psi = [&](float (&dst)[8], int q, uint32_t p0) {
        const uint32_t oa = 4u * (base + p0);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const MQW_GLOBAL char* pr = reinterpret_cast<const MQW_GLOBAL char*>(gview(s_psi[8 * q + 4 * half + e]));
            dst[2 * e] = *reinterpret_cast<const MQW_GLOBAL float*>(pr + oa);
            dst[2 * e + 1] = *reinterpret_cast<const MQW_GLOBAL float*>(pr + (oa + 4u * N));
        }
    };
    float ring[RING][8];
    mqw_f32x16 acc[2];
    // X block gg (class q & 1 = gg & 1) from ring[q] into the products
    auto consume = [&](int q, int gg) {
        mqw_h8 xh, xl;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float xs = ring[q][i] * kXScale;
            const _Float16 h = (_Float16)xs;
            xh[i] = h;
            xl[i] = (_Float16)(xs - (float)h);
        }
        const int cls = q & 1, b = gg >> 1; // b < nbc <= KB
        const _Float16* Bb = s_b + (cls * KB + b) * BLK;
        const mqw_h8 ah = *reinterpret_cast<const mqw_h8*>(Bb + (half * NSP + col) * 8);
        const mqw_h8 al = *reinterpret_cast<const mqw_h8*>(Bb + NSP * 16 + (half * NSP + col) * 8);
        acc[cls] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, xh, acc[cls], 0, 0, 0);
        acc[cls] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, xh, acc[cls], 0, 0, 0);
        acc[cls] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, xl, acc[cls], 0, 0, 0);
    };
    // the loads stay where they are written (in the order of their use)
    auto fence = [] { asm volatile("" ::: "memory"); };
    // chunk c of a pass sits in ring[c % 4] (a pass has nbt + 4 chunks, a
    // multiple of 4) and is issued while chunk c - 3 is used
    issue_x(ring[0], 0, 0);
    issue_x(ring[1], 1, 0);
    issue_x(ring[2], 2, 0);
    fence();
    double dsum = 0.0;
    for (uint32_t p0 = 0; p0 < C; p0 += 256) {
        const uint32_t pn = p0 + 256 < C ? p0 + 256 : p0; // the next pass (the last: this one again)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[c][r] = 0.0f;
        // every group of blocks but the last looks ahead to X blocks only
        int g = 0;
        for (; g + RING < nbt; g += RING) {
#pragma unroll
            for (int q = 0; q < RING; q++) {
                fence();
                issue_x(ring[(q + AHEAD) % RING], g + q + AHEAD, p0);
                fence();
                consume(q, g + q);
            }
        }
        // the last group: X block nbt - 1, then psi of samples 0-7, 8-15, 16-23
        fence();
        issue_x(ring[3], nbt - 1, p0);
        fence();
        consume(0, g);
        issue_psi(ring[0], 0, p0);
        fence();
        consume(1, g + 1);
        issue_psi(ring[1], 1, p0);
        fence();
        consume(2, g + 2);
        issue_psi(ring[2], 2, p0);
        fence();
        consume(3, g + 3);
        // psi of samples 8 q + 4 half + e (e < 4) from ring[q] (q = 3 issued
        // here, then the next pass's first three X blocks): phi psi (complex,
        // no conjugate) as v[2 r + {0: Re, 1: Im}], r = 4 q + e
        float v[32];
#pragma unroll
        for (int q = 0; q < RING; q++) {
            fence();
            if (q == 0)
                issue_psi(ring[3], 3, p0);
            else
                issue_x(ring[q - 1], q - 1, pn);
            fence();
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = 4 * q + e;
                const float pr = acc[0][r] * osc, pi = acc[1][r] * osc;
                const float yr = ring[q][2 * e], yi = ring[q][2 * e + 1];
                v[2 * r] = pr * yr - pi * yi;
                v[2 * r + 1] = pr * yi + pi * yr;
            }
        }
        reduce_scatter_step<16>(v, col);
        reduce_scatter_step<8>(v, col);
        reduce_scatter_step<4>(v, col);
        reduce_scatter_step<2>(v, col);
        reduce_scatter_step<1>(v, col);
        dsum += (double)v[0];
    }
    {
        const int r = col >> 1;
        s_w[w][8 * (r >> 2) + 4 * half + (r & 3)][col & 1] = dsum;
    }
    __syncthreads();
    if (t < 2 * ns) {
        const int j = t >> 1, q = t & 1;
        double sum = 0.0;
        for (int v = 0; v < 8; v++) sum += s_w[v][j][q];
        dots[((uint64_t)tile * nchunk + chunk) * kReadTile * 2 + 2 * (j0 + j) + q] = sum;
    }
}

int launch_readout_dots_mfma(const ReadTile* tiles_dev, uint32_t ntiles, const ReadItem* items_dev,
                             const float* coefs_dev, int n, uint32_t nchunk, double* dots, void* stream)
{
    if (ntiles == 0) return 0;
    static_assert(kReadTile == 64 && kReadMaxK % 32 == 0, "tiles of two 32-sample blocks, whole row blocks");
    const uint64_t N = 1ull << n;
    if (nchunk == 0 || N % nchunk != 0 || (N / nchunk) % 256 != 0) return -2;
    if (n > kReadMfmaMaxN) return -2; // 32-bit byte offsets within a block of 32 basis rows
    // 32 samples per workgroup; the workgroups of a tile and chunk read its X
    // rows through one L2
    constexpr int SUB = 2;
    const uint64_t grid = ((uint64_t)ntiles * nchunk * SUB + 7) / 8 * 8; // a multiple of the 8 XCDs
    if (grid > 0x7fffffffull) return -3;
    hipLaunchKernelGGL((readout_mfma_kernel<4>), dim3((uint32_t)grid), dim3(512), 0, (hipStream_t)stream,
                       tiles_dev, ntiles, items_dev, coefs_dev, n, nchunk, dots);
    return check_launch("readout_dot_mfma");
}

// Pair dots of the two-stage adjoint readout: one workgroup per (tile, chunk,
// item), sum_a phi[a] psi[a] over the chunk (complex, no conjugate).  The
// products of FP32 values are exact in FP64 and summed in FP64 per thread,
// then by a cross-lane tree and the 4 waves in order (deterministic).  The
// chunk is a multiple of 1024 amplitudes (readout_ops sizes chunks of at
// least 4096).
__global__ void __launch_bounds__(256)
pair_dot_kernel(const ReadTile* __restrict__ tiles, uint32_t ntiles, const ReadItem* __restrict__ items, int n,
                uint32_t nchunk, double* __restrict__ dots)
{
    __shared__ double s_w[4][2];
    const uint32_t j = blockIdx.x % kReadTile;
    const uint32_t tc = blockIdx.x / kReadTile;
    const uint32_t chunk = tc % nchunk, tile = tc / nchunk;
    if (tile >= ntiles) return;
    const ReadTile& T = tiles[tile];
    if (j >= T.ns) return;
    const ReadItem& I = items[T.item_off + j];
    const uint64_t N = 1ull << n;
    const uint64_t C = N / nchunk;
    const uint64_t base = (uint64_t)chunk * C;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    typedef float f4 __attribute__((ext_vector_type(4)));
    const MQW_GLOBAL f4* pr = gview(reinterpret_cast<const f4*>(I.phi + base));
    const MQW_GLOBAL f4* pi = gview(reinterpret_cast<const f4*>(I.phi + N + base));
    const MQW_GLOBAL f4* sr = gview(reinterpret_cast<const f4*>(I.psi + base));
    const MQW_GLOBAL f4* si = gview(reinterpret_cast<const f4*>(I.psi + N + base));
    double re = 0.0, im = 0.0;
    for (uint64_t q = (uint64_t)t; q < C / 4; q += 256) {
        const f4 a = __builtin_nontemporal_load(pr + q), b = __builtin_nontemporal_load(pi + q);
        const f4 x = __builtin_nontemporal_load(sr + q), y = __builtin_nontemporal_load(si + q);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            re = fma((double)a[e], (double)x[e], re);
            re = fma(-(double)b[e], (double)y[e], re);
            im = fma((double)a[e], (double)y[e], im);
            im = fma((double)b[e], (double)x[e], im);
        }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        re += __shfl_xor(re, d, 64);
        im += __shfl_xor(im, d, 64);
    }
    if (lane == 0) {
        s_w[w][0] = re;
        s_w[w][1] = im;
    }
    __syncthreads();
    if (t < 2) dots[((uint64_t)tile * nchunk + chunk) * kReadTile * 2 + 2 * j + t] =
        ((s_w[0][t] + s_w[1][t]) + s_w[2][t]) + s_w[3][t];
}

int launch_pair_dots(const ReadTile* tiles_dev, uint32_t ntiles, const ReadItem* items_dev, int n, uint32_t nchunk,
                     double* dots, void* stream)
{
    if (ntiles == 0) return 0;
    const uint64_t N = 1ull << n;
    if (nchunk == 0 || N % nchunk != 0 || (N / nchunk) % 1024 != 0) return -2;
    const uint64_t grid = (uint64_t)ntiles * nchunk * kReadTile;
    if (grid > 0x7fffffffull) return -3;
    hipLaunchKernelGGL(pair_dot_kernel, dim3((uint32_t)grid), dim3(256), 0, (hipStream_t)stream, tiles_dev, ntiles,
                       items_dev, n, nchunk, dots);
    return check_launch("pair_dots");
}

// One wave per item: lane l sums the chunk partials l, l + 64, ... and the
// norm partials the same way, then a fixed-order tree over the lanes (the same
// order on every run).
__global__ void __launch_bounds__(256)
readout_finalize_kernel(const ReadTile* __restrict__ tiles, uint32_t ntiles, const ReadItem* __restrict__ items,
                        uint32_t nchunk, uint32_t nparts, const double* __restrict__ dots)
{
    const uint32_t tile = blockIdx.x / (kReadTile / 4);
    const uint32_t j = (blockIdx.x % (kReadTile / 4)) * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (tile >= ntiles) return;
    const ReadTile& T = tiles[tile];
    if (j >= T.ns) return; // uniform per wave
    const ReadItem& I = items[T.item_off + j];
    double re = 0.0, im = 0.0, nrm = 0.0;
    for (uint32_t c = lane; c < nchunk; c += 64) {
        const double* d = dots + ((uint64_t)tile * nchunk + c) * kReadTile * 2 + 2 * j;
        re += d[0];
        im += d[1];
    }
    for (uint32_t q = lane; q < nparts; q += 64) nrm += I.norm_part[q];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        re += __shfl_xor(re, d, 64);
        im += __shfl_xor(im, d, 64);
        nrm += __shfl_xor(nrm, d, 64);
    }
    if (lane == 0) *I.prob = (re * re + im * im) / nrm;
}

int launch_readout_finalize(const ReadTile* tiles_dev, uint32_t ntiles, const ReadItem* items_dev, uint32_t nchunk,
                            uint32_t nparts, const double* dots, void* stream)
{
    if (ntiles == 0) return 0;
    static_assert(kReadTile % 4 == 0, "four items per workgroup");
    const uint64_t grid = (uint64_t)ntiles * (kReadTile / 4);
    if (grid > 0x7fffffffull) return -3;
    hipLaunchKernelGGL(readout_finalize_kernel, dim3((uint32_t)grid), dim3(256), 0, (hipStream_t)stream, tiles_dev,
                       ntiles, items_dev, nchunk, nparts, dots);
    return check_launch("readout_finalize");
}

// ===========================================================================
// Stage reduction: norm, next-stage scale, probability readout
// ===========================================================================
template <typename real>
__global__ void __launch_bounds__(256) reduce_kernel(const ReduceJob* __restrict__ jobs)
{
    __shared__ double red[256];
    const ReduceJob& J = jobs[blockIdx.x];
    const int t = threadIdx.x;
    double s = 0.0;
    for (uint32_t k = t; k < J.ntiles; k += 256) s += J.norm_part[k];
    red[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) {
        const double tot = red[0];
        if (J.psi_scale) *J.psi_scale = 1.0 / sqrt(tot);
        if (J.prob) {
            const real* o = (const real*)J.out;
            const double re = (double)o[J.eloc], im = (double)o[J.N + J.eloc];
            *J.prob = (re * re + im * im) / tot;
        }
    }
}

int launch_reduce(int precision, const ReduceJob* jobs_dev, uint32_t njobs, void* stream)
{
    if (njobs == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (precision == MQW_FP32)
        hipLaunchKernelGGL(reduce_kernel<float>, dim3(njobs), dim3(256), 0, st, jobs_dev);
    else
        hipLaunchKernelGGL(reduce_kernel<double>, dim3(njobs), dim3(256), 0, st, jobs_dev);
    return check_launch("reduce");
}

// ===========================================================================
// One real-half Clenshaw step on caller device buffers (the function-level
// seam of MultiQW.cpp:88): b2 <- alpha (hp .* b1) - beta H_G b1 - b2 + coef psi
// ===========================================================================
__global__ void __launch_bounds__(256)
half_step_kernel(const float* __restrict__ b1, float* __restrict__ b2, const float* __restrict__ hp,
                 const float* __restrict__ psi, int n, float alpha, float beta, float coef)
{
    const uint64_t N = 1ull << n;
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= N) return;
    float acc = 0.0f;
    for (int i = 0; i < n; i++) acc += b1[k ^ (1ull << i)];
    float v = alpha * (hp[k] * b1[k]) - beta * acc - b2[k];
    if (coef != 0.0f) v += coef * psi[k];
    b2[k] = v;
}

int launch_half_step(const float* b1, float* b2, const float* hp, const float* psi, int n, float alpha, float beta,
                     float coef, void* stream)
{
    const uint64_t N = 1ull << n;
    hipLaunchKernelGGL(half_step_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, b1,
                       b2, hp, psi, n, alpha, beta, coef);
    return check_launch("half_step");
}

} // namespace mqw

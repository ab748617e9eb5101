// Small-n trajectories, one workgroup each, state resident on chip.
//
// For n <= 13 the whole walk of one trajectory (every stage, every Clenshaw
// term, the renormalisations and the readouts of MultiQW.cpp:432-451) runs in
// one workgroup.  Each thread owns R = 2^RB consecutive amplitudes and keeps
// b_{k+1}, b_{k+2}, psi and the diagonal E for them in registers; only b_{k+1}
// is published to LDS (double-buffered, one barrier per term) so the other
// threads can read its flips along the bits RB..n-1.  Global memory is touched
// for E once per trajectory, for the stage coefficients, and for the readouts.
// Chunk c of thread t holds elements c*(4 NT) + 4t + [0, 4) ("plane layout"),
// so every wave-wide LDS or global access is contiguous.
//
// Term recurrence (MultiQW.cpp:88-173,177-236), complex, planar per thread:
//   b_{K-1} = c_{K-1} i^{(K-1)&1} psi
//   b_r     = a_r (E b_{r+1} - gamma H_G b_{r+1}) - b_{r+2} + c_r i^{r&1} psi
// with a_r = 2/scale (r > 0) and 1/scale (r = 0); the result b_0 is the new
// state, renormalised in FP64 (the scale is folded into the next stage's
// coefficients rather than written back).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "mqw_internal.h"

namespace mqw {

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

template <int NB>
struct LdsShape {
    static constexpr int RB = NB <= 10 ? 2 : 3;
    static constexpr int R = 1 << RB;
    static constexpr int NT = 1 << (NB - RB);
    static constexpr int NRED = NT >= 64 ? NT / 64 : NT;
};

template <int NB>
__global__ void __launch_bounds__(LdsShape<NB>::NT)
lds_kernel(const LdsTraj* __restrict__ trajs, uint32_t ntraj, const LdsSeg* __restrict__ segs,
           const float* __restrict__ coefs, double* __restrict__ probs)
{
    using S = LdsShape<NB>;
    constexpr int RB = S::RB, R = S::R, NT = S::NT, NRED = S::NRED;
    constexpr int N = 1 << NB;
    // plane layout: chunk c of thread t holds elements c*(4 NT) + 4t + [0, 4)
    constexpr int PS = 4 * NT;
    __shared__ float lre[2][N];
    __shared__ float lim[2][N];
    __shared__ double red[NRED];
    __shared__ double bcast;

    const int t = threadIdx.x;
    const float psi0 = (float)(1.0 / sqrt((double)N));

    for (uint32_t tr = blockIdx.x; tr < ntraj; tr += gridDim.x) {
        const LdsTraj T = trajs[tr];
        float e[R];
#pragma unroll
        for (int q = 0; q < R; q += 4) {
            const float4 v = *reinterpret_cast<const float4*>(T.E + (q / 4) * PS + t * 4);
            e[q] = v.x;
            e[q + 1] = v.y;
            e[q + 2] = v.z;
            e[q + 3] = v.w;
        }
        float pr[R], pim[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            pr[q] = psi0;
            pim[q] = 0.0f;
        }
        float ps = 1.0f; // 1/||psi|| of the current stage input
        int buf = 0;
        for (uint32_t s = 0; s < T.nseg; s++) {
            const LdsSeg G = segs[T.seg_off + s];
            const float* c = coefs + G.coef_off;
            const int K = (int)G.K;
            float xr[R], xi[R], yr[R], yi[R];
            {
                const float top = c[K - 1] * ps;
                const bool odd = ((K - 1) & 1) != 0;
#pragma unroll
                for (int q = 0; q < R; q++) {
                    xr[q] = odd ? -top * pim[q] : top * pr[q];
                    xi[q] = odd ? top * pr[q] : top * pim[q];
                    yr[q] = 0.0f;
                    yi[q] = 0.0f;
                }
            }
            for (int r = K - 2; r >= 0; r--) {
                float* wr = lre[buf];
                float* wi = lim[buf];
#pragma unroll
                for (int q = 0; q < R; q += 4) {
                    const int a = (q / 4) * PS + t * 4;
                    *reinterpret_cast<float4*>(wr + a) = make_float4(xr[q], xr[q + 1], xr[q + 2], xr[q + 3]);
                    *reinterpret_cast<float4*>(wi + a) = make_float4(xi[q], xi[q + 1], xi[q + 2], xi[q + 3]);
                }
                __syncthreads();
                float ar[R], ai[R];
#pragma unroll
                for (int q = 0; q < R; q++) {
                    ar[q] = 0.0f;
                    ai[q] = 0.0f;
                }
#pragma unroll
                for (int b = 0; b < RB; b++)
#pragma unroll
                    for (int q = 0; q < R; q++) {
                        ar[q] += xr[q ^ (1 << b)];
                        ai[q] += xi[q ^ (1 << b)];
                    }
#pragma unroll
                for (int b = 0; b < NB - RB; b++) {
                    const int pt = t ^ (1 << b);
#pragma unroll
                    for (int q = 0; q < R; q += 4) {
                        const int a = (q / 4) * PS + pt * 4;
                        const float4 u = *reinterpret_cast<const float4*>(wr + a);
                        const float4 v = *reinterpret_cast<const float4*>(wi + a);
                        ar[q] += u.x;
                        ar[q + 1] += u.y;
                        ar[q + 2] += u.z;
                        ar[q + 3] += u.w;
                        ai[q] += v.x;
                        ai[q + 1] += v.y;
                        ai[q + 2] += v.z;
                        ai[q + 3] += v.w;
                    }
                }
                const float a = r == 0 ? 0.5f * G.alpha : G.alpha;
                const float bb = r == 0 ? 0.5f * G.beta : G.beta;
                const float cc = c[r] * ps;
                const bool odd = (r & 1) != 0;
                const float cr = odd ? 0.0f : cc, ci = odd ? cc : 0.0f;
#pragma unroll
                for (int q = 0; q < R; q++) {
                    const float zr = a * e[q] * xr[q] - bb * ar[q] - yr[q] + (cr * pr[q] - ci * pim[q]);
                    const float zi = a * e[q] * xi[q] - bb * ai[q] - yi[q] + (cr * pim[q] + ci * pr[q]);
                    yr[q] = xr[q];
                    yi[q] = xi[q];
                    xr[q] = zr;
                    xi[q] = zi;
                }
                buf ^= 1;
            }
            if (K == 1) {
                const float c0 = c[0] * ps;
#pragma unroll
                for (int q = 0; q < R; q++) {
                    xr[q] = c0 * pr[q];
                    xi[q] = c0 * pim[q];
                }
            }
            // FP64 norm of the stage output
            double nrm = 0.0;
#pragma unroll
            for (int q = 0; q < R; q++) nrm += (double)xr[q] * (double)xr[q] + (double)xi[q] * (double)xi[q];
            const double tsum = wg_sum_t0<NT>(nrm, red);
            if (t == 0) bcast = tsum;
            __syncthreads();
            const double tot = bcast;
            if (G.prob_slot >= 0) {
                // owner of eloc under the plane layout: element w of chunk c of thread tt
                const uint32_t w = T.eloc & 3u, tt = (T.eloc >> 2) & (uint32_t)(NT - 1);
                const uint32_t own = (T.eloc / (uint32_t)PS) * 4u + w;
                if (tt == (uint32_t)t) {
                    double re = 0.0, im = 0.0;
#pragma unroll
                    for (int q = 0; q < R; q++)
                        if ((uint32_t)q == own) {
                            re = xr[q];
                            im = xi[q];
                        }
                    probs[G.prob_slot] = (re * re + im * im) / tot;
                }
            }
            ps = (float)(1.0 / sqrt(tot));
#pragma unroll
            for (int q = 0; q < R; q++) {
                pr[q] = xr[q];
                pim[q] = xi[q];
            }
        }
        __syncthreads(); // LDS and `red` are reused by the next trajectory
    }
}

int lds_supported(int n) { return n >= 4 && n <= 13; }

template <int NB>
static int launch_nb(const LdsTraj* trajs, uint32_t ntraj, const LdsSeg* segs, const float* coefs, double* probs,
                     int grid, hipStream_t st)
{
    hipLaunchKernelGGL(lds_kernel<NB>, dim3(grid), dim3(LdsShape<NB>::NT), 0, st, trajs, ntraj, segs, coefs, probs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NB>
static int grid_nb(int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lds_kernel<NB>, LdsShape<NB>::NT, 0) != hipSuccess)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 256;
    return (per_cu > 0 ? per_cu : 1) * cus;
}

int lds_grid(int n, int device)
{
    switch (n) {
    case 4: return grid_nb<4>(device);
    case 5: return grid_nb<5>(device);
    case 6: return grid_nb<6>(device);
    case 7: return grid_nb<7>(device);
    case 8: return grid_nb<8>(device);
    case 9: return grid_nb<9>(device);
    case 10: return grid_nb<10>(device);
    case 11: return grid_nb<11>(device);
    case 12: return grid_nb<12>(device);
    case 13: return grid_nb<13>(device);
    }
    return 0;
}

int launch_lds(int n, const LdsTraj* trajs, uint32_t ntraj, const LdsSeg* segs, const float* coefs, double* probs,
               uint32_t* counter, int grid, void* stream)
{
    (void)counter;
    hipStream_t st = (hipStream_t)stream;
    if (ntraj == 0) return 0;
    if (grid > (int)ntraj) grid = (int)ntraj;
    switch (n) {
    case 4: return launch_nb<4>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 5: return launch_nb<5>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 6: return launch_nb<6>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 7: return launch_nb<7>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 8: return launch_nb<8>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 9: return launch_nb<9>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 10: return launch_nb<10>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 11: return launch_nb<11>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 12: return launch_nb<12>(trajs, ntraj, segs, coefs, probs, grid, st);
    case 13: return launch_nb<13>(trajs, ntraj, segs, coefs, probs, grid, st);
    }
    return -2;
}

} // namespace mqw

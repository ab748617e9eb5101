// Basis recurrences with the state on chip (n = 15..20).
//
// A chain s_{k+1} = alpha E s_k - beta M s_k - s_{k-1} (M = sum_i X_i, E the
// diagonal; step 0: s_1 = alpha0 E s_0 - beta0 M s_0) is run from its first
// step to its last by one group of G = 2^(n-15) workgroups, one per CU, that
// all sit on one XCD.  Each workgroup (512 threads of 64 amplitudes) holds
// 2^15 amplitudes: alpha E in registers for the whole chain, s_k in LDS from
// the start of step k to the start of step k + 1 (the in-workgroup flips, and
// s_{k-1} for the next step), and the sums in registers.  The partners of the
// top n - 15 bits are the other members' outputs, which the XCD's L2 still
// holds (4 MiB at n = 20).  Per step and amplitude HBM sees the output written
// and nothing read: 4 bytes, against 16 for a launch per step (s_k, s_{k-1},
// E read, s_{k+1} written) and the L2 misses of its partner tiles.  (At n = 20
// a step's s_k and s_{k+1} are twice the L2, and partners that miss it add
// about 5 bytes of reads.)
//
// Workgroup b runs on XCD b % 8 (the round-robin dispatch of a 256-CU part):
// the 32 of one XCD form 32 / G groups of consecutive slots.  Every member
// signals a group counter after its step's stores have completed (they are in
// the XCD's L2) and waits for all G signals before it reads partners; the
// waits are bounded.  A wait that runs out, or a group whose members report
// different XCD ids, sets the launch's abort flag (every member then runs out
// its steps without waiting, all accesses inside its regions) and the error
// word the host reads after the plan; the host then discards the outputs and
// runs the recurrences one launch per step.  The arithmetic and its
// order are the per-step kernel's with 2^14-amplitude tiles (half_term_kernel
// mode 3): the outputs are bit-identical to it.
#include <hip/hip_runtime.h>

#include "mqw_internal.h"

using namespace mqw;

namespace {

#define MQW_GLOBAL __attribute__((address_space(1)))
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 512;
constexpr int kTBits = 9;               // thread index bits: amplitude bits 2 .. 10
constexpr int kWgBits = 15;             // amplitudes per workgroup, log2
constexpr int kPer = 1 << (kWgBits - kTBits); // per thread: 64
constexpr int kPlane = 4 << kTBits;     // element c * 2048 + 4 t + i of thread t is its value 4 c + i
constexpr int kHalf = kPer / 8;          // float4 loads of half a thread's values
constexpr int kParked = 4;               // planes of s_{k-1} kept in LDS beside the exchange buffer
constexpr int kKept = kPer / 4 - kParked; // planes of s_{k-1} kept in registers
constexpr uint32_t kSpinMax = 1u << 21; // default polls of a wait (each sleeps about 64 clocks and reads L2)
constexpr uint32_t kSyncStride = 32;    // words between groups' counters (128 bytes)

// Global accesses are buffer loads and stores on one workgroup's region of
// 2^15 floats (a uniform descriptor; the lane's byte offset in one register,
// the plane's in a scalar one).  The hardware checks every access against the
// region: out of it, a load returns 0 and a store is dropped.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr int kRsrcWord3 = 0x00020000; // gfx9 raw buffer: 32-bit data, no format conversion
constexpr int kNonTemporal = 2;        // cache policy bit of a streamed access
__device__ __forceinline__ rsrc_t region(const float* base)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)(4u << kWgBits), kRsrcWord3);
}

// Planes [C0, C0 + NC) of a thread's values (NC float4 loads in flight,
// static indices): v[.] = p[.] (MODE 0), v[.] += p[.] (1), v[.] -= a * p[.]
// (2).  Global: the region r at the lane's byte offset boff (NT: streamed).
// LDS: base is the lane's own address.  The callers put scheduling barriers
// between parts, so a later part's loads are not hoisted above an earlier
// part's uses.
template <int C0, int NC, int MODE, bool NT>
__device__ __forceinline__ void part_glob(float (&v)[kPer], rsrc_t r, uint32_t boff, float a)
{
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int cc = C0 + c;
        const f4 w = __builtin_amdgcn_raw_buffer_load_b128(r, boff, cc * kPlane * 4, NT ? kNonTemporal : 0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (MODE == 0) v[4 * cc + i] = w[i];
            if (MODE == 1) v[4 * cc + i] += w[i];
            if (MODE == 2) v[4 * cc + i] -= a * w[i];
        }
    }
}
template <int C0, int NC>
__device__ __forceinline__ void part_lds_add(float (&v)[kPer], const float* base)
{
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int cc = C0 + c;
        const f4 w = *reinterpret_cast<const f4*>(base + cc * kPlane);
#pragma unroll
        for (int i = 0; i < 4; i++) v[4 * cc + i] += w[i];
    }
}
template <int MODE, bool NT>
__device__ __forceinline__ void glob_halves(float (&v)[kPer], rsrc_t r, uint32_t boff, float a = 0.0f)
{
    part_glob<0, kHalf, MODE, NT>(v, r, boff, a);
    __builtin_amdgcn_sched_barrier(0);
    part_glob<kHalf, kHalf, MODE, NT>(v, r, boff, a);
    __builtin_amdgcn_sched_barrier(0);
}
// The same in quarters (16 registers in flight), where registers are short.
template <int MODE, bool NT>
__device__ __forceinline__ void glob_quarters(float (&v)[kPer], rsrc_t r, uint32_t boff, float a = 0.0f)
{
    constexpr int Q = kPer / 16;
    part_glob<0, Q, MODE, NT>(v, r, boff, a);
    __builtin_amdgcn_sched_barrier(0);
    part_glob<Q, Q, MODE, NT>(v, r, boff, a);
    __builtin_amdgcn_sched_barrier(0);
    part_glob<2 * Q, Q, MODE, NT>(v, r, boff, a);
    __builtin_amdgcn_sched_barrier(0);
    part_glob<3 * Q, Q, MODE, NT>(v, r, boff, a);
    __builtin_amdgcn_sched_barrier(0);
}
__device__ __forceinline__ void lds_quarters_add(float (&v)[kPer], const float* base)
{
    constexpr int Q = kPer / 16;
    part_lds_add<0, Q>(v, base);
    __builtin_amdgcn_sched_barrier(0);
    part_lds_add<Q, Q>(v, base);
    __builtin_amdgcn_sched_barrier(0);
    part_lds_add<2 * Q, Q>(v, base);
    __builtin_amdgcn_sched_barrier(0);
    part_lds_add<3 * Q, Q>(v, base);
    __builtin_amdgcn_sched_barrier(0);
}

// Amplitude bits 11 then 14 (planes c ^ 1, c ^ 8) for the four planes with
// bits 1, 2 equal to B (amplitude bits 12, 13), from this thread's own s_k in
// LDS at own: 16 values in registers.
template <int B>
__device__ __forceinline__ void own_bits(float (&acc)[kPer], const float* own)
{
    float x[16]; // plane c(m) = (m & 1) | B << 1 | (m >> 1) << 3 at 4 m
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int c = (m & 1) | (B << 1) | ((m >> 1) << 3);
        const f4 w = *reinterpret_cast<const f4*>(own + c * kPlane);
#pragma unroll
        for (int i = 0; i < 4; i++) x[4 * m + i] = w[i];
    }
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int c = (m & 1) | (B << 1) | ((m >> 1) << 3);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            acc[4 * c + i] += x[4 * (m ^ 1) + i]; // plane c ^ 1
            acc[4 * c + i] += x[4 * (m ^ 2) + i]; // plane c ^ 8
        }
    }
}

// Half H (planes [8 H, 8 H + 8)) of the step's end, in place in acc:
// a1 (ae s_k - be sum) with ae = (alpha E) hf = al E, the per-step kernel's
// expression and values, then - a2 s_{k-1}: planes below kKept from s2, the
// others from this thread's parked planes in LDS (park).
template <int H>
__device__ __forceinline__ void own_final(float (&acc)[kPer], const float (&e)[kPer], const float (&s2)[4 * kKept],
                                          const float* own, const float* park, float a1, float hf, float be,
                                          float a2)
{
    float x[4 * kHalf];
#pragma unroll
    for (int c = 0; c < kHalf; c++) {
        const f4 w = *reinterpret_cast<const f4*>(own + (H * kHalf + c) * kPlane);
#pragma unroll
        for (int i = 0; i < 4; i++) x[4 * c + i] = w[i];
    }
#pragma unroll
    for (int c = 0; c < kHalf; c++) {
        const int cc = H * kHalf + c;
        f4 y;
        if (cc >= kKept) y = *reinterpret_cast<const f4*>(park + (cc - kKept) * kPlane);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int q = 4 * cc;
            float o = a1 * (e[q + i] * hf * x[4 * c + i] - be * acc[q + i]);
            o -= a2 * (cc < kKept ? s2[q + i] : y[i]);
            acc[q + i] = o;
        }
    }
}

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void abort_launch(uint32_t* sync, uint32_t* err, uint32_t code)
{
    __hip_atomic_store(sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(err, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// This workgroup's stores (and atomics) have completed, then one signal.
__device__ __forceinline__ void group_signal(uint32_t* cnt)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Until the group counter reaches target, or the launch aborts (then every
// later wait returns at once: the workgroup runs out its steps without
// waiting, every access inside its regions, and the host discards the
// launch's outputs), or `polls` reads found it short (then this wait aborts
// the launch; polls = 0 gives up before the first read).  Every wave polls for itself (one request per wave) with
// uniform values, so the loop is no divergent region inside the kernel's
// large live ranges, and no exit leaves the step loop early.
__device__ __forceinline__ void group_wait(uint32_t* sync, uint32_t* err, const uint32_t* cnt, uint32_t target,
                                           uint32_t polls)
{
    for (uint32_t spins = 0;; spins++) {
        if (spins >= polls) {
            abort_launch(sync, err, 1u);
            return;
        }
        if (__builtin_amdgcn_readfirstlane(ld_agent(cnt)) >= target) return;
        if (__builtin_amdgcn_readfirstlane(ld_agent(sync)) != 0u) return;
        __builtin_amdgcn_s_sleep(1);
    }
}

// The in-thread flips of amplitude bits 0, 1, 12, 13 for the planes c of
// parity P (amplitude bit 11 = P): acc[q] = ((v[q^1] + v[q^2]) + v[q^8]) +
// v[q^16] from 0, with v this thread's s_k read from LDS at own (all operands
// have the same parity).
template <int P>
__device__ __forceinline__ void in_thread(float (&acc)[kPer], const float* own)
{
    float x[kPer / 2]; // planes 2 m + P at 4 m
#pragma unroll
    for (int m = 0; m < kPer / 8; m++) {
        const f4 w = *reinterpret_cast<const f4*>(own + (2 * m + P) * kPlane);
#pragma unroll
        for (int i = 0; i < 4; i++) x[4 * m + i] = w[i];
    }
#pragma unroll
    for (int m = 0; m < kPer / 8; m++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int q = 4 * (2 * m + P) + i, r = 4 * m + i;
            // q ^ 1, q ^ 2: r ^ 1, r ^ 2; q ^ 8, 16 (planes c ^ 2, 4): r ^ 4, 8
            acc[q] = 0.0f;
            acc[q] += x[r ^ 1];
            acc[q] += x[r ^ 2];
            acc[q] += x[r ^ 4];
            acc[q] += x[r ^ 8];
        }
}

// One step's output: plane c of the lane at base + c * kPlane + 4 t (cached
// stores: the other members read them from L2 as the next step's partners).
__device__ __forceinline__ void store_step(const float (&v)[kPer], float* base, uint32_t boff)
{
    const rsrc_t r = region(base);
#pragma unroll
    for (int c = 0; c < kPer / 4; c++) {
        f4 w;
#pragma unroll
        for (int i = 0; i < 4; i++) w[i] = v[4 * c + i];
        __builtin_amdgcn_raw_buffer_store_b128(w, r, boff, c * kPlane * 4, 0);
    }
}

// The end of a step: the group's signal (which also orders this step's LDS
// reads before the next step's writes), or a workgroup barrier alone.
template <int G>
__device__ __forceinline__ void step_done(uint32_t* cnt, uint32_t& gen)
{
    if constexpr (G > 1) {
        group_signal(cnt);
        gen++;
    } else {
        __syncthreads();
    }
}

template <int GB>
__global__ void __launch_bounds__(kThreads)
chain_kernel(const ChainJob* __restrict__ jobs, const uint32_t* __restrict__ order, const uint32_t* __restrict__ goff,
             uint32_t* __restrict__ sync, uint32_t* __restrict__ err, uint32_t polls)
{
    constexpr int G = 1 << GB;
    constexpr int n = kWgBits + GB;
    __shared__ __attribute__((aligned(16))) float s_x[1 << kWgBits];
    __shared__ __attribute__((aligned(16))) float s_park[kParked * kPlane]; // 32 KiB: 160 KiB in all
    const uint32_t b = blockIdx.x, xcd = b & 7u, slot = b >> 3;
    const uint32_t grp = xcd * (32u >> GB) + (slot >> GB), mem = slot & (G - 1);
    uint32_t* cnt = sync + kSyncStride * (1u + grp);
    const int t = threadIdx.x;
    const uint64_t N = 1ull << n;
    const uint64_t mbase = (uint64_t)mem << kWgBits;
    const uint32_t boff = 16u * (uint32_t)t; // the lane's byte offset within a plane
    uint32_t gen = 0; // signals of every member so far
    if constexpr (G > 1) {
        // every member on one XCD: the XCD ids of the group, one bit each
        if (t == 0) {
            uint32_t id;
            asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(id));
            __hip_atomic_fetch_or(cnt + 1, 1u << (id & 15u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        group_signal(cnt);
        gen++;
        group_wait(sync, err, cnt, G * gen, polls);
        if (__builtin_amdgcn_readfirstlane(ld_agent(sync)) != 0u) return;
        if (__popc(__builtin_amdgcn_readfirstlane(ld_agent(cnt + 1))) != 1) {
            abort_launch(sync, err, 2u);
            return;
        }
    }
    for (uint32_t ci = goff[grp]; ci < goff[grp + 1]; ci++) {
        const ChainJob J = jobs[order[ci]];
        const bool cst = J.seed == nullptr;
        // e: alpha E for the whole chain (step 0 takes e / 2 = alpha0 E exactly:
        // the host checks alpha = 2 alpha0, and halving commutes with rounding).
        // v: s_k at the start of a step, the step's output at its end.  The LDS
        // buffer holds s_k from the start of step k to the start of step k + 1,
        // where this thread reads its own s_{k-1} back from it before storing
        // s_{k+1}: HBM sees the outputs written and nothing read per step.
        float e[kPer], v[kPer];
        glob_halves<0, true>(e, region(J.E + mbase), boff);
#pragma unroll
        for (int q = 0; q < kPer; q++) e[q] *= J.alpha;
        uint32_t k = 0;
        if (cst) {
            // s_1 = a1 (alpha0 E psi0 - beta0 n psi0): |+> needs no flips
            const float base = J.beta0 * (float)n * J.psi0;
#pragma unroll
            for (int q = 0; q < kPer; q++) v[q] = J.a1 * (e[q] * 0.5f * J.psi0 - base);
            store_step(v, J.out + mbase, boff);
            step_done<G>(cnt, gen);
            k = 1;
        } else {
            glob_halves<0, false>(v, region(J.seed + mbase), boff);
        }
        float* const own = s_x + 4 * t;
        float* const park = s_park + 4 * t;
        for (; k < J.steps; k++) {
            const float hf = k == 0 ? 0.5f : 1.0f, be = k == 0 ? J.beta0 : J.beta;
            const float* src = k == 0 ? J.seed : J.out + (uint64_t)(k - 1) * N;
            // s_{k-1}: this thread's values in LDS since the previous step (the
            // seed at k = 1; |+> is the constant psi0), planes below kKept to
            // registers, the others parked in LDS beside the exchange buffer.
            // At k = 0 there is none: subtracting a2 * 0 leaves every value
            // (signed zeros too) as it is.
            float s2[4 * kKept];
            if (k > 0 && !(k == 1 && cst)) {
#pragma unroll
                for (int c = 0; c < kPer / 4; c++) {
                    const f4 w = *reinterpret_cast<const f4*>(own + c * kPlane);
                    if (c < kKept) {
#pragma unroll
                        for (int i = 0; i < 4; i++) s2[4 * c + i] = w[i];
                    } else {
                        *reinterpret_cast<f4*>(park + (c - kKept) * kPlane) = w;
                    }
                }
            } else {
                const float c2 = k == 1 ? J.psi0 : 0.0f;
#pragma unroll
                for (int q = 0; q < 4 * kKept; q++) s2[q] = c2;
                const f4 w = {c2, c2, c2, c2};
#pragma unroll
                for (int c = 0; c < kParked; c++) *reinterpret_cast<f4*>(park + c * kPlane) = w;
            }
            // s_k to LDS (own positions: no other thread reads them before the barrier)
#pragma unroll
            for (int c = 0; c < kPer / 4; c++) {
                f4 w;
#pragma unroll
                for (int i = 0; i < 4; i++) w[i] = v[4 * c + i];
                *reinterpret_cast<f4*>(own + c * kPlane) = w;
            }
            asm volatile("" ::: "memory");
            float acc[kPer];
            // The flips in the order of the per-step kernel with 2^14-amplitude
            // tiles (its own flips, then its partner tiles in bit order):
            // amplitude bits 0, 1, 12, 13 (registers; value index q = 4 c + i:
            // q bits 0, 1 are amplitude bits 0, 1, q bits 2 .. 5 amplitude bits
            // 11 .. 14), 2 .. 10 (thread bits, through LDS), 11 and 14 (own
            // planes), then the other members' 15 .. n - 1.  Per parity of the
            // plane c (bit 11), from this thread's s_k in LDS.
            in_thread<0>(acc, own);
            __builtin_amdgcn_sched_barrier(0);
            in_thread<1>(acc, own);
#pragma unroll
            for (int q = 0; q < kPer; q++) asm volatile("" : "+v"(acc[q]));
            __builtin_amdgcn_sched_barrier(0);
            __syncthreads();
#pragma unroll 1
            for (int bb = 0; bb < kTBits; bb++) lds_quarters_add(acc, s_x + 4 * (t ^ (1 << bb)));
            own_bits<0>(acc, own);
            __builtin_amdgcn_sched_barrier(0);
            own_bits<1>(acc, own);
            __builtin_amdgcn_sched_barrier(0);
            own_bits<2>(acc, own);
            __builtin_amdgcn_sched_barrier(0);
            own_bits<3>(acc, own);
            __builtin_amdgcn_sched_barrier(0);
            // the other members' s_k: their signals have had the LDS phase to arrive
            // (memory clobbers: no LDS access moves across the wait)
            asm volatile("" ::: "memory");
            if (G > 1 && k > 0) group_wait(sync, err, cnt, G * gen, polls);
            asm volatile("" ::: "memory");
#pragma unroll
            for (int j = 0; j < GB; j++) glob_quarters<1, false>(acc, region(src + (mbase ^ (1ull << (kWgBits + j)))), boff);
            // the output and - a2 s_{k-1}, half by half, in place in acc
            own_final<0>(acc, e, s2, own, park, J.a1, hf, be, J.a2);
            __builtin_amdgcn_sched_barrier(0);
            own_final<1>(acc, e, s2, own, park, J.a1, hf, be, J.a2);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < kPer; q++) v[q] = acc[q];
            store_step(v, J.out + (uint64_t)k * N + mbase, boff);
            step_done<G>(cnt, gen);
        }
    }
}

__global__ void __launch_bounds__(256) chain_sync_clear(uint32_t* sync, uint32_t words)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < words; i += gridDim.x * 256) sync[i] = 0u;
}

template <int GB>
void launch_gb(const ChainJob* jobs, const uint32_t* order, const uint32_t* goff, uint32_t* sync, uint32_t* err,
               uint32_t polls, hipStream_t st)
{
    hipLaunchKernelGGL((chain_kernel<GB>), dim3(kChainWorkgroups), dim3(kThreads), 0, st, jobs, order, goff, sync, err,
                       polls);
}

} // namespace

namespace mqw {

uint32_t chain_sync_words() { return kSyncStride * (1u + kChainWorkgroups); }

int launch_chains(const ChainJob* jobs_dev, const uint32_t* order_dev, const uint32_t* goff_dev, uint32_t* sync,
                  uint32_t* err, int n, int wait_polls, void* stream)
{
    if (n < kChainMinN || n > kChainMaxN) return -2;
    const uint32_t polls = wait_polls == 0 ? kSpinMax : wait_polls < 0 ? 0u : (uint32_t)wait_polls;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(chain_sync_clear, dim3(16), dim3(256), 0, st, sync, chain_sync_words());
    if (check_launch("chain_sync_clear")) return -1;
    switch (n - kWgBits) {
    case 0: launch_gb<0>(jobs_dev, order_dev, goff_dev, sync, err, polls, st); break;
    case 1: launch_gb<1>(jobs_dev, order_dev, goff_dev, sync, err, polls, st); break;
    case 2: launch_gb<2>(jobs_dev, order_dev, goff_dev, sync, err, polls, st); break;
    case 3: launch_gb<3>(jobs_dev, order_dev, goff_dev, sync, err, polls, st); break;
    case 4: launch_gb<4>(jobs_dev, order_dev, goff_dev, sync, err, polls, st); break;
    default: launch_gb<5>(jobs_dev, order_dev, goff_dev, sync, err, polls, st); break;
    }
    return check_launch("chain");
}

} // namespace mqw

// Internal declarations shared by the host engine (mqw_engine.cpp, mqw_host.cpp)
// and the gfx950 kernels (mqw_kernels.hip).  Not part of the public ABI.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/mqw.h"

#if defined(__HIPCC__)
#define MQW_HD __host__ __device__
#else
#define MQW_HD
#endif

namespace mqw {

constexpr int kMaxN = MQW_MAX_N;

// ---------------------------------------------------------------------------
// Host-side pieces of the path (mqw_host.cpp).
// ---------------------------------------------------------------------------
float eigen_sum_f32(const float* x, long size);
void unpack(int n, const double* rec, float* J, float* h);
double ppnd16(double p);
// shape: optional {skewness, excess kurtosis} of the diagonal, for a
// Cornish-Fisher corrected ground-state depth a (nullptr: the reference's).
void schedule_heuristic(int n, int m, const float* J, const float* h, mqw_schedule* out,
                        const double* shape = nullptr);
void sample_times(int m, int samples, const float* short_t, float* times);
// Coefficients c[0..K) of cos(sx)+sin(sx); returns K or -1 when cap is too small.
int cheb_coeffs(double s, double tol, double* c, int cap);
int cheb_coeffs_vec(double s, double tol, std::vector<double>& c);
void generate_sk(int n, int count, uint64_t seed, double* out);
// FP32 and FP64 energy of the all -1 state: (float)(sum J - sum h) as the walk
// starts (MultiQW.cpp:332), and the exact FP64 value from the same couplings.
double e_init_compat(int n, const float* J, const float* h);
double e_init_exact(int n, const float* J, const float* h);

// ---------------------------------------------------------------------------
// Term jobs: one complex Clenshaw term for one trajectory.
//   u   = alpha * E .* s1 - beta * H_G s1            (s1 per SRC1_*)
//   out = w1 * u - w2 * s2 + c * Psi                 (complex scalars)
// with Psi the stage's input state scaled by 1/||Psi||.  That norm comes from
// the per-tile partial sums the previous stage's final term left in norm_in
// (F_NORM_IN; every workgroup reduces them in the same fixed order); terms
// whose s1 or s2 is Psi take the same factor.  flags are uniform per job.
// ---------------------------------------------------------------------------
enum : uint32_t {
    SRC1_ARRAY = 0u,
    SRC1_PSI = 1u,
    SRC1_ZERO = 2u,
    SRC2_ARRAY = 0u << 2,
    SRC2_PSI = 1u << 2,
    SRC2_ZERO = 2u << 2,
    F_PSI_CONST = 1u << 4, // Psi is |+>: psi0 on every real part, no array
    F_FINAL = 1u << 5,     // write the tile's sum |out|^2 to norm_part[tile]
    F_NORM_IN = 1u << 6,   // Psi's norm from norm_in (ntiles partial sums)
    F_READOUT = 1u << 7,   // tile 0 writes |Psi[eloc]|^2 / ||Psi||^2 to *prob
    F_HSUM = 1u << 8,      // s1's flips over its top `hi` bits are read from hsum
    F_LATE = 1u << 9,      // real halves (mode 3): partner tiles after the tile's own flips,
                           // the order of the on-chip recurrences (launch_chains)
};

template <typename real>
struct TermJob {
    const real* b1;          // s1 array (SRC1_ARRAY)
    const real* b2;          // s2 array (SRC2_ARRAY); may alias out
    real* out;               // planar output
    const real* psi;         // stage input state (planar), unless F_PSI_CONST
    const real* E;           // diagonal H_P (float for FP32, double for FP64)
    const double* norm_in;   // F_NORM_IN: partial sums of |psi|^2, one per tile
    double* norm_part;       // F_FINAL: one double per tile of out
    double* prob;            // F_READOUT
    real* hsum;              // F_HSUM: planar sums of s1 over flips of the top hi bits
    real alpha, beta;
    real w1r, w1i, w2r, w2i;
    real cr, ci;
    real psi0;
    uint32_t flags;
    uint32_t eloc;
};

struct ReduceJob {
    const double* norm_part; // ntiles partial sums of |out|^2
    const void* out;         // planar state (float or double)
    uint64_t N;
    uint32_t ntiles;
    uint32_t eloc;
    double* psi_scale;       // nullable: receives 1/||out||
    double* prob;            // nullable: receives |out[eloc]|^2 / ||out||^2
};

// Shared first stage.  Every m > 1 sample of an instance starts its first
// stage from |+> under the same H, so the Chebyshev vectors V_k = T_k(Ht)|+>
// (real, Ht = H / onenorm) are computed once per instance (real-half term
// jobs, launch_half_real) and each sample's stage output is the combination
//     psi_j = sum_k c_j[k] i^(k&1) V_k
// of its own Jacobi-Anger coefficients: the quantity the Clenshaw terms of
// that stage evaluate, by the forward recurrence instead of the backward one.
constexpr int kCombineMaxK = 64;   // basis vectors per instance, V_0 included
constexpr int kCombineMaxKX = 128; // the same for the compression of a last-stage adjoint basis X
constexpr int kAdjMaxK = 254;      // a compressed adjoint basis: two combine jobs of at most 128 vectors
constexpr int kCombineBatch = 64;  // samples per combine job
struct CombineItem {
    float* out;          // planar output (2N floats)
    double* norm_part;   // nparts partial sums of |out|^2, one per workgroup
    uint32_t coef_off;   // kc coefficients (zero padded) in the coefficient table
    uint32_t pad;
};
struct CombineJob {
    const float* V;      // V_1 .. V_{kc-1}, N floats each, contiguous
    const float* V0;     // V_0 when stored (nullptr: psi0 or unit1 below)
    uint32_t kc;         // basis vectors including V_0 (<= kCombineMaxKX)
    uint32_t ns;         // samples (items) of this job (<= kCombineBatch)
    uint32_t item_off;   // first item
    float psi0;          // V_0 = psi0 * ones ...
    uint32_t unit1;      // ... unless unit1 > 0: V_0 = the unit vector at index unit1 - 1 (an adjoint basis X)
    uint32_t accum;      // 1 = add to the outputs (the second job of a basis wider than kCombineMaxKX);
                         // its norm partials are not written
};
// Compressed first-stage basis for the shared second stage.  The samples'
// first-stage states psi_j = sum_k c_j[k] i^(k&1) V_k span far fewer
// directions than kc: T_k(Ht)|+> for k < kc are numerically dependent, since
// the spectral measure of |+> under Ht is narrow.  Per parity class (k even:
// real parts, k odd: imaginary parts) the Gram matrix G = V V^T (FP64 sums of
// the FP32 vectors' products) and the samples' coefficients give, by an SVD
// of A G^(1/2), orthonormal combinations Q_r = sum_k B[r][k] V_k and sample
// coordinates u_j[r] with psi_j = sum_r u_j[r] Q_r to within a set relative
// error per sample; the second stage's basis is then T_l(Ht_2) Q_r.
struct GramJob {
    const float* V;     // V_1 .. V_{kc-1}, N floats each
    const float* V0;    // V_0 when stored (nullptr: psi0 or unit1 below)
    double* out;        // gram_pairs(kc) pair sums, written by the reduce launch
    uint32_t kc;
    float psi0;         // V_0 = psi0 * ones, unless unit1 > 0: the unit vector at index unit1 - 1
    uint32_t unit1;
    uint32_t pad;
};
// pairs (k1 <= k2, k1 = k2 mod 2) ordered by k2 then k1: pair_start(k2) + k1 / 2
inline uint32_t gram_pair_start(uint32_t k2) { return k2 == 0 ? 0u : k2 + ((k2 - 1) * (k2 - 1)) / 4; }
inline uint32_t gram_pairs(uint32_t kc) { return gram_pair_start(kc); }
constexpr int kGramMaxK = 256;
// pairs per partial sum of a launch whose largest basis has kc_max vectors
inline uint32_t gram_pmax(uint32_t kc_max)
{
    return gram_pairs(kc_max <= 32 ? 32u : kc_max <= 64 ? 64u : kc_max <= 128 ? 128u : 256u);
}
// Pair sums of every job's vectors: njobs x nchunk partial sums of
// gram_pmax(kc_max) doubles in `partial`, then their fixed-order sum per job.
// gram_nchunk: amplitude chunks of 8192 amplitudes, at least 64 chunks (where
// N allows) and at most 512, so a single large basis still spreads over the
// device.  It depends on n alone: the sums, and so the compressions, are the
// same however the jobs are grouped into launches.
inline uint32_t gram_nchunk(int n)
{
    const uint64_t N = 1ull << n;
    const uint64_t lo = std::min<uint64_t>(64, N / 64);
    return (uint32_t)std::max<uint64_t>(lo, std::min<uint64_t>(512, N / 8192));
}
int launch_gram(const GramJob* jobs_dev, uint32_t njobs, uint32_t kc_max, int n, uint32_t nchunk, double* partial,
                void* stream);

// Basis recurrences with the state on chip (mqw_chain.hip): a chain
//   s_1 = a1 (alpha0 E s_0 - beta0 M s_0),
//   s_{k+1} = a1 (alpha E s_k - beta M s_k) - a2 s_{k-1}  (k >= 1)
// (M = sum_i X_i; a1 = a2 = 1, kept as values so the arithmetic is that of
// the per-step jobs) runs all its steps in one launch on one group of
// 2^(n-15) workgroups of one XCD.  kChainWorkgroups workgroups in all, one
// per CU: groups g = xcd * (32 >> (n - 15)) + slot / 2^(n-15) for workgroup
// b = 8 slot + xcd; order[goff[g] .. goff[g + 1]) lists group g's chains.
struct ChainJob {
    const float* E;     // the instance's diagonal (N floats)
    const float* seed;  // s_0, or nullptr: the constant psi0 (|+>)
    float* out;         // s_{k+1} at out + k N
    float alpha0, beta0, alpha, beta;
    float a1, a2;
    float psi0;
    uint32_t steps;     // outputs, >= 1
};
constexpr int kChainMinN = 15, kChainMaxN = 20;
constexpr int kChainWorkgroups = 256; // 8 XCDs x 32 CUs
uint32_t chain_sync_words();
// sync: chain_sync_words() words the launch clears first; err: a word the
// kernel ORs 1 (a wait ran out) or 2 (a group's members on several XCDs) into;
// wait_polls: reads of a group counter before a wait runs out (0: 2^21; < 0:
// none, every wait between workgroups runs out and the launch aborts)
int launch_chains(const ChainJob* jobs_dev, const uint32_t* order_dev, const uint32_t* goff_dev, uint32_t* sync,
                  uint32_t* err, int n, int wait_polls, void* stream);
// The compression of one instance from its Gram pair sums and its samples'
// first-stage coefficients (each row kc long, zero padded): rank R per class
// (the same for both), B (2R x kc, row 2r + cls combines class cls) and u
// (samples x 2R); false when 2R would not be below kc.
struct FirstStageCompression {
    int R = 0;
    std::vector<double> B;
    std::vector<double> u;
};
bool compress_first_stage(int kc, const double* gram, const std::vector<const double*>& coefs, double rel_tol,
                          FirstStageCompression& out);

int launch_combine(const CombineJob* jobs_dev, uint32_t njobs, const CombineItem* items_dev, const float* coefs_dev,
                   int n, uint32_t nparts, uint32_t kc_max, void* stream);

// Shared second stage.  The second stage's output of sample j is
//     psi2_j = sum_{l,k} d_jl c_jk i^(l&1) i^(k&1) W_lk,  W_lk = T_l(Ht_2) V_k,
// with c_j, d_j the sample's first- and second-stage coefficients and V_k
// the shared first-stage basis: one basis W of kc1 x kc2 real vectors per
// instance (real-half recurrences from each V_k), and per sample a GEMM row.
// The rows (l, k) split into a real class ((l&1) == (k&1); sign -1 when both
// are odd) and an imaginary class.  The scale of psi2 is irrelevant: every
// later stage and the readout normalise their input.
constexpr int kC2MaxK = 64;   // kc1 and kc2 limit of the FP32 kernel
constexpr int kC2MaxK2 = 128; // kc1 and kc2 limit of the split-FP16 kernel (rows at most kC2MaxRows)
constexpr int kC2Tile = 64;   // samples per workgroup of the FP32 kernel (2 MFMA row blocks of 32)
constexpr int kC2TileH = 128; // samples per workgroup of the split-FP16 kernel (4 row blocks of 32)
constexpr int kC2RowPad = 16; // each class's rows padded to a multiple of this (one FP16 MFMA k block)
constexpr int kC2MaxRows = kC2MaxK * kC2MaxK2 + 2 * kC2RowPad;
struct Combine2Item {
    float* out;           // planar output (2N floats)
    double* norm_part;    // nparts partial sums of |out|^2, one per workgroup
    uint32_t c1_off;      // kc1 first-stage coefficients (zero padded)
    uint32_t c2_off;      // kc2 second-stage coefficients (zero padded)
};
struct Combine2Tile {
    const float* plus;    // |+> as N floats: W_00
    const float* V;       // V_1 .. V_{kc1-1}: W_0k
    const float* W;       // W_lk for l >= 1 at W + (k (kc2 - 1) + l - 1) N
    const uint32_t* qtab; // q0 real rows then q1 imaginary rows: l | k << 8 | s << 16, s = 1 plus, 2 minus, 0 pad
    uint32_t q0, q1;      // row counts, multiples of kC2RowPad
    uint32_t kc1, kc2;
    uint32_t item_off, ns;
    // split-FP16 kernel: the tile's coefficient matrix (written by the prep
    // launch) at A + a_off halves, scaled by ascale (a power of two putting
    // its largest entry near 2^14); outputs are scaled back by oscale.  The
    // offset is 64-bit: a plan of many tiles passes 2^32 halves.
    uint64_t a_off;
    float ascale, oscale;
};
int launch_combine2(const Combine2Tile* tiles_dev, uint32_t ntiles, const Combine2Item* items_dev,
                    const float* coefs_dev, int n, uint32_t nparts, uint32_t ns_max, void* stream);
// Same product with FP32 accuracy from FP16 matrix ops: each operand split
// into high and low FP16 parts, three products per 16-row block, FP32
// accumulation.  A_dev: scratch of FP16 values (bit patterns), combine2h_nsp(ns_max)
// * 2 * (q0 + q1) per tile, placed at each tile's a_off.
int launch_combine2h(const Combine2Tile* tiles_dev, uint32_t ntiles, const Combine2Item* items_dev,
                     const float* coefs_dev, uint16_t* A_dev, int n, uint32_t nparts, uint32_t ns_max, void* stream);
// padded samples per tile of the split kernel for a launch whose largest tile has ns_max samples
inline uint32_t combine2h_nsp(uint32_t ns_max) { return ns_max <= 32 ? 32u : ns_max <= 64 ? 64u : 128u; }

// Adjoint readout of the last stage.  The reference keeps only
// |psi_m[eloc]|^2 / ||psi_m||^2 of the last stage's output psi_m = P psi
// (P = sum_k a_k T_k(Ht), a_k = c_k i^(k&1)).  T_k(Ht) is real symmetric, so
//     psi_m[eloc] = sum_k a_k <X_k, psi>,   X_k = T_k(Ht) e_eloc,
// with one real basis X per instance (the same recurrence as the shared first
// stage, started from the unit vector e_eloc), and ||psi_m|| = ||psi|| to
// within the series' truncation (the tail sum of |c_k| beyond K, at most the
// coefficient tolerance), so the probability is |sum_k a_k <X_k, psi>|^2 / ||psi||^2.
constexpr int kReadTile = 64;   // samples per dot workgroup
constexpr int kReadMaxK = 256;    // basis vectors per instance (X_0 included), matrix-core dots
constexpr int kReadMaxKF32 = 128; // the same for the FP32 vector dots (values in registers)
constexpr int kReadMfmaMaxN = 25; // matrix-core dots: byte offsets below 2^32 within 32 basis rows
struct ReadItem {
    const float* psi;         // the last stage's input state (planar 2N, unnormalised)
    const float* phi;         // pair dots only: the other factor (planar 2N)
    const double* norm_part;  // its nparts partial sums of |psi|^2
    double* prob;             // receives the probability
    uint32_t coef_off;        // kx coefficients (zero padded) of the last stage
    uint32_t pad;
};
struct ReadTile {
    const float* X;           // X_0 .. X_{kx-1}, N floats each
    uint32_t kx;              // basis size (<= kReadMaxK; kReadMaxKF32 for the FP32 dots)
    uint32_t ns;              // items of the tile (<= kReadTile)
    uint32_t item_off;
    uint32_t pad;
    float ascale, oscale;     // matrix-core dots: coefficient scale (a power of two) and its inverse x 2^-12
};
// x[i] = v for i != k, x[k] = vk (k >= N: every element v).
int launch_fill(float* x, uint64_t N, float v, uint64_t k, float vk, void* stream);
// Partial amplitudes sum_k a_jk <X_k, psi_j> over nchunk amplitude chunks:
//   dots[((tile * nchunk + chunk) * kReadTile + j) * 2 + {0: Re, 1: Im}]
int launch_readout_dots(const ReadTile* tiles_dev, uint32_t ntiles, const ReadItem* items_dev, const float* coefs_dev,
                        int n, uint32_t nchunk, uint32_t kx_max, double* dots, void* stream);
// The same dots on the FP16 matrix cores with split operands (FP32 accuracy):
// phi = X a as a product over the basis index, then sum_a phi psi per lane.
int launch_readout_dots_mfma(const ReadTile* tiles_dev, uint32_t ntiles, const ReadItem* items_dev,
                             const float* coefs_dev, int n, uint32_t nchunk, double* dots, void* stream);
// Two-stage adjoint readout (m >= 4).  For the last two stages, with P_i the
// stage propagators and e = e_eloc,
//     psi_m[eloc] = e^T P_m P_{m-1} psi = (P_{m-1}^T P_m^T e)^T psi,
// and P^T = sum_k a_k T_k(Ht) (T_k real symmetric), so the adjoint state
// phi_j = P_m^T e = sum_k a_jk X_k lies in the span of the last stage's basis
// X, which compresses like the first stage's V (the samples' phi_j span few
// directions Z_r).  phi'_j = P_{m-1}^T phi_j is then a second-stage GEMM
// over Y_lr = T_l(Ht_{m-1}) Z_r, and the amplitude the bilinear dot
// sum_a phi'_j[a] psi_j[a] with psi_j the input of stage m - 1; its norm
// stands for the output's, as for the one-stage readout.
// Partial amplitudes sum_a phi[a] psi[a] (complex, no conjugate) of every
// item's pair, per amplitude chunk, in the layout of launch_readout_dots.
int launch_pair_dots(const ReadTile* tiles_dev, uint32_t ntiles, const ReadItem* items_dev, int n, uint32_t nchunk,
                     double* dots, void* stream);
// Per item: sum the chunks and divide |amplitude|^2 by the norm.
int launch_readout_finalize(const ReadTile* tiles_dev, uint32_t ntiles, const ReadItem* items_dev, uint32_t nchunk,
                            uint32_t nparts, const double* dots, void* stream);

struct EnergyDev {
    double e0, emax;
    uint32_t eloc;
    float shift, e_abs;
    double center; // (emax + e0) / 2 in FP64
};

// ---------------------------------------------------------------------------
// Kernel launchers (mqw_kernels.hip).  All asynchronous on `stream`.
// ---------------------------------------------------------------------------
struct EnergyScratch {
    double* block_sum;   // nblocks
    double* block_min;   // nblocks
    uint32_t* block_arg; // nblocks
    double* block_max;   // nblocks
};
size_t energy_blocks(int n);
// Energies into hp (float, shifted) and optionally e64 (double, shifted by the
// FP64 centre when e64_shift, else unshifted), info on the device.
int launch_energy(int n, const float* J_dev, const float* h_dev, double e_init, int exact, float* hp,
                  double* e64, int e64_shift, const EnergyScratch& scr, EnergyDev* info_dev, void* stream);
// The same for cnt problems of one n in one set of launches (the same kernels
// and results): couplings then fields of problem p at Jh_dev + p (n^2 + n),
// its initial energy at e_inits_dev[p], its outputs at hp + p N (e64 + p N)
// and info_dev + p, its block tables at p energy_blocks(n) in each of scr's.
int launch_energies(int n, uint32_t cnt, const float* Jh_dev, const double* e_inits_dev, int exact, float* hp,
                    double* e64, int e64_shift, const EnergyScratch& scr, EnergyDev* info_dev, void* stream);

int term_tile_bits_supported(int precision, int tb);
int launch_term(int precision, int tb, int variant, const void* jobs_dev, uint32_t njobs, int n, int lc,
                void* stream);
int term_variants();
// Real-half sweep (FP32): hb bits of one real half per workgroup; partial
// sums are laid out [half][tile], 2 * 2^(n-hb) per output.
int half_tile_bits_supported(int hb);
int half_variants(int hb);
// hi > 0: jobs flagged F_HSUM take the flips of the top hi bits from hsum
// (written by launch_high_pass) instead of partner-tile loads.
int launch_half_term(int hb, int variant, const void* jobs_dev, uint32_t njobs, int n, int lc, int hi, void* stream);
// The real half only, one workgroup per (job, tile): real recurrences whose
// arrays are N floats (jobs with real scalars and no F_HSUM).
int launch_half_real(int hb, int variant, const void* jobs_dev, uint32_t njobs, int n, int lc, void* stream);
// High-bit pass (FP32): for every job flagged F_HSUM, hsum = sum over the top
// hi bits b of s1[k ^ 2^b], both real halves, in pencils of 64 * 2^wb B rows;
// 4 <= hi, hi + wb <= 10, wb <= 2, n - hi >= 4 + wb.
int high_pass_supported(int n, int hi, int wb);
int launch_high_pass(int hi, int wb, const void* jobs_dev, uint32_t njobs, int n, void* stream);
// X-string mixer sweep (FP32/FP64): sum over nmask flip masks instead of the
// hypercube; 2^mask_tile_bits(n) amplitudes (one partial norm) per workgroup.
// xy: every mask is a pair 2^a + 2^b and contributes X_a X_b + Y_a Y_b, i.e.
// 2 psi[x ^ m] where bits a and b of x differ, else nothing.
int mask_tile_bits(int n);
int launch_mask_term(int precision, const void* jobs_dev, uint32_t njobs, int n, const uint32_t* masks_dev, int nmask,
                     bool xy, void* stream);
int launch_reduce(int precision, const ReduceJob* jobs_dev, uint32_t njobs, void* stream);
int launch_half_step(const float* b1, float* b2, const float* hp, const float* psi, int n, float alpha,
                     float beta, float coef, void* stream);

// LDS-resident trajectory kernel (whole trajectory inside one workgroup).
struct LdsSeg {
    float alpha, beta;   // 2/scale, 2 gamma/scale
    uint32_t coef_off;   // offset of K float coefficients in the coefficient table
    uint32_t K;
    int32_t prob_slot;   // >= 0: write the readout probability there
    uint32_t pad;
};
struct LdsTraj {
    const float* E;       // diagonal (2^n floats)
    uint32_t seg_off;     // first segment in the segment table
    uint32_t nseg;
    uint32_t eloc;
    uint32_t pad;
};
int lds_supported(int n);
int launch_lds(int n, const LdsTraj* trajs, uint32_t ntraj, const LdsSeg* segs, const float* coefs,
               double* probs, uint32_t* counter, int grid, void* stream);
int lds_grid(int n, int device);

const char* last_kernel_error();
// After a launch: 0, or -1 with the runtime's error kept for last_kernel_error.
int check_launch(const char* what);
// Counts of the shifted diagonal in `bins` equal bins over [-r, r] (counts zeroed by the caller).
int diag_histogram(const float* hp, uint64_t N, float r, int bins, uint32_t* counts, void* stream);
// Power sums k = 1..4 of (hp + shift - mu): 4 doubles per block into part.
int diag_moment_blocks(uint64_t N);
int diag_power_sums(const float* hp, uint64_t N, double shift, double mu, double* part, void* stream);

// ---------------------------------------------------------------------------
// Context (opaque in the C ABI) and helpers shared by the host translation units.
// ---------------------------------------------------------------------------
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
};

struct SweepCfg {
    bool half = false; // real-half kernel (FP32) instead of complex tiles
    int tb = 12;       // tile bits of that kernel
    int lc = 0;        // XCD chunk bits
    int variant = 0;   // loads-in-flight variant of the real-half kernel
    bool chains = false; // basis recurrences on chip (launch_chains) where they apply
    int chain_polls = 0; // reads per wait of theirs (0: the kernel's default; < 0: none)
    uint32_t parts = 1; // partial norm sums per output
    int hi = 0;        // top bits summed by the high-bit pass (0 = none)
    int hwb = 0;       // its pencil rows are 64 * 2^hwb bytes
    const uint32_t* masks = nullptr; // X-string or XY mixer on the device (nmask > 0), else the hypercube
    int nmask = 0;
    bool xy = false;                 // the masks are XY pairs
};

} // namespace mqw

struct mqw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr; // the shared bases' recurrences, overlapping the main stream's work
    std::string err;
    std::map<std::string, mqw::DevBuf> bufs;
    mqw_stats stats{};
    int tuned_tb[MQW_MAX_N + 1];
    int tuned_lc[MQW_MAX_N + 1];
    int tuned_var[MQW_MAX_N + 1];
    int tuned_hp[MQW_MAX_N + 1]; // high_pass option chosen by the tuner, -1 = none
    std::string arch;
    bool chain_capable = false;   // kChainWorkgroups CUs: the on-chip recurrences can run
    uint32_t* chain_err = nullptr; // their error word (device), 0 when all went well
};

namespace mqw {
// Timing events and captured graphs, released on every return path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t create()
    {
        hipError_t e = hipEventCreate(&a);
        return e != hipSuccess ? e : hipEventCreate(&b);
    }
    EventPair() = default;
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
struct GraphGuard {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    GraphGuard() = default;
    GraphGuard(const GraphGuard&) = delete;
    GraphGuard& operator=(const GraphGuard&) = delete;
    ~GraphGuard()
    {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
    }
};

// The entry point that allocates nothing (mqw_clenshaw_noalloc) carves its
// device buffers from one fixed array per device (mqw_noalloc.hip): 40 MiB,
// which holds the diagonal and the four planar states of n <= 20 (36 MiB)
// plus partial norms and the job table.
constexpr size_t kNoAllocScratchBytes = (size_t)40 << 20;
constexpr int kNoAllocMaxN = 20;
// Looks up this device's copy of the array (once per device; the library's
// context set-up calls it) and returns the lookup's result.
hipError_t noalloc_scratch_init(int device);
// Device address of this device's copy, or nullptr before a successful init.
void* noalloc_scratch(int device);

mqw_status fail(mqw_ctx* c, mqw_status s, const char* fmt, ...);
void* dev_get(mqw_ctx* c, const std::string& name, size_t bytes);
void release_bufs(mqw_ctx* c);
SweepCfg choose_sweep(const mqw_ctx* c, int precision, int n, int tile_bits, int chunk_bits, int variant = -1,
                      int high_pass = 0);
} // namespace mqw

#define HIPCHK(ctx, expr)                                                                                   \
    do {                                                                                                    \
        hipError_t _e = (expr);                                                                             \
        if (_e != hipSuccess) return mqw::fail((ctx), MQW_EDEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

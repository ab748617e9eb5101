/*
 * mqw.h -- C ABI of libmqw, the MI355X (gfx950) multi-stage quantum-walk
 * simulator.  Drop-in for the hot path of Asa-Hopkins/MultiStageQW
 * (MultiQW.cpp): Ising energies + ground-state search, the per-stage
 * gamma/t heuristic, the Chebyshev/Clenshaw exp(iHt) propagator with a
 * matrix-free H|psi>, renormalisation and the ground-state probability.
 *
 * Conventions
 *   - Every function returns mqw_status (0 = MQW_OK).  Functions that take a
 *     context leave a message in mqw_last_error(ctx) on failure.
 *   - Complex states are planar, like the reference (MultiQW.cpp:191-194):
 *     N = 2^n real parts followed by N imaginary parts.
 *   - Pointers named *_dev are device pointers (hipMalloc'ed on the
 *     context's device); every other pointer is caller-owned host memory.
 *   - A context owns one HIP stream and a grow-only device workspace; use one
 *     context per host thread.
 *   - H = H_P - gamma * sum_i X_i (the reference's sign, MultiQW.cpp:92), and
 *     the propagator applies exp(+iHt) (odd Chebyshev terms carry +i,
 *     MultiQW.cpp:206-217).
 */
#ifndef MQW_H
#define MQW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MQW_ABI_VERSION 18
#define MQW_DEFAULT_SAMPLES 100 /* MultiQW.cpp:257 */
#define MQW_DEFAULT_TOL 1e-6    /* MultiQW.cpp:440 */
#define MQW_DEFAULT_PROBLEMS 2000 /* MultiQW.cpp:252 */
#define MQW_MAX_STAGES 64
#define MQW_MIN_N 4  /* the reference's 16-wide loads need N >= 16 (MultiQW.cpp:117-119) */
#define MQW_MAX_N 30

typedef enum {
    MQW_OK = 0,
    MQW_EINVAL = 1,  /* bad argument (n, m, sizes, NaN parameters ...) */
    MQW_ENOMEM = 2,  /* host or device allocation failed */
    MQW_EIO = 3,     /* file missing, short or unwritable (CLI helpers) */
    MQW_EDEVICE = 4, /* HIP runtime error */
    MQW_ENOTSUP = 5
} mqw_status;

typedef enum { MQW_ENERGY_COMPAT = 0, MQW_ENERGY_EXACT = 1 } mqw_energy_mode;
typedef enum { MQW_FP32 = 0, MQW_FP64 = 1 } mqw_precision;
typedef enum { MQW_GEMM_SPLIT_FP16 = 0, MQW_GEMM_FP32 = 1 } mqw_gemm;
typedef enum { MQW_MIXER_XSTRING = 0, MQW_MIXER_XY = 1 } mqw_mixer;

typedef struct mqw_ctx mqw_ctx;

/* Result of the energy build (MultiQW.cpp:325-362). */
typedef struct {
    double e0;        /* lowest energy of the walk */
    double emax;      /* highest energy */
    uint32_t eloc;    /* ground-state index; the first in Gray order on ties */
    float shift;      /* (float)((emax+e0)/2), subtracted from H_P */
    float e_abs;      /* (float)((emax-e0)/2), the diagonal's spectral radius */
    double e_init;    /* energy of the all -1 state */
} mqw_energy_info;

/* Heuristic schedule (MultiQW.cpp:365-423). */
typedef struct {
    int n, m;
    float a;       /* expected maximum of 2^n Gaussians, in sigma units */
    float hp2;     /* mean_k E(k)^2 = 2 sum J^2 + sum h^2 */
    float delta2;  /* 16 sum J^2 + 4 sum h^2 */
    float total_t; /* sum of short_t, printed before each result */
    float gamma[MQW_MAX_STAGES];
    float short_t[MQW_MAX_STAGES];
} mqw_schedule;

/* Options of a solve.  Initialise with mqw_opts_default. */
typedef struct {
    int samples;       /* Monte-Carlo samples per problem (100) */
    double tol;        /* Chebyshev tail tolerance (1e-6) */
    int precision;     /* mqw_precision: state/recurrence in FP32 (default) or FP64 */
    int energy_mode;   /* mqw_energy_mode for H_P and E_loc (COMPAT = the reference walk) */
    int tile_bits;     /* 0 = tuned/default; else tile bits: complex tiles 4..12, real-half 13..14 (FP32) */
    int chunk_bits;    /* -1 = tuned/default; XCD chunk of consecutive tiles, log2 */
    int small_n_max;   /* -1 = default; largest n run by the LDS-resident kernel (0 = never) */
    int use_graph;     /* 1 = capture each stage plan in a hipGraph (default 1) */
    size_t max_device_bytes; /* 0 = no cap; else a cap on the trajectory buffers of a group */
    size_t group_bytes;      /* 0 = default (32 GiB): trajectory buffers of one group; a group's
                                trajectories share every sweep launch */
    int variant;             /* -1 = tuned/default; else a sweep-kernel variant (loads in flight) */
    float* sample_probs;     /* optional output: count*samples per-sample probabilities */
    int high_pass;           /* -1 = by n (default); 0 = off; 1, 2, 3 = on with LDS pencil rows of 64,
                                128, 256 B: before each sweep a pass sums the flips over the top 10, 9,
                                8 bits in on-chip pencils, so the sweep loads no partner tiles for them;
                                4 = on, register pencils (no LDS) for the 4..7 bits beyond the sweep's
                                XCD chunk (FP32 real-half sweeps, n >= 17) */
    const uint32_t* mixer_masks; /* NULL = the reference's mixer sum_i X_i; else mixer_count flip masks:
                                    H = H_P - gamma * sum_k prod_{i in m_k} X_i (multi-X mixers, the
                                    reference's README.md:41 to-do); spectral bound E_abs + gamma * count.
                                    The heuristic schedule still assumes sum_i X_i: pass schedules
                                    through mqw_instance.gamma/short_t. */
    int mixer_count;         /* 1..64 masks, each nonzero and < 2^n */
    int fixed_times;         /* 0 = Monte-Carlo stage times U(t, 2t) as MultiQW.cpp:417-430; 1 = every
                                sample runs each stage for exactly short_t (deterministic schedules,
                                e.g. a time-dependent anneal split into stages) */
    int shared_first_stage;  /* 1 (default) = the m > 1 samples of an instance share the Chebyshev
                                vectors T_k(H/onenorm)|+> of their common first stage, each sample's
                                output is its own coefficient combination of them (FP32 real-half
                                sweep, hypercube mixer); 0 = every sample runs its own Clenshaw terms */
    int final_readout;       /* 1 (default) = the last stage of an m > 1 sample is read out without
                                propagating it: psi_m[E_loc] = sum_k a_k <T_k(H/onenorm) e_Eloc, psi>
                                from one real basis per instance, and ||psi_m|| = ||psi|| (the truncated
                                series is unitary to within its coefficient tail, <= tol), so the
                                probability moves by at most 2 tol relative (MultiQW.cpp:441-446
                                renormalises the propagated state); 0 = the last stage runs as
                                Clenshaw terms and its output is renormalised, as the reference */
    int shared_second_stage; /* 1 (default) = for m >= 3, where it costs less than the samples' own terms,
                                the second stage too comes from a shared basis: W_lk = T_l(H2/onenorm2) V_k
                                per instance (kc1 x kc2 real vectors) and one matrix-core GEMM row of
                                coefficient products per sample; 0 = per-sample Clenshaw terms */
    int overlap_bases;       /* 1 (default) = when a solve holds several batches of shared bases, the
                                next batch's basis recurrences run on a second stream while this
                                batch's GEMM, readout and sweeps run (two basis buffers); 0 = one batch
                                after the other on one stream.  Results are identical either way. */
    int second_stage_gemm;   /* arithmetic of the shared second stage's GEMM: MQW_GEMM_SPLIT_FP16 (default)
                                = FP16 matrix ops on operands split into high and low FP16 parts
                                (three products, FP32 accumulation: 22-bit operands, FP32-level
                                error); MQW_GEMM_FP32 = FP32 matrix ops */
    int mixer_kind;          /* with mixer_masks: MQW_MIXER_XSTRING (default) = sum_k prod_{i in m_k} X_i;
                                MQW_MIXER_XY = sum_k (X_a X_b + Y_a Y_b) over pairs m_k = 2^a + 2^b
                                (exactly two bits each; the reference README.md:41 "XY" to-do):
                                (H_M psi)[x] = sum_k 2 psi[x ^ m_k] where bits a, b of x differ.  It
                                keeps Hamming weight; spectral bound E_abs + gamma * 2 * count. */
    int compress_bases;      /* 1 (default) = where the second stage is shared, its basis is built from a
                                compressed first-stage basis: the samples' first-stage states span far
                                fewer directions than the kc1 Chebyshev vectors T_k(H1/onenorm1)|+>, so an
                                SVD (weighted by the vectors' Gram matrix, per parity class) keeps the
                                smallest set of orthonormal combinations Q_r that holds every sample's
                                state to within 2e-7 relative (a quarter of the FP32 rounding of the
                                basis itself), and W_lr = T_l(H2/onenorm2) Q_r replaces W_lk (about 18
                                instead of 37 chains at n = 20); 0 = W from every V_k */
    int direct_m1;           /* 1 (default) = an m == 1 instance's chained samples (MultiQW.cpp:426-430: each
                                sample continues from the previous one by the next time increment) are
                                evaluated directly: sample j is exp(iH T_j)|+> with T_j the sum of the
                                increments, one Chebyshev series in the instance's basis T_k(H/onenorm)|+>,
                                its amplitude and norm from V_k[E_loc] and the basis's Gram matrix (no term
                                sweeps).  It differs from the chain by the chain's accumulated truncation,
                                measured below 4e-7 in probability; 0 = the chain, as the reference */
    int adjoint_two_stages;  /* 1 (default) = for m >= 4, where it costs less than the samples' own stage
                                m - 1 terms, the last two stages are read out together: the samples'
                                adjoint states phi_j = P_m^T e_Eloc (P_i the stage propagators, sums of
                                T_k(H_i/onenorm_i) with the sample's coefficients; T_k is symmetric) lie
                                in the span of the last stage's basis T_k(H_m/onenorm_m) e_Eloc, which is
                                compressed like the first stage's (2e-7 per sample); P_{m-1}^T phi_j is a
                                second-stage GEMM over T_l(H_{m-1}/onenorm_{m-1}) Z_r, and the amplitude
                                psi_m[E_loc] = (P_{m-1}^T phi_j)^T psi_{m-2} a dot with the input of stage
                                m - 1, whose norm stands for the output's (two truncated series: at most
                                4 tol relative in probability); 0 = stage m - 1 as per-sample terms */
    int shared_third_stage;  /* 1 (default) = for m >= 4, where the second stage is shared, the third stage is
                                neither read out (last, or one of the two above) and it pays: the samples'
                                second-stage states are compressed (their real parts and imaginary parts,
                                2e-7 per sample) to directions U_r, and the third stage is a GEMM over
                                T_l(H3/onenorm3) U_r like the second; 0 = per-sample terms */
    int chain_recurrences;   /* 1 (default) = for n = 15..20 on a 256-CU device, the shared bases' Chebyshev
                                recurrences keep their state on chip: one launch runs every step of a
                                chain on a group of workgroups of one XCD, each holding 2^15 amplitudes of
                                s_k and of the diagonal in registers, the partner amplitudes of the top
                                n - 15 bits read from the XCD's L2 (8 instead of about 20 bytes of HBM
                                traffic per vector and amplitude).  Results are bit-identical to 0 = one
                                launch per recurrence step.  Batches of bases then run one after the
                                other (no second stream: two such launches must not share the device). */
    int chain_wait_polls;    /* 0 = default (2^21): reads of a group counter by one wait between the
                                workgroups of an on-chip recurrence launch before it gives up; the launch
                                then ends without waiting, and the context runs its recurrences one launch
                                per step from then on.  -1 = no read: every such wait gives up at once,
                                which forces that fallback (its tests); below -1 = MQW_EINVAL. */
} mqw_opts;

/* One problem of a batch; gamma/short_t override the heuristic when both are
 * non-NULL (length m): the learned-schedule path.  Override gammas may be any
 * finite value (the spectral bound uses |gamma|); times must be positive. */
typedef struct {
    int n, m;
    const double* rec;    /* n(n+1)/2 doubles in the reference file layout */
    const float* gamma;
    const float* short_t;
} mqw_instance;

/* ---- context ------------------------------------------------------------ */
mqw_status mqw_create(int device, mqw_ctx** out);
void mqw_destroy(mqw_ctx* ctx);
const char* mqw_last_error(const mqw_ctx* ctx);
const char* mqw_status_string(mqw_status s);
int mqw_abi_version(void);
void mqw_opts_default(mqw_opts* o);

/* ---- host-side pieces of the path (no device) ---------------------------- */
/* Unpack one record (MultiQW.cpp:303-323): J symmetric n*n with the
 * reference's halving, h of length n, both float. */
mqw_status mqw_unpack(int n, const double* rec, float* J, float* h);
/* Heuristic gamma and times per stage (MultiQW.cpp:365-423). */
mqw_status mqw_schedule_heuristic(int n, int m, const double* rec, mqw_schedule* out);
/* The same heuristic with the ground-state depth a taken from the diagonal's
 * own shape: the 1/N and 1/(eN) lower-tail quantiles corrected by
 * Cornish-Fisher for its skewness and excess kurtosis (from mqw_diag_moments)
 * instead of a Gaussian's.  No counterpart in the reference, whose skewness
 * and kurtosis terms are commented out (MultiQW.cpp:379-380); zero moments
 * give mqw_schedule_heuristic's schedule exactly.  Pass the result as an
 * override (mqw_instance.gamma/short_t). */
mqw_status mqw_schedule_heuristic_moments(int n, int m, const double* rec, double skewness, double excess_kurtosis,
                                          mqw_schedule* out);
/* Monte-Carlo times times[i*samples+j] (MultiQW.cpp:387,417-430): MT19937
 * re-seeded per problem, U(short_t, 2 short_t) stage-major; m == 1 sorts and
 * differences.  short_t has m entries. */
mqw_status mqw_sample_times(int m, int samples, const float* short_t, float* times);
/* Chebyshev coefficients of cos(sx)+sin(sx) on [-1,1] in FP64, truncated at
 * tail sum <= tol (replaces ApproxTools RCF_odd_even, MultiQW.cpp:438-440). */
mqw_status mqw_chebyshev_coeffs(double s, double tol, double* c, int cap, int* K);
/* Seeded SK instances in the reference file layout: count records of
 * n(n+1)/2 N(0,1) doubles (SplitMix64-seeded xoshiro256**, Marsaglia polar). */
mqw_status mqw_generate_sk(int n, int count, uint64_t seed, double* out);

/* ---- reference file surface ---------------------------------------------- */
/* Number of complete n(n+1)/2-double records in a file. */
mqw_status mqw_count_instances(const char* path, int n, long* count);
/* Read count records starting at record `start` (MultiQW.cpp:268-271,303-304);
 * MQW_EIO if the file is missing or holds fewer records. */
mqw_status mqw_read_instances(const char* path, int n, long start, long count, double* out);
/* Append count float32 results to <dir>/output_<n>_<m> (MultiQW.cpp:264-266,457);
 * dir NULL or "" = the working directory. */
mqw_status mqw_append_output(const char* dir, int n, int m, const float* probs, long count);

/* ---- device pieces -------------------------------------------------------- */
/* Energy build + ground-state search on the GPU (MultiQW.cpp:325-362).
 * hp (nullable, 2^n floats) receives the shifted float diagonal H_P;
 * e (nullable, 2^n doubles) receives the unshifted FP64 energies. */
mqw_status mqw_energy(mqw_ctx* ctx, int n, const double* rec, int mode, float* hp, double* e,
                      mqw_energy_info* info);

/* Moments of the diagonal energies the walk uses (H_P + shift, float):
 * out[0..3] = mean, variance, skewness, excess kurtosis, from FP64 sums on
 * the device.  The reference computes the variance in closed form (HP2,
 * MultiQW.cpp:375) and leaves skewness and kurtosis commented out
 * (MultiQW.cpp:379-380). */
mqw_status mqw_diag_moments(mqw_ctx* ctx, int n, const double* rec, int mode, double* out);

/* One Clenshaw step on one real half, device pointers (the seam at
 * MultiQW.cpp:88-173): b2 <- (2/scale)(H_P - gamma H_G) b1 - b2 + coef psi. */
mqw_status mqw_clenshaw_step_dev(mqw_ctx* ctx, const float* b1_dev, float* b2_dev, const float* hp_dev,
                                 const float* psi_dev, unsigned n, float scale, float gamma, float coef);

/* Clenshaw evaluation of sum_r c_r i^(r odd) T_r(H/scale) psi on host arrays
 * (the seam at MultiQW.cpp:177-236).  psi: 2^(n+1) floats, replaced by the
 * (unnormalised) result; hp: 2^n floats. */
mqw_status mqw_clenshaw(mqw_ctx* ctx, int n, const float* coeffs, int K, float* psi, const float* hp,
                        float gamma, float scale, int psi_real);

/* mqw_clenshaw without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as mqw_clenshaw, but every device buffer (the diagonal, the states,
 * partial norms, the job table) comes from one fixed 40 MiB array that each
 * GPU holds until the process ends.  Limits: n <= 20, and K small enough that
 * the job table fits beside the states (about 30000 terms at n = 20); larger
 * calls return MQW_EINVAL.  Calls on one GPU share that GPU's array: calls
 * that overlap on one GPU (several host threads or contexts) race. */
mqw_status mqw_clenshaw_noalloc(mqw_ctx* ctx, int n, const float* coeffs, int K, float* psi, const float* hp,
                                float gamma, float scale, int psi_real);

/* The per-problem loop of MultiQW.cpp:296-455 for count problems of one
 * (n, m): recs holds count*n(n+1)/2 doubles.  prob and total_t receive one
 * float per problem.  opts may be NULL (defaults). */
mqw_status mqw_solve(mqw_ctx* ctx, int n, int m, const double* recs, int count, const mqw_opts* opts,
                     float* prob, float* total_t);

/* Uneven batch: problems of different (n, m) in one call. */
mqw_status mqw_solve_batch(mqw_ctx* ctx, const mqw_instance* inst, int count, const mqw_opts* opts,
                           float* prob, float* total_t);

/* Autotuner: for n on this device, times every sweep configuration (kernel
 * kind and tile bits, loads-in-flight variant, XCD chunk bits) on synthetic
 * jobs, then runs real solves of a generated instance with the fastest few,
 * the default and (from n = 22) the high-bit pass, and keeps the fastest for
 * later solves on the context.  With cache_path, a recorded choice for this
 * GPU architecture and n is reused, and a new one is appended (set
 * MQW_RETUNE=1 to measure again).  Outputs may be NULL; *high_pass receives the
 * mqw_opts.high_pass value chosen (0 = off, 1..3 = on). */
mqw_status mqw_tune(mqw_ctx* ctx, int n, const char* cache_path, int* tile_bits, int* chunk_bits, int* variant,
                    int* high_pass);

/* ---- learned schedule ------------------------------------------------------ */
/* Instance features (32 floats each): n/20, m/5, mean, variance and max |.| of
 * the couplings and of the fields, then a 24-bin histogram (fractions) of the
 * shifted diagonal H_P over [-E_abs, E_abs], computed on the GPU. */
#define MQW_FEATURES 32
mqw_status mqw_features(mqw_ctx* ctx, const mqw_instance* inst, int count, float* out);

/* MLP 32 -> 128 -> 128 -> 10 with GELU, trained with hand-written HIP kernels.
 * Outputs: per stage i < 5, (log gamma_i, log t_i) offsets from the heuristic. */
#define MQW_MLP_IN 32
#define MQW_MLP_HIDDEN 128
#define MQW_MLP_OUT 10
#define MQW_MLP_PARAMS 22026 /* W1 b1 W2 b2 W3 b3, row-major W[in][out] */
typedef struct mqw_mlp mqw_mlp;
mqw_status mqw_mlp_create(mqw_ctx* ctx, uint64_t seed, mqw_mlp** out);
void mqw_mlp_destroy(mqw_mlp* mlp);
mqw_status mqw_mlp_get_params(mqw_mlp* mlp, float* out, size_t count);
mqw_status mqw_mlp_set_params(mqw_mlp* mlp, const float* in, size_t count); /* also resets Adam state */
/* Y[B x 10] = f(X[B x 32]). */
mqw_status mqw_mlp_forward(mqw_mlp* mlp, const float* X, int B, float* Y);
/* Masked MSE  L = sum M (Y - T)^2 / sum M  and its gradient (kept on the
 * device for the next Adam step; copied to grads when non-NULL). */
mqw_status mqw_mlp_loss_grad(mqw_mlp* mlp, const float* X, const float* T, const float* M, int B, float* loss,
                             float* grads);
mqw_status mqw_mlp_adam_step(mqw_mlp* mlp, float lr, float beta1, float beta2, float eps, int step);

/* Kernel statistics of the last solve on this context. */
typedef struct {
    uint64_t sweeps;         /* term-kernel launches (global-memory path) */
    uint64_t term_units;     /* complex H-applies executed, summed over trajectories */
    uint64_t amp_terms;      /* sum over applies of 2^n (amplitude-terms) */
    double algorithmic_bytes;/* 36 B per amplitude-term (FP32) or 72 (FP64), executed */
    double gpu_ms;           /* device time of the propagation (events) */
    double sweep_ms;         /* device time of the term sweeps: term, high-bit pass and reduce launches */
    uint64_t lds_trajectories; /* trajectories run by the LDS-resident kernel */
    uint64_t psi_const_amp_terms; /* amplitude-terms whose psi is the constant |+> (no psi read) */
    double required_bytes;   /* algorithmic_bytes less the psi reads the |+> terms skip */
    double aux_ms;           /* device time of the shared bases, combine and readout launches (in gpu_ms) */
    double aux_bytes;        /* algorithmic bytes of the basis recursions, combine and readout launches */
    uint64_t aux_launches;   /* basis, combine and readout launches (not in sweeps) */
    double basis_ms;         /* of aux_ms: the shared bases' recurrences (real-half term kernel) */
    double basis_bytes;      /* their algorithmic bytes: 16 B per vector and amplitude (less for |+>, e_Eloc) */
    double gemm_ms;          /* of aux_ms: the second-stage GEMM launches (matrix cores, FP32) */
    double gemm_flops;       /* their executed flops (2 per multiply-add, row blocks of 32 samples) */
    double gemm_useful_flops;/* of which on real samples */
    uint64_t basis_launches; /* launches of the recurrences */
    uint64_t gemm_launches;  /* launches of the GEMM */
    double gemm_bytes;       /* the GEMM's algorithmic bytes: every basis row read once per tile, outputs written */
    uint64_t basis_batches;  /* problem batches of shared bases (each fills the basis buffer once) */
    uint64_t overlap_batches; /* of those, batches whose recurrences ran on the second stream beside the main
                                 stream's work (overlap_bases); their event times overlap, so gpu_ms then
                                 exceeds the device's busy time */
    uint64_t compressed_instances; /* instances whose second-stage basis came from a compressed first stage */
    uint64_t compressed_vectors;   /* their kept first-stage vectors (2R each), summed */
    uint64_t uncompressed_vectors; /* their first-stage vectors before compression (kc1 each), summed */
    uint64_t direct_chains;        /* m == 1 instances evaluated directly (direct_m1) */
    uint64_t first_stage_chunks;   /* chunks of the first-stage bases computed before the batches (for
                                      compression and direct m == 1 chains), each within the V share of
                                      max_device_bytes */
    uint64_t adjoint2_instances;   /* instances whose last two stages were read out together (adjoint_two_stages) */
    uint64_t adjoint2_vectors;     /* their kept adjoint directions (2R each), summed */
    uint64_t adjoint2_uncompressed; /* their last-stage basis vectors before compression (kx each), summed */
    uint64_t third_stage_instances; /* instances whose third stage came from compressed second-stage states */
    uint64_t third_stage_vectors;   /* their kept directions (2R each), summed */
    uint64_t third_stage_states;    /* their second-stage state halves before compression (2 per sample), summed */
    uint64_t chain_launches;        /* on-chip recurrence launches (chain_recurrences) */
    uint64_t chain_steps;           /* recurrence steps (vectors) they made */
} mqw_stats;
mqw_status mqw_last_stats(const mqw_ctx* ctx, mqw_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* MQW_H */

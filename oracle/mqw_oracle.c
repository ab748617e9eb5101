/*
 * mqw_oracle.c -- CPU restatement of the multi-stage quantum-walk hot path of
 * MultiQW.cpp (Asa-Hopkins/MultiStageQW).
 *
 * THIS FILE IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker (or, for
 * the CPU baseline, as the thing timed on the host).  The product library
 * (multistageqw_amd/csrc) never links, loads or calls it.
 *
 * Every function cites the reference line range it restates.  Where the
 * reference's arithmetic depends on code that is absent from the reference
 * tree (Eigen reductions, VCL2 vectors, the ApproxTools Chebyshev fit) the
 * restatement documents the assumption it makes:
 *
 *   - Eigen float reductions (.sum(), .norm()): restated as Eigen 3.4's
 *     linear vectorised reduction with 16-float packets (two packet
 *     accumulators, halving tree, scalar tail).  Assumed order; the
 *     arithmetic is pinned only by the closed forms the tests check.
 *   - GCC -O3 -march=native contraction: the reference's build line
 *     (README.md:20) contracts a*b+c into FMA.  The draw t = u*s + s is
 *     pinned bit-for-bit by golden vectors produced with g++'s libstdc++;
 *     the two FMAs in Clenshaw_step follow what g++ emits for the
 *     AVX-512 intrinsics VCL2 wraps (vfmsub for pass 1, vfmadd for pass 2).
 *   - ApproxTools Chebyshev<double>::RCF_odd_even(f, 1e-6): absent.  The
 *     coefficients of cos(sx)+sin(sx) are the Jacobi-Anger closed form
 *     c_0 = J_0(s), c_k = 2(-1)^floor(k/2) J_k(s); J_k by Miller's backward
 *     recurrence (J.C.P. Miller 1952; Abramowitz & Stegun 9.12), truncated at
 *     the first k whose tail sum of |c_j| is <= tol.  Values pinned against
 *     scipy.special.jv; the truncation rule is parity unpinned.
 *   - my_normcdfinvf (MultiQW.cpp:39-86): semantics only.  Restated as
 *     -Phi^{-1}(p) via Wichura's AS241 (PPND16, Applied Statistics 37, 1988)
 *     in double, rounded to float.  Pinned against scipy.special.ndtri.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define ORACLE_MAX_N 30
#define ORACLE_MAX_M 64
#define PI_D 3.1415926535897932384626

/* ------------------------------------------------------------------------ */
/* Eigen-style float sum (assumed order, see header).                         */
/* ------------------------------------------------------------------------ */
float oracle_eigen_sum_f32(const float* x, long size)
{
    enum { P = 16 };
    if (size <= 0) return 0.0f;
    long a2 = (size / (2 * P)) * (2 * P);
    long a1 = (size / P) * P;
    if (a1 == 0) {
        float r = x[0];
        for (long i = 1; i < size; i++) r += x[i];
        return r;
    }
    float p0[P], p1[P];
    for (int l = 0; l < P; l++) p0[l] = x[l];
    if (a1 > P) {
        for (int l = 0; l < P; l++) p1[l] = x[P + l];
        for (long i = 2 * P; i < a2; i += 2 * P)
            for (int l = 0; l < P; l++) {
                p0[l] += x[i + l];
                p1[l] += x[i + P + l];
            }
        for (int l = 0; l < P; l++) p0[l] += p1[l];
        if (a1 > a2)
            for (int l = 0; l < P; l++) p0[l] += x[a2 + l];
    }
    for (int k = P / 2; k >= 1; k >>= 1)
        for (int l = 0; l < k; l++) p0[l] += p0[l + k];
    float r = p0[0];
    for (long i = a1; i < size; i++) r += x[i];
    return r;
}

/* ------------------------------------------------------------------------ */
/* a-1: instance unpack.  MultiQW.cpp:303-323.                                */
/* rec holds n(n+1)/2 doubles: J lower triangle row-major (i=1..n-1, j<i)    */
/* then h.  J is returned symmetric (n*n, J[i*n+j] == J[j*n+i]) with the     */
/* reference's halving (:318-319); it is identical in row- and column-major. */
/* ------------------------------------------------------------------------ */
void oracle_unpack(int n, const double* rec, float* J, float* h)
{
    for (int i = 0; i < n * n; i++) J[i] = 0.0f;
    int k = 0;
    for (int i = 1; i < n; i++)
        for (int j = 0; j < i; j++) {
            float v = (float)rec[k++];
            J[i * n + j] = v;
        }
    /* J += J^T; J /= 2 */
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) {
            float v = J[i * n + j];
            J[i * n + j] = (v + 0.0f) / 2;
            J[j * n + i] = (0.0f + v) / 2;
        }
    for (int i = 0; i < n; i++) h[i] = (float)rec[n * (n - 1) / 2 + i];
}

/* ------------------------------------------------------------------------ */
/* a-2/a-3: Gray-code energy walk.  MultiQW.cpp:325-362 (+ bit helpers :15-34) */
/* hp (float, 2^n) receives the shifted diagonal; e_out (double, nullable)   */
/* receives the unshifted walk energies indexed by state.                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    double e0, emax;
    uint32_t eloc;
    float shift;   /* (float)((E_max+E_0)/2), subtracted from H_P (:361) */
    float e_abs;   /* (float)((E_max-E_0)/2) (:362) */
    double e_init; /* energy of the all -1 state (:332) */
} oracle_energy_info;

void oracle_energy_compat(int n, const float* J, const float* h, float* hp, double* e_out,
                          oracle_energy_info* info)
{
    const uint32_t N = 1u << n;
    float st[ORACLE_MAX_N * ORACLE_MAX_N];
    for (int i = 0; i < n * n; i++) st[i] = 1.0f;
    float js = oracle_eigen_sum_f32(J, (long)n * n);
    float hs = oracle_eigen_sum_f32(h, n);
    double E = (double)(js - hs);
    info->e_init = E;
    hp[0] = (float)E;
    if (e_out) e_out[0] = E;
    double e0 = E, emax = E;
    uint32_t eloc = 0;
    for (uint32_t i = 1; i < N; i++) {
        int f = __builtin_ctz(i);
        for (int j = 0; j < n; j++) st[f * n + j] = -st[f * n + j];
        for (int j = 0; j < n; j++) st[j * n + f] = -st[j * n + f];
        st[f * n + f] = -st[f * n + f];
        /* (J.row(f)*state.row(f)).sum(): non-contiguous row, sequential redux */
        float S = J[f * n + 0] * st[f * n + 0];
        for (int j = 1; j < n; j++) S += J[f * n + j] * st[f * n + j];
        float d = 4 * S - 2 * h[f] * st[f * n + f];
        E += (double)d;
        uint32_t g = i ^ (i >> 1);
        hp[g] = (float)E;
        if (e_out) e_out[g] = E;
        if (E < e0) {
            e0 = E;
            eloc = g;
        }
        if (E > emax) emax = E;
    }
    float shift = (float)((emax + e0) / 2);
    for (uint32_t k = 0; k < N; k++) hp[k] = hp[k] - shift;
    info->e0 = e0;
    info->emax = emax;
    info->eloc = eloc;
    info->shift = shift;
    info->e_abs = (float)((emax - e0) / 2);
}

/* Exact FP64 energy of every state from the float-rounded couplings:
 * E(k) = sum_{i<j} t_ij s_i s_j + sum_i h_i s_i, s_i = +1 iff bit i of k.
 * Brute force O(n^2 2^n): the check for the walk above. */
void oracle_energy_exact(int n, const float* J, const float* h, double* e_out)
{
    const uint32_t N = 1u << n;
    for (uint32_t k = 0; k < N; k++) {
        double E = 0.0;
        for (int i = 0; i < n; i++) {
            double si = ((k >> i) & 1u) ? 1.0 : -1.0;
            E += (double)h[i] * si;
            for (int j = 0; j < i; j++) {
                double sj = ((k >> j) & 1u) ? 1.0 : -1.0;
                E += 2.0 * (double)J[i * n + j] * si * sj;
            }
        }
        e_out[k] = E;
    }
}

/* ------------------------------------------------------------------------ */
/* a-4: -Phi^{-1}(p) (semantics of my_normcdfinvf, MultiQW.cpp:80-86).        */
/* Wichura AS241 PPND16 in double.                                            */
/* ------------------------------------------------------------------------ */
double oracle_ppnd16(double p)
{
    double q = p - 0.5, r, v;
    if (fabs(q) <= 0.425) {
        r = 0.180625 - q * q;
        v = q * (((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r +
                      6.7265770927008700853e+4) * r + 4.5921953931549871457e+4) * r +
                    1.3731693765509461125e+4) * r + 1.9715909503065514427e+3) * r +
                  1.3314166789178437745e+2) * r + 3.3871328727963666080e+0) /
            (((((((5.2264952788528545610e+3 * r + 2.8729085735721942674e+4) * r +
                  3.9307895800092710610e+4) * r + 2.1213794301586595867e+4) * r +
                5.3941960214247511077e+3) * r + 6.8718700749205790830e+2) * r +
              4.2313330701600911252e+1) * r + 1.0);
        return v;
    }
    r = (q < 0) ? p : 1.0 - p;
    r = sqrt(-log(r));
    if (r <= 5.0) {
        r -= 1.6;
        v = (((((((7.74545014278341407640e-4 * r + 2.27238449892691845833e-2) * r +
                  2.41780725177450611770e-1) * r + 1.27045825245236838258e+0) * r +
                3.64784832476320460504e+0) * r + 5.76949722146069140550e+0) * r +
              4.63033784615654529590e+0) * r + 1.42343711074968357734e+0) /
            (((((((1.05075007164441684324e-9 * r + 5.47593808499534494600e-4) * r +
                  1.51986665636164571966e-2) * r + 1.48103976427480074590e-1) * r +
                6.89767334985100004550e-1) * r + 1.67638483018380384940e+0) * r +
              2.05319162663775882187e+0) * r + 1.0);
    } else {
        r -= 5.0;
        v = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r +
                  1.24266094738807843860e-3) * r + 2.65321895265761230930e-2) * r +
                2.96560571828504891230e-1) * r + 1.78482653991729133580e+0) * r +
              5.46378491116411436990e+0) * r + 6.65790464350110377720e+0) /
            (((((((2.04426310338993978564e-15 * r + 1.42151175831644588870e-7) * r +
                  1.84631831751005468180e-5) * r + 7.86869131145613259100e-4) * r +
                1.48753612908506148525e-2) * r + 1.36929880922735805310e-1) * r +
              5.99832206555887937690e-1) * r + 1.0);
    }
    return (q < 0.0) ? -v : v;
}

/* my_normcdfinvf(a) returns -Phi^{-1}(a) (its erfcinv returns -erfcinv, :77). */
float oracle_neg_normcdfinv_f32(float a) { return (float)(-oracle_ppnd16((double)a)); }

/* ------------------------------------------------------------------------ */
/* a-5/a-6: moments and heuristic schedule.  MultiQW.cpp:365-423.             */
/* ------------------------------------------------------------------------ */
typedef struct {
    float a, hp2, delta2, total_t;
    float gamma[ORACLE_MAX_M];
    float short_t[ORACLE_MAX_M];
} oracle_schedule;

void oracle_schedule_heuristic(int n, int m, const float* J, const float* h, oracle_schedule* out)
{
    const unsigned N = 1u << n;
    const unsigned un = (unsigned)n, um = (unsigned)m;
    float b = oracle_neg_normcdfinv_f32(1 / (float)N);
    float e_m = 0.577215664901532860;
    float e = 2.718281828459045;
    float c = oracle_neg_normcdfinv_f32(1 / (e * N));
    /* (1-e_m)*b + e_m*c : g++ -march=native fuses the first product */
    float a = fmaf(1 - e_m, b, e_m * c);
    float jj[ORACLE_MAX_N * ORACLE_MAX_N], hh[ORACLE_MAX_N];
    for (int i = 0; i < n * n; i++) jj[i] = J[i] * J[i];
    for (int i = 0; i < n; i++) hh[i] = h[i] * h[i];
    float sjj = oracle_eigen_sum_f32(jj, (long)n * n);
    float shh = oracle_eigen_sum_f32(hh, n);
    float HP2 = fmaf(2, sjj, shh);
    float delta2 = fmaf(16, sjj, 4 * shh);
    out->a = a;
    out->hp2 = HP2;
    out->delta2 = delta2;
    for (int i = 0; i < m; i++)
        out->gamma[i] = (float)((double)(a * sqrtf(HP2)) / tan(PI_D * (i + 1) / (2 * um + 2)) / un);
    float total_t = 0;
    for (int i = 0; i < m; i++) {
        float last_denom = (float)(sqrt(2 / PI_D) * (double)out->gamma[i] * (double)sqrtf(un * delta2));
        float gl = (i == 0) ? 1.0f
                            : out->gamma[i - 1] / sqrtf(fmaf(out->gamma[i - 1], out->gamma[i - 1], 1.0f));
        float gn = (i + 1 == m) ? 0.0f
                                : out->gamma[i + 1] / sqrtf(fmaf(out->gamma[i + 1], out->gamma[i + 1], 1.0f));
        float dH = gl - gn;
        float first_t = sqrtf((float)(4 * un) * dH / delta2);
        float last_t = sqrtf((float)(2 * un) * dH / last_denom);
        float short_t = (i == 0) ? first_t : (first_t > last_t ? first_t : last_t);
        out->short_t[i] = short_t;
        total_t += short_t;
    }
    out->total_t = total_t;
}

/* ------------------------------------------------------------------------ */
/* a-7: Monte-Carlo times.  MultiQW.cpp:387,417-430.                          */
/* MT19937 (Matsumoto & Nishimura 1998) seeded with 29552825458725 mod 2^32; */
/* uniform_real_distribution<float>(s, 2s) as libstdc++ computes it:         */
/* u = float(x) * 2^-32 (clamped below 1), t = fma(u, s, s).                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t mt[624];
    int idx;
} oracle_mt19937;

void oracle_mt_seed(oracle_mt19937* g, uint32_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

uint32_t oracle_mt_next(oracle_mt19937* g)
{
    if (g->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static int cmp_float(const void* a, const void* b)
{
    float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

/* times[i*samples + j]; m==1 sorts and differences row 0 (:426-430). */
void oracle_sample_times(int m, int samples, const float* short_t, float* times)
{
    oracle_mt19937 g;
    oracle_mt_seed(&g, (uint32_t)(29552825458725ULL & 0xffffffffULL));
    for (int i = 0; i < m; i++) {
        float s = short_t[i];
        float w = 2 * s - s;
        for (int j = 0; j < samples; j++) {
            float u = (float)oracle_mt_next(&g) * 0x1p-32f;
            if (u >= 1.0f) u = nextafterf(1.0f, 0.0f);
            times[i * samples + j] = fmaf(u, w, s);
        }
    }
    if (m == 1) {
        qsort(times, (size_t)samples, sizeof(float), cmp_float);
        for (int k = samples - 1; k >= 1; k--) times[k] = times[k] - times[k - 1];
    }
}

/* ------------------------------------------------------------------------ */
/* a-8: Chebyshev coefficients of cos(s x) + sin(s x) on [-1,1]               */
/* (replaces ApproxTools RCF_odd_even at MultiQW.cpp:438-440).               */
/* Returns K (coefficients kept), or -1 if cap is too small.                 */
/* ------------------------------------------------------------------------ */
int oracle_cheb_coeffs(double s, double tol, double* c, int cap)
{
    if (!(s > 0.0)) {
        if (cap < 1) return -1;
        c[0] = 1.0;
        return 1;
    }
    int M = (int)(s + 12.0 * cbrt(s) + 40.0);
    M += M & 1;
    double* J = (double*)malloc(sizeof(double) * (size_t)(M + 2));
    J[M + 1] = 0.0;
    J[M] = 1e-300;
    for (int k = M; k >= 1; k--) {
        J[k - 1] = (2.0 * k / s) * J[k] - J[k + 1];
        if (fabs(J[k - 1]) > 1e250)
            for (int q = k - 1; q <= M + 1; q++) J[q] *= 1e-250;
    }
    double norm = J[0];
    for (int k = 2; k <= M; k += 2) norm += 2.0 * J[k];
    for (int k = 0; k <= M; k++) J[k] /= norm;
    /* tail sums of |c_k| from the top */
    int K = M + 1;
    double tail = 0.0;
    for (int k = M; k >= 1; k--) {
        double ck = (k == 0 ? 1.0 : 2.0) * fabs(J[k]);
        if (tail + ck > tol) break;
        tail += ck;
        K = k;
    }
    if (K > cap) {
        free(J);
        return -1;
    }
    for (int k = 0; k < K; k++) {
        double sign = ((k >> 1) & 1) ? -1.0 : 1.0;
        c[k] = (k == 0 ? 1.0 : 2.0) * sign * J[k];
    }
    free(J);
    return K;
}

/* ------------------------------------------------------------------------ */
/* a-10: Clenshaw_step, FP32, reference pass split and operation order.      */
/* MultiQW.cpp:88-173.  b2 <- (2/scale)(H_P - gamma H_G) b1 - b2 + coef psi   */
/* on one real half of length N = 2^n (n >= 4).                              */
/* ------------------------------------------------------------------------ */
void oracle_clenshaw_step_f32(const float* b1, float* b2, const float* hp, const float* psi, unsigned n,
                              float scale, float gamma, float coef)
{
    const uint32_t N = 1u << n;
    const int max_cache = 15;
    float ns = -2.0f * gamma / scale;
    float nsi = 1.0f / ns;
    float mnsi = -nsi;
    for (uint32_t i = 0; i < N; i += 16) {
        float b[16];
        for (int l = 0; l < 16; l++) {
            float q = (b1[i + l] * hp[i + l]) / gamma;
            b[l] = fmaf(b2[i + l], mnsi, -q);
        }
        for (int l = 0; l < 16; l++) b[l] += b1[i + (l ^ 1)];
        for (int l = 0; l < 16; l++) b[l] += b1[i + (l ^ 2)];
        for (int l = 0; l < 16; l++) b[l] += b1[i + (l ^ 4)];
        for (int l = 0; l < 16; l++) b[l] += b1[i + (l ^ 8)];
        if ((int)n > max_cache)
            for (uint32_t hb = 1u << max_cache; hb < N; hb <<= 1) {
                const float* p = b1 + (i ^ hb);
                for (int l = 0; l < 16; l++) b[l] += p[l];
            }
        for (int l = 0; l < 16; l++) b2[i + l] = b[l];
    }
    uint32_t end = N < (1u << max_cache) ? N : (1u << max_cache);
    for (uint32_t i = 0; i < N; i += 16) {
        float b[16];
        for (int l = 0; l < 16; l++) b[l] = b2[i + l];
        for (uint32_t hb = 16; hb < end; hb <<= 1) {
            const float* p = b1 + (i ^ hb);
            for (int l = 0; l < 16; l++) b[l] += p[l];
        }
        if (coef != 0) {
            for (int l = 0; l < 16; l++) b[l] = fmaf(b[l], ns, psi[i + l] * coef);
        } else {
            for (int l = 0; l < 16; l++) b[l] = b[l] * ns;
        }
        for (int l = 0; l < 16; l++) b2[i + l] = b[l];
    }
}

/* ------------------------------------------------------------------------ */
/* a-9: Clenshaw driver, FP32.  MultiQW.cpp:177-236.  psi is planar (2N);    */
/* the result overwrites psi.  scratch must hold 4N floats.                  */
/* ------------------------------------------------------------------------ */
void oracle_clenshaw_f32(const float* coeffs, int K, float* psi, const float* hp, unsigned n, float gamma,
                         float scale, int psi_real, float* scratch)
{
    const uint32_t N = 1u << n;
    float* b1 = scratch;
    float* b2 = scratch + 2 * (size_t)N;
    memset(b1, 0, sizeof(float) * 2 * N);
    memset(b2, 0, sizeof(float) * 2 * N);
    int first = 1;
    float im_coef = psi_real ? 0.0f : 1.0f;
    for (int r = K - 1; r > 0; --r) {
        if (!first) {
            if (r & 1) {
                oracle_clenshaw_step_f32(b1, b2, hp, psi + N, n, scale, gamma, -im_coef * coeffs[r]);
                oracle_clenshaw_step_f32(b1 + N, b2 + N, hp, psi, n, scale, gamma, coeffs[r]);
            } else {
                oracle_clenshaw_step_f32(b1, b2, hp, psi, n, scale, gamma, coeffs[r]);
                oracle_clenshaw_step_f32(b1 + N, b2 + N, hp, psi + N, n, scale, gamma, im_coef * coeffs[r]);
            }
        } else {
            if (r & 1) {
                if (!psi_real)
                    for (uint32_t k = 0; k < N; k++) b2[k] -= psi[N + k] * coeffs[r];
                for (uint32_t k = 0; k < N; k++) b2[N + k] += psi[k] * coeffs[r];
            } else {
                if (!psi_real)
                    for (uint32_t k = 0; k < N; k++) b2[N + k] += psi[N + k] * coeffs[r];
                for (uint32_t k = 0; k < N; k++) b2[k] += psi[k] * coeffs[r];
            }
        }
        first = 0;
        float* t = b1;
        b1 = b2;
        b2 = t;
    }
    oracle_clenshaw_step_f32(b1, b2, hp, psi, n, (float)(2.0 * scale), gamma, coeffs[0]);
    oracle_clenshaw_step_f32(b1 + N, b2 + N, hp, psi + N, n, (float)(2.0 * scale), gamma, im_coef * coeffs[0]);
    memcpy(psi, b2, sizeof(float) * 2 * N);
}

/* FP64 Clenshaw: the same recurrence, plain order, double everywhere.
 * psi planar double (2N); hp double (N).  scratch holds 4N doubles. */
void oracle_clenshaw_f64(const double* coeffs, int K, double* psi, const double* hp, unsigned n, double gamma,
                         double scale, double* scratch)
{
    const uint32_t N = 1u << n;
    double* b1 = scratch;
    double* b2 = scratch + 2 * (size_t)N;
    memset(b1, 0, sizeof(double) * 2 * N);
    memset(b2, 0, sizeof(double) * 2 * N);
    for (int r = K - 1; r >= 0; --r) {
        double mult = (r == 0) ? 1.0 / scale : 2.0 / scale;
        double cre = (r & 1) ? 0.0 : coeffs[r];
        double cim = (r & 1) ? coeffs[r] : 0.0;
        for (int half = 0; half < 2; half++) {
            const double* x = b1 + half * N;
            double* y = b2 + half * N;
            for (uint32_t k = 0; k < N; k++) {
                double acc = 0.0;
                for (unsigned i = 0; i < n; i++) acc += x[k ^ (1u << i)];
                double inj = half == 0 ? cre * psi[k] - cim * psi[N + k] : cre * psi[N + k] + cim * psi[k];
                y[k] = mult * (hp[k] * x[k] - gamma * acc) - y[k] + inj;
            }
        }
        double* t = b1;
        b1 = b2;
        b2 = t;
    }
    memcpy(psi, b1, sizeof(double) * 2 * N);
}

/* ------------------------------------------------------------------------ */
/* a-11 + driver: the per-problem loop of main, MultiQW.cpp:296-455.          */
/* recs: count problems of n(n+1)/2 doubles.  Outputs per problem: the mean  */
/* ground-state probability and total_t.  samples = 100 in the reference.    */
/* precision 0: FP32 restatement (reference order); 1: FP64 state.          */
/*                                                                            */
/* Work is split in two phases so that threads can share one problem:        */
/*  1. per problem: unpack, energies, schedule, Monte-Carlo times            */
/*     (MultiQW.cpp:303-430);                                                 */
/*  2. per (problem, run of samples): the stage loop (:432-451).  For m > 1  */
/*     every sample restarts from |+> (:447-451), so samples are independent */
/*     and runs of them go to different threads; an m == 1 problem chains    */
/*     its samples (:426-430) and is one run.                                */
/* The mean over samples is taken after both phases (:453), in the           */
/* reference's float sum order, so results do not depend on the threads.     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, m, samples, precision;
    int norm_f64;           /* FP32 path: 0 = the reference's float norm (Eigen order), 1 = FP64 sum */
    double tol;
    const double* rec;
    const float* gamma_ov;  /* nullable: m per-stage gammas replacing the heuristic */
    const float* short_t_ov;/* nullable: m per-stage times replacing the heuristic */
    /* phase 1 results */
    float* hp;
    oracle_energy_info info;
    oracle_schedule sch;
    float* times;
    float* sp;              /* samples per-sample probabilities */
} oracle_problem;

typedef struct {
    oracle_problem* pr;
    int j0, j1;             /* samples [j0, j1) */
} oracle_run;

static void problem_setup(oracle_problem* pr)
{
    const int n = pr->n, m = pr->m;
    const uint32_t N = 1u << n;
    float J[ORACLE_MAX_N * ORACLE_MAX_N], h[ORACLE_MAX_N];
    oracle_unpack(n, pr->rec, J, h);
    pr->hp = (float*)malloc(sizeof(float) * N);
    oracle_energy_compat(n, J, h, pr->hp, NULL, &pr->info);
    oracle_schedule_heuristic(n, m, J, h, &pr->sch);
    if (pr->gamma_ov && pr->short_t_ov) {
        float tot = 0;
        for (int i = 0; i < m; i++) {
            pr->sch.gamma[i] = pr->gamma_ov[i];
            pr->sch.short_t[i] = pr->short_t_ov[i];
            tot += pr->sch.short_t[i];
        }
        pr->sch.total_t = tot;
    }
    pr->times = (float*)malloc(sizeof(float) * (size_t)m * pr->samples);
    oracle_sample_times(m, pr->samples, pr->sch.short_t, pr->times);
    pr->sp = (float*)malloc(sizeof(float) * pr->samples);
}

static void problem_free(oracle_problem* pr)
{
    free(pr->hp);
    free(pr->times);
    free(pr->sp);
}

/* Samples [j0, j1) of one problem, from |+> (an m == 1 problem runs all of
 * its samples here, chaining the state). */
static void run_samples(const oracle_run* rn)
{
    const oracle_problem* pr = rn->pr;
    const int n = pr->n, m = pr->m, samples = pr->samples;
    const uint32_t N = 1u << n;
    const float* hp = pr->hp;
    const float* times = pr->times;
    float* sp = pr->sp;
    int cap = 4096;
    double* cd = (double*)malloc(sizeof(double) * cap);
    const float psi0 = (float)(1 / sqrt((double)N));
    if (pr->precision == 0) {
        float* psi = (float*)malloc(sizeof(float) * 2 * N);
        float* scratch = (float*)malloc(sizeof(float) * 4 * N);
        float* cf = (float*)malloc(sizeof(float) * cap);
        float* sq = (float*)malloc(sizeof(float) * 2 * N);
        for (uint32_t k = 0; k < N; k++) {
            psi[k] = psi0;
            psi[N + k] = 0.0f;
        }
        int psi_real = 1;
        for (int j = rn->j0; j < rn->j1; j++) {
            for (int i = 0; i < m; i++) {
                float gamma = pr->sch.gamma[i];
                float onenorm = fmaf(gamma, (float)(unsigned)n, pr->info.e_abs);
                double scale = (double)(onenorm * times[i * samples + j]);
                int K = oracle_cheb_coeffs(scale, pr->tol, cd, cap);
                for (int k = 0; k < K; k++) cf[k] = (float)cd[k];
                oracle_clenshaw_f32(cf, K, psi, hp, (unsigned)n, gamma, onenorm, psi_real, scratch);
                for (uint32_t k = 0; k < 2 * N; k++) sq[k] = psi[k] * psi[k];
                float nrm;
                if (pr->norm_f64) {
                    double s2 = 0.0;
                    for (uint32_t k = 0; k < 2 * N; k++) s2 += (double)sq[k];
                    nrm = (float)sqrt(s2);
                } else {
                    nrm = sqrtf(oracle_eigen_sum_f32(sq, 2 * (long)N));
                }
                for (uint32_t k = 0; k < 2 * N; k++) psi[k] = psi[k] / nrm;
                psi_real = 0;
            }
            float re = psi[pr->info.eloc], im = psi[pr->info.eloc + N];
            sp[j] = fmaf(re, re, im * im);
            if (m != 1) {
                for (uint32_t k = 0; k < N; k++) {
                    psi[k] = psi0;
                    psi[N + k] = 0.0f;
                }
                psi_real = 1;
            }
        }
        free(psi);
        free(scratch);
        free(cf);
        free(sq);
    } else {
        double* psi = (double*)malloc(sizeof(double) * 2 * N);
        double* scratch = (double*)malloc(sizeof(double) * 4 * N);
        double* hpd = (double*)malloc(sizeof(double) * N);
        for (uint32_t k = 0; k < N; k++) hpd[k] = (double)hp[k];
        const double p0 = 1 / sqrt((double)N);
        for (uint32_t k = 0; k < N; k++) {
            psi[k] = p0;
            psi[N + k] = 0.0;
        }
        for (int j = rn->j0; j < rn->j1; j++) {
            for (int i = 0; i < m; i++) {
                double gamma = pr->sch.gamma[i];
                float onenorm = fmaf(pr->sch.gamma[i], (float)(unsigned)n, pr->info.e_abs);
                double scale = (double)(onenorm * times[i * samples + j]);
                int K = oracle_cheb_coeffs(scale, pr->tol, cd, cap);
                oracle_clenshaw_f64(cd, K, psi, hpd, (unsigned)n, gamma, (double)onenorm, scratch);
                double s2 = 0.0;
                for (uint32_t k = 0; k < 2 * N; k++) s2 += psi[k] * psi[k];
                double nrm = sqrt(s2);
                for (uint32_t k = 0; k < 2 * N; k++) psi[k] /= nrm;
            }
            double re = psi[pr->info.eloc], im = psi[pr->info.eloc + N];
            sp[j] = (float)(re * re + im * im);
            if (m != 1)
                for (uint32_t k = 0; k < N; k++) {
                    psi[k] = p0;
                    psi[N + k] = 0.0;
                }
        }
        free(psi);
        free(scratch);
        free(hpd);
    }
    free(cd);
}

typedef struct {
    void* items;
    size_t item_size;
    void (*fn)(void*);
    int count;
    int next;
    pthread_mutex_t mu;
} job_queue;

static void* worker(void* arg)
{
    job_queue* q = (job_queue*)arg;
    for (;;) {
        pthread_mutex_lock(&q->mu);
        int i = q->next++;
        pthread_mutex_unlock(&q->mu);
        if (i >= q->count) break;
        q->fn((char*)q->items + (size_t)i * q->item_size);
    }
    return NULL;
}

/* Runs fn over count items on up to `threads` threads (the run.sh fan-out of
 * MultiQW.cpp, in one process). */
static void parallel_for(void* items, size_t item_size, int count, void (*fn)(void*), int threads)
{
    if (threads <= 1 || count <= 1) {
        for (int i = 0; i < count; i++) fn((char*)items + (size_t)i * item_size);
        return;
    }
    job_queue q;
    q.items = items;
    q.item_size = item_size;
    q.fn = fn;
    q.count = count;
    q.next = 0;
    pthread_mutex_init(&q.mu, NULL);
    pthread_t th[256];
    int T = threads > 256 ? 256 : threads;
    if (T > count) T = count;
    for (int t = 0; t < T; t++) pthread_create(&th[t], NULL, worker, &q);
    for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&q.mu);
}

static void setup_fn(void* p) { problem_setup((oracle_problem*)p); }
static void run_fn(void* p) { run_samples((const oracle_run*)p); }

/* Solve count problems on up to `threads` threads.  sample_probs (nullable)
 * receives count*samples per-sample probabilities.  gamma/short_t (nullable)
 * hold count*m per-stage values replacing the heuristic schedule.  norm_f64
 * (FP32 path): 0 = each stage's norm as the reference takes it (MultiQW.cpp:443,
 * a float sum in Eigen's order), 1 = the same squares summed in double, which
 * isolates the float sum's own rounding from the rest of the arithmetic.
 * Returns 0, or -1 on bad args. */
int oracle_solve_ex(int n, int m, const double* recs, int count, int samples, double tol, int precision,
                    int threads, float* prob, float* total_t, float* sample_probs, const float* gamma,
                    const float* short_t, int norm_f64)
{
    if (n < 4 || n > 26 || m < 1 || m > ORACLE_MAX_M || count < 0 || samples < 1) return -1;
    if (count == 0) return 0;
    const int P = n * (n + 1) / 2;
    oracle_problem* prs = (oracle_problem*)calloc((size_t)count, sizeof(oracle_problem));
    for (int c = 0; c < count; c++) {
        prs[c].n = n;
        prs[c].m = m;
        prs[c].samples = samples;
        prs[c].precision = precision;
        prs[c].norm_f64 = norm_f64;
        prs[c].tol = tol;
        prs[c].rec = recs + (size_t)c * P;
        prs[c].gamma_ov = gamma ? gamma + (size_t)c * m : NULL;
        prs[c].short_t_ov = short_t ? short_t + (size_t)c * m : NULL;
    }
    parallel_for(prs, sizeof(oracle_problem), count, setup_fn, threads);
    /* runs of samples: enough to give every thread work, whole problems for m == 1 */
    int per = samples;
    if (m != 1 && threads > count) {
        int want = (threads + count - 1) / count; /* runs per problem */
        per = (samples + want - 1) / want;
    }
    int nruns = 0;
    for (int c = 0; c < count; c++) nruns += (samples + per - 1) / per;
    oracle_run* runs = (oracle_run*)malloc(sizeof(oracle_run) * (size_t)nruns);
    int r = 0;
    for (int c = 0; c < count; c++)
        for (int j = 0; j < samples; j += per) {
            runs[r].pr = &prs[c];
            runs[r].j0 = j;
            runs[r].j1 = j + per < samples ? j + per : samples;
            r++;
        }
    parallel_for(runs, sizeof(oracle_run), nruns, run_fn, threads);
    for (int c = 0; c < count; c++) {
        prob[c] = oracle_eigen_sum_f32(prs[c].sp, samples) / samples;
        total_t[c] = prs[c].sch.total_t;
        if (sample_probs) memcpy(sample_probs + (size_t)c * samples, prs[c].sp, sizeof(float) * samples);
        problem_free(&prs[c]);
    }
    free(runs);
    free(prs);
    return 0;
}

int oracle_solve_sched(int n, int m, const double* recs, int count, int samples, double tol, int precision,
                       int threads, float* prob, float* total_t, float* sample_probs, const float* gamma,
                       const float* short_t)
{
    return oracle_solve_ex(n, m, recs, count, samples, tol, precision, threads, prob, total_t, sample_probs, gamma,
                           short_t, 0);
}

int oracle_solve(int n, int m, const double* recs, int count, int samples, double tol, int precision, int threads,
                 float* prob, float* total_t, float* sample_probs)
{
    return oracle_solve_sched(n, m, recs, count, samples, tol, precision, threads, prob, total_t, sample_probs, NULL,
                              NULL);
}

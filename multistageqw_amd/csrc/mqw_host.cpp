// Host-side pieces of the walk: record unpacking, the heuristic schedule, the
// Monte-Carlo time draws, Chebyshev coefficients and the SK generator.
//
// Reference: Asa-Hopkins/MultiStageQW MultiQW.cpp.  Float semantics follow the
// reference's types; where its g++ -O3 -march=native build (README.md:20)
// contracts a*b+c into a fused multiply-add, the fused form is written out
// with fmaf so the result does not depend on this compiler's choices.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "mqw_internal.h"

namespace mqw {

static const double kPi = 3.1415926535897932384626;

// Float sum in the order of Eigen's linear vectorised reduction with
// 16-float packets: two packet accumulators over pairs of packets, one
// trailing packet, a halving tree over the 16 lanes, then the scalar tail.
// MultiQW.cpp takes .sum()/.norm() of Eigen float arrays (:332,375,395,443,454).
float eigen_sum_f32(const float* x, long size)
{
    constexpr int P = 16;
    if (size <= 0) return 0.0f;
    const long pairs_end = (size / (2 * P)) * (2 * P);
    const long packs_end = (size / P) * P;
    if (packs_end == 0) {
        float r = x[0];
        for (long i = 1; i < size; i++) r += x[i];
        return r;
    }
    float acc0[P], acc1[P];
    for (int l = 0; l < P; l++) acc0[l] = x[l];
    if (packs_end > P) {
        for (int l = 0; l < P; l++) acc1[l] = x[P + l];
        for (long i = 2 * P; i < pairs_end; i += 2 * P)
            for (int l = 0; l < P; l++) {
                acc0[l] += x[i + l];
                acc1[l] += x[i + P + l];
            }
        for (int l = 0; l < P; l++) acc0[l] += acc1[l];
        if (packs_end > pairs_end)
            for (int l = 0; l < P; l++) acc0[l] += x[pairs_end + l];
    }
    for (int w = P / 2; w > 0; w /= 2)
        for (int l = 0; l < w; l++) acc0[l] += acc0[l + w];
    float r = acc0[0];
    for (long i = packs_end; i < size; i++) r += x[i];
    return r;
}

// MultiQW.cpp:303-323: lower triangle in row-major k-order (i = 1..n-1, j < i),
// then h; J is symmetrised and halved, every value rounded to float.
void unpack(int n, const double* rec, float* J, float* h)
{
    std::fill(J, J + n * n, 0.0f);
    int k = 0;
    for (int i = 1; i < n; i++)
        for (int j = 0; j < i; j++) {
            const float v = (float)rec[k++];
            J[i * n + j] = (v + 0.0f) / 2;
            J[j * n + i] = (0.0f + v) / 2;
        }
    for (int i = 0; i < n; i++) h[i] = (float)rec[n * (n - 1) / 2 + i];
}

double e_init_compat(int n, const float* J, const float* h)
{
    // MultiQW.cpp:332: float J.sum() - float h.sum(), widened to double.
    return (double)(eigen_sum_f32(J, (long)n * n) - eigen_sum_f32(h, n));
}

double e_init_exact(int n, const float* J, const float* h)
{
    double e = 0.0;
    for (int i = 0; i < n; i++) {
        e -= (double)h[i];
        for (int j = 0; j < i; j++) e += 2.0 * (double)J[i * n + j];
    }
    return e;
}

// Inverse standard normal CDF: Wichura, "Algorithm AS 241: The percentage
// points of the normal distribution", Applied Statistics 37 (1988), PPND16.
double ppnd16(double p)
{
    static const double A[8] = {3.3871328727963666080e+0, 1.3314166789178437745e+2, 1.9715909503065514427e+3,
                                1.3731693765509461125e+4, 4.5921953931549871457e+4, 6.7265770927008700853e+4,
                                3.3430575583588128105e+4, 2.5090809287301226727e+3};
    static const double B[8] = {1.0, 4.2313330701600911252e+1, 6.8718700749205790830e+2, 5.3941960214247511077e+3,
                                2.1213794301586595867e+4, 3.9307895800092710610e+4, 2.8729085735721942674e+4,
                                5.2264952788528545610e+3};
    static const double Cc[8] = {1.42343711074968357734e+0, 4.63033784615654529590e+0, 5.76949722146069140550e+0,
                                 3.64784832476320460504e+0, 1.27045825245236838258e+0, 2.41780725177450611770e-1,
                                 2.27238449892691845833e-2, 7.74545014278341407640e-4};
    static const double D[8] = {1.0, 2.05319162663775882187e+0, 1.67638483018380384940e+0, 6.89767334985100004550e-1,
                                1.48103976427480074590e-1, 1.51986665636164571966e-2, 5.47593808499534494600e-4,
                                1.05075007164441684324e-9};
    static const double Ee[8] = {6.65790464350110377720e+0, 5.46378491116411436990e+0, 1.78482653991729133580e+0,
                                 2.96560571828504891230e-1, 2.65321895265761230930e-2, 1.24266094738807843860e-3,
                                 2.71155556874348757815e-5, 2.01033439929228813265e-7};
    static const double F[8] = {1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1, 1.48753612908506148525e-2,
                                7.86869131145613259100e-4, 1.84631831751005468180e-5, 1.42151175831644588870e-7,
                                2.04426310338993978564e-15};
    auto ratio = [](const double* num, const double* den, double r) {
        double a = num[7], b = den[7];
        for (int i = 6; i >= 0; i--) {
            a = a * r + num[i];
            b = b * r + den[i];
        }
        return a / b;
    };
    const double q = p - 0.5;
    if (fabs(q) <= 0.425) return q * ratio(A, B, 0.180625 - q * q);
    double r = sqrt(-log(q < 0 ? p : 1.0 - p));
    const double v = (r <= 5.0) ? ratio(Cc, D, r - 1.6) : ratio(Ee, F, r - 5.0);
    return q < 0 ? -v : v;
}

// MultiQW.cpp:365-423.  my_normcdfinvf(x) evaluates -Phi^{-1}(x) (its erfcinv
// returns the negated value, :77); here Phi^{-1} is AS241 in FP64, rounded.
// Cornish-Fisher correction of a standard normal quantile z for a
// distribution of skewness s and excess kurtosis k (the terms to second order):
// x = z + (z^2 - 1) s/6 + (z^3 - 3z) k/24 - (2z^3 - 5z) s^2/36.
static double cornish_fisher_shift(double z, double s, double k)
{
    return (z * z - 1) * s / 6 + (z * z * z - 3 * z) * k / 24 - (2 * z * z * z - 5 * z) * s * s / 36;
}

void schedule_heuristic(int n, int m, const float* J, const float* h, mqw_schedule* out, const double* shape)
{
    const unsigned N = 1u << n, un = (unsigned)n, um = (unsigned)m;
    float b = (float)(-ppnd16((double)(1 / (float)N)));
    const float e_m = 0.577215664901532860;
    const float e = 2.718281828459045;
    float c = (float)(-ppnd16((double)(1 / (e * N))));
    if (shape) {
        // depths below the mean of the lower-tail quantiles 1/N and 1/(eN) of
        // the energies' own distribution instead of a Gaussian's (b, c are
        // -z at those probabilities); zero moments leave b and c unchanged
        b = (float)((double)b - cornish_fisher_shift(-(double)b, shape[0], shape[1]));
        c = (float)((double)c - cornish_fisher_shift(-(double)c, shape[0], shape[1]));
    }
    const float a = fmaf(1 - e_m, b, e_m * c);
    float jj[kMaxN * kMaxN], hh[kMaxN];
    for (int i = 0; i < n * n; i++) jj[i] = J[i] * J[i];
    for (int i = 0; i < n; i++) hh[i] = h[i] * h[i];
    const float sjj = eigen_sum_f32(jj, (long)n * n), shh = eigen_sum_f32(hh, n);
    const float hp2 = fmaf(2, sjj, shh);
    const float delta2 = fmaf(16, sjj, 4 * shh);
    out->n = n;
    out->m = m;
    out->a = a;
    out->hp2 = hp2;
    out->delta2 = delta2;
    for (int i = 0; i < m; i++)
        out->gamma[i] = (float)((double)(a * sqrtf(hp2)) / tan(kPi * (i + 1) / (2 * um + 2)) / un);
    auto weight = [](float g) { return g / sqrtf(fmaf(g, g, 1.0f)); };
    float total = 0;
    for (int i = 0; i < m; i++) {
        const float last_denom = (float)(sqrt(2 / kPi) * (double)out->gamma[i] * (double)sqrtf(un * delta2));
        const float g_prev = (i == 0) ? 1.0f : weight(out->gamma[i - 1]);
        const float g_next = (i + 1 == m) ? 0.0f : weight(out->gamma[i + 1]);
        const float dH = g_prev - g_next;
        const float first_t = sqrtf((float)(4 * un) * dH / delta2);
        const float last_t = sqrtf((float)(2 * un) * dH / last_denom);
        const float st = (i == 0) ? first_t : std::max(first_t, last_t);
        out->short_t[i] = st;
        total += st;
    }
    out->total_t = total;
}

// MultiQW.cpp:387,417-430.  std::mt19937 re-seeded per problem with
// 29552825458725 (mod 2^32); libstdc++'s uniform_real_distribution<float>
// draws u = float(x) * 2^-32 (kept below 1) and returns u * (2s - s) + s,
// fused by the reference build.
void sample_times(int m, int samples, const float* short_t, float* times)
{
    std::mt19937 gen((uint32_t)(29552825458725ULL & 0xffffffffULL));
    for (int i = 0; i < m; i++) {
        const float s = short_t[i], w = 2 * s - s;
        for (int j = 0; j < samples; j++) {
            float u = (float)gen() * 0x1p-32f;
            if (u >= 1.0f) u = nextafterf(1.0f, 0.0f);
            times[i * samples + j] = fmaf(u, w, s);
        }
    }
    if (m == 1) {
        std::sort(times, times + samples);
        for (int k = samples - 1; k >= 1; k--) times[k] -= times[k - 1];
    }
}

// Jacobi-Anger: cos(sx) + sin(sx) = J0(s) + sum_{k>=1} 2 (-1)^floor(k/2) J_k(s) T_k(x).
// J_k(s) by Miller's backward recurrence J_{k-1} = (2k/s) J_k - J_{k+1} from an
// order well past the turning point k = s, normalised by J_0 + 2 sum J_{2k} = 1.
// K is the first order whose tail sum of |c_k| is <= tol.
int cheb_coeffs_vec(double s, double tol, std::vector<double>& c)
{
    if (!(s > 0.0)) {
        c.assign(1, 1.0);
        return 1;
    }
    int top = (int)(s + 12.0 * cbrt(s) + 40.0);
    top += top & 1;
    thread_local std::vector<double> j;
    j.resize((size_t)top + 2);
    j[top + 1] = 0.0;
    j[top] = 1e-300;
    for (int k = top; k > 0; k--) {
        j[k - 1] = (2.0 * k / s) * j[k] - j[k + 1];
        if (fabs(j[k - 1]) > 1e250)
            for (int q = k - 1; q <= top; q++) j[q] *= 1e-250;
    }
    double norm = j[0];
    for (int k = 2; k <= top; k += 2) norm += 2.0 * j[k];
    const double inv = 1.0 / norm;
    int K = top + 1;
    double tail = 0.0;
    for (int k = top; k > 0; k--) {
        const double ck = 2.0 * fabs(j[k] * inv);
        if (tail + ck > tol) break;
        tail += ck;
        K = k;
    }
    c.resize((size_t)K);
    for (int k = 0; k < K; k++) {
        const double sgn = ((k >> 1) & 1) ? -1.0 : 1.0;
        c[k] = (k == 0 ? 1.0 : 2.0) * sgn * (j[k] * inv);
    }
    return K;
}

int cheb_coeffs(double s, double tol, double* c, int cap)
{
    std::vector<double> v;
    const int K = cheb_coeffs_vec(s, tol, v);
    if (K > cap) return -1;
    std::copy(v.begin(), v.end(), c);
    return K;
}

// SK instances: SplitMix64 (Steele, Lea, Flood 2014) seeds xoshiro256**
// (Blackman & Vigna 2018); N(0,1) by Marsaglia's polar method.
namespace {
struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed)
    {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            s[i] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform_pm1() { return (double)(next() >> 11) * 0x1p-52 - 1.0; }
};
} // namespace

void generate_sk(int n, int count, uint64_t seed, double* out)
{
    Xoshiro g(seed);
    const size_t total = (size_t)count * (size_t)(n * (n + 1) / 2);
    bool have = false;
    double spare = 0.0;
    for (size_t i = 0; i < total; i++) {
        if (have) {
            out[i] = spare;
            have = false;
            continue;
        }
        double u, v, q;
        do {
            u = g.uniform_pm1();
            v = g.uniform_pm1();
            q = u * u + v * v;
        } while (q >= 1.0 || q == 0.0);
        const double f = sqrt(-2.0 * log(q) / q);
        out[i] = u * f;
        spare = v * f;
        have = true;
    }
}

} // namespace mqw

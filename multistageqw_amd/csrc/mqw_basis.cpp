// Compression of the shared first-stage basis (see GramJob in mqw_internal.h).
//
// For one instance and one parity class (k even: the real parts of the
// samples' first-stage states, k odd: the imaginary parts), the samples'
// states are the rows of A V, with A the samples' coefficients and V the
// class's basis vectors.  With G = V V^T = Z diag(w) Z^T (cyclic Jacobi), the
// matrix M = A Z diag(sqrt(w)) has M M^T = A G A^T, so its singular values and
// left singular vectors are those of A V itself.  A one-sided Jacobi SVD of M
// (column rotations, accurate for small singular values) gives columns
// m_i = sigma_i u_i.  Keeping the top R columns reconstructs sample j to
// within sqrt(sum_{i >= R} m_i[j]^2) per class; R is the smallest rank that
// keeps every sample's discarded part, both classes together, within rel_tol
// of its whole norm.  The kept directions are the unit
// vectors Q_r = (1 / sigma_r^2) sum_j m_r[j] (A V)_j, i.e. B = m_r^T A / sigma_r^2,
// and sample j's coordinates are m_r[j].
#include <math.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "mqw_internal.h"

namespace mqw {

namespace {

// Eigen-decomposition of the symmetric n x n matrix a (row-major, destroyed):
// eigenvalues w, eigenvectors as the columns of z (row-major).  Householder
// reduction to tridiagonal form T = Q^T A Q, then implicit symmetric QR steps
// with Wilkinson shifts on T, every Givens rotation also applied to Q (Golub
// and Van Loan, Matrix Computations, algorithms 8.3.1 and 8.3.3): O(n^3) in
// all, where cyclic Jacobi takes O(n^3) per sweep.
void sym_eig(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& z)
{
    auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
    auto Z = [&](int i, int j) -> double& { return z[(size_t)i * n + j]; };
    z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Z(i, i) = 1.0;
    std::vector<double> v(n), p(n);
    // 1. Householder: column k below the subdiagonal to zero, k = 0 .. n - 3
    for (int k = 0; k + 2 < n; k++) {
        double sig = 0.0;
        for (int i = k + 1; i < n; i++) sig += A(i, k) * A(i, k);
        const double xn = sqrt(sig);
        if (xn == 0.0) continue;
        const double alpha = A(k + 1, k) >= 0.0 ? -xn : xn;
        double vn = 0.0;
        for (int i = k + 1; i < n; i++) {
            v[i] = A(i, k) - (i == k + 1 ? alpha : 0.0);
            vn += v[i] * v[i];
        }
        vn = sqrt(vn);
        if (vn == 0.0) continue;
        for (int i = k + 1; i < n; i++) v[i] /= vn;
        // trailing block B <- H B H, H = I - 2 v v^T: with p = B v and
        // q = p - (v^T p) v, B <- B - 2 v q^T - 2 q v^T
        double kk = 0.0;
        for (int i = k + 1; i < n; i++) {
            double acc = 0.0;
            for (int j = k + 1; j < n; j++) acc += A(i, j) * v[j];
            p[i] = acc;
            kk += v[i] * acc;
        }
        for (int i = k + 1; i < n; i++) p[i] -= kk * v[i];
        for (int i = k + 1; i < n; i++)
            for (int j = k + 1; j < n; j++) A(i, j) -= 2.0 * (v[i] * p[j] + p[i] * v[j]);
        A(k + 1, k) = A(k, k + 1) = alpha;
        for (int i = k + 2; i < n; i++) A(i, k) = A(k, i) = 0.0;
        // Q <- Q H
        for (int r = 0; r < n; r++) {
            double acc = 0.0;
            for (int j = k + 1; j < n; j++) acc += Z(r, j) * v[j];
            for (int j = k + 1; j < n; j++) Z(r, j) -= 2.0 * acc * v[j];
        }
    }
    // 2. implicit QR steps on the unreduced trailing block [lo, hi]; a
    //    rotation in the plane (k, k + 1) touches rows and columns k - 1 .. k + 2
    const double eps = 2.220446049250313e-16;
    auto negligible = [&](int i) { return fabs(A(i + 1, i)) <= eps * (fabs(A(i, i)) + fabs(A(i + 1, i + 1))); };
    int hi = n - 1, steps = 0;
    while (hi > 0 && steps < 64 * n) {
        if (negligible(hi - 1)) {
            A(hi, hi - 1) = A(hi - 1, hi) = 0.0;
            hi--;
            continue;
        }
        int lo = hi - 1;
        while (lo > 0 && !negligible(lo - 1)) lo--;
        // Wilkinson shift: the eigenvalue of the trailing 2 x 2 closer to T(hi, hi)
        const double dd = 0.5 * (A(hi - 1, hi - 1) - A(hi, hi)), ee = A(hi, hi - 1);
        const double den = dd + (dd >= 0.0 ? 1.0 : -1.0) * sqrt(dd * dd + ee * ee);
        const double mu = A(hi, hi) - (den != 0.0 ? ee * ee / den : 0.0);
        double x = A(lo, lo) - mu, y = A(lo + 1, lo);
        for (int k = lo; k < hi; k++) {
            // [c s; -s c]^T [x; y] = [r; 0]
            double c = 1.0, s = 0.0;
            if (y != 0.0) {
                if (fabs(y) > fabs(x)) {
                    const double tau = -x / y;
                    s = 1.0 / sqrt(1.0 + tau * tau);
                    c = s * tau;
                } else {
                    const double tau = -y / x;
                    c = 1.0 / sqrt(1.0 + tau * tau);
                    s = c * tau;
                }
            }
            const int j0 = std::max(lo, k - 1), j1 = std::min(hi, k + 2);
            for (int j = j0; j <= j1; j++) { // rows k, k + 1: G^T T
                const double t1 = A(k, j), t2 = A(k + 1, j);
                A(k, j) = c * t1 - s * t2;
                A(k + 1, j) = s * t1 + c * t2;
            }
            for (int i = j0; i <= j1; i++) { // columns k, k + 1: T G
                const double t1 = A(i, k), t2 = A(i, k + 1);
                A(i, k) = c * t1 - s * t2;
                A(i, k + 1) = s * t1 + c * t2;
            }
            for (int r = 0; r < n; r++) { // Q G
                const double t1 = Z(r, k), t2 = Z(r, k + 1);
                Z(r, k) = c * t1 - s * t2;
                Z(r, k + 1) = s * t1 + c * t2;
            }
            if (k + 1 < hi) {
                x = A(k + 1, k);
                y = A(k + 2, k);
            }
        }
        steps++;
    }
    w.resize(n);
    for (int i = 0; i < n; i++) w[i] = A(i, i);
}

// One-sided Jacobi: rotates the columns of m (rows x cols, column-major)
// until they are mutually orthogonal.  Columns below 1e-12 of the largest
// column's norm (far below any kept direction) are left alone: the Gram sums
// give them no accurate direction, and rotating them only costs sweeps.
void orthogonalise_columns(int rows, int cols, std::vector<double>& m)
{
    double big = 0.0;
    for (int p = 0; p < cols; p++) {
        double s = 0.0;
        for (int r = 0; r < rows; r++) s += m[(size_t)p * rows + r] * m[(size_t)p * rows + r];
        big = std::max(big, s);
    }
    const double floor2 = 1e-24 * big;
    for (int sweep = 0; sweep < 64; sweep++) {
        bool rotated = false;
        for (int p = 0; p < cols - 1; p++)
            for (int q = p + 1; q < cols; q++) {
                double* mp = &m[(size_t)p * rows];
                double* mq = &m[(size_t)q * rows];
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int r = 0; r < rows; r++) {
                    al += mp[r] * mp[r];
                    be += mq[r] * mq[r];
                    ga += mp[r] * mq[r];
                }
                if (al < floor2 || be < floor2) continue;
                if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int r = 0; r < rows; r++) {
                    const double x = mp[r], y = mq[r];
                    mp[r] = c * x - s * y;
                    mq[r] = s * x + c * y;
                }
            }
        if (!rotated) break;
    }
}

struct ClassSvd {
    int nc = 0;
    std::vector<int> ks;     // basis indices of the class
    std::vector<double> m;   // samples x nc, column-major, columns by decreasing norm
    std::vector<double> sig; // column norms
};

void class_svd(int kc, int cls, const double* gram, const std::vector<const double*>& coefs, ClassSvd& out)
{
    const int S = (int)coefs.size();
    for (int k = cls; k < kc; k += 2) out.ks.push_back(k);
    const int nc = out.nc = (int)out.ks.size();
    std::vector<double> g((size_t)nc * nc);
    for (int a = 0; a < nc; a++)
        for (int b = 0; b < nc; b++) {
            const uint32_t k1 = (uint32_t)std::min(out.ks[a], out.ks[b]), k2 = (uint32_t)std::max(out.ks[a], out.ks[b]);
            g[(size_t)a * nc + b] = gram[gram_pair_start(k2) + k1 / 2];
        }
    std::vector<double> w, z;
    sym_eig(nc, g, w, z);
    // M = A Z diag(sqrt(w)), column-major
    std::vector<double> m((size_t)S * nc, 0.0);
    for (int i = 0; i < nc; i++) {
        const double sw = w[i] > 0.0 ? sqrt(w[i]) : 0.0;
        for (int j = 0; j < S; j++) {
            double acc = 0.0;
            for (int a = 0; a < nc; a++) acc += coefs[j][out.ks[a]] * z[(size_t)a * nc + i];
            m[(size_t)i * S + j] = acc * sw;
        }
    }
    orthogonalise_columns(S, nc, m);
    std::vector<double> sig(nc);
    for (int i = 0; i < nc; i++) {
        double s2 = 0.0;
        for (int j = 0; j < S; j++) s2 += m[(size_t)i * S + j] * m[(size_t)i * S + j];
        sig[i] = sqrt(s2);
    }
    std::vector<int> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sig[x] > sig[y]; });
    out.m.assign((size_t)S * nc, 0.0);
    out.sig.resize(nc);
    for (int i = 0; i < nc; i++) {
        std::copy(&m[(size_t)order[i] * S], &m[(size_t)order[i] * S] + S, &out.m[(size_t)i * S]);
        out.sig[i] = sig[order[i]];
    }
}

// Smallest rank R (the same for both classes) keeping every sample's
// discarded part, over both classes, within rel_tol of its whole norm.
int joint_rank(const ClassSvd (&c)[2], int S, double rel_tol)
{
    std::vector<double> tot(S, 0.0), tail[2];
    for (int cls = 0; cls < 2; cls++) {
        tail[cls].assign(S, 0.0);
        for (int i = 0; i < c[cls].nc; i++)
            for (int j = 0; j < S; j++) tail[cls][j] += c[cls].m[(size_t)i * S + j] * c[cls].m[(size_t)i * S + j];
        for (int j = 0; j < S; j++) tot[j] += tail[cls][j];
    }
    const int nmax = std::max(c[0].nc, c[1].nc);
    for (int R = 0; R <= nmax; R++) {
        bool ok = true;
        for (int j = 0; j < S && ok; j++) ok = tail[0][j] + tail[1][j] <= rel_tol * rel_tol * tot[j];
        if (ok) return R;
        for (int cls = 0; cls < 2; cls++)
            if (R < c[cls].nc)
                for (int j = 0; j < S; j++) tail[cls][j] -= c[cls].m[(size_t)R * S + j] * c[cls].m[(size_t)R * S + j];
    }
    return nmax;
}

} // namespace

bool compress_first_stage(int kc, const double* gram, const std::vector<const double*>& coefs, double rel_tol,
                          FirstStageCompression& out)
{
    const int S = (int)coefs.size();
    if (kc < 4 || S < 1) return false;
    ClassSvd cs[2];
    for (int cls = 0; cls < 2; cls++) class_svd(kc, cls, gram, coefs, cs[cls]);
    const int R = std::max(1, joint_rank(cs, S, rel_tol));
    if (2 * R >= kc || R > cs[1].nc || R > cs[0].nc) return false;
    out.R = R;
    out.B.assign((size_t)2 * R * kc, 0.0);
    out.u.assign((size_t)S * 2 * R, 0.0);
    for (int cls = 0; cls < 2; cls++) {
        const ClassSvd& c = cs[cls];
        for (int r = 0; r < R; r++) {
            const int row = 2 * r + cls;
            const double s2 = c.sig[r] * c.sig[r];
            if (!(s2 > 0.0)) continue; // a zero direction: Q_row = 0, no weight
            const double* mr = &c.m[(size_t)r * S];
            for (int a = 0; a < c.nc; a++) {
                double acc = 0.0;
                for (int j = 0; j < S; j++) acc += mr[j] * coefs[j][c.ks[a]];
                out.B[(size_t)row * kc + c.ks[a]] = acc / s2;
            }
            for (int j = 0; j < S; j++) out.u[(size_t)j * 2 * R + row] = mr[j];
        }
    }
    return true;
}

} // namespace mqw

// Host dense Hermitian eigensolver for the eigensolver's projected (Rayleigh-Ritz) problems, m <= a few
// hundred. Householder reduction of the Hermitian matrix to a complex tridiagonal one, a diagonal unitary
// scaling that makes the off-diagonal real, then implicit symmetric QR steps (Golub & Van Loan) on the real
// symmetric tridiagonal matrix with the rotations accumulated into the eigenvector matrix.
#include "dense_eig.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace dmrg {

using cx = std::complex<double>;

// Symmetric tridiagonal eigenproblem by the implicit symmetric QR algorithm with Wilkinson shifts: the
// deflation loop of Golub & Van Loan, Matrix Computations (4th ed., 2013), Algorithm 8.3.3, whose step is
// their Algorithm 8.3.2 (one implicit QR step, a chase of the bulge down the unreduced block by Givens
// rotations). d: diagonal (m), e[i]: the coupling of rows i and i+1 (m - 1 used). The rotations are
// accumulated into the columns of Z (m x m, column-major, complex): Z <- Z G_k.
static bool tridiag_qr(int m, std::vector<double> &d, std::vector<double> &e, std::vector<cx> &Z) {
    if(m <= 1) return true;
    e.resize(m);
    const double eps      = std::numeric_limits<double>::epsilon();
    const long   max_step = 60L * m; // total implicit steps; typically about 2 per eigenvalue
    long         steps    = 0;
    auto rotate_columns = [&](int k, double c, double s) { // Z(:, k:k+1) <- Z(:, k:k+1) [c s; -s c]
        cx *zk = Z.data() + static_cast<size_t>(m) * k, *zk1 = zk + m;
        for(int i = 0; i < m; ++i) {
            const cx a = zk[i], b = zk1[i];
            zk[i]      = c * a - s * b;
            zk1[i]     = s * a + c * b;
        }
    };
    int hi = m - 1; // the trailing rows hi+1.. are converged
    while(hi > 0) {
        // Deflation test of Alg. 8.3.3: a coupling below eps (|d_i| + |d_i+1|) is set to zero.
        for(int i = 0; i < hi; ++i)
            if(std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
        if(e[hi - 1] == 0.0) {
            --hi;
            continue;
        }
        int lo = hi - 1; // the unreduced block lo..hi ending at hi
        while(lo > 0 && e[lo - 1] != 0.0) --lo;
        if(++steps > max_step) return false;
        // Wilkinson shift: the eigenvalue of the trailing 2 x 2 block [d_hi-1 e; e d_hi] closer to d_hi.
        const double t  = 0.5 * (d[hi - 1] - d[hi]);
        const double b2 = e[hi - 1] * e[hi - 1];
        const double mu = d[hi] - b2 / (t + std::copysign(std::hypot(t, e[hi - 1]), t));
        // Bulge chase. G_k acts on rows/columns (k, k+1): T <- G_k^T T G_k with G_k = [c s; -s c] chosen so
        // that G_k^T [x; z] = [r; 0], x the entry to keep and z the entry to annihilate (the shifted first
        // column for k = lo, then the bulge at (k+1, k-1)).
        double x = d[lo] - mu, z = e[lo];
        for(int k = lo; k < hi; ++k) {
            const double r = std::hypot(x, z);
            double       c = 1.0, s = 0.0;
            if(r > 0.0) c = x / r, s = -z / r;
            if(k > lo) e[k - 1] = r;
            const double p = d[k], q = d[k + 1], b = e[k];
            d[k]     = c * c * p - 2.0 * c * s * b + s * s * q;
            d[k + 1] = s * s * p + 2.0 * c * s * b + c * c * q;
            e[k]     = c * s * (p - q) + (c * c - s * s) * b;
            rotate_columns(k, c, s);
            if(k + 1 < hi) { // the rotation moves the bulge to (k+2, k)
                x = e[k];
                z = -s * e[k + 1];
                e[k + 1] *= c;
            }
        }
    }
    return true;
}

bool hermitian_eig(int m, const cx *A, int lda, double *w, cx *Z_out) {
    if(m <= 0) return true;
    // Working copy (column-major m x m), Hermitian part.
    std::vector<cx> a(static_cast<size_t>(m) * m);
    for(int j = 0; j < m; ++j)
        for(int i = 0; i < m; ++i) a[i + static_cast<size_t>(m) * j] = 0.5 * (A[i + static_cast<size_t>(lda) * j] + std::conj(A[j + static_cast<size_t>(lda) * i]));
    std::vector<cx> Q(static_cast<size_t>(m) * m, cx(0.0, 0.0));
    for(int i = 0; i < m; ++i) Q[i + static_cast<size_t>(m) * i] = 1.0;
    auto at = [&](int i, int j) -> cx & { return a[i + static_cast<size_t>(m) * j]; };
    std::vector<cx> v(m), p(m), q(m);
    // Householder: zero column k below the subdiagonal, A <- H A H with H = I - 2 v v^H.
    for(int k = 0; k + 2 < m; ++k) {
        double xnorm = 0.0;
        for(int i = k + 1; i < m; ++i) xnorm += std::norm(at(i, k));
        xnorm = std::sqrt(xnorm);
        if(xnorm == 0.0) continue;
        const cx x0    = at(k + 1, k);
        const cx phase = std::abs(x0) > 0 ? x0 / std::abs(x0) : cx(1.0, 0.0);
        const cx alpha = -phase * xnorm;
        std::fill(v.begin(), v.end(), cx(0.0, 0.0));
        for(int i = k + 1; i < m; ++i) v[i] = at(i, k);
        v[k + 1] -= alpha;
        double vn = 0.0;
        for(int i = k + 1; i < m; ++i) vn += std::norm(v[i]);
        vn = std::sqrt(vn);
        if(vn == 0.0) continue;
        for(int i = k + 1; i < m; ++i) v[i] /= vn;
        // p = A v, K = v^H p, q = p - K v; A <- A - 2 v q^H - 2 q v^H
        for(int i = 0; i < m; ++i) {
            cx s(0.0, 0.0);
            for(int j = k + 1; j < m; ++j) s += at(i, j) * v[j];
            p[i] = s;
        }
        cx K(0.0, 0.0);
        for(int i = k + 1; i < m; ++i) K += std::conj(v[i]) * p[i];
        for(int i = 0; i < m; ++i) q[i] = p[i] - K * v[i];
        for(int j = 0; j < m; ++j)
            for(int i = 0; i < m; ++i) at(i, j) -= 2.0 * (v[i] * std::conj(q[j]) + q[i] * std::conj(v[j]));
        // Q <- Q H
        for(int i = 0; i < m; ++i) {
            cx s(0.0, 0.0);
            for(int j = k + 1; j < m; ++j) s += Q[i + static_cast<size_t>(m) * j] * v[j];
            for(int j = k + 1; j < m; ++j) Q[i + static_cast<size_t>(m) * j] -= 2.0 * s * std::conj(v[j]);
        }
    }
    // Tridiagonal: real diagonal, complex subdiagonal t_k = a(k+1,k). Scale by D = diag(phases) so the
    // subdiagonal becomes |t_k|: T' = D^H T D, eigenvectors Q D Z'.
    std::vector<double> d(m), e(m, 0.0);
    std::vector<cx>     ph(m, cx(1.0, 0.0));
    for(int k = 0; k < m; ++k) d[k] = at(k, k).real();
    for(int k = 0; k + 1 < m; ++k) {
        const cx t = at(k + 1, k);
        e[k]       = std::abs(t);
        ph[k + 1]  = e[k] > 0 ? ph[k] * t / e[k] : ph[k];
    }
    for(int j = 0; j < m; ++j)
        for(int i = 0; i < m; ++i) Q[i + static_cast<size_t>(m) * j] *= ph[j];
    if(!tridiag_qr(m, d, e, Q)) return false;
    // Sort ascending.
    std::vector<int> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    for(int j = 0; j < m; ++j) {
        w[j] = d[idx[j]];
        for(int i = 0; i < m; ++i) Z_out[i + static_cast<size_t>(m) * j] = Q[i + static_cast<size_t>(m) * idx[j]];
    }
    return true;
}

} // namespace dmrg

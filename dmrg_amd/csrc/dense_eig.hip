// Host dense Hermitian eigensolver for the eigensolver's projected (Rayleigh-Ritz) problems, m <= a few
// hundred. Householder reduction of the Hermitian matrix to a complex tridiagonal one, a diagonal unitary
// scaling that makes the off-diagonal real, then implicit-shift QL iterations on the real symmetric
// tridiagonal matrix with the rotations accumulated into the eigenvector matrix.
#include "dense_eig.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace dmrg {

using cx = std::complex<double>;

// QL with implicit Wilkinson shifts on the symmetric tridiagonal (d, e); e[i] couples i and i+1.
// Rotations are applied to the columns of Z (m x m, column-major, complex).
static bool tridiag_ql(int m, std::vector<double> &d, std::vector<double> &e, std::vector<cx> &Z) {
    if(m <= 1) return true;
    e.resize(m);
    e[m - 1] = 0.0;
    for(int l = 0; l < m; ++l) {
        int iter = 0;
        while(true) {
            int mm = l;
            for(; mm < m - 1; ++mm) {
                const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if(std::fabs(e[mm]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if(mm == l) break;
            if(++iter > 60) return false;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g        = d[mm] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int    i = mm - 1;
            bool   early = false;
            for(; i >= l; --i) {
                double f = s * e[i];
                double b = c * e[i];
                r        = std::hypot(f, g);
                e[i + 1] = r;
                if(r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    early = true;
                    break;
                }
                s        = f / r;
                c        = g / r;
                g        = d[i + 1] - p;
                r        = (d[i] - g) * s + 2.0 * c * b;
                p        = s * r;
                d[i + 1] = g + p;
                g        = c * r - b;
                for(int k = 0; k < m; ++k) {
                    const cx zi1 = Z[k + static_cast<size_t>(m) * (i + 1)];
                    const cx zi  = Z[k + static_cast<size_t>(m) * i];
                    Z[k + static_cast<size_t>(m) * (i + 1)] = s * zi + c * zi1;
                    Z[k + static_cast<size_t>(m) * i]       = c * zi - s * zi1;
                }
            }
            if(early) continue;
            d[l] -= p;
            e[l]  = g;
            e[mm] = 0.0;
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
    if(!tridiag_ql(m, d, e, Q)) return false;
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

// Generalized Davidson with thick restart on the GPU (see eigs.hpp).
#include "eigs.hpp"

#include "dense_eig.hpp"

#include <algorithm>
#include <deque>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>

namespace dmrg {

using cx = std::complex<double>;

// Layout of the small device/host exchange buffer (doubles).
constexpr int SB_COL  = 0;                 // projected column, 2 * VEC_MAX_M
constexpr int SB_NRM  = 2 * VEC_MAX_M;     // |t|^2 before / after orthogonalisation (2 x 2)
constexpr int SB_RITZ = 2 * VEC_MAX_M + 4; // |r|^2, |x|^2
constexpr int SB_Y    = 2 * VEC_MAX_M + 8; // Ritz coefficients, 2 * VEC_MAX_M
constexpr int SB_FLAG = 4 * VEC_MAX_M + 8; // second-pass flag of the DGKS orthogonalisation (int)
// Refined projection: QR of Z = W - tau V, column by column (two Gram-Schmidt passes and the norm).
constexpr int SB_QC1  = 4 * VEC_MAX_M + 10; // coefficients of pass 1, 2 * VEC_MAX_M
constexpr int SB_QC2  = 6 * VEC_MAX_M + 10; // coefficients of pass 2, 2 * VEC_MAX_M
constexpr int SB_QN   = 8 * VEC_MAX_M + 10; // |z|^2 after both passes
constexpr int SB_TAU  = 8 * VEC_MAX_M + 12; // -tau as a complex number (2 doubles)
constexpr int SB_SIZE = 8 * VEC_MAX_M + 14;

// ------------------------------------------------------------------------------------------- correction solver
constexpr int CG_CHECK = 4; // host look at the CG state every CG_CHECK steps

template <typename T32>
CorrectionSolver<T32>::CorrectionSolver(MatVecOp &o, hipStream_t st) : op(&o), s(st), n(o.D.n()) {
    scratch.reserve(static_cast<size_t>(vec_scratch_doubles()) * sizeof(double));
    vec_scratch_init(scratch.as<double>(), s);
    dev.reserve((8 + CG_STATE) * sizeof(double));
    host = static_cast<double *>(pinned_get(CG_STATE * sizeof(double)));
}

template <typename T32>
CorrectionSolver<T32>::~CorrectionSolver() {
    if(host) {
        (void) hipStreamSynchronize(s); // copies into host may still be queued
        pinned_put(host, CG_STATE * sizeof(double));
    }
}

template <typename T32>
template <typename T64>
int CorrectionSolver<T32>::solve(const T64 *x, const T64 *r, double theta, double sigma, int max_it, double rtol, T64 *t) {
    const size_t vb = static_cast<size_t>(n) * sizeof(T32);
    xp.reserve(2 * vb); // x and p side by side: one pass gives x^H q and p^H q
    q.reserve(vb);
    rv.reserve(vb);
    tv.reserve(vb);
    T32    *X = xp.as<T32>(), *P = X + n, *Q = q.as<T32>(), *R = rv.as<T32>(), *Tt = tv.as<T32>();
    double *d = dev.as<double>(), *dots = d, *rr = d + 4, *c = d + 6, *state = d + 8, *scr = scratch.as<double>();
    convert<T32, T64>(n, x, X, s);
    multidot_x<T32>(n, 0, X, n, X, c, rr, nullptr, scr, s); // |x|^2 of the FP32 copy
    scale_by_inv_norm<T32>(n, X, rr, s);
    convert<T32, T64>(n, r, R, s);
    multidot<T32>(n, 1, X, n, R, c, scr, s);
    cg_init<T32>(n, X, c, sigma, R, P, Tt, rr, state, scr, s);
    const double rtol2 = rtol * rtol;
    int          k     = 0;
    while(k < max_it) {
        op->multax_device(P, n, Q, n, 1, s, true);
        cg_shift<T32>(n, P, Q, theta, sigma, state, k, s);
        multidot<T32>(n, 2, X, n, Q, dots, scr, s);
        cg_step<T32>(n, X, P, Q, dots, Tt, R, rr, state, k, rtol2, scr, s);
        ++k;
        if(k % CG_CHECK == 0 && k < max_it) {
            DMRG_HIP_CHECK(hipMemcpyAsync(host, state, CG_STATE * sizeof(double), hipMemcpyDeviceToHost, s));
            DMRG_HIP_CHECK(hipStreamSynchronize(s));
            if(host[CG_STOP + (k & 1)] != 0.0) break;
        }
    }
    DMRG_HIP_CHECK(hipMemcpyAsync(host, state, CG_STATE * sizeof(double), hipMemcpyDeviceToHost, s));
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
    if(host[CG_ITS] < 1.0) DMRG_HIP_CHECK(hipMemcpyAsync(t, r, static_cast<size_t>(n) * sizeof(T64), hipMemcpyDeviceToDevice, s));
    else convert<T64, T32>(n, Tt, t, s);
    return k;
}

template <typename T>
Davidson<T>::Davidson(MatVecOp &o, hipStream_t st) : op(&o), s(st), n(o.D.n()) {
    scratch.reserve(static_cast<size_t>(vec_scratch_doubles()) * sizeof(double));
    vec_scratch_init(scratch.as<double>(), s);
    small.reserve(SB_SIZE * sizeof(double));
    host = static_cast<double *>(pinned_get(SB_SIZE * sizeof(double)));
}

template <typename T>
Davidson<T>::~Davidson() {
    if(host) {
        (void) hipStreamSynchronize(s); // copies into host may still be queued
        pinned_put(host, SB_SIZE * sizeof(double));
    }
}

template <typename T>
void Davidson<T>::matvec(const T *x, T *y, int nvec) {
    op->multax_device(x, n, y, n, nvec, s, std::is_same_v<typename vtraits<T>::R, float>);
}

template <typename T>
void Davidson<T>::orthonormalize_into(T *t, int m, double *dnrm_pre, double *dnrm_post) {
    double *sm = small.as<double>();
    multidot<T>(n, 1, t, n, t, dnrm_pre, scratch.as<double>(), s);
    for(int pass = 0; pass < 2; ++pass) { // classical Gram-Schmidt, twice
        if(m == 0) break;
        multidot<T>(n, m, V.as<T>(), n, t, sm + SB_COL, scratch.as<double>(), s);
        gemv_update<T>(n, m, V.as<T>(), n, sm + SB_COL, -1.0, 1.0, t, s);
    }
    multidot<T>(n, 1, t, n, t, dnrm_post, scratch.as<double>(), s);
    scale_by_inv_norm<T>(n, t, dnrm_post, s);
}

// t = V[:, m] <- src orthogonalised against V[:, 0:m] and normalised. One classical Gram-Schmidt pass whose
// first projection also yields |src|^2 and whose update yields |t|^2; a second pass runs (decided on the
// device, no host round trip) only when the norm dropped below 1/sqrt(2) of its value (DGKS criterion).
// |src|^2 and the final |t|^2 land in dnrm_pre / dnrm_post for the breakdown test. src may be t itself.
template <typename T>
void Davidson<T>::expand_from(const T *src, T *t, int m, double *dnrm_pre, double *dnrm_post) {
    double *sm   = small.as<double>();
    int    *flag = reinterpret_cast<int *>(sm + SB_FLAG);
    double *scr  = scratch.as<double>();
    multidot_x<T>(n, m, V.as<T>(), n, src, sm + SB_COL, dnrm_pre, nullptr, scr, s);
    gemv_update_norm<T>(n, m, V.as<T>(), n, sm + SB_COL, -1.0, 1.0, src, t, dnrm_post, nullptr, scr, s);
    if(m > 0) {
        multidot_dgks<T>(n, m, V.as<T>(), n, t, sm + SB_COL, dnrm_pre, dnrm_post, 0.5, flag, scr, s);
        gemv_update_norm<T>(n, m, V.as<T>(), n, sm + SB_COL, -1.0, 1.0, t, t, dnrm_post, flag, scr, s);
    }
    scale_by_inv_norm<T>(n, t, dnrm_post, s);
}

// Appends column m to the QR factorisation of Z = W - shift V: z = W[:, m] - shift V[:, m], two classical
// Gram-Schmidt passes against Qr[:, 0:m] (coefficients to SB_QC1 / SB_QC2), Qr[:, m] = z / |z| (|z|^2 to SB_QN).
// The host reads the coefficients at the iteration's synchronisation (column m of R = c1 + c2, R(m, m) = |z|).
template <typename T>
void Davidson<T>::qr_append(int m, double shift) {
    double *sm = small.as<double>(), *scr = scratch.as<double>();
    T      *z  = Qr.as<T>() + static_cast<int64_t>(m) * n;
    host[SB_TAU] = -shift, host[SB_TAU + 1] = 0.0; // pinned; no copy from the device ever covers SB_TAU
    DMRG_HIP_CHECK(hipMemcpyAsync(sm + SB_TAU, host + SB_TAU, 2 * sizeof(double), hipMemcpyHostToDevice, s));
    DMRG_HIP_CHECK(hipMemcpyAsync(z, W.as<T>() + static_cast<int64_t>(m) * n, static_cast<size_t>(n) * sizeof(T),
                                  hipMemcpyDeviceToDevice, s));
    gemv_update<T>(n, 1, V.as<T>() + static_cast<int64_t>(m) * n, n, sm + SB_TAU, 1.0, 1.0, z, s);
    if(m > 0) {
        multidot<T>(n, m, Qr.as<T>(), n, z, sm + SB_QC1, scr, s);
        gemv_update<T>(n, m, Qr.as<T>(), n, sm + SB_QC1, -1.0, 1.0, z, s);
        multidot<T>(n, m, Qr.as<T>(), n, z, sm + SB_QC2, scr, s);
        gemv_update<T>(n, m, Qr.as<T>(), n, sm + SB_QC2, -1.0, 1.0, z, s);
    }
    multidot<T>(n, 1, z, n, z, sm + SB_QN, scr, s);
    scale_by_inv_norm<T>(n, z, sm + SB_QN, s);
}

template <typename T>
void Davidson<T>::transform_basis(T *B, int m, const std::vector<cx> &Q, int k) {
    // B[:, 0:k] <- B[:, 0:m] Q  (Q: m x k, column-major)
    std::vector<T> q(static_cast<size_t>(m) * k);
    for(size_t i = 0; i < q.size(); ++i) {
        if constexpr(vtraits<T>::cx) q[i] = T{static_cast<typename vtraits<T>::R>(Q[i].real()), static_cast<typename vtraits<T>::R>(Q[i].imag())};
        else q[i] = static_cast<T>(Q[i].real());
    }
    Ydev.reserve(q.size() * sizeof(T));
    DMRG_HIP_CHECK(hipMemcpyAsync(Ydev.p, q.data(), q.size() * sizeof(T), hipMemcpyHostToDevice, s));
    GemmDesc g;
    g.dtype = vtraits<T>::dt;
    g.M = static_cast<int32_t>(n), g.N = k, g.K = m, g.batch = 1;
    g.A = B, g.am = plain(1), g.ak = plain(n);
    g.B = Ydev.p, g.bk = plain(1), g.bn = plain(m);
    g.C = tmp.p, g.cm = plain(1), g.cn = plain(n);
    gemm(g, s);
    DMRG_HIP_CHECK(hipMemcpyAsync(B, tmp.p, static_cast<size_t>(n) * k * sizeof(T), hipMemcpyDeviceToDevice, s));
    // The host copy of q must outlive the async upload.
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
}

// Right singular vectors of an m x m complex matrix R (column-major) by the one-sided (Hestenes) Jacobi method on
// its columns, which keeps the small singular values to high relative accuracy (the refined projection needs the
// smallest one of W - tau V, far below eps |A| relative to the largest when tau is near an eigenvalue). Returns
// the singular values ascending and the matching right singular vectors as the columns of Vr.
static void small_svd_right(int m, std::vector<std::complex<double>> A, std::vector<double> &sv,
                            std::vector<std::complex<double>> &Vr) {
    using cx = std::complex<double>;
    Vr.assign(static_cast<size_t>(m) * m, cx(0.0, 0.0));
    for(int i = 0; i < m; ++i) Vr[i + static_cast<size_t>(m) * i] = 1.0;
    auto col = [&](std::vector<cx> &M, int j) { return M.data() + static_cast<size_t>(m) * j; };
    const double tol = std::numeric_limits<double>::epsilon() * std::sqrt(static_cast<double>(m));
    for(int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for(int p = 0; p < m - 1; ++p)
            for(int q = p + 1; q < m; ++q) {
                cx    *ap = col(A, p), *aq = col(A, q);
                double a = 0, b = 0;
                cx     c(0.0, 0.0);
                for(int i = 0; i < m; ++i) a += std::norm(ap[i]), b += std::norm(aq[i]), c += std::conj(ap[i]) * aq[i];
                const double ac = std::abs(c);
                if(ac <= tol * std::sqrt(a * b) || ac == 0.0) continue;
                rotated = true;
                // Rotation that orthogonalises columns p and q: [ap aq] <- [ap aq] [cs, -sn conj(e); sn e, cs]
                const cx     e   = c / ac;
                const double zeta = (b - a) / (2.0 * ac);
                const double t    = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double cs   = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
                for(cx *M : {ap, col(Vr, p)}) {
                    cx *Mq = M == ap ? aq : col(Vr, q);
                    for(int i = 0; i < m; ++i) {
                        const cx x = M[i], y = Mq[i];
                        M[i]       = cs * x - sn * std::conj(e) * y;
                        Mq[i]      = sn * e * x + cs * y;
                    }
                }
            }
        if(!rotated) break;
    }
    std::vector<double> nrm(m);
    for(int j = 0; j < m; ++j) {
        double a = 0;
        for(int i = 0; i < m; ++i) a += std::norm(col(A, j)[i]);
        nrm[j] = std::sqrt(a);
    }
    std::vector<int> ord(m);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nrm[x] < nrm[y]; });
    std::vector<cx> V2(static_cast<size_t>(m) * m);
    sv.resize(m);
    for(int j = 0; j < m; ++j) {
        sv[j] = nrm[ord[j]];
        for(int i = 0; i < m; ++i) V2[i + static_cast<size_t>(m) * j] = col(Vr, ord[j])[i];
    }
    Vr.swap(V2);
}

static void sort_by_target(const EigsParams &p, const std::vector<double> &w, std::vector<int> &ord) {
    ord.resize(w.size());
    std::iota(ord.begin(), ord.end(), 0);
    if(p.target == TARGET_LA) std::reverse(ord.begin(), ord.end());
    else if(p.target == TARGET_CLOSEST)
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return std::fabs(w[a] - p.shift) < std::fabs(w[b] - p.shift); });
}

template <typename T>
EigsStats Davidson<T>::solve(const EigsParams &p, const T *init, int k0, T *out, double *evals, double *rnorms, int nkeep) {
    using clock = std::chrono::steady_clock;
    const auto t0  = clock::now();
    EigsStats  st;
    const int  nev = std::max(1, p.nev);
    DMRG_REQUIRE(nev <= n, DMRG_ERR_ARG, "nev larger than the problem size");
    // Basis size rule of solver.cpp:12-15 and the restart size of solver_primme.cpp:373-376: the basis
    // must hold the restart vectors, the retained previous Ritz vector and one expansion.
    int mmax = p.ncv > 0 ? p.ncv : nev * static_cast<int>(std::ceil(std::log2(static_cast<double>(std::max<int64_t>(n, 2)))));
    mmax     = static_cast<int>(std::clamp<int64_t>(mmax, 2 * nev, n));
    int mmin = p.min_restart > 0 ? p.min_restart : std::max(1, mmax / 2 - 1);
    mmin     = std::max(mmin, nev);
    mmax     = static_cast<int>(std::clamp<int64_t>(mmax, mmin + 1 + std::max(2, p.plus_k), n));
    mmax     = std::min(mmax, VEC_MAX_M);
    mmin     = std::max(1, std::min(mmin, mmax - 1 - p.plus_k));
    nkeep = std::max(nkeep, nev);
    nkeep = std::min<int64_t>(nkeep, n);
    k0    = std::min<int>(k0, mmax);
    const double eps_t  = std::is_same_v<typename vtraits<T>::R, float> ? 1.2e-7 : 2.3e-16;
    const size_t vbytes = static_cast<size_t>(n) * sizeof(T);
    V.reserve(vbytes * mmax);
    W.reserve(vbytes * mmax);
    tmp.reserve(vbytes * mmax);
    rbuf.reserve(vbytes);
    T      *Vp = V.as<T>(), *Wp = W.as<T>(), *r = rbuf.as<T>();
    double *sm = small.as<double>();
    bool pc_built = false;
    const bool refined = p.projection == 1 && p.target == TARGET_CLOSEST;
    std::vector<cx> Rr; // refined: upper-triangular factor of W - shift V = Qr Rr (mmax x mmax)
    if(refined) {
        Qr.reserve(vbytes * mmax);
        Rr.assign(static_cast<size_t>(mmax) * mmax, cx(0.0, 0.0));
    }
    auto Rat = [&](int i, int j) -> cx & { return Rr[i + static_cast<size_t>(mmax) * j]; };
    constexpr int QR_SPAN = SB_TAU - SB_QC1; // the doubles of SB_QC1 .. SB_QN
    // Column j of Rr from the host copy of the QR exchange region (after the synchronisation that brought it).
    auto qr_accept = [&](int j) {
        for(int i = 0; i < j; ++i)
            Rat(i, j) = cx(host[SB_QC1 + 2 * i] + host[SB_QC2 + 2 * i], host[SB_QC1 + 2 * i + 1] + host[SB_QC2 + 2 * i + 1]);
        Rat(j, j) = cx(std::sqrt(std::max(0.0, host[SB_QN])), 0.0);
        for(int i = j + 1; i < mmax; ++i) Rat(i, j) = cx(0.0, 0.0);
    };
    auto qr_fetch = [&]() {
        DMRG_HIP_CHECK(hipMemcpyAsync(host + SB_QC1, sm + SB_QC1, QR_SPAN * sizeof(double), hipMemcpyDeviceToHost, s));
    };
    if(p.precond == 1) {
        // Diagonal of H_eff in the operator's FP64 dtype; the kernel converts on the fly.
        diag.reserve(static_cast<size_t>(n) * dtype_size(op->dtype));
        op->diagonal(diag.p, s);
    }

    // ---- initial basis
    int m = 0;
    if(init && k0 > 0) {
        DMRG_HIP_CHECK(hipMemcpyAsync(Vp, init, vbytes * k0, hipMemcpyDeviceToDevice, s));
        for(int j = 0; j < k0; ++j) {
            if(j != m) DMRG_HIP_CHECK(hipMemcpyAsync(Vp + m * n, Vp + j * n, vbytes, hipMemcpyDeviceToDevice, s));
            orthonormalize_into(Vp + m * n, m, sm + SB_NRM, sm + SB_NRM + 2);
            DMRG_HIP_CHECK(hipMemcpyAsync(host + SB_NRM, sm + SB_NRM, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
            DMRG_HIP_CHECK(hipStreamSynchronize(s));
            if(host[SB_NRM + 2] > 1e4 * eps_t * eps_t * host[SB_NRM] && host[SB_NRM] > 0) ++m;
        }
    }
    if(m == 0) {
        random_fill<T>(n, Vp, p.seed, s);
        orthonormalize_into(Vp, 0, sm + SB_NRM, sm + SB_NRM + 2);
        m = 1;
    }
    matvec(Vp, Wp, m);
    st.matvecs += m;
    std::vector<cx> H(static_cast<size_t>(mmax) * mmax, cx(0.0, 0.0));
    auto            Hat = [&](int i, int j) -> cx & { return H[i + static_cast<size_t>(mmax) * j]; };
    for(int j = 0; j < m; ++j) {
        multidot<T>(n, m, Vp, n, Wp + j * n, sm + SB_COL, scratch.as<double>(), s);
        DMRG_HIP_CHECK(hipMemcpyAsync(host + SB_COL, sm + SB_COL, 2 * m * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        for(int i = 0; i < m; ++i) Hat(i, j) = cx(host[SB_COL + 2 * i], host[SB_COL + 2 * i + 1]);
    }
    if(refined)
        for(int j = 0; j < m; ++j) {
            qr_append(j, p.shift);
            qr_fetch();
            DMRG_HIP_CHECK(hipStreamSynchronize(s));
            qr_accept(j);
        }

    std::vector<double> w(mmax);
    std::vector<cx>     Hm, Z;
    std::vector<int>    ord;
    std::vector<cx>     yprev;
    double              best_r = std::numeric_limits<double>::infinity();
    int                 since_best = 0;
    bool                inject_random = false;

    // Ritz vector x = V y and residual r = W y - theta x for column col of Z; |r|^2, |x|^2 go to SB_RITZ.
    auto ritz_launch = [&](int col_in_H, int m_, T *xout) {
        for(int i = 0; i < m_; ++i) {
            host[SB_Y + 2 * i]     = Z[i + static_cast<size_t>(m_) * col_in_H].real();
            host[SB_Y + 2 * i + 1] = Z[i + static_cast<size_t>(m_) * col_in_H].imag();
        }
        DMRG_HIP_CHECK(hipMemcpyAsync(sm + SB_Y, host + SB_Y, 2 * m_ * sizeof(double), hipMemcpyHostToDevice, s));
        ritz_residual<T>(n, m_, Vp, Wp, n, sm + SB_Y, w[col_in_H], xout, r, sm + SB_RITZ, scratch.as<double>(), s);
    };
    auto ritz_norm = [&]() { // |r| / |x| from the host copy of SB_RITZ
        return std::sqrt(std::max(0.0, host[SB_RITZ])) / std::sqrt(std::max(1e-300, host[SB_RITZ + 1]));
    };
    auto ritz = [&](int col_in_H, int m_, T *xout) -> double {
        ritz_launch(col_in_H, m_, xout);
        DMRG_HIP_CHECK(hipMemcpyAsync(host + SB_RITZ, sm + SB_RITZ, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        return ritz_norm();
    };
    auto rayleigh_ritz = [&]() {
        Hm.assign(static_cast<size_t>(m) * m, cx(0.0, 0.0));
        for(int j = 0; j < m; ++j)
            for(int i = 0; i < m; ++i) Hm[i + static_cast<size_t>(m) * j] = Hat(i, j);
        Z.assign(static_cast<size_t>(m) * m, cx(0.0, 0.0));
        w.assign(m, 0.0);
        if(!hermitian_eig(m, Hm.data(), m, w.data(), Z.data())) throw Error(DMRG_ERR_NOT_CONVERGED, "Rayleigh-Ritz eigensolver failed");
        sort_by_target(p, w, ord);
        for(int i = 0; i < m; ++i) st.anorm = std::max(st.anorm, std::fabs(w[i]));
        if(refined) {
            // Refined vectors: right singular vectors of Rr (ascending singular values, the best first), each
            // with its Rayleigh quotient y^H H y as the value.
            std::vector<cx> R(static_cast<size_t>(m) * m), Y;
            for(int j = 0; j < m; ++j)
                for(int i = 0; i < m; ++i) R[i + static_cast<size_t>(m) * j] = Rat(i, j);
            std::vector<double> sv;
            small_svd_right(m, R, sv, Y);
            Z = Y;
            for(int j = 0; j < m; ++j) {
                cx q(0.0, 0.0);
                for(int b = 0; b < m; ++b)
                    for(int a = 0; a < m; ++a)
                        q += std::conj(Y[a + static_cast<size_t>(m) * j]) * Hm[a + static_cast<size_t>(m) * b] * Y[b + static_cast<size_t>(m) * j];
                w[j] = q.real();
            }
            std::iota(ord.begin(), ord.end(), 0);
        }
    };
    // Convergence of the pair with residual norm rq and value ev (q: its index among the wanted pairs).
    std::deque<double> recent_ev, recent_rn;
    auto converged = [&](double rq, double ev, int q) {
        const double ptol = p.tol * std::max(st.anorm, 1e-300);
        if(p.conv_test != 1) return rq <= ptol;
        // folded_spectrum::convTestFun (optimize_folded_spectrum.cpp:41-93), on the leading pair's history
        if(q == 0) {
            recent_ev.push_back(ev);
            recent_rn.push_back(rq);
            while(recent_ev.size() > 500) recent_ev.pop_front();
            while(recent_rn.size() > 500) recent_rn.pop_front();
        }
        const double dev = recent_ev.size() >= 250 ? recent_ev.back() - recent_ev.front() : -1.0; // sum of differences
        const double drn = recent_rn.size() >= 250 ? recent_rn.back() - recent_rn.front() : -1.0;
        const bool   saturated = dev > -p.tol || drn > -p.tol;
        return (rq < ptol && saturated) || rq < p.tol;
    };
    // FP32 correction solves (FP64 solves of one lowest or largest pair without preconditioner).
    constexpr bool fp64     = std::is_same_v<typename vtraits<T>::R, double>;
    const bool     use_inner = fp64 && p.inner_max > 0 && nev == 1 && p.precond == 0 && p.target != TARGET_CLOSEST;
    if(use_inner && !inner) inner = std::make_unique<CorrectionSolver<typename lower_precision<T>::type>>(*op, s);
    // t = M^-1 r with the configured preconditioner, or the correction t from the FP32 solver (theta, x: the
    // Ritz pair being refined).
    auto apply_precond = [&](double theta, const T *x, T *t) {
        if(use_inner) {
            if constexpr(fp64)
                st.inner_matvecs += inner->solve(x, r, theta, p.target == TARGET_LA ? -1.0 : 1.0, p.inner_max, p.inner_tol, t);
        } else if(p.precond == 2) {
            // Block-Jacobi: built once per operator state (readyCalcPc, matvec_mpos.cpp:1086-1088) at the target
            // shift when there is one, else at the first Ritz value that needs it; the FP64 phase of a mixed solve
            // reuses the FP32 phase's blocks.
            if(!pc_built) {
                op->set_jcb(p.jcb);
                const double pc_shift = p.target == TARGET_CLOSEST ? p.shift : theta;
                if(!op->pc_ready || op->pc_shift != pc_shift) op->block_pc_build(pc_shift, s);
                pc_built = true;
            }
            if constexpr(std::is_same_v<typename vtraits<T>::R, double>) {
                op->block_pc_apply(r, t, s);
            } else { // FP32 phase: the block inverses are FP64
                using TD = std::conditional_t<vtraits<T>::cx, cpx<double>, double>;
                pcw.reserve(2 * static_cast<size_t>(n) * sizeof(TD));
                TD *r64 = pcw.as<TD>(), *t64 = r64 + n;
                convert<TD, T>(n, r, r64, s);
                op->block_pc_apply(r64, t64, s);
                convert<T, TD>(n, t64, t, s);
            }
        } else {
            const double floor = 1e-8 * std::max(st.anorm, 1e-300);
            if constexpr(vtraits<T>::cx) jacobi_precond<T, cpx<double>>(n, r, diag.as<cpx<double>>(), theta, floor, t, s);
            else jacobi_precond<T, double>(n, r, diag.as<double>(), theta, floor, t, s);
        }
    };
    // After the synchronisation that brought the new H column and the norms of the expansion vector to the host:
    // grow the basis, or flag a breakdown (the correction lay in the basis) for a random expansion next time.
    auto accept_expansion = [&](int &mm) {
        const double pre = host[SB_NRM], post = host[SB_NRM + 2];
        if(!(post > 1e4 * eps_t * eps_t * pre) || !std::isfinite(post)) {
            if(!std::isfinite(post) || !std::isfinite(pre)) throw Error(DMRG_ERR_NAN, "eigensolver: non-finite expansion vector");
            inject_random = true;
            return;
        }
        for(int i = 0; i <= mm; ++i) Hat(i, mm) = cx(host[SB_COL + 2 * i], host[SB_COL + 2 * i + 1]);
        for(int i = 0; i < mm; ++i) Hat(mm, i) = std::conj(Hat(i, mm));
        Hat(mm, mm) = cx(Hat(mm, mm).real(), 0.0);
        if(refined) qr_accept(mm);
        ++mm;
    };

    std::vector<double> rn(nev, 0.0);
    while(true) {
        rayleigh_ritz();
        // One wanted pair and no restart due: the residual check is deferred to the single synchronisation
        // of the iteration, and the expansion is computed meanwhile (discarded if the pair has converged).
        // (Not with correction solves: their FP32 matvecs would be lost on a pair that has converged.)
        const bool speculate = !use_inner && nev == 1 && m + 1 <= mmax && m < n && st.matvecs < p.max_matvecs;
        if(speculate) {
            ritz_launch(ord[0], m, out);
            yprev.assign(m, cx(0.0, 0.0));
            for(int i = 0; i < m; ++i) yprev[i] = Z[i + static_cast<size_t>(m) * ord[0]];
            T *t = Vp + static_cast<int64_t>(m) * n;
            const T *src = r;
            if(inject_random) {
                random_fill<T>(n, t, p.seed + 7919ull * (st.iters + 1), s);
                src           = t;
                inject_random = false;
            } else if(p.precond != 0) {
                apply_precond(w[ord[0]], out, t);
                src = t;
            }
            expand_from(src, t, m, sm + SB_NRM, sm + SB_NRM + 2);
            matvec(t, Wp + static_cast<int64_t>(m) * n, 1);
            st.matvecs += 1;
            multidot<T>(n, m + 1, Vp, n, Wp + static_cast<int64_t>(m) * n, sm + SB_COL, scratch.as<double>(), s);
            if(refined) qr_append(m, p.shift);
            DMRG_HIP_CHECK(hipMemcpyAsync(host, sm, (SB_RITZ + 2) * sizeof(double), hipMemcpyDeviceToHost, s));
            if(refined) qr_fetch();
            DMRG_HIP_CHECK(hipStreamSynchronize(s));
            rn[0]    = ritz_norm();
            st.nconv = converged(rn[0], w[ord[0]], 0) ? 1 : 0;
            if(st.nconv == 1) break; // converged: the expansion is dropped (m unchanged)
            if(rn[0] < 0.9 * best_r) {
                best_r     = rn[0];
                since_best = 0;
            } else if(++since_best >= p.stagnation_window && p.stagnation_window > 0) {
                st.stagnated = true;
                break;
            }
            if(p.max_time_s > 0 && std::chrono::duration<double>(clock::now() - t0).count() > p.max_time_s) break;
            accept_expansion(m);
            ++st.iters;
            continue;
        }
        // Residuals of the wanted pairs in order; stop at the first unconverged one (its r drives expansion).
        int pstar = -1;
        for(int q = 0; q < nev && q < m; ++q) {
            rn[q] = ritz(ord[q], m, out + q * n);
            if(!converged(rn[q], w[ord[q]], q)) {
                pstar = q;
                break;
            }
        }
        const int nconv = pstar < 0 ? std::min(nev, m) : pstar;
        st.nconv        = nconv;
        if(pstar < 0 && m >= nev) break;
        if(m >= n) { // the basis spans the whole space: the Ritz pairs are exact
            st.nconv = std::min<int64_t>(nev, n);
            break;
        }
        if(pstar < 0) pstar = m - 1; // fewer basis vectors than nev: keep expanding
        const double rcur = rn[pstar];
        if(rcur < 0.9 * best_r) {
            best_r     = rcur;
            since_best = 0;
        } else if(++since_best >= p.stagnation_window && p.stagnation_window > 0) {
            st.stagnated = true;
            break;
        }
        if(st.matvecs >= p.max_matvecs) break;
        if(p.max_time_s > 0 && std::chrono::duration<double>(clock::now() - t0).count() > p.max_time_s) break;

        // ---- thick restart (GD+k)
        if(m + 1 > mmax) {
            std::vector<std::vector<cx>> cols;
            for(int q = 0; q < mmin && q < m; ++q) {
                std::vector<cx> c(m);
                for(int i = 0; i < m; ++i) c[i] = Z[i + static_cast<size_t>(m) * ord[q]];
                cols.push_back(std::move(c));
            }
            if(p.plus_k > 0 && static_cast<int>(yprev.size()) == m - 1) {
                std::vector<cx> c(m, cx(0.0, 0.0));
                for(int i = 0; i < m - 1; ++i) c[i] = yprev[i];
                cols.push_back(std::move(c));
            }
            // Modified Gram-Schmidt in C^m (twice).
            std::vector<std::vector<cx>> Qc;
            for(auto &c : cols) {
                for(int pass = 0; pass < 2; ++pass)
                    for(auto &q : Qc) {
                        cx d(0.0, 0.0);
                        for(int i = 0; i < m; ++i) d += std::conj(q[i]) * c[i];
                        for(int i = 0; i < m; ++i) c[i] -= d * q[i];
                    }
                double nr = 0;
                for(auto &v : c) nr += std::norm(v);
                nr = std::sqrt(nr);
                if(nr > 1e-10) {
                    for(auto &v : c) v /= nr;
                    Qc.push_back(c);
                }
            }
            // At most mmax - 1 vectors survive, so the next expansion has a free column.
            if(static_cast<int>(Qc.size()) > mmax - 1) Qc.resize(mmax - 1);
            const int       k = static_cast<int>(Qc.size());
            std::vector<cx> Q(static_cast<size_t>(m) * k);
            for(int j = 0; j < k; ++j)
                for(int i = 0; i < m; ++i) Q[i + static_cast<size_t>(m) * j] = Qc[j][i];
            transform_basis(Vp, m, Q, k);
            transform_basis(Wp, m, Q, k);
            std::vector<cx> HQ(static_cast<size_t>(m) * k, cx(0.0, 0.0)), Hn(static_cast<size_t>(k) * k, cx(0.0, 0.0));
            for(int j = 0; j < k; ++j)
                for(int l = 0; l < m; ++l)
                    for(int i = 0; i < m; ++i) HQ[i + static_cast<size_t>(m) * j] += Hat(i, l) * Q[l + static_cast<size_t>(m) * j];
            for(int j = 0; j < k; ++j)
                for(int i = 0; i < k; ++i) {
                    cx acc(0.0, 0.0);
                    for(int l = 0; l < m; ++l) acc += std::conj(Q[l + static_cast<size_t>(m) * i]) * HQ[l + static_cast<size_t>(m) * j];
                    Hn[i + static_cast<size_t>(k) * j] = acc;
                }
            std::fill(H.begin(), H.end(), cx(0.0, 0.0));
            for(int j = 0; j < k; ++j)
                for(int i = 0; i < k; ++i) Hat(i, j) = Hn[i + static_cast<size_t>(k) * j];
            if(refined) {
                // (W - shift V) Q = Qr (Rr Q): QR of the small m x k matrix B = Rr Q (modified Gram-Schmidt, twice),
                // then Qr <- Qr Qs and Rr <- Rs.
                std::vector<cx> B(static_cast<size_t>(m) * k, cx(0.0, 0.0)), Qs(static_cast<size_t>(m) * k, cx(0.0, 0.0));
                std::vector<cx> Rs(static_cast<size_t>(k) * k, cx(0.0, 0.0));
                for(int j = 0; j < k; ++j)
                    for(int l = 0; l < m; ++l)
                        for(int i = 0; i <= l; ++i) B[i + static_cast<size_t>(m) * j] += Rat(i, l) * Q[l + static_cast<size_t>(m) * j];
                for(int j = 0; j < k; ++j) {
                    std::vector<cx> c(B.begin() + static_cast<size_t>(m) * j, B.begin() + static_cast<size_t>(m) * (j + 1));
                    for(int pass = 0; pass < 2; ++pass)
                        for(int i = 0; i < j; ++i) {
                            cx d(0.0, 0.0);
                            for(int l = 0; l < m; ++l) d += std::conj(Qs[l + static_cast<size_t>(m) * i]) * c[l];
                            for(int l = 0; l < m; ++l) c[l] -= d * Qs[l + static_cast<size_t>(m) * i];
                            Rs[i + static_cast<size_t>(k) * j] += d;
                        }
                    double nr = 0;
                    for(auto &v : c) nr += std::norm(v);
                    nr = std::sqrt(nr);
                    Rs[j + static_cast<size_t>(k) * j] = nr;
                    for(int l = 0; l < m; ++l) Qs[l + static_cast<size_t>(m) * j] = nr > 0 ? c[l] / nr : cx(0.0, 0.0);
                }
                transform_basis(Qr.as<T>(), m, Qs, k);
                std::fill(Rr.begin(), Rr.end(), cx(0.0, 0.0));
                for(int j = 0; j < k; ++j)
                    for(int i = 0; i <= j; ++i) Rat(i, j) = Rs[i + static_cast<size_t>(k) * j];
            }
            m = k;
            ++st.restarts;
            yprev.clear();
        } else {
            yprev.assign(m, cx(0.0, 0.0));
            for(int i = 0; i < m; ++i) yprev[i] = Z[i + static_cast<size_t>(m) * ord[0]];
        }

        // ---- expansion vector t = M^-1 r (or r), orthonormalised against V
        DMRG_REQUIRE(m < mmax, DMRG_ERR_ARG, "eigensolver: basis full after restart");
        T       *t   = Vp + static_cast<int64_t>(m) * n;
        const T *src = r;
        if(inject_random) {
            random_fill<T>(n, t, p.seed + 7919ull * (st.iters + 1), s);
            src           = t;
            inject_random = false;
        } else if(p.precond != 0 || use_inner) {
            apply_precond(w[ord[pstar]], out + static_cast<int64_t>(pstar) * n, t);
            src = t;
        }
        expand_from(src, t, m, sm + SB_NRM, sm + SB_NRM + 2);
        matvec(t, Wp + static_cast<int64_t>(m) * n, 1);
        st.matvecs += 1;
        multidot<T>(n, m + 1, Vp, n, Wp + static_cast<int64_t>(m) * n, sm + SB_COL, scratch.as<double>(), s);
        if(refined) qr_append(m, p.shift);
        DMRG_HIP_CHECK(hipMemcpyAsync(host, sm, (SB_NRM + 4) * sizeof(double), hipMemcpyDeviceToHost, s));
        if(refined) qr_fetch();
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        accept_expansion(m);
        ++st.iters;
    }

    // ---- outputs: nkeep Ritz vectors (the first nconv were formed in the last pass already)
    for(int q = 0; q < std::min(nkeep, m); ++q) {
        const double rq = ritz(ord[q], m, out + q * n);
        if(q < nev) {
            evals[q]  = w[ord[q]];
            rnorms[q] = rq;
        }
    }
    for(int q = m; q < nev; ++q) {
        evals[q]  = std::numeric_limits<double>::quiet_NaN();
        rnorms[q] = std::numeric_limits<double>::infinity();
    }
    st.seconds = std::chrono::duration<double>(clock::now() - t0).count();
    return st;
}

template <typename T64, typename T32>
EigsStats eigs_mixed(MatVecOp &op, hipStream_t s, const EigsParams &p, const T64 *init, T64 *out, double *evals,
                     double *rnorms, EigsStats *phase1) {
    const int64_t n    = op.D.n();
    const int     nev  = std::max(1, p.nev);
    const int     keep = std::min<int64_t>(nev + 2, n);
    DevBuf        b32, b32i, b64;
    b32.reserve(static_cast<size_t>(n) * keep * sizeof(T32));
    b64.reserve(static_cast<size_t>(n) * keep * sizeof(T64));
    DMRG_HIP_CHECK(hipMemsetAsync(b32.p, 0, static_cast<size_t>(n) * keep * sizeof(T32), s));
    const T32 *init32 = nullptr;
    if(init) {
        b32i.reserve(static_cast<size_t>(n) * sizeof(T32));
        convert<T32, T64>(n, init, b32i.as<T32>(), s);
        init32 = b32i.as<T32>();
    }
    // Phase 1: FP32 vectors and FP32 matrix-core matvecs. The residual floor of FP32 arithmetic is
    // ~1e-6 |A|; stop at 5e-6 (or at stagnation), or already at 1e-4 when the FP64 phase expands with FP32
    // correction solves, whose steps cost less than FP32 Davidson iterations (measured at chi = 512).
    const bool corrections = p.inner_max > 0 && nev == 1 && p.precond == 0 && p.target != TARGET_CLOSEST;
    EigsParams p1          = p;
    p1.tol                 = std::max(p.tol, corrections ? 1e-4 : 5e-6);
    p1.stagnation_window = 25;
    std::vector<double> e1(nev), r1(nev);
    EigsStats           s1;
    {
        Davidson<T32> d32(op, s);
        s1 = d32.solve(p1, init32, init32 ? 1 : 0, b32.as<T32>(), e1.data(), r1.data(), keep);
    }
    if(phase1) *phase1 = s1;
    convert<T64, T32>(n * keep, b32.as<T32>(), b64.as<T64>(), s);
    // Phase 2: FP64 Davidson from the FP32 Ritz vectors.
    EigsParams p2 = p;
    p2.seed += 1;
    p2.stagnation_window = 0;
    Davidson<T64> d64(op, s);
    EigsStats     s2 = d64.solve(p2, b64.as<T64>(), keep, out, evals, rnorms, nev);
    s2.seconds += s1.seconds;
    return s2;
}

template class CorrectionSolver<float>;
template class CorrectionSolver<cpx<float>>;
template int CorrectionSolver<float>::solve<double>(const double *, const double *, double, double, int, double, double *);
template int CorrectionSolver<cpx<float>>::solve<cpx<double>>(const cpx<double> *, const cpx<double> *, double, double, int, double,
                                                             cpx<double> *);
template class Davidson<double>;
template class Davidson<float>;
template class Davidson<cpx<double>>;
template class Davidson<cpx<float>>;
template EigsStats eigs_mixed<double, float>(MatVecOp &, hipStream_t, const EigsParams &, const double *, double *, double *,
                                             double *, EigsStats *);
template EigsStats eigs_mixed<cpx<double>, cpx<float>>(MatVecOp &, hipStream_t, const EigsParams &, const cpx<double> *,
                                                       cpx<double> *, double *, double *, EigsStats *);

} // namespace dmrg

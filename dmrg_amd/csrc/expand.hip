// Subspace expansion, the optimally mixed block of get_optimally_mixed_block_1site (env.cpp:96-239): a
// Rayleigh-Ritz step on the energy variance over the Krylov space {psi, H1 psi, H2 psi}, with H1 the
// energy operator and H2 the operator on the variance environments (MPO^2). Eight matvecs on the GPU
// (two to span the space, three per operator after orthonormalisation, batched as one 3-vector matvec
// each), block dots in double, and the 3 x 3 eigenproblem on the host.
#include "expand.hpp"

#include "dense_eig.hpp"
#include "vecops.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace dmrg {

using cx = std::complex<double>;

namespace {

// Eigen's selfAdjointView<Lower>(): the lower triangle of K, mirrored.
void lower_hermitian(const cx *K, cx *H) {
    for(int j = 0; j < 3; ++j)
        for(int i = 0; i < 3; ++i) H[i + 3 * j] = i >= j ? K[i + 3 * j] : std::conj(K[j + 3 * i]);
}

int zero_rows(const cx *K, double tol) {
    int c = 0;
    for(int i = 0; i < 3; ++i) {
        double mx = 0;
        for(int j = 0; j < 3; ++j) mx = std::max(mx, std::abs(K[i + 3 * j]));
        c += mx <= tol ? 1 : 0;
    }
    return c;
}

template <typename T>
void mixed_block_impl(MatVecOp &h1, MatVecOp &h2, const T *theta, int mode, T *mixed, MixedBlockResult &res, hipStream_t s) {
    const int64_t n = h1.D.n();
    DevBuf        Vb, HVb, dev, scr;
    Vb.reserve(3 * static_cast<size_t>(n) * sizeof(T));
    HVb.reserve(3 * static_cast<size_t>(n) * sizeof(T));
    dev.reserve(64 * sizeof(double));
    scr.reserve(static_cast<size_t>(vec_scratch_doubles()) * sizeof(double));
    vec_scratch_init(scr.as<double>(), s);
    T      *V = Vb.as<T>(), *HV = HVb.as<T>();
    double *d = dev.as<double>(), *sc = scr.as<double>();
    auto    col = [&](T *B, int j) { return B + static_cast<int64_t>(j) * n; };

    // V = [psi, H1 psi, H2 psi]
    DMRG_HIP_CHECK(hipMemcpyAsync(V, theta, static_cast<size_t>(n) * sizeof(T), hipMemcpyDeviceToDevice, s));
    h1.multax_device(col(V, 0), n, col(V, 1), n, 1, s, false);
    h2.multax_device(col(V, 0), n, col(V, 2), n, 1, s, false);

    // Modified Gram-Schmidt (env.cpp:74-93): a column of norm < eps stays as it is and later columns are not
    // projected against it.
    double nrm2[2];
    for(int i = 0; i < 3; ++i) {
        multidot_x<T>(n, 0, V, n, col(V, i), d, d + 2, nullptr, sc, s);
        DMRG_HIP_CHECK(hipMemcpyAsync(nrm2, d + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        const double rii = std::sqrt(std::max(0.0, nrm2[0]));
        if(rii < std::numeric_limits<double>::epsilon()) continue;
        scale<T>(n, col(V, i), 1.0 / rii, s);
        for(int j = i + 1; j < 3; ++j) {
            multidot<T>(n, 1, col(V, i), n, col(V, j), d + 4, sc, s);              // R(i,j) = q_i^H v_j
            gemv_update<T>(n, 1, col(V, i), n, d + 4, -1.0, 1.0, col(V, j), s);    // v_j -= q_i R(i,j)
        }
    }

    // K1 = V^H H1 V and K2 = V^H H2 V (column j of K from one 3-column dot pass each)
    cx K1[9], K2[9];
    for(int op = 0; op < 2; ++op) {
        (op == 0 ? h1 : h2).multax_device(V, n, HV, n, 3, s, false);
        for(int j = 0; j < 3; ++j) multidot<T>(n, 3, V, n, col(HV, j), d + 8 + 6 * j, sc, s);
        double host[18];
        DMRG_HIP_CHECK(hipMemcpyAsync(host, d + 8, 18 * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        cx K[9];
        for(int j = 0; j < 3; ++j)
            for(int i = 0; i < 3; ++i) K[i + 3 * j] = cx(host[6 * j + 2 * i], host[6 * j + 2 * i + 1]);
        lower_hermitian(K, op == 0 ? K1 : K2);
    }
    // K3 = K2 - K1 K1, then the mode's decoupling (env.cpp:166-176)
    cx K3[9];
    for(int j = 0; j < 3; ++j)
        for(int i = 0; i < 3; ++i) {
            cx acc = K2[i + 3 * j];
            for(int l = 0; l < 3; ++l) acc -= K1[i + 3 * l] * K1[l + 3 * j];
            K3[i + 3 * j] = acc;
        }
    const bool h1_on = (mode & 1) != 0, h2_on = (mode & 2) != 0;
    if(!h1_on && K3[2 + 3 * 2] != 0.0) K3[1 + 3 * 0] = K3[1 + 3 * 1] = K3[2 + 3 * 1] = 0.0;
    if(!h2_on && K3[1 + 3 * 1] != 0.0) K3[2 + 3 * 0] = K3[2 + 3 * 1] = K3[2 + 3 * 2] = 0.0;
    cx Kh[9], Z[9];
    lower_hermitian(K3, Kh);
    double w[3];
    if(!hermitian_eig(3, Kh, 3, w, Z)) throw Error(DMRG_ERR_NOT_CONVERGED, "mixed block: 3x3 eigensolver failed");

    // Choice of the eigenvector (env.cpp:201-226): the lowest, unless a K has a row at the noise level; then
    // the one whose first component (the weight of psi) is largest.
    const double tol = std::clamp(K3[0].real() * K3[0].real(), std::numeric_limits<double>::epsilon(), 1e-10);
    const int    nz  = std::max({zero_rows(K1, tol), zero_rows(K2, tol), zero_rows(K3, tol)});
    int          opt = 0;
    if(nz > 0) {
        double best = -1;
        for(int j = 0; j < 3; ++j)
            if(std::abs(Z[3 * j]) > best) {
                best = std::abs(Z[3 * j]);
                opt  = j;
            }
    }
    // Real coefficients (env.cpp:228): the eigenvector rotated so that its largest component is real positive.
    int kmax = 0;
    for(int i = 1; i < 3; ++i)
        if(std::abs(Z[i + 3 * opt]) > std::abs(Z[kmax + 3 * opt])) kmax = i;
    const cx ph = std::abs(Z[kmax + 3 * opt]) > 0 ? Z[kmax + 3 * opt] / std::abs(Z[kmax + 3 * opt]) : cx(1.0, 0.0);
    double   c[6];
    for(int i = 0; i < 3; ++i) {
        res.alpha[i] = (Z[i + 3 * opt] / ph).real();
        c[2 * i]     = res.alpha[i];
        c[2 * i + 1] = 0.0;
    }
    res.opt_val = w[opt];
    res.opt_idx = opt;
    for(int i = 0; i < 9; ++i) {
        res.K1[2 * i] = K1[i].real(), res.K1[2 * i + 1] = K1[i].imag();
        res.K2[2 * i] = K2[i].real(), res.K2[2 * i + 1] = K2[i].imag();
        res.K3[2 * i] = K3[i].real(), res.K3[2 * i + 1] = K3[i].imag();
    }
    DMRG_HIP_CHECK(hipMemcpyAsync(d + 30, c, 6 * sizeof(double), hipMemcpyHostToDevice, s));
    gemv_update<T>(n, 3, V, n, d + 30, 1.0, 0.0, mixed, s); // mixed = V col
    DMRG_HIP_CHECK(hipStreamSynchronize(s));                  // the local buffers die at return
}

} // namespace

void optimally_mixed_block(MatVecOp &h1, MatVecOp &h2, const void *theta, int mode, void *mixed, MixedBlockResult &res, hipStream_t s) {
    DMRG_REQUIRE(h1.dtype == h2.dtype, DMRG_ERR_ARG, "mixed block: both operators must have the same dtype");
    DMRG_REQUIRE(h1.D.d == h2.D.d && h1.D.chiL == h2.D.chiL && h1.D.chiR == h2.D.chiR, DMRG_ERR_DIMS,
                 "mixed block: operators act on different shapes");
    if(h1.dtype == DMRG_FP64) mixed_block_impl<double>(h1, h2, static_cast<const double *>(theta), mode, static_cast<double *>(mixed), res, s);
    else if(h1.dtype == DMRG_CX64)
        mixed_block_impl<cpx<double>>(h1, h2, static_cast<const cpx<double> *>(theta), mode, static_cast<cpx<double> *>(mixed), res, s);
    else throw Error(DMRG_ERR_ARG, "mixed block: FP64 or CX64 operators only");
}

} // namespace dmrg

// Device SVD (rocSOLVER) and remapping copies used by the sweep.
#include "svd.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

namespace dmrg {

#define DMRG_ROCBLAS_CHECK(expr)                                                                             \
    do {                                                                                                     \
        rocblas_status st_ = (expr);                                                                         \
        if(st_ != rocblas_status_success)                                                                    \
            throw ::dmrg::Error(DMRG_ERR_HIP, std::string(#expr) + " failed: " + rocblas_status_to_string(st_)); \
    } while(0)

DeviceSvd::DeviceSvd(hipStream_t st) : s(st) {
    rocblas_handle h = nullptr;
    DMRG_ROCBLAS_CHECK(rocblas_create_handle(&h));
    DMRG_ROCBLAS_CHECK(rocblas_set_stream(h, s));
    handle = h;
}
void DeviceSvd::set_stream(hipStream_t st) {
    s = st;
    DMRG_ROCBLAS_CHECK(rocblas_set_stream(static_cast<rocblas_handle>(handle), s));
}
DeviceSvd::~DeviceSvd() {
    if(handle) (void) rocblas_destroy_handle(static_cast<rocblas_handle>(handle));
}

template <typename T>
void DeviceSvd::compute(int m, int n, const T *A, std::vector<double> &S, int method) {
    DMRG_REQUIRE(m > 0 && n > 0, DMRG_ERR_DIMS, "svd of an empty matrix");
    DMRG_REQUIRE(method >= 0 && method <= 3, DMRG_ERR_ARG, "svd method must be 0 (gesvd), 1 (gesdd), 2 (gesvdj) or 3 (Jacobi)");
    // rocSOLVER's gesdd and gesvdj obtain the small singular values through A^H A (errors of order
    // eps * S[0]^2 / S[k] were measured: 1e-9 relative to S[0] on graded spectra), which misses the
    // 1e-12 parity of the reference's LAPACK routines; only the bidiagonal QR path and the one-sided Jacobi
    // kernels are offered.
    DMRG_REQUIRE(method == 0 || method == 3, DMRG_ERR_UNSUPPORTED,
                 "svd: methods 1 (gesdd) and 2 (gesvdj) lose the small singular values; use 0 (gesvd) or 3 (Jacobi)");
    if(method == 3) {
        jacobi<T>(m, n, A, S);
        return;
    }
    k_ = std::min(m, n);
    a.reserve(static_cast<size_t>(m) * n * sizeof(T));
    sv.reserve(static_cast<size_t>(k_) * sizeof(double));
    u.reserve(static_cast<size_t>(m) * k_ * sizeof(T));
    vt.reserve(static_cast<size_t>(k_) * n * sizeof(T));
    e.reserve(static_cast<size_t>(std::max(1, k_ - 1)) * sizeof(double));
    info.reserve(sizeof(rocblas_int) * 2);
    res.reserve(sizeof(double));
    DMRG_HIP_CHECK(hipMemcpyAsync(a.p, A, static_cast<size_t>(m) * n * sizeof(T), hipMemcpyDeviceToDevice, s));
    auto        h  = static_cast<rocblas_handle>(handle);
    const auto  sv_ = rocblas_svect_singular;
    rocblas_int *inf = info.as<rocblas_int>();
    if constexpr(std::is_same_v<T, double>) {
        double *pa = a.as<double>(), *pu = u.as<double>(), *pv = vt.as<double>();
        if(method == 0)
            DMRG_ROCBLAS_CHECK(rocsolver_dgesvd(h, sv_, sv_, m, n, pa, m, sv.as<double>(), pu, m, pv, k_, e.as<double>(),
                                                rocblas_outofplace, inf));
        else if(method == 1)
            DMRG_ROCBLAS_CHECK(rocsolver_dgesdd(h, sv_, sv_, m, n, pa, m, sv.as<double>(), pu, m, pv, k_, inf));
        else
            DMRG_ROCBLAS_CHECK(rocsolver_dgesvdj(h, sv_, sv_, m, n, pa, m, 0.0, res.as<double>(), 100, inf + 1,
                                                 sv.as<double>(), pu, m, pv, k_, inf));
    } else {
        using Z = rocblas_double_complex;
        Z *pa = a.as<Z>(), *pu = u.as<Z>(), *pv = vt.as<Z>();
        if(method == 0)
            DMRG_ROCBLAS_CHECK(rocsolver_zgesvd(h, sv_, sv_, m, n, pa, m, sv.as<double>(), pu, m, pv, k_, e.as<double>(),
                                                rocblas_outofplace, inf));
        else if(method == 1)
            DMRG_ROCBLAS_CHECK(rocsolver_zgesdd(h, sv_, sv_, m, n, pa, m, sv.as<double>(), pu, m, pv, k_, inf));
        else
            DMRG_ROCBLAS_CHECK(rocsolver_zgesvdj(h, sv_, sv_, m, n, pa, m, 0.0, res.as<double>(), 100, inf + 1,
                                                 sv.as<double>(), pu, m, pv, k_, inf));
    }
    S.resize(k_);
    rocblas_int hinf = 0;
    DMRG_HIP_CHECK(hipMemcpyAsync(S.data(), sv.p, k_ * sizeof(double), hipMemcpyDeviceToHost, s));
    DMRG_HIP_CHECK(hipMemcpyAsync(&hinf, info.p, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
    DMRG_REQUIRE(hinf == 0, DMRG_ERR_NOT_CONVERGED, "svd: rocSOLVER reported no convergence (info " + std::to_string(hinf) + ")");
    // Values at the rounding level can come back out of order; restore descending order (stable) and
    // permute the singular vectors to match.
    bool sorted = true;
    for(int i = 1; i < k_ && sorted; ++i) sorted = S[i] <= S[i - 1];
    if(!sorted) sort_descending<T>(m, n, S);
}

template <typename E>
__global__ void permute_vectors_kernel(int64_t m, int64_t n, int64_t k, const int *perm, const E *u, E *u_out, const E *vt,
                                       E *vt_out) {
    // u_out[:, j] = u[:, perm[j]] (m x k), vt_out[j, :] = vt[perm[j], :] (k x n, leading dimension k)
    const int64_t tu = m * k, tv = k * n;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < tu + tv;
        t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if(t < tu) {
            const int64_t i = t % m, j = t / m;
            u_out[t]        = u[i + m * perm[j]];
        } else {
            const int64_t q = t - tu, j = q % k, c = q / k;
            vt_out[q]       = vt[perm[j] + k * c];
        }
    }
}

template <typename T>
void DeviceSvd::sort_descending(int m, int n, std::vector<double> &S) {
    std::vector<int> perm(k_);
    for(int i = 0; i < k_; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return S[a] > S[b]; });
    std::vector<double> Ss(k_);
    for(int i = 0; i < k_; ++i) Ss[i] = S[perm[i]];
    S.swap(Ss);
    const size_t es = sizeof(T);
    DevBuf       dperm, u2, vt2;
    dperm.reserve(static_cast<size_t>(k_) * sizeof(int));
    u2.reserve(static_cast<size_t>(m) * k_ * es);
    vt2.reserve(static_cast<size_t>(k_) * n * es);
    DMRG_HIP_CHECK(hipMemcpyAsync(dperm.p, perm.data(), k_ * sizeof(int), hipMemcpyHostToDevice, s));
    const int64_t total = static_cast<int64_t>(m) * k_ + static_cast<int64_t>(k_) * n;
    hipLaunchKernelGGL(permute_vectors_kernel<T>, dim3(static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 8192))), dim3(256),
                       0, s, m, n, k_, dperm.as<int>(), u.as<T>(), u2.as<T>(), vt.as<T>(), vt2.as<T>());
    DMRG_HIP_CHECK(hipGetLastError());
    DMRG_HIP_CHECK(hipMemcpyAsync(sv.p, S.data(), k_ * sizeof(double), hipMemcpyHostToDevice, s));
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
    std::swap(u, u2);
    std::swap(vt, vt2);
}
template void DeviceSvd::compute<double>(int, int, const double *, std::vector<double> &, int);
template void DeviceSvd::compute<cpx<double>>(int, int, const cpx<double> *, std::vector<double> &, int);

template <typename T>
void DeviceSvd::eigh(int n, T *A, int lda, bool vectors, std::vector<double> &w) {
    DMRG_REQUIRE(n > 0 && lda >= n, DMRG_ERR_DIMS, "eigh: need n > 0 and lda >= n");
    sv.reserve(static_cast<size_t>(n) * sizeof(double));
    e.reserve(static_cast<size_t>(n) * sizeof(double));
    info.reserve(sizeof(rocblas_int) * 2);
    auto       h  = static_cast<rocblas_handle>(handle);
    const auto ev = vectors ? rocblas_evect_original : rocblas_evect_none;
    if constexpr(std::is_same_v<T, double>)
        DMRG_ROCBLAS_CHECK(rocsolver_dsyevd(h, ev, rocblas_fill_lower, n, A, lda, sv.as<double>(), e.as<double>(),
                                            info.as<rocblas_int>()));
    else
        DMRG_ROCBLAS_CHECK(rocsolver_zheevd(h, ev, rocblas_fill_lower, n, reinterpret_cast<rocblas_double_complex *>(A), lda,
                                            sv.as<double>(), e.as<double>(), info.as<rocblas_int>()));
    w.resize(n);
    rocblas_int hinf = 0;
    DMRG_HIP_CHECK(hipMemcpyAsync(w.data(), sv.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    DMRG_HIP_CHECK(hipMemcpyAsync(&hinf, info.p, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
    DMRG_REQUIRE(hinf == 0, DMRG_ERR_NOT_CONVERGED, "eigh: rocSOLVER reported no convergence (info " + std::to_string(hinf) + ")");
}
template void DeviceSvd::eigh<double>(int, double *, int, bool, std::vector<double> &);
template void DeviceSvd::eigh<cpx<double>>(int, cpx<double> *, int, bool, std::vector<double> &);

std::pair<int, double> rank_from_truncation_error(const std::vector<double> &S, double lim, int rank_max, int rank_min) {
    const int           k = static_cast<int>(S.size());
    std::vector<double> terr(k + 1, 0.0);
    double              acc = 0.0;
    for(int i = k - 1; i >= 0; --i) {
        acc += S[i] * S[i];
        terr[i] = std::sqrt(acc);
    }
    terr[k]  = 0.0;
    int rank = 0;
    for(int i = 0; i <= k; ++i)
        if(terr[i] >= lim) ++rank;
    int rank_lim = k;
    if(rank_max > 0) rank_lim = std::min(k, rank_max);
    rank = std::min(rank, rank_lim);
    if(rank_min > 0) rank = std::max(rank, std::min(k, rank_min));
    DMRG_REQUIRE(rank > 0, DMRG_ERR_ARG, "rank <= 0");
    return {rank, terr[rank]};
}

template <typename E>
__global__ void remap2d_kernel(int64_t M, int64_t N, const E *in, Idx2 im, Idx2 in_, E *out, Idx2 om, Idx2 on,
                               const double *scale, int rs) {
    const int64_t total = M * N;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < total; t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint32_t i = static_cast<uint32_t>(t % M), j = static_cast<uint32_t>(t / M);
        E              v = in[idx_off(im, i) + idx_off(in_, j)];
        if(rs) {
            const double f = scale[rs == 1 ? i : j];
            if constexpr(std::is_same_v<E, cpx<double>>) {
                v.re *= f;
                v.im *= f;
            } else {
                v *= f;
            }
        }
        out[idx_off(om, i) + idx_off(on, j)] = v;
    }
}

void remap2d(dmrg_dtype t, int64_t M, int64_t N, const void *in, Idx2 im, Idx2 in_, void *out, Idx2 om, Idx2 on,
             const double *d_scale, int rs, hipStream_t s) {
    if(M <= 0 || N <= 0) return;
    const int64_t total = M * N;
    const dim3    g(static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 8192))), b(256);
    if(t == DMRG_FP64)
        hipLaunchKernelGGL(remap2d_kernel<double>, g, b, 0, s, M, N, static_cast<const double *>(in), im, in_,
                           static_cast<double *>(out), om, on, d_scale, rs);
    else if(t == DMRG_CX64)
        hipLaunchKernelGGL(remap2d_kernel<cpx<double>>, g, b, 0, s, M, N, static_cast<const cpx<double> *>(in), im, in_,
                           static_cast<cpx<double> *>(out), om, on, d_scale, rs);
    else
        throw Error(DMRG_ERR_UNSUPPORTED, "remap2d: FP64/CX64 only");
    DMRG_HIP_CHECK(hipGetLastError());
}

} // namespace dmrg

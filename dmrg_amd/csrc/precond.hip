// Block-Jacobi preconditioner of the H_eff operator: explicit inverses of the diagonal blocks of
// (H_eff - shift), factorised with batched LU (rocSOLVER getrf/getri, pivoted) and applied as one batched
// GEMV (rocBLAS). Mirrors MatVecMPOS::CalcPc / MultPc for jcbMaxBlockSize > 1 (matvec_mpos.cpp:1086-1331,
// 1362-1378): blocks at offsets k*jcb, the last one possibly smaller.
//
// Factorizations (the operator's eig::Factorization): LU, LDLT and QR all solve a nonsingular block exactly, so
// they share the LU inverse. LLT (the reference's choice for H^2, optimize_folded_spectrum.cpp:309) factorises
// -B instead of B when Re B(0,0) <= 0 (matvec_mpos.cpp:1150-1154) and falls through to LDLT, i.e. the exact
// inverse of B, when that Cholesky factorisation fails (:1172-1190). Its preconditioner is therefore -B^-1 on the
// blocks with Re B(0,0) <= 0 whose negation is positive definite and B^-1 on all others: the LU inverses, with
// the sign of each block decided by a batched Cholesky factorisation (rocSOLVER potrf) of the sign-adjusted copy.
#include "dense.hpp"
#include "matvec_op.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <vector>

namespace dmrg {

#define DMRG_ROCBLAS_OK(expr)                                                                                \
    do {                                                                                                     \
        rocblas_status st_ = (expr);                                                                         \
        if(st_ != rocblas_status_success)                                                                    \
            throw ::dmrg::Error(DMRG_ERR_HIP, std::string(#expr) + " failed: " + rocblas_status_to_string(st_)); \
    } while(0)

// Inverses of all blocks may not exceed this many bytes (n * jcb elements).
constexpr size_t BLOCK_PC_MAX_BYTES = size_t(16) << 30;

MatVecOp::~MatVecOp() {
    if(blas) (void) rocblas_destroy_handle(static_cast<rocblas_handle>(blas));
    if(staging_ev) {
        (void) hipEventSynchronize(staging_ev);
        (void) hipEventDestroy(staging_ev);
    }
    if(staging) pinned_put(staging, staging_bytes);
}

namespace {

// Copies block k of the diagonal blocks (b x b each) negated when Re B_k(0,0) <= 0, and records that choice.
template <typename E>
__global__ void signed_block_copy(int64_t nb, int64_t b, const E *src, E *dst, int *negated) {
    const int64_t bb = b * b;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < nb * bb;
        t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t k   = t / bb;
        const E       b00 = src[k * bb];
        bool          neg;
        if constexpr(sizeof(E) == 16) neg = reinterpret_cast<const double *>(&b00)[0] <= 0.0;
        else neg = b00 <= 0.0;
        E v = src[t];
        if(neg) {
            if constexpr(sizeof(E) == 16) {
                reinterpret_cast<double *>(&v)[0] = -reinterpret_cast<double *>(&v)[0];
                reinterpret_cast<double *>(&v)[1] = -reinterpret_cast<double *>(&v)[1];
            } else {
                v = -v;
            }
        }
        dst[t] = v;
        if(t % bb == 0) negated[k] = neg ? 1 : 0;
    }
}

// Multiplies block k of the inverses by sign[k] (+1 or -1).
template <typename E>
__global__ void sign_blocks(int64_t nb, int64_t b, const int *sign, E *inv) {
    const int64_t bb = b * b;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < nb * bb;
        t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if(sign[t / bb] >= 0) continue;
        if constexpr(sizeof(E) == 16) {
            double *v = reinterpret_cast<double *>(inv + t);
            v[0] = -v[0], v[1] = -v[1];
        } else {
            inv[t] = -inv[t];
        }
    }
}

} // namespace

void MatVecOp::set_factorization(int f) {
    // eig::Factorization: NONE 0, LLT 1, LDLT 2, LU 3, QR 4, ILUT 5, ILDLT 6 (enums.h:30).
    DMRG_REQUIRE(f >= 0 && f <= 6, DMRG_ERR_UNSUPPORTED,
                 "block-Jacobi factorization " + std::to_string(f) + ": an eig::Factorization value 0..6");
    const int g = f == 1 ? 1 : 3; // every other choice applies the exact (LU) inverse
    if(g != pc_fact) pc_ready = false;
    pc_fact = g;
}

void MatVecOp::set_jcb(int64_t size) {
    DMRG_REQUIRE(size >= 1, DMRG_ERR_ARG, "block size must be >= 1");
    if(size != jcb) pc_ready = false;
    jcb = size;
}

void MatVecOp::block_pc_build(double shift, hipStream_t s) {
    DMRG_REQUIRE(dtype == DMRG_FP64 || dtype == DMRG_CX64, DMRG_ERR_UNSUPPORTED, "block-Jacobi: FP64 or CX64 operator");
    const int64_t n  = D.n();
    const int64_t b  = std::min<int64_t>(jcb, n);
    const int64_t nb = (n + b - 1) / b;
    const size_t  es = dtype_size(dtype);
    DMRG_REQUIRE(static_cast<size_t>(nb * b * b) * es <= BLOCK_PC_MAX_BYTES, DMRG_ERR_OOM,
                 "block-Jacobi inverses would exceed " + std::to_string(BLOCK_PC_MAX_BYTES >> 30) + " GiB; use a smaller block size");
    DMRG_REQUIRE(b < INT32_MAX && nb < INT32_MAX, DMRG_ERR_DIMS, "block-Jacobi: too many blocks");
    if(!blas) {
        rocblas_handle h = nullptr;
        DMRG_ROCBLAS_OK(rocblas_create_handle(&h));
        blas = h;
    }
    auto h = static_cast<rocblas_handle>(blas);
    DMRG_ROCBLAS_OK(rocblas_set_stream(h, s));
    pcinv.reserve(static_cast<size_t>(nb * b * b) * es);
    pcpiv.reserve(static_cast<size_t>(nb * b) * sizeof(rocblas_int));
    pcinfo.reserve(static_cast<size_t>(nb) * sizeof(rocblas_int));
    heff_diag_blocks(dtype, D, mpo.p, envL_p, envR_p, b, shift, pcinv.p, s);
    const rocblas_int  bi = static_cast<rocblas_int>(b), nbi = static_cast<rocblas_int>(nb);
    const rocblas_stride sa = b * b, sp = b;
    std::vector<int> sign;
    if(pc_fact == 1) { // LLT: which blocks take the negated inverse (see the top of this file)
        pcchol.reserve(static_cast<size_t>(nb * b * b) * es);
        pcneg.reserve(static_cast<size_t>(nb) * 2 * sizeof(int));
        int          *neg   = pcneg.as<int>();
        rocblas_int  *cinfo = reinterpret_cast<rocblas_int *>(neg + nb);
        const int64_t total = nb * b * b;
        const dim3    grid(static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 8192)));
        if(dtype == DMRG_FP64) {
            hipLaunchKernelGGL(signed_block_copy<double>, grid, dim3(256), 0, s, nb, b, pcinv.as<double>(), pcchol.as<double>(), neg);
            DMRG_HIP_CHECK(hipGetLastError());
            DMRG_ROCBLAS_OK(rocsolver_dpotrf_strided_batched(h, rocblas_fill_lower, bi, pcchol.as<double>(), bi, sa, cinfo, nbi));
        } else {
            hipLaunchKernelGGL(signed_block_copy<rocblas_double_complex>, grid, dim3(256), 0, s, nb, b,
                               pcinv.as<rocblas_double_complex>(), pcchol.as<rocblas_double_complex>(), neg);
            DMRG_HIP_CHECK(hipGetLastError());
            DMRG_ROCBLAS_OK(rocsolver_zpotrf_strided_batched(h, rocblas_fill_lower, bi, pcchol.as<rocblas_double_complex>(), bi, sa,
                                                             cinfo, nbi));
        }
        std::vector<int> host(2 * nb);
        DMRG_HIP_CHECK(hipMemcpyAsync(host.data(), neg, host.size() * sizeof(int), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        sign.assign(nb, 1);
        bool any = false;
        for(int64_t k = 0; k < nb; ++k)
            if(host[k] && host[nb + k] == 0) sign[k] = -1, any = true; // -B is positive definite
        if(!any) sign.clear();
    }
    if(dtype == DMRG_FP64) {
        DMRG_ROCBLAS_OK(rocsolver_dgetrf_strided_batched(h, bi, bi, pcinv.as<double>(), bi, sa, pcpiv.as<rocblas_int>(), sp,
                                                         pcinfo.as<rocblas_int>(), nbi));
        DMRG_ROCBLAS_OK(rocsolver_dgetri_strided_batched(h, bi, pcinv.as<double>(), bi, sa, pcpiv.as<rocblas_int>(), sp,
                                                         pcinfo.as<rocblas_int>(), nbi));
    } else {
        auto *a = pcinv.as<rocblas_double_complex>();
        DMRG_ROCBLAS_OK(rocsolver_zgetrf_strided_batched(h, bi, bi, a, bi, sa, pcpiv.as<rocblas_int>(), sp, pcinfo.as<rocblas_int>(), nbi));
        DMRG_ROCBLAS_OK(rocsolver_zgetri_strided_batched(h, bi, a, bi, sa, pcpiv.as<rocblas_int>(), sp, pcinfo.as<rocblas_int>(), nbi));
    }
    std::vector<rocblas_int> info(nb);
    DMRG_HIP_CHECK(hipMemcpyAsync(info.data(), pcinfo.p, nb * sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
    for(int64_t k = 0; k < nb; ++k)
        DMRG_REQUIRE(info[k] == 0, DMRG_ERR_NOT_CONVERGED,
                     "block-Jacobi: diagonal block " + std::to_string(k) + " of (H_eff - shift) is singular");
    if(!sign.empty()) {
        int *dsign = pcneg.as<int>(); // the flags are no longer needed
        DMRG_HIP_CHECK(hipMemcpyAsync(dsign, sign.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
        const int64_t total = nb * b * b;
        const dim3    grid(static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 8192)));
        if(dtype == DMRG_FP64) hipLaunchKernelGGL(sign_blocks<double>, grid, dim3(256), 0, s, nb, b, dsign, pcinv.as<double>());
        else hipLaunchKernelGGL(sign_blocks<rocblas_double_complex>, grid, dim3(256), 0, s, nb, b, dsign, pcinv.as<rocblas_double_complex>());
        DMRG_HIP_CHECK(hipGetLastError());
        DMRG_HIP_CHECK(hipStreamSynchronize(s)); // sign dies here
    }
    pc_shift = shift;
    pc_ready = true;
}

void MatVecOp::block_pc_apply(const void *x, void *y, hipStream_t s) {
    DMRG_REQUIRE(pc_ready, DMRG_ERR_ARG, "block-Jacobi preconditioner not built");
    const int64_t n  = D.n();
    const int64_t b  = std::min<int64_t>(jcb, n);
    const int64_t nb = (n + b - 1) / b;
    const size_t  es = dtype_size(dtype);
    auto          h  = static_cast<rocblas_handle>(blas);
    DMRG_ROCBLAS_OK(rocblas_set_stream(h, s));
    const void *xi = x;
    void       *yo = y;
    if(nb * b != n) { // ragged last block: pad to nb*b with zeros
        pcx.reserve(static_cast<size_t>(nb * b) * es);
        pcy.reserve(static_cast<size_t>(nb * b) * es);
        DMRG_HIP_CHECK(hipMemcpyAsync(pcx.p, x, static_cast<size_t>(n) * es, hipMemcpyDeviceToDevice, s));
        DMRG_HIP_CHECK(hipMemsetAsync(pcx.as<char>() + static_cast<size_t>(n) * es, 0, static_cast<size_t>(nb * b - n) * es, s));
        xi = pcx.p;
        yo = pcy.p;
    }
    const rocblas_int    bi = static_cast<rocblas_int>(b), nbi = static_cast<rocblas_int>(nb);
    const rocblas_stride sa = b * b, sv = b;
    if(dtype == DMRG_FP64) {
        const double one = 1.0, zero = 0.0;
        DMRG_ROCBLAS_OK(rocblas_dgemv_strided_batched(h, rocblas_operation_none, bi, bi, &one, pcinv.as<double>(), bi, sa,
                                                      static_cast<const double *>(xi), 1, sv, &zero, static_cast<double *>(yo), 1, sv,
                                                      nbi));
    } else {
        const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
        DMRG_ROCBLAS_OK(rocblas_zgemv_strided_batched(h, rocblas_operation_none, bi, bi, &one, pcinv.as<rocblas_double_complex>(),
                                                      bi, sa, static_cast<const rocblas_double_complex *>(xi), 1, sv, &zero,
                                                      static_cast<rocblas_double_complex *>(yo), 1, sv, nbi));
    }
    if(yo != y) DMRG_HIP_CHECK(hipMemcpyAsync(y, yo, static_cast<size_t>(n) * es, hipMemcpyDeviceToDevice, s));
}

} // namespace dmrg

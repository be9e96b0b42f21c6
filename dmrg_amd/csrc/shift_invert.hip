// Shift-invert for small problems: MatVecMPOS::FactorOP and MultOPv (matvec_mpos.cpp:887-975), which the reference
// uses with ARPACK's shift-invert mode. The dense H_eff of the operator (as set_shift left it) is formed on the
// device, in the orientation whose product with x is the matvec, and factored with rocSOLVER: Cholesky (potrf) for
// LLT, pivoted LU (getrf) for LU and also for LDLT, which solves with the same matrix exactly. MultOPv then solves
// with all of a call's vectors at once (potrs / getrs).
#include "dense.hpp"
#include "matvec_op.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <string>

namespace dmrg {

#define DMRG_ROCBLAS_OK(expr)                                                                                \
    do {                                                                                                     \
        rocblas_status st_ = (expr);                                                                         \
        if(st_ != rocblas_status_success)                                                                    \
            throw ::dmrg::Error(DMRG_ERR_HIP, std::string(#expr) + " failed: " + rocblas_status_to_string(st_)); \
    } while(0)

// The dense factor takes n^2 elements; larger operators belong to the iterative solver.
constexpr size_t FACTOR_OP_MAX_BYTES = static_cast<size_t>(8) << 30;

void MatVecOp::factor_op(int f, hipStream_t s) {
    // eig::Factorization: NONE 0, LLT 1, LDLT 2, LU 3 (enums.h:30); FactorOP has no branch for the others.
    DMRG_REQUIRE(f >= 0 && f <= 3, DMRG_ERR_UNSUPPORTED, "FactorOP: factorization NONE (0), LLT (1), LDLT (2) or LU (3)");
    DMRG_REQUIRE(dtype == DMRG_FP64 || dtype == DMRG_CX64, DMRG_ERR_UNSUPPORTED, "FactorOP: FP64 or CX64 operator");
    op_ready = false;
    op_fact  = f;
    if(f == 0) { // matvec_mpos.cpp:891-894: ready, nothing factored
        op_ready = true;
        return;
    }
    const int64_t n  = D.n();
    const size_t  es = dtype_size(dtype);
    DMRG_REQUIRE(static_cast<size_t>(n) * n * es <= FACTOR_OP_MAX_BYTES && n < INT32_MAX, DMRG_ERR_OOM,
                 "FactorOP: the dense H_eff (" + std::to_string(n) + "^2 elements) exceeds " +
                     std::to_string(FACTOR_OP_MAX_BYTES >> 30) + " GiB");
    if(!blas) {
        rocblas_handle h = nullptr;
        DMRG_ROCBLAS_OK(rocblas_create_handle(&h));
        blas = h;
    }
    auto h = static_cast<rocblas_handle>(blas);
    DMRG_ROCBLAS_OK(rocblas_set_stream(h, s));
    opA.reserve(static_cast<size_t>(n) * n * es);
    opinfo.reserve(sizeof(rocblas_int));
    heff_dense(dtype, D, mpo.p, envL_p, envR_p, opA.p, n, false, s);
    const rocblas_int ni = static_cast<rocblas_int>(n);
    if(f == 1) {
        if(dtype == DMRG_FP64) DMRG_ROCBLAS_OK(rocsolver_dpotrf(h, rocblas_fill_lower, ni, opA.as<double>(), ni, opinfo.as<rocblas_int>()));
        else
            DMRG_ROCBLAS_OK(rocsolver_zpotrf(h, rocblas_fill_lower, ni, opA.as<rocblas_double_complex>(), ni, opinfo.as<rocblas_int>()));
    } else {
        oppiv.reserve(static_cast<size_t>(n) * sizeof(rocblas_int));
        if(dtype == DMRG_FP64)
            DMRG_ROCBLAS_OK(rocsolver_dgetrf(h, ni, ni, opA.as<double>(), ni, oppiv.as<rocblas_int>(), opinfo.as<rocblas_int>()));
        else
            DMRG_ROCBLAS_OK(rocsolver_zgetrf(h, ni, ni, opA.as<rocblas_double_complex>(), ni, oppiv.as<rocblas_int>(),
                                             opinfo.as<rocblas_int>()));
    }
    rocblas_int info = 0;
    DMRG_HIP_CHECK(hipMemcpyAsync(&info, opinfo.p, sizeof(info), hipMemcpyDeviceToHost, s));
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
    DMRG_REQUIRE(info == 0, DMRG_ERR_ARG,
                 f == 1 ? "FactorOP (LLT): H_eff - shift is not positive definite (leading minor " + std::to_string(info) + ")"
                        : "FactorOP (LU): H_eff - shift is singular (zero pivot " + std::to_string(info) + ")");
    op_ready = true;
}

void MatVecOp::mult_opv_host(const void *x, int64_t ldx, void *y, int64_t ldy, int nvec) {
    DMRG_REQUIRE(op_ready, DMRG_ERR_ARG, "MultOPv: call FactorOP first (after any set_shift)");
    DMRG_REQUIRE(op_fact != 0, DMRG_ERR_UNSUPPORTED, "MultOPv: factorization NONE holds no factor to solve with");
    if(nvec == 0) return;
    const int64_t n  = D.n();
    const size_t  es = dtype_size(dtype);
    DMRG_REQUIRE(ldx >= n && ldy >= n, DMRG_ERR_ARG, "leading dimension smaller than rows()");
    hipStream_t s = nullptr;
    auto        h = static_cast<rocblas_handle>(blas);
    DMRG_ROCBLAS_OK(rocblas_set_stream(h, s));
    opB_rhs.reserve(static_cast<size_t>(nvec) * n * es);
    DMRG_HIP_CHECK(hipMemcpy2D(opB_rhs.p, n * es, x, ldx * es, n * es, nvec, hipMemcpyHostToDevice));
    const rocblas_int ni = static_cast<rocblas_int>(n), nr = static_cast<rocblas_int>(nvec);
    if(op_fact == 1) {
        if(dtype == DMRG_FP64) DMRG_ROCBLAS_OK(rocsolver_dpotrs(h, rocblas_fill_lower, ni, nr, opA.as<double>(), ni, opB_rhs.as<double>(), ni));
        else
            DMRG_ROCBLAS_OK(rocsolver_zpotrs(h, rocblas_fill_lower, ni, nr, opA.as<rocblas_double_complex>(), ni,
                                             opB_rhs.as<rocblas_double_complex>(), ni));
    } else {
        if(dtype == DMRG_FP64)
            DMRG_ROCBLAS_OK(rocsolver_dgetrs(h, rocblas_operation_none, ni, nr, opA.as<double>(), ni, oppiv.as<rocblas_int>(),
                                             opB_rhs.as<double>(), ni));
        else
            DMRG_ROCBLAS_OK(rocsolver_zgetrs(h, rocblas_operation_none, ni, nr, opA.as<rocblas_double_complex>(), ni,
                                             oppiv.as<rocblas_int>(), opB_rhs.as<rocblas_double_complex>(), ni));
    }
    DMRG_HIP_CHECK(hipMemcpy2D(y, ldy * es, opB_rhs.p, n * es, n * es, nvec, hipMemcpyDeviceToHost));
    num_op += nvec;
}

} // namespace dmrg

// C++ drop-in for the reference's H_eff operator and contraction API, forwarding to the C ABI of
// libdmrg_amd (include/dmrg_amd.h). Header-only; link with -ldmrg_amd.
//
//   dmrg_amd::MatVecMPOS<T>     replaces MatVecMPOS<T> (source/math/eig/matvec/matvec_mpos.h:23-180):
//                               same constructor inputs (MPO list + env blocks in cx64 storage), rows(),
//                               get_shape_mps(), MultAx(T*, T*), the PRIMME callback
//                               MultAx(void*, int*, void*, int*, int*, primme_params*, int*), set_shift(),
//                               num_mv. Env blocks and MPOs live in HBM for the operator's lifetime.
//   dmrg_amd::contraction::matrix_vector_product<Scalar>
//                               replaces tools::common::contraction::matrix_vector_product
//                               (source/tools/common/contraction.h:31-36) with host pointers.
//   dmrg_amd::contraction::expectation_value<Scalar> (contraction_tblis.cpp:30-97).
//
// MatVecMPOS<T> also carries the members the reference's solver templates read, with the same types:
//   eig::solver::eigs_primme<MatrixProductType> (solver_primme.cpp:297-517): Scalar, can_shift /
//     can_shift_invert / can_precondition, set_shift, isReadyShift, isReadyFactorOp, preconditioner,
//     set_jcbMaxBlockSize(std::optional<long>), rows, num_op, t_multOPv, and the MultPc callback of
//     preconditioner_jacobi (optimize_energy.cpp:34-42; optimize_folded_spectrum.cpp:31-39);
//   eig::solver_arpack<MatrixType> (solver_arpack.cpp:77-244, 320-380): set_mode, set_side, get_form, get_side,
//     rows, MultAx(T*, T*) (also as the member pointer &MatrixType::MultAx) and the fields num_mv, num_op.
// Block-Jacobi MultPc (jcbMaxBlockSize > 1) runs on the GPU with the factorization chosen by the field
// factorization (LLT, or LU for every other choice; matvec_mpos.cpp:1086-1331). Shift-invert: FactorOP factors
// the dense H_eff (with its shift) on the GPU with the same field (LLT: Cholesky; LU and LDLT: pivoted LU; NONE:
// ready without a factor, as the reference) and MultOPv solves with it (matvec_mpos.cpp:887-975), for PRIMME's
// matrixMatvec and for ARPACK's shift-invert mode.
#pragma once
#include "../dmrg_amd.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <complex>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

struct primme_params; // PRIMME's parameter block: only passed through the callback signature

namespace dmrg_amd {

class error : public std::runtime_error {
  public:
    dmrg_status status;
    error(dmrg_status s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(dmrg_status s, const char *where) {
    if(s != DMRG_OK) throw error(s, std::string(where) + ": " + dmrg_last_error());
}

template <typename T>
constexpr dmrg_dtype dtype_of() {
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, std::complex<double>>, "fp64 or cx64");
    return std::is_same_v<T, double> ? DMRG_FP64 : DMRG_CX64;
}

// Accepts assignments from any enumeration (eig::Preconditioner, eig::Factorization, eig::Form ...), keeps the
// value and compares equal to the enumerator of that value, so that the reference's expressions
// (matrix.preconditioner = eig::Preconditioner::JACOBI; matrix.get_form() != Form::SYMM) compile unchanged.
struct EnumValue {
    int value = 0;
    EnumValue() = default;
    explicit EnumValue(int v) : value(v) {}
    template <typename E, typename = std::enable_if_t<std::is_enum_v<E>>>
    EnumValue(E e) : value(static_cast<int>(e)) {}
    template <typename E, typename = std::enable_if_t<std::is_enum_v<E>>>
    EnumValue &operator=(E e) {
        value = static_cast<int>(e);
        return *this;
    }
    template <typename E, typename = std::enable_if_t<std::is_enum_v<E>>>
    bool operator==(E e) const { return value == static_cast<int>(e); }
    template <typename E, typename = std::enable_if_t<std::is_enum_v<E>>>
    bool operator!=(E e) const { return value != static_cast<int>(e); }
    template <typename E, typename = std::enable_if_t<std::is_enum_v<E>>>
    operator E() const { return static_cast<E>(value); }
};

// Accumulating wall-clock timer with the get_time() the reference's tid::ur timers expose.
struct Timer {
    double seconds = 0.0;
    double get_time() const { return seconds; }
};

template <typename T>
class MatVecMPOS {
  public:
    using Scalar                            = T;
    static constexpr bool can_shift         = true;
    static constexpr bool can_shift_invert  = true;
    static constexpr bool can_precondition  = true;
    EnumValue             preconditioner;   // eig::Preconditioner: NONE (0) or JACOBI
    EnumValue             factorization;    // eig::Factorization of the block-Jacobi blocks and of FactorOP
    long                  num_mv = 0, num_op = 0, num_pc = 0; // matvec_mpos.h:139-141 (num_mv: vectors applied)
    std::unique_ptr<Timer> t_multAx  = std::make_unique<Timer>();
    std::unique_ptr<Timer> t_multOPv = std::make_unique<Timer>();
    std::unique_ptr<Timer> t_multPc  = std::make_unique<Timer>();
    std::unique_ptr<Timer> t_factorOP = std::make_unique<Timer>();

    // mpos: host cx64 MPO tensors (DL, DR, d, d); envs: host cx64 blocks (chi, chi, D), as EnvBase stores them.
    MatVecMPOS(const std::vector<const std::complex<double> *> &mpos, const std::vector<std::array<long, 4>> &mpo_dims,
               const std::complex<double> *envL, std::array<long, 3> envL_dims, const std::complex<double> *envR,
               std::array<long, 3> envR_dims) {
        std::vector<const void *> p(mpos.begin(), mpos.end());
        std::vector<int64_t>      d;
        for(const auto &a : mpo_dims) d.insert(d.end(), a.begin(), a.end());
        const int64_t dl[3] = {envL_dims[0], envL_dims[1], envL_dims[2]};
        const int64_t dr[3] = {envR_dims[0], envR_dims[1], envR_dims[2]};
        check(dmrg_matvec_create(&op_, dtype_of<T>(), static_cast<int>(p.size()), p.data(), d.data(), envL, dl, envR, dr),
              "MatVecMPOS");
    }
    // MatVecMPOS(mpos, enva, envb) (matvec_mpos.h:116-120): the A operator as above plus the B operator of MultBx.
    MatVecMPOS(const std::vector<const std::complex<double> *> &mpos, const std::vector<std::array<long, 4>> &mpo_dims,
               const std::complex<double> *envL, std::array<long, 3> envL_dims, const std::complex<double> *envR,
               std::array<long, 3> envR_dims, const std::complex<double> *envL_B, std::array<long, 3> envL_B_dims,
               const std::complex<double> *envR_B, std::array<long, 3> envR_B_dims)
        : MatVecMPOS(mpos, mpo_dims, envL, envL_dims, envR, envR_dims) {
        std::vector<const void *> p(mpos.begin(), mpos.end());
        std::vector<int64_t>      d;
        for(const auto &a : mpo_dims) d.insert(d.end(), a.begin(), a.end());
        const int64_t dl[3] = {envL_B_dims[0], envL_B_dims[1], envL_B_dims[2]};
        const int64_t dr[3] = {envR_B_dims[0], envR_B_dims[1], envR_B_dims[2]};
        check(dmrg_matvec_set_b(op_, static_cast<int>(p.size()), p.data(), d.data(), envL_B, dl, envR_B, dr), "MatVecMPOS");
    }
    ~MatVecMPOS() {
        if(op_) dmrg_matvec_destroy(op_);
    }
    MatVecMPOS(const MatVecMPOS &)            = delete;
    MatVecMPOS &operator=(const MatVecMPOS &) = delete;

    // matvec_mpos.h:127-128: int, as Arpack++ and PRIMME take it.
    int rows() const { return static_cast<int>(dmrg_matvec_rows(op_)); }
    int cols() const { return rows(); }
    std::array<long, 3> get_shape_mps() const {
        int64_t s[3];
        check(dmrg_matvec_shape_mps(op_, s), "get_shape_mps");
        return {static_cast<long>(s[0]), static_cast<long>(s[1]), static_cast<long>(s[2])};
    }
    // MultAx(T*, T*) (matvec_mpos.cpp:963-974): one vector, host buffers of rows() elements. Also the member
    // function Arpack++ calls through &MatrixType::MultAx (solver_arpack.cpp:109,142,169) and by reverse
    // communication (:336-340).
    void MultAx(T *mps_in, T *mps_out) {
        const auto t0 = std::chrono::steady_clock::now();
        check(dmrg_matvec_multax(op_, mps_in, rows(), mps_out, rows(), 1), "MultAx");
        t_multAx->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        num_mv = static_cast<long>(dmrg_matvec_num_mv(op_));
    }

    // PRIMME matrixMatvec callback (matvec_mpos.cpp:976-984): blockSize columns at strides ldx / ldy.
    // Errors are reported through *err (non-zero makes PRIMME stop), never by an exception.
    void MultAx(void *x, int *ldx, void *y, int *ldy, int *blockSize, primme_params *, int *err) {
        const auto        t0 = std::chrono::steady_clock::now();
        const dmrg_status s  = dmrg_matvec_multax(op_, x, *ldx, y, *ldy, *blockSize);
        t_multAx->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        num_mv = static_cast<long>(dmrg_matvec_num_mv(op_));
        *err = (s == DMRG_OK) ? 0 : -1;
    }

    // Arpack++ glue (solver_arpack.cpp:90-91,108): the problem form and eigenvector side are recorded; the
    // operator is the same for every setting (matvec_mpos.cpp:1475-1481).
    template <typename Form>
    void set_mode(Form f) { form_ = f; }
    template <typename Side>
    void set_side(Side s) { side_ = s; }
    const EnumValue &get_form() const { return form_; }
    const EnumValue &get_side() const { return side_; }

    // MultBx (matvec_mpos.cpp:986-1019): the B operator; mps_out is left as it is when there is none. Like
    // MultAx it counts into num_mv and t_multAx.
    void MultBx(T *mps_in, T *mps_out) {
        const auto t0 = std::chrono::steady_clock::now();
        check(dmrg_matvec_multbx(op_, mps_in, rows(), mps_out, rows(), 1), "MultBx");
        t_multAx->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        num_mv = static_cast<long>(dmrg_matvec_num_mv(op_));
    }
    void MultBx(void *x, int *ldx, void *y, int *ldy, int *blockSize, primme_params *, int *err) {
        const auto        t0 = std::chrono::steady_clock::now();
        const dmrg_status s  = dmrg_matvec_multbx(op_, x, *ldx, y, *ldy, *blockSize);
        t_multAx->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        num_mv = static_cast<long>(dmrg_matvec_num_mv(op_));
        *err = (s == DMRG_OK) ? 0 : -1;
    }

    // Jacobi preconditioner, PRIMME applyPreconditioner callback (matvec_mpos.cpp:1333-1361 with
    // jcbMaxBlockSize = 1): y_i = x_i / (diag(H_eff) - primme->ShiftsForPreconditioner[i]). A template on the
    // parameter block so that its fields are only looked up where PRIMME's header is included.
    template <typename P>
    void MultPc(void *x, int *ldx, void *y, int *ldy, int *blockSize, P *primme, int *err) {
        MultPc(x, *ldx, y, *ldy, *blockSize, primme ? primme->ShiftsForPreconditioner : nullptr, err);
    }
    void MultPc(void *x, long ldx, void *y, long ldy, int blockSize, const double *shifts, int *err) {
        if(preconditioner.value == 0) { // eig::Preconditioner::NONE: y is left as PRIMME passed it
            *err = 0;
            return;
        }
        const auto t0 = std::chrono::steady_clock::now();
        if(factorization.value != applied_factorization_) {
            const dmrg_status sf = dmrg_matvec_set_factorization(op_, factorization.value);
            if(sf != DMRG_OK) {
                *err = -1;
                return;
            }
            applied_factorization_ = factorization.value;
        }
        const dmrg_status s = dmrg_matvec_precond(op_, x, ldx, y, ldy, blockSize, shifts);
        t_multPc->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        num_pc += blockSize;
        *err = (s == DMRG_OK) ? 0 : -1;
    }
    // set_jcbMaxBlockSize (matvec_mpos.cpp:1483-1505): 1 selects the Jacobi preconditioner, larger sizes the
    // block-Jacobi one. The size is clamped to [1, rows()] and then lowered in steps of 2 (down to half) to the
    // first size that divides rows() or leaves a last block of more than 80% of the others, keeping the one with
    // the largest remainder seen, so that the blocks are of nearly equal size. No value: unchanged.
    void set_jcbMaxBlockSize(std::optional<long> size) {
        if(!size.has_value()) return;
        const long n   = rows();
        long       jcb = std::clamp(size.value(), 1l, n);
        long       best_rem = n % jcb, best = jcb;
        for(long cand = jcb; cand >= jcb / 2 && cand >= 1; cand -= 2) {
            const long rem = n % cand;
            if(rem > best_rem || rem == 0) best_rem = rem, best = cand;
            if(rem == 0 || rem > 4 * cand / 5) break;
        }
        if(best != jcb) jcb = std::clamp(best, jcb / 2, n);
        check(dmrg_matvec_set_jcb_block_size(op_, jcb), "set_jcbMaxBlockSize");
        jcb_max_block_size_ = jcb;
    }
    long get_jcbMaxBlockSize() const { return jcb_max_block_size_; }

    // set_shift(sigma) (matvec_mpos.cpp:1433-1472).
    void set_shift(std::complex<double> shift) {
        check(dmrg_matvec_set_shift(op_, shift.real(), shift.imag()), "set_shift");
        ready_shift_ = true;
    }
    bool isReadyShift() const { return ready_shift_; }
    bool isReadyFactorOp() const { return dmrg_matvec_factor_ready(op_) != 0; }
    // FactorOP (matvec_mpos.cpp:887-913): once, with the factorization field, on the operator as set_shift left it.
    void FactorOP() {
        if(isReadyFactorOp()) return;
        const auto t0 = std::chrono::steady_clock::now();
        check(dmrg_matvec_factor_op(op_, factorization.value, nullptr), "FactorOP");
        t_factorOP->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    // MultOPv, PRIMME matrixMatvec form (matvec_mpos.cpp:936-975): y_i = (H_eff - shift)^-1 x_i; errors (no
    // factor, factorization NONE) set *err.
    void MultOPv(void *x, int *ldx, void *y, int *ldy, int *blockSize, primme_params *, int *err) {
        const auto        t0 = std::chrono::steady_clock::now();
        const dmrg_status s  = dmrg_matvec_mult_opv(op_, x, *ldx, y, *ldy, *blockSize);
        t_multOPv->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        num_op = static_cast<long>(dmrg_matvec_num_op(op_));
        *err   = (s == DMRG_OK) ? 0 : -1;
    }
    // One vector of rows() elements, the form ARPACK's shift-invert reverse communication applies.
    void MultOPv(T *mps_in, T *mps_out) {
        const auto t0 = std::chrono::steady_clock::now();
        check(dmrg_matvec_mult_opv(op_, mps_in, rows(), mps_out, rows(), 1), "MultOPv");
        t_multOPv->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        num_op = static_cast<long>(dmrg_matvec_num_op(op_));
    }

    dmrg_matvec *handle() { return op_; }

  private:
    dmrg_matvec *op_                    = nullptr;
    long         jcb_max_block_size_    = 1;
    bool         ready_shift_           = false;
    int          applied_factorization_ = 3; // the library's default, eig::Factorization::LU
    EnumValue    form_;                      // eig::Form::SYMM
    EnumValue    side_{1};                   // eig::Side::R, the reference's default (matvec_mpos.h:48)
};

namespace contraction {

template <typename Scalar>
struct DeviceCopy {
    void  *p     = nullptr;
    size_t bytes = 0;
    DeviceCopy(const Scalar *host, size_t n) : bytes(n * sizeof(Scalar)) {
        check(dmrg_malloc(&p, bytes), "dmrg_malloc");
        if(host) check(dmrg_memcpy_htod(p, host, bytes), "dmrg_memcpy_htod");
    }
    ~DeviceCopy() { dmrg_free(p); }
};

// Host-pointer contraction with the reference signature: copies in, contracts on the GPU, copies out.
template <typename Scalar>
void matrix_vector_product(Scalar *res, const Scalar *mps, std::array<long, 3> mps_dims, const Scalar *mpo,
                           std::array<long, 4> mpo_dims, const Scalar *envL, std::array<long, 3> envL_dims,
                           const Scalar *envR, std::array<long, 3> envR_dims) {
    const int64_t ms[3] = {mps_dims[0], mps_dims[1], mps_dims[2]};
    const int64_t wo[4] = {mpo_dims[0], mpo_dims[1], mpo_dims[2], mpo_dims[3]};
    const int64_t el[3] = {envL_dims[0], envL_dims[1], envL_dims[2]};
    const int64_t er[3] = {envR_dims[0], envR_dims[1], envR_dims[2]};
    const size_t  n     = static_cast<size_t>(ms[0] * ms[1] * ms[2]);
    DeviceCopy<Scalar> dm(mps, n), dw(mpo, static_cast<size_t>(wo[0] * wo[1] * wo[2] * wo[3])),
        dl(envL, static_cast<size_t>(el[0] * el[1] * el[2])), dr(envR, static_cast<size_t>(er[0] * er[1] * er[2])), dres(nullptr, n);
    check(dmrg_matrix_vector_product(dtype_of<Scalar>(), dres.p, dm.p, ms, dw.p, wo, dl.p, el, dr.p, er, nullptr),
          "matrix_vector_product");
    check(dmrg_memcpy_dtoh(res, dres.p, n * sizeof(Scalar)), "dmrg_memcpy_dtoh");
}

template <typename Scalar>
Scalar expectation_value(const Scalar *mps, std::array<long, 3> mps_dims, const Scalar *mpo, std::array<long, 4> mpo_dims,
                         const Scalar *envL, std::array<long, 3> envL_dims, const Scalar *envR, std::array<long, 3> envR_dims) {
    const int64_t ms[3] = {mps_dims[0], mps_dims[1], mps_dims[2]};
    const int64_t wo[4] = {mpo_dims[0], mpo_dims[1], mpo_dims[2], mpo_dims[3]};
    const int64_t el[3] = {envL_dims[0], envL_dims[1], envL_dims[2]};
    const int64_t er[3] = {envR_dims[0], envR_dims[1], envR_dims[2]};
    DeviceCopy<Scalar> dm(mps, static_cast<size_t>(ms[0] * ms[1] * ms[2])), dw(mpo, static_cast<size_t>(wo[0] * wo[1] * wo[2] * wo[3])),
        dl(envL, static_cast<size_t>(el[0] * el[1] * el[2])), dr(envR, static_cast<size_t>(er[0] * er[1] * er[2]));
    double r[2];
    check(dmrg_expectation_value(dtype_of<Scalar>(), dm.p, ms, dw.p, wo, dl.p, el, dr.p, er, r, nullptr), "expectation_value");
    if constexpr(std::is_same_v<Scalar, double>) return r[0];
    else return Scalar(r[0], r[1]);
}

} // namespace contraction
} // namespace dmrg_amd

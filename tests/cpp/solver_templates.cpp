// Compiles dmrg_amd::MatVecMPOS<T> inside stand-ins of the reference's two solver templates and runs them:
//   eigs_primme_like<M>  - the member expressions of eig::solver::eigs_primme (solver_primme.cpp:297-517) on the
//                          operator, with the preconditioner callback of preconditioner_jacobi
//                          (optimize_energy.cpp:34-42) and PRIMME's matrixMatvec wrapper (:122-128);
//   solver_arpack_like<M> - those of eig::solver_arpack (solver_arpack.cpp:77-244, 320-380): set_mode / set_side,
//                          get_form, a solver object built from &MatrixType::MultAx, reverse communication through
//                          MultAx(GetVector(), PutVector()), and the num_mv / num_op fields.
// The eig:: settings, PRIMME's parameter block and the Arpack++ classes are reduced here to the members those
// expressions touch. The eigensolver behind each stand-in is a host Lanczos with full reorthogonalisation (the
// small projected problems go to the library's dmrg_host_heev), enough to check that each path applies the
// operator the reference's would.
//
// input : int64 header {cplx, d, chiL, chiR, DL, DR, jcb_request}, then cx64 mpo, envL, envR, then x (n, T)
// output: doubles {E_primme, E_arpack_member, E_arpack_rc}, int64 {matvecs counted by the PRIMME stand-in,
//         num_mv of its operator, num_mv of the ARPACK operator, jcb_applied, Lanczos steps of the PRIMME
//         stand-in, of both ARPACK stand-ins}, then y_pc (n, T): MultPc of x through the PRIMME callback with
//         ShiftsForPreconditioner = {0.5} after factorization = LLT
#include "dmrg_amd/matvec_mpos.hpp"

#include <cmath>
#include <cstdio>
#include <functional>
#include <type_traits>
#include <optional>
#include <vector>

using cx = std::complex<double>;

// ------------------------------------------------------------------ stand-ins of the reference's types
namespace eig {
enum class Form { SYMM, NSYM };
enum class Side { L, R, LR };
enum class Shinv { ON, OFF };
enum class Preconditioner { NONE, JACOBI };
enum class Factorization { NONE, LLT, LDLT, LU, QR, ILUT, ILDLT };
} // namespace eig

struct primme_params {
    long   n = 0;
    int    numEvals = 1;
    void  *matrix = nullptr;
    void (*matrixMatvec)(void *, int *, void *, int *, int *, primme_params *, int *) = nullptr;
    void  *preconditioner = nullptr;
    void (*applyPreconditioner)(void *, int *, void *, int *, int *, primme_params *, int *) = nullptr;
    double *ShiftsForPreconditioner = nullptr;
    struct {
        int precondition = 0;
    } correctionParams;
    struct {
        long numMatvecs = 0;
    } stats;
};

namespace eig {
using primme_callback = void (*)(void *, int *, void *, int *, int *, primme_params *, int *);
struct settings {
    std::optional<std::complex<double>> sigma;
    Shinv                               shift_invert = Shinv::OFF;
    std::optional<long>                 jcbMaxBlockSize;
    std::optional<primme_callback>      primme_preconditioner;
    std::optional<Form>                 form = Form::SYMM;
    std::optional<Side>                 side = Side::R;
    std::optional<long>                 maxNev = 1;
};
} // namespace eig

template <typename T>
static T from_cx(cx c) {
    if constexpr(std::is_same_v<T, double>) return c.real();
    else return c;
}

// Host Lanczos with full reorthogonalisation for the lowest eigenvalue of a Hermitian operator given as a
// y = A x callback; the projected tridiagonal problem is solved with dmrg_host_heev.
template <typename T>
static double lanczos_lowest(long n, int m, const std::function<void(T *, T *)> &apply, long *steps) {
    m = static_cast<int>(std::min<long>(m, n));
    std::vector<std::vector<T>> V;
    std::vector<T>              v(n), w(n);
    for(long i = 0; i < n; ++i) v[i] = T(std::cos(0.37 * i + 0.1) + 0.5);
    auto dot = [&](const std::vector<T> &a, const std::vector<T> &b) {
        cx s = 0;
        for(long i = 0; i < n; ++i) s += std::conj(cx(a[i])) * cx(b[i]);
        return s;
    };
    auto scale = [&](std::vector<T> &a, double f) {
        for(auto &x : a) x *= f;
    };
    scale(v, 1.0 / std::sqrt(dot(v, v).real()));
    std::vector<double> alpha, beta;
    int                 k = 0;
    for(; k < m; ++k) {
        V.push_back(v);
        apply(V.back().data(), w.data());
        alpha.push_back(dot(V.back(), w).real());
        for(int pass = 0; pass < 2; ++pass) // full reorthogonalisation, twice
            for(const auto &u : V) {
                const cx c = dot(u, w);
                for(long i = 0; i < n; ++i) w[i] -= from_cx<T>(c * cx(u[i]));
            }
        const double b = std::sqrt(dot(w, w).real());
        if(b < 1e-14 * std::abs(alpha.back()) || k + 1 == m) {
            ++k;
            break;
        }
        beta.push_back(b);
        v = w;
        scale(v, 1.0 / b);
    }
    *steps = k;
    std::vector<cx> Tm(static_cast<size_t>(k) * k, cx(0)), Z(static_cast<size_t>(k) * k);
    for(int i = 0; i < k; ++i) Tm[i + static_cast<size_t>(k) * i] = alpha[i];
    for(int i = 0; i + 1 < k; ++i) Tm[i + 1 + static_cast<size_t>(k) * i] = Tm[i + static_cast<size_t>(k) * (i + 1)] = beta[i];
    std::vector<double> ev(k);
    dmrg_amd::check(dmrg_host_heev(k, reinterpret_cast<const double *>(Tm.data()), ev.data(), reinterpret_cast<double *>(Z.data())),
                    "dmrg_host_heev");
    return ev[0];
}

// ------------------------------------------------------------------ eig::solver::eigs_primme, reduced
template <typename MatrixProductType>
static void MultAx_wrapper(void *x, int *ldx, void *y, int *ldy, int *blockSize, primme_params *primme, int *ierr) {
    auto matrix_ptr = static_cast<MatrixProductType *>(primme->matrix); // solver_primme.cpp:122-128
    matrix_ptr->MultAx(x, ldx, y, ldy, blockSize, primme, ierr);
}
template <typename MatrixProductType>
static void MultOPv_wrapper(void *x, int *ldx, void *y, int *ldy, int *blockSize, primme_params *primme, int *ierr) {
    static_cast<MatrixProductType *>(primme->matrix)->MultOPv(x, ldx, y, ldy, blockSize, primme, ierr);
}
// preconditioner_jacobi (optimize_energy.cpp:34-42)
template <typename MatrixProductType>
static void preconditioner_jacobi(void *x, int *ldx, void *y, int *ldy, int *blockSize, primme_params *primme, int *ierr) {
    if(x == nullptr || y == nullptr || primme == nullptr) return;
    const auto H_ptr      = static_cast<MatrixProductType *>(primme->matrix);
    H_ptr->preconditioner = eig::Preconditioner::JACOBI;
    H_ptr->MultPc(x, ldx, y, ldy, blockSize, primme, ierr);
}

struct primme_result {
    double energy = 0;
    long   num_mv = 0, num_op = 0, steps = 0;
    double time_op = 0;
};

template <typename MatrixProductType>
static primme_result eigs_primme_like(MatrixProductType &matrix, eig::settings &config) {
    using Scalar = typename MatrixProductType::Scalar;
    if constexpr(MatrixProductType::can_shift) {
        if(config.sigma) {
            matrix.set_shift(config.sigma.value());
            if constexpr(MatrixProductType::can_shift_invert) {
                if(config.shift_invert == eig::Shinv::ON) matrix.FactorOP();
            } else if(config.shift_invert == eig::Shinv::ON)
                throw std::runtime_error("Tried to shift-invert an incompatible matrix");
        }
    } else if(config.sigma) {
        throw std::runtime_error("Tried to apply shift on an incompatible matrix");
    }
    primme_params primme;
    primme.matrix       = &matrix;
    primme.matrixMatvec = MultAx_wrapper<MatrixProductType>;
    if(matrix.isReadyFactorOp()) { primme.matrixMatvec = MultOPv_wrapper<MatrixProductType>; }
    if constexpr(MatrixProductType::can_precondition) {
        if(config.primme_preconditioner) {
            primme.applyPreconditioner           = config.primme_preconditioner.value();
            primme.preconditioner                = &config;
            primme.correctionParams.precondition = true;
            matrix.preconditioner                = eig::Preconditioner::JACOBI;
            matrix.set_jcbMaxBlockSize(config.jcbMaxBlockSize);
        }
    }
    primme.n        = matrix.rows();
    primme.numEvals = static_cast<int>(config.maxNev.value_or(1));
    if(config.sigma and not matrix.isReadyShift()) throw std::runtime_error("shift not applied on the operator");
    primme_result r;
    // the stand-in of dprimme / zprimme: Lanczos through primme.matrixMatvec
    r.energy = lanczos_lowest<Scalar>(primme.n, 400, [&](Scalar *x, Scalar *y) {
        int ld = static_cast<int>(primme.n), bs = 1, err = 1;
        primme.matrixMatvec(x, &ld, y, &ld, &bs, &primme, &err);
        if(err != 0) throw std::runtime_error("matrixMatvec reported an error");
        primme.stats.numMatvecs += bs;
    }, &r.steps);
    r.num_mv  = primme.stats.numMatvecs;
    r.num_op  = matrix.num_op;                    // solver_primme.cpp:497
    r.time_op = matrix.t_multOPv->get_time();     // :501
    return r;
}

// ------------------------------------------------------------------ Arpack++ and eig::solver_arpack, reduced
// Solver classes with the constructor shapes of ARSymStdEig / ARCompStdEig (operator + member pointer) and of the
// reverse-communication ARrc*StdEig (TakeStep / GetIdo / GetVector / PutVector).
template <typename ARTYPE, typename ARFOP>
class ARStdEigStub {
  public:
    ARStdEigStub(int n, int nev, ARFOP *op, void (ARFOP::*MultOPx)(ARTYPE[], ARTYPE[])) : n_(n), nev_(nev), op_(op), mult_(MultOPx) {}
    double FindLowest(long *steps) {
        return lanczos_lowest<ARTYPE>(n_, 400, [&](ARTYPE *x, ARTYPE *y) { (op_->*mult_)(x, y); }, steps);
    }

  private:
    int    n_, nev_;
    ARFOP *op_;
    void (ARFOP::*mult_)(ARTYPE[], ARTYPE[]);
};

template <typename ARTYPE>
class ARrcStdEigStub { // reverse communication: the caller applies the operator between steps
  public:
    explicit ARrcStdEigStub(long n) : n_(n), x_(n), y_(n) {}
    int     GetIdo() const { return ido_; }
    ARTYPE *GetVector() { return x_.data(); }
    ARTYPE *PutVector() { return y_.data(); }
    // One "step": first call hands out the start vector; each later call consumes A x and hands out the next x.
    void TakeStep() {
        if(ido_ == 0) {
            for(long i = 0; i < n_; ++i) x_[i] = ARTYPE(std::cos(0.37 * i + 0.1) + 0.5);
            normalise(x_);
            ido_ = 1;
            return;
        }
        V_.push_back(x_);
        W_.push_back(y_);
        std::vector<ARTYPE> r = y_;
        for(int pass = 0; pass < 2; ++pass)
            for(const auto &u : V_) {
                const cx c = dot(u, r);
                for(long i = 0; i < n_; ++i) r[i] -= from_cx<ARTYPE>(c * cx(u[i]));
            }
        const double b = std::sqrt(dot(r, r).real());
        if(V_.size() >= 400 || static_cast<long>(V_.size()) >= n_ || b < 1e-14) {
            ido_ = 99;
            return;
        }
        x_ = r;
        normalise(x_);
    }
    // Rayleigh-Ritz on the collected basis (V, AV).
    double Lowest() const {
        const int       k = static_cast<int>(V_.size());
        std::vector<cx> H(static_cast<size_t>(k) * k), Z(static_cast<size_t>(k) * k);
        for(int j = 0; j < k; ++j)
            for(int i = 0; i < k; ++i) H[i + static_cast<size_t>(k) * j] = dot(V_[i], W_[j]);
        std::vector<double> ev(k);
        dmrg_amd::check(dmrg_host_heev(k, reinterpret_cast<const double *>(H.data()), ev.data(), reinterpret_cast<double *>(Z.data())),
                        "dmrg_host_heev");
        return ev[0];
    }
    long Steps() const { return static_cast<long>(V_.size()); }

  private:
    cx dot(const std::vector<ARTYPE> &a, const std::vector<ARTYPE> &b) const {
        cx s = 0;
        for(long i = 0; i < n_; ++i) s += std::conj(cx(a[i])) * cx(b[i]);
        return s;
    }
    void normalise(std::vector<ARTYPE> &a) const {
        const double f = 1.0 / std::sqrt(dot(a, a).real());
        for(auto &x : a) x *= f;
    }
    long                             n_;
    int                              ido_ = 0;
    std::vector<ARTYPE>              x_, y_;
    std::vector<std::vector<ARTYPE>> V_, W_;
};

template <typename MatrixType>
struct solver_arpack_like {
    using Scalar = typename MatrixType::Scalar;
    MatrixType    &matrix;
    eig::settings &config;
    double         energy_member = 0, energy_rc = 0;
    long           num_mv = 0, num_op = 0, steps = 0;

    void eigs() {
        matrix.set_mode(config.form.value()); // solver_arpack.cpp:90-91
        matrix.set_side(config.side.value());
        if(matrix.get_form() != eig::Form::SYMM) throw std::runtime_error("ERROR: matrix not SYMMETRIC"); // :108
        const int nev = std::clamp<int>(static_cast<int>(config.maxNev.value()), 1, matrix.rows() / 2);    // :80
        ARStdEigStub<Scalar, MatrixType> solver(matrix.rows(), nev, &matrix, &MatrixType::MultAx);          // :109
        long s1 = 0;
        energy_member = solver.FindLowest(&s1);
        ARrcStdEigStub<Scalar> rc(matrix.rows());
        while(rc.GetIdo() != 99) { // :330-345
            rc.TakeStep();
            if(std::abs(rc.GetIdo()) == 1) {
                if constexpr(MatrixType::can_shift_invert) {
                    if(matrix.isReadyFactorOp()) matrix.MultOPv(rc.GetVector(), rc.PutVector());
                    else matrix.MultAx(rc.GetVector(), rc.PutVector());
                } else
                    matrix.MultAx(rc.GetVector(), rc.PutVector());
            }
        }
        energy_rc = rc.Lowest();
        num_mv    = matrix.num_mv; // :243
        num_op    = matrix.num_op;
        steps     = s1 + rc.Steps();
    }
};

// ------------------------------------------------------------------ driver
template <typename T>
static int run(FILE *in, FILE *out, const int64_t h[7]) {
    const int64_t d = h[1], cL = h[2], cR = h[3], DL = h[4], DR = h[5], jcb_req = h[6];
    const int64_t n = d * cL * cR;
    std::vector<cx> mpo(DL * DR * d * d), eL(cL * cL * DL), eR(cR * cR * DR);
    std::vector<T>  x(n), ypc(n, T(0));
    double          sig[2]; // shift-invert shifts: inside the spectrum (LU) and below it (LLT)
    if(fread(mpo.data(), sizeof(cx), mpo.size(), in) != mpo.size() || fread(eL.data(), sizeof(cx), eL.size(), in) != eL.size() ||
       fread(eR.data(), sizeof(cx), eR.size(), in) != eR.size() || fread(x.data(), sizeof(T), x.size(), in) != x.size() ||
       fread(sig, sizeof(double), 2, in) != 2)
        return 2;
    using M = dmrg_amd::MatVecMPOS<T>;
    eig::settings cfg;
    cfg.primme_preconditioner = preconditioner_jacobi<M>;
    cfg.jcbMaxBlockSize       = jcb_req;
    M             op({mpo.data()}, {{DL, DR, d, d}}, eL.data(), {cL, cL, DL}, eR.data(), {cR, cR, DR});
    const auto    rp        = eigs_primme_like(op, cfg);
    const long    op_num_mv = op.num_mv;
    // The preconditioner callback as PRIMME calls it, with the block factorization of the folded-spectrum solver.
    op.factorization = eig::Factorization::LLT;
    double        shift[1] = {0.5};
    primme_params pp;
    pp.matrix                  = &op;
    pp.ShiftsForPreconditioner = shift;
    int ld = static_cast<int>(n), bs = 1, err = 1;
    cfg.primme_preconditioner.value()(x.data(), &ld, ypc.data(), &ld, &bs, &pp, &err);
    if(err != 0) return 3;
    if(op.preconditioner != eig::Preconditioner::JACOBI) return 4;

    M                     op2({mpo.data()}, {{DL, DR, d, d}}, eL.data(), {cL, cL, DL}, eR.data(), {cR, cR, DR});
    solver_arpack_like<M> arp{op2, cfg};
    arp.eigs();
    if(op2.get_side() != eig::Side::R) return 5;
    const double  e[3] = {rp.energy, arp.energy_member, arp.energy_rc};
    const int64_t k[6] = {rp.num_mv, op_num_mv, arp.num_mv, op.get_jcbMaxBlockSize(), rp.steps, arp.steps};
    fwrite(e, sizeof(double), 3, out);
    fwrite(k, sizeof(int64_t), 6, out);
    fwrite(ypc.data(), sizeof(T), ypc.size(), out);

    // Shift-invert, as eigs_primme sets it up (shift, FactorOP, matrixMatvec = MultOPv): the stand-in's Lanczos
    // finds the lowest eigenvalue of (H - sigma)^-1 with an LU factor of H - sigma for a shift inside the spectrum.
    M             op3({mpo.data()}, {{DL, DR, d, d}}, eL.data(), {cL, cL, DL}, eR.data(), {cR, cR, DR});
    eig::settings cfg3;
    cfg3.sigma        = cx(sig[0], 0.0);
    cfg3.shift_invert = eig::Shinv::ON;
    op3.factorization = eig::Factorization::LU;
    if(op3.isReadyFactorOp()) return 6;
    const auto rsi = eigs_primme_like(op3, cfg3);
    if(!op3.isReadyFactorOp() || rsi.num_op != rsi.num_mv || op3.t_factorOP->get_time() <= 0.0) return 7;
    // ARPACK's shift-invert reverse communication with MultOPv(T*, T*), an LLT factor of H - sigma below the
    // spectrum; the member-pointer solver still applies MultAx, i.e. H - sigma.
    M op4({mpo.data()}, {{DL, DR, d, d}}, eL.data(), {cL, cL, DL}, eR.data(), {cR, cR, DR});
    op4.set_shift(cx(sig[1], 0.0));
    op4.factorization = eig::Factorization::LLT;
    op4.FactorOP();
    solver_arpack_like<M> arp4{op4, cfg};
    arp4.eigs();
    const double  esi[3] = {rsi.energy, arp4.energy_member, arp4.energy_rc};
    const int64_t ksi[2] = {rsi.num_op, arp4.num_op};
    fwrite(esi, sizeof(double), 3, out);
    fwrite(ksi, sizeof(int64_t), 2, out);
    return 0;
}

int main(int argc, char **argv) {
    if(argc != 3) return 1;
    FILE *in = fopen(argv[1], "rb");
    if(!in) return 1;
    int64_t h[7];
    if(fread(h, sizeof(int64_t), 7, in) != 7) return 2;
    FILE *out = fopen(argv[2], "wb");
    if(!out) return 1;
    int rc = 0;
    try {
        rc = h[0] ? run<cx>(in, out, h) : run<double>(in, out, h);
    } catch(const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        rc = 20;
    }
    fclose(in);
    fclose(out);
    return rc;
}

// GPU-resident eigensolver for H_eff: generalized Davidson with thick restart ("GD+k": the previous
// Ritz vector is kept at restart), block size 1, the role PRIMME's GD_Olsen_plusK plays in the
// reference (solver_primme.cpp:297-517). Basis, residuals and projections live in HBM; only the
// m x m Rayleigh-Ritz problem is solved on the host.
#pragma once
#include "matvec_op.hpp"
#include "vecops.hpp"

#include <complex>
#include <memory>
#include <vector>

namespace dmrg {

enum EigsTarget { TARGET_SA = 0, TARGET_LA = 1, TARGET_CLOSEST = 2 };

struct EigsParams {
    int     nev         = 1;
    int     ncv         = 0;  // max basis size; <= 0: nev * ceil(log2 n) clamped to [2 nev, n] (solver.cpp:12-15)
    int     min_restart = 0;  // <= 0: max(1, ncv/2 - 1) (solver_primme.cpp:374)
    int     target      = TARGET_SA;
    double  shift       = 0.0; // sigma for TARGET_CLOSEST
    double  tol         = 1e-12; // converged when |r| <= tol * |A|_est
    int64_t max_matvecs = 100000;
    int     precond     = 0; // 1: Jacobi (diagonal of H_eff), 2: block-Jacobi with blocks of jcb
    int64_t jcb         = 128;
    int     plus_k      = 1; // previous Ritz vectors retained at restart
    double  max_time_s  = 0; // 0: no limit
    uint64_t seed       = 12345;
    // Stagnation stop (used by the FP32 phase of the mixed solver): stop when the residual has not
    // dropped by 10% over this many iterations (0: off).
    int stagnation_window = 0;
    // FP64 solves only: expand with t from at most inner_max FP32 CG steps on the correction equation
    // (Jacobi-Davidson), stopped early at a residual reduction of inner_tol (0: plain Davidson expansion).
    int    inner_max = 0;
    double inner_tol = 1e-2;
    // 1 (with TARGET_CLOSEST): refined projection (primme_proj_refined, the reference's choice for folded-spectrum
    // solves, optimize_folded_spectrum.cpp:295-299): the pair is x = V y with y the right singular vector of the
    // smallest singular value of the triangular factor of W - shift V (kept as a QR factorisation column by
    // column), and its value the Rayleigh quotient y^H H y. 0: Rayleigh-Ritz.
    int projection = 0;
    // 1: the convergence test of folded_spectrum::convTestFun (optimize_folded_spectrum.cpp:41-93): converged when
    // |r| < tol |A| and the last up to 500 Ritz values or residual norms have stopped decreasing (their change
    // over at least 250 iterations is above -tol), or when |r| < tol. 0: |r| <= tol |A|.
    int conv_test = 0;
};

struct EigsStats {
    int64_t matvecs  = 0;
    int     iters    = 0;
    int     restarts = 0;
    int     nconv    = 0;
    double  anorm    = 0.0;
    double  seconds  = 0.0;
    bool    stagnated = false;
    int64_t inner_matvecs = 0; // FP32 applications inside the correction solves
};

template <typename T>
struct lower_precision {
    using type = T;
};
template <>
struct lower_precision<double> {
    using type = float;
};
template <>
struct lower_precision<cpx<double>> {
    using type = cpx<float>;
};

// Approximate solution of the Jacobi-Davidson correction equation for a Ritz pair (theta, x) with residual r,
//   sigma (I - x x^H)(H - theta)(I - x x^H) t = -sigma (I - x x^H) r,   t orthogonal to x,
// by conjugate gradients in FP32 with matrix-core matvecs. sigma = +1 refines the lowest eigenvalue (the
// projected operator is positive definite once theta lies below the second eigenvalue), -1 the largest.
// The iteration stops after max_it steps, at a residual reduction of rtol (checked every few steps) or at
// a breakdown (p^H A p <= 0, the operator is not definite); t = r when no step was taken.
template <typename T32>
class CorrectionSolver {
  public:
    CorrectionSolver(MatVecOp &op, hipStream_t s);
    ~CorrectionSolver();
    CorrectionSolver(const CorrectionSolver &) = delete;
    CorrectionSolver &operator=(const CorrectionSolver &) = delete;
    // x, r, t: device vectors of the FP64 type (length n). Returns the number of FP32 matvecs.
    template <typename T64>
    int solve(const T64 *x, const T64 *r, double theta, double sigma, int max_it, double rtol, T64 *t);
    void rebind(MatVecOp &o) {
        op = &o;
        n  = o.D.n();
    }

  private:
    MatVecOp   *op;
    hipStream_t s;
    int64_t     n;
    DevBuf      xp, q, rv, tv, scratch, dev;
    double     *host = nullptr; // pinned copy of the state
};

template <typename T>
class Davidson {
  public:
    Davidson(MatVecOp &op, hipStream_t s);
    ~Davidson();
    Davidson(const Davidson &) = delete;
    Davidson &operator=(const Davidson &) = delete;
    // init: k0 start vectors (device, type T, columns of length n) or nullptr for a seeded random start.
    // out:  nev Ritz vectors (device, n x nev); evals/rnorms: host arrays of length nev.
    // nkeep > nev: also export the nkeep best Ritz vectors into out (n x nkeep).
    EigsStats solve(const EigsParams &p, const T *init, int k0, T *out, double *evals, double *rnorms, int nkeep = 0);
    // Use another operator (possibly of another size); device buffers are kept and grown as needed.
    void rebind(MatVecOp &o) {
        op = &o;
        n  = o.D.n();
        if(inner) inner->rebind(o);
    }

  private:
    MatVecOp   *op;
    hipStream_t s;
    int64_t     n;
    DevBuf      V, W, tmp, rbuf, scratch, small, Ydev, diag, pcw, Qr; // Qr: orthonormal factor of W - shift V (refined)
    std::unique_ptr<CorrectionSolver<typename lower_precision<T>::type>> inner;
    double     *host = nullptr; // pinned
    void        matvec(const T *x, T *y, int nvec);
    void        orthonormalize_into(T *t, int m, double *dnrm_pre, double *dnrm_post);
    void        expand_from(const T *src, T *t, int m, double *dnrm_pre, double *dnrm_post);
    void        project_column(int m, std::vector<std::complex<double>> &H, int ldh, int col);
    void        transform_basis(T *B, int m, const std::vector<std::complex<double>> &Q, int k);
    void        qr_append(int m, double shift); // Qr[:, m] from W[:, m] - shift V[:, m]; coefficients to SB_QC*
};

// Mixed precision: FP32 matrix-core Davidson to a loose tolerance, then FP64 Davidson seeded with the
// best FP32 Ritz vectors, to the requested tolerance (FP64 Rayleigh-Ritz, orthogonalisation, residuals).
// With inner_max > 0 (one wanted pair at the lowest or largest eigenvalue, no preconditioner) the FP64 phase
// expands with FP32 correction-equation solves, so that most matvecs stay on the FP32 matrix cores.
template <typename T64, typename T32>
EigsStats eigs_mixed(MatVecOp &op, hipStream_t s, const EigsParams &p, const T64 *init, T64 *out, double *evals,
                     double *rnorms, EigsStats *phase1);

} // namespace dmrg

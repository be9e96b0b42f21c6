// GPU-resident eigensolver for H_eff: generalized Davidson with thick restart ("GD+k": the previous
// Ritz vector is kept at restart), block size 1, the role PRIMME's GD_Olsen_plusK plays in the
// reference (solver_primme.cpp:297-517). Basis, residuals and projections live in HBM; only the
// m x m Rayleigh-Ritz problem is solved on the host.
#pragma once
#include "matvec_op.hpp"
#include "vecops.hpp"

#include <complex>
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
};

struct EigsStats {
    int64_t matvecs  = 0;
    int     iters    = 0;
    int     restarts = 0;
    int     nconv    = 0;
    double  anorm    = 0.0;
    double  seconds  = 0.0;
    bool    stagnated = false;
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
    }

  private:
    MatVecOp   *op;
    hipStream_t s;
    int64_t     n;
    DevBuf      V, W, tmp, rbuf, scratch, small, Ydev, diag, pcw;
    double     *host = nullptr; // pinned
    void        matvec(const T *x, T *y, int nvec);
    void        orthonormalize_into(T *t, int m, double *dnrm_pre, double *dnrm_post);
    void        expand_from(const T *src, T *t, int m, double *dnrm_pre, double *dnrm_post);
    void        project_column(int m, std::vector<std::complex<double>> &H, int ldh, int col);
    void        transform_basis(T *B, int m, const std::vector<std::complex<double>> &Q, int k);
};

// Mixed precision: FP32 matrix-core Davidson to a loose tolerance, then FP64 Davidson seeded with the
// best FP32 Ritz vectors, to the requested tolerance (FP64 Rayleigh-Ritz, orthogonalisation, residuals).
template <typename T64, typename T32>
EigsStats eigs_mixed(MatVecOp &op, hipStream_t s, const EigsParams &p, const T64 *init, T64 *out, double *evals,
                     double *rnorms, EigsStats *phase1);

} // namespace dmrg

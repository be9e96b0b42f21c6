// Block-vector kernels of the eigensolver (HBM-bound): projections against a basis, basis updates,
// Ritz vectors and residuals, preconditioning. All reductions accumulate in double and are deterministic.
#pragma once
#include "common.hpp"

#include <complex>
#include <type_traits>

namespace dmrg {

template <typename T>
struct vtraits;
template <>
struct vtraits<double> {
    using R                        = double;
    static constexpr bool       cx = false;
    static constexpr dmrg_dtype dt = DMRG_FP64;
};
template <>
struct vtraits<float> {
    using R                        = float;
    static constexpr bool       cx = false;
    static constexpr dmrg_dtype dt = DMRG_FP32;
};
template <>
struct vtraits<cpx<double>> {
    using R                        = double;
    static constexpr bool       cx = true;
    static constexpr dmrg_dtype dt = DMRG_CX64;
};
template <>
struct vtraits<cpx<float>> {
    using R                        = float;
    static constexpr bool       cx = true;
    static constexpr dmrg_dtype dt = DMRG_CX32;
};

// Scratch for reductions: partial sums of up to VEC_MAX_BLOCKS blocks x VEC_MAX_M columns (complex), then a
// workgroup counter (last 8 doubles) for reductions whose last workgroup folds the partial sums itself. Call
// vec_scratch_init after (re)allocating a scratch buffer: the counter must start at zero (it resets itself).
constexpr int VEC_MAX_BLOCKS = 2048; // enough waves in flight to stream the basis at HBM rate
constexpr int VEC_MAX_M      = 256;
constexpr int64_t vec_scratch_doubles() { return 2LL * VEC_MAX_BLOCKS * VEC_MAX_M + 4 * VEC_MAX_M + 8; }
void vec_scratch_init(double *scratch, hipStream_t s);

// out[j] = sum_i conj(V[i, j]) x[i], j < m (complex double, out on device, 2*m doubles).
template <typename T>
void multidot(int64_t n, int m, const T *V, int64_t ldv, const T *x, double *d_out, double *scratch, hipStream_t s);
// out[j] = sum_i conj(V[i, j]) x[i] for j < m and, when d_self is given, d_self[0] = |x|^2 from the same pass.
// skip (device int, optional): when *skip == 0 the kernels do nothing (outputs keep their values).
template <typename T>
void multidot_x(int64_t n, int m, const T *V, int64_t ldv, const T *x, double *d_out, double *d_self, const int *skip,
                double *scratch, hipStream_t s);
// y = beta x + alpha V h (x may alias y) and d_nrm2[0] = |y|^2; skip as above.
template <typename T>
void gemv_update_norm(int64_t n, int m, const T *V, int64_t ldv, const double *d_h, double alpha, double beta, const T *x, T *y,
                      double *d_nrm2, const int *skip, double *scratch, hipStream_t s);
// The projection of a second orthogonalisation pass with its DGKS test folded in: when d_post[0] < eta2 *
// d_pre[0], out[j] = sum_i conj(V[i, j]) x[i] (j < m) as in multidot_x; otherwise out is left as it is. Either way
// d_flag[0] receives the decision, for the kernels of the pass that follow (their skip argument).
template <typename T>
void multidot_dgks(int64_t n, int m, const T *V, int64_t ldv, const T *x, double *d_out, const double *d_pre, const double *d_post,
                   double eta2, int *d_flag, double *scratch, hipStream_t s);
// x = beta x + alpha V h   (h: complex double, m entries, device).
template <typename T>
void gemv_update(int64_t n, int m, const T *V, int64_t ldv, const double *d_h, double alpha, double beta, T *x, hipStream_t s);
// x = V y, r = W y - theta x, d_out[0] = |r|^2, d_out[1] = |x|^2 (y: complex double, device).
template <typename T>
void ritz_residual(int64_t n, int m, const T *V, const T *W, int64_t ld, const double *d_y, double theta, T *x, T *r,
                   double *d_out, double *scratch, hipStream_t s);
// x *= 1 / sqrt(d_nrm2[0]) (the squared norm read on the device).
template <typename T>
void scale_by_inv_norm(int64_t n, T *x, const double *d_nrm2, hipStream_t s);
template <typename T>
void scale(int64_t n, T *x, double alpha, hipStream_t s);
// t = r / (diag - theta), with |diag - theta| floored at floor (Jacobi preconditioner, MultPc jcb = 1).
template <typename T, typename D>
void jacobi_precond(int64_t n, const T *r, const D *diag, double theta, double floor, T *t, hipStream_t s);
// Seeded pseudo-random vector with entries uniform in [-1, 1] (both parts for complex).
template <typename T>
void random_fill(int64_t n, T *x, uint64_t seed, hipStream_t s);
template <typename TO, typename TI>
void convert(int64_t n, const TI *in, TO *out, hipStream_t s);

// Conjugate gradients on A t = b with A = sigma P (H - theta) P, P = I - x x^H (|x| = 1), the matvec H p
// supplied by the caller between the steps. Device state (doubles) at these offsets:
constexpr int CG_RHO   = 0; // |r|^2 of the current iteration (two slots, iteration k uses slot k & 1)
constexpr int CG_STOP  = 2; // stop flag (two slots): converged to rtol or broke down
constexpr int CG_RHO0  = 4; // |b|^2
constexpr int CG_ITS   = 5; // completed steps
constexpr int CG_RR    = 6; // |r|^2 after the last completed step
constexpr int CG_STATE = 8;
// r holds the FP32 residual of the Ritz pair on entry: b = -sigma P r goes to r and p, t = 0; d_c = x^H r.
template <typename T>
void cg_init(int64_t n, const T *x, const double *d_c, double sigma, T *r, T *p, T *t, double *d_rr, double *state, double *scratch,
             hipStream_t s);
// q = sigma (q - theta p), after q = H p (skipped once stopped).
template <typename T>
void cg_shift(int64_t n, const T *p, T *q, double theta, double sigma, const double *state, int k, hipStream_t s);
// One CG step of iteration k given d_dots = {x^H q, p^H q} (complex): t, r, p and the state advance.
template <typename T>
void cg_step(int64_t n, const T *x, T *p, const T *q, const double *d_dots, T *t, T *r, double *d_rr, double *state, int k,
             double rtol2, double *scratch, hipStream_t s);

} // namespace dmrg

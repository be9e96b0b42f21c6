/*
 * dmrg_amd — MI355X (gfx950) implementation of the DMRG++ effective-Hamiltonian hot path.
 *
 * C ABI. Plain pointers and sizes only; no exception crosses this boundary. Every call
 * returns a dmrg_status; on failure dmrg_last_error() gives a message (thread-local).
 *
 * Tensor layouts are those of the reference (Eigen column-major, complex interleaved
 * as std::complex<double>):
 *   mps / theta  (d, chiL, chiR)        element [s,a,b]   at s + d*(a + chiL*b)
 *   env block    (chi, chi, D)          element [a,a',w]  at a + chi*(a' + chi*w)   (index 0 = ket)
 *   mpo          (DL, DR, d, d)         element [w,w',s,s'] at w + DL*(w' + DR*(s + d*s'))
 * Pointers named d* or documented "device" are device (HBM) pointers; "host" are host.
 * A NULL stream means the default stream.
 */
#ifndef DMRG_AMD_H
#define DMRG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dmrg_status {
    DMRG_OK               = 0,
    DMRG_ERR_DIMS         = 1, /* dimension mismatch (reference: except::runtime_error, contraction_tblis.cpp:230-234) */
    DMRG_ERR_ARG          = 2, /* invalid argument */
    DMRG_ERR_OOM          = 3, /* device allocation failed */
    DMRG_ERR_HIP          = 4, /* HIP runtime error */
    DMRG_ERR_NAN          = 5, /* NaN/Inf detected */
    DMRG_ERR_NOT_CONVERGED = 6, /* eigensolver hit its iteration or time limit */
    DMRG_ERR_UNSUPPORTED  = 7
} dmrg_status;

typedef enum dmrg_dtype {
    DMRG_FP64 = 0, /* double                                      (reference fp64) */
    DMRG_CX64 = 1, /* std::complex<double>, interleaved            (reference cx64) */
    DMRG_FP32 = 2, /* float: inner precision of the mixed path */
    DMRG_CX32 = 3  /* std::complex<float>: inner precision of the mixed path */
} dmrg_dtype;

/* ---------------------------------------------------------------- runtime */
const char *dmrg_last_error(void);
int         dmrg_version(void);
dmrg_status dmrg_device_count(int *count);
dmrg_status dmrg_set_device(int device);
dmrg_status dmrg_device_info(char *name, size_t name_len, int *cu_count, size_t *hbm_bytes);
dmrg_status dmrg_malloc(void **dptr, size_t bytes);
dmrg_status dmrg_free(void *dptr);
dmrg_status dmrg_memcpy_htod(void *dst, const void *src, size_t bytes);
dmrg_status dmrg_memcpy_dtoh(void *dst, const void *src, size_t bytes);
dmrg_status dmrg_memcpy_dtod(void *dst, const void *src, size_t bytes);
dmrg_status dmrg_memset(void *dptr, int value, size_t bytes);
dmrg_status dmrg_synchronize(void);

/* Stream and event helpers (default priority streams only). */
dmrg_status dmrg_stream_create(void **stream);
dmrg_status dmrg_stream_destroy(void *stream);
dmrg_status dmrg_stream_synchronize(void *stream);
dmrg_status dmrg_event_create(void **event);
dmrg_status dmrg_event_destroy(void *event);
dmrg_status dmrg_event_record(void *event, void *stream);
dmrg_status dmrg_event_elapsed_ms(float *ms, void *start, void *stop);

/* ---------------------------------------------------------------- contractions (device pointers) */

/* out[s',a',b'] = sum L[a,a',w] W[w,w',s,s'] R[b,b',w'] theta[s,a,b].
 * Replaces tools::common::contraction::matrix_vector_product<Scalar>(res, mps, mps_dims, mpo, mpo_dims,
 * envL, envL_dims, envR, envR_dims)  (contraction.h:31-36; contraction_tblis.cpp:213-292), Scalar in
 * {fp64, cx64}; FP32/CX32 run the same contraction on FP32 matrix cores. Same dimension checks
 * (contraction_tblis.cpp:230-234) -> DMRG_ERR_DIMS. Workspace is held by the library; the stateless
 * entry points (this one, dmrg_env_build_block, dmrg_expectation_value, dmrg_dot, dmrg_nrm2) share it, so
 * calls of them on different streams are ordered one after the other on the device. */
dmrg_status dmrg_matrix_vector_product(dmrg_dtype dtype, void *res, const void *mps, const int64_t mps_dims[3],
                                       const void *mpo, const int64_t mpo_dims[4], const void *envL,
                                       const int64_t envL_dims[3], const void *envR, const int64_t envR_dims[3],
                                       void *stream);

/* Allocation-free form of dmrg_matrix_vector_product: same arguments, result and dimension checks. No counterpart
 * in the reference; it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The two intermediates (d*chiL*chiR*(DL+DR) elements) live in a
 * fixed device array of DMRG_NOALLOC_SCRATCH_BYTES per GPU (held from the first call on that GPU until the process
 * ends); larger problems are refused with DMRG_ERR_ARG. No leg is split over K and GEMM plans are never timed (a
 * plan tuned earlier for the same shape is used, else the size heuristic's). The call allocates and frees
 * nothing. Calls that overlap on one GPU (other streams or host threads) share the array and race. */
#define DMRG_NOALLOC_SCRATCH_BYTES (128LL << 20)
dmrg_status dmrg_matrix_vector_product_noalloc(dmrg_dtype dtype, void *res, const void *mps, const int64_t mps_dims[3],
                                               const void *mpo, const int64_t mpo_dims[4], const void *envL,
                                               const int64_t envL_dims[3], const void *envR, const int64_t envR_dims[3],
                                               void *stream);
/* DMRG_NOALLOC_SCRATCH_BYTES as compiled into the library. */
int64_t     dmrg_matrix_vector_product_noalloc_max_bytes(void);

/* Environment update: side 'L': block[b,b',w'] = sum theta[s,a,b] other[a,a',w] W[w,w',s,s'] conj(theta[s',a',b']);
 *                     side 'R': block[a,a',w]  = sum theta[s,a,b] other[b,b',w'] W[w,w',s,s'] conj(theta[s',a',b']).
 * Replaces EnvBase::build_block(otherblock, mps, mpo) (EnvBase.cpp:67-151); the reference is always cx64
 * (EnvBase.h:22-26); DMRG_FP64 is provided for real problems. block_out has shape (chiR,chiR,DR) for 'L'
 * and (chiL,chiL,DL) for 'R'. Dimension checks of EnvBase.cpp:100-108,137-145 -> DMRG_ERR_DIMS. */
dmrg_status dmrg_env_build_block(dmrg_dtype dtype, char side, void *block_out, const void *otherblock,
                                 const int64_t other_dims[3], const void *mps, const int64_t mps_dims[3],
                                 const void *mpo, const int64_t mpo_dims[4], void *stream);

/* Subspace-expansion term P = alpha * (block . mps . mpo). Replaces EnvBase::get_expansion_term(mps, mpo, alpha)
 * (EnvBase.cpp:303-361); the pre-truncation of its MpsSite overload (:363-455) is an SVD of mps that callers
 * compose with dmrg_svd_truncate. side 'L': block (chiL, chiL, DL) of the left env (energy block with the MPO,
 * variance block with the MPO^2), P (d, chiL, chiR*DR), P[s',a',b + chiR wR] = sum block[a,a',w] mps[s,a,b]
 * W[w,wR,s,s']. side 'R': block (chiR, chiR, DR), P (d, chiL*DL, chiR) holding, at storage offset
 * s' + d(b' + chiR(a + chiL wL)), sum block[b,b',w] mps[s,a,b] W[wL,w,s,s'] (the reference's shuffle {3,0,1,2}
 * followed by a plain reshape). alpha = 0 gives zeros; a NaN alpha is DMRG_ERR_ARG (EnvBase.cpp:358). With
 * DMRG_CX64 data whose mps and mpo have no imaginary part, the term is computed in real arithmetic from the real
 * part of the block and returned as complex, as the reference does (EnvBase.cpp:305-307). */
dmrg_status dmrg_env_expansion_term(dmrg_dtype dtype, char side, void *P, const void *block, const int64_t block_dims[3],
                                    const void *mps, const int64_t mps_dims[3], const void *mpo, const int64_t mpo_dims[4],
                                    double alpha, void *stream);

/* <theta|H_eff|theta> = sum conj(theta) . (H_eff theta).
 * Replaces tools::common::contraction::expectation_value (contraction_tblis.cpp:30-97). result[0..1] = re, im. */
dmrg_status dmrg_expectation_value(dmrg_dtype dtype, const void *mps, const int64_t mps_dims[3], const void *mpo,
                                   const int64_t mpo_dims[4], const void *envL, const int64_t envL_dims[3],
                                   const void *envR, const int64_t envR_dims[3], double result[2], void *stream);

/* sum_i conj(x_i) y_i. Replaces contract_mps_mps_overlap / contract_mps_overlap (contraction_tblis.cpp:518-533). */
dmrg_status dmrg_dot(dmrg_dtype dtype, int64_t n, const void *x, const void *y, double result[2], void *stream);
/* sqrt(sum |x_i|^2). Replaces contract_mps_norm (contraction.h:284-301). */
dmrg_status dmrg_nrm2(dmrg_dtype dtype, int64_t n, const void *x, double *result, void *stream);

/* ---------------------------------------------------------------- MatVecMPOS operator (device resident) */

typedef struct dmrg_matvec dmrg_matvec;

/* Builds the operator from host data in the reference's storage: nmpo MPOs (cx64, each (DL,DR,d,d), dims in
 * mpo_dims[4*i..4*i+3]) and the env blocks (cx64, (chiL,chiL,DL) and (chiR,chiR,DR)). With dtype DMRG_FP64 the
 * data are cast to real, as MatVecMPOS<fp64> does (matvec_mpos.cpp:54-65); DMRG_CX64 keeps them complex.
 * Up to 5 MPOs are pre-merged into one (DL, DR, prod d, prod d) MPO (matvec_mpos.cpp:104-119); longer chains
 * are applied one MPO at a time (contraction_tblis.cpp:307-420).
 * Replaces MatVecMPOS<T>::MatVecMPOS(mpos, envs) (matvec_mpos.cpp:40-130). Data are copied to the device once. */
dmrg_status dmrg_matvec_create(dmrg_matvec **op, dmrg_dtype dtype, int nmpo, const void *const *mpos,
                               const int64_t *mpo_dims, const void *envL, const int64_t envL_dims[3],
                               const void *envR, const int64_t envR_dims[3]);
dmrg_status dmrg_matvec_destroy(dmrg_matvec *op);
/* The B operator of MatVecMPOS(mpos, enva, envb) (matvec_mpos.h:116-120, members mpos_B, envL_B, envR_B): MPOs
 * and env blocks in the layout of dmrg_matvec_create, copied to the device; its theta shape must equal op's.
 * Replaces the envb half of that constructor (matvec_mpos.cpp:40-130). Setting it again replaces it. */
dmrg_status dmrg_matvec_set_b(dmrg_matvec *op, int nmpo, const void *const *mpos, const int64_t *mpo_dims, const void *envL,
                              const int64_t envL_dims[3], const void *envR, const int64_t envR_dims[3]);
/* MultBx on host buffers (matvec_mpos.cpp:986-1019): y[:, i] = B x[:, i] for i < nvec; like the reference it
 * returns without writing y when no B operator was set, and counts nvec products in num_mv. */
dmrg_status dmrg_matvec_multbx(dmrg_matvec *op, const void *x, int64_t ldx, void *y, int64_t ldy, int nvec);
/* MultBx on device buffers, stream-ordered. */
dmrg_status dmrg_matvec_multbx_device(dmrg_matvec *op, const void *dx, int64_t ldx, void *dy, int64_t ldy, int nvec,
                                      void *stream);
/* MatVecMPOS::rows(): n = prod(d) * chiL * chiR. */
int64_t     dmrg_matvec_rows(const dmrg_matvec *op);
/* MatVecMPOS::get_shape_mps(): {prod d, chiL, chiR}. */
dmrg_status dmrg_matvec_shape_mps(const dmrg_matvec *op, int64_t shape[3]);
/* MatVecMPOS::num_mv counter (matvec_mpos.cpp:973). */
int64_t     dmrg_matvec_num_mv(const dmrg_matvec *op);

/* y[:,i] = H_eff x[:,i], i < nvec, columns at stride ldx / ldy elements, HOST buffers.
 * Replaces MatVecMPOS<T>::MultAx(void *x, int *ldx, void *y, int *ldy, int *blockSize, primme_params*, int *err)
 * (matvec_mpos.cpp:976-984) and MultAx(T*, T*) (matvec_mpos.cpp:963-974): the PRIMME / ARPACK callback path. */
dmrg_status dmrg_matvec_multax(dmrg_matvec *op, const void *x, int64_t ldx, void *y, int64_t ldy, int nvec);
/* Same with DEVICE buffers, stream ordered; nvec vectors are applied in one batched launch per leg. */
dmrg_status dmrg_matvec_multax_device(dmrg_matvec *op, const void *dx, int64_t ldx, void *dy, int64_t ldy, int nvec,
                                      void *stream);
/* Mixed-precision inner product: DEVICE FP32 vectors (DMRG_FP32 / DMRG_CX32 of the operator's kind) through
 * FP32 matrix cores with FP32 accumulation. No counterpart in the reference (it declares T32 but never uses
 * it, matvec_mpos.h:29,35). */
dmrg_status dmrg_matvec_multax_device_fp32(dmrg_matvec *op, const void *dx, int64_t ldx, void *dy, int64_t ldy,
                                           int nvec, void *stream);
/* Replaces MatVecMPOS<T>::set_shift(cx64 sigma) (matvec_mpos.cpp:1433-1472): adds (sigma_old - sigma)/nmpo * I to
 * the bottom-left corner W[DL-1,0,:,:] of every MPO of the operator. */
dmrg_status dmrg_matvec_set_shift(dmrg_matvec *op, double re, double im);
/* Diagonal of H_eff (length rows), DEVICE output, operator dtype. Replaces MatVecMPOS::get_diagonal_new
 * (matvec_mpos.cpp:399-407); feeds the Jacobi preconditioner (matvec_mpos.cpp:1333-1416, jcb = 1). */
dmrg_status dmrg_matvec_get_diagonal(dmrg_matvec *op, void *ddiag, void *stream);
/* Jacobi preconditioner on host buffers, the PRIMME applyPreconditioner callback MatVecMPOS::MultPc with
 * jcbMaxBlockSize = 1 (matvec_mpos.cpp:1333-1361): y[:, i] = x[:, i] / (diag(H_eff) - shifts[i]) for i < nvec
 * (shifts = PRIMME's ShiftsForPreconditioner; NULL means 0). The diagonal is computed once per operator state
 * (again after set_shift). */
/* MatVecMPOS::set_jcbMaxBlockSize (matvec_mpos.h): 1 selects the Jacobi (diagonal) preconditioner of
 * dmrg_matvec_precond, > 1 the block-Jacobi one (inverses of the diagonal blocks of (H_eff - shift) of this size,
 * built at the next dmrg_matvec_precond call with its first shift, and again after set_shift or a size change). */
dmrg_status dmrg_matvec_set_jcb_block_size(dmrg_matvec *op, int64_t size);
/* The MatVecMPOS::factorization of the block-Jacobi blocks (matvec_mpos.cpp:1143-1240), as eig::Factorization
 * values 0 NONE, 1 LLT, 2 LDLT, 3 LU (the default), 4 QR, 5 ILUT, 6 ILDLT; others DMRG_ERR_UNSUPPORTED. LLT keeps
 * the reference's sign rule: a block with Re B(0,0) <= 0 whose negation is positive definite is applied as -B^-1,
 * every other block as B^-1. Every other value applies the exact inverse B^-1 of each nonsingular block: LDLT, LU
 * and QR solve with it; NONE falls back to it (the reference's ColPivHouseholderQR inverse); ILUT and ILDLT, which
 * the reference applies as incomplete factorizations with iterative block solves to 1e-3, get the exact inverse
 * instead. The choice matters only for block sizes > 1. */
dmrg_status dmrg_matvec_set_factorization(dmrg_matvec *op, int factorization);
/* MatVecMPOS::FactorOP (matvec_mpos.cpp:887-913), for shift-invert: forms the dense H_eff of the operator as it
 * stands (with the shift of dmrg_matvec_set_shift), in the orientation whose product with x is the matvec, and
 * factors it on the device: factorization 1 (LLT) with Cholesky, 3 (LU) and 2 (LDLT) with pivoted LU (an exact
 * solve with the same matrix); 0 (NONE) marks the operator ready without a factor, as the reference does; other
 * values DMRG_ERR_UNSUPPORTED. A matrix that is not positive definite (LLT) or is singular (LU) is DMRG_ERR_ARG.
 * n^2 elements of device memory, at most 8 GiB (DMRG_ERR_OOM beyond). set_shift afterwards clears it. */
dmrg_status dmrg_matvec_factor_op(dmrg_matvec *op, int factorization, void *stream);
/* MatVecMPOS::isReadyFactorOp. */
int         dmrg_matvec_factor_ready(const dmrg_matvec *op);
/* MatVecMPOS::MultOPv (matvec_mpos.cpp:936-975): y_i = (H_eff)^-1 x_i for nvec host vectors (columns at strides
 * ldx / ldy) with the factor of dmrg_matvec_factor_op; counted in dmrg_matvec_num_op. */
dmrg_status dmrg_matvec_mult_opv(dmrg_matvec *op, const void *x, int64_t ldx, void *y, int64_t ldy, int nvec);
int64_t     dmrg_matvec_num_op(const dmrg_matvec *op);
dmrg_status dmrg_matvec_precond(dmrg_matvec *op, const void *x, int64_t ldx, void *y, int64_t ldy, int nvec,
                                const double *shifts);

/* ---------------------------------------------------------------- eigensolver */

#define DMRG_EIGS_MAX_NEV 16
typedef enum dmrg_eigs_target {
    DMRG_TARGET_SA      = 0, /* smallest algebraic (primme_smallest; fDMRG ground state) */
    DMRG_TARGET_LA      = 1, /* largest algebraic  (primme_largest) */
    DMRG_TARGET_CLOSEST = 2  /* closest to shift   (primme_closest_abs; xDMRG on H^2) */
} dmrg_eigs_target;

/* Subset of eig::settings (source/math/eig/settings.h:12-70) that the GPU Davidson honours. */
typedef struct dmrg_eigs_config {
    int      nev;         /* maxNev, default 1 */
    int      ncv;         /* maxNcv (basis size); <= 0: nev*ceil(log2 n) clamped to [2 nev, n] (solver.cpp:12-15) */
    int      min_restart; /* primme_minRestartSize; <= 0: max(1, ncv/2 - 1) (solver_primme.cpp:374) */
    int      target;      /* dmrg_eigs_target */
    double   shift;       /* target shift for DMRG_TARGET_CLOSEST (primme_targetShifts) */
    double   tol;         /* eps: converged when |r| <= tol * |A|est (|A|est = largest |Ritz value| seen) */
    int64_t  max_matvecs; /* default 100000 (eigs_iter_max, settings.h:251) */
    int      precond;     /* 0: none, 1: Jacobi (diagonal of H_eff, MultPc with jcb = 1), 2: block-Jacobi with
                             blocks of jcb_block (MultPc with jcbMaxBlockSize > 1, LU), built at the target shift
                             for DMRG_TARGET_CLOSEST, else at the first Ritz value */
    int      mixed;       /* 0: FP64 throughout; 1: FP32 matrix-core phase then FP64 refinement */
    double   max_time_s;  /* 0: unlimited (reference: config.maxTime) */
    uint64_t seed;        /* seed of the random start vector when no initial guess is given */
    int64_t  jcb_block;   /* block size of precond = 2 (jcbMaxBlockSize, default 128, settings.h:261) */
    int      inner_max;   /* mixed = 1, one pair at SA or LA, precond = 0: the FP64 phase expands with at most this
                             many FP32 conjugate-gradient steps on the Jacobi-Davidson correction equation
                             (default 20; 0: plain FP64 Davidson refinement) */
    double   inner_tol;   /* those steps stop at this reduction of the correction residual (default 1e-2) */
    int      projection;  /* 0: Rayleigh-Ritz; 1 (with DMRG_TARGET_CLOSEST): refined projection, the reference's
                             primme_proj_refined for folded-spectrum solves (optimize_folded_spectrum.cpp:295-299):
                             x = V y with y the right singular vector of the smallest singular value of the
                             triangular factor of W - shift V, its value the Rayleigh quotient */
    int      conv_test;   /* 0: |r| <= tol |A|; 1: folded_spectrum::convTestFun (optimize_folded_spectrum.cpp:41-93):
                             |r| < tol |A| and the Ritz values or residual norms of the last up to 500 iterations
                             (at least 250) no longer decreasing by more than tol, or |r| < tol */
} dmrg_eigs_config;

typedef struct dmrg_eigs_result {
    double  eigval[DMRG_EIGS_MAX_NEV];
    double  rnorm[DMRG_EIGS_MAX_NEV];
    int     nconv;        /* converged pairs; < nev when the matvec/time limit was reached */
    int     iters;
    int     restarts;
    int     stagnated;
    int64_t matvecs_fp64; /* FP64 H_eff applications (block matvecs count each vector) */
    int64_t matvecs_fp32; /* FP32 applications (mixed mode: FP32 phase plus correction solves) */
    double  seconds;
    double  anorm;
} dmrg_eigs_result;

void        dmrg_eigs_default_config(dmrg_eigs_config *cfg);
/* Eigenpairs of the operator. x0: optional initial guess (DEVICE, n elements of the operator dtype, or NULL);
 * evecs: DEVICE output, n x nev column-major. Replaces eig::solver::eigs(MatVecMPOS<T>&) with the PRIMME backend
 * (solver.cpp:491-509, solver_primme.cpp:297-517): same problem (H_eff x = lambda x, nev pairs nearest the target),
 * a GPU Davidson instead of PRIMME's GD_Olsen_plusK. Returns DMRG_OK also when not converged (res->nconv < nev). */
dmrg_status dmrg_eigs(dmrg_matvec *op, const dmrg_eigs_config *cfg, const void *x0, void *evecs, dmrg_eigs_result *res,
                      void *stream);

/* Optimally mixed block of subspace expansion. Replaces tools::finite::env::get_optimally_mixed_block_1site
 * (env.cpp:96-239): V = MGS([theta, H1 theta, H2 theta]) (env.cpp:74-93), K1 = V^H H1 V, K2 = V^H H2 V (lower
 * triangles), K3 = K2 - K1 K1 with the mode's decoupling, the eigenvector of K3's lowest eigenvalue (or, when a
 * K has a row below tol, the one with the largest weight on theta), mixed = V * col. h1: the energy operator,
 * h2: the operator on the variance envs (MPO^2), same shape and dtype (FP64 or CX64). mode: bit 1 = H1,
 * bit 2 = H2 (EnvExpandMode values). theta, mixed: DEVICE, n elements. alphas = {alpha_mps, alpha_h1v,
 * alpha_h2v}; opt_val = the chosen K3 eigenvalue. The eigenvector is rotated so that its largest component is
 * real positive before its real part is taken (the reference keeps Eigen's unnormalised phase): alphas and
 * mixed match the reference up to one overall sign for real problems. */
dmrg_status dmrg_optimally_mixed_block(dmrg_matvec *h1, dmrg_matvec *h2, const void *theta, int mode, void *mixed,
                                       double alphas[3], double *opt_val, void *stream);

/* Folded-spectrum executor. Replaces eigs_manager_folded_spectrum + eigs_folded_spectrum_executor
 * (optimize_folded_spectrum.cpp:235-342): the eigensolve of dmrg_eigs on h2, the operator on the variance envs
 * with the MPO^2 of H - energy_shift (cfg->target DMRG_TARGET_CLOSEST with cfg->shift = eigv_target is the
 * reference's primme_closest_abs for OptRitz::SM; cfg->precond = 1 its Jacobi preconditioner), warm-started from
 * x0, then per returned vector the results extract_results reports (opt_extract_results.cpp:19-90):
 * energy[q] = <x|H|x>/<x|x> with h, variance[q] = <x|(H - E_shift)^2|x>/<x|x> - (energy - E_shift)^2. */
dmrg_status dmrg_eigs_folded_spectrum(dmrg_matvec *h, dmrg_matvec *h2, const dmrg_eigs_config *cfg, const void *x0,
                                      double energy_shift, void *evecs, double *energy, double *variance,
                                      dmrg_eigs_result *res, void *stream);

/* Batched eigensolves: the lowest (DMRG_TARGET_SA) or highest (LA) eigenpair of each of k operators of uneven
 * shapes (one dtype, FP64 or CX64), advanced together in Davidson steps (each problem with its own basis size and
 * restarts): per step one grouped matvec of all operators (replayed from a HIP graph), grouped block-vector
 * kernels and one host synchronisation for the whole batch; a converged problem leaves the batch. No counterpart in the reference (eig::solver::eigs solves
 * one operator, solver_primme.cpp:297-517; the method and the basis rule are the same as dmrg_eigs): for request
 * loops with many small resident problems. cfg: nev 1, precond 0, target SA or LA; tol, ncv, max_matvecs,
 * seed as in dmrg_eigs. cfg->mixed = 1: each problem first runs in an FP32 batch (FP32 vectors, grouped FP32
 * matrix-core matvecs, double reductions) to 5e-6 |A| or until its residual stops dropping by 10% over 25 steps,
 * then its FP32 Ritz vector starts it in the FP64 batch, which converges it to tol; results report both matvec
 * counts. Problems with n >= 131072 fill the GPU with their own matvecs: they are solved first, one after
 * another, by the single-problem solver of dmrg_eigs (with cfg->mixed, its FP32 phase and cfg->inner_max
 * correction-equation steps), and only the smaller ones run in the lockstep batch. x0s: NULL or k device start
 * vectors (entries may be NULL); evecs: k device outputs. */
dmrg_status dmrg_eigs_batched(dmrg_matvec *const *ops, int k, const dmrg_eigs_config *cfg, const void *const *x0s,
                              void *const *evecs, dmrg_eigs_result *results, void *stream);

/* ---------------------------------------------------------------- grouped matvecs for serving */

/* Matvecs of k operators of uneven shapes (one dtype) as one launch per contraction leg: y_i = H_i x_i.
 * No counterpart in the reference (its MultAx serves one operator at a time); for request loops that keep many
 * small operators resident, where one operator's matvec fills only a few CUs. fp32 = 1 runs the FP32 matrix-core
 * form (xs/ys in the single-precision dtype). capture_graph = 1 records the launches once into a HIP graph that
 * dmrg_matvec_group_run replays. The group holds pointers into the operators' buffers and into xs/ys: keep
 * those (and the operators' shift/MPO state) unchanged while the group is used; destroy it first. With
 * dmrg_autotune_enable on (the default), creating a group of a new shape class times the tile configurations of
 * each leg on these operands (so ys receives H_i x_i during creation) and caches the fastest. Creation
 * synchronises the device before it tunes, so xs must be written (by work already enqueued on any stream)
 * when it is called, and no other work may use ys meanwhile. */
typedef struct dmrg_matvec_group dmrg_matvec_group;
dmrg_status dmrg_matvec_group_create(dmrg_matvec *const *ops, int k, const void *const *xs, void *const *ys, int fp32,
                                     int capture_graph, dmrg_matvec_group **out);
/* Enqueue one grouped matvec of every operator on stream (stream-ordered; no host synchronisation). */
dmrg_status dmrg_matvec_group_run(dmrg_matvec_group *group, void *stream);
/* Algorithmic FLOP of one run (8 per complex MAC, 2 per real MAC). */
double      dmrg_matvec_group_flop(const dmrg_matvec_group *group);
void        dmrg_matvec_group_destroy(dmrg_matvec_group *group);

/* ---------------------------------------------------------------- finite DMRG sweep (ising_tf_rf) */

#define DMRG_FDMRG_MAX_L 256
/* Ground state of H = -sum J sz_i sz_{i+1} - sum (h_tran + h_rand[i]) sx_i on an open chain (ising_tf_rf,
 * IsingRandomField.cpp:93-136; the random field sits on the transverse sx term, J = 0 on the last site).
 * Fields are passed explicitly (h_rand[L]). */
typedef struct dmrg_fdmrg_config {
    int           L;
    double        J;               /* J1 */
    double        h_tran;          /* g */
    const double *h_rand;          /* host, L entries */
    int           chi_max;         /* bond_max */
    int           max_sweeps;      /* single-site sweeps after the two-site phase (default 20) */
    int           two_site_sweeps; /* two-site sweeps that grow the bonds (default 10, stop early on |dE| < 1e-12) */
    double        var_tol;         /* variance_convergence_threshold (settings.h:275), default 1e-12 */
    double        trunc_lim;       /* svd_truncation_min (settings.h:265), default 1e-14 */
    double        eigs_tol;        /* eigensolver tolerance, default 1e-12 */
    int           eigs_ncv;        /* 0: basis size rule */
    int           mixed;           /* 1: FP32 matrix-core eigensolves with FP64 refinement for problems of size >= mixed_min_n */
    int           complex_;        /* 1: complex128 tensors (the model is real; for testing the complex path) */
    uint64_t      seed;            /* initial random product state */
    int64_t       mixed_min_n;     /* <= 0: 131072 (smaller warm-started solves converge faster in FP64 alone) */
    int           xdmrg;           /* 1: excited-state search (xDMRG): every solve minimises <(H - energy_target)^2> with the
                                      folded-spectrum executor on the MPO^2 operator (optimize_folded_spectrum.cpp:235-342,
                                      closest_abs to 0), the shift fixed at the target (xdmrg.cpp:487-501) */
    double        energy_target;   /* the target energy E_tgt of xdmrg = 1 (absolute, not a density) */
    int64_t       eig_max_size;    /* ground-state solves of problems up to this size use the dense solver, as the
                                      reference's OptSolver::EIG for problem_size <= eig_max_size (fdmrg.cpp:307,
                                      AlgorithmFinite.cpp:318; the reference's default is 4096); 0 (the default):
                                      the iterative solver at every size, which is faster on the GPU */
    int           expand_mode;     /* EnvExpandMode bits (config/enums.h:224-234; dmrg_env_expand_mode, settings.h:228):
                                      0 (NONE): two-site sweeps grow the bonds, then single-site sweeps; H1 (1) and/or
                                      H2 (2) with NSITE (8) | FORWARD (32), e.g. the reference's default 43 (also
                                      dmrg_fdmrg_default_config's): single-site
                                      sweeps from the product state, each step preceded by subspace expansion of the active
                                      site and the next one in the sweep direction (env.cpp:1285-1380, the optimally mixed
                                      block of env.cpp:96-239, split with chi_max and trunc_lim). SSITE and BACKWARD:
                                      DMRG_ERR_UNSUPPORTED */
} dmrg_fdmrg_config;

typedef struct dmrg_fdmrg_result {
    double  energy;
    double  variance;
    double  entropies[DMRG_FDMRG_MAX_L + 1]; /* von Neumann entropy per bond 0..L (measure.cpp:134-172) */
    int64_t chi[DMRG_FDMRG_MAX_L + 1];
    int     sweeps;
    int     eigensolves;
    int     converged;                       /* variance < var_tol */
    int64_t matvecs;
    int64_t matvecs_fp32;
    double  seconds;
    double  t_eigs, t_env, t_svd;
    double  trunc_error_max;                 /* max over bonds of |S[discarded]|_2 (svd.cpp:257-280) */
    double  gflop;                           /* algorithmic matrix-core work: matvecs (FP64 and FP32) and env updates */
    int     expansions;                      /* subspace expansions (expand_mode != 0) */
    double  t_expand;                        /* seconds in the optimally mixed blocks of the expansions */
} dmrg_fdmrg_result;

void        dmrg_fdmrg_default_config(dmrg_fdmrg_config *cfg);
/* Replaces fdmrg::run_algorithm (fdmrg.cpp:170-195) for ising_tf_rf: sweeps with all tensors on the GPU; with
 * xdmrg = 1 the xdmrg::run_algorithm loop (xdmrg.cpp:297-341) on the folded spectrum around energy_target. */
dmrg_status dmrg_fdmrg_ising_tf_rf(const dmrg_fdmrg_config *cfg, dmrg_fdmrg_result *res, void *stream);

/* ---------------------------------------------------------------- serving: many independent solves on one GPU */

typedef struct dmrg_request {
    int           L;
    int           chi_max;
    const double *h_rand;    /* host, L fields of this disorder realisation */
    double        arrival_s; /* arrival time relative to the start of dmrg_serve */
    uint64_t      seed;      /* initial state seed */
} dmrg_request;

typedef struct dmrg_request_result {
    double  energy;
    double  variance;
    double  entropy_mid;
    double  arrival_s, start_s, finish_s, latency_s; /* latency = finish - arrival */
    int     worker;
    int     status; /* DMRG_OK, DMRG_ERR_NOT_CONVERGED (variance above var_tol) or an error */
    int     sweeps;
    int64_t matvecs;
    int64_t footprint_bytes;
    double  gflop; /* algorithmic matrix-core work of the request (dmrg_fdmrg_result.gflop) */
} dmrg_request_result;

typedef struct dmrg_serve_config {
    int     workers;          /* concurrent requests (one HIP stream each) */
    int64_t hbm_budget_bytes; /* admission limit on the summed HBM footprint of running requests */
    int     mixed;
    int     complex_;
    double  var_tol, eigs_tol;
    int     max_sweeps, two_site_sweeps;
    double  J, h_tran;
    int     batch_eigs; /* 1: the requests' iterative FP64 energy solves run together in one batch (grouped matvec
                           launches, one synchronisation per Davidson step for all of them); 0: each on its own stream */
    int     expand_mode; /* the requests' dmrg_fdmrg_config.expand_mode: 0 (the zero default here) grows the bonds with
                            two-site sweeps; 43 is the reference's subspace expansion */
} dmrg_serve_config;

typedef struct dmrg_serve_stats {
    double  makespan_s;
    double  solves_per_s;
    double  p50_latency_s, p99_latency_s, mean_latency_s;
    int     completed, failed;
    int64_t peak_admitted_bytes;
    int     peak_concurrency;
    int64_t matvecs;
    int64_t batch_steps; /* batched eigensolver steps (one grouped matvec + one synchronisation each); 0 unbatched */
    int     peak_batch;  /* most eigensolves in one batched step */
    double  batch_step_s, batch_rebuild_s, batch_rr_s; /* batcher thread: time in steps, in rebuilds of the
                                                          grouped launches, in the host Rayleigh-Ritz solves */
} dmrg_serve_stats;

/* Runs the requests as they arrive (real time), admitting each when a worker and HBM budget are free, and
 * fills out[nreq] and stats. The reference's counterpart is one process per disorder seed
 * (slurm/batches_fdmrg.py:8-27); here the solves share one GPU. */
dmrg_status dmrg_serve(const dmrg_serve_config *cfg, const dmrg_request *reqs, int nreq, dmrg_request_result *out,
                       dmrg_serve_stats *stats);
/* Estimated HBM footprint (bytes) that admission charges for a request. */
int64_t     dmrg_request_footprint(int L, int chi_max, int complex_);

/* Continuous batching of eigensolve requests on resident operators: each request joins the running batch of
 * dmrg_eigs_batched between two Davidson steps once it has arrived (arrival_s, real time from the call) and its
 * working memory (basis and image vectors, restart temporaries and matvec workspace: footprint_bytes) fits in
 * hbm_budget_bytes beside that of the running requests, and while fewer than max_batch problems run
 * (max_batch <= 0: no limit); it leaves when converged (or at max_matvecs). Requests are admitted in arrival
 * order. Each problem advances as it would alone (its own basis size, phase and restarts), so a request's
 * result is the one it gets alone up to rounding (the grouped products' tile shapes depend on the set). No
 * counterpart in the reference (one eigensolve per call, solver_primme.cpp:297-517): for request loops. cfg as in
 * dmrg_eigs_batched; the operators must stay alive until the call returns. A request whose footprint alone
 * exceeds hbm_budget_bytes fails the whole call with DMRG_ERR_OOM before any work starts (as dmrg_serve).
 * stats: latency percentiles over all requests; completed counts converged ones and solves_per_s is
 * completed / makespan (as dmrg_serve). */
typedef struct dmrg_eig_request {
    dmrg_matvec *op;
    const void  *x0;        /* device start vector or NULL (seeded random start) */
    void        *evec;      /* device output, n elements */
    double       arrival_s; /* arrival time relative to the start of dmrg_serve_eigs */
} dmrg_eig_request;

typedef struct dmrg_eig_request_result {
    dmrg_eigs_result res;
    double           arrival_s, start_s, finish_s, latency_s; /* latency = finish - arrival */
    int64_t          footprint_bytes;
} dmrg_eig_request_result;

dmrg_status dmrg_serve_eigs(const dmrg_eigs_config *cfg, const dmrg_eig_request *reqs, int nreq, int64_t hbm_budget_bytes,
                            int max_batch, dmrg_eig_request_result *out, dmrg_serve_stats *stats, void *stream);

/* Host dense Hermitian eigensolver used for the eigensolver's Rayleigh-Ritz step (the role LAPACK ?heevd
 * plays inside PRIMME). A, Z: m x m column-major complex interleaved; w ascending. */
dmrg_status dmrg_host_heev(int m, const double *A, double *w, double *Z);

/* GEMM autotuner (no counterpart in the reference). Every contraction leg is a GEMM whose launch plan (tile
 * configuration, K slices) is chosen by timing the candidates with HIP events the first time a problem
 * signature is seen on a stream that is not being captured, and cached for the process. */
dmrg_status dmrg_autotune_enable(int enable);
dmrg_status dmrg_autotune_reset(void);
/* Only GEMMs of at least macs multiply-adds (M N K batch; default 2^29: the chi = 512 matvec legs) or of at least
 * macs / 16 bytes of operands (the chi = 512 MPO leg) are tuned; smaller ones take a size heuristic, since a sweep
 * or a request stream meets thousands of small shapes once each and their tuning costs more than it saves.
 * 0 tunes every GEMM. */
dmrg_status dmrg_autotune_min_macs(int64_t macs);
int         dmrg_autotune_entries(void);
/* Testing hook: force every GEMM to tile configuration cfg (0..5: 64x64, 32x64, 64x32, 16x128, 128x64, 32x32;
 * 6: the column-streaming kernel for tiny K, which falls back to 64x64 beyond K = 32 or M = 64; 7: 32x32 with a
 * deep K stage of 128 (4-byte elements) or 64 elements) with ksplit K slices; cfg < 0 restores the
 * tuned/heuristic choice. */
dmrg_status dmrg_gemm_plan_override(int cfg, int ksplit);
/* Testing hook: force every GEMM to a hybrid plan of configuration cfg (0..2): output columns [0, nsplit) in one
 * launch without K slicing, the rest in a second launch with ksplit2 K slices (GEMMs without split-K scratch,
 * or whose column maps do not split at nsplit, run as one launch). Cleared by dmrg_gemm_plan_override(-1, 0). */
dmrg_status dmrg_gemm_hybrid_override(int cfg, int nsplit, int ksplit2);

/* Profiling: runs nrep device matvecs (one vector, x -> y) on stream and returns average durations measured with
 * HIP events on that stream: ms[0] leg 1 kernel, ms[1] leg 2 (MPO) kernel, ms[2] leg 3 main kernel, ms[3] leg 3
 * split-K reduce (0 when leg 3 is not split). No counterpart in the reference (its per-operator timers t_multAx,
 * matvec_mpos.h:174-179, time the whole call). */
dmrg_status dmrg_matvec_time_legs(dmrg_matvec *op, const void *dx, void *dy, int nrep, int fp32, float ms[4],
                                  void *stream);
/* Profiling of the eigensolver's HBM-bound block-vector kernels on a working set of n-element vectors with m basis
 * columns (allocated and filled by the call): average ms per launch (HIP events) and algorithmic bytes per launch
 * of [0] the projections V^H x (m + 1 vectors read), [1] the update y - V c (m + 1 read, 1 written), [2] the Ritz
 * vector and residual V y, W y - theta V y (2m read, 2 written). No counterpart in the reference (PRIMME's
 * orthogonalisation and restarts run inside its own BLAS calls). */
dmrg_status dmrg_block_vector_time(dmrg_dtype dtype, int64_t n, int m, int nrep, float ms[3], double bytes[3]);

/* ------------------------------------------------------------ dense path, edges, truncated SVD */

/* Explicit H_eff matrix H (n x n, leading dimension ldh >= n, device, dtype FP64 or CX64) of one MPO
 * (DL,DR,d,d) and env blocks (chiL,chiL,DL), (chiR,chiR,DR); n = d*chiL*chiR <= 65535.
 * transpose = 0: H[(s',a',b'),(s,a,b)] = sum L[a,a',w] W[w,w',s,s'] R[b,b',w'] so that H x is the matvec; the
 * orientation of contract_mpo_env / TensorsFinite::get_effective_hamiltonian (TensorsFinite.cpp:119-142).
 * transpose = 1: its transpose, the orientation of MatVecMPO::get_tensor/get_matrix (matvec_mpo.cpp:313-332). */
dmrg_status dmrg_heff_dense(dmrg_dtype dtype, void *H, int64_t ldh, const void *mpo, const int64_t mpo_dims[4],
                            const void *envL, const int64_t envL_dims[3], const void *envR, const int64_t envR_dims[3],
                            int transpose, void *stream);
/* The same matrix for an operator (merged MPO with its shift, its env blocks): MatVecMPO::get_matrix
 * (matvec_mpo.cpp:329-332) with transpose = 1. */
dmrg_status dmrg_matvec_get_matrix(dmrg_matvec *op, void *H, int64_t ldh, int transpose, void *stream);

/* All eigenvalues (ascending, host w[n]) and optionally eigenvectors (overwriting A, device, column-major)
 * of a Hermitian n x n device matrix, lower triangle referenced (rocSOLVER syevd / heevd). Replaces the dense
 * eig::solver::eig path (solver.cpp:205-245, dsyevr/zheevr) used for n <= 4096 (optimize_energy_eig.cpp:18-58).
 * Synchronises the stream. */
dmrg_status dmrg_dense_eigh(dmrg_dtype dtype, int64_t n, void *A, int64_t lda, double *w, int vectors, void *stream);

/* Chain-end environment block (chi, chi, mpo_dim), device: zero except block[i,i,:] = edge[:] for all i < chi,
 * edge a host vector of mpo_dim <= 128 elements of dtype (FP64 or CX64). Replaces EnvBase::set_edge_dims
 * (EnvBase.cpp:254-285); the edge itself is MpoSite::get_MPO_edge_left/right (MpoSite.cpp:450-547): the unit
 * vector on the last MPO row for the left edge, on the first column for the right edge. */
dmrg_status dmrg_env_set_edge(dmrg_dtype dtype, void *block_out, int64_t chi, int64_t mpo_dim, const void *edge,
                              void *stream);

/* Truncated SVD A = U S Vt of a device m x n column-major matrix (FP64 or CX64), the step of split::split_mps
 * (split.cpp:21-346) done by svd::solver (svd_lapacke.cpp:82-330). rank = #{k : ||S[k:]||_2 >= trunc_lim}
 * capped at rank_max (> 0; 0 = no cap), at least 1 (svd.cpp:257-280). Outputs: U (device, m x rank, ld m),
 * Vt (device, rank x n, ld rank), S (host, all min(m,n) singular values, descending), rank,
 * truncation_error = ||S[rank:]||_2 (the reference's definition, a norm) and discarded_weight = sum S[rank:]^2
 * (either may be NULL). U and Vt must hold m*min(m,n) and min(m,n)*n elements. method: 0 = gesvd (bidiagonal
 * QR, rocSOLVER); 3 = one-sided Jacobi on the GPU (this library's kernels; the gesvj routine that the reference's
 * automatic choice takes for small problems, svd.cpp:122-124, and the routine the sweep uses). Methods 1 (gesdd)
 * and 2 (gesvdj) return DMRG_ERR_UNSUPPORTED because rocSOLVER's versions lose the small singular values.
 * Synchronises the stream. */
dmrg_status dmrg_svd_truncate(dmrg_dtype dtype, int64_t m, int64_t n, const void *A, double trunc_lim, int64_t rank_max,
                              int method, void *U, double *S, void *Vt, int64_t *rank, double *truncation_error,
                              double *discarded_weight, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DMRG_AMD_H */

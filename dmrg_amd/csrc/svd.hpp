// Truncated SVD of a device matrix (split of theta after an optimisation step) and the
// index-remapping copies the sweep needs between tensor and matrix views.
#pragma once
#include "matvec_op.hpp"

#include <vector>

namespace dmrg {

// Thin SVD A = U diag(S) Vt of a device m x n column-major matrix (double or complex double) on rocSOLVER.
// U: m x k, Vt: k x n (k = min(m, n)), S: descending, copied to the host. A is left untouched.
// method: 0 bidiagonal QR (rocSOLVER gesvd), 3 one-sided Jacobi (this library's kernels; the gesvj routine the
// reference picks for small problems, svd.cpp:122-124). The reference also selects gesdd (svd_lapacke.cpp:162-290);
// rocSOLVER's 1 gesdd and 2 gesvdj are rejected because they lose the small singular values.
constexpr int SVD_GESVD  = 0;
constexpr int SVD_JACOBI = 3;

class DeviceSvd {
  public:
    explicit DeviceSvd(hipStream_t s);
    ~DeviceSvd();
    DeviceSvd(const DeviceSvd &) = delete;
    DeviceSvd &operator=(const DeviceSvd &) = delete;
    void set_stream(hipStream_t s);
    template <typename T>
    void compute(int m, int n, const T *A, std::vector<double> &S, int method = 0);
    // Hermitian eigendecomposition in place (rocSOLVER syevd / heevd): A (n x n, lda) is overwritten by the
    // eigenvectors when vectors is set; w receives the n eigenvalues in ascending order.
    template <typename T>
    void eigh(int n, T *A, int lda, bool vectors, std::vector<double> &w);
    template <typename T>
    void jacobi(int m, int n, const T *A, std::vector<double> &S);
    void *U() const { return u.p; }
    void *Vt() const { return vt.p; }
    int   k() const { return k_; }

  private:
    template <typename T>
    void        sort_descending(int m, int n, std::vector<double> &S);
    hipStream_t s;
    void       *handle = nullptr;
    DevBuf      a, sv, u, vt, e, info, res;
    DevBuf      bj_part, bj_q, bj_rot; // block Jacobi: partial Gram matrices, rotations, per-pair flags
    int         k_ = 0;
};

// Rank after truncation, svd.cpp:257-280: the largest r such that the discarded norm |S[r:]|_2 is still >= lim
// counts as kept, capped by rank_max (> 0) and at least min(size, rank_min). Returns (rank, |S[rank:]|_2).
std::pair<int, double> rank_from_truncation_error(const std::vector<double> &S, double lim, int rank_max, int rank_min = 1);

// out[off(om, i) + off(on, j)] = scale * in[off(im, i) + off(in_, j)] for i < M, j < N, where scale is
// 1, or d_scale_row[i] (rs = 1) or d_scale_col[j] (rs = 2) from a device array of doubles.
void remap2d(dmrg_dtype t, int64_t M, int64_t N, const void *in, Idx2 im, Idx2 in_, void *out, Idx2 om, Idx2 on,
             const double *d_scale, int rs, hipStream_t s);

} // namespace dmrg

// Host dense Hermitian eigensolver for small projected problems.
#pragma once
#include <complex>
#include <vector>

namespace dmrg {
// All eigenpairs of the Hermitian part of A (m x m, column-major, leading dimension lda).
// w: eigenvalues ascending; Z: eigenvectors (m x m, column-major). Returns false if QL did not converge.
bool hermitian_eig(int m, const std::complex<double> *A, int lda, double *w, std::complex<double> *Z);
} // namespace dmrg

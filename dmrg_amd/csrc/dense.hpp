// Dense small-problem path and environment edges:
//   - the explicit H_eff matrix of an (MPO, env L, env R) triple (TensorsFinite.cpp:119-142 contract_mpo_env,
//     and MatVecMPO::get_matrix, matvec_mpo.cpp:313-332, which is its transpose);
//   - chain-end environment blocks (EnvBase::set_edge_dims, EnvBase.cpp:254-285).
#pragma once
#include "contract.hpp"

namespace dmrg {

// H[r, c] (ldh) with r = s' + d(a' + chiL b') and c = s + d(a + chiL b):
//   H[r, c] = sum_{w, w'} L[a, a', w] W[w, w', s, s'] R[b, b', w'],
// so that H x equals the matvec. transpose writes H[c, r] instead (the MatVecMPO::get_tensor order).
void heff_dense(dmrg_dtype t, const Dims &D, const void *mpo, const void *envL, const void *envR, void *H, int64_t ldh,
                bool transpose, hipStream_t s);

// Diagonal blocks of (H_eff - shift I) for the block-Jacobi preconditioner (MatVecMPOS::get_diagonal_block,
// matvec_mpos.cpp:500-771): nb = ceil(n / jcb) blocks of jcb x jcb at offsets k*jcb, stored one after the other
// (column-major, leading dimension jcb). Rows/columns past n in the last block hold the identity.
void heff_diag_blocks(dmrg_dtype t, const Dims &D, const void *mpo, const void *envL, const void *envR, int64_t jcb,
                      double shift, void *out, hipStream_t s);

// Edge block (chi, chi, Dm): zero except block[i, i, :] = edge[:] for every i < chi. Dm <= EDGE_MAX.
constexpr int EDGE_MAX = 128;
void env_set_edge(dmrg_dtype t, void *out, int64_t chi, int64_t Dm, const void *edge_host, hipStream_t s);

} // namespace dmrg

// Subspace expansion on the GPU: the optimally mixed block (env.cpp:96-239); the expansion term itself is
// env_expansion_term in contract.hpp.
#pragma once
#include "matvec_op.hpp"

namespace dmrg {

struct MixedBlockResult {
    double alpha[3] = {1.0, 0.0, 0.0}; // alpha_mps, alpha_h1v, alpha_h2v
    double opt_val  = 0.0;              // K3 eigenvalue of the chosen vector (variance estimate)
    int    opt_idx  = 0;
    double K1[18] = {}, K2[18] = {}, K3[18] = {}; // 3 x 3 complex, column-major (re, im)
};

// mode: bit 0 = H1 (EnvExpandMode::H1), bit 1 = H2. theta and mixed: device vectors of h1's dtype (n elements).
void optimally_mixed_block(MatVecOp &h1, MatVecOp &h2, const void *theta, int mode, void *mixed, MixedBlockResult &res, hipStream_t s);

} // namespace dmrg

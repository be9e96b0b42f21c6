// Batched eigensolver for many small H_eff problems: k independent operators (uneven shapes, one dtype) are
// advanced in lockstep Davidson steps (the method of the single-problem solver: generalized Davidson with thick
// restart, one wanted pair per problem), so that one iteration of all k problems costs a fixed number of
// launches and one host synchronisation: the matvecs of all problems run as one grouped launch per contraction
// leg (replayed from a HIP graph), and the block-vector work of all problems runs in grouped kernels whose
// workgroups each serve one problem.
#pragma once
#include "eigs.hpp"

#include <vector>

namespace dmrg {

struct BatchEigsResult {
    double  eigval = 0.0, rnorm = 0.0, anorm = 0.0;
    int     nconv = 0, iters = 0;
    int64_t matvecs = 0;
};

// x0s[i]: optional start vector (device, may be null); evecs[i]: device output (n_i elements). p.nev must be 1,
// p.target SA or LA, p.precond 0. Converged problems leave the batch; the grouped launches are rebuilt when
// the active set has halved.
void eigs_batched(MatVecOp *const *ops, int k, const EigsParams &p, const void *const *x0s, void *const *evecs,
                  std::vector<BatchEigsResult> &res, hipStream_t s);

} // namespace dmrg

// Batched eigensolver for many small H_eff problems: k independent operators (uneven shapes, one dtype) are
// advanced together in Davidson steps (the method of the single-problem solver: generalized Davidson with thick
// restart, one wanted pair per problem; each problem keeps its own basis size), so that one iteration of all k
// problems costs a fixed number of launches and one host synchronisation: the matvecs of all problems run as one grouped launch per contraction
// leg (replayed from a HIP graph), and the block-vector work of all problems runs in grouped kernels whose
// workgroups each serve one problem.
#pragma once
#include "eigs.hpp"

#include <memory>
#include <vector>

namespace dmrg {

struct BatchEigsResult {
    double  eigval = 0.0, rnorm = 0.0, anorm = 0.0;
    int     nconv = 0, iters = 0;
    int64_t matvecs = 0;      // FP64 matvecs
    int64_t matvecs_fp32 = 0; // FP32 matvecs of the mixed-precision batch's first phase
};

// x0s[i]: optional start vector (device, may be null); evecs[i]: device output (n_i elements). p.nev must be 1,
// p.target SA or LA, p.precond 0. Converged problems leave the batch; the grouped launches are rebuilt when
// the set changes.
// mixed: FP32 phase (FP32 vectors and matrix-core matvecs) to 5e-6 |A| or stagnation, then FP64 to p.tol.
// Problems with n >= 2^17 fill the GPU alone: they are solved one after another with the single-problem solver
// (eigs_mixed in mixed precision, with p.inner_max correction-equation steps), the rest in the lockstep batch.
void eigs_batched(MatVecOp *const *ops, int k, const EigsParams &p, const void *const *x0s, void *const *evecs,
                  std::vector<BatchEigsResult> &res, hipStream_t s, bool mixed = false);

// Continuous batching of eigensolve requests: a request joins the running batch between steps once it has
// arrived (arrival_s, seconds after the call) and its working memory fits in the budget (and the batch has
// fewer than max_batch problems, when max_batch > 0); it leaves when converged.
struct EigRequest {
    MatVecOp   *op   = nullptr;
    const void *x0   = nullptr;
    void       *evec = nullptr;
    double      arrival_s = 0.0;
};
struct EigRequestResult {
    BatchEigsResult res;
    double          arrival_s = 0, start_s = 0, finish_s = 0, latency_s = 0;
    int64_t         footprint_bytes = 0;
};
struct ServeSummary {
    double  makespan_s          = 0;
    int64_t peak_admitted_bytes = 0;
    int     peak_concurrency    = 0;
    int64_t steps               = 0;
};
void serve_eigs(const std::vector<EigRequest> &req, const EigsParams &p, int64_t budget, int max_batch,
                std::vector<EigRequestResult> &out, ServeSummary &sum, hipStream_t s, bool mixed = false);

// Eigensolve batcher shared by concurrent callers (the sweep workers of a serving loop): solve() blocks the
// calling thread while its problem runs, on the service's own engine thread and stream, in one batch with the
// problems of the other callers; problems join between steps and leave as soon as they converge. The operator
// must not change until solve() returns; x0 must be complete on the device when solve() is called (the caller
// has synchronised the stream that wrote it); evec is complete on the device when solve() returns.
class EigService {
  public:
    EigService(dmrg_dtype t, const EigsParams &p);
    ~EigService();
    EigService(const EigService &)            = delete;
    EigService &operator=(const EigService &) = delete;
    BatchEigsResult solve(MatVecOp *op, const void *x0, void *evec);
    struct Stats {
        int64_t steps = 0, rebuilds = 0, solves = 0;
        int     peak_batch = 0;
        double  t_rebuild = 0, t_rr = 0, t_step = 0; // engine-thread seconds: rebuilds, Rayleigh-Ritz, whole steps
    };
    Stats stats() const;
    struct Impl;

  private:
    std::unique_ptr<Impl> impl_;
};

} // namespace dmrg

// Finite-chain DMRG sweep driver (ground state of the random transverse-field Ising chain).
#pragma once
#include "common.hpp"

#include <complex>
#include <vector>

namespace dmrg {
class EigService;
// svc: optional shared eigensolve batcher; the sweep's iterative FP64 energy solves then run in its batch.
void fdmrg_ising(const dmrg_fdmrg_config &cfg, dmrg_fdmrg_result &res, hipStream_t s, EigService *svc = nullptr);
// MPO^2 of one site MPO (cx64, dims d[4]) -> out dims (MpoSite.cpp:57-69).
std::vector<std::complex<double>> mpo_squared(const std::vector<std::complex<double>> &W, const int64_t d[4], int64_t out[4]);
void    serve(const dmrg_serve_config &cfg, const dmrg_request *req, int nreq, dmrg_request_result *out, dmrg_serve_stats &st);
int64_t request_footprint(int L, int chi, bool cplx);
} // namespace dmrg

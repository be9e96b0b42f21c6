// Finite-chain DMRG sweep driver (ground state of the random transverse-field Ising chain).
#pragma once
#include "common.hpp"

#include <complex>
#include <vector>

namespace dmrg {
void fdmrg_ising(const dmrg_fdmrg_config &cfg, dmrg_fdmrg_result &res, hipStream_t s);
// MPO^2 of one site MPO (cx64, dims d[4]) -> out dims (MpoSite.cpp:57-69).
std::vector<std::complex<double>> mpo_squared(const std::vector<std::complex<double>> &W, const int64_t d[4], int64_t out[4]);
} // namespace dmrg

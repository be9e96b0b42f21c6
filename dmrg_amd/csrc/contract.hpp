// Contraction plans of the H_eff hot path, each expressed as a short chain of
// generic MFMA GEMMs over the reference's column-major layouts.
#pragma once
#include "common.hpp"

namespace dmrg {

// Problem dimensions: theta (d, chiL, chiR), envL (chiL, chiL, DL), envR (chiR, chiR, DR), mpo (DL, DR, d, d).
struct Dims {
    int64_t d = 0, chiL = 0, chiR = 0, DL = 0, DR = 0;
    int64_t n() const { return d * chiL * chiR; }
    // Intermediate sizes (elements per vector): U = theta.R (d,chiL,chiR,DR), V = W.U (chiL,DL,chiR,d).
    int64_t u_elems() const { return d * chiL * chiR * DR; }
    int64_t v_elems() const { return chiL * DL * chiR * d; }
    // Multiply-accumulates of one matvec / env update (the three legs).
    double macs() const {
        return static_cast<double>(d) * chiL * chiL * chiR * DL + static_cast<double>(chiL) * chiR * DL * DR * d * d +
               static_cast<double>(d) * chiL * chiR * chiR * DR;
    }
};

// Checks of contraction_tblis.cpp:230-234 (matvec argument dims).
Dims check_matvec_dims(const int64_t mps[3], const int64_t mpo[4], const int64_t envL[3], const int64_t envR[3]);

// Workspace elements (of the operator dtype) one matvec needs per vector.
int64_t matvec_ws_elems(const Dims &D);

// out[:, v] = H_eff theta[:, v] for v < nvec (vectors at stride ldx / ldy elements).
// ws must hold nvec * matvec_ws_elems(D) elements of dtype t.
// If ev is non-null, ev[0..4] are recorded on s before leg 1, after legs 1 and 2, after the main kernel of
// leg 3 and after leg 3 (its split-K reduce, when there is one).
// The three GEMMs of one matvec (legs 1-3 in launch order), as matvec launches them.
// partials = false: ws holds only the two intermediates (nvec * matvec_ws_elems_nopart elements) and no leg is
// split over K.
void matvec_descs(dmrg_dtype t, const Dims &D, const void *theta, int64_t ldx, void *out, int64_t ldy, int nvec,
                  const void *mpo, const void *envL, const void *envR, void *ws, GemmDesc leg[3], bool partials = true);
void matvec(dmrg_dtype t, const Dims &D, const void *theta, int64_t ldx, void *out, int64_t ldy, int nvec,
            const void *mpo, const void *envL, const void *envR, void *ws, hipStream_t s, hipEvent_t *ev = nullptr,
            bool partials = true);
int64_t matvec_ws_elems_nopart(const Dims &D);

// One MPO of a chain applied one at a time: device array (DL, DR, d, d) in the matvec's dtype.
struct ChainMpo {
    const void *W  = nullptr;
    int64_t     DL = 0, DR = 0, d = 0;
};
// MPO-chain matvec (contraction_tblis.cpp:307-420): D.d is the product of the chain's physical dims, D.DL the
// first MPO's left bond and D.DR the last one's right bond. ws holds nvec * chain_ws_elems elements.
double  chain_macs(const Dims &D, int nmpo, const ChainMpo *mpos);
int64_t chain_ws_elems(const Dims &D, int nmpo, const ChainMpo *mpos);
void    matvec_chain(dmrg_dtype t, const Dims &D, int nmpo, const ChainMpo *mpos, const void *theta, int64_t ldx, void *out,
                     int64_t ldy, int nvec, const void *envL, const void *envR, void *ws, hipStream_t s);

// Environment update (EnvBase::build_block). For side 'L' the output is (chiR, chiR, DR); for 'R' (chiL, chiL, DL).
int64_t env_ws_elems(const Dims &D);
void    env_build_block(dmrg_dtype t, char side, const Dims &D, void *out, const void *other, const void *theta,
                        const void *mpo, void *ws, hipStream_t s);

// Subspace-expansion term without the factor alpha (EnvBase.cpp:303-361); ws holds d*chiL*chiR*D elements,
// D = DL for side 'L' and DR for side 'R'.
void env_expansion_term(dmrg_dtype t, char side, const Dims &D, void *P, const void *block, const void *mps, const void *mpo,
                        void *ws, hipStream_t s);

// Grow-only device workspace shared by the stateless entry points (one per device).
void *scratch_workspace(size_t bytes);

// BLAS-1 style reductions (deterministic two-pass).
void dot_device(dmrg_dtype t, int64_t n, const void *x, const void *y, double *d_result2, hipStream_t s);
void nrm2sq_device(dmrg_dtype t, int64_t n, const void *x, double *d_result, hipStream_t s);

} // namespace dmrg

namespace dmrg {
void convert_f64_to_f32(int64_t nreal, const double *src, float *dst, hipStream_t s);
void convert_f32_to_f64(int64_t nreal, const float *src, double *dst, hipStream_t s);
// ws must hold chiL*DR*d elements; out holds d*chiL*chiR elements (operator dtype, FP64/CX64).
void heff_diagonal(dmrg_dtype t, const Dims &D, const void *mpo, const void *envL, const void *envR, void *ws, void *out,
                   hipStream_t s);
constexpr int DOT_BUFFER_DOUBLES = 2 + 2 * 512;
} // namespace dmrg

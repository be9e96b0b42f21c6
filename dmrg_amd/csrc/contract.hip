// Contraction plans for the H_eff matvec and the environment updates.
//
// Each contraction is three GEMMs over the reference's column-major tensors:
//   two O(chi^3 d D) legs that run on the matrix cores, and a middle "W leg" that
//   contracts the MPO over (physical, MPO-bond) pairs (K = d*D, memory bound).
// Index maps are written out per leg; m/k/n name the GEMM dimensions and the
// tensor index each one fuses.
#include "contract.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace dmrg {

Dims check_matvec_dims(const int64_t mps[3], const int64_t mpo[4], const int64_t envL[3], const int64_t envR[3]) {
    // contraction_tblis.cpp:230-234
    DMRG_REQUIRE(mps[1] == envL[0], DMRG_ERR_DIMS, "Dimension mismatch mps and envL");
    DMRG_REQUIRE(mps[2] == envR[0], DMRG_ERR_DIMS, "Dimension mismatch mps and envR");
    DMRG_REQUIRE(mps[0] == mpo[2], DMRG_ERR_DIMS, "Dimension mismatch mps and mpo");
    DMRG_REQUIRE(envL[2] == mpo[0], DMRG_ERR_DIMS, "Dimension mismatch envL and mpo");
    DMRG_REQUIRE(envR[2] == mpo[1], DMRG_ERR_DIMS, "Dimension mismatch envR and mpo");
    DMRG_REQUIRE(envL[0] == envL[1] && envR[0] == envR[1], DMRG_ERR_DIMS, "env blocks must be square in the bond indices");
    DMRG_REQUIRE(mpo[2] == mpo[3], DMRG_ERR_DIMS, "mpo must have equal physical dims");
    for(int i = 0; i < 3; ++i) DMRG_REQUIRE(mps[i] > 0 && envL[i] > 0 && envR[i] > 0, DMRG_ERR_DIMS, "zero-sized dimension");
    Dims D;
    D.d    = mps[0];
    D.chiL = mps[1];
    D.chiR = mps[2];
    D.DL   = mpo[0];
    D.DR   = mpo[1];
    DMRG_REQUIRE(D.DL > 0 && D.DR > 0, DMRG_ERR_DIMS, "zero-sized mpo");
    return D;
}

// Split-K slices the final leg may use (its partial slabs live in the workspace).
constexpr int64_t MAX_KSPLIT = 8;

int64_t matvec_ws_elems(const Dims &D) { return D.d * D.chiL * D.chiR * (D.DL + D.DR) + MAX_KSPLIT * D.n(); }
int64_t matvec_ws_elems_nopart(const Dims &D) { return D.d * D.chiL * D.chiR * (D.DL + D.DR); }

static GemmDesc make(dmrg_dtype t, int64_t M, int64_t N, int64_t K, int nvec) {
    DMRG_REQUIRE(M < (int64_t(1) << 31) && N < (int64_t(1) << 31) && K < (int64_t(1) << 31), DMRG_ERR_DIMS,
                 "contraction too large");
    GemmDesc g;
    g.dtype = t;
    g.M     = static_cast<int32_t>(M);
    g.N     = static_cast<int32_t>(N);
    g.K     = static_cast<int32_t>(K);
    g.batch = nvec;
    return g;
}

void matvec_descs(dmrg_dtype t, const Dims &D, const void *theta, int64_t ldx, void *out, int64_t ldy, int nvec,
                  const void *mpo, const void *envL, const void *envR, void *ws, GemmDesc leg[3], bool partials) {
    const int64_t d = D.d, cL = D.chiL, cR = D.chiR, DL = D.DL, DR = D.DR;
    const size_t  es  = dtype_size(t);
    char         *w0  = static_cast<char *>(ws);
    const int64_t nT1 = (cL >= cR) ? d * cR * cL * DL : d * cL * cR * DR; // first intermediate per vector
    const int64_t nT2 = (cL >= cR) ? cR * DR * d * cL : cL * DL * cR * d; // second intermediate per vector
    void         *T1  = w0;
    void         *T2  = w0 + static_cast<size_t>(nvec) * nT1 * es;
    void         *PT  = partials ? w0 + static_cast<size_t>(nvec) * (nT1 + nT2) * es : nullptr; // split-K partials
    const int64_t PC  = partials ? MAX_KSPLIT * D.n() * nvec : 0;                            // (leg 1, then leg 3)

    if(cL >= cR) {
        // Order of contraction_tblis.cpp:279-283 (theta.L first).
        // Leg 1: T1[s,b,a',w] = sum_a theta[s,a,b] L[a,a',w]           m=(s,b) k=a n=(a',w)
        GemmDesc g1 = make(t, d * cR, cL * DL, cL, nvec);
        g1.A = theta, g1.am = fused(1, d * cL, d), g1.ak = plain(d), g1.a_batch = ldx;
        g1.B = envL, g1.bk = plain(1), g1.bn = plain(cL);
        g1.C = T1, g1.cm = plain(1), g1.cn = plain(d * cR), g1.c_batch = nT1;
        g1.part = PT, g1.part_capacity = PC; // leg 3's partial slabs are free until leg 3
        leg[0] = g1;
        // Leg 2: T2[b,w',s',a'] = sum_{s,w} W[w,w',s,s'] T1[s,b,a',w]  m=(w',s') k=(s,w) n=(b,a')
        GemmDesc g2 = make(t, DR * d, cR * cL, d * DL, nvec);
        g2.A = mpo, g2.am = fused(DL, DL * DR * d, DR), g2.ak = fused(DL * DR, 1, d);
        g2.B = T1, g2.bk = fused(1, d * cR * cL, d), g2.bn = fused(d, d * cR, cR), g2.b_batch = nT1;
        g2.C = T2, g2.cm = fused(cR, cR * DR, DR), g2.cn = fused(1, cR * DR * d, cR), g2.c_batch = nT2;
        leg[1] = g2;
        // Leg 3: out[s',a',b'] = sum_{b,w'} T2[b,w',s',a'] R[b,b',w']   m=(s',a') k=(b,w') n=b'
        GemmDesc g3 = make(t, d * cL, cR, cR * DR, nvec);
        g3.A = T2, g3.am = plain(cR * DR), g3.ak = plain(1), g3.a_batch = nT2;
        g3.B = envR, g3.bk = fused(1, cR * cR, cR), g3.bn = plain(cR);
        g3.C = out, g3.cm = plain(1), g3.cn = plain(d * cL), g3.c_batch = ldy;
        g3.part = PT, g3.part_capacity = PC;
        leg[2] = g3;
    } else {
        // Order of contraction_tblis.cpp:285-289 (theta.R first).
        // Leg 1: U[s,a,b',w'] = sum_b theta[s,a,b] R[b,b',w']            m=(s,a) k=b n=(b',w')
        GemmDesc g1 = make(t, d * cL, cR * DR, cR, nvec);
        g1.A = theta, g1.am = plain(1), g1.ak = plain(d * cL), g1.a_batch = ldx;
        g1.B = envR, g1.bk = plain(1), g1.bn = plain(cR);
        g1.C = T1, g1.cm = plain(1), g1.cn = plain(d * cL), g1.c_batch = nT1;
        g1.part = PT, g1.part_capacity = PC;
        leg[0] = g1;
        // Leg 2: V[a,w,b',s'] = sum_{s,w'} W[w,w',s,s'] U[s,a,b',w']    m=(w,s') k=(s,w') n=(a,b')
        GemmDesc g2 = make(t, DL * d, cL * cR, d * DR, nvec);
        g2.A = mpo, g2.am = fused(1, DL * DR * d, DL), g2.ak = fused(DL * DR, DL, d);
        g2.B = T1, g2.bk = fused(1, d * cL * cR, d), g2.bn = fused(d, d * cL, cL), g2.b_batch = nT1;
        g2.C = T2, g2.cm = fused(cL, cL * DL * cR, DL), g2.cn = fused(1, cL * DL, cL), g2.c_batch = nT2;
        leg[1] = g2;
        // Leg 3: out[s',a',b'] = sum_{a,w} L[a,a',w] V[a,w,b',s']       m=a' k=(a,w) n=(b',s')
        GemmDesc g3 = make(t, cL, cR * d, cL * DL, nvec);
        g3.A = envL, g3.am = plain(cL), g3.ak = fused(1, cL * cL, cL);
        g3.B = T2, g3.bk = plain(1), g3.bn = plain(cL * DL), g3.b_batch = nT2;
        g3.C = out, g3.cm = plain(d), g3.cn = fused(d * cL, 1, cR), g3.c_batch = ldy;
        g3.part = PT, g3.part_capacity = PC;
        leg[2] = g3;
    }
}

void matvec(dmrg_dtype t, const Dims &D, const void *theta, int64_t ldx, void *out, int64_t ldy, int nvec,
            const void *mpo, const void *envL, const void *envR, void *ws, hipStream_t s, hipEvent_t *ev, bool partials) {
    GemmDesc leg[3];
    matvec_descs(t, D, theta, ldx, out, ldy, nvec, mpo, envL, envR, ws, leg, partials);
    auto mark = [&](int i) {
        if(ev) DMRG_HIP_CHECK(hipEventRecord(ev[i], s));
    };
    mark(0);
    gemm(leg[0], s);
    mark(1);
    gemm(leg[1], s);
    mark(2);
    if(ev) leg[2].ev_after_main = ev[3];
    gemm(leg[2], s);
    mark(4);
}

// ------------------------------------------------------------------------------------------
// MPO-chain matvec (more than 5 MPOs, contraction_tblis.cpp:307-420): theta.L, then one MPO at a time over its
// physical index and its left bond, then .R. The intermediates keep one layout, X[s, b, a', w] with the fused
// physical index s = (s_0, ..., s_{n-1}) (site 0 fastest); MPO i turns s_i into s_i' and w_i into w_{i+1}, so the
// sites before i hold outputs and the sites after it inputs, as the reference's reshapes do (:380-406).

double chain_macs(const Dims &D, int nmpo, const ChainMpo *mpos) {
    double m = static_cast<double>(D.d) * D.chiL * D.chiL * D.chiR * D.DL + static_cast<double>(D.d) * D.chiL * D.chiR * D.chiR * D.DR;
    for(int i = 0; i < nmpo; ++i) m += static_cast<double>(D.d) * mpos[i].d * D.chiL * D.chiR * mpos[i].DL * mpos[i].DR;
    return m;
}

int64_t chain_ws_elems(const Dims &D, int nmpo, const ChainMpo *mpos) {
    int64_t maxD = std::max(D.DL, D.DR);
    for(int i = 0; i < nmpo; ++i) maxD = std::max({maxD, mpos[i].DL, mpos[i].DR});
    return 2 * D.d * D.chiL * D.chiR * maxD + MAX_KSPLIT * D.n();
}

void matvec_chain(dmrg_dtype t, const Dims &D, int nmpo, const ChainMpo *mpos, const void *theta, int64_t ldx, void *out,
                  int64_t ldy, int nvec, const void *envL, const void *envR, void *ws, hipStream_t s) {
    const int64_t d = D.d, cL = D.chiL, cR = D.chiR;
    int64_t       maxD = std::max(D.DL, D.DR);
    for(int i = 0; i < nmpo; ++i) maxD = std::max({maxD, mpos[i].DL, mpos[i].DR});
    const size_t  es = dtype_size(t);
    const int64_t nX = d * cR * cL * maxD; // one intermediate per vector
    char         *w0 = static_cast<char *>(ws);
    void         *X[2] = {w0, w0 + static_cast<size_t>(nvec) * nX * es};
    void         *PT   = w0 + static_cast<size_t>(2 * nvec) * nX * es;
    const int64_t PC   = MAX_KSPLIT * D.n() * nvec;
    // theta.L (contraction_tblis.cpp:358): X0[s,b,a',w] = sum_a theta[s,a,b] L[a,a',w]     m=(s,b) k=a n=(a',w)
    GemmDesc g1 = make(t, d * cR, cL * D.DL, cL, nvec);
    g1.A = theta, g1.am = fused(1, d * cL, d), g1.ak = plain(d), g1.a_batch = ldx;
    g1.B = envL, g1.bk = plain(1), g1.bn = plain(cL);
    g1.C = X[0], g1.cm = plain(1), g1.cn = plain(d * cR), g1.c_batch = nX;
    g1.part = PT, g1.part_capacity = PC;
    gemm(g1, s);
    // MPO i (:380-406): X'[pre,s_i',post,b,a',w'] = sum_{s_i,w} W_i[w,w',s_i,s_i'] X[pre,s_i,post,b,a',w]
    //   m=(s_i',w') k=(s_i,w) n=(pre,(post,b,a')); pre spans the P = d_0...d_{i-1} outputs already made.
    int64_t P = 1;
    for(int i = 0; i < nmpo; ++i) {
        const ChainMpo &w  = mpos[i];
        const int64_t   di = w.d, DLi = w.DL, DRi = w.DR;
        GemmDesc        g  = make(t, di * DRi, (d / di) * cR * cL, di * DLi, nvec);
        g.A = w.W, g.am = fused(DLi * DRi * di, DLi, di), g.ak = fused(DLi * DRi, 1, di);
        g.B = X[i & 1], g.bk = fused(P, d * cR * cL, di), g.bn = fused(1, P * di, P), g.b_batch = nX;
        g.C = X[(i + 1) & 1], g.cm = fused(P, d * cR * cL, di), g.cn = fused(1, P * di, P), g.c_batch = nX;
        gemm(g, s);
        P *= di;
    }
    // .R (:417-419): out[s',a',b'] = sum_{b,w'} X[s',b,a',w'] R[b,b',w']     m=(s',a') k=(b,w') n=b'
    GemmDesc g3 = make(t, d * cL, cR, cR * D.DR, nvec);
    g3.A = X[nmpo & 1], g3.am = fused(1, d * cR, d), g3.ak = fused(d, d * cR * cL, cR), g3.a_batch = nX;
    g3.B = envR, g3.bk = fused(1, cR * cR, cR), g3.bn = plain(cR);
    g3.C = out, g3.cm = plain(1), g3.cn = plain(d * cL), g3.c_batch = ldy;
    g3.part = PT, g3.part_capacity = PC;
    gemm(g3, s);
}

int64_t env_ws_elems(const Dims &D) {
    return D.d * D.chiL * D.chiR * (D.DL + D.DR) + MAX_KSPLIT * std::max(D.chiR * D.chiR * D.DR, D.chiL * D.chiL * D.DL);
}

void env_build_block(dmrg_dtype t, char side, const Dims &D, void *out, const void *other, const void *theta,
                     const void *mpo, void *ws, hipStream_t s) {
    const int64_t d = D.d, cL = D.chiL, cR = D.chiR, DL = D.DL, DR = D.DR;
    const size_t  es = dtype_size(t);
    char         *w0 = static_cast<char *>(ws);
    if(side == 'L') {
        // L'[b,b',w'] = sum theta[s,a,b] L[a,a',w] W[w,w',s,s'] conj theta[s',a',b']   (EnvBase.cpp:109-112)
        void *X = w0;                                       // X[a,w,s',b']  (chiL*DL) x (d*chiR)
        void *Y = w0 + static_cast<size_t>(cL * DL * d * cR) * es; // Y[s,a,b',w']  (d*chiL) x (chiR*DR)
        // X[(a,w),(s',b')] = sum_a' L[a,a',w] conj theta[s',a',b']       m=(a,w) k=a' n=(s',b')
        GemmDesc e1 = make(t, cL * DL, d * cR, cL, 1);
        e1.A = other, e1.am = fused(1, cL * cL, cL), e1.ak = plain(cL);
        e1.B = theta, e1.bk = plain(d), e1.bn = fused(1, d * cL, d), e1.conjB = 1;
        e1.C = X, e1.cm = plain(1), e1.cn = plain(cL * DL);
        gemm(e1, s);
        // Y[(s,a),(b',w')] = sum_{s',w} W[w,w',s,s'] X[a,w,s',b']        m=(s,w') k=(s',w) n=(a,b')
        GemmDesc e2 = make(t, d * DR, cL * cR, d * DL, 1);
        e2.A = mpo, e2.am = fused(DL * DR, DL, d), e2.ak = fused(DL * DR * d, 1, d);
        e2.B = X, e2.bk = fused(cL * DL, cL, d), e2.bn = fused(1, cL * DL * d, cL);
        e2.C = Y, e2.cm = fused(1, d * cL * cR, d), e2.cn = fused(d, d * cL, cL);
        gemm(e2, s);
        // L'[b,(b',w')] = sum_{(s,a)} theta[s,a,b] Y[(s,a),(b',w')]        m=b k=(s,a) n=(b',w')
        GemmDesc e3 = make(t, cR, cR * DR, d * cL, 1);
        e3.A = theta, e3.am = plain(d * cL), e3.ak = plain(1);
        e3.B = Y, e3.bk = plain(1), e3.bn = plain(d * cL);
        e3.C = out, e3.cm = plain(1), e3.cn = plain(cR);
        e3.part = w0 + static_cast<size_t>(d * cL * cR * (DL + DR)) * es, e3.part_capacity = MAX_KSPLIT * cR * cR * DR;
        gemm(e3, s);
    } else if(side == 'R') {
        // R'[a,a',w] = sum theta[s,a,b] R[b,b',w'] W[w,w',s,s'] conj theta[s',a',b']   (EnvBase.cpp:146-149)
        void *X = w0;                                       // X[s',a',b,w']  (d*chiL) x (chiR*DR)
        void *Y = w0 + static_cast<size_t>(d * cL * cR * DR) * es; // Y[s,b,a',w]   (d*chiR) x (chiL*DL)
        // X[(s',a'),(b,w')] = sum_b' conj theta[s',a',b'] R[b,b',w']      m=(s',a') k=b' n=(b,w')
        GemmDesc e1 = make(t, d * cL, cR * DR, cR, 1);
        e1.A = theta, e1.am = plain(1), e1.ak = plain(d * cL), e1.conjA = 1;
        e1.B = other, e1.bk = plain(cR), e1.bn = fused(1, cR * cR, cR);
        e1.C = X, e1.cm = plain(1), e1.cn = plain(d * cL);
        gemm(e1, s);
        // Y[(s,b),(a',w)] = sum_{s',w'} W[w,w',s,s'] X[s',a',b,w']        m=(s,w) k=(s',w') n=(b,a')
        GemmDesc e2 = make(t, d * DL, cR * cL, d * DR, 1);
        e2.A = mpo, e2.am = fused(DL * DR, 1, d), e2.ak = fused(DL * DR * d, DL, d);
        e2.B = X, e2.bk = fused(1, d * cL * cR, d), e2.bn = fused(d * cL, d, cR);
        e2.C = Y, e2.cm = fused(1, d * cR * cL, d), e2.cn = fused(d, d * cR, cR);
        gemm(e2, s);
        // R'[a,(a',w)] = sum_{(s,b)} theta[s,a,b] Y[(s,b),(a',w)]          m=a k=(s,b) n=(a',w)
        GemmDesc e3 = make(t, cL, cL * DL, d * cR, 1);
        e3.A = theta, e3.am = plain(d), e3.ak = fused(1, d * cL, d);
        e3.B = Y, e3.bk = plain(1), e3.bn = plain(d * cR);
        e3.C = out, e3.cm = plain(1), e3.cn = plain(cL);
        e3.part = w0 + static_cast<size_t>(d * cL * cR * (DL + DR)) * es, e3.part_capacity = MAX_KSPLIT * cL * cL * DL;
        gemm(e3, s);
    } else {
        throw Error(DMRG_ERR_ARG, "env side must be 'L' or 'R'");
    }
}

// Expansion term of subspace expansion (EnvBase.cpp:303-361), alpha applied by the caller:
//   side 'L': P[s',a',b,wR] = sum_{a,w,s} block[a,a',w] mps[s,a,b] W[w,wR,s,s'], P dims (d, chiL, chiR*DR);
//   side 'R': the reference forms R2[b',a,wL,s'] = sum_{b,w,s} block[b,b',w] mps[s,a,b] W[wL,w,s,s'],
//             shuffles it to [s',b',a,wL] and reshapes that buffer to (d, chiL*DL, chiR) as it lies, so P's
//             storage offset is s' + d(b' + chiR(a + chiL wL)).
void env_expansion_term(dmrg_dtype t, char side, const Dims &D, void *P, const void *block, const void *mps, const void *mpo,
                        void *ws, hipStream_t s) {
    const int64_t d = D.d, cL = D.chiL, cR = D.chiR, DL = D.DL, DR = D.DR;
    if(side == 'L') {
        // X[(a',w),(s,b)] = sum_a block[a,a',w] mps[s,a,b]          m=(a',w) k=a n=(s,b)
        GemmDesc e1 = make(t, cL * DL, d * cR, cL, 1);
        e1.A = block, e1.am = fused(cL, cL * cL, cL), e1.ak = plain(1);
        e1.B = mps, e1.bk = plain(d), e1.bn = fused(1, d * cL, d);
        e1.C = ws, e1.cm = plain(1), e1.cn = plain(cL * DL);
        gemm(e1, s);
        // P[s',a',b,wR] = sum_{w,s} W[w,wR,s,s'] X[(a',w),(s,b)]     m=(s',wR) k=(w,s) n=(a',b)
        GemmDesc e2 = make(t, d * DR, cL * cR, DL * d, 1);
        e2.A = mpo, e2.am = fused(DL * DR * d, DL, d), e2.ak = fused(1, DL * DR, DL);
        e2.B = ws, e2.bk = fused(cL, cL * DL, DL), e2.bn = fused(1, cL * DL * d, cL);
        e2.C = P, e2.cm = fused(1, d * cL * cR, d), e2.cn = fused(d, d * cL, cL);
        gemm(e2, s);
    } else if(side == 'R') {
        // X[(b',w),(s,a)] = sum_b block[b,b',w] mps[s,a,b]          m=(b',w) k=b n=(s,a)
        GemmDesc e1 = make(t, cR * DR, d * cL, cR, 1);
        e1.A = block, e1.am = fused(cR, cR * cR, cR), e1.ak = plain(1);
        e1.B = mps, e1.bk = plain(d * cL), e1.bn = plain(1);
        e1.C = ws, e1.cm = plain(1), e1.cn = plain(cR * DR);
        gemm(e1, s);
        // P at s' + d(b' + chiR(a + chiL wL)) = sum_{w,s} W[wL,w,s,s'] X[(b',w),(s,a)]   m=(s',wL) k=(w,s) n=(b',a)
        GemmDesc e2 = make(t, d * DL, cR * cL, DR * d, 1);
        e2.A = mpo, e2.am = fused(DL * DR * d, 1, d), e2.ak = fused(DL, DL * DR, DR);
        e2.B = ws, e2.bk = fused(cR, cR * DR, DR), e2.bn = fused(1, cR * DR * d, cR);
        e2.C = P, e2.cm = fused(1, d * cR * cL, d), e2.cn = fused(d, d * cR, cR);
        gemm(e2, s);
    } else {
        throw Error(DMRG_ERR_ARG, "env side must be 'L' or 'R'");
    }
}

// ------------------------------------------------------------------------------------------
// Stateless workspace

void *scratch_workspace(size_t bytes) {
    static std::mutex          mu;
    static std::vector<void *> bufs;
    static std::vector<size_t> sizes;
    std::lock_guard<std::mutex> lock(mu);
    int                         dev = 0;
    DMRG_HIP_CHECK(hipGetDevice(&dev));
    if(static_cast<int>(bufs.size()) <= dev) {
        bufs.resize(dev + 1, nullptr);
        sizes.resize(dev + 1, 0);
    }
    if(sizes[dev] < bytes) {
        if(bufs[dev]) {
            // Other streams may still be reading the old buffer.
            DMRG_HIP_CHECK(hipDeviceSynchronize());
            DMRG_HIP_CHECK(hipFree(bufs[dev]));
            bufs[dev]  = nullptr;
            sizes[dev] = 0;
        }
        void *p = nullptr;
        if(hipMalloc(&p, bytes) != hipSuccess) throw Error(DMRG_ERR_OOM, "workspace allocation failed");
        bufs[dev]  = p;
        sizes[dev] = bytes;
    }
    return bufs[dev];
}

} // namespace dmrg

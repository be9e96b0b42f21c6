// Generic batched GEMM on CDNA4 matrix cores (gfx950), FP64 and FP32, real and complex.
//
//   C[b][m,n] (+)= sum_k op(A)[b][m,k] * op(B)[b][k,n]
//
// Every operand index maps to memory through a two-level stride (Idx2), so the
// column-major tensors of the H_eff contraction (theta (d,chiL,chiR), env (chi,chi,D),
// MPO (DL,DR,d,d)) are read and written in place: no transposes or reshapes are
// materialised between the legs of a contraction.
//
// Kernel structure: a workgroup owns a BM x BN tile of C and loops over K in
// steps of BK. Tiles of A and B go global -> registers -> LDS (double buffered,
// one barrier per K step); the next tile's global loads are in flight while the
// waves run MFMA on the current one. Each wave owns a (BM/WM) x (BN/WN) sub-tile
// built from 16x16 blocks:
//   FP64: v_mfma_f64_16x16x4_f64,  FP32: v_mfma_f32_16x16x4_f32.
// Complex products use the 4-real-product form (4 MFMAs per block per k-step),
// which keeps FP64 rounding at the level of a real GEMM.
// The "swap" variant feeds B as the MFMA's first operand so that the lane index
// of the accumulator runs along M; it is picked when C is unit-stride in M, so
// the epilogue stores are coalesced either way.
#include "common.hpp"

#include <algorithm>
#include <mutex>
#include <type_traits>

namespace dmrg {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float  v4f __attribute__((ext_vector_type(4)));

template <typename R>
struct mfma_t;
template <>
struct mfma_t<double> {
    using acc                                  = v4d;
    __device__ static inline acc op(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct mfma_t<float> {
    using acc                                = v4f;
    __device__ static inline acc op(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

template <typename R, bool CPLX>
using elem_t = std::conditional_t<CPLX, cpx<R>, R>;

// LDS row padding (in elements) that makes the MFMA fragment reads conflict free:
// a fragment read puts 16 consecutive elements of one row on each group of 16 lanes;
// the groups that the LDS serves in one pass must land on disjoint bank ranges.
constexpr int lds_pad(int row_elems, int elem_bytes) {
    const int seg_bytes = 16 * elem_bytes;                 // bytes one 16-lane group reads
    const int groups    = seg_bytes >= 256 ? 1 : 256 / seg_bytes; // groups per 256-byte pass
    for(int p = 0; p <= 64; ++p) {
        const int row_bytes = (row_elems + p) * elem_bytes;
        if(row_bytes % 16 != 0 && elem_bytes >= 16) continue;
        bool ok = true;
        for(int g = 0; g < groups && ok; ++g)
            for(int h = 0; h < g && ok; ++h) {
                const int og = (g * row_bytes) % 256, oh = (h * row_bytes) % 256;
                const int d  = og > oh ? og - oh : oh - og;
                if(d < seg_bytes || 256 - d < seg_bytes) ok = false;
            }
        if(ok) return p;
    }
    return 0;
}

template <typename E>
__device__ inline E load_elem(const E *base, int64_t off, bool valid, bool conj) {
    if constexpr(std::is_same_v<E, cpx<double>> || std::is_same_v<E, cpx<float>>) {
        E v{0, 0};
        if(valid) {
            v = base[off];
            if(conj) v.im = -v.im;
        }
        return v;
    } else {
        return valid ? base[off] : E(0);
    }
}

template <typename R, bool CPLX, int BM, int BN, int BK, int WM, int WN, bool SWAP>
__global__ __launch_bounds__(64 * WM *WN) void gemm_kernel(const GemmDesc g) {
    using E               = elem_t<R, CPLX>;
    using MF              = mfma_t<R>;
    using acc_t           = typename MF::acc;
    constexpr int NT      = 64 * WM * WN;
    constexpr int TM      = BM / (16 * WM);
    constexpr int TN      = BN / (16 * WN);
    constexpr int LA      = BM * BK / NT;
    constexpr int LB      = BK * BN / NT;
    constexpr int PA      = lds_pad(BM, sizeof(E));
    constexpr int PB      = lds_pad(BN, sizeof(E));
    static_assert(TM >= 1 && TN >= 1, "wave tile too small");
    static_assert(LA >= 1 && LB >= 1 && (BM * BK) % NT == 0 && (BK * BN) % NT == 0, "bad load split");
    static_assert(BK % 4 == 0, "BK must be a multiple of the MFMA k");

    __shared__ E As[2][BK][BM + PA];
    __shared__ E Bs[2][BK][BN + PB];

    const int tid  = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm   = wave % WM;
    const int wn   = wave / WM;

    // Tile index. Workgroups are dealt round-robin over the 8 XCDs; remap so that each
    // XCD works a contiguous range of tiles (shared B columns stay in one L2).
    const int tilesM = (g.M + BM - 1) / BM;
    const int tilesN = (g.N + BN - 1) / BN;
    const int total  = tilesM * tilesN;
    int       bid    = blockIdx.x;
    if(g.xcd_remap && (total % 8) == 0) bid = (bid % 8) * (total / 8) + bid / 8;
    const int mt = bid % tilesM;
    const int nt = bid / tilesM;
    const int m0 = mt * BM;
    const int n0 = nt * BN;
    const int b  = blockIdx.z;

    const E *A = static_cast<const E *>(g.A) + b * g.a_batch;
    const E *B = static_cast<const E *>(g.B) + b * g.b_batch;
    E       *C = static_cast<E *>(g.C) + b * g.c_batch;

    // Per-thread load coordinates and the fixed (non-K) part of their offsets.
    int     a_mm[LA], a_kk[LA], b_kk[LB], b_nn[LB];
    int64_t a_moff[LA], b_noff[LB];
    bool    a_mval[LA], b_nval[LB];
#pragma unroll
    for(int e = 0; e < LA; ++e) {
        const int idx = tid + e * NT;
        if(g.a_fast_m) {
            a_mm[e] = idx % BM;
            a_kk[e] = idx / BM;
        } else {
            a_kk[e] = idx % BK;
            a_mm[e] = idx / BK;
        }
        const int gm = m0 + a_mm[e];
        a_mval[e]    = gm < g.M;
        a_moff[e]    = a_mval[e] ? idx_off(g.am, gm) : 0;
    }
#pragma unroll
    for(int e = 0; e < LB; ++e) {
        const int idx = tid + e * NT;
        if(g.b_fast_k) {
            b_kk[e] = idx % BK;
            b_nn[e] = idx / BK;
        } else {
            b_nn[e] = idx % BN;
            b_kk[e] = idx / BN;
        }
        const int gn = n0 + b_nn[e];
        b_nval[e]    = gn < g.N;
        b_noff[e]    = b_nval[e] ? idx_off(g.bn, gn) : 0;
    }

    E ra[LA], rb[LB];
    auto load_tiles = [&](int k0) {
#pragma unroll
        for(int e = 0; e < LA; ++e) {
            const int  gk = k0 + a_kk[e];
            const bool v  = a_mval[e] && gk < g.K;
            ra[e]         = load_elem(A, a_moff[e] + (v ? idx_off(g.ak, gk) : 0), v, g.conjA);
        }
#pragma unroll
        for(int e = 0; e < LB; ++e) {
            const int  gk = k0 + b_kk[e];
            const bool v  = b_nval[e] && gk < g.K;
            rb[e]         = load_elem(B, b_noff[e] + (v ? idx_off(g.bk, gk) : 0), v, g.conjB);
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for(int e = 0; e < LA; ++e) As[buf][a_kk[e]][a_mm[e]] = ra[e];
#pragma unroll
        for(int e = 0; e < LB; ++e) Bs[buf][b_kk[e]][b_nn[e]] = rb[e];
    };

    acc_t acc_re[TM][TN];
    acc_t acc_im[CPLX ? TM : 1][CPLX ? TN : 1];
#pragma unroll
    for(int i = 0; i < TM; ++i)
#pragma unroll
        for(int j = 0; j < TN; ++j) {
            acc_re[i][j] = acc_t{0, 0, 0, 0};
            if constexpr(CPLX) acc_im[i][j] = acc_t{0, 0, 0, 0};
        }

    const int numK = (g.K + BK - 1) / BK;
    load_tiles(0);
    store_tiles(0);
    __syncthreads();

    const int lrow = lane >> 4; // k row within the 4-deep MFMA step
    const int lcol = lane & 15; // m (or n) within a 16-wide block
    for(int kt = 0; kt < numK; ++kt) {
        const int buf = kt & 1;
        if(kt + 1 < numK) load_tiles((kt + 1) * BK);
#pragma unroll
        for(int ks = 0; ks < BK / 4; ++ks) {
            const int kr = ks * 4 + lrow;
            E         fa[TM], fb[TN];
#pragma unroll
            for(int i = 0; i < TM; ++i) fa[i] = As[buf][kr][(wm * TM + i) * 16 + lcol];
#pragma unroll
            for(int j = 0; j < TN; ++j) fb[j] = Bs[buf][kr][(wn * TN + j) * 16 + lcol];
#pragma unroll
            for(int i = 0; i < TM; ++i)
#pragma unroll
                for(int j = 0; j < TN; ++j) {
                    if constexpr(CPLX) {
                        // (ar + i ai)(br + i bi) = (ar br - ai bi) + i (ar bi + ai br)
                        const R ar = fa[i].re, ai = fa[i].im, br = fb[j].re, bi = fb[j].im;
                        if constexpr(!SWAP) {
                            acc_re[i][j] = MF::op(ar, br, acc_re[i][j]);
                            acc_re[i][j] = MF::op(-ai, bi, acc_re[i][j]);
                            acc_im[i][j] = MF::op(ar, bi, acc_im[i][j]);
                            acc_im[i][j] = MF::op(ai, br, acc_im[i][j]);
                        } else {
                            acc_re[i][j] = MF::op(br, ar, acc_re[i][j]);
                            acc_re[i][j] = MF::op(bi, -ai, acc_re[i][j]);
                            acc_im[i][j] = MF::op(bi, ar, acc_im[i][j]);
                            acc_im[i][j] = MF::op(br, ai, acc_im[i][j]);
                        }
                    } else {
                        if constexpr(!SWAP) acc_re[i][j] = MF::op(fa[i], fb[j], acc_re[i][j]);
                        else acc_re[i][j] = MF::op(fb[j], fa[i], acc_re[i][j]);
                    }
                }
        }
        if(kt + 1 < numK) store_tiles(buf ^ 1);
        __syncthreads();
    }

    // Epilogue. 16x16 accumulator layout (measured on gfx950): lane l, register v holds column l%16 of
    // row 4*(l/16)+v for v_mfma_f32_16x16x4_f32 and of row (l/16)+4*v for v_mfma_f64_16x16x4_f64.
    // Without swap rows run along M; with swap they run along N.
#pragma unroll
    for(int i = 0; i < TM; ++i)
#pragma unroll
        for(int j = 0; j < TN; ++j)
#pragma unroll
            for(int v = 0; v < 4; ++v) {
                const int row = std::is_same_v<R, double> ? (lrow + 4 * v) : (4 * lrow + v);
                int       gm, gn;
                if constexpr(!SWAP) {
                    gm = m0 + (wm * TM + i) * 16 + row;
                    gn = n0 + (wn * TN + j) * 16 + lcol;
                } else {
                    gm = m0 + (wm * TM + i) * 16 + lcol;
                    gn = n0 + (wn * TN + j) * 16 + row;
                }
                if(gm < g.M && gn < g.N) {
                    const int64_t off = idx_off(g.cm, gm) + idx_off(g.cn, gn);
                    if constexpr(CPLX) {
                        E val{acc_re[i][j][v], acc_im[i][j][v]};
                        if(g.accumulate) {
                            const E old = C[off];
                            val.re += old.re;
                            val.im += old.im;
                        }
                        C[off] = val;
                    } else {
                        R val = acc_re[i][j][v];
                        if(g.accumulate) val += C[off];
                        C[off] = val;
                    }
                }
            }
}

// ---------------------------------------------------------------------------------------
// Host launcher

struct CfgShape {
    int BM, BN;
};
static CfgShape cfg_shape(TileCfg c) {
    switch(c) {
        case TileCfg::T64x64: return {64, 64};
        case TileCfg::T32x64: return {32, 64};
        case TileCfg::T64x32: return {64, 32};
        case TileCfg::T16x128: return {16, 128};
        case TileCfg::T128x64: return {128, 64};
        case TileCfg::T32x32: return {32, 32};
        default: return {64, 64};
    }
}

int device_cu_count() {
    static int cus = [] {
        int dev = 0, n = 0;
        if(hipGetDevice(&dev) != hipSuccess) return 256;
        if(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
        return n;
    }();
    return cus;
}

static TileCfg choose_cfg(const GemmDesc &g) {
    if(g.M <= 16) return TileCfg::T16x128;
    const int64_t cus  = device_cu_count();
    auto          nblk = [&](int bm, int bn) { return static_cast<int64_t>((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn) * g.batch; };
    if(nblk(64, 64) >= 2 * cus) return TileCfg::T64x64;
    if(nblk(32, 64) >= cus) return TileCfg::T32x64;
    if(nblk(64, 32) >= cus) return TileCfg::T64x32;
    return TileCfg::T32x32;
}

template <typename R, bool CPLX, int BM, int BN, int BK, int WM, int WN>
static void launch_cfg(const GemmDesc &g, bool swap, hipStream_t s) {
    const int  tiles = ((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
    const dim3 grid(tiles, 1, g.batch);
    const dim3 block(64 * WM * WN);
    if(swap) hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, BK, WM, WN, true>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, BK, WM, WN, false>), grid, block, 0, s, g);
}

template <typename R, bool CPLX>
static void launch_typed(const GemmDesc &g, TileCfg cfg, bool swap, hipStream_t s) {
    // K depth per LDS stage, sized so both double-buffered tiles fit in 64 KiB of LDS:
    // 16 elements for 8- and 16-byte elements, 32 for 4-byte ones; halved for the
    // 16-byte element in the tall/wide configurations.
    constexpr int E  = static_cast<int>(sizeof(elem_t<R, CPLX>));
    constexpr int BK = E == 4 ? 32 : 16;
    constexpr int BH = E == 16 ? 8 : BK;
    switch(cfg) {
        case TileCfg::T64x64: launch_cfg<R, CPLX, 64, 64, BK, 2, 2>(g, swap, s); break;
        case TileCfg::T32x64: launch_cfg<R, CPLX, 32, 64, BK, 1, 2>(g, swap, s); break;
        case TileCfg::T64x32: launch_cfg<R, CPLX, 64, 32, BK, 2, 1>(g, swap, s); break;
        case TileCfg::T16x128: launch_cfg<R, CPLX, 16, 128, BH, 1, 2>(g, swap, s); break;
        case TileCfg::T128x64: launch_cfg<R, CPLX, 128, 64, BH, 2, 2>(g, swap, s); break;
        case TileCfg::T32x32: launch_cfg<R, CPLX, 32, 32, BK, 1, 1>(g, swap, s); break;
        default: throw Error(DMRG_ERR_ARG, "bad tile config");
    }
}

static int fast_score(const Idx2 &x) { return x.s0 == 1 ? static_cast<int>(std::min<uint32_t>(x.c, 1u << 20)) : 0; }

void gemm(const GemmDesc &gin, hipStream_t stream, TileCfg cfg) {
    if(gin.M <= 0 || gin.N <= 0 || gin.batch <= 0) return;
    GemmDesc g = gin;
    // K == 0 is legal: every load is masked, the K loop is empty and C is written with zeros.
    if(g.K < 0) g.K = 0;
    if(g.K == 0 && g.accumulate) return;
    DMRG_REQUIRE(static_cast<int64_t>(g.M) * g.N < (int64_t(1) << 31), DMRG_ERR_DIMS, "gemm: output too large for 32-bit tile indexing");
    g.a_fast_m = fast_score(g.am) >= fast_score(g.ak) ? 1 : 0;
    g.b_fast_k = fast_score(g.bk) >= fast_score(g.bn) ? 1 : 0;
    const bool swap = (g.cm.s0 == 1 && g.cm.c >= 16) && !(g.cn.s0 == 1 && g.cn.c >= 16);
    if(cfg == TileCfg::AUTO) cfg = choose_cfg(g);
    switch(g.dtype) {
        case DMRG_FP64: launch_typed<double, false>(g, cfg, swap, stream); break;
        case DMRG_CX64: launch_typed<double, true>(g, cfg, swap, stream); break;
        case DMRG_FP32: launch_typed<float, false>(g, cfg, swap, stream); break;
        case DMRG_CX32: launch_typed<float, true>(g, cfg, swap, stream); break;
    }
    DMRG_HIP_CHECK(hipGetLastError());
}

} // namespace dmrg

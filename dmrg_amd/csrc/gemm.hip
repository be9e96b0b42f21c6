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
// Complex products use the 3-real-product (Gauss) form, 3 MFMAs per block per k-step:
//   P1 = sum ar br, P2 = sum ai bi, P3 = sum (ar + ai)(br + bi);  re = P1 - P2, im = P3 - P1 - P2.
// Its rounding error is bounded by a small multiple of eps * sum |a||b|, as for the 4-product form.
// The "swap" variant feeds B as the MFMA's first operand so that the lane index
// of the accumulator runs along M; it is picked when C is unit-stride in M, so
// the epilogue stores are coalesced either way.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

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

// LDS row pads (elements) per load direction. A tile stored [k][r] is written by lanes that walk r
// (operand contiguous along r: rows stay contiguous, the read-friendly pad applies) or k (operand contiguous
// along K: consecutive lanes write consecutive rows, so the row pitch must spread them over the banks).
// 16-byte elements: one pad of 1 serves both. The K-direction pads were chosen with a bank model of
// 64 banks x 4 bytes (and 32 x 4) over the fragment reads and the tile writes.
constexpr int pad_rfast(int row_elems, int elem_bytes) { return elem_bytes >= 16 ? 1 : lds_pad(row_elems, elem_bytes); }
constexpr int pad_kfast(int row_elems, int elem_bytes) { return (elem_bytes == 4 && row_elems >= 64) ? 9 : 1; }
constexpr int imax(int a, int b) { return a > b ? a : b; }

template <typename E>
__device__ inline E zero_elem() {
    if constexpr(std::is_same_v<E, cpx<double>> || std::is_same_v<E, cpx<float>>) return E{0, 0};
    else return E(0);
}

// Element loads that never wait: the aligned loader reads through a raw buffer resource whose base is
// the tile origin; an element outside the matrix gets an offset past the buffer's range, which the
// hardware answers with zero and without touching memory. No branch or use of the value follows the
// load, so the wait for it lands at the LDS store, after the MFMA work of the current K step.
constexpr uint32_t BUF_RANGE = 0x7fffffffu; // bytes addressable through one tile resource
constexpr uint32_t BUF_OOB   = 0x80000000u; // offset that reads as zero
constexpr int      BUF_FLAGS = 0x00020000;  // dword 3 of a gfx9 buffer resource (32-bit data format)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), static_cast<short>(0), static_cast<int>(BUF_RANGE), BUF_FLAGS);
}
template <typename E>
__device__ __forceinline__ E buffer_load_elem(__amdgpu_buffer_rsrc_t r, uint32_t off_bytes) {
    if constexpr(sizeof(E) == 16) return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(off_bytes), 0, 0));
    else if constexpr(sizeof(E) == 8) return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(r, static_cast<int>(off_bytes), 0, 0));
    else return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(off_bytes), 0, 0));
}
// Conjugation by a sign mask (0 or the sign bit of the imaginary part's high word): one XOR per element.
template <typename E>
__device__ __forceinline__ E conj_by_mask(E v, uint32_t imsign) {
    if constexpr(std::is_same_v<E, cpx<double>>) {
        const uint64_t b = __builtin_bit_cast(uint64_t, v.im) ^ (static_cast<uint64_t>(imsign) << 32);
        v.im             = __builtin_bit_cast(double, b);
    } else if constexpr(std::is_same_v<E, cpx<float>>) {
        v.im = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v.im) ^ imsign);
    }
    return v;
}
template <typename E>
__device__ __forceinline__ E conj_elem(E v, bool conj) {
    if constexpr(std::is_same_v<E, cpx<double>> || std::is_same_v<E, cpx<float>>) v.im = conj ? -v.im : v.im;
    return v;
}

// Walks the K index of one loaded element tile by tile without divisions:
// holds gk, gk % c and the element offset (gk % c) * s0 + (gk / c) * s1.
struct KCursor {
    int32_t  gk;
    uint32_t r;
    int64_t  off;
};
struct KStep {
    uint32_t dr, c;
    int64_t  delta, wrap;
};
__device__ inline KStep make_kstep(const Idx2 &x, int bk) {
    KStep st;
    st.c     = x.c;
    st.dr    = static_cast<uint32_t>(bk) % x.c;
    const int64_t dq = static_cast<int64_t>(static_cast<uint32_t>(bk) / x.c);
    st.delta = st.dr * x.s0 + dq * x.s1;
    st.wrap  = x.s1 - static_cast<int64_t>(x.c) * x.s0;
    return st;
}
__device__ inline KCursor make_kcursor(const Idx2 &x, int gk) {
    KCursor k;
    k.gk  = gk;
    k.r   = static_cast<uint32_t>(gk) % x.c;
    k.off = idx_off(x, static_cast<uint32_t>(gk));
    return k;
}
__device__ inline void advance(KCursor &k, const KStep &st, int bk) {
    k.gk += bk;
    k.r += st.dr;
    k.off += st.delta;
    if(k.r >= st.c) {
        k.r -= st.c;
        k.off += st.wrap;
    }
}

// Loads one operand tile (RT rows of the non-K dimension x BK) per K step into registers and
// then into LDS laid out [k][r]. Two address schemes:
//  ALIGNED: every chunk of the operand's index maps is a multiple or a divisor of the tile
//           extent, so offset(r0 + rr, k0 + kk) = offset(r0, k0) + offset(rr, kk): the per-element
//           part is a static 32-bit value and the tile base is uniform (scalar registers).
//  generic: per-element K cursors (KCursor) for arbitrary chunked index maps and ragged edges.
template <typename E, int RT, int BK, int NT, bool ALIGNED>
struct TileLoader {
    static constexpr int L = (RT * BK + NT - 1) / NT;
    int                  rr[L], kk[L];
    bool                 rval[L];
    bool                 lval[L];  // generic: validity of the element loaded last
    uint32_t             eoff[L];  // ALIGNED: static in-tile offset (bytes)
    uint32_t             eoffv[L]; // ALIGNED: eoff, or the zero-reading offset for rows outside the matrix
    int64_t              roff[L];  // generic: offset of the non-K coordinate
    KCursor              cur[L];   // generic: K cursors
    KCursor              kc;       // ALIGNED: K position of the next tile (uniform)
    KStep                st;
    const E             *tile_base; // ALIGNED: operand + offset of (r0, 0)
    Idx2                 kdim;
    int                  kbase;

    __device__ __forceinline__ void init(const E *src, const Idx2 &rdim, const Idx2 &kd, int r0, int rmax, int kbeg, bool fast_r, int tid) {
        kdim  = kd;
        kbase = kbeg;
#pragma unroll
        for(int e = 0; e < L; ++e) {
            const int  idx  = tid + e * NT;
            const bool used = idx < RT * BK;
            rr[e]           = fast_r ? idx % RT : (idx / BK) % RT;
            kk[e]           = fast_r ? (idx / RT) % BK : idx % BK;
            rval[e]         = used && (r0 + rr[e]) < rmax;
            if constexpr(ALIGNED) {
                eoff[e]  = static_cast<uint32_t>((idx_off(rdim, rr[e]) + idx_off(kd, kk[e])) * static_cast<int64_t>(sizeof(E)));
                eoffv[e] = rval[e] ? eoff[e] : BUF_OOB;
            } else {
                roff[e] = rval[e] ? idx_off(rdim, r0 + rr[e]) : 0;
                cur[e]  = make_kcursor(kd, kbeg + kk[e]);
            }
        }
        st = make_kstep(kd, BK);
        if constexpr(ALIGNED) {
            tile_base = src + idx_off(rdim, r0);
            kc        = make_kcursor(kd, kbeg);
        } else {
            tile_base = src;
        }
    }

    // Issues the loads of the next tile (K offset k0; tiles are loaded in K order from kbeg on) and advances.
    // Elements outside the matrix read as zero (ALIGNED) or are zeroed at store (generic). ALIGNED: the tile
    // origin advances by a cursor (no division per step) and each element costs one compare against the K
    // extent left (row validity is folded into its offset at init).
    __device__ __forceinline__ void load(E (&reg)[L], int k0, int kend) {
        if constexpr(ALIGNED) {
            const auto rs  = tile_rsrc(tile_base + kc.off);
            const int  rem = kend - k0;
            advance(kc, st, BK);
#pragma unroll
            for(int e = 0; e < L; ++e) reg[e] = buffer_load_elem<E>(rs, kk[e] < rem ? eoffv[e] : BUF_OOB);
        } else {
#pragma unroll
            for(int e = 0; e < L; ++e) {
                const bool v = rval[e] && cur[e].gk < kend;
                lval[e]      = v;
                reg[e]       = tile_base[v ? roff[e] + cur[e].off : 0];
                advance(cur[e], st, BK);
            }
        }
    }

    // Writes the tile into LDS laid out [k][r] with a row pitch of `pitch` elements.
    // The registers are updated in place (conjugation), so that each element goes to LDS without a copy.
    __device__ __forceinline__ void store(E *lds, int pitch, E (&reg)[L], int tid, uint32_t imsign) const {
#pragma unroll
        for(int e = 0; e < L; ++e)
            if(L * NT == RT * BK || tid + e * NT < RT * BK) {
                if constexpr(!ALIGNED) reg[e] = lval[e] ? reg[e] : zero_elem<E>();
                reg[e]                     = conj_by_mask(reg[e], imsign);
                lds[kk[e] * pitch + rr[e]] = reg[e];
            }
    }
};

// One output tile of g: bx is the tile's position in the launch (before the XCD remap), split the K slice
// and b the batch index.
template <typename R, bool CPLX, int BM, int BN, int BK, int WM, int WN, bool SWAP, bool ALIGNED>
__device__ __forceinline__ void gemm_body(const GemmDesc &g, int bx, int split, int b) {
    using E               = elem_t<R, CPLX>;
    using MF              = mfma_t<R>;
    using acc_t           = typename MF::acc;
    constexpr int NT      = 64 * WM * WN;
    constexpr int TM      = BM / (16 * WM);
    constexpr int TN      = BN / (16 * WN);
    constexpr int EB      = static_cast<int>(sizeof(E));
    constexpr int LDA     = BM + imax(pad_rfast(BM, EB), pad_kfast(BM, EB)); // allocated row length of A's tile
    constexpr int LDB     = BN + imax(pad_rfast(BN, EB), pad_kfast(BN, EB));
    constexpr int NACC    = CPLX ? 3 : 1; // complex: sum ar*br, sum ai*bi, sum (ar+ai)*(br+bi)
    static_assert(TM >= 1 && TN >= 1, "wave tile too small");
    static_assert(BK % 4 == 0, "BK must be a multiple of the MFMA k");

    __shared__ E As[2][BK * LDA];
    __shared__ E Bs[2][BK * LDB];

    const int tid  = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm   = wave % WM;
    const int wn   = wave / WM;

    // Tile index. Workgroups are dealt round-robin over the 8 XCDs (workgroup x runs on XCD x % 8), each with its
    // own L2. When the tile grid splits into 2 x 4 blocks, XCD x takes block (x / 4, x % 4): half of the tile rows
    // and a quarter of the tile columns, walked column by column, so that the operand panels an XCD reads are half
    // of A and a quarter of B. (Giving each XCD a range of whole tile columns instead makes every XCD stream all of
    // A; at chi = 512 that fetched 2.2x the HBM bytes for leg 1 at the same kernel time.) Otherwise each XCD
    // takes a contiguous range of tiles.
    const int tilesM = (g.M + BM - 1) / BM;
    const int tilesN = (g.N + BN - 1) / BN;
    const int total  = tilesM * tilesN;
    int       mt, nt;
    if(g.xcd_remap && (tilesM % 2) == 0 && (tilesN % 4) == 0) {
        const int x = bx % 8, j = bx / 8, rm = tilesM / 2;
        mt = (x / 4) * rm + j % rm;
        nt = (x % 4) * (tilesN / 4) + j / rm;
    } else {
        int bid = bx;
        if(g.xcd_remap && (total % 8) == 0) bid = (bid % 8) * (total / 8) + bid / 8;
        mt = bid % tilesM;
        nt = bid / tilesM;
    }
    const int m0 = mt * BM;
    const int n0 = nt * BN;

    // K range of this workgroup (split-K slice).
    const int kbeg  = split * g.kchunk;
    const int kend  = min(g.K, kbeg + g.kchunk);

    // Row pitches: A is loaded along M when g.a_fast_m, B along N when !g.b_fast_k.
    const int pa = g.a_fast_m ? BM + pad_rfast(BM, EB) : BM + pad_kfast(BM, EB);
    const int pb = !g.b_fast_k ? BN + pad_rfast(BN, EB) : BN + pad_kfast(BN, EB);
    TileLoader<E, BM, BK, NT, ALIGNED> la;
    TileLoader<E, BN, BK, NT, ALIGNED> lb;
    la.init(static_cast<const E *>(g.A) + b * g.a_batch, g.am, g.ak, m0, g.M, kbeg, g.a_fast_m, tid);
    lb.init(static_cast<const E *>(g.B) + b * g.b_batch, g.bn, g.bk, n0, g.N, kbeg, !g.b_fast_k, tid);
    E ra[TileLoader<E, BM, BK, NT, ALIGNED>::L], rb[TileLoader<E, BN, BK, NT, ALIGNED>::L];

    acc_t acc[NACC][TM][TN];
#pragma unroll
    for(int c = 0; c < NACC; ++c)
#pragma unroll
        for(int i = 0; i < TM; ++i)
#pragma unroll
            for(int j = 0; j < TN; ++j) acc[c][i][j] = acc_t{0, 0, 0, 0};

    const int      numK = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
    const uint32_t sgnA = g.conjA ? 0x80000000u : 0u, sgnB = g.conjB ? 0x80000000u : 0u; // conjugation masks
    const int lrow = lane >> 4; // k row within the 4-deep MFMA step
    const int lcol = lane & 15; // m (or n) within a 16-wide block
    if(numK > 0) {
        la.load(ra, kbeg, kend);
        lb.load(rb, kbeg, kend);
        la.store(As[0], pa, ra, tid, sgnA);
        lb.store(Bs[0], pb, rb, tid, sgnB);
        __syncthreads();
    }
    for(int kt = 0; kt < numK; ++kt) {
        const int buf = kt & 1;
        if(kt + 1 < numK) {
            la.load(ra, kbeg + (kt + 1) * BK, kend);
            lb.load(rb, kbeg + (kt + 1) * BK, kend);
        }
#pragma unroll
        for(int ks = 0; ks < BK / 4; ++ks) {
            const int kr = ks * 4 + lrow;
            E         fa[TM], fb[TN];
#pragma unroll
            for(int i = 0; i < TM; ++i) fa[i] = As[buf][kr * pa + (wm * TM + i) * 16 + lcol];
#pragma unroll
            for(int j = 0; j < TN; ++j) fb[j] = Bs[buf][kr * pb + (wn * TN + j) * 16 + lcol];
            R sa[CPLX ? TM : 1], sb[CPLX ? TN : 1];
            if constexpr(CPLX) {
#pragma unroll
                for(int i = 0; i < TM; ++i) sa[i] = fa[i].re + fa[i].im;
#pragma unroll
                for(int j = 0; j < TN; ++j) sb[j] = fb[j].re + fb[j].im;
            }
#pragma unroll
            for(int i = 0; i < TM; ++i)
#pragma unroll
                for(int j = 0; j < TN; ++j) {
                    if constexpr(CPLX) {
                        // (ar + i ai)(br + i bi): P1 += ar br, P2 += ai bi, P3 += (ar + ai)(br + bi).
                        if constexpr(!SWAP) {
                            acc[0][i][j] = MF::op(fa[i].re, fb[j].re, acc[0][i][j]);
                            acc[1][i][j] = MF::op(fa[i].im, fb[j].im, acc[1][i][j]);
                            acc[2][i][j] = MF::op(sa[i], sb[j], acc[2][i][j]);
                        } else {
                            acc[0][i][j] = MF::op(fb[j].re, fa[i].re, acc[0][i][j]);
                            acc[1][i][j] = MF::op(fb[j].im, fa[i].im, acc[1][i][j]);
                            acc[2][i][j] = MF::op(sb[j], sa[i], acc[2][i][j]);
                        }
                    } else {
                        if constexpr(!SWAP) acc[0][i][j] = MF::op(fa[i], fb[j], acc[0][i][j]);
                        else acc[0][i][j] = MF::op(fb[j], fa[i], acc[0][i][j]);
                    }
                }
        }
        if(kt + 1 < numK) {
            la.store(As[buf ^ 1], pa, ra, tid, sgnA);
            lb.store(Bs[buf ^ 1], pb, rb, tid, sgnB);
        }
        __syncthreads();
    }

    // Epilogue. 16x16 accumulator layout (measured on gfx950): lane l, register v holds column l%16 of
    // row 4*(l/16)+v for v_mfma_f32_16x16x4_f32 and of row (l/16)+4*v for v_mfma_f64_16x16x4_f64.
    // Without swap rows run along M; with swap they run along N.
    // Split-K partials go to g.part as plain column-major M x N slabs, one per (batch, split).
    E *C = static_cast<E *>(g.C) + b * g.c_batch;
    E *P = g.ksplit > 1 ? static_cast<E *>(g.part) + (static_cast<int64_t>(b) * g.ksplit + split) * g.M * g.N : nullptr;
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
                    E val;
                    if constexpr(CPLX) {
                        const R p1 = acc[0][i][j][v], p2 = acc[1][i][j][v];
                        val        = E{p1 - p2, acc[2][i][j][v] - p1 - p2};
                    }
                    else val = acc[0][i][j][v];
                    if(P) {
                        P[gm + static_cast<int64_t>(g.M) * gn] = val;
                    } else {
                        const int64_t off = idx_off(g.cm, gm) + idx_off(g.cn, gn);
                        if(g.accumulate) {
                            const E old = C[off];
                            if constexpr(CPLX) {
                                val.re += old.re;
                                val.im += old.im;
                            } else {
                                val += old;
                            }
                        }
                        C[off] = val;
                    }
                }
            }
}

template <typename R, bool CPLX, int BM, int BN, int BK, int WM, int WN, bool SWAP, bool ALIGNED>
__global__ __launch_bounds__(64 * WM *WN) void gemm_kernel(const GemmDesc g) {
    gemm_body<R, CPLX, BM, BN, BK, WM, WN, SWAP, ALIGNED>(g, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Grouped GEMM: one launch for several independent GEMMs of one tile configuration. Workgroup x belongs to
// the GEMM whose tile range [tile_start[i], tile_start[i + 1]) holds x (no split-K, batch 1 each); the
// descriptors and the prefix sums live in device memory, so a launch can be captured once and replayed.
template <typename R, bool CPLX, int BM, int BN, int BK, int WM, int WN, bool SWAP, bool ALIGNED>
__global__ __launch_bounds__(64 * WM *WN) void gemm_grouped_kernel(const GemmDesc *descs, const int *tile_start, int ngemm) {
    const int x  = blockIdx.x;
    int       lo = 0, hi = ngemm - 1; // last i with tile_start[i] <= x
    while(lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if(tile_start[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    gemm_body<R, CPLX, BM, BN, BK, WM, WN, SWAP, ALIGNED>(descs[lo], x - tile_start[lo], 0, 0);
}

// GEMM with tiny K and M (K <= SMALLK_MAX_K, M <= SMALLK_MAX_M): A sits in LDS, every thread streams one column
// n of B (K elements, kept in registers) and writes the M elements of C's column n. Consecutive threads take
// consecutive n, so B and C are read and written along their n index maps; the kernel is HBM-bound.
template <typename R, bool CPLX, int KMAX>
__global__ __launch_bounds__(256) void smallk_kernel(const GemmDesc g) {
    using E = elem_t<R, CPLX>;
    __shared__ E As[SMALLK_MAX_M * SMALLK_MAX_K];
    const int    b  = blockIdx.y;
    const E     *A  = static_cast<const E *>(g.A) + b * g.a_batch;
    const E     *B  = static_cast<const E *>(g.B) + b * g.b_batch;
    E           *C  = static_cast<E *>(g.C) + b * g.c_batch;
    for(int t = threadIdx.x; t < g.M * g.K; t += blockDim.x) {
        const int m = t % g.M, k = t / g.M;
        As[t]       = conj_elem(A[idx_off(g.am, m) + idx_off(g.ak, k)], g.conjA != 0);
    }
    __syncthreads();
    for(int64_t n = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; n < g.N; n += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t boff = idx_off(g.bn, static_cast<uint32_t>(n));
        const int64_t coff = idx_off(g.cn, static_cast<uint32_t>(n));
        E             bv[KMAX];
#pragma unroll
        for(int k = 0; k < KMAX; ++k)
            if(k < g.K) bv[k] = conj_elem(B[boff + idx_off(g.bk, k)], g.conjB != 0);
        for(int m = 0; m < g.M; ++m) {
            R ar = 0, ai = 0;
#pragma unroll
            for(int k = 0; k < KMAX; ++k)
                if(k < g.K) {
                    const E a = As[m + g.M * k];
                    if constexpr(CPLX) {
                        ar = fma(a.re, bv[k].re, fma(-a.im, bv[k].im, ar));
                        ai = fma(a.re, bv[k].im, fma(a.im, bv[k].re, ai));
                    } else {
                        ar = fma(a, bv[k], ar);
                    }
                }
            E *dst = C + coff + idx_off(g.cm, m);
            if constexpr(CPLX) {
                if(g.accumulate) ar += dst->re, ai += dst->im;
                *dst = E{ar, ai};
            } else {
                if(g.accumulate) ar += *dst;
                *dst = ar;
            }
        }
    }
}

// Sums the split-K partial slabs in a fixed order and writes C through its index map.
template <typename E>
__global__ void splitk_reduce(const GemmDesc g) {
    const int64_t MN    = static_cast<int64_t>(g.M) * g.N;
    const int64_t total = MN * g.batch;
    const E      *P     = static_cast<const E *>(g.part);
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < total;
        t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b   = t / MN;
        const int64_t mn  = t - b * MN;
        const int     m   = static_cast<int>(mn % g.M);
        const int     n   = static_cast<int>(mn / g.M);
        const E      *src = P + b * g.ksplit * MN + mn;
        E             s   = src[0];
        for(int k = 1; k < g.ksplit; ++k) {
            const E v = src[k * MN];
            if constexpr(std::is_same_v<E, cpx<double>> || std::is_same_v<E, cpx<float>>) {
                s.re += v.re;
                s.im += v.im;
            } else {
                s += v;
            }
        }
        E            *C   = static_cast<E *>(g.C) + b * g.c_batch;
        const int64_t off = idx_off(g.cm, m) + idx_off(g.cn, n);
        if(g.accumulate) {
            const E old = C[off];
            if constexpr(std::is_same_v<E, cpx<double>> || std::is_same_v<E, cpx<float>>) {
                s.re += old.re;
                s.im += old.im;
            } else {
                s += old;
            }
        }
        C[off] = s;
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
        case TileCfg::T32x32K: return {32, 32};
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

// A plan: tile configuration and K slices; with nsplit > 0 the columns [0, nsplit) run as one launch without
// K slicing and the tail [nsplit, N) as a second launch with ksplit2 K slices, so that the tail (the part of
// the grid that would occupy the GPU only partially) is spread over all CUs.
struct LaunchPlan {
    TileCfg cfg;
    int     ksplit;
    int     nsplit  = 0;
    int     ksplit2 = 1;
};

static LaunchPlan heuristic_plan(const GemmDesc &g) {
    const int64_t cus  = device_cu_count();
    auto          nblk = [&](int bm, int bn) { return static_cast<int64_t>((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn) * g.batch; };
    LaunchPlan    p{TileCfg::T64x64, 1};
    if(g.M <= 16) p.cfg = TileCfg::T16x128;
    else if(g.M >= 48 && g.N >= 48) p.cfg = TileCfg::T64x64;
    else p.cfg = g.M < g.N ? TileCfg::T32x64 : TileCfg::T64x32;
    // Fewer 64x64 tiles than CUs: 32x32 tiles, four times as many.
    if(p.cfg == TileCfg::T64x64 && nblk(64, 64) < cus) p.cfg = TileCfg::T32x32;
    const CfgShape sh    = cfg_shape(p.cfg);
    const int64_t  tiles = nblk(sh.BM, sh.BN);
    // Small grids with a long K: slice K (slices of at least 128) so that about two workgroups per CU are in
    // flight. These are the plans the tuner picks for the chi <= 256 matvec legs.
    if(tiles < 2 * cus && g.K >= 256 && p.cfg != TileCfg::T16x128) {
        int64_t ks = std::min<int64_t>({(2 * cus + tiles - 1) / tiles, g.K / 128, 8});
        if(g.part && ks > 1) ks = std::min<int64_t>(ks, g.part_capacity / (static_cast<int64_t>(g.M) * g.N * g.batch));
        else ks = 1;
        p.ksplit = static_cast<int>(std::max<int64_t>(1, ks));
    }
    return p;
}

// Tile-separable index map: offset(t0 + t) = offset(t0) + offset(t) for every tile origin t0 (a multiple
// of ext) and 0 <= t < ext, which holds when each chunk is a multiple or a divisor of ext.
static bool separable(const Idx2 &x, int ext) {
    return x.c >= 0x7fffffffu || (x.c % static_cast<uint32_t>(ext)) == 0 || (static_cast<uint32_t>(ext) % x.c) == 0;
}
// Largest in-tile offset (bytes) must stay inside the aligned loader's buffer range.
static bool small_tile_offsets(const Idx2 &r, int rt, const Idx2 &k, int bk, int elem_bytes) {
    auto mx = [](const Idx2 &x, int ext) {
        int64_t m = 0;
        for(int t = 0; t < ext; ++t) m = std::max<int64_t>(m, std::llabs(idx_off(x, t)));
        return m;
    };
    return (mx(r, rt) + mx(k, bk)) * elem_bytes < static_cast<int64_t>(BUF_RANGE) && r.s0 >= 0 && r.s1 >= 0 && k.s0 >= 0 &&
           k.s1 >= 0;
}

template <typename R, bool CPLX, int BM, int BN, int BK, int WM, int WN>
static void launch_cfg(const GemmDesc &g, bool swap, hipStream_t s) {
    const int  tiles = ((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
    const dim3 grid(tiles, g.ksplit, g.batch);
    const dim3 block(64 * WM * WN);
    const bool aligned = separable(g.am, BM) && separable(g.ak, BK) && separable(g.bn, BN) && separable(g.bk, BK) &&
                         (g.kchunk % BK) == 0 && small_tile_offsets(g.am, BM, g.ak, BK, static_cast<int>(sizeof(elem_t<R, CPLX>))) &&
                         small_tile_offsets(g.bn, BN, g.bk, BK, static_cast<int>(sizeof(elem_t<R, CPLX>)));
    if(aligned) {
        if(swap) hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, BK, WM, WN, true, true>), grid, block, 0, s, g);
        else hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, BK, WM, WN, false, true>), grid, block, 0, s, g);
    } else {
        if(swap) hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, BK, WM, WN, true, false>), grid, block, 0, s, g);
        else hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, BK, WM, WN, false, false>), grid, block, 0, s, g);
    }
    if(g.ev_after_main) DMRG_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(g.ev_after_main), s));
    if(g.ksplit > 1) {
        const int64_t total = static_cast<int64_t>(g.M) * g.N * g.batch;
        const dim3    rg(static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 8192)));
        hipLaunchKernelGGL((splitk_reduce<elem_t<R, CPLX>>), rg, dim3(256), 0, s, g);
    }
}

template <typename R, bool CPLX>
static void launch_typed(const GemmDesc &g, TileCfg cfg, bool swap, hipStream_t s) {
    // K depth per LDS stage, sized so both double-buffered tiles fit in 64 KiB of LDS:
    // 16 elements for 8- and 16-byte elements, 32 for 4-byte ones; halved for the
    // 16-byte element in the tall/wide configurations. All configurations run 4 waves.
    constexpr int E  = static_cast<int>(sizeof(elem_t<R, CPLX>));
    constexpr int BK = E == 4 ? 32 : 16;
    constexpr int BH = E == 16 ? 8 : BK;
    constexpr int BD = E == 4 ? 128 : 64; // deep K stage of T32x32K (LDS: at most 132 KiB of the 160 KiB)
    switch(cfg) {
        case TileCfg::T64x64: launch_cfg<R, CPLX, 64, 64, BK, 2, 2>(g, swap, s); break;
        case TileCfg::T32x64: launch_cfg<R, CPLX, 32, 64, BK, 2, 2>(g, swap, s); break;
        case TileCfg::T64x32: launch_cfg<R, CPLX, 64, 32, BK, 2, 2>(g, swap, s); break;
        case TileCfg::T16x128: launch_cfg<R, CPLX, 16, 128, BH, 1, 4>(g, swap, s); break;
        case TileCfg::T128x64: launch_cfg<R, CPLX, 128, 64, BH, 2, 2>(g, swap, s); break;
        case TileCfg::T32x32: launch_cfg<R, CPLX, 32, 32, BK, 2, 2>(g, swap, s); break;
        case TileCfg::T32x32K: launch_cfg<R, CPLX, 32, 32, BD, 2, 2>(g, swap, s); break;
        case TileCfg::SMALLK: {
            if(g.K > SMALLK_MAX_K || g.M > SMALLK_MAX_M) { // outside the kernel's limits
                launch_cfg<R, CPLX, 64, 64, BK, 2, 2>(g, swap, s);
                break;
            }
            const int64_t blocks = std::min<int64_t>((static_cast<int64_t>(g.N) + 255) / 256, 8192);
            const dim3    grid(static_cast<unsigned>(blocks), g.batch);
            if(g.K <= 8) hipLaunchKernelGGL((smallk_kernel<R, CPLX, 8>), grid, dim3(256), 0, s, g);
            else if(g.K <= 16) hipLaunchKernelGGL((smallk_kernel<R, CPLX, 16>), grid, dim3(256), 0, s, g);
            else hipLaunchKernelGGL((smallk_kernel<R, CPLX, SMALLK_MAX_K>), grid, dim3(256), 0, s, g);
            if(g.ev_after_main) DMRG_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(g.ev_after_main), s));
            break;
        }
        default: throw Error(DMRG_ERR_ARG, "bad tile config");
    }
}

static int fast_score(const Idx2 &x) { return x.s0 == 1 ? static_cast<int>(std::min<uint32_t>(x.c, 1u << 20)) : 0; }

// ---------------------------------------------------------------------------------------
// Autotuner: times candidate (tile configuration, K slices) plans of a GEMM with HIP events on
// its own stream and caches the fastest per problem signature (dtype, sizes, index maps).

static std::mutex                        g_tune_mu;
static std::map<std::string, LaunchPlan> g_tuned;
static int                               g_autotune     = 1;
// GEMMs below this many (complex or real) multiply-adds, and with fewer than a sixteenth as many bytes of operands,
// take the size heuristic instead of being tuned.
static int64_t                           g_tune_min_macs = int64_t(1) << 29;
static int                               g_override_cfg = -1, g_override_ks = 0;
static int                               g_override_nsplit = 0, g_override_ks2 = 1;

static std::string tune_key(const GemmDesc &g) {
    char buf[512];
    auto I = [](const Idx2 &x) { return std::to_string(x.s0) + "," + std::to_string(x.s1) + "," + std::to_string(x.c); };
    std::snprintf(buf, sizeof(buf), "%d|%d|%d|%d|%d|%d|%d|", static_cast<int>(g.dtype), g.M, g.N, g.K, g.batch,
                  g.part ? 1 : 0, static_cast<int>(std::min<int64_t>(g.part_capacity / std::max<int64_t>(1, int64_t(g.M) * g.N * g.batch), 8)));
    return std::string(buf) + I(g.am) + "|" + I(g.ak) + "|" + I(g.bk) + "|" + I(g.bn) + "|" + I(g.cm) + "|" + I(g.cn);
}

static void launch_plan_single(GemmDesc g, const LaunchPlan &p0, hipStream_t stream);

// K depth of one LDS stage rounded up to 32 (every configuration's BK divides 32, except T32x32K's deep stage).
static int deep_bk(TileCfg c, dmrg_dtype t) { return c == TileCfg::T32x32K ? (dtype_size(t) == 4 ? 128 : 64) : 32; }

// Columns [n0, n0 + nn) of g as a GEMM of their own (valid when the B and C column maps are separable at n0).
static GemmDesc column_slice(const GemmDesc &g, int n0, int nn) {
    GemmDesc h     = g;
    const size_t es = dtype_size(g.dtype);
    h.B            = static_cast<const char *>(g.B) + idx_off(g.bn, static_cast<uint32_t>(n0)) * static_cast<int64_t>(es);
    h.C            = static_cast<char *>(g.C) + idx_off(g.cn, static_cast<uint32_t>(n0)) * static_cast<int64_t>(es);
    h.N            = nn;
    h.ev_after_main = nullptr;
    return h;
}
static bool column_separable(const Idx2 &x, int n0) { return x.c >= 0x7fffffffu || (static_cast<uint32_t>(n0) % x.c) == 0; }

static void launch_plan(const GemmDesc &g, const LaunchPlan &p, hipStream_t stream) {
    if(p.nsplit > 0 && p.nsplit < g.N && g.batch == 1 && column_separable(g.bn, p.nsplit) && column_separable(g.cn, p.nsplit)) {
        GemmDesc head = column_slice(g, 0, p.nsplit), tail = column_slice(g, p.nsplit, g.N - p.nsplit);
        launch_plan_single(head, LaunchPlan{p.cfg, 1}, stream);
        tail.ev_after_main = g.ev_after_main;
        launch_plan_single(tail, LaunchPlan{p.cfg, p.ksplit2}, stream);
        return;
    }
    launch_plan_single(g, LaunchPlan{p.cfg, p.ksplit}, stream);
}

static void launch_plan_single(GemmDesc g, const LaunchPlan &p0, hipStream_t stream) {
    LaunchPlan p = p0;
    if(p.ksplit > 1) {
        const int64_t need = static_cast<int64_t>(g.M) * g.N * g.batch * p.ksplit;
        if(!g.part || g.part_capacity < need) p.ksplit = 1;
    }
    g.ksplit = 1;
    const int kr = deep_bk(p.cfg, g.dtype); // K extent per slice: a multiple of the configuration's BK
    g.kchunk = (g.K + kr - 1) / kr * kr;
    if(p.ksplit > 1) {
        const int kc = (((g.K + p.ksplit - 1) / p.ksplit) + kr - 1) / kr * kr;
        g.kchunk     = kc;
        g.ksplit     = (g.K + kc - 1) / kc;
    }
    if(g.kchunk <= 0) g.kchunk = 32;
    const bool swap = (g.cm.s0 == 1 && g.cm.c >= 16) && !(g.cn.s0 == 1 && g.cn.c >= 16);
    switch(g.dtype) {
        case DMRG_FP64: launch_typed<double, false>(g, p.cfg, swap, stream); break;
        case DMRG_CX64: launch_typed<double, true>(g, p.cfg, swap, stream); break;
        case DMRG_FP32: launch_typed<float, false>(g, p.cfg, swap, stream); break;
        case DMRG_CX32: launch_typed<float, true>(g, p.cfg, swap, stream); break;
    }
    DMRG_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// Grouped launches

static int group_bk(dmrg_dtype t, TileCfg cfg) {
    const int e = static_cast<int>(dtype_size(t));
    const int bk = e == 4 ? 32 : 16;
    return (cfg == TileCfg::T16x128 || cfg == TileCfg::T128x64) && e == 16 ? 8 : bk;
}

int grouped_tiles(const GemmDesc &g, TileCfg cfg) {
    const CfgShape sh = cfg_shape(cfg);
    return ((g.M + sh.BM - 1) / sh.BM) * ((g.N + sh.BN - 1) / sh.BN);
}

// Four launch kinds at most, each with ngemm descriptors and ngemm + 1 starts, 256-byte aligned.
size_t grouped_store_bytes(size_t ngemm) {
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    return 4 * (al(sizeof(GemmDesc) * ngemm) + al(sizeof(int) * (ngemm + 1)));
}

std::vector<GroupedGemm> grouped_prepare(const std::vector<GemmDesc> &gs, TileCfg cfg, void *store, size_t store_bytes) {
    DMRG_REQUIRE(!store || store_bytes >= grouped_store_bytes(gs.size()), DMRG_ERR_ARG, "grouped GEMM: descriptor store too small");
    size_t store_off = 0;
    DMRG_REQUIRE(cfg == TileCfg::T64x64 || cfg == TileCfg::T32x32 || cfg == TileCfg::T16x128, DMRG_ERR_ARG,
                 "grouped GEMM: tile configuration not available");
    const CfgShape sh = cfg_shape(cfg);
    const int      bk = group_bk(gs.empty() ? DMRG_FP64 : gs[0].dtype, cfg);
    // (swap, aligned) -> member GEMMs
    std::vector<GemmDesc> kinds[4];
    for(GemmDesc g : gs) {
        if(g.M <= 0 || g.N <= 0) continue;
        DMRG_REQUIRE(g.batch == 1 && !g.accumulate && g.K > 0, DMRG_ERR_ARG, "grouped GEMM: batch 1, K > 0, no accumulation");
        DMRG_REQUIRE(g.dtype == gs[0].dtype, DMRG_ERR_ARG, "grouped GEMM: one dtype per group");
        DMRG_REQUIRE(static_cast<int64_t>(g.M) * g.N < (int64_t(1) << 31), DMRG_ERR_DIMS, "gemm: output too large");
        g.a_fast_m      = fast_score(g.am) >= fast_score(g.ak) ? 1 : 0;
        g.b_fast_k      = fast_score(g.bk) >= fast_score(g.bn) ? 1 : 0;
        g.ksplit        = 1;
        g.kchunk        = (g.K + 31) / 32 * 32;
        g.xcd_remap     = 0;
        g.part          = nullptr;
        g.part_capacity = 0;
        g.ev_after_main = nullptr;
        const bool swap = (g.cm.s0 == 1 && g.cm.c >= 16) && !(g.cn.s0 == 1 && g.cn.c >= 16);
        const int  es   = static_cast<int>(dtype_size(g.dtype));
        const bool aligned = separable(g.am, sh.BM) && separable(g.ak, bk) && separable(g.bn, sh.BN) && separable(g.bk, bk) &&
                             (g.kchunk % bk) == 0 && small_tile_offsets(g.am, sh.BM, g.ak, bk, es) &&
                             small_tile_offsets(g.bn, sh.BN, g.bk, bk, es);
        kinds[(swap ? 2 : 0) + (aligned ? 1 : 0)].push_back(g);
    }
    std::vector<GroupedGemm> out;
    for(int k = 0; k < 4; ++k) {
        if(kinds[k].empty()) continue;
        GroupedGemm gg;
        gg.dtype   = kinds[k][0].dtype;
        gg.cfg     = cfg;
        gg.swap    = (k & 2) != 0;
        gg.aligned = (k & 1) != 0;
        gg.ngemm   = static_cast<int>(kinds[k].size());
        std::vector<int> st(gg.ngemm + 1, 0);
        for(int i = 0; i < gg.ngemm; ++i) {
            const int64_t next = static_cast<int64_t>(st[i]) + grouped_tiles(kinds[k][i], cfg);
            DMRG_REQUIRE(next < (int64_t(1) << 31), DMRG_ERR_DIMS, "grouped GEMM: too many tiles");
            st[i + 1] = static_cast<int>(next);
        }
        gg.tiles = st[gg.ngemm];
        if(store) {
            auto al    = [](size_t b) { return (b + 255) / 256 * 256; };
            char *base = static_cast<char *>(store);
            gg.owned   = false;
            gg.descs   = base + store_off;
            store_off += al(sizeof(GemmDesc) * gg.ngemm);
            gg.starts = reinterpret_cast<int *>(base + store_off);
            store_off += al(sizeof(int) * (gg.ngemm + 1));
        } else if(hipMalloc(&gg.descs, sizeof(GemmDesc) * gg.ngemm) != hipSuccess ||
                  hipMalloc(reinterpret_cast<void **>(&gg.starts), sizeof(int) * (gg.ngemm + 1)) != hipSuccess) {
            (void) hipGetLastError();
            out.push_back(gg);
            grouped_free(out);
            throw Error(DMRG_ERR_OOM, "grouped GEMM: descriptor allocation failed");
        }
        DMRG_HIP_CHECK(hipMemcpy(gg.descs, kinds[k].data(), sizeof(GemmDesc) * gg.ngemm, hipMemcpyHostToDevice));
        DMRG_HIP_CHECK(hipMemcpy(gg.starts, st.data(), sizeof(int) * (gg.ngemm + 1), hipMemcpyHostToDevice));
        out.push_back(gg);
    }
    return out;
}

template <typename R, bool CPLX, int BM, int BN, int BK, int WM, int WN>
static void grouped_launch_cfg(const GroupedGemm &g, hipStream_t s) {
    const dim3 grid(g.tiles), block(64 * WM * WN);
    const auto *d = static_cast<const GemmDesc *>(g.descs);
    if(g.aligned) {
        if(g.swap) hipLaunchKernelGGL((gemm_grouped_kernel<R, CPLX, BM, BN, BK, WM, WN, true, true>), grid, block, 0, s, d, g.starts, g.ngemm);
        else hipLaunchKernelGGL((gemm_grouped_kernel<R, CPLX, BM, BN, BK, WM, WN, false, true>), grid, block, 0, s, d, g.starts, g.ngemm);
    } else {
        if(g.swap) hipLaunchKernelGGL((gemm_grouped_kernel<R, CPLX, BM, BN, BK, WM, WN, true, false>), grid, block, 0, s, d, g.starts, g.ngemm);
        else hipLaunchKernelGGL((gemm_grouped_kernel<R, CPLX, BM, BN, BK, WM, WN, false, false>), grid, block, 0, s, d, g.starts, g.ngemm);
    }
}

template <typename R, bool CPLX>
static void grouped_launch_typed(const GroupedGemm &g, hipStream_t s) {
    constexpr int E  = static_cast<int>(sizeof(elem_t<R, CPLX>));
    constexpr int BK = E == 4 ? 32 : 16;
    constexpr int BH = E == 16 ? 8 : BK;
    switch(g.cfg) {
        case TileCfg::T64x64: grouped_launch_cfg<R, CPLX, 64, 64, BK, 2, 2>(g, s); break;
        case TileCfg::T32x32: grouped_launch_cfg<R, CPLX, 32, 32, BK, 2, 2>(g, s); break;
        case TileCfg::T16x128: grouped_launch_cfg<R, CPLX, 16, 128, BH, 1, 4>(g, s); break;
        default: throw Error(DMRG_ERR_ARG, "grouped GEMM: tile configuration not available");
    }
}

void grouped_launch(const GroupedGemm &g, hipStream_t s) {
    if(g.tiles <= 0) return;
    switch(g.dtype) {
        case DMRG_FP64: grouped_launch_typed<double, false>(g, s); break;
        case DMRG_CX64: grouped_launch_typed<double, true>(g, s); break;
        case DMRG_FP32: grouped_launch_typed<float, false>(g, s); break;
        case DMRG_CX32: grouped_launch_typed<float, true>(g, s); break;
    }
    DMRG_HIP_CHECK(hipGetLastError());
}

void grouped_free(std::vector<GroupedGemm> &gg) {
    for(auto &g : gg) {
        if(g.owned && g.descs) (void) hipFree(g.descs);
        if(g.owned && g.starts) (void) hipFree(g.starts);
        g.descs  = nullptr;
        g.starts = nullptr;
    }
    gg.clear();
}

// Workgroups of configuration t that one CU holds at once (aligned, swapped variant as the representative).
template <typename R, bool CPLX, int BM, int BN, int BK, int WM, int WN>
static int occupancy_of() {
    int nb = 0;
    if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gemm_kernel<R, CPLX, BM, BN, BK, WM, WN, true, true>, 64 * WM * WN, 0) != hipSuccess)
        return 0;
    return nb;
}
template <typename R, bool CPLX>
static int occupancy_typed(TileCfg t) {
    constexpr int E  = static_cast<int>(sizeof(elem_t<R, CPLX>));
    constexpr int BK = E == 4 ? 32 : 16;
    switch(t) {
        case TileCfg::T64x64: return occupancy_of<R, CPLX, 64, 64, BK, 2, 2>();
        case TileCfg::T32x64: return occupancy_of<R, CPLX, 32, 64, BK, 2, 2>();
        case TileCfg::T64x32: return occupancy_of<R, CPLX, 64, 32, BK, 2, 2>();
        default: return 0;
    }
}
static int resident_blocks(dmrg_dtype dt, TileCfg t) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(static_cast<int>(dt), static_cast<int>(t));
    auto       it  = cache.find(key);
    if(it != cache.end()) return it->second;
    int v = 0;
    switch(dt) {
        case DMRG_FP64: v = occupancy_typed<double, false>(t); break;
        case DMRG_CX64: v = occupancy_typed<double, true>(t); break;
        case DMRG_FP32: v = occupancy_typed<float, false>(t); break;
        case DMRG_CX32: v = occupancy_typed<float, true>(t); break;
    }
    cache[key] = v;
    return v;
}

static std::vector<LaunchPlan> candidates(const GemmDesc &g) {
    std::vector<LaunchPlan> c;
    const TileCfg cfgs[] = {TileCfg::T64x64, TileCfg::T32x64, TileCfg::T64x32, TileCfg::T32x32, TileCfg::T128x64, TileCfg::T16x128};
    if(g.K <= SMALLK_MAX_K && g.M <= SMALLK_MAX_M) c.push_back({TileCfg::SMALLK, 1});
    // Hybrid plans: whole waves of tiles without K slicing, the remaining tile columns sliced over K.
    if(g.part && g.batch == 1 && g.K >= 128) {
        const TileCfg hcfgs[] = {TileCfg::T64x64, TileCfg::T32x64, TileCfg::T64x32};
        for(TileCfg t : hcfgs) {
            const CfgShape sh    = cfg_shape(t);
            const int      tm    = (g.M + sh.BM - 1) / sh.BM, tn = (g.N + sh.BN - 1) / sh.BN;
            const int64_t  slots = static_cast<int64_t>(device_cu_count()) * resident_blocks(g.dtype, t);
            if(slots <= 0 || static_cast<int64_t>(tm) * tn <= slots) continue;
            const int64_t full_cols = (static_cast<int64_t>(tm) * tn / slots) * slots / tm; // tile columns in whole waves
            if(full_cols <= 0 || full_cols >= tn) continue;
            const int nsplit = static_cast<int>(full_cols * sh.BN);
            for(int ks2 : {2, 3, 4, 8}) {
                if(g.K < 64 * ks2 || g.part_capacity < static_cast<int64_t>(g.M) * (g.N - nsplit) * ks2) continue;
                c.push_back({t, 1, nsplit, ks2});
            }
        }
    }
    for(TileCfg t : cfgs) {
        if(t == TileCfg::T16x128 && g.M > 32) continue;
        if(t == TileCfg::T128x64 && g.M < 128) continue;
        for(int ks : {1, 2, 3, 4, 8}) {
            if(ks > 1 && (g.K < 256 * ks || !g.part || g.part_capacity < static_cast<int64_t>(g.M) * g.N * g.batch * ks)) continue;
            c.push_back({t, ks});
        }
    }
    // Deep-K 32x32 tiles for grids that fill at most two workgroups per CU; K slices down to one stage each.
    const int64_t tiles32 = static_cast<int64_t>((g.M + 31) / 32) * ((g.N + 31) / 32) * g.batch;
    if(tiles32 <= 2 * device_cu_count()) {
        const int bd = deep_bk(TileCfg::T32x32K, g.dtype);
        for(int ks : {1, 2, 4, 8}) {
            if(ks > 1 && (g.K < bd * ks || !g.part || g.part_capacity < static_cast<int64_t>(g.M) * g.N * g.batch * ks)) continue;
            c.push_back({TileCfg::T32x32K, ks});
        }
    }
    return c;
}

static LaunchPlan tune(const GemmDesc &g, hipStream_t stream) {
    const auto cands = candidates(g);
    hipEvent_t e0, e1;
    DMRG_HIP_CHECK(hipEventCreate(&e0));
    DMRG_HIP_CHECK(hipEventCreate(&e1));
    LaunchPlan best = heuristic_plan(g);
    float      best_ms = 1e30f;
    try {
        for(const auto &p : cands) {
            launch_plan(g, p, stream); // warm-up (first-use code load, caches)
            // Fastest of 5 timed launches, so that one launch slowed by a clock change does not decide the plan
            // (with the mean of 3 launches, the MPO leg at chi = 512 went to a plan 20 % slower in some runs).
            float tmin = 1e30f;
            for(int r = 0; r < 5; ++r) {
                DMRG_HIP_CHECK(hipEventRecord(e0, stream));
                launch_plan(g, p, stream);
                DMRG_HIP_CHECK(hipEventRecord(e1, stream));
                DMRG_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0;
                DMRG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                tmin = std::min(tmin, ms);
            }
            if(tmin < best_ms) {
                best_ms = tmin;
                best    = p;
            }
        }
    } catch(...) {
        (void) hipEventDestroy(e0);
        (void) hipEventDestroy(e1);
        throw;
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    return best;
}

// Optional persistent cache: with DMRG_TUNE_CACHE=<file>, tuned plans are loaded at first use and each new
// plan is appended, so that separate processes (e.g. profiler passes) launch the same kernels.
static const char *tune_cache_path() {
    static const char *p = std::getenv("DMRG_TUNE_CACHE");
    return (p && *p) ? p : nullptr;
}
static void load_tune_cache_locked() {
    static bool loaded = false;
    if(loaded) return;
    loaded = true;
    const char *path = tune_cache_path();
    if(!path) return;
    FILE *f = std::fopen(path, "r");
    if(!f) return;
    char line[1024];
    while(std::fgets(line, sizeof(line), f)) {
        std::string l(line);
        const size_t sp = l.rfind(' ');
        if(sp == std::string::npos) continue;
        int cfg = 0, ks = 1, nsplit = 0, ks2 = 1;
        if(std::sscanf(l.c_str() + sp + 1, "%d:%d:%d:%d", &cfg, &ks, &nsplit, &ks2) < 2) continue;
        g_tuned[l.substr(0, sp)] = LaunchPlan{static_cast<TileCfg>(cfg), ks, nsplit, ks2};
    }
    std::fclose(f);
}
static void append_tune_cache_locked(const std::string &key, const LaunchPlan &p) {
    const char *path = tune_cache_path();
    if(!path) return;
    if(FILE *f = std::fopen(path, "a")) {
        std::fprintf(f, "%s %d:%d:%d:%d\n", key.c_str(), static_cast<int>(p.cfg), p.ksplit, p.nsplit, p.ksplit2);
        std::fclose(f);
    }
}

void set_autotune(int on) { g_autotune = on; }
void set_tune_min_macs(int64_t macs) { g_tune_min_macs = macs; }
int  autotune_enabled() { return g_autotune; }
void set_plan_override(int cfg, int ks, int nsplit, int ks2) {
    g_override_cfg    = cfg;
    g_override_ks     = ks;
    g_override_nsplit = nsplit;
    g_override_ks2    = ks2;
}
void clear_tuning() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned.clear();
}
int tuned_entries() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    return static_cast<int>(g_tuned.size());
}

static thread_local bool g_tls_no_tune = false;
NoTuneScope::NoTuneScope() : prev_(g_tls_no_tune) { g_tls_no_tune = true; }
NoTuneScope::~NoTuneScope() { g_tls_no_tune = prev_; }

void gemm(const GemmDesc &gin, hipStream_t stream, TileCfg cfg, int ksplit) {
    if(gin.M <= 0 || gin.N <= 0 || gin.batch <= 0) return;
    GemmDesc g = gin;
    // K == 0 is legal: every load is skipped, the K loop is empty and C is written with zeros.
    if(g.K < 0) g.K = 0;
    if(g.K == 0 && g.accumulate) return;
    DMRG_REQUIRE(static_cast<int64_t>(g.M) * g.N < (int64_t(1) << 31), DMRG_ERR_DIMS, "gemm: output too large for 32-bit tile indexing");
    g.a_fast_m = fast_score(g.am) >= fast_score(g.ak) ? 1 : 0;
    g.b_fast_k = fast_score(g.bk) >= fast_score(g.bn) ? 1 : 0;
    LaunchPlan p;
    if(g_override_cfg >= 0) {
        p = {static_cast<TileCfg>(g_override_cfg), std::max(1, g_override_ks), g_override_nsplit, std::max(1, g_override_ks2)};
    } else if(cfg != TileCfg::AUTO) {
        p = {cfg, std::max(1, ksplit)};
    } else {
        // Tuning needs synchronisation, so it is skipped for accumulating GEMMs (repeated launches would
        // add up) and while the stream is being captured into a graph.
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        DMRG_HIP_CHECK(hipStreamIsCapturing(stream, &cs));
        // Only large GEMMs are timed (at least g_tune_min_macs multiply-adds, or g_tune_min_macs / 16 bytes of
        // operands for the memory-bound ones): a sweep or a request stream meets thousands of small shapes once
        // each, whose tuning (about 180 timed, synchronised launches) costs more than any plan saves.
        const double macs  = static_cast<double>(g.M) * g.N * std::max(g.K, 1) * g.batch;
        const double bytes = (static_cast<double>(g.M) * g.K + static_cast<double>(g.K) * g.N + static_cast<double>(g.M) * g.N) *
                             g.batch * static_cast<double>(dtype_size(g.dtype));
        const bool   large = macs >= static_cast<double>(g_tune_min_macs) || bytes * 16.0 >= static_cast<double>(g_tune_min_macs);
        const bool   can_tune = g_autotune && !g.accumulate && cs == hipStreamCaptureStatusNone && !g_tls_no_tune && large;
        const std::string key = tune_key(g);
        bool              found = false;
        {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            load_tune_cache_locked();
            auto                        it = g_tuned.find(key);
            if(it != g_tuned.end()) {
                p     = it->second;
                found = true;
            }
        }
        if(!found) {
            p = can_tune ? tune(g, stream) : heuristic_plan(g);
            if(can_tune) {
                std::lock_guard<std::mutex> lock(g_tune_mu);
                g_tuned[key] = p;
                append_tune_cache_locked(key, p);
            }
        }
        if(ksplit > 0) p = LaunchPlan{p.cfg, ksplit};
    }
    launch_plan(g, p, stream);
}

// ---------------------------------------------------------------------------------------
// Block one-sided Jacobi SVD steps on the matrix cores. Blocks of 16 columns of W (m x npad, ld m) and V
// (n x npad, ld n) are paired by the circle method (round r over nbp players, pair = blockIdx.x, the lower block
// first); a pair's 32 columns are X = [block P, block Q], addressed as one index map with a gap between the two
// blocks. The descriptors are built in the kernel, so a round costs one launch per step and no host work.

namespace {

__device__ inline void bj_block_pair(int nbp, int r, int i, int &P, int &Q) {
    if(i == 0) {
        P = r % (nbp - 1);
        Q = nbp - 1;
    } else {
        P = (r + i) % (nbp - 1);
        Q = (r - i + (nbp - 1)) % (nbp - 1);
    }
    if(P > Q) {
        const int t = P;
        P           = Q;
        Q           = t;
    }
}

constexpr uint32_t BJ_PLAIN = 0x7fffffffu;

// Partial Gram matrices G = X^H X over the row chunks [y rows_per, (y + 1) rows_per): part[(pair ks + y)][i + 32 j].
template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void bj_gram_kernel(int m, int nb, int nbp, int r, const void *W, int ks, int rows_per, void *part) {
    int P, Q;
    bj_block_pair(nbp, r, blockIdx.x, P, Q);
    if(Q >= nb) return; // the padding block of an odd block count
    using E = elem_t<R, CPLX>;
    GemmDesc      g;
    const int64_t gap = static_cast<int64_t>(Q - P) * 16 * m;
    g.M = 32, g.N = 32, g.K = m, g.batch = 1;
    g.A     = static_cast<const E *>(W) + static_cast<int64_t>(P) * 16 * m;
    g.am    = Idx2{m, gap, 16u};
    g.ak    = Idx2{1, 0, BJ_PLAIN};
    g.conjA = CPLX ? 1 : 0;
    g.B = g.A, g.bk = g.ak, g.bn = g.am;
    g.a_fast_m = 0, g.b_fast_k = 1, g.xcd_remap = 0;
    g.ksplit = ks, g.kchunk = rows_per;
    g.part   = static_cast<E *>(part) + static_cast<int64_t>(blockIdx.x) * ks * 32 * 32;
    g.C      = g.part;
    g.cm     = Idx2{1, 0, BJ_PLAIN};
    g.cn     = Idx2{32, 0, BJ_PLAIN};
    gemm_body<R, CPLX, 32, 32, 16, 2, 2, false, false>(g, 0, blockIdx.y, 0);
}

// X <- X Q in place (Q: qbuf[pair], 32 x 32 column-major) for the pairs whose inner solve rotated: row tiles
// [0, tw) of W, then those of V. A workgroup owns its 64 rows of all 32 columns and reads them (both K stages)
// before it writes them.
template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void bj_update_kernel(int m, int n, int nb, int nbp, int r, void *W, void *V, const void *qbuf,
                                                        const int *rotated, int tw) {
    int P, Q;
    bj_block_pair(nbp, r, blockIdx.x, P, Q);
    if(Q >= nb || !rotated[blockIdx.x]) return;
    using E = elem_t<R, CPLX>;
    const bool    on_w = static_cast<int>(blockIdx.y) < tw;
    const int     ld   = on_w ? m : n;
    E            *base = static_cast<E *>(on_w ? W : V) + static_cast<int64_t>(P) * 16 * ld;
    const int64_t gap  = static_cast<int64_t>(Q - P) * 16 * ld;
    GemmDesc      g;
    g.M = ld, g.N = 32, g.K = 32, g.batch = 1;
    g.A  = base;
    g.am = Idx2{1, 0, BJ_PLAIN};
    g.ak = Idx2{ld, gap, 16u};
    g.B  = static_cast<const E *>(qbuf) + static_cast<int64_t>(blockIdx.x) * 32 * 32;
    g.bk = Idx2{1, 0, BJ_PLAIN};
    g.bn = Idx2{32, 0, BJ_PLAIN};
    g.C  = base;
    g.cm = Idx2{1, 0, BJ_PLAIN};
    g.cn = Idx2{ld, gap, 16u};
    g.a_fast_m = 1, g.b_fast_k = 1, g.xcd_remap = 0, g.ksplit = 1, g.kchunk = 32;
    gemm_body<R, CPLX, 64, 32, 16, 2, 2, false, false>(g, on_w ? blockIdx.y : blockIdx.y - tw, 0, 0);
}

} // namespace

void bj_gram_mfma(dmrg_dtype t, int m, int npad, int nbp, int r, const void *W, int ks, int rows_per, void *part, hipStream_t s) {
    const dim3 grid(nbp / 2, ks), block(256);
    const int  nb = npad / 16;
    if(t == DMRG_CX64) hipLaunchKernelGGL((bj_gram_kernel<double, true>), grid, block, 0, s, m, nb, nbp, r, W, ks, rows_per, part);
    else hipLaunchKernelGGL((bj_gram_kernel<double, false>), grid, block, 0, s, m, nb, nbp, r, W, ks, rows_per, part);
}

void bj_update_mfma(dmrg_dtype t, int m, int n, int npad, int nbp, int r, void *W, void *V, const void *qbuf, const int *rotated,
                    hipStream_t s) {
    const int  tw = (m + 63) / 64, tv = (n + 63) / 64, nb = npad / 16;
    const dim3 grid(nbp / 2, tw + tv), block(256);
    if(t == DMRG_CX64)
        hipLaunchKernelGGL((bj_update_kernel<double, true>), grid, block, 0, s, m, n, nb, nbp, r, W, V, qbuf, rotated, tw);
    else hipLaunchKernelGGL((bj_update_kernel<double, false>), grid, block, 0, s, m, n, nb, nbp, r, W, V, qbuf, rotated, tw);
}

} // namespace dmrg

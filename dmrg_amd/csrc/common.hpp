// Shared host/device definitions for the dmrg_amd HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dmrg_amd.h"

namespace dmrg {

// Complex element stored interleaved (re, im), the layout of std::complex<T>.
template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R re, im;
};

// Error carrying a dmrg_status across the C++ layers. Never crosses the C-ABI:
// every extern "C" entry point catches it and returns the status.
struct Error : std::runtime_error {
    dmrg_status status;
    Error(dmrg_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};

void set_last_error(const std::string &msg);

#define DMRG_HIP_CHECK(expr)                                                                                   \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if(e_ != hipSuccess)                                                                                   \
            throw ::dmrg::Error(DMRG_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_));      \
    } while(0)

#define DMRG_REQUIRE(cond, status, msg)                                                                        \
    do {                                                                                                       \
        if(!(cond)) throw ::dmrg::Error((status), (msg));                                                      \
    } while(0)

inline size_t dtype_size(dmrg_dtype t) {
    switch(t) {
        case DMRG_FP64: return 8;
        case DMRG_CX64: return 16;
        case DMRG_FP32: return 4;
        case DMRG_CX32: return 8;
    }
    throw Error(DMRG_ERR_ARG, "unknown dtype");
}
inline bool dtype_is_complex(dmrg_dtype t) { return t == DMRG_CX64 || t == DMRG_CX32; }
inline bool dtype_is_single(dmrg_dtype t) { return t == DMRG_FP32 || t == DMRG_CX32; }
inline dmrg_dtype dtype_single_of(dmrg_dtype t) { return dtype_is_complex(t) ? DMRG_CX32 : DMRG_FP32; }
inline dmrg_dtype dtype_double_of(dmrg_dtype t) { return dtype_is_complex(t) ? DMRG_CX64 : DMRG_FP64; }

// Two-level strided index: off(i) = (i % c) * s0 + (i / c) * s1.
// A plain stride is {s, 0, INT32_MAX}. Lets one GEMM read the reference's
// column-major tensors in place when a GEMM dimension fuses two tensor indices.
struct Idx2 {
    int64_t  s0 = 0;
    int64_t  s1 = 0;
    uint32_t c  = 0x7fffffffu;
};
inline Idx2 plain(int64_t s) { return Idx2{s, 0, 0x7fffffffu}; }
inline Idx2 fused(int64_t s0, int64_t s1, int64_t c) { return Idx2{s0, s1, static_cast<uint32_t>(c)}; }

__host__ __device__ inline int64_t idx_off(const Idx2 &x, uint32_t i) {
    return static_cast<int64_t>(i % x.c) * x.s0 + static_cast<int64_t>(i / x.c) * x.s1;
}

// One batched GEMM C[b][m,n] (+)= sum_k op(A)[b][m,k] * op(B)[b][k,n].
struct GemmDesc {
    dmrg_dtype  dtype = DMRG_FP64;
    int32_t     M = 0, N = 0, K = 0, batch = 1;
    const void *A = nullptr;
    Idx2        am, ak;
    int64_t     a_batch = 0;
    int32_t     conjA   = 0;
    const void *B       = nullptr;
    Idx2        bk, bn;
    int64_t     b_batch = 0;
    int32_t     conjB   = 0;
    void       *C       = nullptr;
    Idx2        cm, cn;
    int64_t     c_batch    = 0;
    int32_t     accumulate = 0;
    // Split-K scratch supplied by the caller: room for part_capacity elements (0: no split-K).
    void   *part          = nullptr;
    int64_t part_capacity = 0;
    // Filled by the launcher.
    int32_t a_fast_m  = 1; // 1: consecutive threads walk M when loading A, 0: walk K
    int32_t b_fast_k  = 1; // 1: consecutive threads walk K when loading B, 0: walk N
    int32_t xcd_remap = 1;
    int32_t ksplit    = 1; // number of K slices (grid y)
    int32_t kchunk    = 0; // K extent of one slice (multiple of 32)
    // Profiling: if set, this hipEvent_t is recorded right after the main kernel (before a split-K reduce).
    void *ev_after_main = nullptr;
};

// Tile configuration id (see gemm.hip for the table).
// SMALLK: a column-streaming kernel for GEMMs whose K (and M) are tiny, such as the MPO leg of the matvec
// (K = d * D); it is HBM-bound and skips the tile machinery. GEMMs outside its limits fall back to T64x64.
// T32x32K: 32x32 tiles with one deep K stage per step (128 elements of 4 bytes, 64 of 8 or 16 bytes) for GEMMs too
// small to fill the GPU, where a K step's MFMA work is shorter than a global load's latency.
enum class TileCfg : int { T64x64 = 0, T32x64 = 1, T64x32 = 2, T16x128 = 3, T128x64 = 4, T32x32 = 5, SMALLK = 6, T32x32K = 7, AUTO = -1 };
constexpr int SMALLK_MAX_K = 32;
constexpr int SMALLK_MAX_M = 64;

// Launch one GEMM. cfg/ksplit = AUTO/0 pick from the tuning table or the size heuristic.
void gemm(const GemmDesc &g, hipStream_t stream, TileCfg cfg = TileCfg::AUTO, int ksplit = 0);
// While an object of this type lives, gemm() on this thread never times candidate plans (no events, no
// synchronisation): it launches a plan already tuned for the signature, else the size heuristic's.
struct NoTuneScope {
    NoTuneScope();
    ~NoTuneScope();
    NoTuneScope(const NoTuneScope &)            = delete;
    NoTuneScope &operator=(const NoTuneScope &) = delete;

  private:
    bool prev_;
};

// Grouped GEMM: independent GEMMs (batch 1, no split-K) of one tile configuration run as one launch. The
// descriptors and the tile prefix sums are device arrays filled by grouped_prepare (one group per
// (swap, aligned) kind of GEMM); they stay valid as long as the GEMMs' buffers do, so a grouped launch can be
// captured into a graph and replayed. Tile configurations: T64x64, T32x32, T16x128.
struct GroupedGemm {
    dmrg_dtype dtype   = DMRG_FP64;
    TileCfg    cfg     = TileCfg::T64x64;
    bool       swap    = false;
    bool       aligned = false;
    int        ngemm   = 0;
    int        tiles   = 0;
    void      *descs   = nullptr; // device, ngemm GemmDesc
    int       *starts  = nullptr; // device, ngemm + 1 tile offsets
    bool       owned   = true;    // descs/starts allocated by grouped_prepare (else inside a caller's store)
};
// Tiles of g under configuration cfg.
int grouped_tiles(const GemmDesc &g, TileCfg cfg);
// Whether GEMM plans are tuned by timing (dmrg_autotune_enable).
int autotune_enabled();

// Block one-sided Jacobi SVD steps (FP64 or CX64) on W (m x npad, ld m) and V (n x npad, ld n), npad a multiple
// of 16, for the block pairs of circle-method round r over nbp blocks of 16 columns: partial Gram matrices
// part[(pair ks + chunk)][32 x 32] over row chunks of rows_per rows (a multiple of 32), and X <- X qbuf[pair] on
// both matrices for the pairs flagged in rotated.
void bj_gram_mfma(dmrg_dtype t, int m, int npad, int nbp, int r, const void *W, int ks, int rows_per, void *part, hipStream_t s);
void bj_update_mfma(dmrg_dtype t, int m, int n, int npad, int nbp, int r, void *W, void *V, const void *qbuf, const int *rotated,
                    hipStream_t s);
// Builds the launch groups of gs; device storage for the descriptors is allocated with hipMalloc and owned by
// the returned groups (release with grouped_free). With store (a device buffer of store_bytes, grown by the
// caller to at least grouped_store_bytes(gs.size())), the descriptors are written into it instead and nothing
// is allocated; the store must not be in use by queued launches.
std::vector<GroupedGemm> grouped_prepare(const std::vector<GemmDesc> &gs, TileCfg cfg, void *store = nullptr,
                                         size_t store_bytes = 0);
size_t                   grouped_store_bytes(size_t ngemm);
void                     grouped_launch(const GroupedGemm &g, hipStream_t s);
void                     grouped_free(std::vector<GroupedGemm> &gg);
int  device_cu_count();
// Autotuner controls: on/off, a forced plan for every GEMM (cfg < 0 clears it), cache reset and size.
void set_autotune(int on);
void set_tune_min_macs(int64_t macs);
void set_plan_override(int cfg, int ksplit, int nsplit = 0, int ksplit2 = 1);
void clear_tuning();
int  tuned_entries();

} // namespace dmrg

// Device-resident H_eff operator: the GPU counterpart of MatVecMPOS<T> (matvec_mpos.h:23-180).
#pragma once
#include "contract.hpp"

#include <algorithm>
#include <complex>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace dmrg {

// Device memory recycled by size class (powers of two and 1.5 times them, from 4 KiB): blocks go back to the pool
// instead of to hipFree, which would synchronise the whole device (and with it every other stream's work). A
// block handed out again is reused by work that is ordered after the work of its previous holder (the same
// stream, or after a host synchronisation); every pool below is used that way.
struct DevPool {
    DevPool() = default;
    DevPool(const DevPool &)            = delete;
    DevPool &operator=(const DevPool &) = delete;
    ~DevPool();
    static size_t cls(size_t b);
    void         *get(size_t b, size_t &c); // a block of class c >= b
    void          put(void *p, size_t c);
    size_t        held_bytes() const;

  private:
    mutable std::mutex                    mu_;
    std::map<size_t, std::vector<void *>> free_;
    std::vector<std::pair<void *, size_t>> all_;
};

// The pool of the calling thread (null: none). While a DevPoolScope lives on a thread, DevBufs that grow on it take
// their blocks from a pool of that thread, and a block goes back to the pool it came from when its DevBuf grows
// again or dies; the pool frees its memory when the scope has ended and the last DevBuf holding one of its blocks
// is gone. For threads that issue all their device work on one stream (the request loop's workers): their
// buffers then grow and die without any hipFree, hence without device-wide synchronisation.
std::shared_ptr<DevPool> &thread_dev_pool();
struct DevPoolScope {
    DevPoolScope();
    ~DevPoolScope();
    DevPoolScope(const DevPoolScope &)            = delete;
    DevPoolScope &operator=(const DevPoolScope &) = delete;

  private:
    std::shared_ptr<DevPool> prev_;
};

// Owning device buffer (grow-only).
struct DevBuf {
    void  *p     = nullptr;
    size_t bytes = 0;
    DevBuf()     = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), pool_(std::move(o.pool_)), cls_(o.cls_) { o.p = nullptr, o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(pool_, o.pool_);
        std::swap(cls_, o.cls_);
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if(p) {
            if(pool_) pool_->put(p, cls_);
            else (void) hipFree(p);
        }
        p = nullptr, bytes = 0, cls_ = 0;
        pool_.reset();
    }
    // Grow-only; growth is geometric (x1.5) so that buffers that follow a growing bond dimension are
    // reallocated O(log chi) times. Without a thread pool the old block is freed after a device synchronisation
    // (queued kernels may still read it; hipFree synchronises the device anyway); with one it goes back to its
    // pool, to be reused by later work of the same stream.
    void *reserve(size_t b) {
        if(b > bytes) {
            const size_t nb = std::max(b, bytes + bytes / 2);
            if(std::shared_ptr<DevPool> tp = thread_dev_pool()) {
                size_t c  = 0;
                void  *np = tp->get(nb, c);
                if(p) {
                    if(pool_) pool_->put(p, cls_);
                    else {
                        (void) hipDeviceSynchronize();
                        (void) hipFree(p);
                    }
                }
                p = np, bytes = c, cls_ = c, pool_ = std::move(tp);
                return p;
            }
            if(p && !pool_) (void) hipDeviceSynchronize();
            release();
            if(hipMalloc(&p, nb) != hipSuccess) {
                (void) hipGetLastError();
                p = nullptr;
                throw Error(DMRG_ERR_OOM, "device allocation of " + std::to_string(nb) + " bytes failed");
            }
            bytes = nb;
        }
        return p;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }

  private:
    std::shared_ptr<DevPool> pool_; // the pool the block came from (null: hipMalloc)
    size_t                   cls_ = 0;
};

// Pinned host blocks cached per thread (hipHostFree synchronises the device): get returns a block of at least
// bytes, put hands it back to the calling thread's cache.
void *pinned_get(size_t bytes);
void  pinned_put(void *p, size_t bytes);

// Merge a chain of MPOs (cx64, each (DL_i, DR_i, d_i, d_i)) into one (DL_0, DR_last, prod d, prod d) with the
// first site's physical index fastest (matvec_mpos.cpp:104-119: contract {1},{0}, shuffle {0,3,1,4,2,5}).
std::vector<std::complex<double>> merge_mpos(int nmpo, const std::complex<double> *const *mpos, const int64_t *dims,
                                             int64_t out_dims[4]);

struct MatVecOp {
    dmrg_dtype                         dtype;
    Dims                               D;
    int                                nmpo_in = 1;
    std::vector<std::complex<double>>  mpo_host; // merged MPO in cx64 (shift applied), source of the device copies
    DevBuf                             mpo, envL, envR;       // operator dtype (owned copies)
    const void                        *envL_p = nullptr, *envR_p = nullptr; // env blocks used: owned or borrowed views
    DevBuf                             mpo32, envL32, envR32; // FP32 shadows for the mixed path
    bool                               have32 = false;
    DevBuf                             ws, ws32, xbuf, ybuf, diagws, diagc;
    bool                               diag_valid = false; // diagc holds the current operator's diagonal
    // More than 5 MPOs (matvec_mpos.cpp:120-123): the matvec applies them one at a time (matvec_chain); the merged
    // MPO above still serves the diagonal, the dense matrix and the block preconditioner.
    bool                               chain = false;
    std::vector<std::vector<std::complex<double>>> chain_host; // per MPO, cx64, shift applied
    std::vector<int64_t>               chain_dims;            // 4 per MPO
    std::vector<DevBuf>                chain_dev, chain_dev32;
    // When set, MPO uploads (rebind, set_shift) are enqueued on this stream from a pinned staging block instead of
    // blocking copies; the operator must then be used on this stream.
    hipStream_t                        upload_stream = nullptr;
    void                              *staging       = nullptr;
    size_t                             staging_bytes = 0;
    hipEvent_t                         staging_ev    = nullptr;
    std::complex<double>               sigma{0.0, 0.0};
    bool                               ready_shift = false;
    int64_t                            num_mv      = 0;
    double                             flops       = 0.0; // algorithmic FLOP of every matvec applied (8 per complex MAC)

    MatVecOp(dmrg_dtype t, int nmpo, const void *const *mpos, const int64_t *mpo_dims, const void *envL_h,
             const int64_t envL_dims[3], const void *envR_h, const int64_t envR_dims[3]);
    // Device-view variant: env blocks are device arrays of dtype t that stay owned by the caller and must outlive
    // the operator; the MPOs (host, cx64) are merged and copied.
    MatVecOp(dmrg_dtype t, int nmpo, const void *const *mpos, const int64_t *mpo_dims, const void *envL_d,
             const int64_t envL_dims[3], const void *envR_d, const int64_t envR_dims[3], bool device_envs);
    // Re-point a device-view operator at new MPOs and env blocks; workspaces are kept (grow-only).
    void rebind(int nmpo, const void *const *mpos, const int64_t *mpo_dims, const void *envL_d, const int64_t envL_dims[3],
                const void *envR_d, const int64_t envR_dims[3]);

    // Multiply-accumulates of one matvec (the chain's own count when the MPOs are applied one at a time).
    double macs() const;
    std::vector<ChainMpo> chain_view(bool fp32) const;
    void upload_mpo();
    // Merged MPO, dims, and (for more than 5 MPOs) the chain copies, from host cx64 data.
    void load_mpos(int nmpo, const void *const *mpos, const int64_t *mpo_dims, const int64_t envL_dims[3],
                   const int64_t envR_dims[3]);
    void ensure_fp32(hipStream_t s);
    void multax_device(const void *dx, int64_t ldx, void *dy, int64_t ldy, int nvec, hipStream_t s, bool fp32);
    void multax_host(const void *x, int64_t ldx, void *y, int64_t ldy, int nvec);
    void set_shift(std::complex<double> shift);
    void diagonal(void *out, hipStream_t s);
    // Preconditioner on host buffers: Jacobi y[:, i] = x[:, i] / (diag(H_eff) - shifts[i]) when jcb == 1,
    // block-Jacobi (below) when jcb > 1.
    void precond_host(const void *x, int64_t ldx, void *y, int64_t ldy, int nvec, const double *shifts);

    // Block-Jacobi preconditioner (jcbMaxBlockSize > 1, matvec_mpos.cpp:1086-1331 with LU factorisation): the
    // inverses of the diagonal blocks of (H_eff - shift) at offsets k*jcb, built once at the first use after the
    // operator, its shift or the block size changed (readyCalcPc), with the shift of that first use.
    int64_t jcb      = 1;
    bool    pc_ready = false;
    double  pc_shift = 0.0;
    DevBuf  pcinv, pcx, pcy, pcpiv, pcinfo, pcchol, pcneg;
    int     pc_fact = 3; // eig::Factorization of the blocks: 1 LLT (sign rule of the reference), 3 LU (also LDLT, QR)
    void    set_factorization(int f);
    void   *blas = nullptr; // rocblas_handle, created on first use
    void    set_jcb(int64_t size);
    void    block_pc_build(double shift, hipStream_t s);
    // y = P x for one vector in the operator's dtype (device buffers of n elements).
    void    block_pc_apply(const void *x, void *y, hipStream_t s);
    // Shift-invert (MatVecMPOS::FactorOP / MultOPv, matvec_mpos.cpp:887-975): the dense H_eff of the operator as it
    // stands (with its shift), factored on the device; op_fact is the eig::Factorization applied (0 NONE: marked
    // ready without a factorization, as the reference does), op_ready false until factor_op and after set_shift.
    bool    op_ready = false;
    int     op_fact  = 0;
    int64_t num_op   = 0;
    DevBuf  opA, oppiv, opinfo, opB_rhs;
    void    factor_op(int f, hipStream_t s);
    // y[:, i] = (H_eff)^-1 x[:, i] for nvec host vectors with the factorization of factor_op.
    void    mult_opv_host(const void *x, int64_t ldx, void *y, int64_t ldy, int nvec);
    // Second operator for MultBx (mpos_B, envL_B, envR_B; matvec_mpos.h:47), null when none was given.
    std::unique_ptr<MatVecOp> opB;

    ~MatVecOp();
    void time_legs(const void *dx, void *dy, int nrep, bool fp32, float ms[4], hipStream_t s);
};

} // namespace dmrg

// C-ABI entry points: runtime helpers, stateless contractions and the MatVecMPOS operator.
// No C++ exception crosses this boundary: each entry point converts errors to a dmrg_status
// and records the message for dmrg_last_error().
#include "contract.hpp"
#include "matvec_op.hpp"
#include "dense_eig.hpp"
#include "eigs.hpp"
#include "expand.hpp"
#include "group.hpp"
#include "batch_eigs.hpp"

#include <chrono>
#include "sweep.hpp"
#include "dense.hpp"
#include "svd.hpp"
#include "vecops.hpp"

#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace dmrg {

static thread_local std::string g_last_error;
void                            set_last_error(const std::string &msg) { g_last_error = msg; }

template <typename F>
static dmrg_status guard(F &&f) {
    try {
        f();
        return DMRG_OK;
    } catch(const Error &e) {
        set_last_error(e.what());
        return e.status;
    } catch(const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return DMRG_ERR_OOM;
    } catch(const std::exception &e) {
        set_last_error(e.what());
        return DMRG_ERR_ARG;
    } catch(...) {
        set_last_error("unknown error");
        return DMRG_ERR_ARG;
    }
}

// Serialises the stateless entry points that share the per-device workspace.
static std::mutex g_stateless_mu;

// Held for the duration of one stateless call: takes the lock, makes the call's stream wait for the
// previous stateless call's kernels on this device (they may still be using the shared workspace),
// and marks this call's end on its stream. Stateless calls on different streams therefore run one
// after the other on the device; operators (dmrg_matvec) own their workspaces and are not affected.
class StatelessScope {
  public:
    explicit StatelessScope(hipStream_t s) : lock_(g_stateless_mu), s_(s) {
        int dev = 0;
        DMRG_HIP_CHECK(hipGetDevice(&dev));
        static std::vector<hipEvent_t> evs;
        if(static_cast<int>(evs.size()) <= dev) evs.resize(dev + 1, nullptr);
        if(!evs[dev]) DMRG_HIP_CHECK(hipEventCreateWithFlags(&evs[dev], hipEventDisableTiming));
        ev_ = evs[dev];
        DMRG_HIP_CHECK(hipStreamWaitEvent(s_, ev_, 0));
    }
    ~StatelessScope() { (void) hipEventRecord(ev_, s_); }
    StatelessScope(const StatelessScope &)            = delete;
    StatelessScope &operator=(const StatelessScope &) = delete;

  private:
    std::lock_guard<std::mutex> lock_;
    hipStream_t                 s_;
    hipEvent_t                  ev_ = nullptr;
};

// Small per-device buffer for reduction results.
static double *reduction_buffer() {
    static std::mutex            mu;
    static std::vector<double *> bufs;
    std::lock_guard<std::mutex>  lock(mu);
    int                          dev = 0;
    DMRG_HIP_CHECK(hipGetDevice(&dev));
    if(static_cast<int>(bufs.size()) <= dev) bufs.resize(dev + 1, nullptr);
    if(!bufs[dev]) {
        void *p = nullptr;
        if(hipMalloc(&p, DOT_BUFFER_DOUBLES * sizeof(double)) != hipSuccess) throw Error(DMRG_ERR_OOM, "reduction buffer");
        bufs[dev] = static_cast<double *>(p);
    }
    return bufs[dev];
}

// rocSOLVER-backed SVD / eigensolver objects, one per (host thread, device, stream); each owns a rocBLAS
// handle and grow-only work buffers, released when the thread ends.
static DeviceSvd &solver_for(hipStream_t s) {
    thread_local std::map<std::pair<int, hipStream_t>, std::unique_ptr<DeviceSvd>> cache;
    int dev = 0;
    DMRG_HIP_CHECK(hipGetDevice(&dev));
    auto &slot = cache[{dev, s}];
    if(!slot) slot = std::make_unique<DeviceSvd>(s);
    slot->set_stream(s);
    return *slot;
}

} // namespace dmrg

using namespace dmrg;

namespace {
dmrg::EigsParams batched_params(const dmrg_eigs_config *cfg) {
    DMRG_REQUIRE(cfg->nev == 1 && cfg->precond == 0 && (cfg->target == DMRG_TARGET_SA || cfg->target == DMRG_TARGET_LA),
                 DMRG_ERR_UNSUPPORTED, "batched eigs: nev 1, no preconditioner, target SA or LA");
    DMRG_REQUIRE(cfg->tol > 0, DMRG_ERR_ARG, "tol must be positive");
    DMRG_REQUIRE(cfg->ncv <= 0 || (cfg->ncv >= 4 && cfg->ncv <= 64), DMRG_ERR_ARG, "batched eigs: ncv in [4, 64] or 0");
    EigsParams p;
    p.nev         = 1;
    p.ncv         = cfg->ncv;
    p.target      = cfg->target;
    p.tol         = std::max(cfg->tol, std::numeric_limits<double>::epsilon());
    p.max_matvecs = cfg->max_matvecs > 0 ? cfg->max_matvecs : 100000;
    p.seed        = cfg->seed;
    // Used by the problems that dmrg_eigs_batched solves alone in mixed precision.
    DMRG_REQUIRE(cfg->inner_max >= 0 && cfg->inner_tol > 0 && cfg->inner_tol < 1, DMRG_ERR_ARG, "bad inner_max / inner_tol");
    p.inner_max = cfg->mixed ? cfg->inner_max : 0;
    p.inner_tol = cfg->inner_tol;
    return p;
}
// Times the eigensolver's block-vector kernels on an n x (2m + 4) working set of its own, filled with random data.
template <typename T>
void time_block_vectors(int64_t n, int m, int nrep, float ms[3], double bytes[3]) {
    const size_t es = sizeof(T);
    DevBuf       buf, scr, dev;
    buf.reserve(static_cast<size_t>(n) * (2 * m + 4) * es);
    scr.reserve(static_cast<size_t>(vec_scratch_doubles()) * sizeof(double));
    dev.reserve(sizeof(double) * (4 * VEC_MAX_M + 8));
    T      *V = buf.as<T>(), *W = V + n * m, *x = W + n * m, *r = x + n, *y = r + n;
    double *d = dev.as<double>();
    hipStream_t s = nullptr;
    hipEvent_t  e0 = nullptr, e1 = nullptr;
    DMRG_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    try {
        DMRG_HIP_CHECK(hipEventCreate(&e0));
        DMRG_HIP_CHECK(hipEventCreate(&e1));
        random_fill<T>(n * (2 * m + 4), V, 7, s);
        vec_scratch_init(scr.as<double>(), s);
        std::vector<double> h(2 * VEC_MAX_M, 1.0 / m);
        DMRG_HIP_CHECK(hipMemcpyAsync(d, h.data(), sizeof(double) * 2 * m, hipMemcpyHostToDevice, s));
        auto timed = [&](auto &&launch) {
            launch(); // warm-up
            DMRG_HIP_CHECK(hipEventRecord(e0, s));
            for(int k = 0; k < nrep; ++k) launch();
            DMRG_HIP_CHECK(hipEventRecord(e1, s));
            DMRG_HIP_CHECK(hipEventSynchronize(e1));
            float t = 0;
            DMRG_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
            return t / nrep;
        };
        // projections c = V^H x: reads m + 1 vectors
        ms[0]    = timed([&] { multidot<T>(n, m, V, n, x, d + 2 * VEC_MAX_M, scr.as<double>(), s); });
        bytes[0] = static_cast<double>(m + 1) * n * es;
        // update y = y - V c: reads m + 1, writes 1
        ms[1]    = timed([&] { gemv_update<T>(n, m, V, n, d, -1.0, 1.0, y, s); });
        bytes[1] = static_cast<double>(m + 2) * n * es;
        // Ritz vector and residual x = V y, r = W y - theta x: reads 2m, writes 2
        ms[2]    = timed([&] { ritz_residual<T>(n, m, V, W, n, d, 0.5, x, r, d + 2 * VEC_MAX_M, scr.as<double>(), s); });
        bytes[2] = static_cast<double>(2 * m + 2) * n * es;
    } catch(...) {
        if(e0) (void) hipEventDestroy(e0);
        if(e1) (void) hipEventDestroy(e1);
        (void) hipStreamSynchronize(s);
        (void) hipStreamDestroy(s);
        throw;
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
    DMRG_HIP_CHECK(hipStreamDestroy(s));
}

dmrg_eigs_result batched_result(const BatchEigsResult &r, double sec) {
    dmrg_eigs_result o{};
    o.eigval[0]    = r.eigval;
    o.rnorm[0]     = r.rnorm;
    o.nconv        = r.nconv;
    o.iters        = r.iters;
    o.matvecs_fp64 = r.matvecs;
    o.matvecs_fp32 = r.matvecs_fp32;
    o.anorm        = r.anorm;
    o.seconds      = sec;
    return o;
}
} // namespace

extern "C" {

const char *dmrg_last_error(void) { return g_last_error.c_str(); }
int         dmrg_version(void) { return 100; }

dmrg_status dmrg_device_count(int *count) {
    return guard([&] {
        DMRG_REQUIRE(count, DMRG_ERR_ARG, "null count");
        int n = 0;
        if(hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

dmrg_status dmrg_set_device(int device) { return guard([&] { DMRG_HIP_CHECK(hipSetDevice(device)); }); }

dmrg_status dmrg_device_info(char *name, size_t name_len, int *cu_count, size_t *hbm_bytes) {
    return guard([&] {
        int dev = 0;
        DMRG_HIP_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        DMRG_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        if(name && name_len) {
            std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
        }
        if(cu_count) *cu_count = prop.multiProcessorCount;
        if(hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    });
}

dmrg_status dmrg_malloc(void **dptr, size_t bytes) {
    return guard([&] {
        DMRG_REQUIRE(dptr, DMRG_ERR_ARG, "null pointer");
        *dptr = nullptr;
        if(bytes == 0) return;
        if(hipMalloc(dptr, bytes) != hipSuccess) {
            (void) hipGetLastError();
            throw Error(DMRG_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
    });
}
dmrg_status dmrg_free(void *dptr) {
    return guard([&] {
        if(dptr) DMRG_HIP_CHECK(hipFree(dptr));
    });
}
dmrg_status dmrg_memcpy_htod(void *dst, const void *src, size_t bytes) {
    return guard([&] {
        if(bytes) DMRG_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    });
}
dmrg_status dmrg_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    return guard([&] {
        if(bytes) DMRG_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}
dmrg_status dmrg_memcpy_dtod(void *dst, const void *src, size_t bytes) {
    return guard([&] {
        if(bytes) DMRG_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    });
}
dmrg_status dmrg_memset(void *dptr, int value, size_t bytes) {
    return guard([&] {
        if(bytes) DMRG_HIP_CHECK(hipMemset(dptr, value, bytes));
    });
}
dmrg_status dmrg_synchronize(void) { return guard([&] { DMRG_HIP_CHECK(hipDeviceSynchronize()); }); }

dmrg_status dmrg_stream_create(void **stream) {
    return guard([&] {
        hipStream_t s = nullptr;
        DMRG_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}
dmrg_status dmrg_stream_destroy(void *stream) {
    return guard([&] {
        if(stream) DMRG_HIP_CHECK(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    });
}
dmrg_status dmrg_stream_synchronize(void *stream) {
    return guard([&] { DMRG_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream))); });
}
dmrg_status dmrg_event_create(void **event) {
    return guard([&] {
        hipEvent_t e = nullptr;
        DMRG_HIP_CHECK(hipEventCreate(&e));
        *event = e;
    });
}
dmrg_status dmrg_event_destroy(void *event) {
    return guard([&] {
        if(event) DMRG_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(event)));
    });
}
dmrg_status dmrg_event_record(void *event, void *stream) {
    return guard([&] { DMRG_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream))); });
}
dmrg_status dmrg_event_elapsed_ms(float *ms, void *start, void *stop) {
    return guard([&] {
        DMRG_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
        DMRG_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    });
}

// ------------------------------------------------------------------------------- contractions

dmrg_status dmrg_matrix_vector_product(dmrg_dtype dtype, void *res, const void *mps, const int64_t mps_dims[3],
                                       const void *mpo, const int64_t mpo_dims[4], const void *envL,
                                       const int64_t envL_dims[3], const void *envR, const int64_t envR_dims[3],
                                       void *stream) {
    return guard([&] {
        DMRG_REQUIRE(res && mps && mpo && envL && envR && mps_dims && mpo_dims && envL_dims && envR_dims, DMRG_ERR_ARG,
                     "null argument");
        const Dims                  D = check_matvec_dims(mps_dims, mpo_dims, envL_dims, envR_dims);
        StatelessScope scope(static_cast<hipStream_t>(stream));
        void *ws = scratch_workspace(static_cast<size_t>(matvec_ws_elems(D)) * dtype_size(dtype));
        matvec(dtype, D, mps, D.n(), res, D.n(), 1, mpo, envL, envR, ws, static_cast<hipStream_t>(stream));
    });
}

dmrg_status dmrg_env_build_block(dmrg_dtype dtype, char side, void *block_out, const void *otherblock,
                                 const int64_t other_dims[3], const void *mps, const int64_t mps_dims[3],
                                 const void *mpo, const int64_t mpo_dims[4], void *stream) {
    return guard([&] {
        DMRG_REQUIRE(block_out && otherblock && mps && mpo && other_dims && mps_dims && mpo_dims, DMRG_ERR_ARG, "null argument");
        DMRG_REQUIRE(side == 'L' || side == 'R', DMRG_ERR_ARG, "side must be 'L' or 'R'");
        DMRG_REQUIRE(mps_dims[0] == mpo_dims[2] && mpo_dims[2] == mpo_dims[3], DMRG_ERR_DIMS, "env: mps dim 0 != mpo dim 2");
        DMRG_REQUIRE(other_dims[0] == other_dims[1], DMRG_ERR_DIMS, "env: block must be square in its bond indices");
        Dims D;
        D.d    = mps_dims[0];
        D.chiL = mps_dims[1];
        D.chiR = mps_dims[2];
        D.DL   = mpo_dims[0];
        D.DR   = mpo_dims[1];
        if(side == 'L') {
            // EnvBase.cpp:100-108
            DMRG_REQUIRE(mps_dims[1] == other_dims[0], DMRG_ERR_DIMS, "envL: mps dim 1 != left-block dim 0");
            DMRG_REQUIRE(mpo_dims[0] == other_dims[2], DMRG_ERR_DIMS, "envL: mpo dim 0 != left-block dim 2");
        } else {
            // EnvBase.cpp:137-145
            DMRG_REQUIRE(mps_dims[2] == other_dims[0], DMRG_ERR_DIMS, "envR: mps dim 2 != right-block dim 0");
            DMRG_REQUIRE(mpo_dims[1] == other_dims[2], DMRG_ERR_DIMS, "envR: mpo dim 1 != right-block dim 2");
        }
        StatelessScope scope(static_cast<hipStream_t>(stream));
        void *ws = scratch_workspace(static_cast<size_t>(env_ws_elems(D)) * dtype_size(dtype));
        env_build_block(dtype, side, D, block_out, otherblock, mps, mpo, ws, static_cast<hipStream_t>(stream));
    });
}

namespace dmrg {
namespace {
// Real-cast branch of the expansion term (EnvBase.cpp:305-307): helpers on complex double arrays.
__global__ void any_imag_kernel(int64_t n, const cpx<double> *x, int *flag) {
    bool nz = false;
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        nz |= x[i].im != 0.0;
    if(nz) atomicOr(flag, 1);
}
__global__ void real_part_kernel(int64_t n, const cpx<double> *x, double *y) {
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        y[i] = x[i].re;
}
__global__ void cx_from_real_kernel(int64_t n, const double *x, double alpha, cpx<double> *y) {
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        y[i] = cpx<double>{alpha * x[i], 0.0};
}
dim3 ew_blocks(int64_t n) { return dim3(static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)))); }
} // namespace
} // namespace dmrg

dmrg_status dmrg_env_expansion_term(dmrg_dtype dtype, char side, void *P, const void *block, const int64_t block_dims[3],
                                    const void *mps, const int64_t mps_dims[3], const void *mpo, const int64_t mpo_dims[4],
                                    double alpha, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(P && block && mps && mpo && block_dims && mps_dims && mpo_dims, DMRG_ERR_ARG, "null argument");
        DMRG_REQUIRE(side == 'L' || side == 'R', DMRG_ERR_ARG, "side must be 'L' or 'R'");
        DMRG_REQUIRE(!std::isnan(alpha), DMRG_ERR_ARG, "the mixing factor alpha is nan");
        DMRG_REQUIRE(mps_dims[0] == mpo_dims[2] && mpo_dims[2] == mpo_dims[3], DMRG_ERR_DIMS, "expansion: mps dim 0 != mpo dim 2");
        DMRG_REQUIRE(block_dims[0] == block_dims[1], DMRG_ERR_DIMS, "expansion: block must be square in its bond indices");
        if(side == 'L') {
            DMRG_REQUIRE(block_dims[0] == mps_dims[1], DMRG_ERR_DIMS, "expansion L: block dim 0 != mps dim 1");
            DMRG_REQUIRE(block_dims[2] == mpo_dims[0], DMRG_ERR_DIMS, "expansion L: block dim 2 != mpo dim 0");
        } else {
            DMRG_REQUIRE(block_dims[0] == mps_dims[2], DMRG_ERR_DIMS, "expansion R: block dim 0 != mps dim 2");
            DMRG_REQUIRE(block_dims[2] == mpo_dims[1], DMRG_ERR_DIMS, "expansion R: block dim 2 != mpo dim 1");
        }
        Dims D;
        D.d = mps_dims[0], D.chiL = mps_dims[1], D.chiR = mps_dims[2], D.DL = mpo_dims[0], D.DR = mpo_dims[1];
        const int64_t nP = D.d * D.chiL * D.chiR * (side == 'L' ? D.DR : D.DL);
        auto          s  = static_cast<hipStream_t>(stream);
        StatelessScope scope(s);
        if(nP == 0) return;
        if(alpha == 0.0) {
            DMRG_HIP_CHECK(hipMemsetAsync(P, 0, static_cast<size_t>(nP) * dtype_size(dtype), s));
            return;
        }
        const int64_t nX = D.d * D.chiL * D.chiR * (side == 'L' ? D.DL : D.DR); // the contraction's intermediate
        if(dtype == DMRG_CX64) {
            // EnvBase.cpp:305-307: when mps and mpo are real, the term is contracted in real arithmetic with the
            // real part of the env block and cast back to complex.
            const int64_t nB = block_dims[0] * block_dims[1] * block_dims[2], nM = D.n(), nW = D.DL * D.DR * D.d * D.d;
            double       *w0 = static_cast<double *>(scratch_workspace(static_cast<size_t>(2 + nB + nM + nW + nP + nX) * sizeof(double)));
            int          *flag = reinterpret_cast<int *>(w0);
            DMRG_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
            hipLaunchKernelGGL(any_imag_kernel, ew_blocks(nM), dim3(256), 0, s, nM, static_cast<const cpx<double> *>(mps), flag);
            hipLaunchKernelGGL(any_imag_kernel, ew_blocks(nW), dim3(256), 0, s, nW, static_cast<const cpx<double> *>(mpo), flag);
            DMRG_HIP_CHECK(hipGetLastError());
            int host_flag = 1;
            DMRG_HIP_CHECK(hipMemcpyAsync(&host_flag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
            DMRG_HIP_CHECK(hipStreamSynchronize(s));
            if(host_flag == 0) {
                double *Br = w0 + 2, *Mr = Br + nB, *Wr = Mr + nM, *Pr = Wr + nW, *Xr = Pr + nP;
                hipLaunchKernelGGL(real_part_kernel, ew_blocks(nB), dim3(256), 0, s, nB, static_cast<const cpx<double> *>(block), Br);
                hipLaunchKernelGGL(real_part_kernel, ew_blocks(nM), dim3(256), 0, s, nM, static_cast<const cpx<double> *>(mps), Mr);
                hipLaunchKernelGGL(real_part_kernel, ew_blocks(nW), dim3(256), 0, s, nW, static_cast<const cpx<double> *>(mpo), Wr);
                DMRG_HIP_CHECK(hipGetLastError());
                env_expansion_term(DMRG_FP64, side, D, Pr, Br, Mr, Wr, Xr, s);
                hipLaunchKernelGGL(cx_from_real_kernel, ew_blocks(nP), dim3(256), 0, s, nP, Pr, alpha, static_cast<cpx<double> *>(P));
                DMRG_HIP_CHECK(hipGetLastError());
                return;
            }
        }
        void *ws = scratch_workspace(static_cast<size_t>(nX) * dtype_size(dtype));
        env_expansion_term(dtype, side, D, P, block, mps, mpo, ws, s);
        if(alpha != 1.0) {
            switch(dtype) {
                case DMRG_FP64: scale<double>(nP, static_cast<double *>(P), alpha, s); break;
                case DMRG_CX64: scale<cpx<double>>(nP, static_cast<cpx<double> *>(P), alpha, s); break;
                case DMRG_FP32: scale<float>(nP, static_cast<float *>(P), alpha, s); break;
                case DMRG_CX32: scale<cpx<float>>(nP, static_cast<cpx<float> *>(P), alpha, s); break;
                default: throw Error(DMRG_ERR_ARG, "bad dtype");
            }
        }
    });
}

dmrg_status dmrg_optimally_mixed_block(dmrg_matvec *h1, dmrg_matvec *h2, const void *theta, int mode, void *mixed,
                                       double alphas[3], double *opt_val, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(h1 && h2 && theta && mixed && alphas, DMRG_ERR_ARG, "null argument");
        MixedBlockResult r;
        optimally_mixed_block(*reinterpret_cast<MatVecOp *>(h1), *reinterpret_cast<MatVecOp *>(h2), theta, mode, mixed, r,
                              static_cast<hipStream_t>(stream));
        for(int i = 0; i < 3; ++i) alphas[i] = r.alpha[i];
        if(opt_val) *opt_val = r.opt_val;
    });
}

dmrg_status dmrg_dot(dmrg_dtype dtype, int64_t n, const void *x, const void *y, double result[2], void *stream) {
    return guard([&] {
        DMRG_REQUIRE(result && (n == 0 || (x && y)) && n >= 0, DMRG_ERR_ARG, "bad argument");
        StatelessScope scope(static_cast<hipStream_t>(stream));
        double *buf = reduction_buffer();
        auto    s   = static_cast<hipStream_t>(stream);
        dot_device(dtype, n, x, y, buf, s);
        DMRG_HIP_CHECK(hipMemcpyAsync(result, buf, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        if(!dtype_is_complex(dtype)) result[1] = 0.0;
    });
}

dmrg_status dmrg_nrm2(dmrg_dtype dtype, int64_t n, const void *x, double *result, void *stream) {
    return guard([&] {
        double r[2];
        StatelessScope scope(static_cast<hipStream_t>(stream));
        double *buf = reduction_buffer();
        auto    s   = static_cast<hipStream_t>(stream);
        nrm2sq_device(dtype, n, x, buf, s);
        DMRG_HIP_CHECK(hipMemcpyAsync(r, buf, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        *result = std::sqrt(std::max(0.0, r[0]));
    });
}

dmrg_status dmrg_expectation_value(dmrg_dtype dtype, const void *mps, const int64_t mps_dims[3], const void *mpo,
                                   const int64_t mpo_dims[4], const void *envL, const int64_t envL_dims[3],
                                   const void *envR, const int64_t envR_dims[3], double result[2], void *stream) {
    return guard([&] {
        DMRG_REQUIRE(result && mps && mpo && envL && envR, DMRG_ERR_ARG, "null argument");
        const Dims                  D = check_matvec_dims(mps_dims, mpo_dims, envL_dims, envR_dims);
        StatelessScope scope(static_cast<hipStream_t>(stream));
        const size_t es = dtype_size(dtype);
        char *ws = static_cast<char *>(scratch_workspace(static_cast<size_t>(matvec_ws_elems(D) + D.n()) * es));
        void *hv = ws + static_cast<size_t>(matvec_ws_elems(D)) * es;
        auto  s  = static_cast<hipStream_t>(stream);
        matvec(dtype, D, mps, D.n(), hv, D.n(), 1, mpo, envL, envR, ws, s);
        double *buf = reduction_buffer();
        dot_device(dtype, D.n(), mps, hv, buf, s);
        DMRG_HIP_CHECK(hipMemcpyAsync(result, buf, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        if(!dtype_is_complex(dtype)) result[1] = 0.0;
    });
}

// ------------------------------------------------------------------------------- MatVecMPOS

dmrg_status dmrg_matvec_create(dmrg_matvec **op, dmrg_dtype dtype, int nmpo, const void *const *mpos,
                               const int64_t *mpo_dims, const void *envL, const int64_t envL_dims[3],
                               const void *envR, const int64_t envR_dims[3]) {
    return guard([&] {
        DMRG_REQUIRE(op, DMRG_ERR_ARG, "null op");
        *op = nullptr;
        DMRG_REQUIRE(dtype == DMRG_FP64 || dtype == DMRG_CX64, DMRG_ERR_ARG, "operator dtype must be FP64 or CX64");
        *op = reinterpret_cast<dmrg_matvec *>(new MatVecOp(dtype, nmpo, mpos, mpo_dims, envL, envL_dims, envR, envR_dims));
    });
}
dmrg_status dmrg_matvec_destroy(dmrg_matvec *op) {
    return guard([&] { delete reinterpret_cast<MatVecOp *>(op); });
}
dmrg_status dmrg_matvec_set_b(dmrg_matvec *op, int nmpo, const void *const *mpos, const int64_t *mpo_dims, const void *envL,
                              const int64_t envL_dims[3], const void *envR, const int64_t envR_dims[3]) {
    return guard([&] {
        DMRG_REQUIRE(op, DMRG_ERR_ARG, "null op");
        auto *a = reinterpret_cast<MatVecOp *>(op);
        auto  b = std::make_unique<MatVecOp>(a->dtype, nmpo, mpos, mpo_dims, envL, envL_dims, envR, envR_dims);
        DMRG_REQUIRE(b->D.d == a->D.d && b->D.chiL == a->D.chiL && b->D.chiR == a->D.chiR, DMRG_ERR_DIMS,
                     "B operator: theta shape differs from the A operator's");
        a->opB = std::move(b);
    });
}
dmrg_status dmrg_matvec_multbx(dmrg_matvec *op, const void *x, int64_t ldx, void *y, int64_t ldy, int nvec) {
    return guard([&] {
        DMRG_REQUIRE(op && x && y && nvec >= 0, DMRG_ERR_ARG, "bad argument");
        auto *a = reinterpret_cast<MatVecOp *>(op);
        if(!a->opB) return;
        a->opB->multax_host(x, ldx, y, ldy, nvec);
        a->num_mv += nvec;
    });
}
dmrg_status dmrg_matvec_multbx_device(dmrg_matvec *op, const void *dx, int64_t ldx, void *dy, int64_t ldy, int nvec,
                                      void *stream) {
    return guard([&] {
        DMRG_REQUIRE(op && dx && dy && nvec >= 0, DMRG_ERR_ARG, "bad argument");
        auto *a = reinterpret_cast<MatVecOp *>(op);
        if(!a->opB) return;
        a->opB->multax_device(dx, ldx, dy, ldy, nvec, static_cast<hipStream_t>(stream), false);
        a->num_mv += nvec;
    });
}
int64_t dmrg_matvec_rows(const dmrg_matvec *op) { return op ? reinterpret_cast<const MatVecOp *>(op)->D.n() : -1; }
dmrg_status dmrg_matvec_shape_mps(const dmrg_matvec *op, int64_t shape[3]) {
    return guard([&] {
        DMRG_REQUIRE(op && shape, DMRG_ERR_ARG, "null argument");
        const auto *o = reinterpret_cast<const MatVecOp *>(op);
        shape[0] = o->D.d, shape[1] = o->D.chiL, shape[2] = o->D.chiR;
    });
}
int64_t dmrg_matvec_num_mv(const dmrg_matvec *op) { return op ? reinterpret_cast<const MatVecOp *>(op)->num_mv : -1; }

dmrg_status dmrg_matvec_multax(dmrg_matvec *op, const void *x, int64_t ldx, void *y, int64_t ldy, int nvec) {
    return guard([&] {
        DMRG_REQUIRE(op && x && y && nvec >= 0, DMRG_ERR_ARG, "bad argument");
        reinterpret_cast<MatVecOp *>(op)->multax_host(x, ldx, y, ldy, nvec);
    });
}
dmrg_status dmrg_matvec_multax_device(dmrg_matvec *op, const void *dx, int64_t ldx, void *dy, int64_t ldy, int nvec,
                                      void *stream) {
    return guard([&] {
        DMRG_REQUIRE(op && dx && dy && nvec >= 0, DMRG_ERR_ARG, "bad argument");
        reinterpret_cast<MatVecOp *>(op)->multax_device(dx, ldx, dy, ldy, nvec, static_cast<hipStream_t>(stream), false);
    });
}
dmrg_status dmrg_matvec_multax_device_fp32(dmrg_matvec *op, const void *dx, int64_t ldx, void *dy, int64_t ldy,
                                           int nvec, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(op && dx && dy && nvec >= 0, DMRG_ERR_ARG, "bad argument");
        reinterpret_cast<MatVecOp *>(op)->multax_device(dx, ldx, dy, ldy, nvec, static_cast<hipStream_t>(stream), true);
    });
}
dmrg_status dmrg_matvec_set_shift(dmrg_matvec *op, double re, double im) {
    return guard([&] {
        DMRG_REQUIRE(op, DMRG_ERR_ARG, "null op");
        reinterpret_cast<MatVecOp *>(op)->set_shift(std::complex<double>(re, im));
    });
}
dmrg_status dmrg_matvec_get_diagonal(dmrg_matvec *op, void *ddiag, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(op && ddiag, DMRG_ERR_ARG, "null argument");
        reinterpret_cast<MatVecOp *>(op)->diagonal(ddiag, static_cast<hipStream_t>(stream));
    });
}

dmrg_status dmrg_matvec_set_jcb_block_size(dmrg_matvec *op, int64_t size) {
    return guard([&] {
        DMRG_REQUIRE(op, DMRG_ERR_ARG, "null operator");
        reinterpret_cast<MatVecOp *>(op)->set_jcb(size);
    });
}

dmrg_status dmrg_matvec_set_factorization(dmrg_matvec *op, int factorization) {
    return guard([&] {
        DMRG_REQUIRE(op, DMRG_ERR_ARG, "null operator");
        reinterpret_cast<MatVecOp *>(op)->set_factorization(factorization);
    });
}

dmrg_status dmrg_matvec_factor_op(dmrg_matvec *op, int factorization, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(op, DMRG_ERR_ARG, "null operator");
        reinterpret_cast<MatVecOp *>(op)->factor_op(factorization, static_cast<hipStream_t>(stream));
    });
}
int dmrg_matvec_factor_ready(const dmrg_matvec *op) { return op && reinterpret_cast<const MatVecOp *>(op)->op_ready ? 1 : 0; }
dmrg_status dmrg_matvec_mult_opv(dmrg_matvec *op, const void *x, int64_t ldx, void *y, int64_t ldy, int nvec) {
    return guard([&] {
        DMRG_REQUIRE(op && (nvec == 0 || (x && y)) && nvec >= 0, DMRG_ERR_ARG, "bad argument");
        reinterpret_cast<MatVecOp *>(op)->mult_opv_host(x, ldx, y, ldy, nvec);
    });
}
int64_t dmrg_matvec_num_op(const dmrg_matvec *op) { return op ? reinterpret_cast<const MatVecOp *>(op)->num_op : -1; }

dmrg_status dmrg_matvec_precond(dmrg_matvec *op, const void *x, int64_t ldx, void *y, int64_t ldy, int nvec,
                                const double *shifts) {
    return guard([&] {
        DMRG_REQUIRE(op && (nvec == 0 || (x && y)) && nvec >= 0, DMRG_ERR_ARG, "bad argument");
        reinterpret_cast<MatVecOp *>(op)->precond_host(x, ldx, y, ldy, nvec, shifts);
    });
}

void dmrg_eigs_default_config(dmrg_eigs_config *cfg) {
    if(!cfg) return;
    cfg->nev         = 1;
    cfg->ncv         = 0;
    cfg->min_restart = 0;
    cfg->target      = DMRG_TARGET_SA;
    cfg->shift       = 0.0;
    cfg->tol         = 1e-12;
    cfg->max_matvecs = 100000;
    cfg->precond     = 0;
    cfg->mixed       = 0;
    cfg->projection  = 0;
    cfg->conv_test   = 0;
    cfg->max_time_s  = 0.0;
    cfg->seed        = 12345;
    cfg->jcb_block   = 128;
    cfg->inner_max   = 20;
    cfg->inner_tol   = 1e-2;
}

dmrg_status dmrg_eigs(dmrg_matvec *opp, const dmrg_eigs_config *cfg, const void *x0, void *evecs, dmrg_eigs_result *res,
                      void *stream) {
    return guard([&] {
        DMRG_REQUIRE(opp && cfg && evecs && res, DMRG_ERR_ARG, "null argument");
        DMRG_REQUIRE(cfg->nev >= 1 && cfg->nev <= DMRG_EIGS_MAX_NEV, DMRG_ERR_ARG, "nev out of range");
        DMRG_REQUIRE(cfg->target >= 0 && cfg->target <= 2, DMRG_ERR_ARG, "bad target");
        DMRG_REQUIRE(cfg->tol > 0, DMRG_ERR_ARG, "tol must be positive");
        auto      &op = *reinterpret_cast<MatVecOp *>(opp);
        EigsParams p;
        p.nev         = cfg->nev;
        p.ncv         = cfg->ncv;
        p.min_restart = cfg->min_restart;
        p.target      = cfg->target;
        p.shift       = cfg->shift;
        p.tol         = std::max(cfg->tol, std::numeric_limits<double>::epsilon());
        p.max_matvecs = cfg->max_matvecs > 0 ? cfg->max_matvecs : 100000;
        p.precond     = cfg->precond;
        p.max_time_s  = cfg->max_time_s;
        p.seed        = cfg->seed;
        p.jcb         = cfg->jcb_block;
        DMRG_REQUIRE(cfg->inner_max >= 0 && cfg->inner_tol > 0 && cfg->inner_tol < 1, DMRG_ERR_ARG, "bad inner_max / inner_tol");
        p.inner_max = cfg->mixed ? cfg->inner_max : 0;
        p.inner_tol = cfg->inner_tol;
        DMRG_REQUIRE(cfg->projection == 0 || cfg->projection == 1, DMRG_ERR_ARG, "projection: 0 Rayleigh-Ritz or 1 refined");
        DMRG_REQUIRE(cfg->conv_test == 0 || cfg->conv_test == 1, DMRG_ERR_ARG, "conv_test: 0 or 1");
        p.projection = cfg->projection;
        p.conv_test  = cfg->conv_test;
        auto      s   = static_cast<hipStream_t>(stream);
        EigsStats st, s1;
        *res          = dmrg_eigs_result{};
        const bool cx = op.dtype == DMRG_CX64;
        if(cfg->mixed) {
            if(cx) st = eigs_mixed<cpx<double>, cpx<float>>(op, s, p, static_cast<const cpx<double> *>(x0), static_cast<cpx<double> *>(evecs), res->eigval, res->rnorm, &s1);
            else st = eigs_mixed<double, float>(op, s, p, static_cast<const double *>(x0), static_cast<double *>(evecs), res->eigval, res->rnorm, &s1);
            res->matvecs_fp32 = s1.matvecs + st.inner_matvecs;
            res->iters        = s1.iters;
            res->restarts     = s1.restarts;
        } else if(cx) {
            Davidson<cpx<double>> d(op, s);
            st = d.solve(p, static_cast<const cpx<double> *>(x0), x0 ? 1 : 0, static_cast<cpx<double> *>(evecs), res->eigval, res->rnorm);
        } else {
            Davidson<double> d(op, s);
            st = d.solve(p, static_cast<const double *>(x0), x0 ? 1 : 0, static_cast<double *>(evecs), res->eigval, res->rnorm);
        }
        res->nconv        = st.nconv;
        res->iters += st.iters;
        res->restarts += st.restarts;
        res->stagnated    = st.stagnated ? 1 : 0;
        res->matvecs_fp64 = st.matvecs;
        res->seconds      = st.seconds;
        res->anorm        = st.anorm;
    });
}

dmrg_status dmrg_eigs_folded_spectrum(dmrg_matvec *h, dmrg_matvec *h2, const dmrg_eigs_config *cfg, const void *x0,
                                      double energy_shift, void *evecs, double *energy, double *variance,
                                      dmrg_eigs_result *res, void *stream) {
    const dmrg_status st = dmrg_eigs(h2, cfg, x0, evecs, res, stream);
    if(st != DMRG_OK) return st;
    return guard([&] {
        DMRG_REQUIRE(h && energy && variance, DMRG_ERR_ARG, "null argument");
        auto &H  = *reinterpret_cast<MatVecOp *>(h);
        auto &H2 = *reinterpret_cast<MatVecOp *>(h2);
        DMRG_REQUIRE(H.dtype == H2.dtype && H.D.n() == H2.D.n(), DMRG_ERR_DIMS, "folded spectrum: H and H^2 act on different spaces");
        auto          s  = static_cast<hipStream_t>(stream);
        const int64_t n  = H.D.n();
        const size_t  es = dtype_size(H.dtype);
        DevBuf        y;
        y.reserve(static_cast<size_t>(n) * es + 8 * sizeof(double));
        double *red = reinterpret_cast<double *>(static_cast<char *>(y.p) + static_cast<size_t>(n) * es);
        for(int q = 0; q < cfg->nev; ++q) {
            const void *x = static_cast<const char *>(evecs) + static_cast<size_t>(q) * n * es;
            H.multax_device(x, n, y.p, n, 1, s, false);
            dot_device(H.dtype, n, x, y.p, red, s);
            H2.multax_device(x, n, y.p, n, 1, s, false);
            dot_device(H.dtype, n, x, y.p, red + 2, s);
            dot_device(H.dtype, n, x, x, red + 4, s);
            double r[6];
            DMRG_HIP_CHECK(hipMemcpyAsync(r, red, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
            DMRG_HIP_CHECK(hipStreamSynchronize(s));
            energy[q]       = r[0] / r[4];
            const double dE = energy[q] - energy_shift;
            variance[q]     = r[2] / r[4] - dE * dE; // <(H - E_shift)^2> - (E - E_shift)^2 (xdmrg.cpp:488-496)
        }
    });
}


dmrg_status dmrg_eigs_batched(dmrg_matvec *const *ops, int k, const dmrg_eigs_config *cfg, const void *const *x0s,
                              void *const *evecs, dmrg_eigs_result *results, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(ops && cfg && evecs && results && k >= 1, DMRG_ERR_ARG, "null argument");
        const EigsParams p = batched_params(cfg);
        std::vector<MatVecOp *> v(k);
        for(int i = 0; i < k; ++i) {
            DMRG_REQUIRE(ops[i] && evecs[i], DMRG_ERR_ARG, "null operator or output");
            v[i] = reinterpret_cast<MatVecOp *>(ops[i]);
        }
        const auto                   t0 = std::chrono::steady_clock::now();
        std::vector<BatchEigsResult> r;
        eigs_batched(v.data(), k, p, x0s, evecs, r, static_cast<hipStream_t>(stream), cfg->mixed != 0);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for(int i = 0; i < k; ++i) results[i] = batched_result(r[i], sec);
    });
}

dmrg_status dmrg_serve_eigs(const dmrg_eigs_config *cfg, const dmrg_eig_request *reqs, int nreq, int64_t hbm_budget_bytes,
                            int max_batch, dmrg_eig_request_result *out, dmrg_serve_stats *stats, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(cfg && reqs && out && stats && nreq >= 1, DMRG_ERR_ARG, "null argument");
        DMRG_REQUIRE(hbm_budget_bytes > 0, DMRG_ERR_ARG, "serve eigs: budget must be positive");
        const EigsParams        p = batched_params(cfg);
        std::vector<EigRequest> rq(nreq);
        for(int i = 0; i < nreq; ++i) {
            DMRG_REQUIRE(reqs[i].op && reqs[i].evec && reqs[i].arrival_s >= 0, DMRG_ERR_ARG, "serve eigs: bad request");
            rq[i].op        = reinterpret_cast<MatVecOp *>(reqs[i].op);
            rq[i].x0        = reqs[i].x0;
            rq[i].evec      = reqs[i].evec;
            rq[i].arrival_s = reqs[i].arrival_s;
        }
        std::vector<EigRequestResult> r;
        ServeSummary                  sum;
        serve_eigs(rq, p, hbm_budget_bytes, max_batch, r, sum, static_cast<hipStream_t>(stream), cfg->mixed != 0);
        std::vector<double> lat(nreq);
        double              mean = 0;
        int                 conv = 0;
        for(int i = 0; i < nreq; ++i) {
            out[i]                 = dmrg_eig_request_result{};
            out[i].res             = batched_result(r[i].res, r[i].finish_s - r[i].start_s);
            out[i].arrival_s       = r[i].arrival_s;
            out[i].start_s         = r[i].start_s;
            out[i].finish_s        = r[i].finish_s;
            out[i].latency_s       = r[i].latency_s;
            out[i].footprint_bytes = r[i].footprint_bytes;
            lat[i]                 = r[i].latency_s;
            mean += lat[i] / nreq;
            conv += r[i].res.nconv > 0;
        }
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) { // linear interpolation between order statistics, as dmrg_serve
            const double pos = q * (lat.size() - 1);
            const size_t lo  = static_cast<size_t>(std::floor(pos)), hi = std::min(lat.size() - 1, lo + 1);
            return lat[lo] + (pos - lo) * (lat[hi] - lat[lo]);
        };
        *stats                     = dmrg_serve_stats{};
        stats->makespan_s          = sum.makespan_s;
        stats->solves_per_s        = conv / std::max(sum.makespan_s, 1e-12); // converged solves, as dmrg_serve
        stats->p50_latency_s       = pct(0.5);
        stats->p99_latency_s       = pct(0.99);
        stats->mean_latency_s      = mean;
        stats->completed           = conv;
        stats->failed              = nreq - conv;
        stats->peak_admitted_bytes = sum.peak_admitted_bytes;
        stats->peak_concurrency    = sum.peak_concurrency;
        stats->batch_steps         = sum.steps;
        stats->peak_batch          = sum.peak_concurrency;
        for(int i = 0; i < nreq; ++i) stats->matvecs += r[i].res.matvecs;
    });
}

dmrg_status dmrg_matvec_group_create(dmrg_matvec *const *ops, int k, const void *const *xs, void *const *ys, int fp32,
                                     int capture_graph, dmrg_matvec_group **out) {
    return guard([&] {
        DMRG_REQUIRE(ops && xs && ys && out && k >= 1, DMRG_ERR_ARG, "bad argument");
        *out = nullptr;
        std::vector<MatVecOp *> v(k);
        for(int i = 0; i < k; ++i) {
            DMRG_REQUIRE(ops[i], DMRG_ERR_ARG, "null operator");
            v[i] = reinterpret_cast<MatVecOp *>(ops[i]);
        }
        *out = reinterpret_cast<dmrg_matvec_group *>(matvec_group_create(v.data(), k, xs, ys, fp32 != 0, capture_graph != 0));
    });
}

dmrg_status dmrg_matvec_group_run(dmrg_matvec_group *group, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(group, DMRG_ERR_ARG, "null group");
        reinterpret_cast<MatvecGroup *>(group)->run(static_cast<hipStream_t>(stream));
    });
}

double dmrg_matvec_group_flop(const dmrg_matvec_group *group) {
    return group ? reinterpret_cast<const MatvecGroup *>(group)->flop : 0.0;
}

void dmrg_matvec_group_destroy(dmrg_matvec_group *group) { delete reinterpret_cast<MatvecGroup *>(group); }

void dmrg_fdmrg_default_config(dmrg_fdmrg_config *cfg) {
    if(!cfg) return;
    *cfg                 = dmrg_fdmrg_config{};
    cfg->L               = 16;
    cfg->J               = 1.0;
    cfg->h_tran          = 1.0;
    cfg->chi_max         = 64;
    cfg->max_sweeps      = 20;
    cfg->two_site_sweeps = 10;
    cfg->var_tol         = 1e-12;
    cfg->trunc_lim       = 1e-14;
    cfg->eigs_tol        = 1e-12;
    cfg->seed            = 0;
    cfg->expand_mode     = 1 | 2 | 8 | 32; // H1 | H2 | NSITE | FORWARD, dmrg_env_expand_mode's default (settings.h:228)
}

dmrg_status dmrg_fdmrg_ising_tf_rf(const dmrg_fdmrg_config *cfg, dmrg_fdmrg_result *res, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(cfg && res && cfg->h_rand, DMRG_ERR_ARG, "null argument");
        DMRG_REQUIRE(cfg->L >= 2 && cfg->L <= DMRG_FDMRG_MAX_L, DMRG_ERR_ARG, "L out of range");
        fdmrg_ising(*cfg, *res, static_cast<hipStream_t>(stream));
    });
}

dmrg_status dmrg_serve(const dmrg_serve_config *cfg, const dmrg_request *reqs, int nreq, dmrg_request_result *out,
                       dmrg_serve_stats *stats) {
    return guard([&] {
        DMRG_REQUIRE(cfg && (nreq == 0 || (reqs && out)) && stats && nreq >= 0, DMRG_ERR_ARG, "null argument");
        serve(*cfg, reqs, nreq, out, *stats);
    });
}
int64_t dmrg_request_footprint(int L, int chi_max, int complex_) { return request_footprint(L, chi_max, complex_ != 0); }

dmrg_status dmrg_host_heev(int m, const double *A, double *w, double *Z) {
    return guard([&] {
        DMRG_REQUIRE(m >= 0 && (m == 0 || (A && w && Z)), DMRG_ERR_ARG, "bad argument");
        if(!hermitian_eig(m, reinterpret_cast<const std::complex<double> *>(A), m, w, reinterpret_cast<std::complex<double> *>(Z)))
            throw Error(DMRG_ERR_NOT_CONVERGED, "tridiagonal QL did not converge");
    });
}

dmrg_status dmrg_autotune_enable(int enable) {
    return guard([&] { set_autotune(enable); });
}
dmrg_status dmrg_autotune_reset(void) {
    return guard([&] { clear_tuning(); });
}
dmrg_status dmrg_autotune_min_macs(int64_t macs) {
    return guard([&] {
        DMRG_REQUIRE(macs >= 0, DMRG_ERR_ARG, "autotune threshold must be >= 0");
        set_tune_min_macs(macs);
    });
}
int         dmrg_autotune_entries(void) { return tuned_entries(); }
dmrg_status dmrg_gemm_hybrid_override(int cfg, int nsplit, int ksplit2) {
    return guard([&] {
        DMRG_REQUIRE(cfg >= 0 && cfg <= static_cast<int>(TileCfg::T64x32) && nsplit > 0 && ksplit2 >= 1 && ksplit2 <= 8,
                     DMRG_ERR_ARG, "bad hybrid plan override");
        set_plan_override(cfg, 1, nsplit, ksplit2);
    });
}
dmrg_status dmrg_gemm_plan_override(int cfg, int ksplit) {
    return guard([&] {
        DMRG_REQUIRE(cfg <= static_cast<int>(TileCfg::T32x32K) && ksplit >= 0 && ksplit <= 8, DMRG_ERR_ARG, "bad plan override");
        set_plan_override(cfg, ksplit);
    });
}

dmrg_status dmrg_matvec_time_legs(dmrg_matvec *op, const void *dx, void *dy, int nrep, int fp32, float ms_per_leg[4],
                                  void *stream) {
    return guard([&] {
        DMRG_REQUIRE(op && dx && dy && ms_per_leg && nrep > 0, DMRG_ERR_ARG, "bad argument");
        reinterpret_cast<MatVecOp *>(op)->time_legs(dx, dy, nrep, fp32 != 0, ms_per_leg, static_cast<hipStream_t>(stream));
    });
}

dmrg_status dmrg_block_vector_time(dmrg_dtype dtype, int64_t n, int m, int nrep, float ms[3], double bytes[3]) {
    return guard([&] {
        DMRG_REQUIRE(n > 0 && m >= 1 && m <= VEC_MAX_M && nrep > 0 && ms && bytes, DMRG_ERR_ARG, "bad argument");
        switch(dtype) {
            case DMRG_FP64: time_block_vectors<double>(n, m, nrep, ms, bytes); break;
            case DMRG_CX64: time_block_vectors<cpx<double>>(n, m, nrep, ms, bytes); break;
            case DMRG_FP32: time_block_vectors<float>(n, m, nrep, ms, bytes); break;
            case DMRG_CX32: time_block_vectors<cpx<float>>(n, m, nrep, ms, bytes); break;
        }
    });
}

// ------------------------------------------------------------------------- dense path, edges, SVD

dmrg_status dmrg_heff_dense(dmrg_dtype dtype, void *H, int64_t ldh, const void *mpo, const int64_t mpo_dims[4],
                            const void *envL, const int64_t envL_dims[3], const void *envR, const int64_t envR_dims[3],
                            int transpose, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(H && mpo && envL && envR && mpo_dims && envL_dims && envR_dims, DMRG_ERR_ARG, "null argument");
        const int64_t mps_dims[3] = {mpo_dims[2], envL_dims[0], envR_dims[0]};
        const Dims    D           = check_matvec_dims(mps_dims, mpo_dims, envL_dims, envR_dims);
        heff_dense(dtype, D, mpo, envL, envR, H, ldh, transpose != 0, static_cast<hipStream_t>(stream));
    });
}

dmrg_status dmrg_matvec_get_matrix(dmrg_matvec *opp, void *H, int64_t ldh, int transpose, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(opp && H, DMRG_ERR_ARG, "null argument");
        auto *op = reinterpret_cast<MatVecOp *>(opp);
        heff_dense(op->dtype, op->D, op->mpo.p, op->envL_p, op->envR_p, H, ldh, transpose != 0, static_cast<hipStream_t>(stream));
    });
}

dmrg_status dmrg_dense_eigh(dmrg_dtype dtype, int64_t n, void *A, int64_t lda, double *w, int vectors, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(A && w, DMRG_ERR_ARG, "null argument");
        DMRG_REQUIRE(n > 0 && n < INT32_MAX && lda >= n && lda < INT32_MAX, DMRG_ERR_DIMS, "eigh: bad n / lda");
        auto               &solver = solver_for(static_cast<hipStream_t>(stream));
        std::vector<double> ev;
        if(dtype == DMRG_FP64) solver.eigh(static_cast<int>(n), static_cast<double *>(A), static_cast<int>(lda), vectors != 0, ev);
        else if(dtype == DMRG_CX64)
            solver.eigh(static_cast<int>(n), static_cast<cpx<double> *>(A), static_cast<int>(lda), vectors != 0, ev);
        else throw Error(DMRG_ERR_UNSUPPORTED, "eigh: FP64 or CX64");
        std::memcpy(w, ev.data(), static_cast<size_t>(n) * sizeof(double));
    });
}

dmrg_status dmrg_env_set_edge(dmrg_dtype dtype, void *block_out, int64_t chi, int64_t mpo_dim, const void *edge,
                              void *stream) {
    return guard([&] {
        DMRG_REQUIRE(block_out && edge, DMRG_ERR_ARG, "null argument");
        env_set_edge(dtype, block_out, chi, mpo_dim, edge, static_cast<hipStream_t>(stream));
    });
}

dmrg_status dmrg_svd_truncate(dmrg_dtype dtype, int64_t m, int64_t n, const void *A, double trunc_lim, int64_t rank_max,
                              int method, void *U, double *S, void *Vt, int64_t *rank, double *truncation_error,
                              double *discarded_weight, void *stream) {
    return guard([&] {
        DMRG_REQUIRE(A && U && S && Vt && rank, DMRG_ERR_ARG, "null argument");
        DMRG_REQUIRE(m > 0 && n > 0 && m < INT32_MAX && n < INT32_MAX && m * n < (int64_t(1) << 40), DMRG_ERR_DIMS,
                     "svd: bad matrix size");
        DMRG_REQUIRE(dtype == DMRG_FP64 || dtype == DMRG_CX64, DMRG_ERR_UNSUPPORTED, "svd: FP64 or CX64");
        auto                s      = static_cast<hipStream_t>(stream);
        auto               &solver = solver_for(s);
        std::vector<double> sv;
        if(dtype == DMRG_FP64) solver.compute(static_cast<int>(m), static_cast<int>(n), static_cast<const double *>(A), sv, method);
        else solver.compute(static_cast<int>(m), static_cast<int>(n), static_cast<const cpx<double> *>(A), sv, method);
        const int64_t k     = solver.k();
        const auto    rk    = rank_from_truncation_error(sv, trunc_lim, rank_max > 0 ? static_cast<int>(std::min<int64_t>(rank_max, k)) : 0, 1);
        const int64_t r     = rk.first;
        const size_t  es    = dtype_size(dtype);
        DMRG_HIP_CHECK(hipMemcpyAsync(U, solver.U(), static_cast<size_t>(m * r) * es, hipMemcpyDeviceToDevice, s));
        DMRG_HIP_CHECK(hipMemcpy2DAsync(Vt, static_cast<size_t>(r) * es, solver.Vt(), static_cast<size_t>(k) * es,
                                        static_cast<size_t>(r) * es, static_cast<size_t>(n), hipMemcpyDeviceToDevice, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        double disc = 0.0;
        for(int64_t i = k - 1; i >= r; --i) disc += sv[i] * sv[i];
        std::memcpy(S, sv.data(), static_cast<size_t>(k) * sizeof(double));
        *rank = r;
        if(truncation_error) *truncation_error = rk.second;
        if(discarded_weight) *discarded_weight = disc;
    });
}

} // extern "C"

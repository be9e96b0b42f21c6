// Block-vector kernels of the eigensolver. Every kernel streams its vectors once from HBM with
// coalesced accesses; reductions go through per-block partials folded in a fixed order.
#include "vecops.hpp"

#include <algorithm>

namespace dmrg {

constexpr int VT_THREADS = 256;

template <typename T>
__device__ inline void ld(const T &v, double &re, double &im) {
    if constexpr(vtraits<T>::cx) {
        re = static_cast<double>(v.re);
        im = static_cast<double>(v.im);
    } else {
        re = static_cast<double>(v);
        im = 0.0;
    }
}
template <typename T>
__device__ inline T mk(double re, double im) {
    using R = typename vtraits<T>::R;
    if constexpr(vtraits<T>::cx) return T{static_cast<R>(re), static_cast<R>(im)};
    else return static_cast<R>(re);
}

__device__ inline double wsum(double v) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block reduction of NV values; results valid in thread 0.
template <int NV>
__device__ inline void bsum(double (&v)[NV], double (*red)[VT_THREADS / 64]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for(int i = 0; i < NV; ++i) v[i] = wsum(v[i]);
    __syncthreads();
    if(lane == 0)
#pragma unroll
        for(int i = 0; i < NV; ++i) red[i][wave] = v[i];
    __syncthreads();
    if(threadIdx.x == 0)
#pragma unroll
        for(int i = 0; i < NV; ++i) {
            double s = 0;
            for(int w = 0; w < VT_THREADS / 64; ++w) s += red[i][w];
            v[i] = s;
        }
}

static int grid_for(int64_t n) {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(VEC_MAX_BLOCKS, (n + VT_THREADS - 1) / VT_THREADS)));
}

// ------------------------------------------------------------------------------------------- fused folds
// A reduction over few workgroups folds its partial sums in its last workgroup instead of a second launch:
// every workgroup publishes its partials (release fence) and counts itself in; the one that counts last
// reads all partials past the non-coherent L1 and sums them in workgroup order (the order of the separate
// fold kernel's operands, so the result stays deterministic), then resets the counter for the next use.
constexpr int FUSE_MAX_BLOCKS = 128; // more partials than this per value are folded by a separate launch
constexpr int FUSE_MAX_VALUES = VT_THREADS;

static unsigned *scratch_counter(double *scratch) { return reinterpret_cast<unsigned *>(scratch + vec_scratch_doubles() - 8); }

void vec_scratch_init(double *scratch, hipStream_t s) {
    DMRG_HIP_CHECK(hipMemsetAsync(scratch_counter(scratch), 0, 8 * sizeof(double), s));
}

// partial: [block][nv]; value v goes to out[v], or self_out[v & 1] for the column v / 2 == j_self (when given).
__device__ inline void fold_in_last_block(const double *partial, int nv, double *out, int j_self, double *self_out, unsigned *counter) {
    __shared__ int last;
    __threadfence();
    __syncthreads();
    if(threadIdx.x == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1 ? 1 : 0;
    __syncthreads();
    if(!last) return;
    __threadfence();
    // Thread t sums value t % nv over the workgroups b = t / nv, t / nv + per, ...; then thread v adds the per
    // sums of value v in order. nv <= blockDim.x.
    __shared__ double red[VT_THREADS];
    const int         per = blockDim.x / nv;
    const int         t   = threadIdx.x;
    if(t < per * nv) {
        const int v = t % nv;
        double    a = 0.0;
        for(unsigned b = t / nv; b < gridDim.x; b += per)
            a += __hip_atomic_load(partial + static_cast<int64_t>(b) * nv + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        red[t] = a;
    }
    __syncthreads();
    if(t < nv) {
        double acc = 0.0;
        for(int k = 0; k < per; ++k) acc += red[t + k * nv];
        double *o = ((t >> 1) == j_self && self_out) ? self_out + (t & 1) : out + t;
        *o        = acc;
    }
    if(t == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------------------------------- multidot
constexpr int MD_CH        = 16;  // basis columns per pass (32 double accumulators per thread)
constexpr int MD_PER_THREAD = 4;  // rows per thread and pass at full grid: amortises the block reduction

// Reduce-scatter of NV values across the 64 lanes of a wavefront: after the call lane l holds in v[0] the
// wave total of value index rs_index<NV>(l) (valid in the lanes whose low bits below the last exchange are
// zero). Halving the set exchanged at each step costs NV - 1 + (6 - log2 NV) shuffles instead of 6 NV.
template <int H, int O, int NV>
__device__ inline void rs_step(double (&v)[NV], int lane) {
    if constexpr(H >= 1) {
        const bool up = (lane & O) != 0;
#pragma unroll
        for(int k = 0; k < H; ++k) {
            const double send = up ? v[k] : v[k + H];
            const double keep = up ? v[k + H] : v[k];
            v[k]              = keep + __shfl_xor(send, O, 64);
        }
        rs_step<H / 2, O / 2, NV>(v, lane);
    } else {
#pragma unroll
        for(int o = O; o >= 1; o >>= 1) v[0] += __shfl_xor(v[0], o, 64); // remaining low lane bits
    }
}
template <int NV>
__device__ inline int rs_index(int lane) {
    int idx = 0;
#pragma unroll
    for(int h = NV / 2, o = 32; h >= 1; h >>= 1, o >>= 1) idx += (lane & o) ? h : 0;
    return idx;
}
template <int NV>
__device__ inline bool rs_owner(int lane) {
    return (lane & (64 / NV - 1)) == 0;
}

// DGKS test evaluated by the projection kernel of a second orthogonalisation pass itself: the pass runs when
// post[0] < eta2 * pre[0] (|t|^2 fell by more than that factor in the first pass); block 0 records the decision
// in flag[0] for the kernels of the pass that follow. pre and post are not written during the kernel, so every
// block reaches the same decision.
struct DgksTest {
    const double *pre  = nullptr;
    const double *post = nullptr;
    double        eta2 = 0.0;
    int          *flag = nullptr;
};

// Columns j < m are V[:, j]; column m (when with_self) is x itself, so x^H x comes out of the same pass.
template <typename T>
__global__ __launch_bounds__(VT_THREADS) void multidot_partial(int64_t n, int m, int mc, const T *V, int64_t ldv, const T *x,
                                                               const int *skip, double *partial, unsigned *counter, double *out,
                                                               int m_out, double *self_out, DgksTest dg) {
    if(dg.flag) {
        const bool run = dg.post[0] < dg.eta2 * dg.pre[0];
        if(blockIdx.x == 0 && threadIdx.x == 0) dg.flag[0] = run ? 1 : 0;
        if(!run) return;
    } else if(skip && *skip == 0) {
        return;
    }
    constexpr int     NV = 2 * MD_CH;
    __shared__ double red[NV][VT_THREADS / 64];
    const int         lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for(int j0 = 0; j0 < mc; j0 += MD_CH) {
        double acc[NV];
#pragma unroll
        for(int c = 0; c < NV; ++c) acc[c] = 0.0;
        for(int64_t i = blockIdx.x * static_cast<int64_t>(VT_THREADS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * VT_THREADS) {
            double xr, xi;
            ld(x[i], xr, xi);
#pragma unroll
            for(int c = 0; c < MD_CH; ++c) {
                if(j0 + c < mc) {
                    double vr, vi;
                    if(j0 + c < m) ld(V[i + (j0 + c) * ldv], vr, vi);
                    else vr = xr, vi = xi;
                    acc[2 * c] += vr * xr + vi * xi; // conj(v) x
                    acc[2 * c + 1] += vr * xi - vi * xr;
                }
            }
        }
        rs_step<NV / 2, 32, NV>(acc, lane);
        if(rs_owner<NV>(lane)) red[rs_index<NV>(lane)][wave] = acc[0];
        __syncthreads();
        if(threadIdx.x < NV) {
            const int c = threadIdx.x >> 1;
            if(j0 + c < mc) {
                double t = 0;
#pragma unroll
                for(int w = 0; w < VT_THREADS / 64; ++w) t += red[threadIdx.x][w];
                partial[2 * (static_cast<int64_t>(blockIdx.x) * mc + j0 + c) + (threadIdx.x & 1)] = t;
            }
        }
        __syncthreads();
    }
    if(counter) fold_in_last_block(partial, 2 * mc, out, m_out, self_out, counter);
}

static int md_grid(int64_t n) {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(VEC_MAX_BLOCKS, (n + VT_THREADS * MD_PER_THREAD - 1) / (VT_THREADS * MD_PER_THREAD))));
}

// out[j] = sum_b partial[b][j] for j < m (column j == m_out goes to self_out instead, when given): one block
// per column, each thread a fixed strided subset of the partials, then a fixed-order tree (deterministic).
__global__ __launch_bounds__(VT_THREADS) void fold_columns(int nblk, int m, const double *partial, double *out, int m_out,
                                                           double *self_out, const int *skip) {
    if(skip && *skip == 0) return;
    __shared__ double red[2][VT_THREADS / 64];
    const int         j   = blockIdx.x;
    double            acc[2] = {0.0, 0.0};
    for(int b = threadIdx.x; b < nblk; b += VT_THREADS) {
        acc[0] += partial[2 * (static_cast<int64_t>(b) * m + j)];
        acc[1] += partial[2 * (static_cast<int64_t>(b) * m + j) + 1];
    }
    bsum<2>(acc, red);
    if(threadIdx.x == 0) {
        double *o = (j == m_out && self_out) ? self_out : out + 2 * j;
        o[0]      = acc[0];
        o[1]      = acc[1];
    }
}

template <typename T>
void multidot(int64_t n, int m, const T *V, int64_t ldv, const T *x, double *d_out, double *scratch, hipStream_t s) {
    if(m <= 0) return;
    DMRG_REQUIRE(m <= VEC_MAX_M, DMRG_ERR_ARG, "multidot: too many columns");
    const int g = md_grid(n);
    if(g <= FUSE_MAX_BLOCKS && 2 * m <= FUSE_MAX_VALUES) {
        hipLaunchKernelGGL(multidot_partial<T>, dim3(g), dim3(VT_THREADS), 0, s, n, m, m, V, ldv, x, nullptr, scratch,
                           scratch_counter(scratch), d_out, -1, nullptr, DgksTest{});
    } else {
        hipLaunchKernelGGL(multidot_partial<T>, dim3(g), dim3(VT_THREADS), 0, s, n, m, m, V, ldv, x, nullptr, scratch, nullptr,
                           nullptr, -1, nullptr, DgksTest{});
        hipLaunchKernelGGL(fold_columns, dim3(m), dim3(VT_THREADS), 0, s, g, m, scratch, d_out, -1, nullptr, nullptr);
    }
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename T>
static void multidot_x_impl(int64_t n, int m, const T *V, int64_t ldv, const T *x, double *d_out, double *d_self, const int *skip,
                            const DgksTest &dg, double *scratch, hipStream_t s) {
    const int mc = m + (d_self ? 1 : 0);
    if(mc <= 0) return;
    DMRG_REQUIRE(mc <= VEC_MAX_M, DMRG_ERR_ARG, "multidot: too many columns");
    const int  g     = md_grid(n);
    const int *fskip = dg.flag ? dg.flag : skip; // the fold follows the pass's decision
    if(g <= FUSE_MAX_BLOCKS && 2 * mc <= FUSE_MAX_VALUES) {
        hipLaunchKernelGGL(multidot_partial<T>, dim3(g), dim3(VT_THREADS), 0, s, n, m, mc, V, ldv, x, skip, scratch,
                           scratch_counter(scratch), d_out, d_self ? m : -1, d_self, dg);
    } else {
        hipLaunchKernelGGL(multidot_partial<T>, dim3(g), dim3(VT_THREADS), 0, s, n, m, mc, V, ldv, x, skip, scratch, nullptr,
                           nullptr, -1, nullptr, dg);
        hipLaunchKernelGGL(fold_columns, dim3(mc), dim3(VT_THREADS), 0, s, g, mc, scratch, d_out, d_self ? m : -1, d_self, fskip);
    }
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void multidot_x(int64_t n, int m, const T *V, int64_t ldv, const T *x, double *d_out, double *d_self, const int *skip,
                double *scratch, hipStream_t s) {
    multidot_x_impl<T>(n, m, V, ldv, x, d_out, d_self, skip, DgksTest{}, scratch, s);
}

template <typename T>
void multidot_dgks(int64_t n, int m, const T *V, int64_t ldv, const T *x, double *d_out, const double *d_pre, const double *d_post,
                   double eta2, int *d_flag, double *scratch, hipStream_t s) {
    DgksTest dg;
    dg.pre  = d_pre;
    dg.post = d_post;
    dg.eta2 = eta2;
    dg.flag = d_flag;
    multidot_x_impl<T>(n, m, V, ldv, x, d_out, nullptr, nullptr, dg, scratch, s);
}

// ------------------------------------------------------------------------------------------- gemv
template <typename T>
__global__ __launch_bounds__(VT_THREADS) void gemv_update_kernel(int64_t n, int m, const T *V, int64_t ldv, const double *h,
                                                                 double alpha, double beta, T *x) {
    __shared__ double hs[2 * VEC_MAX_M];
    for(int j = threadIdx.x; j < 2 * m; j += VT_THREADS) hs[j] = h[j];
    __syncthreads();
    for(int64_t i = blockIdx.x * static_cast<int64_t>(VT_THREADS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * VT_THREADS) {
        double ar = 0, ai = 0;
        for(int j = 0; j < m; ++j) {
            double vr, vi;
            ld(V[i + j * ldv], vr, vi);
            ar += vr * hs[2 * j] - vi * hs[2 * j + 1];
            ai += vr * hs[2 * j + 1] + vi * hs[2 * j];
        }
        double xr = 0, xi = 0;
        if(beta != 0.0) ld(x[i], xr, xi);
        x[i] = mk<T>(beta * xr + alpha * ar, beta * xi + alpha * ai);
    }
}

// y = beta x + alpha V h (x may alias y) and per-block partials of |y|^2.
template <typename T>
__global__ __launch_bounds__(VT_THREADS) void gemv_update_norm_kernel(int64_t n, int m, const T *V, int64_t ldv, const double *h,
                                                                      double alpha, double beta, const T *x, T *y,
                                                                      const int *skip, double *partial, unsigned *counter,
                                                                      double *out) {
    if(skip && *skip == 0) return;
    __shared__ double hs[2 * VEC_MAX_M];
    __shared__ double red[2][VT_THREADS / 64];
    for(int j = threadIdx.x; j < 2 * m; j += VT_THREADS) hs[j] = h[j];
    __syncthreads();
    double acc[2] = {0.0, 0.0};
    for(int64_t i = blockIdx.x * static_cast<int64_t>(VT_THREADS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * VT_THREADS) {
        double ar = 0, ai = 0;
        for(int j = 0; j < m; ++j) {
            double vr, vi;
            ld(V[i + j * ldv], vr, vi);
            ar += vr * hs[2 * j] - vi * hs[2 * j + 1];
            ai += vr * hs[2 * j + 1] + vi * hs[2 * j];
        }
        double xr = 0, xi = 0;
        if(beta != 0.0) ld(x[i], xr, xi);
        const T yv = mk<T>(beta * xr + alpha * ar, beta * xi + alpha * ai);
        y[i]       = yv;
        double yr, yi;
        ld(yv, yr, yi); // the norm of the stored (possibly FP32-rounded) vector
        acc[0] += yr * yr + yi * yi;
    }
    bsum<2>(acc, red);
    if(threadIdx.x == 0) {
        partial[2 * blockIdx.x]     = acc[0];
        partial[2 * blockIdx.x + 1] = 0.0;
    }
    if(counter) fold_in_last_block(partial, 2, out, -1, nullptr, counter);
}

static int ew_grid(int64_t n) { return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(8192, (n + VT_THREADS - 1) / VT_THREADS))); }

template <typename T>
void gemv_update(int64_t n, int m, const T *V, int64_t ldv, const double *d_h, double alpha, double beta, T *x, hipStream_t s) {
    DMRG_REQUIRE(m <= VEC_MAX_M, DMRG_ERR_ARG, "gemv: too many columns");
    hipLaunchKernelGGL(gemv_update_kernel<T>, dim3(ew_grid(n)), dim3(VT_THREADS), 0, s, n, m, V, ldv, d_h, alpha, beta, x);
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void gemv_update_norm(int64_t n, int m, const T *V, int64_t ldv, const double *d_h, double alpha, double beta, const T *x, T *y,
                      double *d_nrm2, const int *skip, double *scratch, hipStream_t s) {
    DMRG_REQUIRE(m <= VEC_MAX_M, DMRG_ERR_ARG, "gemv: too many columns");
    const int g = grid_for(n);
    if(g <= FUSE_MAX_BLOCKS) {
        hipLaunchKernelGGL(gemv_update_norm_kernel<T>, dim3(g), dim3(VT_THREADS), 0, s, n, m, V, ldv, d_h, alpha, beta, x, y, skip, scratch,
                           scratch_counter(scratch), d_nrm2);
    } else {
        hipLaunchKernelGGL(gemv_update_norm_kernel<T>, dim3(g), dim3(VT_THREADS), 0, s, n, m, V, ldv, d_h, alpha, beta, x, y, skip, scratch,
                           nullptr, nullptr);
        hipLaunchKernelGGL(fold_columns, dim3(1), dim3(VT_THREADS), 0, s, g, 1, scratch, d_nrm2, -1, nullptr, skip);
    }
    DMRG_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------------------------------- Ritz vector + residual
template <typename T>
__global__ __launch_bounds__(VT_THREADS) void ritz_kernel(int64_t n, int m, const T *V, const T *W, int64_t ldv, const double *y,
                                                          double theta, T *x, T *r, double *partial, unsigned *counter, double *out) {
    __shared__ double ys[2 * VEC_MAX_M];
    __shared__ double red[2][VT_THREADS / 64];
    for(int j = threadIdx.x; j < 2 * m; j += VT_THREADS) ys[j] = y[j];
    __syncthreads();
    double acc[2] = {0.0, 0.0};
    for(int64_t i = blockIdx.x * static_cast<int64_t>(VT_THREADS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * VT_THREADS) {
        double xr = 0, xi = 0, wr = 0, wi = 0;
        for(int j = 0; j < m; ++j) {
            double vr, vi, ur, ui;
            ld(V[i + j * ldv], vr, vi);
            ld(W[i + j * ldv], ur, ui);
            const double yr = ys[2 * j], yi = ys[2 * j + 1];
            xr += vr * yr - vi * yi;
            xi += vr * yi + vi * yr;
            wr += ur * yr - ui * yi;
            wi += ur * yi + ui * yr;
        }
        const double rr = wr - theta * xr, ri = wi - theta * xi;
        x[i] = mk<T>(xr, xi);
        r[i] = mk<T>(rr, ri);
        acc[0] += rr * rr + ri * ri;
        acc[1] += xr * xr + xi * xi;
    }
    bsum<2>(acc, red);
    if(threadIdx.x == 0) {
        partial[2 * blockIdx.x]     = acc[0];
        partial[2 * blockIdx.x + 1] = acc[1];
    }
    if(counter) fold_in_last_block(partial, 2, out, -1, nullptr, counter);
}

template <typename T>
void ritz_residual(int64_t n, int m, const T *V, const T *W, int64_t ld, const double *d_y, double theta, T *x, T *r,
                   double *d_out, double *scratch, hipStream_t s) {
    DMRG_REQUIRE(m <= VEC_MAX_M, DMRG_ERR_ARG, "ritz: too many columns");
    const int g = grid_for(n);
    if(g <= FUSE_MAX_BLOCKS) {
        hipLaunchKernelGGL(ritz_kernel<T>, dim3(g), dim3(VT_THREADS), 0, s, n, m, V, W, ld, d_y, theta, x, r, scratch,
                           scratch_counter(scratch), d_out);
    } else {
        hipLaunchKernelGGL(ritz_kernel<T>, dim3(g), dim3(VT_THREADS), 0, s, n, m, V, W, ld, d_y, theta, x, r, scratch, nullptr, nullptr);
        hipLaunchKernelGGL(fold_columns, dim3(1), dim3(VT_THREADS), 0, s, g, 1, scratch, d_out, -1, nullptr, nullptr);
    }
    DMRG_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------------------------------- elementwise
template <typename T>
__global__ void scale_inv_norm_kernel(int64_t n, T *x, const double *nrm2) {
    const double a = nrm2[0] > 0.0 ? 1.0 / sqrt(nrm2[0]) : 0.0; // a zero vector stays zero (breakdown tests see it)
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double re, im;
        ld(x[i], re, im);
        x[i] = mk<T>(a * re, a * im);
    }
}
template <typename T>
void scale_by_inv_norm(int64_t n, T *x, const double *d_nrm2, hipStream_t s) {
    hipLaunchKernelGGL(scale_inv_norm_kernel<T>, dim3(ew_grid(n)), dim3(VT_THREADS), 0, s, n, x, d_nrm2);
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ void scale_kernel(int64_t n, T *x, double a) {
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double re, im;
        ld(x[i], re, im);
        x[i] = mk<T>(a * re, a * im);
    }
}
template <typename T>
void scale(int64_t n, T *x, double alpha, hipStream_t s) {
    hipLaunchKernelGGL(scale_kernel<T>, dim3(ew_grid(n)), dim3(VT_THREADS), 0, s, n, x, alpha);
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename T, typename D>
__global__ void jacobi_kernel(int64_t n, const T *r, const D *diag, double theta, double floor, T *t) {
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double rr, ri, dr, di;
        ld(r[i], rr, ri);
        ld(diag[i], dr, di);
        double den = dr - theta;
        if(fabs(den) < floor) den = den < 0 ? -floor : floor;
        t[i] = mk<T>(rr / den, ri / den);
    }
}
template <typename T, typename D>
void jacobi_precond(int64_t n, const T *r, const D *diag, double theta, double floor, T *t, hipStream_t s) {
    hipLaunchKernelGGL((jacobi_kernel<T, D>), dim3(ew_grid(n)), dim3(VT_THREADS), 0, s, n, r, diag, theta, floor, t);
    DMRG_HIP_CHECK(hipGetLastError());
}

__device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ inline double u11(uint64_t z) { return static_cast<double>(z >> 11) * (2.0 / 9007199254740992.0) - 1.0; }

template <typename T>
__global__ void random_kernel(int64_t n, T *x, uint64_t seed) {
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint64_t h = splitmix64(seed * 0x100000001B3ull + static_cast<uint64_t>(i));
        x[i]             = mk<T>(u11(h), vtraits<T>::cx ? u11(splitmix64(h)) : 0.0);
    }
}
template <typename T>
void random_fill(int64_t n, T *x, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(random_kernel<T>, dim3(ew_grid(n)), dim3(VT_THREADS), 0, s, n, x, seed);
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename TO, typename TI>
__global__ void convert_kernel(int64_t n, const TI *in, TO *out) {
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double re, im;
        ld(in[i], re, im);
        out[i] = mk<TO>(re, im);
    }
}
template <typename TO, typename TI>
void convert(int64_t n, const TI *in, TO *out, hipStream_t s) {
    hipLaunchKernelGGL((convert_kernel<TO, TI>), dim3(ew_grid(n)), dim3(VT_THREADS), 0, s, n, in, out);
    DMRG_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------------------------------- projected CG
// Kernels of the FP32 correction-equation solver (CG on A = sigma P (H - theta) P, P = I - x x^H). The scalars
// live on the device: dots = {x^H q (2), p^H q (2)}; state as laid out by the CG_* offsets. Every block of a
// kernel derives the same step from the same inputs; only block 0 writes the state, into the slot of the
// next iteration (k + 1) & 1, which no block of the current kernel reads.

template <typename T>
__global__ __launch_bounds__(VT_THREADS) void cg_shift_kernel(int64_t n, const T *p, T *q, double theta, double sigma, const double *state,
                                                              int k) {
    if(state[CG_STOP + (k & 1)] != 0.0) return;
    for(int64_t i = blockIdx.x * static_cast<int64_t>(VT_THREADS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * VT_THREADS) {
        double qr, qi, pr, pi;
        ld(q[i], qr, qi);
        ld(p[i], pr, pi);
        q[i] = mk<T>(sigma * (qr - theta * pr), sigma * (qi - theta * pi));
    }
}

// r = -sigma (r - x c) with c = x^H r (dots[0..1]); p = r; t = 0; partial |r|^2 per block.
template <typename T>
__global__ __launch_bounds__(VT_THREADS) void cg_init_kernel(int64_t n, const T *x, const double *c, double sigma, T *r, T *p, T *t,
                                                             double *partial) {
    __shared__ double red[1][VT_THREADS / 64];
    const double      cr = c[0], ci = c[1];
    double            acc[1] = {0.0};
    for(int64_t i = blockIdx.x * static_cast<int64_t>(VT_THREADS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * VT_THREADS) {
        double rr, ri, xr, xi;
        ld(r[i], rr, ri);
        ld(x[i], xr, xi);
        const double vr = -sigma * (rr - (xr * cr - xi * ci)), vi = -sigma * (ri - (xr * ci + xi * cr));
        const T      v  = mk<T>(vr, vi);
        r[i]            = v;
        p[i]            = v;
        t[i]            = mk<T>(0.0, 0.0);
        double wr, wi;
        ld(v, wr, wi);
        acc[0] += wr * wr + wi * wi;
    }
    bsum<1>(acc, red);
    if(threadIdx.x == 0) {
        partial[2 * blockIdx.x]     = acc[0];
        partial[2 * blockIdx.x + 1] = 0.0;
    }
}

// state <- start of the iteration: rho_0 = |r|^2 (rr[0]); stopped when r = 0.
__global__ void cg_start_kernel(const double *rr, double *state) {
    state[CG_RHO]      = rr[0];
    state[CG_RHO + 1]  = 0.0;
    state[CG_RHO0]     = rr[0];
    state[CG_STOP]     = rr[0] > 0.0 ? 0.0 : 1.0;
    state[CG_STOP + 1] = 0.0;
    state[CG_ITS]      = 0.0;
    state[CG_RR]       = rr[0];
}

// alpha = rho / p^H q; t += alpha p; r -= alpha (q - x (x^H q)); partial |r|^2. No update on a breakdown (p^H q <= 0).
template <typename T>
__global__ __launch_bounds__(VT_THREADS) void cg_update_kernel(int64_t n, const T *x, const T *p, const T *q, const double *dots,
                                                               const double *state, int k, T *t, T *r, double *partial) {
    __shared__ double red[1][VT_THREADS / 64];
    const double      pq   = dots[2];
    const bool        live = state[CG_STOP + (k & 1)] == 0.0 && pq > 0.0;
    const double      a    = live ? state[CG_RHO + (k & 1)] / pq : 0.0;
    const double      cr = dots[0], ci = dots[1];
    double            acc[1] = {0.0};
    if(live) {
        for(int64_t i = blockIdx.x * static_cast<int64_t>(VT_THREADS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * VT_THREADS) {
            double xr, xi, pr, pi, qr, qi, tr, ti, rr, ri;
            ld(x[i], xr, xi);
            ld(p[i], pr, pi);
            ld(q[i], qr, qi);
            ld(t[i], tr, ti);
            ld(r[i], rr, ri);
            t[i]            = mk<T>(tr + a * pr, ti + a * pi);
            const double yr = qr - (xr * cr - xi * ci), yi = qi - (xr * ci + xi * cr);
            const T      v  = mk<T>(rr - a * yr, ri - a * yi);
            r[i]            = v;
            double wr, wi;
            ld(v, wr, wi);
            acc[0] += wr * wr + wi * wi;
        }
    }
    bsum<1>(acc, red);
    if(threadIdx.x == 0) {
        partial[2 * blockIdx.x]     = acc[0];
        partial[2 * blockIdx.x + 1] = 0.0;
    }
}

// beta = |r|^2 / rho; p = r + beta p; block 0 advances the state (rho, stop flag, step count).
template <typename T>
__global__ __launch_bounds__(VT_THREADS) void cg_direction_kernel(int64_t n, const T *r, const double *dots, const double *rr,
                                                                  double *state, int k, double rtol2, T *p) {
    const bool   stopped = state[CG_STOP + (k & 1)] != 0.0;
    const bool   live    = !stopped && dots[2] > 0.0;
    const double rho     = state[CG_RHO + (k & 1)];
    if(blockIdx.x == 0 && threadIdx.x == 0) {
        const int nx          = (k + 1) & 1;
        state[CG_RHO + nx]    = live ? rr[0] : rho;
        state[CG_STOP + nx]   = (!live || rr[0] <= rtol2 * state[CG_RHO0]) ? 1.0 : 0.0;
        state[CG_ITS]         = state[CG_ITS] + (live ? 1.0 : 0.0);
        if(live) state[CG_RR] = rr[0];
    }
    if(!live) return;
    const double b = rr[0] / rho;
    for(int64_t i = blockIdx.x * static_cast<int64_t>(VT_THREADS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * VT_THREADS) {
        double rr_, ri, pr, pi;
        ld(r[i], rr_, ri);
        ld(p[i], pr, pi);
        p[i] = mk<T>(rr_ + b * pr, ri + b * pi);
    }
}

template <typename T>
void cg_init(int64_t n, const T *x, const double *d_c, double sigma, T *r, T *p, T *t, double *d_rr, double *state, double *scratch,
             hipStream_t s) {
    const int g = grid_for(n);
    hipLaunchKernelGGL(cg_init_kernel<T>, dim3(g), dim3(VT_THREADS), 0, s, n, x, d_c, sigma, r, p, t, scratch);
    hipLaunchKernelGGL(fold_columns, dim3(1), dim3(VT_THREADS), 0, s, g, 1, scratch, d_rr, -1, nullptr, nullptr);
    hipLaunchKernelGGL(cg_start_kernel, dim3(1), dim3(1), 0, s, d_rr, state);
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void cg_shift(int64_t n, const T *p, T *q, double theta, double sigma, const double *state, int k, hipStream_t s) {
    hipLaunchKernelGGL(cg_shift_kernel<T>, dim3(grid_for(n)), dim3(VT_THREADS), 0, s, n, p, q, theta, sigma, state, k);
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void cg_step(int64_t n, const T *x, T *p, const T *q, const double *d_dots, T *t, T *r, double *d_rr, double *state, int k,
             double rtol2, double *scratch, hipStream_t s) {
    const int g = grid_for(n);
    hipLaunchKernelGGL(cg_update_kernel<T>, dim3(g), dim3(VT_THREADS), 0, s, n, x, p, q, d_dots, state, k, t, r, scratch);
    hipLaunchKernelGGL(fold_columns, dim3(1), dim3(VT_THREADS), 0, s, g, 1, scratch, d_rr, -1, nullptr, nullptr);
    hipLaunchKernelGGL(cg_direction_kernel<T>, dim3(g), dim3(VT_THREADS), 0, s, n, r, d_dots, d_rr, state, k, rtol2, p);
    DMRG_HIP_CHECK(hipGetLastError());
}

#define DMRG_VEC_INST(T)                                                                                                 \
    template void multidot<T>(int64_t, int, const T *, int64_t, const T *, double *, double *, hipStream_t);          \
    template void gemv_update<T>(int64_t, int, const T *, int64_t, const double *, double, double, T *, hipStream_t); \
    template void ritz_residual<T>(int64_t, int, const T *, const T *, int64_t, const double *, double, T *, T *,      \
                                   double *, double *, hipStream_t);                                                   \
    template void scale_by_inv_norm<T>(int64_t, T *, const double *, hipStream_t);                                    \
    template void scale<T>(int64_t, T *, double, hipStream_t);                                                         \
    template void random_fill<T>(int64_t, T *, uint64_t, hipStream_t);                                               \
    template void multidot_x<T>(int64_t, int, const T *, int64_t, const T *, double *, double *, const int *, double *,  \
                                hipStream_t);                                                                          \
    template void multidot_dgks<T>(int64_t, int, const T *, int64_t, const T *, double *, const double *, const double *, \
                                   double, int *, double *, hipStream_t);                                              \
    template void gemv_update_norm<T>(int64_t, int, const T *, int64_t, const double *, double, double, const T *, T *,  \
                                      double *, const int *, double *, hipStream_t);
DMRG_VEC_INST(double)
DMRG_VEC_INST(float)
#define DMRG_CG_INST(T)                                                                                                  \
    template void cg_init<T>(int64_t, const T *, const double *, double, T *, T *, T *, double *, double *, double *,     \
                             hipStream_t);                                                                               \
    template void cg_shift<T>(int64_t, const T *, T *, double, double, const double *, int, hipStream_t);               \
    template void cg_step<T>(int64_t, const T *, T *, const T *, const double *, T *, T *, double *, double *, int, double, \
                             double *, hipStream_t);
DMRG_CG_INST(float)
DMRG_CG_INST(cpx<float>)
DMRG_VEC_INST(cpx<double>)
DMRG_VEC_INST(cpx<float>)
template void jacobi_precond<double, double>(int64_t, const double *, const double *, double, double, double *, hipStream_t);
template void jacobi_precond<float, double>(int64_t, const float *, const double *, double, double, float *, hipStream_t);
template void jacobi_precond<cpx<double>, cpx<double>>(int64_t, const cpx<double> *, const cpx<double> *, double, double,
                                                       cpx<double> *, hipStream_t);
template void jacobi_precond<cpx<float>, cpx<double>>(int64_t, const cpx<float> *, const cpx<double> *, double, double,
                                                      cpx<float> *, hipStream_t);
template void convert<float, double>(int64_t, const double *, float *, hipStream_t);
template void convert<double, float>(int64_t, const float *, double *, hipStream_t);
template void convert<cpx<float>, cpx<double>>(int64_t, const cpx<double> *, cpx<float> *, hipStream_t);
template void convert<cpx<double>, cpx<float>>(int64_t, const cpx<float> *, cpx<double> *, hipStream_t);
template void convert<double, double>(int64_t, const double *, double *, hipStream_t);
template void convert<cpx<double>, cpx<double>>(int64_t, const cpx<double> *, cpx<double> *, hipStream_t);

} // namespace dmrg

// Vector reductions and elementwise kernels (HBM bound).
// Reductions are deterministic: a fixed grid writes per-block partials that one
// block folds in a fixed order, so repeated calls give bit-identical results.
#include "contract.hpp"

#include <algorithm>

namespace dmrg {

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS  = 512;

template <typename R>
__device__ inline R wave_sum(R v) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of NV values per thread; result valid in thread 0.
template <typename R, int NV>
__device__ inline void block_sum(R (&v)[NV]) {
    __shared__ R red[NV][RED_THREADS / 64];
    const int    lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for(int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if(lane == 0)
#pragma unroll
        for(int i = 0; i < NV; ++i) red[i][wave] = v[i];
    __syncthreads();
    if(threadIdx.x == 0)
#pragma unroll
        for(int i = 0; i < NV; ++i) {
            R s = 0;
            for(int w = 0; w < RED_THREADS / 64; ++w) s += red[i][w];
            v[i] = s;
        }
}

// partial[blk] = sum conj(x) y over a grid-stride slice (complex: 2 values, real: 1).
template <typename R, bool CPLX>
__global__ __launch_bounds__(RED_THREADS) void dot_partial(int64_t n, const R *x, const R *y, double *partial) {
    double acc[2] = {0.0, 0.0};
    for(int64_t i = blockIdx.x * static_cast<int64_t>(RED_THREADS) + threadIdx.x; i < n;
        i += static_cast<int64_t>(gridDim.x) * RED_THREADS) {
        if constexpr(CPLX) {
            const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
            acc[0] += xr * yr + xi * yi;
            acc[1] += xr * yi - xi * yr;
        } else {
            acc[0] += static_cast<double>(x[i]) * static_cast<double>(y[i]);
        }
    }
    block_sum<double, 2>(acc);
    if(threadIdx.x == 0) {
        partial[2 * blockIdx.x]     = acc[0];
        partial[2 * blockIdx.x + 1] = acc[1];
    }
}

__global__ __launch_bounds__(RED_THREADS) void fold_partials(int nblk, const double *partial, double *out, int nout) {
    double acc[2] = {0.0, 0.0};
    for(int i = threadIdx.x; i < nblk; i += RED_THREADS) {
        acc[0] += partial[2 * i];
        acc[1] += partial[2 * i + 1];
    }
    block_sum<double, 2>(acc);
    if(threadIdx.x == 0) {
        out[0] = acc[0];
        if(nout > 1) out[1] = acc[1];
    }
}

static int red_blocks(int64_t n) {
    const int64_t b = (n + RED_THREADS - 1) / RED_THREADS;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, b)));
}

// d_out must hold 2 + 2*RED_BLOCKS doubles: [0..1] result, then partials.
void dot_device(dmrg_dtype t, int64_t n, const void *x, const void *y, double *d_out, hipStream_t s) {
    const int nb      = red_blocks(n);
    double   *partial = d_out + 2;
    switch(t) {
        case DMRG_FP64:
            hipLaunchKernelGGL((dot_partial<double, false>), dim3(nb), dim3(RED_THREADS), 0, s, n,
                               static_cast<const double *>(x), static_cast<const double *>(y), partial);
            break;
        case DMRG_CX64:
            hipLaunchKernelGGL((dot_partial<double, true>), dim3(nb), dim3(RED_THREADS), 0, s, n,
                               static_cast<const double *>(x), static_cast<const double *>(y), partial);
            break;
        case DMRG_FP32:
            hipLaunchKernelGGL((dot_partial<float, false>), dim3(nb), dim3(RED_THREADS), 0, s, n,
                               static_cast<const float *>(x), static_cast<const float *>(y), partial);
            break;
        case DMRG_CX32:
            throw Error(DMRG_ERR_UNSUPPORTED, "dot: CX32 not supported");
    }
    hipLaunchKernelGGL(fold_partials, dim3(1), dim3(RED_THREADS), 0, s, nb, partial, d_out, 2);
    DMRG_HIP_CHECK(hipGetLastError());
}

void nrm2sq_device(dmrg_dtype t, int64_t n, const void *x, double *d_out, hipStream_t s) { dot_device(t, n, x, x, d_out, s); }

} // namespace dmrg

namespace dmrg {

// ---------------------------------------------------------------------------------------
// Precision conversion (FP64 <-> FP32), n real elements (a complex vector is 2n reals).

__global__ void f64_to_f32_kernel(int64_t n, const double *src, float *dst) {
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dst[i] = static_cast<float>(src[i]);
}
__global__ void f32_to_f64_kernel(int64_t n, const float *src, double *dst) {
    for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dst[i] = static_cast<double>(src[i]);
}

static dim3 ew_grid(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return dim3(static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, 4096))));
}

void convert_f64_to_f32(int64_t nreal, const double *src, float *dst, hipStream_t s) {
    if(nreal <= 0) return;
    hipLaunchKernelGGL(f64_to_f32_kernel, ew_grid(nreal), dim3(256), 0, s, nreal, src, dst);
    DMRG_HIP_CHECK(hipGetLastError());
}
void convert_f32_to_f64(int64_t nreal, const float *src, double *dst, hipStream_t s) {
    if(nreal <= 0) return;
    hipLaunchKernelGGL(f32_to_f64_kernel, ew_grid(nreal), dim3(256), 0, s, nreal, src, dst);
    DMRG_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// Diagonal of H_eff: diag[s,a,b] = sum_{w,w'} L[a,a,w] W[w,w',s,s] R[b,b,w'].
// Pass 1: P[a,w',s] = sum_w L[a,a,w] W[w,w',s,s];  pass 2: diag[s,a,b] = sum_w' P[a,w',s] R[b,b,w'].

template <typename R, bool CPLX>
struct ScalarOps;
template <typename R>
struct ScalarOps<R, false> {
    using T = R;
    __device__ static T zero() { return R(0); }
    __device__ static T mul(T a, T b) { return a * b; }
    __device__ static T add(T a, T b) { return a + b; }
};
template <typename R>
struct ScalarOps<R, true> {
    using T = cpx<R>;
    __device__ static T zero() { return T{R(0), R(0)}; }
    __device__ static T mul(T a, T b) { return T{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
    __device__ static T add(T a, T b) { return T{a.re + b.re, a.im + b.im}; }
};

template <typename R, bool CPLX>
__global__ void diag_pass1(int64_t d, int64_t cL, int64_t DL, int64_t DR, const typename ScalarOps<R, CPLX>::T *L,
                           const typename ScalarOps<R, CPLX>::T *W, typename ScalarOps<R, CPLX>::T *P) {
    using O         = ScalarOps<R, CPLX>;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; // i = a + cL*(w' + DR*s)
    if(i >= cL * DR * d) return;
    const int64_t a = i % cL, wp = (i / cL) % DR, s = i / (cL * DR);
    auto          acc = O::zero();
    for(int64_t w = 0; w < DL; ++w)
        acc = O::add(acc, O::mul(L[a + cL * (a + cL * w)], W[w + DL * (wp + DR * (s + d * s))]));
    P[i] = acc;
}

template <typename R, bool CPLX>
__global__ void diag_pass2(int64_t d, int64_t cL, int64_t cR, int64_t DR, const typename ScalarOps<R, CPLX>::T *P,
                           const typename ScalarOps<R, CPLX>::T *Rb, typename ScalarOps<R, CPLX>::T *out) {
    using O         = ScalarOps<R, CPLX>;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; // i = s + d*(a + cL*b)
    if(i >= d * cL * cR) return;
    const int64_t s = i % d, a = (i / d) % cL, b = i / (d * cL);
    auto          acc = O::zero();
    for(int64_t wp = 0; wp < DR; ++wp) acc = O::add(acc, O::mul(P[a + cL * (wp + DR * s)], Rb[b + cR * (b + cR * wp)]));
    out[i] = acc;
}

void heff_diagonal(dmrg_dtype t, const Dims &D, const void *mpo, const void *envL, const void *envR, void *ws, void *out,
                   hipStream_t s) {
    const int64_t np = D.chiL * D.DR * D.d, nd = D.n();
    const dim3    g1((np + 255) / 256), g2((nd + 255) / 256), blk(256);
    switch(t) {
        case DMRG_FP64:
            hipLaunchKernelGGL((diag_pass1<double, false>), g1, blk, 0, s, D.d, D.chiL, D.DL, D.DR, static_cast<const double *>(envL),
                               static_cast<const double *>(mpo), static_cast<double *>(ws));
            hipLaunchKernelGGL((diag_pass2<double, false>), g2, blk, 0, s, D.d, D.chiL, D.chiR, D.DR, static_cast<const double *>(ws),
                               static_cast<const double *>(envR), static_cast<double *>(out));
            break;
        case DMRG_CX64:
            hipLaunchKernelGGL((diag_pass1<double, true>), g1, blk, 0, s, D.d, D.chiL, D.DL, D.DR,
                               static_cast<const cpx<double> *>(envL), static_cast<const cpx<double> *>(mpo),
                               static_cast<cpx<double> *>(ws));
            hipLaunchKernelGGL((diag_pass2<double, true>), g2, blk, 0, s, D.d, D.chiL, D.chiR, D.DR,
                               static_cast<const cpx<double> *>(ws), static_cast<const cpx<double> *>(envR),
                               static_cast<cpx<double> *>(out));
            break;
        default: throw Error(DMRG_ERR_UNSUPPORTED, "diagonal: FP64/CX64 only");
    }
    DMRG_HIP_CHECK(hipGetLastError());
}

} // namespace dmrg

// Explicit H_eff matrices (for the dense eigensolver branch) and chain-end environment blocks.
#include "dense.hpp"

#include <algorithm>
#include <cstring>

namespace dmrg {

namespace {

template <typename T>
__device__ __forceinline__ void fma_acc(T &acc, T a, T b) {
    if constexpr(std::is_same_v<T, double>) acc = fma(a, b, acc);
    else {
        acc.re = fma(a.re, b.re, fma(-a.im, b.im, acc.re));
        acc.im = fma(a.re, b.im, fma(a.im, b.re, acc.im));
    }
}

// H_eff[r, c] = sum_{w, w'} L[a, a', w] W[w, w', s, s'] R[b, b', w'], r = (s', a', b'), c = (s, a, b).
template <typename T>
__device__ __forceinline__ T heff_element(int64_t d, int64_t cL, int64_t cR, int64_t DL, int64_t DR, const T *__restrict__ W,
                                          const T *__restrict__ Lb, const T *__restrict__ Rb, int64_t r, int64_t c) {
    const int64_t sp = r % d, ap = (r / d) % cL, bp = r / (d * cL);
    const int64_t s  = c % d, a = (c / d) % cL, b = c / (d * cL);
    T             acc{};
    for(int64_t wr = 0; wr < DR; ++wr) {
        T lw{};
        for(int64_t wl = 0; wl < DL; ++wl)
            fma_acc(lw, Lb[a + cL * (ap + cL * wl)], W[wl + DL * (wr + DR * (s + d * sp))]);
        fma_acc(acc, lw, Rb[b + cR * (bp + cR * wr)]);
    }
    return acc;
}

// One thread per output element; consecutive threads walk down a column of the output, so stores coalesce.
template <typename T>
__global__ void __launch_bounds__(256) heff_dense_kernel(int64_t d, int64_t cL, int64_t cR, int64_t DL, int64_t DR,
                                                         const T *__restrict__ W, const T *__restrict__ Lb,
                                                         const T *__restrict__ Rb, T *__restrict__ H, int64_t ldh,
                                                         int transpose) {
    const int64_t n  = d * cL * cR;
    const int64_t i  = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; // output row
    const int64_t j  = blockIdx.y;                                                   // output column
    if(i >= n) return;
    const int64_t r  = transpose ? j : i; // (s', a', b')
    const int64_t c  = transpose ? i : j; // (s, a, b)
    H[i + ldh * j]   = heff_element(d, cL, cR, DL, DR, W, Lb, Rb, r, c);
}

template <typename T>
__global__ void __launch_bounds__(256) diag_blocks_kernel(int64_t d, int64_t cL, int64_t cR, int64_t DL, int64_t DR,
                                                          const T *__restrict__ W, const T *__restrict__ Lb,
                                                          const T *__restrict__ Rb, int64_t jcb, int64_t nb, double shift,
                                                          T *__restrict__ out) {
    const int64_t n     = d * cL * cR;
    const int64_t bsz   = jcb * jcb;
    const int64_t total = nb * bsz;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < total;
        t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t blk = t / bsz, e = t - blk * bsz;
        const int64_t i = e % jcb, j = e / jcb;
        const int64_t r = blk * jcb + i, c = blk * jcb + j;
        T             v{};
        if(r < n && c < n) {
            v = heff_element(d, cL, cR, DL, DR, W, Lb, Rb, r, c);
            if(i == j) {
                if constexpr(std::is_same_v<T, double>) v -= shift;
                else v.re -= shift;
            }
        } else if(i == j) {
            if constexpr(std::is_same_v<T, double>) v = 1.0;
            else v.re = 1.0;
        }
        out[t] = v;
    }
}

template <typename T>
struct EdgeArg {
    T v[EDGE_MAX];
};

template <typename T>
__global__ void __launch_bounds__(256) edge_kernel(int64_t chi, int64_t Dm, EdgeArg<T> e, T *__restrict__ out) {
    const int64_t total = chi * chi * Dm;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < total;
        t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t a = t % chi, ap = (t / chi) % chi, w = t / (chi * chi);
        out[t]          = (a == ap) ? e.v[w] : T{};
    }
}

template <typename T>
void launch_dense(const Dims &D, const void *mpo, const void *envL, const void *envR, void *H, int64_t ldh, bool tr,
                  hipStream_t s) {
    const int64_t n = D.n();
    const dim3    g(static_cast<unsigned>((n + 255) / 256), static_cast<unsigned>(n)), b(256);
    hipLaunchKernelGGL(heff_dense_kernel<T>, g, b, 0, s, D.d, D.chiL, D.chiR, D.DL, D.DR, static_cast<const T *>(mpo),
                       static_cast<const T *>(envL), static_cast<const T *>(envR), static_cast<T *>(H), ldh, tr ? 1 : 0);
    DMRG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_edge(void *out, int64_t chi, int64_t Dm, const void *edge_host, hipStream_t s) {
    EdgeArg<T> e{};
    std::memcpy(e.v, edge_host, static_cast<size_t>(Dm) * sizeof(T));
    const int64_t total = chi * chi * Dm;
    const dim3    g(static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 8192))), b(256);
    hipLaunchKernelGGL(edge_kernel<T>, g, b, 0, s, chi, Dm, e, static_cast<T *>(out));
    DMRG_HIP_CHECK(hipGetLastError());
}

} // namespace

void heff_diag_blocks(dmrg_dtype t, const Dims &D, const void *mpo, const void *envL, const void *envR, int64_t jcb,
                      double shift, void *out, hipStream_t s) {
    DMRG_REQUIRE(jcb >= 1, DMRG_ERR_ARG, "block size must be >= 1");
    const int64_t n = D.n(), nb = (n + jcb - 1) / jcb, total = nb * jcb * jcb;
    const dim3    g(static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 65536))), b(256);
    if(t == DMRG_FP64)
        hipLaunchKernelGGL(diag_blocks_kernel<double>, g, b, 0, s, D.d, D.chiL, D.chiR, D.DL, D.DR, static_cast<const double *>(mpo),
                           static_cast<const double *>(envL), static_cast<const double *>(envR), jcb, nb, shift,
                           static_cast<double *>(out));
    else if(t == DMRG_CX64)
        hipLaunchKernelGGL(diag_blocks_kernel<cpx<double>>, g, b, 0, s, D.d, D.chiL, D.chiR, D.DL, D.DR,
                           static_cast<const cpx<double> *>(mpo), static_cast<const cpx<double> *>(envL),
                           static_cast<const cpx<double> *>(envR), jcb, nb, shift, static_cast<cpx<double> *>(out));
    else throw Error(DMRG_ERR_UNSUPPORTED, "diagonal blocks: FP64 or CX64");
    DMRG_HIP_CHECK(hipGetLastError());
}

void heff_dense(dmrg_dtype t, const Dims &D, const void *mpo, const void *envL, const void *envR, void *H, int64_t ldh,
                bool transpose, hipStream_t s) {
    const int64_t n = D.n();
    DMRG_REQUIRE(ldh >= n, DMRG_ERR_DIMS, "ldh must be >= rows of H_eff");
    DMRG_REQUIRE(n <= 65535, DMRG_ERR_DIMS, "explicit H_eff is limited to n <= 65535 rows");
    if(t == DMRG_FP64) launch_dense<double>(D, mpo, envL, envR, H, ldh, transpose, s);
    else if(t == DMRG_CX64) launch_dense<cpx<double>>(D, mpo, envL, envR, H, ldh, transpose, s);
    else throw Error(DMRG_ERR_UNSUPPORTED, "explicit H_eff: FP64 or CX64");
}

void env_set_edge(dmrg_dtype t, void *out, int64_t chi, int64_t Dm, const void *edge_host, hipStream_t s) {
    DMRG_REQUIRE(chi > 0 && Dm > 0, DMRG_ERR_DIMS, "edge block needs chi > 0 and D > 0");
    DMRG_REQUIRE(Dm <= EDGE_MAX, DMRG_ERR_DIMS, "edge vector longer than " + std::to_string(EDGE_MAX));
    if(t == DMRG_FP64) launch_edge<double>(out, chi, Dm, edge_host, s);
    else if(t == DMRG_CX64) launch_edge<cpx<double>>(out, chi, Dm, edge_host, s);
    else throw Error(DMRG_ERR_UNSUPPORTED, "edge block: FP64 or CX64");
}

} // namespace dmrg

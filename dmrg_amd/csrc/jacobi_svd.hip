// One-sided (Hestenes) Jacobi SVD on the GPU, the method of LAPACK's gesvj that the reference can select
// (svd_lapacke.cpp:193-212). The columns of W (m x n, m >= n) are rotated pairwise until they are mutually
// orthogonal; then S = column norms, U = normalised columns and V = the accumulated rotations. One-sided
// Jacobi keeps high relative accuracy for small singular values.
//
// Each sweep runs the n(n-1)/2 pairs in n-1 rounds of disjoint pairs (circle method). When W and V fit in
// LDS, one workgroup runs every sweep with both matrices resident and a barrier between rounds. Mid-size
// problems launch one kernel per round with one workgroup (1 to 4 wavefronts, by column length) per pair; large
// ones pair blocks of 16 columns instead (block one-sided Jacobi, three launches per round of block pairs). The
// only host round trip per sweep reads whether any rotation was applied.
#include "svd.hpp"

#include <algorithm>
#include <cmath>

namespace dmrg {

namespace {

template <typename T>
struct jtraits;
template <>
struct jtraits<double> {
    static constexpr bool cx = false;
};
template <>
struct jtraits<cpx<double>> {
    static constexpr bool cx = true;
};

__device__ inline double wave_sum(double v) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Pair of round r (circle method over np players; player np-1 stays fixed).
__device__ inline void round_pair(int np, int r, int i, int &a, int &b) {
    if(i == 0) {
        a = r % (np - 1);
        b = np - 1;
    } else {
        a = (r + i) % (np - 1);
        b = (r - i + (np - 1)) % (np - 1);
    }
}

// One round: workgroup w of the grid (NT threads, a multiple of 64) handles pair w. flag[0] is raised when a
// rotation is applied. Several wavefronts per pair shorten the dependent walk down long columns.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void jacobi_round(int m, int n, int np, int r, T *W, T *V, double tol, unsigned *flag) {
    int p, q;
    round_pair(np, r, blockIdx.x, p, q);
    if(p >= n || q >= n) return; // padding player (odd n); uniform over the workgroup
    if(p > q) {
        const int t = p;
        p           = q;
        q           = t;
    }
    T        *wp = W + static_cast<int64_t>(p) * m, *wq = W + static_cast<int64_t>(q) * m;
    const int tid = threadIdx.x;
    double    al = 0, be = 0, gr = 0, gi = 0;
#pragma unroll 4
    for(int i = tid; i < m; i += NT) {
        if constexpr(jtraits<T>::cx) {
            const T x = wp[i], y = wq[i];
            al += x.re * x.re + x.im * x.im;
            be += y.re * y.re + y.im * y.im;
            gr += x.re * y.re + x.im * y.im; // conj(x) y
            gi += x.re * y.im - x.im * y.re;
        } else {
            const double x = wp[i], y = wq[i];
            al += x * x;
            be += y * y;
            gr += x * y;
        }
    }
    al = wave_sum(al);
    be = wave_sum(be);
    gr = wave_sum(gr);
    gi = wave_sum(gi);
    if constexpr(NT > 64) {
        __shared__ double red[4][NT / 64];
        const int         w = tid >> 6;
        if((tid & 63) == 0) red[0][w] = al, red[1][w] = be, red[2][w] = gr, red[3][w] = gi;
        __syncthreads();
        al = be = gr = gi = 0;
#pragma unroll
        for(int k = 0; k < NT / 64; ++k) al += red[0][k], be += red[1][k], gr += red[2][k], gi += red[3][k];
    }
    const double g = sqrt(gr * gr + gi * gi);
    if(!(g > tol * sqrt(al * be)) || g == 0.0) return; // same decision in every thread of the workgroup
    // Align the phase of column q so that the off-diagonal is |g|, then a real rotation zeroes it:
    // t = sign(z) / (|z| + sqrt(1 + z^2)), z = (be - al) / (2 |g|); w_p' = c w_p - s w_q^, w_q' = s w_p + c w_q^.
    const double z  = (be - al) / (2.0 * g);
    const double t  = (z >= 0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
    const double c  = 1.0 / sqrt(1.0 + t * t), s = c * t;
    const double er = gr / g, ei = -gi / g; // e^{-i phi}
    if(tid == 0) atomicOr(flag, 1u);
    auto rotate = [&](T *x, T *y, int len) {
#pragma unroll 4
        for(int i = tid; i < len; i += NT) {
            if constexpr(jtraits<T>::cx) {
                const T      xv = x[i], yv = y[i];
                const double yr = yv.re * er - yv.im * ei, yi = yv.re * ei + yv.im * er; // e^{-i phi} y
                x[i]            = T{c * xv.re - s * yr, c * xv.im - s * yi};
                y[i]            = T{s * xv.re + c * yr, s * xv.im + c * yi};
            } else {
                const double xv = x[i], yv = y[i] * er;
                x[i]            = c * xv - s * yv;
                y[i]            = s * xv + c * yv;
            }
        }
    };
    rotate(wp, wq, m);
    rotate(V + static_cast<int64_t>(p) * n, V + static_cast<int64_t>(q) * n, n);
}

// ------------------------------------------------------------------------------------------------------------
// Block one-sided Jacobi: columns in blocks of BB = B2 / 2 (W and V padded with zero columns to whole blocks);
// each round pairs the blocks (circle method) and, for every block pair, (1) forms the Gram matrix G = X^H X of
// its B2 columns X on the matrix cores (partial sums over row chunks, bj_gram_mfma), (2) runs cyclic two-sided
// Jacobi on G in LDS, accumulating Q (bj_eig; the 2 x 2 rotation and its threshold are those of jacobi_round,
// computed from the same inner products), (3) rotates X <- X Q and V[:, cols] <- V[:, cols] Q on the matrix cores
// (bj_update_mfma). A sweep meets every column pair, as the scalar sweep does, in n / BB - 1 rounds of 3
// launches instead of n - 1 launches.

template <typename T>
__device__ inline T jconj(const T &a) {
    if constexpr(jtraits<T>::cx) return T{a.re, -a.im};
    else return a;
}
template <typename T>
__device__ inline T jmul(const T &a, const T &b) {
    if constexpr(jtraits<T>::cx) return T{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    else return a * b;
}
template <typename T>
__device__ inline T jzero() {
    if constexpr(jtraits<T>::cx) return T{0.0, 0.0};
    else return 0.0;
}
template <typename T>
__device__ inline T jscale(double a, const T &b) {
    if constexpr(jtraits<T>::cx) return T{a * b.re, a * b.im};
    else return a * b;
}
template <typename T>
__device__ inline T jadd(const T &a, const T &b) {
    if constexpr(jtraits<T>::cx) return T{a.re + b.re, a.im + b.im};
    else return a + b;
}
template <typename T>
__device__ inline T jreal(double a) {
    if constexpr(jtraits<T>::cx) return T{a, 0.0};
    else return a;
}
template <typename T>
__device__ inline double jre(const T &a) {
    if constexpr(jtraits<T>::cx) return a.re;
    else return a;
}

constexpr int BJ_THREADS = 256;
constexpr int BJ_ROWS    = 64; // granularity of the Gram kernel's row chunks

// Sums the partial Gram matrices of a pair (fixed order), diagonalises G by cyclic Jacobi sweeps in LDS and
// writes the accumulated rotation Q (B2 x B2, column-major) and whether any rotation was applied.
template <typename T, int B2>
__global__ __launch_bounds__(BJ_THREADS) void bj_eig(int n, int nbp, int r, int ks, const T *part, T *qbuf, int *rotated,
                                                      double tol, int max_inner, unsigned *flag) {
    constexpr int BB = B2 / 2, NP = B2 / 2;
    int           P, Q;
    round_pair(nbp, r, blockIdx.x, P, Q);
    const int tid = threadIdx.x;
    if(P * BB >= n || Q * BB >= n) {
        if(tid == 0) rotated[blockIdx.x] = 0;
        return;
    }
    __shared__ T      G[B2][B2 + 1], Qs[B2][B2 + 1];
    __shared__ double rc[NP], rs[NP];
    __shared__ T      re[NP];
    __shared__ int    rp[NP], rq[NP], ron[NP];
    __shared__ int    any_sweep, any_total;
    const T          *src = part + static_cast<int64_t>(blockIdx.x) * ks * B2 * B2;
    for(int idx = tid; idx < B2 * B2; idx += BJ_THREADS) {
        const int i = idx % B2, j = idx / B2;
        T         a = jzero<T>();
        for(int c = 0; c < ks; ++c) a = jadd(a, src[static_cast<int64_t>(c) * B2 * B2 + idx]);
        G[i][j]  = a;
        Qs[i][j] = jreal<T>(i == j ? 1.0 : 0.0);
    }
    if(tid == 0) any_total = 0;
    __syncthreads();
    // Hermitian: keep the upper triangle, real diagonal.
    for(int idx = tid; idx < B2 * B2; idx += BJ_THREADS) {
        const int i = idx % B2, j = idx / B2;
        if(i > j) G[i][j] = jconj(G[j][i]);
        else if(i == j) G[i][i] = jreal<T>(jre(G[i][i]));
    }
    __syncthreads();
    for(int sw = 0; sw < max_inner; ++sw) {
        if(tid == 0) any_sweep = 0;
        __syncthreads();
        for(int rr = 0; rr < B2 - 1; ++rr) {
            if(tid < NP) {
                int p, q;
                round_pair(B2, rr, tid, p, q);
                if(p > q) {
                    const int t = p;
                    p           = q;
                    q           = t;
                }
                rp[tid]         = p;
                rq[tid]         = q;
                const double al = jre(G[p][p]), be = jre(G[q][q]);
                const T      gm = G[p][q];
                double       gr, gi;
                if constexpr(jtraits<T>::cx) gr = gm.re, gi = gm.im;
                else gr = gm, gi = 0.0;
                const double g  = sqrt(gr * gr + gi * gi);
                const bool   on = g > tol * sqrt(al * be) && g != 0.0;
                ron[tid]        = on ? 1 : 0;
                if(on) {
                    const double z = (be - al) / (2.0 * g);
                    const double t = (z >= 0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
                    const double c = 1.0 / sqrt(1.0 + t * t);
                    rc[tid]        = c;
                    rs[tid]        = c * t;
                    if constexpr(jtraits<T>::cx) re[tid] = T{gr / g, -gi / g}; // e^{-i phi}
                    else re[tid] = gr / g;
                    any_sweep = 1;
                }
            }
            __syncthreads();
            // rows p, q of G <- J^H G
            for(int k = tid; k < NP * B2; k += BJ_THREADS) {
                const int u = k / B2, col = k % B2;
                if(!ron[u]) continue;
                const int    p = rp[u], q = rq[u];
                const double c = rc[u], sn = rs[u];
                const T      ec = jconj(re[u]);
                const T      a = G[p][col], b = jmul(ec, G[q][col]);
                G[p][col]      = jadd(jscale(c, a), jscale(-sn, b));
                G[q][col]      = jadd(jscale(sn, a), jscale(c, b));
            }
            __syncthreads();
            // columns p, q of G and Q <- . J
            for(int k = tid; k < 2 * NP * B2; k += BJ_THREADS) {
                const int u = (k / B2) % NP, row = k % B2;
                if(!ron[u]) continue;
                const int    p = rp[u], q = rq[u];
                const double c = rc[u], sn = rs[u];
                const bool onG = k < NP * B2;
                T(*M)[B2 + 1]  = onG ? G : Qs;
                const T a = M[row][p], b = jmul(re[u], M[row][q]);
                const T np_ = jadd(jscale(c, a), jscale(-sn, b)), nq_ = jadd(jscale(sn, a), jscale(c, b));
                // the rotation annihilates G[p][q] and G[q][p]: store the exact zero there
                M[row][p] = (onG && row == q) ? jzero<T>() : np_;
                M[row][q] = (onG && row == p) ? jzero<T>() : nq_;
            }
            __syncthreads();
        }
        const int done = !any_sweep;
        if(tid == 0 && any_sweep) any_total = 1;
        __syncthreads();
        if(done) break;
    }
    if(tid == 0) {
        rotated[blockIdx.x] = any_total;
        if(any_total) atomicOr(flag, 1u);
    }
    if(any_total) {
        T *qo = qbuf + static_cast<int64_t>(blockIdx.x) * B2 * B2;
        for(int idx = tid; idx < B2 * B2; idx += BJ_THREADS) qo[idx] = Qs[idx % B2][idx / B2];
    }
}

template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
    for(int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

// Whole Jacobi iteration in one workgroup with W (m x n) and V (n x n) resident in LDS, for small problems
// where a kernel launch per round would dominate. A group of G lanes handles one pair; the groups take the
// pairs of a round in turn, a barrier separates the rounds, and the sweeps stop after a sweep that applied no
// rotation (at most max_sweeps; the number of sweeps run goes to sweeps_out).
template <typename T, int G>
__global__ __launch_bounds__(1024) void jacobi_lds(int m, int n, int np, T *W, T *V, double tol, int max_sweeps, int *sweeps_out) {
    extern __shared__ unsigned char lds_raw[];
    T            *Ws = reinterpret_cast<T *>(lds_raw);
    T            *Vs = Ws + static_cast<int64_t>(m) * n;
    __shared__ int rotated;
    const int     tid = threadIdx.x, lane = tid % G, grp = tid / G, ngrp = blockDim.x / G;
    for(int t = tid; t < m * n; t += blockDim.x) Ws[t] = W[t];
    for(int t = tid; t < n * n; t += blockDim.x) {
        const bool d = (t % n) == (t / n);
        if constexpr(jtraits<T>::cx) Vs[t] = T{d ? 1.0 : 0.0, 0.0};
        else Vs[t] = d ? 1.0 : 0.0;
    }
    int sweep = 0;
    for(; sweep < max_sweeps; ++sweep) {
        if(tid == 0) rotated = 0;
        __syncthreads();
        for(int r = 0; r < np - 1; ++r) {
            // every group runs the same trip count so that the shuffles see whole groups
            for(int i0 = 0; i0 < np / 2; i0 += ngrp) {
                const int i = i0 + grp;
                int       p = 0, q = 0;
                bool      live = i < np / 2;
                if(live) {
                    round_pair(np, r, i, p, q);
                    live = p < n && q < n;
                }
                if(p > q) {
                    const int tt = p;
                    p            = q;
                    q            = tt;
                }
                T     *wp = Ws + p * m, *wq = Ws + q * m;
                double al = 0, be = 0, gr = 0, gi = 0;
                if(live) {
                    for(int k = lane; k < m; k += G) {
                        if constexpr(jtraits<T>::cx) {
                            const T x = wp[k], y = wq[k];
                            al += x.re * x.re + x.im * x.im;
                            be += y.re * y.re + y.im * y.im;
                            gr += x.re * y.re + x.im * y.im;
                            gi += x.re * y.im - x.im * y.re;
                        } else {
                            const double x = wp[k], y = wq[k];
                            al += x * x;
                            be += y * y;
                            gr += x * y;
                        }
                    }
                }
                al = group_sum<G>(al);
                be = group_sum<G>(be);
                gr = group_sum<G>(gr);
                if constexpr(jtraits<T>::cx) gi = group_sum<G>(gi);
                const double g = sqrt(gr * gr + gi * gi);
                if(!live || !(g > tol * sqrt(al * be)) || g == 0.0) continue;
                const double z  = (be - al) / (2.0 * g);
                const double t  = (z >= 0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
                const double c  = 1.0 / sqrt(1.0 + t * t), s = c * t;
                const double er = gr / g, ei = -gi / g;
                if(lane == 0) rotated = 1;
                auto rot = [&](T *x, T *y, int len) {
                    for(int k = lane; k < len; k += G) {
                        if constexpr(jtraits<T>::cx) {
                            const T      xv = x[k], yv = y[k];
                            const double yr = yv.re * er - yv.im * ei, yi = yv.re * ei + yv.im * er;
                            x[k]            = T{c * xv.re - s * yr, c * xv.im - s * yi};
                            y[k]            = T{s * xv.re + c * yr, s * xv.im + c * yi};
                        } else {
                            const double xv = x[k], yv = y[k] * er;
                            x[k]            = c * xv - s * yv;
                            y[k]            = s * xv + c * yv;
                        }
                    }
                };
                rot(wp, wq, m);
                rot(Vs + p * n, Vs + q * n, n);
            }
            __syncthreads();
        }
        const int any = rotated;
        __syncthreads();
        if(any == 0) break;
    }
    for(int t = tid; t < m * n; t += blockDim.x) W[t] = Ws[t];
    for(int t = tid; t < n * n; t += blockDim.x) V[t] = Vs[t];
    if(tid == 0) sweeps_out[0] = sweep;
}

template <typename T>
__global__ void identity_kernel(int n, T *V) {
    const int64_t total = static_cast<int64_t>(n) * n;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < total; t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const bool d = (t % n) == (t / n);
        if constexpr(jtraits<T>::cx) V[t] = T{d ? 1.0 : 0.0, 0.0};
        else V[t] = d ? 1.0 : 0.0;
    }
}

// out[j + n i] = conj(in[i + m j]) (conjugate transpose of an m x n matrix).
template <typename T>
__global__ void adjoint_kernel(int m, int n, const T *in, T *out) {
    const int64_t total = static_cast<int64_t>(m) * n;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < total; t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        T             v = in[t];
        if constexpr(jtraits<T>::cx) v.im = -v.im;
        out[j + static_cast<int64_t>(n) * i] = v;
    }
}

__global__ void column_norms_kernel(int m, int n, const double *W, int cx, double *out) {
    // out[j] = |W[:, j]| (W viewed as doubles; complex columns are 2m doubles long)
    const int j = blockIdx.x;
    if(j >= n) return;
    const int64_t len = static_cast<int64_t>(m) * (cx ? 2 : 1);
    double        acc = 0;
    for(int64_t i = threadIdx.x; i < len; i += 64) {
        const double v = W[j * len + i];
        acc += v * v;
    }
    acc = wave_sum(acc);
    if(threadIdx.x == 0) out[j] = sqrt(acc);
}

// Assemble the outputs in descending order of S: Uo[:, k] = W[:, perm[k]] / S[k] (m x n),
// Vo[k, :] = conj(V[:, perm[k]])^T (n x n, leading dimension n): A = Uo diag(S) Vo.
template <typename T>
__global__ void assemble_kernel(int m, int n, const int *perm, const double *S, const T *W, const T *V, T *Uo, T *Vo) {
    const int64_t tu = static_cast<int64_t>(m) * n, tv = static_cast<int64_t>(n) * n;
    for(int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < tu + tv; t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if(t < tu) {
            const int64_t i = t % m, k = t / m;
            const double  sk = S[k];
            const double  f  = sk > 0 ? 1.0 / sk : 0.0;
            T             v  = W[i + static_cast<int64_t>(m) * perm[k]];
            if constexpr(jtraits<T>::cx) v = T{v.re * f, v.im * f};
            else v *= f;
            Uo[t] = v;
        } else {
            const int64_t q = t - tu, k = q % n, j = q / n;
            T             v = V[j + static_cast<int64_t>(n) * perm[k]];
            if constexpr(jtraits<T>::cx) v.im = -v.im;
            Vo[q] = v;
        }
    }
}

// LDS budget of the single-workgroup path (gfx950 has 160 KiB per CU).
constexpr size_t JACOBI_LDS_BYTES = 152 * 1024;
constexpr int    JACOBI_LDS_MAX_N = 32;
// Block path from this many columns on (measured on gfx950, complex n x n, ms per SVD, per-pair rounds / block
// rounds: 64: 2.86 / 2.4, 128: 7.4 / 5.8, 256: 19.5 / 14.4, 512: 61 / 35, 2048: 1325 / 289); B2 columns (two
// blocks) per inner problem. From 33 columns there are at least 3 blocks, so every round has work.
constexpr int JACOBI_BLOCK_MIN_N = 33;
constexpr int JACOBI_BLOCK_B2    = 32;
// Inner Jacobi sweeps per block pair and round: one sweep leaves the outer sweep count unchanged (the outer
// iteration finishes what an inexact inner solve leaves) at a fraction of the cost of converging each one.
constexpr int JACOBI_BLOCK_INNER = 1;

unsigned grid_of(int64_t total) { return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 8192))); }

} // namespace

template <typename T>
void DeviceSvd::jacobi(int m, int n, const T *A, std::vector<double> &S) {
    // Work on W = A (m >= n) or W = A^H (m < n); W has mm >= nn.
    const bool tr = m < n;
    const int  mm = tr ? n : m, nn = tr ? m : n;
    k_            = nn;
    const size_t es = sizeof(T);
    const size_t lds_need = (static_cast<size_t>(mm) * nn + static_cast<size_t>(nn) * nn) * es;
    const bool   resident = nn > 1 && nn <= JACOBI_LDS_MAX_N && lds_need <= JACOBI_LDS_BYTES;
    const bool   blocked  = !resident && nn >= JACOBI_BLOCK_MIN_N;
    const int    npad     = blocked ? (nn + 15) / 16 * 16 : nn; // whole 16-column blocks (zero columns)
    a.reserve(static_cast<size_t>(mm) * npad * es);  // W (m x npad)
    e.reserve(static_cast<size_t>(nn) * npad * es);  // V (n x npad)
    sv.reserve(static_cast<size_t>(nn) * sizeof(double));
    info.reserve(sizeof(unsigned) * 2 + static_cast<size_t>(nn) * sizeof(int));
    res.reserve(static_cast<size_t>(nn) * sizeof(double));
    T *W = a.as<T>(), *V = e.as<T>();
    if(tr) hipLaunchKernelGGL(adjoint_kernel<T>, dim3(grid_of(static_cast<int64_t>(m) * n)), dim3(256), 0, s, m, n, A, W);
    else DMRG_HIP_CHECK(hipMemcpyAsync(W, A, static_cast<size_t>(m) * n * es, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(identity_kernel<T>, dim3(grid_of(static_cast<int64_t>(nn) * nn)), dim3(256), 0, s, nn, V);
    DMRG_HIP_CHECK(hipGetLastError());
    if(npad > nn) {
        DMRG_HIP_CHECK(hipMemsetAsync(W + static_cast<int64_t>(mm) * nn, 0, static_cast<size_t>(mm) * (npad - nn) * es, s));
        DMRG_HIP_CHECK(hipMemsetAsync(V + static_cast<int64_t>(nn) * nn, 0, static_cast<size_t>(nn) * (npad - nn) * es, s));
    }
    unsigned    *flag = info.as<unsigned>();
    const int    np   = nn + (nn & 1);
    const double tol  = std::sqrt(static_cast<double>(mm)) * 2.3e-16; // gesvj's threshold, sqrt(m) eps
    unsigned     hflag = 0;
    int          sweep = 0;
    const size_t lds   = lds_need;
    // A round costs about the same few microseconds inside one workgroup as one launch per round over the
    // whole chip (measured on gfx950), so the resident path pays off only while one workgroup takes a whole
    // round with one wavefront per pair and the launches it saves dominate.
    if(resident) {
        int *dsweeps = reinterpret_cast<int *>(flag + 1);
        // n <= 32: at most 16 pairs a round; 16 lanes a pair for columns up to 128 (4 wavefronts share each
        // round's barrier), one wavefront a pair beyond
        if(mm <= 128) {
            DMRG_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&jacobi_lds<T, 16>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>(JACOBI_LDS_BYTES)));
            hipLaunchKernelGGL((jacobi_lds<T, 16>), dim3(1), dim3(256), lds, s, mm, nn, np, W, V, tol, 60, dsweeps);
        } else {
            DMRG_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&jacobi_lds<T, 64>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>(JACOBI_LDS_BYTES)));
            hipLaunchKernelGGL((jacobi_lds<T, 64>), dim3(1), dim3(1024), lds, s, mm, nn, np, W, V, tol, 60, dsweeps);
        }
        DMRG_HIP_CHECK(hipGetLastError());
        int hs = 0;
        DMRG_HIP_CHECK(hipMemcpyAsync(&hs, dsweeps, sizeof(int), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        DMRG_REQUIRE(hs < 60, DMRG_ERR_NOT_CONVERGED, "Jacobi SVD: no convergence in 60 sweeps");
        sweep = 60; // skip the per-round path below
    }
    // Block path: column blocks of JACOBI_BLOCK_B2 / 2, three launches per round of block pairs.
    if(blocked) {
        constexpr int    B2     = JACOBI_BLOCK_B2;
        static_assert(B2 == 32, "the matrix-core Gram and update kernels take pairs of 16-column blocks");
        const dmrg_dtype dt     = jtraits<T>::cx ? DMRG_CX64 : DMRG_FP64;
        const int        nb     = npad / (B2 / 2);
        const int     nbp    = nb + (nb & 1);
        const int     npairs = nbp / 2;
        const int     ks     = static_cast<int>(std::clamp<int64_t>((512 + npairs - 1) / npairs, 1, std::max(1, mm / BJ_ROWS)));
        const int     rows_per = ((mm + ks - 1) / ks + BJ_ROWS - 1) / BJ_ROWS * BJ_ROWS;
        const int     ksn      = (mm + rows_per - 1) / rows_per; // chunks actually holding rows
        bj_part.reserve(static_cast<size_t>(npairs) * ksn * B2 * B2 * es);
        bj_q.reserve(static_cast<size_t>(npairs) * B2 * B2 * es);
        bj_rot.reserve(static_cast<size_t>(npairs) * sizeof(int));
        T   *part = bj_part.as<T>(), *qb = bj_q.as<T>();
        int *rot  = bj_rot.as<int>();
        for(; sweep < 60; ++sweep) {
            DMRG_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(unsigned), s));
            for(int r = 0; r < nbp - 1; ++r) {
                bj_gram_mfma(dt, mm, npad, nbp, r, W, ksn, rows_per, part, s);
                hipLaunchKernelGGL((bj_eig<T, B2>), dim3(npairs), dim3(BJ_THREADS), 0, s, npad, nbp, r, ksn, part, qb, rot, tol, JACOBI_BLOCK_INNER, flag);
                bj_update_mfma(dt, mm, nn, npad, nbp, r, W, V, qb, rot, s);
            }
            DMRG_HIP_CHECK(hipGetLastError());
            DMRG_HIP_CHECK(hipMemcpyAsync(&hflag, flag, sizeof(unsigned), hipMemcpyDeviceToHost, s));
            DMRG_HIP_CHECK(hipStreamSynchronize(s));
            if(hflag == 0) break;
        }
        sweep = 60; // skip the per-pair rounds below
    }
    for(; sweep < 60 && nn > 1; ++sweep) {
        DMRG_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(unsigned), s));
        for(int r = 0; r < np - 1; ++r) {
            if(mm > 256) hipLaunchKernelGGL((jacobi_round<T, 256>), dim3(np / 2), dim3(256), 0, s, mm, nn, np, r, W, V, tol, flag);
            else if(mm > 128) hipLaunchKernelGGL((jacobi_round<T, 128>), dim3(np / 2), dim3(128), 0, s, mm, nn, np, r, W, V, tol, flag);
            else hipLaunchKernelGGL((jacobi_round<T, 64>), dim3(np / 2), dim3(64), 0, s, mm, nn, np, r, W, V, tol, flag);
        }
        DMRG_HIP_CHECK(hipGetLastError());
        DMRG_HIP_CHECK(hipMemcpyAsync(&hflag, flag, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        if(hflag == 0) break;
    }
    DMRG_REQUIRE(nn <= 1 || hflag == 0 || resident, DMRG_ERR_NOT_CONVERGED, "Jacobi SVD: no convergence in 60 sweeps");
    hipLaunchKernelGGL(column_norms_kernel, dim3(nn), dim3(64), 0, s, mm, nn, reinterpret_cast<const double *>(W),
                       jtraits<T>::cx ? 1 : 0, res.as<double>());
    std::vector<double> norms(nn);
    DMRG_HIP_CHECK(hipMemcpyAsync(norms.data(), res.p, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int> perm(nn);
    for(int i = 0; i < nn; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return norms[x] > norms[y]; });
    S.resize(nn);
    for(int i = 0; i < nn; ++i) S[i] = norms[perm[i]];
    int *dperm = reinterpret_cast<int *>(flag + 2);
    DMRG_HIP_CHECK(hipMemcpyAsync(dperm, perm.data(), nn * sizeof(int), hipMemcpyHostToDevice, s));
    DMRG_HIP_CHECK(hipMemcpyAsync(sv.p, S.data(), nn * sizeof(double), hipMemcpyHostToDevice, s));
    // W = Uw S Vw^H. Without transpose: U = Uw (m x k), Vt = Vw^H (k x n). With transpose (A = W^H):
    // U = Vw (m x k) and Vt = Uw^H (k x n).
    u.reserve(static_cast<size_t>(mm) * nn * es);
    vt.reserve(static_cast<size_t>(nn) * nn * es);
    hipLaunchKernelGGL(assemble_kernel<T>, dim3(grid_of(static_cast<int64_t>(mm) * nn + static_cast<int64_t>(nn) * nn)), dim3(256), 0, s, mm,
                       nn, dperm, sv.as<double>(), W, V, u.as<T>(), vt.as<T>());
    DMRG_HIP_CHECK(hipGetLastError());
    if(tr) {
        // U_A (m x k) = Vw[:, perm] = adjoint of vt (k x m); Vt_A (k x n) = Uw^H = adjoint of u (n x k).
        DevBuf t1, t2;
        t1.reserve(static_cast<size_t>(m) * k_ * es);
        t2.reserve(static_cast<size_t>(k_) * n * es);
        hipLaunchKernelGGL(adjoint_kernel<T>, dim3(grid_of(static_cast<int64_t>(k_) * m)), dim3(256), 0, s, k_, m, vt.as<T>(), t1.as<T>());
        hipLaunchKernelGGL(adjoint_kernel<T>, dim3(grid_of(static_cast<int64_t>(n) * k_)), dim3(256), 0, s, n, k_, u.as<T>(), t2.as<T>());
        DMRG_HIP_CHECK(hipGetLastError());
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        std::swap(u, t1);
        std::swap(vt, t2);
    }
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
}
template void DeviceSvd::jacobi<double>(int, int, const double *, std::vector<double> &);
template void DeviceSvd::jacobi<cpx<double>>(int, int, const cpx<double> *, std::vector<double> &);

} // namespace dmrg

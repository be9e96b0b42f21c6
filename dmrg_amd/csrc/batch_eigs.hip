// Batched lockstep Davidson (see batch_eigs.hpp).
//
// Per iteration, for every active problem p with basis V_p, W_p = H_p V_p of m columns (the same m for all):
//   host:   Rayleigh-Ritz of the m x m projected matrix -> Ritz coefficients y_p and value theta_p
//   device: x = V y, r = W y - theta x and their norms            (ritz, one grouped launch + fold)
//           t = r - V (V^H r), twice (classical Gram-Schmidt x2)  (dot, gemv: grouped launches + folds)
//           H t for all problems                                  (grouped matvec, replayed graph)
//           V[:, m] = t / |t|, W[:, m] = H t / |t|               (finish)
//           new column V^H W[:, m] of the projected matrix         (dot + fold)
//   one device-to-host copy of every problem's scalars, one synchronisation.
// A problem whose Ritz residual is below tol |A| leaves the batch with x as its eigenvector.
#include "batch_eigs.hpp"

#include "dense_eig.hpp"
#include "group.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>

namespace dmrg {

namespace {

using cx = std::complex<double>;

constexpr int BT     = 256; // threads per workgroup
constexpr int BMAX_M = 64;  // largest basis
constexpr int PST    = 2 * (BMAX_M + 10); // partial sums per workgroup (doubles; room for a whole chunk of 8 columns)

// Per-problem exchange buffer (doubles).
constexpr int IO_Y     = 0;                      // Ritz coefficients (complex, BMAX_M)
constexpr int IO_THETA = 2 * BMAX_M;             // Ritz value
constexpr int IO_C     = IO_THETA + 2;           // projection coefficients (complex, BMAX_M + 1)
constexpr int IO_HCOL  = IO_C + 2 * (BMAX_M + 1); // new column of the projected matrix (complex, BMAX_M + 1)
constexpr int IO_NRM   = IO_HCOL + 2 * (BMAX_M + 1); // |r|^2, |x|^2, |src|^2, |t|^2 pass 1, |t|^2 pass 2
constexpr int IO_SIZE  = IO_NRM + 8;

template <typename T>
struct BProb {
    int64_t n;
    int64_t blk0;
    int     nblk;
    int     active;
    T      *V, *W, *x, *r, *X, *Y;
    double *io;
};

template <typename T>
struct bt;
template <>
struct bt<double> {
    static constexpr bool cx = false;
};
template <>
struct bt<cpx<double>> {
    static constexpr bool cx = true;
};

template <typename T>
__device__ inline void ldc(const T &v, double &re, double &im) {
    if constexpr(bt<T>::cx) re = v.re, im = v.im;
    else re = v, im = 0.0;
}
template <typename T>
__device__ inline T mkc(double re, double im) {
    if constexpr(bt<T>::cx) return T{re, im};
    else return re;
}

__device__ inline double wsum64(double v) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum NV values over the workgroup and store them in partial[0..NV) (thread 0 writes).
template <int NV>
__device__ inline void block_store(double (&v)[NV], double *partial) {
    __shared__ double red[NV][BT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for(int i = 0; i < NV; ++i) v[i] = wsum64(v[i]);
    __syncthreads();
    if(lane == 0)
#pragma unroll
        for(int i = 0; i < NV; ++i) red[i][wave] = v[i];
    __syncthreads();
    if(threadIdx.x < NV) {
        double s = 0;
#pragma unroll
        for(int w = 0; w < BT / 64; ++w) s += red[threadIdx.x][w];
        partial[threadIdx.x] = s;
    }
}

// x = V y, r = W y - theta x; partials |r|^2, |x|^2 (columns 0, 1).
template <typename T>
__global__ __launch_bounds__(BT) void b_ritz(const BProb<T> *pr, const int *blkmap, int m, double *partial) {
    const BProb<T> &P = pr[blkmap[blockIdx.x]];
    if(!P.active) return;
    __shared__ double ys[2 * BMAX_M + 2];
    for(int j = threadIdx.x; j < 2 * m + 2; j += BT) ys[j] = j < 2 * m ? P.io[IO_Y + j] : P.io[IO_THETA + (j - 2 * m)];
    __syncthreads();
    const double th = ys[2 * m];
    const int64_t lb = blockIdx.x - P.blk0;
    double        acc[2] = {0.0, 0.0};
    for(int64_t i = lb * BT + threadIdx.x; i < P.n; i += static_cast<int64_t>(P.nblk) * BT) {
        double xr = 0, xi = 0, wr = 0, wi = 0;
        for(int j = 0; j < m; ++j) {
            double vr, vi, ur, ui;
            ldc(P.V[i + j * P.n], vr, vi);
            ldc(P.W[i + j * P.n], ur, ui);
            const double yr = ys[2 * j], yi = ys[2 * j + 1];
            xr += vr * yr - vi * yi;
            xi += vr * yi + vi * yr;
            wr += ur * yr - ui * yi;
            wi += ur * yi + ui * yr;
        }
        const double rr = wr - th * xr, ri = wi - th * xi;
        P.x[i] = mkc<T>(xr, xi);
        P.r[i] = mkc<T>(rr, ri);
        acc[0] += rr * rr + ri * ri;
        acc[1] += xr * xr + xi * xi;
    }
    block_store<2>(acc, partial + static_cast<int64_t>(blockIdx.x) * PST);
}

// Source vector of a pass: the residual r (src 0), the staging vector X (src 1) or W[:, m] (src 2).
template <typename T>
__device__ inline const T *src_of(const BProb<T> &P, int src, int m) {
    return src == 0 ? P.r : src == 1 ? P.X : P.W + static_cast<int64_t>(m) * P.n;
}

// c_j = V_j^H s for j < mc (partials, complex: slots 2j, 2j+1), and |s|^2 (slots 2 mc, 2 mc + 1) when self.
template <typename T>
__global__ __launch_bounds__(BT) void b_dot(const BProb<T> *pr, const int *blkmap, int mc, int src, int msrc, int self,
                                            double *partial) {
    const BProb<T> &P = pr[blkmap[blockIdx.x]];
    if(!P.active) return;
    const T      *s   = src_of(P, src, msrc);
    const int64_t lb  = blockIdx.x - P.blk0;
    double       *out = partial + static_cast<int64_t>(blockIdx.x) * PST;
    constexpr int CH  = 8;
    const int     nc  = mc + (self ? 1 : 0);
    for(int j0 = 0; j0 < nc; j0 += CH) {
        double acc[2 * CH];
#pragma unroll
        for(int c = 0; c < 2 * CH; ++c) acc[c] = 0.0;
        for(int64_t i = lb * BT + threadIdx.x; i < P.n; i += static_cast<int64_t>(P.nblk) * BT) {
            double xr, xi;
            ldc(s[i], xr, xi);
#pragma unroll
            for(int c = 0; c < CH; ++c) {
                const int j = j0 + c;
                if(j < nc) {
                    double vr, vi;
                    if(j < mc) ldc(P.V[i + j * P.n], vr, vi);
                    else vr = xr, vi = xi;
                    acc[2 * c] += vr * xr + vi * xi; // conj(v) s
                    acc[2 * c + 1] += vr * xi - vi * xr;
                }
            }
        }
        block_store<2 * CH>(acc, out + 2 * j0); // 2 j0 + 2 CH <= PST
    }
}

// X = s - V c (c from io[IO_C]); partial |X|^2 (slot 0).
template <typename T>
__global__ __launch_bounds__(BT) void b_gemv(const BProb<T> *pr, const int *blkmap, int m, int src, double *partial) {
    const BProb<T> &P = pr[blkmap[blockIdx.x]];
    if(!P.active) return;
    __shared__ double cs[2 * BMAX_M];
    for(int j = threadIdx.x; j < 2 * m; j += BT) cs[j] = P.io[IO_C + j];
    __syncthreads();
    const T      *s  = src_of(P, src, m);
    const int64_t lb = blockIdx.x - P.blk0;
    double        acc[1] = {0.0};
    for(int64_t i = lb * BT + threadIdx.x; i < P.n; i += static_cast<int64_t>(P.nblk) * BT) {
        double ar = 0, ai = 0;
        for(int j = 0; j < m; ++j) {
            double vr, vi;
            ldc(P.V[i + j * P.n], vr, vi);
            ar += vr * cs[2 * j] - vi * cs[2 * j + 1];
            ai += vr * cs[2 * j + 1] + vi * cs[2 * j];
        }
        double sr, si;
        ldc(s[i], sr, si);
        const double tr = sr - ar, ti = si - ai;
        P.X[i]          = mkc<T>(tr, ti);
        acc[0] += tr * tr + ti * ti;
    }
    block_store<1>(acc, partial + static_cast<int64_t>(blockIdx.x) * PST);
}

// V[:, m] = X / |t|, W[:, m] = Y / |t| with |t|^2 = io[IO_NRM + 4].
template <typename T>
__global__ __launch_bounds__(BT) void b_finish(const BProb<T> *pr, const int *blkmap, int m) {
    const BProb<T> &P = pr[blkmap[blockIdx.x]];
    if(!P.active) return;
    const double  nr = P.io[IO_NRM + 4];
    const double  a  = nr > 0 ? 1.0 / sqrt(nr) : 0.0;
    const int64_t lb = blockIdx.x - P.blk0;
    T            *v = P.V + static_cast<int64_t>(m) * P.n, *w = P.W + static_cast<int64_t>(m) * P.n;
    for(int64_t i = lb * BT + threadIdx.x; i < P.n; i += static_cast<int64_t>(P.nblk) * BT) {
        double xr, xi, yr, yi;
        ldc(P.X[i], xr, xi);
        ldc(P.Y[i], yr, yi);
        v[i] = mkc<T>(a * xr, a * xi);
        w[i] = mkc<T>(a * yr, a * yi);
    }
}

// io[dst + c] = sum over the problem's workgroups of partial slot c, c < nval, in workgroup order.
template <typename T>
__global__ __launch_bounds__(BT) void b_fold(const BProb<T> *pr, const double *partial, int nval, int dst) {
    const BProb<T> &P = pr[blockIdx.y];
    const int       c = blockIdx.x;
    if(!P.active || c >= nval) return;
    double acc[1] = {0.0};
    for(int b = threadIdx.x; b < P.nblk; b += BT) acc[0] += partial[(P.blk0 + b) * PST + c];
    __shared__ double res[1];
    block_store<1>(acc, res);
    __syncthreads();
    if(threadIdx.x == 0) P.io[dst + c] = res[0];
}

// Seeded start vector into X (entries uniform in [-1, 1]).
__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
template <typename T>
__global__ __launch_bounds__(BT) void b_random(const BProb<T> *pr, const int *blkmap, uint64_t seed) {
    const BProb<T> &P  = pr[blkmap[blockIdx.x]];
    const int64_t   lb = blockIdx.x - P.blk0;
    for(int64_t i = lb * BT + threadIdx.x; i < P.n; i += static_cast<int64_t>(P.nblk) * BT) {
        const uint64_t h = mix64(seed * 0x100000001B3ull + static_cast<uint64_t>(i));
        const double   a = static_cast<double>(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        const double   b = static_cast<double>(mix64(h) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        P.X[i]           = mkc<T>(a, bt<T>::cx ? b : 0.0);
    }
}

template <typename T>
struct Batch {
    hipStream_t                  s;
    std::vector<MatVecOp *>      ops;
    std::vector<BProb<T>>        hp;   // host copy of the descriptors
    DevBuf                       dprobs, dblk, dpart, dio, mem;
    std::vector<int>             blkmap;
    int                          nblk_total = 0;
    double                      *hio        = nullptr; // pinned, k * IO_SIZE
    std::unique_ptr<MatvecGroup> group;

    Batch(hipStream_t st) : s(st) {}
    ~Batch() {
        if(hio) (void) hipHostFree(hio);
    }
    int k() const { return static_cast<int>(ops.size()); }

    void upload_probs() {
        DMRG_HIP_CHECK(hipMemcpyAsync(dprobs.p, hp.data(), sizeof(BProb<T>) * hp.size(), hipMemcpyHostToDevice, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s)); // hp may change right after
    }

    // Grouped matvec X -> Y of the active problems (rebuilt when the active set changes).
    void rebuild_group() {
        std::vector<MatVecOp *> a;
        std::vector<const void *> xs;
        std::vector<void *>       ys;
        for(int p = 0; p < k(); ++p)
            if(hp[p].active) a.push_back(ops[p]), xs.push_back(hp[p].X), ys.push_back(hp[p].Y);
        group.reset();
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        if(!a.empty()) group.reset(matvec_group_create(a.data(), static_cast<int>(a.size()), xs.data(), ys.data(), false, true));
    }

    void fold(int nval, int dst) {
        hipLaunchKernelGGL(b_fold<T>, dim3(nval, k()), dim3(BT), 0, s, dprobs.as<BProb<T>>(), dpart.as<double>(), nval, dst);
    }
    // t = src - V (V^H src), twice when m > 0, into X; |src|^2 in IO_C + 2m, the final |t|^2 in IO_NRM + 4.
    void orthogonalize(int m, int src) {
        const dim3 g(nblk_total), b(BT);
        auto      *P = dprobs.as<BProb<T>>();
        auto      *M = dblk.as<int>();
        double    *Q = dpart.as<double>();
        hipLaunchKernelGGL(b_dot<T>, g, b, 0, s, P, M, m, src, m, 1, Q);
        fold(2 * (m + 1), IO_C);
        hipLaunchKernelGGL(b_gemv<T>, g, b, 0, s, P, M, m, src, Q);
        fold(1, m > 0 ? IO_NRM + 3 : IO_NRM + 4);
        if(m > 0) {
            hipLaunchKernelGGL(b_dot<T>, g, b, 0, s, P, M, m, 1, m, 0, Q);
            fold(2 * m, IO_C);
            hipLaunchKernelGGL(b_gemv<T>, g, b, 0, s, P, M, m, 1, Q);
            fold(1, IO_NRM + 4);
        }
        DMRG_HIP_CHECK(hipGetLastError());
    }
};

template <typename T>
void batched_impl(MatVecOp *const *ops, int k, const EigsParams &p, const void *const *x0s, void *const *evecs,
                  std::vector<BatchEigsResult> &res, hipStream_t s) {
    DMRG_REQUIRE(k >= 1, DMRG_ERR_ARG, "batched eigs: no problems");
    DMRG_REQUIRE(std::max(1, p.nev) == 1 && p.precond == 0 && (p.target == TARGET_SA || p.target == TARGET_LA), DMRG_ERR_UNSUPPORTED,
                 "batched eigs: one pair per problem at SA or LA, no preconditioner");
    Batch<T> B(s);
    B.ops.assign(ops, ops + k);
    int64_t nmin = INT64_MAX, nmax = 0;
    for(int q = 0; q < k; ++q) nmin = std::min(nmin, ops[q]->D.n()), nmax = std::max(nmax, ops[q]->D.n());
    // Basis rule of the single-problem solver (solver.cpp:12-15), common to the batch.
    int mmax = p.ncv > 0 ? p.ncv : static_cast<int>(std::ceil(std::log2(static_cast<double>(std::max<int64_t>(nmax, 2)))));
    mmax     = static_cast<int>(std::clamp<int64_t>(mmax, 4, std::min<int64_t>(nmin, BMAX_M)));
    const int mmin = std::max(1, std::min(mmax / 2 - 1, mmax - 2));
    DMRG_REQUIRE(nmin >= 2, DMRG_ERR_ARG, "batched eigs: problems of size < 2");
    mmax = static_cast<int>(std::min<int64_t>(mmax, nmin));
    // Device storage: per problem V, W (n x mmax), x, r, X, Y (n each).
    size_t total = 0;
    for(int q = 0; q < k; ++q) total += static_cast<size_t>(ops[q]->D.n()) * (2 * mmax + 4);
    B.mem.reserve(total * sizeof(T));
    B.hp.resize(k);
    T      *base = B.mem.template as<T>();
    int64_t blk  = 0;
    for(int q = 0; q < k; ++q) {
        const int64_t n = ops[q]->D.n();
        BProb<T>     &P = B.hp[q];
        P.n             = n;
        P.V = base, base += n * mmax;
        P.W = base, base += n * mmax;
        P.x = base, base += n;
        P.r = base, base += n;
        P.X = base, base += n;
        P.Y = base, base += n;
        P.nblk   = static_cast<int>(std::clamp<int64_t>((n + BT * 4 - 1) / (BT * 4), 1, 256));
        P.blk0   = blk;
        P.active = 1;
        blk += P.nblk;
        for(int b = 0; b < P.nblk; ++b) B.blkmap.push_back(q);
    }
    B.nblk_total = static_cast<int>(blk);
    B.dprobs.reserve(sizeof(BProb<T>) * k);
    B.dblk.reserve(sizeof(int) * B.blkmap.size());
    B.dpart.reserve(sizeof(double) * PST * B.blkmap.size());
    B.dio.reserve(sizeof(double) * IO_SIZE * k);
    DMRG_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&B.hio), sizeof(double) * IO_SIZE * k, hipHostMallocDefault));
    std::fill(B.hio, B.hio + static_cast<size_t>(IO_SIZE) * k, 0.0);
    for(int q = 0; q < k; ++q) B.hp[q].io = B.dio.template as<double>() + static_cast<int64_t>(q) * IO_SIZE;
    DMRG_HIP_CHECK(hipMemcpyAsync(B.dblk.p, B.blkmap.data(), sizeof(int) * B.blkmap.size(), hipMemcpyHostToDevice, s));
    B.upload_probs();

    // Start vectors into X: the caller's, else seeded random ones.
    hipLaunchKernelGGL(b_random<T>, dim3(B.nblk_total), dim3(BT), 0, s, B.dprobs.template as<BProb<T>>(), B.dblk.template as<int>(), p.seed);
    for(int q = 0; q < k; ++q)
        if(x0s && x0s[q]) DMRG_HIP_CHECK(hipMemcpyAsync(B.hp[q].X, x0s[q], static_cast<size_t>(B.hp[q].n) * sizeof(T), hipMemcpyDeviceToDevice, s));
    B.rebuild_group();

    res.assign(k, BatchEigsResult{});
    std::vector<std::vector<cx>> H(k, std::vector<cx>(static_cast<size_t>(mmax) * mmax, cx(0.0, 0.0)));
    std::vector<double>          anorm(k, 0.0);
    std::vector<std::vector<cx>> Z(k);
    std::vector<std::vector<double>> w(k);
    std::vector<int>             opt(k, 0);
    int                          active = k, group_size = k;
    const dim3                   G(B.nblk_total), TB(BT);
    auto                        *dP = B.dprobs.template as<BProb<T>>();
    auto                        *dM = B.dblk.template as<int>();
    double                      *dQ = B.dpart.template as<double>();

    // Expansion from X (m = 0: the start vector) or from the residual; then matvec and the new column.
    auto expand = [&](int m, int src) {
        B.orthogonalize(m, src);
        B.group->run(s);
        hipLaunchKernelGGL(b_finish<T>, G, TB, 0, s, dP, dM, m);
        hipLaunchKernelGGL(b_dot<T>, G, TB, 0, s, dP, dM, m + 1, 2, m, 0, dQ);
        B.fold(2 * (m + 1), IO_HCOL);
        DMRG_HIP_CHECK(hipGetLastError());
    };
    auto fetch = [&]() {
        DMRG_HIP_CHECK(hipMemcpyAsync(B.hio, B.dio.p, sizeof(double) * IO_SIZE * k, hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
    };
    auto Hat = [&](int q, int i, int j) -> cx & { return H[q][i + static_cast<size_t>(mmax) * j]; };
    auto take_column = [&](int q, int m) {
        const double *io = B.hio + static_cast<size_t>(q) * IO_SIZE;
        for(int i = 0; i <= m; ++i) Hat(q, i, m) = cx(io[IO_HCOL + 2 * i], io[IO_HCOL + 2 * i + 1]);
        for(int i = 0; i < m; ++i) Hat(q, m, i) = std::conj(Hat(q, i, m));
        Hat(q, m, m) = cx(Hat(q, m, m).real(), 0.0);
        res[q].matvecs += 1;
    };

    expand(0, 1);
    fetch();
    for(int q = 0; q < k; ++q) take_column(q, 0);
    int m = 1;
    while(active > 0) {
        // Rayleigh-Ritz per active problem; Ritz coefficients to the device.
        for(int q = 0; q < k; ++q) {
            if(!B.hp[q].active) continue;
            std::vector<cx> Hm(static_cast<size_t>(m) * m);
            for(int j = 0; j < m; ++j)
                for(int i = 0; i < m; ++i) Hm[i + static_cast<size_t>(m) * j] = Hat(q, i, j);
            Z[q].assign(static_cast<size_t>(m) * m, cx(0.0, 0.0));
            w[q].assign(m, 0.0);
            if(!hermitian_eig(m, Hm.data(), m, w[q].data(), Z[q].data())) throw Error(DMRG_ERR_NOT_CONVERGED, "batched eigs: Rayleigh-Ritz failed");
            for(int i = 0; i < m; ++i) anorm[q] = std::max(anorm[q], std::fabs(w[q][i]));
            opt[q]      = p.target == TARGET_LA ? m - 1 : 0;
            double *io  = B.hio + static_cast<size_t>(q) * IO_SIZE;
            for(int i = 0; i < m; ++i) {
                io[IO_Y + 2 * i]     = Z[q][i + static_cast<size_t>(m) * opt[q]].real();
                io[IO_Y + 2 * i + 1] = Z[q][i + static_cast<size_t>(m) * opt[q]].imag();
            }
            io[IO_THETA] = w[q][opt[q]];
        }
        DMRG_HIP_CHECK(hipMemcpyAsync(B.dio.p, B.hio, sizeof(double) * IO_SIZE * k, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(b_ritz<T>, G, TB, 0, s, dP, dM, m, dQ);
        B.fold(2, IO_NRM);
        const bool restart = m == mmax;
        if(!restart) expand(m, 0);
        fetch();
        // Convergence, expansion, stopping.
        bool changed = false;
        for(int q = 0; q < k; ++q) {
            if(!B.hp[q].active) continue;
            const double *io = B.hio + static_cast<size_t>(q) * IO_SIZE;
            const double  rn = std::sqrt(std::max(0.0, io[IO_NRM])) / std::sqrt(std::max(1e-300, io[IO_NRM + 1]));
            res[q].eigval    = w[q][opt[q]];
            res[q].anorm     = anorm[q];
            res[q].rnorm     = rn;
            res[q].iters += 1;
            const bool conv  = rn <= p.tol * std::max(anorm[q], 1e-300);
            const bool limit = res[q].matvecs >= p.max_matvecs;
            const bool broke = !restart && !(io[IO_NRM + 4] > 1e-26 * std::max(io[IO_C + 2 * m], 1e-300));
            if(conv || limit || broke) {
                res[q].nconv = conv ? 1 : 0;
                DMRG_HIP_CHECK(hipMemcpyAsync(evecs[q], B.hp[q].x, static_cast<size_t>(B.hp[q].n) * sizeof(T), hipMemcpyDeviceToDevice, s));
                B.hp[q].active = 0;
                --active;
                changed = true;
                continue;
            }
            if(!restart) take_column(q, m);
        }
        if(active == 0) break;
        if(restart) {
            // Thick restart to the mmin Ritz vectors nearest the target: V <- V Q, W <- W Q, H <- diag.
            for(int q = 0; q < k; ++q) {
                if(!B.hp[q].active) continue;
                std::vector<int> ord(m);
                std::iota(ord.begin(), ord.end(), 0);
                if(p.target == TARGET_LA) std::reverse(ord.begin(), ord.end());
                std::vector<T> Qh(static_cast<size_t>(m) * mmin);
                for(int j = 0; j < mmin; ++j)
                    for(int i = 0; i < m; ++i) {
                        const cx z = Z[q][i + static_cast<size_t>(m) * ord[j]];
                        if constexpr(bt<T>::cx) Qh[i + static_cast<size_t>(m) * j] = T{z.real(), z.imag()};
                        else Qh[i + static_cast<size_t>(m) * j] = z.real();
                    }
                DevBuf Qd, tmp;
                Qd.reserve(Qh.size() * sizeof(T));
                tmp.reserve(static_cast<size_t>(B.hp[q].n) * mmin * sizeof(T));
                DMRG_HIP_CHECK(hipMemcpyAsync(Qd.p, Qh.data(), Qh.size() * sizeof(T), hipMemcpyHostToDevice, s));
                for(T *Bm : {B.hp[q].V, B.hp[q].W}) {
                    GemmDesc g;
                    g.dtype = ops[q]->dtype;
                    g.M = static_cast<int32_t>(B.hp[q].n), g.N = mmin, g.K = m, g.batch = 1;
                    g.A = Bm, g.am = plain(1), g.ak = plain(B.hp[q].n);
                    g.B = Qd.p, g.bk = plain(1), g.bn = plain(m);
                    g.C = tmp.p, g.cm = plain(1), g.cn = plain(B.hp[q].n);
                    gemm(g, s);
                    DMRG_HIP_CHECK(hipMemcpyAsync(Bm, tmp.p, static_cast<size_t>(B.hp[q].n) * mmin * sizeof(T), hipMemcpyDeviceToDevice, s));
                }
                DMRG_HIP_CHECK(hipStreamSynchronize(s)); // Qh and the temporaries die here
                std::fill(H[q].begin(), H[q].end(), cx(0.0, 0.0));
                for(int j = 0; j < mmin; ++j) Hat(q, j, j) = cx(w[q][ord[j]], 0.0);
            }
            m = mmin;
        } else {
            ++m;
        }
        if(changed) {
            B.upload_probs();
            if(2 * active <= group_size) {
                B.rebuild_group();
                group_size = active;
            }
        }
    }
    DMRG_HIP_CHECK(hipStreamSynchronize(s));
}

} // namespace

void eigs_batched(MatVecOp *const *ops, int k, const EigsParams &p, const void *const *x0s, void *const *evecs,
                  std::vector<BatchEigsResult> &res, hipStream_t s) {
    DMRG_REQUIRE(ops && evecs && k >= 1, DMRG_ERR_ARG, "batched eigs: bad arguments");
    const dmrg_dtype t = ops[0]->dtype;
    for(int q = 0; q < k; ++q) DMRG_REQUIRE(ops[q] && ops[q]->dtype == t && evecs[q], DMRG_ERR_ARG, "batched eigs: one dtype, outputs");
    if(t == DMRG_FP64) batched_impl<double>(ops, k, p, x0s, evecs, res, s);
    else if(t == DMRG_CX64) batched_impl<cpx<double>>(ops, k, p, x0s, evecs, res, s);
    else throw Error(DMRG_ERR_ARG, "batched eigs: FP64 or CX64 operators");
}

} // namespace dmrg

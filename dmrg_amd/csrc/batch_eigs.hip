// Batched Davidson over a changing set of problems (see batch_eigs.hpp).
//
// Each problem p carries its own basis size m_p and phase: START (expand from its start vector), ITER (Ritz pair,
// then expand from the residual) or RESTART (Ritz pair only, then a thick restart on the host side). One step of
// all problems:
//   host:   Rayleigh-Ritz of each m_p x m_p projected matrix -> Ritz coefficients y_p and value theta_p
//   device: x = V y, r = W y - theta x and their norms            (ritz: grouped launch + fold)
//           t = s - V (V^H s), twice, s = r or the start vector     (dot, gemv: grouped launches + folds)
//           H t for all problems                                  (grouped matvec, replayed graph)
//           V[:, m] = t / |t|, W[:, m] = H t / |t|               (finish)
//           new column V^H W[:, m] of the projected matrix         (dot + fold)
//   one device-to-host copy of every problem's scalars, one synchronisation.
// A problem whose Ritz residual is below tol |A| leaves with x as its eigenvector; new problems join between
// steps (continuous batching), and the grouped launches are rebuilt when the set changes. The engine runs in FP64
// or, as the first phase of the mixed-precision batch (MixedEngine), in FP32.
#include "batch_eigs.hpp"

#include "dense_eig.hpp"
#include "group.hpp"
#include "vecops.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <unordered_map>
#include <map>
#include <numeric>
#include <thread>

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

enum Phase : int { PH_NONE = 0, PH_START = 1, PH_ITER = 2, PH_RESTART = 3 };

template <typename T>
struct BProb {
    int64_t n;
    int64_t blk0;
    int     nblk;
    int     phase;
    int     m;
    int     pad;
    T      *V, *W, *x, *r, *X, *Y;
    double *io;
};

// Element traits: complex or real, the real type, and whether the engine runs in single precision (FP32 vectors
// and FP32 matrix-core matvecs; every reduction still accumulates in double).
template <typename T>
struct bt;
template <>
struct bt<double> {
    static constexpr bool cx = false, single = false;
    using R                  = double;
};
template <>
struct bt<cpx<double>> {
    static constexpr bool cx = true, single = false;
    using R                  = double;
};
template <>
struct bt<float> {
    static constexpr bool cx = false, single = true;
    using R                  = float;
};
template <>
struct bt<cpx<float>> {
    static constexpr bool cx = true, single = true;
    using R                  = float;
};

template <typename T>
__device__ inline void ldc(const T &v, double &re, double &im) {
    if constexpr(bt<T>::cx) re = v.re, im = v.im;
    else re = v, im = 0.0;
}
template <typename T>
__device__ inline T mkc(double re, double im) {
    using R = typename bt<T>::R;
    if constexpr(bt<T>::cx) return T{static_cast<R>(re), static_cast<R>(im)};
    else return static_cast<T>(re);
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

// Stages of the grouped kernels and folds; per problem they fix the column count and the destination.
enum Stage : int { ST_RITZ = 0, ST_DOT1 = 1, ST_GEMV1 = 2, ST_DOT2 = 3, ST_GEMV2 = 4, ST_HCOL = 5 };

template <typename T>
__device__ inline bool stage_on(const BProb<T> &P, int st) {
    switch(st) {
        case ST_RITZ: return P.phase == PH_ITER || P.phase == PH_RESTART;
        case ST_DOT2:
        case ST_GEMV2: return P.phase == PH_ITER;
        default: return P.phase == PH_START || P.phase == PH_ITER;
    }
}
// Basis columns used by a stage (the new column's index for ST_HCOL).
template <typename T>
__device__ inline int stage_m(const BProb<T> &P) {
    return P.phase == PH_START ? 0 : P.m;
}
// Source of a dot/gemv stage: the start vector X (START, and pass 2), the residual r (pass 1 of ITER), W[:, m] (HCOL).
template <typename T>
__device__ inline const T *stage_src(const BProb<T> &P, int st) {
    if(st == ST_HCOL) return P.W + static_cast<int64_t>(stage_m(P)) * P.n;
    if(st == ST_DOT2 || st == ST_GEMV2 || P.phase == PH_START) return P.X;
    return P.r;
}
// Values folded by a stage and where they go in io.
template <typename T>
__device__ inline void stage_fold(const BProb<T> &P, int st, int &nval, int &dst) {
    const int m = stage_m(P);
    switch(st) {
        case ST_RITZ: nval = 2, dst = IO_NRM; break;
        case ST_DOT1: nval = 2 * (m + 1), dst = IO_C; break; // c, then |s|^2 at IO_C + 2m
        case ST_GEMV1: nval = 1, dst = P.phase == PH_START ? IO_NRM + 4 : IO_NRM + 3; break;
        case ST_DOT2: nval = 2 * m, dst = IO_C; break;
        case ST_GEMV2: nval = 1, dst = IO_NRM + 4; break;
        default: nval = 2 * (m + 1), dst = IO_HCOL; break;
    }
}

// x = V y, r = W y - theta x; partials |r|^2, |x|^2.
template <typename T>
__global__ __launch_bounds__(BT) void b_ritz(const BProb<T> *pr, const int *blkmap, double *partial) {
    const BProb<T> &P = pr[blkmap[blockIdx.x]];
    if(!stage_on(P, ST_RITZ)) return;
    const int         m = P.m;
    __shared__ double ys[2 * BMAX_M + 2];
    for(int j = threadIdx.x; j < 2 * m + 2; j += BT) ys[j] = j < 2 * m ? P.io[IO_Y + j] : P.io[IO_THETA + (j - 2 * m)];
    __syncthreads();
    const double  th = ys[2 * m];
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

// Dots of a stage: c_j = V_j^H s for the stage's columns, and |s|^2 after them in ST_DOT1.
template <typename T>
__global__ __launch_bounds__(BT) void b_dot(const BProb<T> *pr, const int *blkmap, int st, double *partial) {
    const BProb<T> &P = pr[blkmap[blockIdx.x]];
    if(!stage_on(P, st)) return;
    const T      *s   = stage_src(P, st);
    const int     m   = stage_m(P);
    const int     mc  = st == ST_HCOL ? m + 1 : m;
    const int     nc  = mc + (st == ST_DOT1 ? 1 : 0);
    const int64_t lb  = blockIdx.x - P.blk0;
    double       *out = partial + static_cast<int64_t>(blockIdx.x) * PST;
    constexpr int CH  = 8;
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

// X = s - V c (c from io[IO_C]); partial |X|^2.
template <typename T>
__global__ __launch_bounds__(BT) void b_gemv(const BProb<T> *pr, const int *blkmap, int st, double *partial) {
    const BProb<T> &P = pr[blkmap[blockIdx.x]];
    if(!stage_on(P, st)) return;
    const int         m = stage_m(P);
    __shared__ double cs[2 * BMAX_M];
    for(int j = threadIdx.x; j < 2 * m; j += BT) cs[j] = P.io[IO_C + j];
    __syncthreads();
    const T      *s  = stage_src(P, st);
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
__global__ __launch_bounds__(BT) void b_finish(const BProb<T> *pr, const int *blkmap) {
    const BProb<T> &P = pr[blkmap[blockIdx.x]];
    if(!stage_on(P, ST_GEMV1)) return;
    const int     m  = stage_m(P);
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

// io[dst + c] = sum over the problem's workgroups of partial slot c, in workgroup order (c < the stage's count).
template <typename T>
__global__ __launch_bounds__(BT) void b_fold(const BProb<T> *pr, const double *partial, int st) {
    const BProb<T> &P = pr[blockIdx.y];
    if(!stage_on(P, st)) return;
    int nval, dst;
    stage_fold(P, st, nval, dst);
    const int c = blockIdx.x;
    if(c >= nval) return;
    double acc[1] = {0.0};
    for(int b = threadIdx.x; b < P.nblk; b += BT) acc[0] += partial[(P.blk0 + b) * PST + c];
    __shared__ double res[1];
    block_store<1>(acc, res);
    __syncthreads();
    if(threadIdx.x == 0) P.io[dst + c] = res[0];
}

template <typename T>
struct Slot {
    MatVecOp       *op = nullptr;
    void           *mem = nullptr, *ws = nullptr; // pooled: vectors (layout in Engine::add), matvec workspace
    size_t          mem_cls = 0, ws_cls = 0;
    int64_t         n = 0;
    int             m = 0, phase = PH_START, mmax = 4, mmin = 1, opt = 0;
    std::vector<cx> H, Z;
    std::vector<double> w;
    double          anorm = 0.0;
    double          best_rn = 1e300; // stagnation rule: smallest residual seen and steps since it dropped by 10%
    int             since_best = 0;
    BatchEigsResult res;
    void           *evec = nullptr;
    int             id   = 0;
    int64_t         footprint = 0;
    T *V() const { return static_cast<T *>(mem); }
    T *restart_tmp() const { return V() + n * (2 * mmax + 4); }                 // n x mmin
    T *restart_q() const { return restart_tmp() + n * mmin; }                    // mmax x mmin
    cx &Hat(int i, int j) { return H[i + static_cast<size_t>(mmax) * j]; }
};

struct Finished {
    int             id;
    BatchEigsResult res;
    int64_t         footprint;
};

template <typename T>
struct Engine {
    hipStream_t                           s;
    EigsParams                            p;
    std::vector<std::unique_ptr<Slot<T>>> slots;
    std::vector<BProb<T>>                 hp;
    DevBuf                                dprobs, dblk, dpart, dio;
    double                               *hio     = nullptr;
    size_t                                hio_cap = 0;
    int                                   nblk_total = 0, maxm = 0, nactive = 0;
    std::unique_ptr<MatvecGroup>          group;
    bool                                  dirty = true;
    bool                                  capture; // record each step's launches into a HIP graph at each rebuild
    hipGraph_t                            step_graph = nullptr;
    hipGraphExec_t                        step_exec  = nullptr;
    bool                                  eager_retire = false; // rebuild as soon as a problem finishes (its operator may change)
    DevPool                               pool;
    DevBuf                                gstore[3]; // grouped GEMM descriptors, one store per leg
    int64_t                               steps = 0, rebuilds = 0;
    double                                t_rebuild = 0, t_rr = 0, t_step = 0; // host seconds: rebuilds, Rayleigh-Ritz, whole steps

    Engine(hipStream_t st, const EigsParams &pp, bool capture_graph = true) : s(st), p(pp), capture(capture_graph) {}
    ~Engine() {
        drop_step_graph();
        group.reset();
        if(s) (void) hipStreamSynchronize(s);
        for(auto &q : slots) release(*q);
        if(hio) (void) hipHostFree(hio);
    }
    void release(Slot<T> &S) {
        pool.put(S.mem, S.mem_cls);
        pool.put(S.ws, S.ws_cls);
        S.mem = S.ws = nullptr;
    }
    int k() const { return static_cast<int>(slots.size()); }

    // Vectors: V, W (n x mmax), x, r, X, Y, restart temporaries (n x mmin, mmax x mmin); plus the matvec workspace.
    static int64_t vec_elems(int64_t n, int mmax, int mmin) { return n * (2 * mmax + 4) + n * mmin + int64_t(mmax) * mmin; }
    static int64_t footprint_of(const MatVecOp &op, int mmax, int mmin) {
        return vec_elems(op.D.n(), mmax, mmin) * static_cast<int64_t>(sizeof(T)) +
               static_cast<int64_t>(matvec_group_ws_bytes(op, bt<T>::single));
    }
    static void basis_rule(const EigsParams &p, int64_t n, int &mmax, int &mmin) {
        // Basis rule of the single-problem solver (solver.cpp:12-15, solver_primme.cpp:374).
        mmax = p.ncv > 0 ? p.ncv : static_cast<int>(std::ceil(std::log2(static_cast<double>(std::max<int64_t>(n, 2)))));
        // At least 4, at most min(n, BMAX_M): for n < 4 the upper bound wins, so that add() and the admission
        // footprint of serve_impl use the same basis size.
        mmax = static_cast<int>(std::min<int64_t>(std::max(mmax, 4), std::min<int64_t>(n, BMAX_M)));
        mmin = std::max(1, std::min(mmax / 2 - 1, mmax - 2));
    }

    // anorm0: an estimate of |A| carried over from an earlier phase (0: none).
    void add(MatVecOp *op, const void *x0, void *evec, int id, double anorm0 = 0.0) {
        DMRG_REQUIRE(op->D.n() >= 2, DMRG_ERR_ARG, "batched eigs: problems of size < 2");
        DMRG_REQUIRE(!op->chain, DMRG_ERR_UNSUPPORTED, "batched eigs: operators with more than 5 MPOs (applied in sequence)");
        auto sl  = std::make_unique<Slot<T>>();
        sl->op   = op;
        sl->n    = op->D.n();
        sl->evec = evec;
        sl->id   = id;
        sl->anorm = anorm0;
        basis_rule(p, sl->n, sl->mmax, sl->mmin);
        sl->H.assign(static_cast<size_t>(sl->mmax) * sl->mmax, cx(0.0, 0.0));
        sl->footprint = footprint_of(*op, sl->mmax, sl->mmin);
        sl->mem       = pool.get(static_cast<size_t>(vec_elems(sl->n, sl->mmax, sl->mmin)) * sizeof(T), sl->mem_cls);
        try {
            sl->ws = pool.get(matvec_group_ws_bytes(*op, bt<T>::single), sl->ws_cls);
        } catch(...) {
            release(*sl);
            throw;
        }
        T *X = sl->V() + sl->n * (2 * sl->mmax + 2);
        if(x0) DMRG_HIP_CHECK(hipMemcpyAsync(X, x0, static_cast<size_t>(sl->n) * sizeof(T), hipMemcpyDeviceToDevice, s));
        else random_fill<T>(sl->n, X, p.seed, s);
        slots.push_back(std::move(sl));
        ++nactive;
        dirty = true;
    }

    // Device descriptors, block map, exchange buffers and the grouped matvec of the current set.
    // Finished problems stay in the set (their matvecs still run inside the grouped launches) until the next
    // rebuild, which happens when problems join or when half of the set has finished.
    void rebuild() {
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        drop_step_graph();
        group.reset();
        ++rebuilds;
        for(auto &q : slots)
            if(q->phase == PH_NONE) release(*q);
        slots.erase(std::remove_if(slots.begin(), slots.end(), [](const auto &q) { return q->phase == PH_NONE; }), slots.end());
        hp.assign(k(), BProb<T>{});
        std::vector<int> blkmap;
        int64_t          blk = 0;
        for(int q = 0; q < k(); ++q) {
            Slot<T>  &S = *slots[q];
            BProb<T> &P = hp[q];
            T        *b = S.V();
            P.n         = S.n;
            P.V = b, P.W = b + S.n * S.mmax, P.x = b + S.n * 2 * S.mmax, P.r = P.x + S.n, P.X = P.r + S.n, P.Y = P.X + S.n;
            P.nblk  = static_cast<int>(std::clamp<int64_t>((S.n + BT * 4 - 1) / (BT * 4), 1, 256));
            P.blk0  = blk;
            P.phase = S.phase, P.m = S.m;
            blk += P.nblk;
            for(int b2 = 0; b2 < P.nblk; ++b2) blkmap.push_back(q);
        }
        nblk_total = static_cast<int>(blk);
        if(k() == 0) return;
        dprobs.reserve(sizeof(BProb<T>) * k());
        dblk.reserve(sizeof(int) * blkmap.size());
        dpart.reserve(sizeof(double) * PST * blkmap.size());
        dio.reserve(sizeof(double) * IO_SIZE * k());
        const size_t need = static_cast<size_t>(IO_SIZE) * k();
        if(need > hio_cap) {
            if(hio) (void) hipHostFree(hio);
            hio = nullptr;
            DMRG_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&hio), sizeof(double) * need, hipHostMallocDefault));
            hio_cap = need;
        }
        for(int q = 0; q < k(); ++q) hp[q].io = dio.template as<double>() + static_cast<int64_t>(q) * IO_SIZE;
        DMRG_HIP_CHECK(hipMemcpyAsync(dblk.p, blkmap.data(), sizeof(int) * blkmap.size(), hipMemcpyHostToDevice, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<MatVecOp *>   ops(k());
        std::vector<const void *> xs(k());
        std::vector<void *>       ys(k()), wss(k());
        for(int q = 0; q < k(); ++q) ops[q] = slots[q]->op, xs[q] = hp[q].X, ys[q] = hp[q].Y, wss[q] = slots[q]->ws;
        GroupStorage gs;
        gs.wss = wss.data();
        for(int j = 0; j < 3; ++j) {
            gs.stores[j]      = gstore[j].reserve(grouped_store_bytes(k()));
            gs.store_bytes[j] = gstore[j].bytes;
        }
        // The grouped matvec goes into the step graph as plain kernel nodes.
        group.reset(matvec_group_create(ops.data(), k(), xs.data(), ys.data(), bt<T>::single, false, &gs));
        if(capture) capture_step();
        dirty = false;
    }

    void upload_phases() {
        maxm = 0;
        for(int q = 0; q < k(); ++q) {
            hp[q].phase = slots[q]->phase;
            hp[q].m     = slots[q]->m;
            maxm        = std::max(maxm, slots[q]->m);
        }
        DMRG_HIP_CHECK(hipMemcpyAsync(dprobs.p, hp.data(), sizeof(BProb<T>) * hp.size(), hipMemcpyHostToDevice, s));
    }

    // fold_rows: blocks per problem of the fold (enough for 2 (m + 2) values; the surplus returns at once).
    void stage(int st, hipStream_t cs, int fold_rows) {
        auto     *P = dprobs.template as<BProb<T>>();
        auto     *M = dblk.template as<int>();
        double   *Q = dpart.template as<double>();
        const dim3 g(nblk_total), b(BT);
        switch(st) {
            case ST_RITZ: hipLaunchKernelGGL(b_ritz<T>, g, b, 0, cs, P, M, Q); break;
            case ST_GEMV1:
            case ST_GEMV2: hipLaunchKernelGGL(b_gemv<T>, g, b, 0, cs, P, M, st, Q); break;
            default: hipLaunchKernelGGL(b_dot<T>, g, b, 0, cs, P, M, st, Q); break;
        }
        hipLaunchKernelGGL(b_fold<T>, dim3(fold_rows, k()), dim3(BT), 0, cs, P, Q, st);
        DMRG_HIP_CHECK(hipGetLastError());
    }

    // The device part of a step: every launch reads the problems' phase, basis size and scalars from device
    // memory, so the sequence is the same for every step between two rebuilds.
    void device_step(hipStream_t cs, int fold_rows) {
        stage(ST_RITZ, cs, fold_rows);
        stage(ST_DOT1, cs, fold_rows);
        stage(ST_GEMV1, cs, fold_rows);
        stage(ST_DOT2, cs, fold_rows);
        stage(ST_GEMV2, cs, fold_rows);
        group->run(cs);
        hipLaunchKernelGGL(b_finish<T>, dim3(nblk_total), dim3(BT), 0, cs, dprobs.template as<BProb<T>>(), dblk.template as<int>());
        stage(ST_HCOL, cs, fold_rows);
    }

    void drop_step_graph() {
        if(step_exec) (void) hipGraphExecDestroy(step_exec);
        if(step_graph) (void) hipGraphDestroy(step_graph);
        step_exec  = nullptr;
        step_graph = nullptr;
    }

    // Records device_step into a HIP graph (kernel nodes only), replayed by every step until the next rebuild.
    void capture_step() {
        hipStream_t cs = nullptr;
        DMRG_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        try {
            DMRG_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            device_step(cs, 2 * (BMAX_M + 2));
            DMRG_HIP_CHECK(hipStreamEndCapture(cs, &step_graph));
            DMRG_HIP_CHECK(hipGraphInstantiate(&step_exec, step_graph, nullptr, nullptr, 0));
        } catch(...) {
            hipGraph_t partial = nullptr;
            (void) hipStreamEndCapture(cs, &partial); // leave capture mode before the stream goes
            if(partial) (void) hipGraphDestroy(partial);
            (void) hipStreamDestroy(cs);
            drop_step_graph();
            throw;
        }
        DMRG_HIP_CHECK(hipStreamDestroy(cs));
    }

    // One step of every active problem; returns the problems that finished in it.
    std::vector<Finished> step() {
        using clk     = std::chrono::steady_clock;
        const auto t0 = clk::now();
        auto       since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
        if(dirty) {
            rebuild();
            t_rebuild += since(t0);
        }
        ++steps;
        struct StepTimer {
            double &acc;
            clk::time_point t;
            ~StepTimer() { acc += std::chrono::duration<double>(clk::now() - t).count(); }
        } step_timer{t_step, t0};
        const auto t_rr0 = clk::now();
        std::vector<Finished> done;
        if(nactive == 0) return done;
        // Rayleigh-Ritz of the problems that have a basis.
        for(int q = 0; q < k(); ++q) {
            Slot<T> &S = *slots[q];
            if(S.phase != PH_ITER && S.phase != PH_RESTART) continue;
            const int       m = S.m;
            std::vector<cx> Hm(static_cast<size_t>(m) * m);
            for(int j = 0; j < m; ++j)
                for(int i = 0; i < m; ++i) Hm[i + static_cast<size_t>(m) * j] = S.Hat(i, j);
            S.Z.assign(static_cast<size_t>(m) * m, cx(0.0, 0.0));
            S.w.assign(m, 0.0);
            if(!hermitian_eig(m, Hm.data(), m, S.w.data(), S.Z.data())) throw Error(DMRG_ERR_NOT_CONVERGED, "batched eigs: Rayleigh-Ritz failed");
            for(int i = 0; i < m; ++i) S.anorm = std::max(S.anorm, std::fabs(S.w[i]));
            S.opt      = p.target == TARGET_LA ? m - 1 : 0;
            double *io = hio + static_cast<size_t>(q) * IO_SIZE;
            for(int i = 0; i < m; ++i) {
                io[IO_Y + 2 * i]     = S.Z[i + static_cast<size_t>(m) * S.opt].real();
                io[IO_Y + 2 * i + 1] = S.Z[i + static_cast<size_t>(m) * S.opt].imag();
            }
            io[IO_THETA] = S.w[S.opt];
        }
        t_rr += since(t_rr0);
        upload_phases();
        DMRG_HIP_CHECK(hipMemcpyAsync(dio.p, hio, sizeof(double) * IO_SIZE * k(), hipMemcpyHostToDevice, s));
        if(step_exec) DMRG_HIP_CHECK(hipGraphLaunch(step_exec, s));
        else device_step(s, 2 * (maxm + 2));
        DMRG_HIP_CHECK(hipMemcpyAsync(hio, dio.p, sizeof(double) * IO_SIZE * k(), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));

        for(int q = 0; q < k(); ++q) {
            Slot<T>      &S  = *slots[q];
            if(S.phase == PH_NONE) continue;
            const double *io = hio + static_cast<size_t>(q) * IO_SIZE;
            auto take_column = [&](int m) {
                for(int i = 0; i <= m; ++i) S.Hat(i, m) = cx(io[IO_HCOL + 2 * i], io[IO_HCOL + 2 * i + 1]);
                for(int i = 0; i < m; ++i) S.Hat(m, i) = std::conj(S.Hat(i, m));
                S.Hat(m, m) = cx(S.Hat(m, m).real(), 0.0);
                S.res.matvecs += 1;
            };
            if(S.phase == PH_START) {
                take_column(0);
                S.m     = 1;
                S.phase = PH_ITER;
                continue;
            }
            const double rn = std::sqrt(std::max(0.0, io[IO_NRM])) / std::sqrt(std::max(1e-300, io[IO_NRM + 1]));
            S.res.eigval    = S.w[S.opt];
            S.res.rnorm     = rn;
            S.res.anorm     = S.anorm;
            S.res.iters += 1;
            const bool conv  = rn <= p.tol * std::max(S.anorm, 1e-300);
            const bool limit = S.res.matvecs >= p.max_matvecs;
            const bool broke = S.phase == PH_ITER && !(io[IO_NRM + 4] > 1e-26 * std::max(io[IO_C + 2 * S.m], 1e-300));
            if(rn < 0.9 * S.best_rn) S.best_rn = rn, S.since_best = 0;
            else ++S.since_best;
            const bool stagnated = p.stagnation_window > 0 && S.since_best >= p.stagnation_window;
            if(conv || limit || broke || stagnated) {
                S.res.nconv = conv ? 1 : 0;
                DMRG_HIP_CHECK(hipMemcpyAsync(S.evec, hp[q].x, static_cast<size_t>(S.n) * sizeof(T), hipMemcpyDeviceToDevice, s));
                S.phase = PH_NONE;
                --nactive;
                done.push_back(Finished{S.id, S.res, S.footprint});
                continue;
            }
            if(S.phase == PH_ITER) {
                take_column(S.m);
                if(++S.m == S.mmax) S.phase = PH_RESTART;
                continue;
            }
            // Thick restart to the mmin Ritz vectors nearest the target: V <- V Q, W <- W Q, H <- diag.
            const int        m = S.m;
            std::vector<int> ord(m);
            std::iota(ord.begin(), ord.end(), 0);
            if(p.target == TARGET_LA) std::reverse(ord.begin(), ord.end());
            std::vector<T> Qh(static_cast<size_t>(m) * S.mmin);
            for(int j = 0; j < S.mmin; ++j)
                for(int i = 0; i < m; ++i) {
                    const cx z = S.Z[i + static_cast<size_t>(m) * ord[j]];
                    using R = typename bt<T>::R;
                    if constexpr(bt<T>::cx) Qh[i + static_cast<size_t>(m) * j] = T{static_cast<R>(z.real()), static_cast<R>(z.imag())};
                    else Qh[i + static_cast<size_t>(m) * j] = static_cast<R>(z.real());
                }
            T *Qd = S.restart_q(), *tmp = S.restart_tmp();
            DMRG_HIP_CHECK(hipMemcpyAsync(Qd, Qh.data(), Qh.size() * sizeof(T), hipMemcpyHostToDevice, s));
            for(T *Bm : {hp[q].V, hp[q].W}) {
                GemmDesc g;
                g.dtype = bt<T>::single ? dtype_single_of(S.op->dtype) : S.op->dtype;
                g.M = static_cast<int32_t>(S.n), g.N = S.mmin, g.K = m, g.batch = 1;
                g.A = Bm, g.am = plain(1), g.ak = plain(S.n);
                g.B = Qd, g.bk = plain(1), g.bn = plain(m);
                g.C = tmp, g.cm = plain(1), g.cn = plain(S.n);
                gemm(g, s);
                DMRG_HIP_CHECK(hipMemcpyAsync(Bm, tmp, static_cast<size_t>(S.n) * S.mmin * sizeof(T), hipMemcpyDeviceToDevice, s));
            }
            DMRG_HIP_CHECK(hipStreamSynchronize(s)); // Qh dies here
            std::fill(S.H.begin(), S.H.end(), cx(0.0, 0.0));
            for(int j = 0; j < S.mmin; ++j) S.Hat(j, j) = cx(S.w[ord[j]], 0.0);
            S.m     = S.mmin;
            S.phase = PH_ITER;
        }
        if(!done.empty()) {
            DMRG_HIP_CHECK(hipStreamSynchronize(s)); // eigenvectors in place before the problems are reported
            if(eager_retire || 2 * nactive <= k()) dirty = true;
        }
        return done;
    }
};

// Mixed precision over the batch: each problem first runs in an FP32 engine (FP32 vectors, FP32 matrix-core grouped
// matvecs, double reductions) to 5e-6 |A| or until its residual stagnates (no 10% drop over 25 steps), the
// residual floor of FP32 arithmetic being ~1e-6 |A|; its FP32 Ritz vector then starts it in an FP64 engine that
// converges it to tol (the two phases of the single-problem mixed solver, eigs_mixed, without the correction
// solves). The two engines step one after the other on the same stream, so a problem moves to the FP64 phase as
// soon as its FP32 phase ends while the others go on.
template <typename T64, typename T32>
struct MixedEngine {
    using Lo = Engine<T32>;
    using Hi = Engine<T64>;
    struct Rec {
        MatVecOp       *op   = nullptr;
        void           *evec = nullptr;
        void           *x32 = nullptr, *y32 = nullptr;
        size_t          cx32 = 0, cy32 = 0;
        BatchEigsResult res1;
        int64_t         footprint = 0;
    };
    hipStream_t                  s;
    EigsParams                   p;
    Lo                           lo;
    Hi                           hi;
    DevPool                      pool; // FP32 start vectors and FP32 phase results
    std::unordered_map<int, Rec> rec;
    int                          nactive = 0;
    int64_t                      steps   = 0;

    static EigsParams phase1(const EigsParams &p) {
        EigsParams p1        = p;
        p1.tol               = std::max(p.tol, 5e-6);
        p1.stagnation_window = 25;
        return p1;
    }
    static EigsParams phase2(const EigsParams &p) {
        EigsParams p2        = p;
        p2.seed              = p.seed + 1;
        p2.stagnation_window = 0;
        return p2;
    }
    MixedEngine(hipStream_t st, const EigsParams &pp, bool capture_graph = true)
        : s(st), p(pp), lo(st, phase1(pp), capture_graph), hi(st, phase2(pp), capture_graph) {}

    static void basis_rule(const EigsParams &p, int64_t n, int &mmax, int &mmin) { Hi::basis_rule(p, n, mmax, mmin); }
    // Both phases' vectors (the engines keep their pooled memory) and the FP32 staging vectors.
    static int64_t footprint_of(const MatVecOp &op, int mmax, int mmin) {
        return Hi::footprint_of(op, mmax, mmin) + Lo::footprint_of(op, mmax, mmin) + 2 * op.D.n() * static_cast<int64_t>(sizeof(T32));
    }

    void add(MatVecOp *op, const void *x0, void *evec, int id) {
        Rec r;
        r.op   = op;
        r.evec = evec;
        const int64_t n = op->D.n();
        int           mmax, mmin;
        basis_rule(p, n, mmax, mmin);
        r.footprint = footprint_of(*op, mmax, mmin);
        r.y32       = pool.get(static_cast<size_t>(n) * sizeof(T32), r.cy32);
        if(x0) {
            r.x32 = pool.get(static_cast<size_t>(n) * sizeof(T32), r.cx32);
            convert<T32, T64>(n, static_cast<const T64 *>(x0), static_cast<T32 *>(r.x32), s);
        }
        lo.add(op, r.x32, r.y32, id);
        rec[id] = r;
        ++nactive;
    }

    std::vector<Finished> step() {
        std::vector<Finished> done;
        if(lo.nactive > 0)
            for(const auto &d : lo.step()) { // FP32 phase over: its Ritz vector starts the FP64 phase
                Rec          &r = rec.at(d.id);
                const int64_t n = r.op->D.n();
                convert<T64, T32>(n, static_cast<const T32 *>(r.y32), static_cast<T64 *>(r.evec), s);
                r.res1 = d.res;
                hi.add(r.op, r.evec, r.evec, d.id, d.res.anorm);
                pool.put(r.x32, r.cx32); // later users of these blocks are ordered after the conversion on s
                pool.put(r.y32, r.cy32);
                r.x32 = r.y32 = nullptr;
            }
        if(hi.nactive > 0)
            for(const auto &d : hi.step()) {
                auto it = rec.find(d.id);
                Finished f{d.id, d.res, it->second.footprint};
                f.res.matvecs_fp32 = it->second.res1.matvecs;
                f.res.iters += it->second.res1.iters;
                rec.erase(it);
                --nactive;
                done.push_back(f);
            }
        steps = lo.steps + hi.steps;
        return done;
    }
};

// Problems at least this large (n = d chi^2 >= 2^17: chi >= 256 at d = 2) fill the GPU with their own matvecs:
// dmrg_eigs_batched solves them one after another with the single-problem solver (in mixed precision with its
// FP32 correction-equation expansions) rather than in the lockstep engine, whose grouped launches pay off only
// for problems that leave most of the GPU idle.
constexpr int64_t SOLO_MIN_N = 131072;

// One problem alone: Davidson in FP64, or the mixed solver when T32 is not void.
template <typename T, typename T32>
BatchEigsResult solve_alone(MatVecOp &op, const EigsParams &p, const void *x0, void *evec, hipStream_t s) {
    BatchEigsResult r;
    double          ev[DMRG_EIGS_MAX_NEV] = {}, rn[DMRG_EIGS_MAX_NEV] = {};
    EigsStats       st;
    if constexpr(std::is_void_v<T32>) {
        Davidson<T> d(op, s);
        st = d.solve(p, static_cast<const T *>(x0), x0 ? 1 : 0, static_cast<T *>(evec), ev, rn);
    } else {
        EigsStats s1;
        st             = eigs_mixed<T, T32>(op, s, p, static_cast<const T *>(x0), static_cast<T *>(evec), ev, rn, &s1);
        r.matvecs_fp32 = s1.matvecs + st.inner_matvecs;
        r.iters        = s1.iters;
    }
    r.eigval  = ev[0];
    r.rnorm   = rn[0];
    r.anorm   = st.anorm;
    r.nconv   = st.nconv;
    r.iters  += st.iters;
    r.matvecs = st.matvecs;
    return r;
}

template <typename T, typename Eng = Engine<T>, typename T32 = void>
void batched_impl(MatVecOp *const *ops, int k, const EigsParams &p, const void *const *x0s, void *const *evecs,
                  std::vector<BatchEigsResult> &res, hipStream_t s) {
    res.assign(k, BatchEigsResult{});
    for(int q = 0; q < k; ++q)
        if(ops[q]->D.n() >= SOLO_MIN_N) res[q] = solve_alone<T, T32>(*ops[q], p, x0s ? x0s[q] : nullptr, evecs[q], s);
    Eng E(s, p);
    for(int q = 0; q < k; ++q)
        if(ops[q]->D.n() < SOLO_MIN_N) E.add(ops[q], x0s ? x0s[q] : nullptr, evecs[q], q);
    while(E.nactive > 0)
        for(const auto &d : E.step()) res[d.id] = d.res;
}

template <typename T, typename Eng = Engine<T>>
void serve_impl(const std::vector<EigRequest> &req, const EigsParams &p, int64_t budget, int max_batch,
                std::vector<EigRequestResult> &out, ServeSummary &sum, hipStream_t s) {
    using clock   = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto       now = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
    Eng        E(s, p);
    const int  n = static_cast<int>(req.size());
    out.assign(n, EigRequestResult{});
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });
    // A request that cannot fit in the budget even alone is an error of the whole call, as in dmrg_serve.
    for(int i = 0; i < n; ++i) {
        int mmax, mmin;
        Eng::basis_rule(p, req[i].op->D.n(), mmax, mmin);
        DMRG_REQUIRE(Eng::footprint_of(*req[i].op, mmax, mmin) <= budget, DMRG_ERR_OOM,
                     "request " + std::to_string(i) + " does not fit in the HBM budget");
    }
    size_t  next = 0;
    int64_t admitted = 0;
    int     finished = 0;
    while(finished < n) {
        // Admit, in arrival order, the requests that have arrived and fit in the budget and the batch.
        const double t = now();
        while(next < order.size()) {
            const EigRequest &r = req[order[next]];
            if(r.arrival_s > t) break;
            int mmax, mmin;
            Eng::basis_rule(p, r.op->D.n(), mmax, mmin);
            const int64_t fp = Eng::footprint_of(*r.op, mmax, mmin);
            if(admitted + fp > budget || (max_batch > 0 && E.nactive >= max_batch)) break;
            E.add(r.op, r.x0, r.evec, order[next]);
            EigRequestResult &o = out[order[next]];
            o.arrival_s         = r.arrival_s;
            o.start_s           = t;
            o.footprint_bytes   = fp;
            admitted += fp;
            sum.peak_admitted_bytes = std::max(sum.peak_admitted_bytes, admitted);
            sum.peak_concurrency    = std::max(sum.peak_concurrency, E.nactive);
            ++next;
        }
        if(E.nactive == 0) { // idle until the next arrival
            const double wait = req[order[next]].arrival_s - now();
            if(wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
            continue;
        }
        for(const auto &d : E.step()) {
            EigRequestResult &o = out[d.id];
            o.res               = d.res;
            o.finish_s          = now();
            o.latency_s         = o.finish_s - o.arrival_s;
            admitted -= d.footprint;
            ++finished;
        }
    }
    sum.makespan_s = now();
    sum.steps      = E.steps;
}

} // namespace

struct EigService::Impl {
    struct Req {
        MatVecOp       *op   = nullptr;
        const void     *x0   = nullptr;
        void           *evec = nullptr;
        BatchEigsResult res;
        bool            done   = false;
        int             status = DMRG_OK;
        std::string     err;
    };
    dmrg_dtype              dtype;
    EigsParams              p;
    mutable std::mutex      mu;
    std::condition_variable cv, cv_done;
    std::deque<Req *>       queue;
    bool                    stop = false;
    Stats                   st;
    std::thread             th;

    template <typename T>
    void run() {
        hipStream_t s = nullptr;
        if(hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
        std::unordered_map<int, Req *> live;
        int                            next_id = 0;
        auto fail_all = [&](int status, const std::string &msg) {
            std::lock_guard<std::mutex> lk(mu);
            for(auto &kv : live) kv.second->status = status, kv.second->err = msg, kv.second->done = true;
            live.clear();
            cv_done.notify_all();
        };
        auto E = std::make_unique<Engine<T>>(s, p, false);
        E->eager_retire = true;
        while(true) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !queue.empty() || E->nactive > 0; });
                if(stop && queue.empty() && E->nactive == 0) break;
                while(!queue.empty()) {
                    Req *r = queue.front();
                    queue.pop_front();
                    try {
                        E->add(r->op, r->x0, r->evec, next_id);
                        live[next_id++] = r;
                    } catch(const Error &e) {
                        r->status = e.status, r->err = e.what(), r->done = true;
                        cv_done.notify_all();
                    }
                }
                st.peak_batch = std::max(st.peak_batch, E->nactive);
            }
            try {
                const auto done = E->step();
                std::lock_guard<std::mutex> lk(mu);
                st.steps     = E->steps;
                st.rebuilds  = E->rebuilds;
                st.t_rebuild = E->t_rebuild;
                st.t_rr      = E->t_rr;
                st.t_step    = E->t_step;
                for(const auto &d : done) {
                    auto it = live.find(d.id);
                    if(it == live.end()) continue;
                    it->second->res  = d.res;
                    it->second->done = true;
                    live.erase(it);
                    ++st.solves;
                }
                if(!done.empty()) cv_done.notify_all();
            } catch(const Error &e) {
                fail_all(e.status, e.what());
                E.reset();
                E = std::make_unique<Engine<T>>(s, p, false);
                E->eager_retire = true;
            } catch(const std::exception &e) {
                fail_all(DMRG_ERR_HIP, e.what());
                E.reset();
                E = std::make_unique<Engine<T>>(s, p, false);
                E->eager_retire = true;
            }
        }
        E.reset();
        if(s) (void) hipStreamDestroy(s);
    }
};

EigService::EigService(dmrg_dtype t, const EigsParams &p) : impl_(std::make_unique<Impl>()) {
    DMRG_REQUIRE(t == DMRG_FP64 || t == DMRG_CX64, DMRG_ERR_ARG, "eigensolve service: FP64 or CX64 operators");
    DMRG_REQUIRE(std::max(1, p.nev) == 1 && p.precond == 0 && (p.target == TARGET_SA || p.target == TARGET_LA), DMRG_ERR_UNSUPPORTED,
                 "eigensolve service: one pair per problem at SA or LA, no preconditioner");
    impl_->dtype = t;
    impl_->p     = p;
    Impl *im     = impl_.get();
    if(t == DMRG_FP64) impl_->th = std::thread([im] { im->run<double>(); });
    else impl_->th = std::thread([im] { im->run<cpx<double>>(); });
}

EigService::~EigService() {
    {
        std::lock_guard<std::mutex> lk(impl_->mu);
        impl_->stop = true;
    }
    impl_->cv.notify_all();
    if(impl_->th.joinable()) impl_->th.join();
}

BatchEigsResult EigService::solve(MatVecOp *op, const void *x0, void *evec) {
    DMRG_REQUIRE(op && evec && op->dtype == impl_->dtype, DMRG_ERR_ARG, "eigensolve service: bad problem");
    Impl::Req r;
    r.op   = op;
    r.x0   = x0;
    r.evec = evec;
    {
        std::lock_guard<std::mutex> lk(impl_->mu);
        DMRG_REQUIRE(!impl_->stop, DMRG_ERR_ARG, "eigensolve service: stopped");
        impl_->queue.push_back(&r);
    }
    impl_->cv.notify_all();
    {
        std::unique_lock<std::mutex> lk(impl_->mu);
        impl_->cv_done.wait(lk, [&] { return r.done; });
    }
    if(r.status != DMRG_OK) throw Error(static_cast<dmrg_status>(r.status), r.err);
    return r.res;
}

EigService::Stats EigService::stats() const {
    std::lock_guard<std::mutex> lk(impl_->mu);
    return impl_->st;
}

void eigs_batched(MatVecOp *const *ops, int k, const EigsParams &p, const void *const *x0s, void *const *evecs,
                  std::vector<BatchEigsResult> &res, hipStream_t s, bool mixed) {
    DMRG_REQUIRE(ops && evecs && k >= 1, DMRG_ERR_ARG, "batched eigs: bad arguments");
    DMRG_REQUIRE(std::max(1, p.nev) == 1 && p.precond == 0 && (p.target == TARGET_SA || p.target == TARGET_LA), DMRG_ERR_UNSUPPORTED,
                 "batched eigs: one pair per problem at SA or LA, no preconditioner");
    const dmrg_dtype t = ops[0]->dtype;
    for(int q = 0; q < k; ++q) DMRG_REQUIRE(ops[q] && ops[q]->dtype == t && evecs[q], DMRG_ERR_ARG, "batched eigs: one dtype, outputs");
    if(t == DMRG_FP64) {
        if(mixed) batched_impl<double, MixedEngine<double, float>, float>(ops, k, p, x0s, evecs, res, s);
        else batched_impl<double>(ops, k, p, x0s, evecs, res, s);
    } else if(t == DMRG_CX64) {
        if(mixed) batched_impl<cpx<double>, MixedEngine<cpx<double>, cpx<float>>, cpx<float>>(ops, k, p, x0s, evecs, res, s);
        else batched_impl<cpx<double>>(ops, k, p, x0s, evecs, res, s);
    } else {
        throw Error(DMRG_ERR_ARG, "batched eigs: FP64 or CX64 operators");
    }
}

void serve_eigs(const std::vector<EigRequest> &req, const EigsParams &p, int64_t budget, int max_batch,
                std::vector<EigRequestResult> &out, ServeSummary &sum, hipStream_t s, bool mixed) {
    DMRG_REQUIRE(!req.empty(), DMRG_ERR_ARG, "serve eigs: no requests");
    DMRG_REQUIRE(std::max(1, p.nev) == 1 && p.precond == 0 && (p.target == TARGET_SA || p.target == TARGET_LA), DMRG_ERR_UNSUPPORTED,
                 "serve eigs: one pair per problem at SA or LA, no preconditioner");
    const dmrg_dtype t = req[0].op->dtype;
    for(const auto &r : req) DMRG_REQUIRE(r.op && r.op->dtype == t && r.evec, DMRG_ERR_ARG, "serve eigs: one dtype, outputs");
    if(t == DMRG_FP64) {
        if(mixed) serve_impl<double, MixedEngine<double, float>>(req, p, budget, max_batch, out, sum, s);
        else serve_impl<double>(req, p, budget, max_batch, out, sum, s);
    } else if(t == DMRG_CX64) {
        if(mixed) serve_impl<cpx<double>, MixedEngine<cpx<double>, cpx<float>>>(req, p, budget, max_batch, out, sum, s);
        else serve_impl<cpx<double>>(req, p, budget, max_batch, out, sum, s);
    } else {
        throw Error(DMRG_ERR_ARG, "serve eigs: FP64 or CX64 operators");
    }
}

} // namespace dmrg

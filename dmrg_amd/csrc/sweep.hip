// Finite-chain DMRG sweep driver for the random transverse-field Ising model, built on the GPU
// contraction, eigensolver and SVD paths. Counterpart of fdmrg::run_algorithm / update_state
// (fdmrg.cpp:170-376) with the ising_tf_rf model (IsingRandomField.cpp:93-177):
//   - all MPS tensors, env blocks (H and H^2) and intermediates stay in HBM;
//   - two-site sweeps grow the bonds up to chi_max (truncation rule of svd.cpp:257-280),
//   - single-site sweeps then converge the state to an energy variance below var_tol,
//     with H^2 built from energy-shifted MPOs as set_energy_shift_mpo does (MpoSite.cpp:57-69).
#include "sweep.hpp"

#include "batch_eigs.hpp"

#include "dense.hpp"

#include "eigs.hpp"
#include "expand.hpp"
#include "svd.hpp"

#include <chrono>
#include <cmath>
#include <complex>
#include <memory>

namespace dmrg {

using cx = std::complex<double>;

// Ising MPO of one site (IsingRandomField::get_mpo, rows wL, cols wR) with the chain edges applied
// (MpoSite::apply_edge_left/right: site 0 keeps row D-1, site L-1 keeps column 0).
static std::vector<cx> ising_site_mpo(int i, int L, double J, double gi, double eshift, int64_t dims[4]) {
    const cx I2[2][2] = {{1, 0}, {0, 1}}, SX[2][2] = {{0, 1}, {1, 0}}, SZ[2][2] = {{1, 0}, {0, -1}};
    const double Ji = (i == L - 1) ? 0.0 : J;
    std::vector<cx> W(3 * 3 * 2 * 2, cx(0.0, 0.0));
    auto at = [&](int a, int b, int s, int t) -> cx & { return W[a + 3 * (b + 3 * (s + 2 * t))]; };
    for(int s = 0; s < 2; ++s)
        for(int t = 0; t < 2; ++t) {
            at(0, 0, s, t) = I2[s][t];
            at(1, 0, s, t) = SZ[s][t];
            at(2, 0, s, t) = -gi * SX[s][t] - eshift * I2[s][t];
            at(2, 1, s, t) = -Ji * SZ[s][t];
            at(2, 2, s, t) = I2[s][t];
        }
    const int r0 = (i == 0) ? 2 : 0, r1 = (i == 0) ? 3 : 3;
    const int c0 = 0, c1 = (i == L - 1) ? 1 : 3;
    dims[0] = r1 - r0, dims[1] = c1 - c0, dims[2] = 2, dims[3] = 2;
    std::vector<cx> out(dims[0] * dims[1] * 4);
    for(int a = r0; a < r1; ++a)
        for(int b = c0; b < c1; ++b)
            for(int s = 0; s < 2; ++s)
                for(int t = 0; t < 2; ++t) out[(a - r0) + dims[0] * ((b - c0) + dims[1] * (s + 2 * t))] = at(a, b, s, t);
    return out;
}

// W2[(w1,w2),(w1',w2'),s,s''] = sum_s' W[w1,w1',s,s'] conj(W[w2,w2',s',s'']) (MpoSite.cpp:57-69).
std::vector<cx> mpo_squared(const std::vector<cx> &W, const int64_t d[4], int64_t out[4]) {
    const int64_t DL = d[0], DR = d[1], p = d[2];
    out[0] = DL * DL, out[1] = DR * DR, out[2] = p, out[3] = p;
    std::vector<cx> W2(out[0] * out[1] * p * p, cx(0.0, 0.0));
    for(int64_t s2 = 0; s2 < p; ++s2)
        for(int64_t s = 0; s < p; ++s)
            for(int64_t w2p = 0; w2p < DR; ++w2p)
                for(int64_t w1p = 0; w1p < DR; ++w1p)
                    for(int64_t w2 = 0; w2 < DL; ++w2)
                        for(int64_t w1 = 0; w1 < DL; ++w1) {
                            cx acc(0.0, 0.0);
                            for(int64_t sp = 0; sp < p; ++sp)
                                acc += W[w1 + DL * (w1p + DR * (s + p * sp))] * std::conj(W[w2 + DL * (w2p + DR * (sp + p * s2))]);
                            W2[(w1 + DL * w2) + out[0] * ((w1p + DR * w2p) + out[1] * (s + p * s2))] = acc;
                        }
    return W2;
}

namespace {

// settings::precision::eig_max_size (source/config/settings.h:249): largest problem for the dense solver.
constexpr int64_t EIG_MAX_SIZE = 4096;

template <typename T>
struct Fdmrg {
    using clock = std::chrono::steady_clock;
    static constexpr dmrg_dtype DT = std::is_same_v<T, double> ? DMRG_FP64 : DMRG_CX64;
    const dmrg_fdmrg_config &cfg;
    hipStream_t              s;
    int                      L, d = 2;
    std::vector<std::vector<cx>>              W, W2;        // host MPOs (cx64)
    std::vector<std::array<int64_t, 4>>       Wd, W2d;
    std::vector<DevBuf>                       Wdev, W2dev;   // device MPOs in dtype T (env updates)
    std::vector<DevBuf>                       M;             // MPS site tensors (d, chi[i], chi[i+1])
    std::vector<int64_t>                      chi;           // chi[0..L]
    std::vector<DevBuf>                       EL, ER, EL2, ER2; // EL[i]: left of site i; ER[i]: right of site i
    DevBuf                                    theta, theta2, xbuf, ws, tmp, sdev, dense;
    DeviceSvd                                 svd;
    std::unique_ptr<MatVecOp>                 hop, h2op;
    std::unique_ptr<Davidson<T>>              dav;
    std::vector<double>                       entropy, terr;
    double                                    eshift = 0.0;
    bool                                      folded = false; // xDMRG: solves on (H - E_tgt)^2
    dmrg_fdmrg_result                        *res;
    EigService                               *svc = nullptr; // shared batcher for the iterative FP64 energy solves
    double t_eigs = 0, t_env = 0, t_svd = 0, t_expand = 0;
    double env_flop = 0; // algorithmic FLOP of the env updates

    Fdmrg(const dmrg_fdmrg_config &c, hipStream_t st, dmrg_fdmrg_result *r) : cfg(c), s(st), L(c.L), svd(st), res(r) {}

    size_t es() const { return sizeof(T); }

    void upload(DevBuf &b, const std::vector<cx> &h) {
        b.reserve(h.size() * es());
        if constexpr(std::is_same_v<T, double>) {
            std::vector<double> re(h.size());
            for(size_t i = 0; i < h.size(); ++i) re[i] = h[i].real();
            DMRG_HIP_CHECK(hipMemcpy(b.p, re.data(), re.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            DMRG_HIP_CHECK(hipMemcpy(b.p, h.data(), h.size() * sizeof(cx), hipMemcpyHostToDevice));
        }
    }

    void build_mpos(double e_per_site) {
        W.assign(L, {});
        W2.assign(L, {});
        Wd.assign(L, {});
        W2d.assign(L, {});
        Wdev.resize(L);
        W2dev.resize(L);
        for(int i = 0; i < L; ++i) {
            int64_t dd[4];
            // H itself is built without shift; the squared MPO carries the shift E/L per site so that
            // <H^2> measured on it is the variance around E without cancellation.
            W[i] = ising_site_mpo(i, L, cfg.J, cfg.h_tran + cfg.h_rand[i], 0.0, dd);
            std::copy(dd, dd + 4, Wd[i].begin());
            int64_t         ds[4];
            std::vector<cx> Ws = ising_site_mpo(i, L, cfg.J, cfg.h_tran + cfg.h_rand[i], e_per_site, ds);
            int64_t         d2[4];
            W2[i] = mpo_squared(Ws, ds, d2);
            std::copy(d2, d2 + 4, W2d[i].begin());
            upload(Wdev[i], W[i]);
            upload(W2dev[i], W2[i]);
        }
    }

    Dims dims_site(int i, bool sq) const {
        Dims D;
        D.d    = d;
        D.chiL = chi[i];
        D.chiR = chi[i + 1];
        D.DL   = sq ? W2d[i][0] : Wd[i][0];
        D.DR   = sq ? W2d[i][1] : Wd[i][1];
        return D;
    }

    void set_edge(DevBuf &b) {
        b.reserve(es());
        T one;
        if constexpr(std::is_same_v<T, double>) one = 1.0;
        else one = T{1.0, 0.0};
        DMRG_HIP_CHECK(hipMemcpyAsync(b.p, &one, es(), hipMemcpyHostToDevice, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
    }

    void env_left(int i) { // EL[i+1] from EL[i], M[i] (left-normalised), W[i]
        const auto t0 = clock::now();
        for(int sq = 0; sq < 2; ++sq) {
            const Dims D  = dims_site(i, sq);
            DevBuf    &in = sq ? EL2[i] : EL[i];
            DevBuf    &o  = sq ? EL2[i + 1] : EL[i + 1];
            o.reserve(static_cast<size_t>(D.chiR * D.chiR * D.DR) * es());
            void *w = ws.reserve(static_cast<size_t>(env_ws_elems(D)) * es());
            env_build_block(DT, 'L', D, o.p, in.p, M[i].p, (sq ? W2dev : Wdev)[i].p, w, s);
            env_flop += D.macs() * (std::is_same_v<T, double> ? 2.0 : 8.0);
        }
        t_env += std::chrono::duration<double>(clock::now() - t0).count();
    }
    void env_right(int i) { // ER[i-1] from ER[i], M[i] (right-normalised), W[i]
        const auto t0 = clock::now();
        for(int sq = 0; sq < 2; ++sq) {
            const Dims D  = dims_site(i, sq);
            DevBuf    &in = sq ? ER2[i] : ER[i];
            DevBuf    &o  = sq ? ER2[i - 1] : ER[i - 1];
            o.reserve(static_cast<size_t>(D.chiL * D.chiL * D.DL) * es());
            void *w = ws.reserve(static_cast<size_t>(env_ws_elems(D)) * es());
            env_build_block(DT, 'R', D, o.p, in.p, M[i].p, (sq ? W2dev : Wdev)[i].p, w, s);
            env_flop += D.macs() * (std::is_same_v<T, double> ? 2.0 : 8.0);
        }
        t_env += std::chrono::duration<double>(clock::now() - t0).count();
    }

    void init_state() {
        chi.assign(L + 1, 1);
        M.resize(L);
        // Random product state (seeded): each site a normalised random d-vector.
        uint64_t x = cfg.seed * 6364136223846793005ull + 1442695040888963407ull;
        auto     rnd = [&]() {
            x ^= x >> 12, x ^= x << 25, x ^= x >> 27;
            return static_cast<double>((x * 2685821657736338717ull) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        };
        for(int i = 0; i < L; ++i) {
            std::vector<T> v(d);
            double         nrm = 0;
            for(int k = 0; k < d; ++k) {
                if constexpr(std::is_same_v<T, double>) v[k] = rnd() + (k == 0 ? 1.5 : 0.0);
                else v[k] = T{rnd() + (k == 0 ? 1.5 : 0.0), rnd()};
                if constexpr(std::is_same_v<T, double>) nrm += v[k] * v[k];
                else nrm += v[k].re * v[k].re + v[k].im * v[k].im;
            }
            for(auto &e : v) {
                if constexpr(std::is_same_v<T, double>) e /= std::sqrt(nrm);
                else e = T{e.re / std::sqrt(nrm), e.im / std::sqrt(nrm)};
            }
            M[i].reserve(d * es());
            DMRG_HIP_CHECK(hipMemcpy(M[i].p, v.data(), d * es(), hipMemcpyHostToDevice));
        }
        EL.resize(L + 1), ER.resize(L + 1), EL2.resize(L + 1), ER2.resize(L + 1);
        set_edge(EL[0]), set_edge(EL2[0]), set_edge(ER[L - 1]), set_edge(ER2[L - 1]);
        for(int i = L - 1; i >= 1; --i) env_right(i);
    }

    struct Solve {
        double E;
        double rnorm;
    };

    // Point hop (H, or with sq the shifted H^2 on the variance envs) at the active sites [i, i+ns).
    MatVecOp &bind_op(std::unique_ptr<MatVecOp> &h, int i, int ns, bool sq) {
        const auto               &Ws  = sq ? W2 : W;
        const auto               &Wds = sq ? W2d : Wd;
        std::vector<const void *> mp;
        std::vector<int64_t>      md;
        for(int k = 0; k < ns; ++k) {
            mp.push_back(Ws[i + k].data());
            md.insert(md.end(), Wds[i + k].begin(), Wds[i + k].end());
        }
        const int64_t dl[3] = {chi[i], chi[i], Wds[i][0]};
        const int64_t dr[3] = {chi[i + ns], chi[i + ns], Wds[i + ns - 1][1]};
        const void   *el = (sq ? EL2 : EL)[i].p, *er = (sq ? ER2 : ER)[i + ns - 1].p;
        if(!h) {
            h                = std::make_unique<MatVecOp>(DT, ns, mp.data(), md.data(), el, dl, er, dr, true);
            h->upload_stream = s; // later rebinds upload their MPO on the sweep's stream
        } else {
            h->rebind(ns, mp.data(), md.data(), el, dl, er, dr);
        }
        return *h;
    }

    // <theta|H|theta> / <theta|theta> on the active sites [i, i+ns).
    double energy_of(int i, int ns, const T *th) {
        MatVecOp     &op = bind_op(hop, i, ns, false);
        const int64_t n  = op.D.n();
        tmp.reserve(static_cast<size_t>(n) * es() + 64 * sizeof(double));
        op.multax_device(th, n, tmp.p, n, 1, s, false);
        double *red = reinterpret_cast<double *>(static_cast<char *>(tmp.p) + static_cast<size_t>(n) * es());
        dot_device(DT, n, th, tmp.p, red, s);
        dot_device(DT, n, th, th, red + 2, s);
        double r[4];
        DMRG_HIP_CHECK(hipMemcpyAsync(r, red, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        return r[0] / r[2];
    }

    // Eigenvector of H_eff on the active sites [i, i+ns) with theta as the start vector; theta <- eigenvector.
    // Ground state (SA) of H, or in the folded mode the eigenvalue of (H - E_tgt)^2 closest to 0 (the
    // folded-spectrum executor with primme_closest_abs at target 0, optimize_folded_spectrum.cpp:272-342),
    // reported with its energy <H>.
    Solve solve(int i, int ns, T *th) {
        const auto    t0 = clock::now();
        MatVecOp     &op = folded ? bind_op(h2op, i, ns, true) : bind_op(hop, i, ns, false);
        const int64_t n  = op.D.n();
        EigsParams    p;
        p.target      = folded ? TARGET_CLOSEST : TARGET_SA;
        p.shift       = 0.0;
        p.tol         = cfg.eigs_tol > 0 ? cfg.eigs_tol : 1e-12;
        p.ncv         = cfg.eigs_ncv;
        p.max_matvecs = 20000;
        p.seed        = cfg.seed + 17;
        double    ev = 0, rn = 0;
        EigsStats st;
        xbuf.reserve(static_cast<size_t>(n) * es());
        const int64_t mixed_min = cfg.mixed_min_n > 0 ? cfg.mixed_min_n : 131072;
        if(folded && n <= EIG_MAX_SIZE) {
            // Small folded problems go to the dense solver, as xDMRG picks OptSolver::EIG for problem sizes up to
            // eig_max_size (xdmrg.cpp:367, optimize_folded_spectrum_eig.cpp:17-36): the folded spectrum's gaps are
            // tiny relative to its norm, where an iterative solver can settle on the second-closest level.
            dense.reserve(static_cast<size_t>(n) * n * es());
            heff_dense(DT, op.D, op.mpo.p, op.envL_p, op.envR_p, dense.p, n, false, s);
            std::vector<double> w;
            svd.eigh<T>(static_cast<int>(n), dense.as<T>(), static_cast<int>(n), true, w);
            int k = 0;
            for(int j = 1; j < static_cast<int>(n); ++j)
                if(std::fabs(w[j]) < std::fabs(w[k])) k = j; // closest_abs to the target 0
            DMRG_HIP_CHECK(hipMemcpyAsync(th, dense.as<T>() + static_cast<int64_t>(k) * n, static_cast<size_t>(n) * es(),
                                          hipMemcpyDeviceToDevice, s));
            res->eigensolves += 1;
            ev = energy_of(i, ns, th);
            t_eigs += std::chrono::duration<double>(clock::now() - t0).count();
            return {ev, 0.0};
        }
        if(!folded && cfg.eig_max_size > 0 && n <= cfg.eig_max_size) {
            // The reference's dense branch for small problems (OptSolver::EIG when problem_size <= eig_max_size,
            // fdmrg.cpp:307, optimize_energy_eig.cpp:18-58): explicit H_eff, lowest eigenpair.
            dense.reserve(static_cast<size_t>(n) * n * es());
            heff_dense(DT, op.D, op.mpo.p, op.envL_p, op.envR_p, dense.p, n, false, s);
            std::vector<double> w;
            svd.eigh<T>(static_cast<int>(n), dense.as<T>(), static_cast<int>(n), true, w);
            DMRG_HIP_CHECK(hipMemcpyAsync(th, dense.as<T>(), static_cast<size_t>(n) * es(), hipMemcpyDeviceToDevice, s));
            res->eigensolves += 1;
            t_eigs += std::chrono::duration<double>(clock::now() - t0).count();
            return {w[0], 0.0};
        }
        if(svc && !folded && !(cfg.mixed && n >= mixed_min)) {
            // Batched with the solves of the other requests (same method, basis rule and tolerance).
            DMRG_HIP_CHECK(hipStreamSynchronize(s)); // theta and the operator's blocks are complete
            const BatchEigsResult r = svc->solve(&op, th, xbuf.p);
            ev = r.eigval, rn = r.rnorm;
            st.matvecs = r.matvecs;
            op.num_mv += r.matvecs;
            op.flops += static_cast<double>(r.matvecs) * op.macs() * (dtype_is_complex(op.dtype) ? 8.0 : 2.0);
        } else if(cfg.mixed && !folded && n >= mixed_min) {
            EigsStats s1;
            p.inner_max = 20; // FP32 correction solves in the FP64 phase
            st = eigs_mixed<T, std::conditional_t<std::is_same_v<T, double>, float, cpx<float>>>(op, s, p, th, xbuf.as<T>(), &ev, &rn, &s1);
            res->matvecs_fp32 += s1.matvecs + st.inner_matvecs;
        } else {
            if(!dav) dav = std::make_unique<Davidson<T>>(op, s);
            dav->rebind(op);
            st = dav->solve(p, th, 1, xbuf.as<T>(), &ev, &rn);
        }
        DMRG_HIP_CHECK(hipMemcpyAsync(th, xbuf.p, static_cast<size_t>(n) * es(), hipMemcpyDeviceToDevice, s));
        res->matvecs += st.matvecs;
        res->eigensolves += 1;
        if(folded) ev = energy_of(i, ns, th);
        t_eigs += std::chrono::duration<double>(clock::now() - t0).count();
        return {ev, rn};
    }

    // <theta|(H - E_shift)^2|theta> on site i (single-site, H^2 envs).
    double expval_sq(int i, const T *th) {
        MatVecOp     &op = bind_op(h2op, i, 1, true);
        const int64_t n  = op.D.n();
        tmp.reserve(static_cast<size_t>(n) * es() + 64 * sizeof(double) + 2 * 512 * sizeof(double) + 64);
        op.multax_device(th, n, tmp.p, n, 1, s, false);
        double *red = reinterpret_cast<double *>(static_cast<char *>(tmp.p) + static_cast<size_t>(n) * es());
        dot_device(DT, n, th, tmp.p, red, s);
        double r[2];
        DMRG_HIP_CHECK(hipMemcpyAsync(r, red, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        DMRG_HIP_CHECK(hipStreamSynchronize(s));
        return r[0];
    }

    // theta[(s1 + d s2), a, b] = sum_k M[i][s1, a, k] M[i+1][s2, k, b]
    void merge_two(int i, T *th) {
        const int64_t a = chi[i], k = chi[i + 1], b = chi[i + 2];
        GemmDesc      g;
        g.dtype = DT;
        g.M = static_cast<int32_t>(d * a), g.N = static_cast<int32_t>(d * b), g.K = static_cast<int32_t>(k), g.batch = 1;
        g.A = M[i].p, g.am = plain(1), g.ak = plain(d * a);
        g.B = M[i + 1].p, g.bk = plain(d), g.bn = fused(1, d * k, d);
        g.C = th, g.cm = fused(1, d * d, d), g.cn = fused(d, d * d * a, d);
        gemm(g, s);
    }

    static double entropy_of(const std::vector<double> &S, int r) {
        double e = 0;
        for(int k = 0; k < r; ++k) {
            const double p = S[k] * S[k];
            if(p > 0) e -= p * std::log(p);
        }
        return e;
    }

    // Two-site update of sites (i, i+1); moving right leaves the centre on i+1, moving left on i.
    double two_site(int i, bool right) {
        const int64_t a = chi[i], b = chi[i + 2];
        const int64_t n = d * d * a * b;
        theta.reserve(static_cast<size_t>(n) * es());
        merge_two(i, theta.as<T>());
        const Solve sol = solve(i, 2, theta.as<T>());
        split_two(i, right, theta.as<T>());
        return sol.E;
    }

    // Truncated split of the two-site tensor th of sites (i, i+1) (mps::merge_multisite_mps, split.cpp:21-346):
    // centre_right puts the singular values on site i+1 (A_i = U) and rebuilds the left envs of site i+1,
    // otherwise on site i (B_{i+1} = Vt) and rebuilds the right envs of site i.
    void split_two(int i, bool right, const T *th) {
        const int64_t a = chi[i], b = chi[i + 2];
        const int64_t n = d * d * a * b;
        // X[(s1, a), (s2, b)] = theta[s1 + d s2, a, b]
        theta2.reserve(static_cast<size_t>(n) * es());
        remap2d(DT, d * a, d * b, th, fused(1, d * d, d), fused(d, d * d * a, d), theta2.p, plain(1), plain(d * a), nullptr, 0, s);
        const auto          t0 = clock::now();
        std::vector<double> S;
        svd.compute<T>(static_cast<int>(d * a), static_cast<int>(d * b), theta2.as<T>(), S, SVD_JACOBI);
        auto [r, te] = rank_from_truncation_error(S, cfg.trunc_lim > 0 ? cfg.trunc_lim : 1e-14, cfg.chi_max);
        double nrm = 0;
        for(int k = 0; k < r; ++k) nrm += S[k] * S[k];
        nrm = std::sqrt(nrm);
        for(int k = 0; k < r; ++k) S[k] /= nrm;
        sdev.reserve(S.size() * sizeof(double));
        DMRG_HIP_CHECK(hipMemcpyAsync(sdev.p, S.data(), r * sizeof(double), hipMemcpyHostToDevice, s));
        const int     k  = svd.k();
        const int64_t mU = d * a;
        M[i].reserve(static_cast<size_t>(d * a * r) * es());
        M[i + 1].reserve(static_cast<size_t>(d * r * b) * es());
        if(right) {
            // A_i = U[:, :r]; M_{i+1}[s2, k, b] = S_k Vt[k, s2 + d b]
            DMRG_HIP_CHECK(hipMemcpyAsync(M[i].p, svd.U(), static_cast<size_t>(mU * r) * es(), hipMemcpyDeviceToDevice, s));
            remap2d(DT, r, d * b, svd.Vt(), plain(1), plain(k), M[i + 1].p, plain(d), fused(1, d * r, d), sdev.as<double>(), 1, s);
        } else {
            // M_i[s1, a, k] = U[(s1,a), k] S_k; B_{i+1}[s2, k, b] = Vt[k, s2 + d b]
            remap2d(DT, mU, r, svd.U(), plain(1), plain(mU), M[i].p, plain(1), plain(mU), sdev.as<double>(), 2, s);
            remap2d(DT, r, d * b, svd.Vt(), plain(1), plain(k), M[i + 1].p, plain(d), fused(1, d * r, d), nullptr, 0, s);
        }
        chi[i + 1]      = r;
        entropy[i + 1]  = entropy_of(S, r);
        terr[i + 1]     = te;
        t_svd += std::chrono::duration<double>(clock::now() - t0).count();
        if(right) env_left(i);
        else env_right(i + 1);
    }

    // Subspace expansion before the single-site step at site c (env.cpp:1285-1380 with NSITE | FORWARD): the
    // active site and the next one in the sweep direction form a two-site block; its optimally mixed block
    // (the combination of theta, H theta and (H - E_s)^2 theta of lowest K3 value, env.cpp:96-239; H1 / H2 bits of
    // the mode) replaces it, split back with the bond limit and the truncation limit and the centre left on c.
    void expand_step(int c, bool right) {
        const int i = right ? c : c - 1;
        if(i < 0 || i + 1 >= L) return;
        const int64_t n = d * d * chi[i] * chi[i + 2];
        theta.reserve(static_cast<size_t>(n) * es());
        merge_two(i, theta.as<T>());
        MatVecOp &H1 = bind_op(hop, i, 2, false);
        MatVecOp &H2 = bind_op(h2op, i, 2, true);
        xbuf.reserve(static_cast<size_t>(n) * es());
        MixedBlockResult mr;
        const auto       t0 = clock::now();
        optimally_mixed_block(H1, H2, theta.p, cfg.expand_mode & 3, xbuf.p, mr, s);
        t_expand += std::chrono::duration<double>(clock::now() - t0).count();
        ++res->expansions;
        split_two(i, !right, xbuf.as<T>());
    }

    // Moves the centre from site c (tensor th, shape (d, chi[c], chi[c+1])) to site c + 1: th as (d a) x b =
    // U S Vt, A_c = U, M_{c+1} <- (S Vt) M_{c+1}; records the Schmidt values' entropy at bond c + 1.
    void move_right(int c, T *th) {
        const int64_t       a = chi[c], b = chi[c + 1];
        std::vector<double> S;
        svd.compute<T>(static_cast<int>(d * a), static_cast<int>(b), th, S, SVD_JACOBI);
        const int k = svd.k();
        sdev.reserve(S.size() * sizeof(double));
        DMRG_HIP_CHECK(hipMemcpyAsync(sdev.p, S.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
        M[c].reserve(static_cast<size_t>(d * a * k) * es());
        DMRG_HIP_CHECK(hipMemcpyAsync(M[c].p, svd.U(), static_cast<size_t>(d * a * k) * es(), hipMemcpyDeviceToDevice, s));
        theta2.reserve(static_cast<size_t>(k * b) * es());
        remap2d(DT, k, b, svd.Vt(), plain(1), plain(k), theta2.p, plain(1), plain(k), sdev.as<double>(), 1, s); // S Vt
        const int64_t e = chi[c + 2];
        tmp.reserve(static_cast<size_t>(d * k * e) * es());
        GemmDesc g;
        g.dtype = DT;
        g.M = k, g.N = static_cast<int32_t>(d * e), g.K = static_cast<int32_t>(b), g.batch = 1;
        g.A = theta2.p, g.am = plain(1), g.ak = plain(k);
        g.B = M[c + 1].p, g.bk = plain(d), g.bn = fused(1, d * b, d);
        g.C = tmp.p, g.cm = plain(d), g.cn = fused(1, d * k, d);
        gemm(g, s);
        M[c + 1].reserve(static_cast<size_t>(d * k * e) * es());
        DMRG_HIP_CHECK(hipMemcpyAsync(M[c + 1].p, tmp.p, static_cast<size_t>(d * k * e) * es(), hipMemcpyDeviceToDevice, s));
        chi[c + 1]     = k;
        entropy[c + 1] = entropy_of(S, k);
    }

    // Single-site update of site c, then move the centre one site (right or left).
    std::pair<double, double> one_site(int c, bool right) {
        if(cfg.expand_mode != 0) expand_step(c, right);
        const int64_t a = chi[c], b = chi[c + 1];
        const int64_t n = d * a * b;
        theta.reserve(static_cast<size_t>(n) * es());
        DMRG_HIP_CHECK(hipMemcpyAsync(theta.p, M[c].p, static_cast<size_t>(n) * es(), hipMemcpyDeviceToDevice, s));
        const Solve sol = solve(c, 1, theta.as<T>());
        DMRG_HIP_CHECK(hipMemcpyAsync(M[c].p, theta.p, static_cast<size_t>(n) * es(), hipMemcpyDeviceToDevice, s));
        // Variance around the shift: <(H - Es)^2> - (E - Es)^2.
        const double e2  = expval_sq(c, theta.as<T>());
        const double dE  = sol.E - eshift * L;
        const double var = e2 - dE * dE;
        const auto   t0  = clock::now();
        std::vector<double> S;
        if(right && c + 1 < L) {
            move_right(c, theta.as<T>());
            t_svd += std::chrono::duration<double>(clock::now() - t0).count();
            env_left(c);
        } else if(!right && c > 0) {
            // theta as a x (d b): X[a, (s, b)] = theta[s + d a + d a_dim b]; B_c = Vt, M_{c-1} <- M_{c-1} (U S)
            theta2.reserve(static_cast<size_t>(n) * es());
            remap2d(DT, a, d * b, theta.p, plain(d), fused(1, d * a, d), theta2.p, plain(1), plain(a), nullptr, 0, s);
            svd.compute<T>(static_cast<int>(a), static_cast<int>(d * b), theta2.as<T>(), S, SVD_JACOBI);
            const int k = svd.k();
            sdev.reserve(S.size() * sizeof(double));
            DMRG_HIP_CHECK(hipMemcpyAsync(sdev.p, S.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
            M[c].reserve(static_cast<size_t>(d * k * b) * es());
            remap2d(DT, k, d * b, svd.Vt(), plain(1), plain(k), M[c].p, plain(d), fused(1, d * k, d), nullptr, 0, s);
            tmp.reserve(static_cast<size_t>(a * k) * es());
            remap2d(DT, a, k, svd.U(), plain(1), plain(a), tmp.p, plain(1), plain(a), sdev.as<double>(), 2, s); // U S
            const int64_t z = chi[c - 1];
            theta2.reserve(static_cast<size_t>(d * z * k) * es());
            GemmDesc g;
            g.dtype = DT;
            g.M = static_cast<int32_t>(d * z), g.N = k, g.K = static_cast<int32_t>(a), g.batch = 1;
            g.A = M[c - 1].p, g.am = plain(1), g.ak = plain(d * z);
            g.B = tmp.p, g.bk = plain(1), g.bn = plain(a);
            g.C = theta2.p, g.cm = plain(1), g.cn = plain(d * z);
            gemm(g, s);
            M[c - 1].reserve(static_cast<size_t>(d * z * k) * es());
            DMRG_HIP_CHECK(hipMemcpyAsync(M[c - 1].p, theta2.p, static_cast<size_t>(d * z * k) * es(), hipMemcpyDeviceToDevice, s));
            chi[c]     = k;
            entropy[c] = entropy_of(S, k);
            t_svd += std::chrono::duration<double>(clock::now() - t0).count();
            env_right(c);
        }
        return {sol.E, var};
    }

    void run() {
        const auto t0 = clock::now();
        DMRG_REQUIRE(L >= 2 && L <= DMRG_FDMRG_MAX_L, DMRG_ERR_ARG, "L out of range");
        DMRG_REQUIRE(cfg.chi_max >= 1, DMRG_ERR_ARG, "chi_max must be positive");
        entropy.assign(L + 1, 0.0);
        terr.assign(L + 1, 0.0);
        folded = cfg.xdmrg != 0;
        // xDMRG: the MPO^2 carries the fixed target shift from the start (set_energy_shift_mpo, xdmrg.cpp:487-501).
        eshift = folded ? cfg.energy_target / L : 0.0;
        build_mpos(eshift);
        init_state();
        // expand_mode: EnvExpandMode bits (config/enums.h:224-234). With NSITE | FORWARD the bonds grow by
        // subspace expansion in the single-site sweeps (the reference's default); otherwise two-site sweeps grow them.
        const bool expand = cfg.expand_mode != 0;
        if(expand) {
            DMRG_REQUIRE((cfg.expand_mode & ~(1 | 2 | 8 | 32)) == 0 && (cfg.expand_mode & 8) && (cfg.expand_mode & 32) &&
                             (cfg.expand_mode & 3),
                         DMRG_ERR_UNSUPPORTED, "expand_mode: H1 and/or H2 with NSITE | FORWARD (SSITE and BACKWARD are not built)");
        }
        // Two-site sweeps: grow bonds and converge the energy.
        double E = 0, Eprev = 1e300;
        int    sweeps = 0;
        const int two_max = expand ? 0 : (cfg.two_site_sweeps > 0 ? cfg.two_site_sweeps : 10);
        for(int sw = 0; sw < two_max; ++sw) {
            for(int i = 0; i + 1 < L; ++i) E = two_site(i, true);
            for(int i = L - 2; i >= 0; --i) E = two_site(i, false);
            ++sweeps;
            if(std::fabs(E - Eprev) <= 1e-12 * std::max(1.0, std::fabs(E))) break;
            Eprev = E;
        }
        // Single-site sweeps with H^2 shifted by the current energy (by the target in xDMRG); stop at
        // variance < var_tol.
        if(!folded) {
            eshift = E / L;
            build_mpos(eshift);
            // Rebuild the H^2 envs for the current (centre at site 0, right-normalised) state.
            for(int i = L - 1; i >= 1; --i) env_right(i);
        }
        double var = 1e300;
        const int max_sw = cfg.max_sweeps > 0 ? cfg.max_sweeps : 20;
        int       one_sw = 0;
        for(; one_sw < max_sw; ++one_sw) {
            double vmax = 0;
            for(int c = 0; c < L; ++c) {
                auto [e, v] = one_site(c, true);
                E           = e;
                var         = v;
                vmax        = std::max(vmax, v);
            }
            for(int c = L - 1; c >= 0; --c) {
                auto [e, v] = one_site(c, false);
                E           = e;
                var         = v;
            }
            ++sweeps;
            if(var < cfg.var_tol) break;
            if(expand && !folded && std::fabs(E / L - eshift) > 1e-10 * std::max(1.0, std::fabs(E / L))) {
                // The bonds grew from a product state: move the H^2 shift to the current energy (the centre is on
                // site 0, every other site right-normalised) so that the variance needs no large cancellation.
                eshift = E / L;
                build_mpos(eshift);
                for(int i = L - 1; i >= 1; --i) env_right(i);
            }
        }
        if(expand) {
            // In the expansion sweeps a bond's Schmidt values are recorded when the centre crosses it, and every
            // later expansion and solve of the half-sweep still changes the state. One pass of the centre from
            // site 0 (where the left-moving half-sweep ended) to the right end, without solves, records the
            // Schmidt values of the final state at every bond; the state itself is unchanged.
            const auto t1 = clock::now();
            for(int c = 0; c + 1 < L; ++c) move_right(c, M[c].as<T>());
            t_svd += std::chrono::duration<double>(clock::now() - t1).count();
        }
        res->energy    = E;
        res->variance  = var;
        res->converged = var < cfg.var_tol ? 1 : 0;
        res->sweeps    = sweeps;
        for(int b = 0; b <= L; ++b) {
            res->entropies[b] = entropy[b];
            res->chi[b]       = chi[b];
            res->trunc_error_max = std::max(res->trunc_error_max, terr[b]);
        }
        res->gflop   = ((hop ? hop->flops : 0.0) + (h2op ? h2op->flops : 0.0) + env_flop) * 1e-9;
        res->t_eigs  = t_eigs;
        res->t_env   = t_env;
        res->t_svd   = t_svd;
        res->t_expand = t_expand;
        res->seconds = std::chrono::duration<double>(clock::now() - t0).count();
    }
};

} // namespace

void fdmrg_ising(const dmrg_fdmrg_config &cfg, dmrg_fdmrg_result &res, hipStream_t s, EigService *svc) {
    res = dmrg_fdmrg_result{};
    if(cfg.complex_) {
        Fdmrg<cpx<double>> f(cfg, s, &res);
        f.svc = svc;
        f.run();
    } else {
        Fdmrg<double> f(cfg, s, &res);
        f.svc = svc;
        f.run();
    }
}

} // namespace dmrg

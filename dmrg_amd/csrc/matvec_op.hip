// MatVecOp: device-resident H_eff operator (MatVecMPOS<T>, matvec_mpos.cpp:40-130, 963-984, 1433-1472).
#include "matvec_op.hpp"
#include "vecops.hpp"

#include <cstring>

namespace dmrg {

using cx = std::complex<double>;

// ------------------------------------------------------------------------------------------ memory pools

DevPool::~DevPool() {
    for(auto &b : all_) (void) hipFree(b.first);
}

size_t DevPool::cls(size_t b) {
    for(int k = 12; k < 62; ++k) {
        const size_t p2 = size_t(1) << k;
        if(p2 >= b) return p2;
        if(p2 + p2 / 2 >= b) return p2 + p2 / 2;
    }
    return b;
}

void *DevPool::get(size_t b, size_t &c) {
    c = cls(b);
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto                        it = free_.find(c);
        if(it != free_.end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            return p;
        }
    }
    void *p = nullptr;
    if(hipMalloc(&p, c) != hipSuccess) {
        (void) hipGetLastError();
        throw Error(DMRG_ERR_OOM, "device allocation of " + std::to_string(c) + " bytes failed");
    }
    std::lock_guard<std::mutex> lk(mu_);
    all_.emplace_back(p, c);
    return p;
}

void DevPool::put(void *p, size_t c) {
    if(!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    free_[c].push_back(p);
}

size_t DevPool::held_bytes() const {
    std::lock_guard<std::mutex> lk(mu_);
    size_t                      t = 0;
    for(const auto &b : all_) t += b.second;
    return t;
}

std::shared_ptr<DevPool> &thread_dev_pool() {
    static thread_local std::shared_ptr<DevPool> pool;
    return pool;
}

DevPoolScope::DevPoolScope() : prev_(thread_dev_pool()) { thread_dev_pool() = std::make_shared<DevPool>(); }
DevPoolScope::~DevPoolScope() { thread_dev_pool() = std::move(prev_); }

namespace {
struct PinnedCache {
    std::vector<std::pair<void *, size_t>> free_;
    ~PinnedCache() {
        for(auto &b : free_) (void) hipHostFree(b.first);
    }
};
PinnedCache &pinned_cache() {
    static thread_local PinnedCache c;
    return c;
}
} // namespace

void *pinned_get(size_t bytes) {
    auto &f = pinned_cache().free_;
    for(size_t i = 0; i < f.size(); ++i)
        if(f[i].second >= bytes) {
            void *p = f[i].first;
            f.erase(f.begin() + static_cast<std::ptrdiff_t>(i));
            return p;
        }
    void *p = nullptr;
    DMRG_HIP_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    return p;
}

void pinned_put(void *p, size_t bytes) {
    if(p) pinned_cache().free_.emplace_back(p, bytes);
}

std::vector<cx> merge_mpos(int nmpo, const cx *const *mpos, const int64_t *dims, int64_t out[4]) {
    DMRG_REQUIRE(nmpo >= 1 && mpos && dims, DMRG_ERR_ARG, "need at least one MPO");
    for(int i = 0; i < nmpo; ++i) {
        DMRG_REQUIRE(mpos[i], DMRG_ERR_ARG, "null MPO");
        DMRG_REQUIRE(dims[4 * i] > 0 && dims[4 * i + 1] > 0 && dims[4 * i + 2] > 0, DMRG_ERR_DIMS, "zero-sized MPO");
        DMRG_REQUIRE(dims[4 * i + 2] == dims[4 * i + 3], DMRG_ERR_DIMS, "MPO has different spin dimensions up and down");
        if(i > 0) DMRG_REQUIRE(dims[4 * i] == dims[4 * (i - 1) + 1], DMRG_ERR_DIMS, "MPO chain bond mismatch");
    }
    int64_t         DL = dims[0], DR = dims[1], d = dims[2];
    std::vector<cx> cur(mpos[0], mpos[0] + DL * DR * d * d);
    for(int i = 1; i < nmpo; ++i) {
        const int64_t DRn = dims[4 * i + 1], dn = dims[4 * i + 2], Dm = DR;
        const cx     *B   = mpos[i];
        const int64_t dd  = d * dn;
        std::vector<cx> nxt(static_cast<size_t>(DL * DRn * dd * dd), cx(0.0, 0.0));
        // new[wL,wR,(s1,s2),(s1',s2')] = sum_w A[wL,w,s1,s1'] B[w,wR,s2,s2'], s1 fastest in the fused index.
        for(int64_t s2p = 0; s2p < dn; ++s2p)
            for(int64_t s1p = 0; s1p < d; ++s1p)
                for(int64_t s2 = 0; s2 < dn; ++s2)
                    for(int64_t s1 = 0; s1 < d; ++s1)
                        for(int64_t wR = 0; wR < DRn; ++wR)
                            for(int64_t wL = 0; wL < DL; ++wL) {
                                cx acc(0.0, 0.0);
                                for(int64_t w = 0; w < Dm; ++w)
                                    acc += cur[wL + DL * (w + Dm * (s1 + d * s1p))] * B[w + Dm * (wR + DRn * (s2 + dn * s2p))];
                                nxt[wL + DL * (wR + DRn * ((s1 + d * s2) + dd * (s1p + d * s2p)))] = acc;
                            }
        cur = std::move(nxt);
        DR  = DRn;
        d   = dd;
    }
    out[0] = DL, out[1] = DR, out[2] = d, out[3] = d;
    return cur;
}

template <typename T>
static void upload_as(DevBuf &buf, const cx *src, int64_t n, dmrg_dtype t) {
    const size_t es = dtype_size(t);
    buf.reserve(static_cast<size_t>(n) * es);
    if(t == DMRG_CX64) {
        DMRG_HIP_CHECK(hipMemcpy(buf.p, src, static_cast<size_t>(n) * es, hipMemcpyHostToDevice));
    } else {
        std::vector<double> re(static_cast<size_t>(n));
        for(int64_t i = 0; i < n; ++i) re[i] = src[i].real();
        DMRG_HIP_CHECK(hipMemcpy(buf.p, re.data(), static_cast<size_t>(n) * es, hipMemcpyHostToDevice));
    }
}

MatVecOp::MatVecOp(dmrg_dtype t, int nmpo, const void *const *mpos, const int64_t *mpo_dims, const void *envL_h,
                   const int64_t envL_dims[3], const void *envR_h, const int64_t envR_dims[3])
    : dtype(t), nmpo_in(nmpo) {
    DMRG_REQUIRE(envL_h && envR_h && envL_dims && envR_dims, DMRG_ERR_ARG, "null env");
    load_mpos(nmpo, mpos, mpo_dims, envL_dims, envR_dims);
    upload_mpo();
    upload_as<double>(envL, static_cast<const cx *>(envL_h), D.chiL * D.chiL * D.DL, t);
    upload_as<double>(envR, static_cast<const cx *>(envR_h), D.chiR * D.chiR * D.DR, t);
    envL_p = envL.p;
    envR_p = envR.p;
}

MatVecOp::MatVecOp(dmrg_dtype t, int nmpo, const void *const *mpos, const int64_t *mpo_dims, const void *envL_d,
                   const int64_t envL_dims[3], const void *envR_d, const int64_t envR_dims[3], bool)
    : dtype(t), nmpo_in(nmpo) {
    rebind(nmpo, mpos, mpo_dims, envL_d, envL_dims, envR_d, envR_dims);
}

void MatVecOp::rebind(int nmpo, const void *const *mpos, const int64_t *mpo_dims, const void *envL_d,
                      const int64_t envL_dims[3], const void *envR_d, const int64_t envR_dims[3]) {
    DMRG_REQUIRE(envL_d && envR_d && envL_dims && envR_dims, DMRG_ERR_ARG, "null env");
    load_mpos(nmpo, mpos, mpo_dims, envL_dims, envR_dims);
    nmpo_in     = nmpo;
    sigma       = cx(0.0, 0.0);
    ready_shift = false;
    upload_mpo();
    envL_p = envL_d;
    envR_p = envR_d;
}

void MatVecOp::load_mpos(int nmpo, const void *const *mpos, const int64_t *mpo_dims, const int64_t envL_dims[3],
                         const int64_t envR_dims[3]) {
    int64_t md[4];
    mpo_host = merge_mpos(nmpo, reinterpret_cast<const cx *const *>(mpos), mpo_dims, md);
    const int64_t mps_dims[3] = {md[2], envL_dims[0], envR_dims[0]};
    D                         = check_matvec_dims(mps_dims, md, envL_dims, envR_dims);
    // MatVecMPOS<fp64> requires real MPOs (matvec_mpos.cpp:59); env blocks are cast with .real().
    auto require_real = [&](const cx *v, size_t n) {
        if(dtype != DMRG_FP64) return;
        for(size_t i = 0; i < n; ++i) DMRG_REQUIRE(v[i].imag() == 0.0, DMRG_ERR_ARG, "mpo is not real");
    };
    require_real(mpo_host.data(), mpo_host.size());
    chain = nmpo > 5; // matvec_mpos.cpp:104-123: at most 5 MPOs are merged, longer lists applied in sequence
    chain_host.clear();
    chain_dims.clear();
    if(chain) {
        for(int i = 0; i < nmpo; ++i) {
            const int64_t *dm = mpo_dims + 4 * i;
            const cx      *src = static_cast<const cx *>(mpos[i]);
            chain_host.emplace_back(src, src + dm[0] * dm[1] * dm[2] * dm[3]);
            chain_dims.insert(chain_dims.end(), dm, dm + 4);
            require_real(src, chain_host.back().size());
        }
        chain_dev.resize(nmpo);
        chain_dev32.resize(nmpo);
    }
}

double MatVecOp::macs() const {
    if(!chain) return D.macs();
    const auto v = chain_view(false);
    return chain_macs(D, static_cast<int>(v.size()), v.data());
}

std::vector<ChainMpo> MatVecOp::chain_view(bool fp32) const {
    std::vector<ChainMpo> v(chain_host.size());
    for(size_t i = 0; i < v.size(); ++i) {
        v[i].W  = fp32 ? chain_dev32[i].p : chain_dev[i].p;
        v[i].DL = chain_dims[4 * i], v[i].DR = chain_dims[4 * i + 1], v[i].d = chain_dims[4 * i + 2];
    }
    return v;
}

void MatVecOp::upload_mpo() {
    if(upload_stream) {
        // Stream-ordered upload through a pinned staging block: the host never waits for the copy, only (rarely)
        // for the previous upload's copy to have left the staging block before it is overwritten.
        const size_t es    = dtype_size(dtype);
        size_t       total = mpo_host.size();
        for(const auto &c : chain_host) total += c.size();
        total *= es;
        if(staging_ev) DMRG_HIP_CHECK(hipEventSynchronize(staging_ev));
        else DMRG_HIP_CHECK(hipEventCreateWithFlags(&staging_ev, hipEventDisableTiming));
        if(staging_bytes < total) {
            if(staging) pinned_put(staging, staging_bytes);
            staging       = pinned_get(total);
            staging_bytes = total;
        }
        char *st = static_cast<char *>(staging);
        auto  put = [&](DevBuf &dst, const std::vector<cx> &src) {
            const size_t bytes = src.size() * es;
            if(dtype == DMRG_CX64) std::memcpy(st, src.data(), bytes);
            else
                for(size_t i = 0; i < src.size(); ++i) reinterpret_cast<double *>(st)[i] = src[i].real();
            dst.reserve(bytes);
            DMRG_HIP_CHECK(hipMemcpyAsync(dst.p, st, bytes, hipMemcpyHostToDevice, upload_stream));
            st += bytes;
        };
        put(mpo, mpo_host);
        for(size_t i = 0; i < chain_host.size(); ++i) put(chain_dev[i], chain_host[i]);
        DMRG_HIP_CHECK(hipEventRecord(staging_ev, upload_stream));
    } else {
        upload_as<double>(mpo, mpo_host.data(), static_cast<int64_t>(mpo_host.size()), dtype);
        for(size_t i = 0; i < chain_host.size(); ++i)
            upload_as<double>(chain_dev[i], chain_host[i].data(), static_cast<int64_t>(chain_host[i].size()), dtype);
    }
    have32     = false;
    diag_valid = false;
    pc_ready   = false;
}

void MatVecOp::ensure_fp32(hipStream_t s) {
    if(have32) return;
    const int64_t cf = dtype_is_complex(dtype) ? 2 : 1;
    const int64_t nm = D.DL * D.DR * D.d * D.d * cf, nl = D.chiL * D.chiL * D.DL * cf, nr = D.chiR * D.chiR * D.DR * cf;
    mpo32.reserve(nm * sizeof(float));
    envL32.reserve(nl * sizeof(float));
    envR32.reserve(nr * sizeof(float));
    convert_f64_to_f32(nm, mpo.as<double>(), mpo32.as<float>(), s);
    convert_f64_to_f32(nl, static_cast<const double *>(envL_p), envL32.as<float>(), s);
    convert_f64_to_f32(nr, static_cast<const double *>(envR_p), envR32.as<float>(), s);
    for(size_t i = 0; i < chain_host.size(); ++i) {
        const int64_t nc = static_cast<int64_t>(chain_host[i].size()) * cf;
        chain_dev32[i].reserve(nc * sizeof(float));
        convert_f64_to_f32(nc, chain_dev[i].as<double>(), chain_dev32[i].as<float>(), s);
    }
    have32 = true;
}

void MatVecOp::multax_device(const void *dx, int64_t ldx, void *dy, int64_t ldy, int nvec, hipStream_t s, bool fp32) {
    if(nvec == 0) return;
    const int64_t n = D.n();
    DMRG_REQUIRE(ldx >= n && ldy >= n, DMRG_ERR_ARG, "leading dimension smaller than rows()");
    if(fp32) ensure_fp32(s);
    const dmrg_dtype t  = fp32 ? dtype_single_of(dtype) : dtype;
    DevBuf          &wb = fp32 ? ws32 : ws;
    if(chain) {
        const auto v = chain_view(fp32);
        void      *w = wb.reserve(static_cast<size_t>(nvec) * chain_ws_elems(D, static_cast<int>(v.size()), v.data()) * dtype_size(t));
        matvec_chain(t, D, static_cast<int>(v.size()), v.data(), dx, ldx, dy, ldy, nvec, fp32 ? envL32.p : envL_p,
                     fp32 ? envR32.p : envR_p, w, s);
    } else {
        void *w = wb.reserve(static_cast<size_t>(nvec) * matvec_ws_elems(D) * dtype_size(t));
        matvec(t, D, dx, ldx, dy, ldy, nvec, fp32 ? mpo32.p : mpo.p, fp32 ? envL32.p : envL_p, fp32 ? envR32.p : envR_p, w, s);
    }
    num_mv += nvec;
    flops += static_cast<double>(nvec) * macs() * (dtype_is_complex(dtype) ? 8.0 : 2.0);
}

void MatVecOp::multax_host(const void *x, int64_t ldx, void *y, int64_t ldy, int nvec) {
    if(nvec == 0) return;
    const int64_t n  = D.n();
    const size_t  es = dtype_size(dtype);
    DMRG_REQUIRE(ldx >= n && ldy >= n, DMRG_ERR_ARG, "leading dimension smaller than rows()");
    xbuf.reserve(static_cast<size_t>(nvec) * n * es);
    ybuf.reserve(static_cast<size_t>(nvec) * n * es);
    DMRG_HIP_CHECK(hipMemcpy2D(xbuf.p, n * es, x, ldx * es, n * es, nvec, hipMemcpyHostToDevice));
    multax_device(xbuf.p, n, ybuf.p, n, nvec, nullptr, false);
    DMRG_HIP_CHECK(hipMemcpy2D(y, ldy * es, ybuf.p, n * es, n * es, nvec, hipMemcpyDeviceToHost));
}

void MatVecOp::set_shift(cx shift) {
    // matvec_mpos.cpp:1437-1471: applied once (readyShift); (sigma - shift) / N_mpo times the identity is added to
    // the bottom-left corner W[DL-1, 0] of each of the N_mpo MPOs the operator holds: the one merged MPO of a list
    // of up to 5, or every MPO of a longer chain.
    if(ready_shift) return;
    if(sigma == shift) return;
    auto add_corner = [&](std::vector<cx> &W, int64_t DL, int64_t DR, int64_t d, cx delta) {
        for(int64_t s = 0; s < d; ++s) {
            cx &v = W[(DL - 1) + DL * (0 + DR * (s + d * s))];
            v += (dtype == DMRG_FP64) ? cx(delta.real(), 0.0) : delta;
        }
    };
    if(chain) {
        const int n     = static_cast<int>(chain_host.size());
        const cx  delta = (sigma - shift) / static_cast<double>(n);
        std::vector<const void *> ptrs(n);
        for(int i = 0; i < n; ++i) {
            add_corner(chain_host[i], chain_dims[4 * i], chain_dims[4 * i + 1], chain_dims[4 * i + 2], delta);
            ptrs[i] = chain_host[i].data();
        }
        int64_t md[4];
        mpo_host = merge_mpos(n, reinterpret_cast<const cx *const *>(ptrs.data()), chain_dims.data(), md);
    } else {
        add_corner(mpo_host, D.DL, D.DR, D.d, sigma - shift);
    }
    sigma = shift;
    upload_mpo();
    ready_shift = true;
    op_ready    = false; // a factorization of the unshifted operator no longer applies
}

void MatVecOp::diagonal(void *out, hipStream_t s) {
    void *w = diagws.reserve(static_cast<size_t>(D.chiL * D.DR * D.d) * dtype_size(dtype));
    heff_diagonal(dtype, D, mpo.p, envL_p, envR_p, w, out, s);
}

void MatVecOp::precond_host(const void *x, int64_t ldx, void *y, int64_t ldy, int nvec, const double *shifts) {
    // MultPc with jcbMaxBlockSize = 1 (matvec_mpos.cpp:1333-1361): out = in / (diag(A) - shift), no floor.
    if(nvec == 0) return;
    const int64_t n  = D.n();
    const size_t  es = dtype_size(dtype);
    DMRG_REQUIRE(ldx >= n && ldy >= n, DMRG_ERR_ARG, "leading dimension smaller than rows()");
    if(jcb > 1) {
        xbuf.reserve(static_cast<size_t>(nvec) * n * es);
        ybuf.reserve(static_cast<size_t>(nvec) * n * es);
        DMRG_HIP_CHECK(hipMemcpy2D(xbuf.p, n * es, x, ldx * es, n * es, nvec, hipMemcpyHostToDevice));
        if(!pc_ready) block_pc_build(shifts ? shifts[0] : 0.0, nullptr);
        for(int i = 0; i < nvec; ++i)
            block_pc_apply(xbuf.as<char>() + static_cast<size_t>(i) * n * es, ybuf.as<char>() + static_cast<size_t>(i) * n * es, nullptr);
        DMRG_HIP_CHECK(hipMemcpy2D(y, ldy * es, ybuf.p, n * es, n * es, nvec, hipMemcpyDeviceToHost));
        return;
    }
    if(!diag_valid) {
        diagc.reserve(static_cast<size_t>(n) * es);
        diagonal(diagc.p, nullptr);
        diag_valid = true;
    }
    xbuf.reserve(static_cast<size_t>(nvec) * n * es);
    ybuf.reserve(static_cast<size_t>(nvec) * n * es);
    DMRG_HIP_CHECK(hipMemcpy2D(xbuf.p, n * es, x, ldx * es, n * es, nvec, hipMemcpyHostToDevice));
    for(int i = 0; i < nvec; ++i) {
        const size_t off = static_cast<size_t>(i) * n * es;
        if(dtype == DMRG_FP64)
            jacobi_precond<double, double>(n, reinterpret_cast<const double *>(xbuf.as<char>() + off), diagc.as<double>(),
                                           shifts ? shifts[i] : 0.0, 0.0, reinterpret_cast<double *>(ybuf.as<char>() + off),
                                           nullptr);
        else
            jacobi_precond<cpx<double>, cpx<double>>(n, reinterpret_cast<const cpx<double> *>(xbuf.as<char>() + off),
                                                     diagc.as<cpx<double>>(), shifts ? shifts[i] : 0.0, 0.0,
                                                     reinterpret_cast<cpx<double> *>(ybuf.as<char>() + off), nullptr);
    }
    DMRG_HIP_CHECK(hipMemcpy2D(y, ldy * es, ybuf.p, n * es, n * es, nvec, hipMemcpyDeviceToHost));
}

void MatVecOp::time_legs(const void *dx, void *dy, int nrep, bool fp32, float ms[4], hipStream_t s) {
    DMRG_REQUIRE(!chain, DMRG_ERR_UNSUPPORTED, "leg timing of an operator with an MPO chain");
    hipEvent_t ev[5];
    for(auto &e : ev) DMRG_HIP_CHECK(hipEventCreate(&e));
    double acc[4] = {0, 0, 0, 0};
    try {
        if(fp32) ensure_fp32(s);
        const dmrg_dtype t  = fp32 ? dtype_single_of(dtype) : dtype;
        DevBuf          &wb = fp32 ? ws32 : ws;
        void            *w  = wb.reserve(static_cast<size_t>(matvec_ws_elems(D)) * dtype_size(t));
        for(int r = 0; r < nrep; ++r) {
            matvec(t, D, dx, D.n(), dy, D.n(), 1, fp32 ? mpo32.p : mpo.p, fp32 ? envL32.p : envL_p,
                   fp32 ? envR32.p : envR_p, w, s, ev);
            DMRG_HIP_CHECK(hipEventSynchronize(ev[4]));
            for(int i = 0; i < 4; ++i) {
                float t_ms = 0;
                DMRG_HIP_CHECK(hipEventElapsedTime(&t_ms, ev[i], ev[i + 1]));
                acc[i] += t_ms;
            }
        }
        num_mv += nrep;
    } catch(...) {
        for(auto &e : ev) (void) hipEventDestroy(e);
        throw;
    }
    for(auto &e : ev) (void) hipEventDestroy(e);
    for(int i = 0; i < 4; ++i) ms[i] = static_cast<float>(acc[i] / nrep);
}

} // namespace dmrg

// Request loop for independent DMRG solves on one GPU (continuous batching at the stream level).
//
// Requests (disorder realisations of ising_tf_rf with their own L, chi_max and fields) arrive at given
// times. Worker threads, each owning a HIP stream (default priority), admit the first pending request
// that has arrived and whose HBM footprint fits in the remaining budget, run its sweep on their stream
// and release the footprint when done. Requests that arrive while others run are admitted as soon as a
// worker and memory free up; the kernels of concurrent requests overlap on the device. The reference's
// scale-out is one OS process per disorder seed (slurm/batches_fdmrg.py:8-27); here they share the GPU.
#include "sweep.hpp"

#include "batch_eigs.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace dmrg {

// HBM estimate of one request (bytes): MPS, H and H^2 env blocks for every site, and the largest
// two-site problem's eigensolver basis (V, W, scratch) and contraction workspaces.
int64_t request_footprint(int L, int chi, bool cplx) {
    const double es   = cplx ? 16.0 : 8.0;
    const double c2   = static_cast<double>(chi) * chi;
    const double n2   = 4.0 * c2;                                  // two-site problem size (d = 2)
    const int    mmax = static_cast<int>(std::ceil(std::log2(std::max(2.0, n2)))) + 4;
    const double mps  = L * 2.0 * c2;
    const double envs = L * c2 * (3 + 9) * 2.0;                    // (H, H^2) x (L, R), D = 3, D^2 = 9
    const double eig  = n2 * (3.0 * mmax + 4);                     // V, W, restart tmp, vectors
    const double ws   = n2 * (3 + 3 + 8) + c2 * 9 * 2 * (8 + 2);   // matvec / env workspaces incl. split-K
    return static_cast<int64_t>(es * (mps + envs + eig + ws)) + (8ll << 20);
}

void serve(const dmrg_serve_config &cfg, const dmrg_request *req, int nreq, dmrg_request_result *out, dmrg_serve_stats &st) {
    using clock = std::chrono::steady_clock;
    const int     W      = std::max(1, cfg.workers);
    const int64_t budget = cfg.hbm_budget_bytes > 0 ? cfg.hbm_budget_bytes : (int64_t(64) << 30);
    std::vector<int64_t> fp(nreq);
    for(int i = 0; i < nreq; ++i) {
        fp[i] = request_footprint(req[i].L, req[i].chi_max, cfg.complex_ != 0);
        DMRG_REQUIRE(fp[i] <= budget, DMRG_ERR_OOM, "request " + std::to_string(i) + " does not fit in the HBM budget");
        DMRG_REQUIRE(req[i].h_rand && req[i].L >= 2 && req[i].L <= DMRG_FDMRG_MAX_L, DMRG_ERR_ARG, "bad request");
    }
    std::vector<int> order(nreq);
    for(int i = 0; i < nreq; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });

    // Shared eigensolve batcher: the energy solves of all running requests advance together.
    std::unique_ptr<EigService> svc;
    if(cfg.batch_eigs) {
        EigsParams p;
        p.target      = TARGET_SA;
        p.tol         = cfg.eigs_tol > 0 ? cfg.eigs_tol : 1e-12;
        p.max_matvecs = 20000;
        svc           = std::make_unique<EigService>(cfg.complex_ ? DMRG_CX64 : DMRG_FP64, p);
    }

    std::mutex              mu;
    std::condition_variable cv;
    std::vector<char>       taken(nreq, 0);
    int64_t                 admitted = 0, peak = 0;
    int                     running = 0, peak_running = 0, done = 0;
    const auto              t0      = clock::now();
    auto                    now_s   = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };

    auto worker = [&](int wid) {
        // All of a worker's device work goes to its stream, so its buffers are recycled in a pool of its own:
        // no hipFree (a device-wide synchronisation that would stall every other worker) while requests run.
        DevPoolScope pool_scope;
        hipStream_t  s = nullptr;
        if(hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
        while(true) {
            int idx = -1;
            {
                std::unique_lock<std::mutex> lock(mu);
                while(true) {
                    const double t       = now_s();
                    double       next_at = 1e300;
                    bool         pending = false;
                    for(int k : order) { // arrival order; skip over arrived requests that do not fit
                        if(taken[k]) continue;
                        pending = true;
                        if(req[k].arrival_s > t) {
                            next_at = req[k].arrival_s;
                            break;
                        }
                        if(admitted + fp[k] <= budget) {
                            idx = k;
                            break;
                        }
                    }
                    if(idx >= 0 || !pending) break;
                    if(next_at < 1e300) cv.wait_for(lock, std::chrono::duration<double>(std::max(0.0, next_at - t)));
                    else cv.wait(lock); // woken when a running request releases memory
                }
                if(idx < 0) break;
                taken[idx] = 1;
                admitted += fp[idx];
                ++running;
                peak         = std::max(peak, admitted);
                peak_running = std::max(peak_running, running);
            }
            dmrg_request_result &r = out[idx];
            r                      = dmrg_request_result{};
            r.arrival_s            = req[idx].arrival_s;
            r.start_s              = now_s();
            r.worker               = wid;
            r.footprint_bytes      = fp[idx];
            dmrg_fdmrg_config c{};
            c.L               = req[idx].L;
            c.J               = cfg.J;
            c.h_tran          = cfg.h_tran;
            c.h_rand          = req[idx].h_rand;
            c.chi_max         = req[idx].chi_max;
            c.max_sweeps      = cfg.max_sweeps > 0 ? cfg.max_sweeps : 20;
            c.two_site_sweeps = cfg.two_site_sweeps > 0 ? cfg.two_site_sweeps : 10;
            c.var_tol         = cfg.var_tol > 0 ? cfg.var_tol : 1e-12;
            c.trunc_lim       = 1e-14;
            c.eigs_tol        = cfg.eigs_tol > 0 ? cfg.eigs_tol : 1e-12;
            c.mixed           = cfg.mixed;
            c.complex_        = cfg.complex_;
            c.seed            = req[idx].seed;
            c.expand_mode     = cfg.expand_mode;
            try {
                dmrg_fdmrg_result fr;
                fdmrg_ising(c, fr, s, svc.get());
                r.energy      = fr.energy;
                r.variance    = fr.variance;
                r.entropy_mid = fr.entropies[c.L / 2];
                r.sweeps      = fr.sweeps;
                r.matvecs     = fr.matvecs + fr.matvecs_fp32;
                r.gflop       = fr.gflop;
                r.status      = fr.converged ? DMRG_OK : DMRG_ERR_NOT_CONVERGED;
            } catch(const Error &e) {
                r.status = e.status;
            } catch(...) {
                r.status = DMRG_ERR_ARG;
            }
            r.finish_s  = now_s();
            r.latency_s = r.finish_s - r.arrival_s;
            {
                std::lock_guard<std::mutex> lock(mu);
                admitted -= fp[idx];
                --running;
                ++done;
            }
            cv.notify_all();
        }
        if(s) (void) hipStreamDestroy(s);
        cv.notify_all();
    };

    std::vector<std::thread> threads;
    for(int w = 0; w < W; ++w) threads.emplace_back(worker, w);
    for(auto &t : threads) t.join();

    st               = dmrg_serve_stats{};
    st.makespan_s    = now_s();
    std::vector<double> lat;
    for(int i = 0; i < nreq; ++i) {
        if(out[i].status == DMRG_OK) {
            ++st.completed;
            lat.push_back(out[i].latency_s);
        } else {
            ++st.failed;
        }
        st.matvecs += out[i].matvecs;
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        if(lat.empty()) return 0.0;
        const double pos = q * (lat.size() - 1);
        const size_t lo  = static_cast<size_t>(std::floor(pos)), hi = std::min(lat.size() - 1, lo + 1);
        return lat[lo] + (pos - lo) * (lat[hi] - lat[lo]);
    };
    st.p50_latency_s       = pct(0.50);
    st.p99_latency_s       = pct(0.99);
    double sum             = 0;
    for(double v : lat) sum += v;
    st.mean_latency_s      = lat.empty() ? 0 : sum / lat.size();
    st.solves_per_s        = st.makespan_s > 0 ? st.completed / st.makespan_s : 0;
    st.peak_admitted_bytes = peak;
    st.peak_concurrency    = peak_running;
    if(svc) {
        const EigService::Stats bs = svc->stats();
        st.batch_steps             = bs.steps;
        st.peak_batch              = bs.peak_batch;
        st.batch_step_s            = bs.t_step;
        st.batch_rebuild_s         = bs.t_rebuild;
        st.batch_rr_s              = bs.t_rr;
    }
}

} // namespace dmrg

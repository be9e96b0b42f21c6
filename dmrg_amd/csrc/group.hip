// Grouped H_eff matvecs (see group.hpp). The three legs of every operator's matvec are the GEMMs matvec()
// launches one by one; here leg j of all k operators goes into one grouped launch, so k small problems that
// each fill a few CUs share the whole GPU in 3 launches (plus at most 3 more when the GEMMs differ in kind).
#include "group.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <string>

namespace dmrg {

MatvecGroup::~MatvecGroup() {
    if(exec) (void) hipGraphExecDestroy(exec);
    if(graph) (void) hipGraphDestroy(graph);
    for(auto &l : legs) grouped_free(l);
}

void MatvecGroup::run(hipStream_t s) {
    if(exec) {
        DMRG_HIP_CHECK(hipGraphLaunch(exec, s));
        return;
    }
    for(auto &l : legs)
        for(auto &g : l) grouped_launch(g, s);
}

namespace {

// Grouped-leg autotuner: the candidate tile configurations of a leg are timed on the group's own operands
// (HIP events, 3 launches after a warm-up) and the fastest is cached by dtype, leg, the leg's 64x64 tile count
// (rounded to a power of two) and its largest M (rounded likewise), so that sets of similar shape reuse it.
std::mutex                 g_gtune_mu;
std::map<std::string, int> g_gtune;

int pow2_bucket(int64_t v) { return v <= 1 ? 0 : static_cast<int>(std::ceil(std::log2(static_cast<double>(v)))); }

TileCfg tuned_leg_cfg(int leg, const std::vector<GemmDesc> &descs, const GroupStorage *storage, TileCfg fallback) {
    if(descs.empty() || !autotune_enabled()) return fallback;
    int64_t tiles = 0;
    int     mmax  = 0;
    for(const auto &g : descs) tiles += grouped_tiles(g, TileCfg::T64x64), mmax = std::max(mmax, g.M);
    const std::string key = std::to_string(static_cast<int>(descs[0].dtype)) + "|" + std::to_string(leg) + "|" +
                            std::to_string(pow2_bucket(tiles)) + "|" + std::to_string(pow2_bucket(mmax));
    {
        std::lock_guard<std::mutex> lk(g_gtune_mu);
        auto                        it = g_gtune.find(key);
        if(it != g_gtune.end()) return static_cast<TileCfg>(it->second);
    }
    hipStream_t ts = nullptr;
    hipEvent_t  e0 = nullptr, e1 = nullptr;
    DMRG_HIP_CHECK(hipStreamCreateWithFlags(&ts, hipStreamNonBlocking));
    TileCfg best = fallback;
    float   best_ms = 1e30f;
    try {
        DMRG_HIP_CHECK(hipEventCreate(&e0));
        DMRG_HIP_CHECK(hipEventCreate(&e1));
        std::vector<TileCfg> cands = {TileCfg::T64x64, TileCfg::T32x32};
        if(mmax <= 32) cands.push_back(TileCfg::T16x128);
        for(TileCfg c : cands) {
            std::vector<GroupedGemm> gg = storage ? grouped_prepare(descs, c, storage->stores[leg], storage->store_bytes[leg])
                                                  : grouped_prepare(descs, c);
            for(auto &g : gg) grouped_launch(g, ts); // warm-up
            DMRG_HIP_CHECK(hipEventRecord(e0, ts));
            for(int r = 0; r < 3; ++r)
                for(auto &g : gg) grouped_launch(g, ts);
            DMRG_HIP_CHECK(hipEventRecord(e1, ts));
            DMRG_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            DMRG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            grouped_free(gg);
            if(ms < best_ms) best_ms = ms, best = c;
        }
    } catch(...) {
        if(e0) (void) hipEventDestroy(e0);
        if(e1) (void) hipEventDestroy(e1);
        (void) hipStreamSynchronize(ts);
        (void) hipStreamDestroy(ts);
        throw;
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    DMRG_HIP_CHECK(hipStreamDestroy(ts));
    std::lock_guard<std::mutex> lk(g_gtune_mu);
    g_gtune[key] = static_cast<int>(best);
    return best;
}

} // namespace

size_t matvec_group_ws_bytes(const MatVecOp &op, bool fp32) {
    const dmrg_dtype t = fp32 ? dtype_single_of(op.dtype) : op.dtype;
    return static_cast<size_t>(matvec_ws_elems(op.D)) * dtype_size(t);
}

MatvecGroup *matvec_group_create(MatVecOp *const *ops, int k, const void *const *xs, void *const *ys, bool fp32, bool capture,
                                 const GroupStorage *storage) {
    DMRG_REQUIRE(k >= 1 && ops && xs && ys, DMRG_ERR_ARG, "matvec group: bad arguments");
    auto                  grp = std::make_unique<MatvecGroup>();
    std::vector<GemmDesc> leg[3];
    grp->ws.resize(k);
    grp->k = k;
    const dmrg_dtype t0 = ops[0]->dtype;
    for(int i = 0; i < k; ++i) {
        MatVecOp &op = *ops[i];
        DMRG_REQUIRE(op.dtype == t0, DMRG_ERR_ARG, "matvec group: operators of one dtype");
        DMRG_REQUIRE(!op.chain, DMRG_ERR_UNSUPPORTED, "matvec group: operators with more than 5 MPOs (applied in sequence)");
        DMRG_REQUIRE(xs[i] && ys[i], DMRG_ERR_ARG, "matvec group: null vector");
        if(fp32) {
            op.ensure_fp32(nullptr);
            DMRG_HIP_CHECK(hipStreamSynchronize(nullptr));
        }
        const dmrg_dtype t  = fp32 ? dtype_single_of(op.dtype) : op.dtype;
        void            *wp = nullptr;
        if(storage && storage->wss) {
            wp = storage->wss[i];
            DMRG_REQUIRE(wp, DMRG_ERR_ARG, "matvec group: null workspace");
        } else {
            wp = grp->ws[i].reserve(matvec_group_ws_bytes(op, fp32));
        }
        GemmDesc g[3];
        matvec_descs(t, op.D, xs[i], op.D.n(), ys[i], op.D.n(), 1, fp32 ? op.mpo32.p : op.mpo.p, fp32 ? op.envL32.p : op.envL_p,
                     fp32 ? op.envR32.p : op.envR_p, wp, g);
        for(int j = 0; j < 3; ++j) leg[j].push_back(g[j]);
        grp->flop += op.D.macs() * (dtype_is_complex(t) ? 8.0 : 2.0);
    }
    // The tuner launches on a stream of its own, reading xs and writing ys and the workspaces: in the standalone
    // path (no caller-owned storage) every stream's work on them must be finished first. The batched engine owns
    // its vectors and synchronises its stream before it rebuilds a group.
    if(!storage && autotune_enabled()) DMRG_HIP_CHECK(hipDeviceSynchronize());
    // Tiles: the MPO leg (M = d D, K = d D) in 16 x 128 tiles while M is small; the big legs in 64 x 64 tiles
    // when that gives about two waves of workgroups on the GPU, else in 32 x 32 tiles.
    for(int j = 0; j < 3; ++j) {
        // Size rule (used when tuning is off): MPO leg with small M in 16x128 tiles, else 64x64 tiles when there
        // are enough of them to fill the chip twice, 32x32 otherwise; with tuning the timed fastest.
        TileCfg cfg;
        if(j == 1) {
            int mmax = 0;
            for(const auto &g : leg[j]) mmax = std::max(mmax, g.M);
            cfg = mmax <= 32 ? TileCfg::T16x128 : TileCfg::T32x32;
        } else {
            int64_t tiles = 0;
            for(const auto &g : leg[j]) tiles += grouped_tiles(g, TileCfg::T64x64);
            cfg = tiles >= 2 * device_cu_count() ? TileCfg::T64x64 : TileCfg::T32x32;
        }
        cfg          = tuned_leg_cfg(j, leg[j], storage, cfg);
        grp->legs[j] = storage ? grouped_prepare(leg[j], cfg, storage->stores[j], storage->store_bytes[j]) : grouped_prepare(leg[j], cfg);
    }
    if(capture) {
        // Captured on a stream of its own at default priority; the graph holds only kernel nodes.
        hipStream_t cs = nullptr;
        DMRG_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        try {
            DMRG_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            for(auto &l : grp->legs)
                for(auto &g : l) grouped_launch(g, cs);
            DMRG_HIP_CHECK(hipStreamEndCapture(cs, &grp->graph));
            DMRG_HIP_CHECK(hipGraphInstantiate(&grp->exec, grp->graph, nullptr, nullptr, 0));
        } catch(...) {
            hipGraph_t partial = nullptr;
            (void) hipStreamEndCapture(cs, &partial); // leave capture mode before the stream goes
            if(partial) (void) hipGraphDestroy(partial);
            (void) hipStreamDestroy(cs);
            throw;
        }
        DMRG_HIP_CHECK(hipStreamDestroy(cs));
    }
    return grp.release();
}

} // namespace dmrg

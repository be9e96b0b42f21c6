// Grouped H_eff matvecs for serving many small operators at once: the matvecs of k operators (uneven
// shapes) run as one launch per contraction leg, optionally captured once into a HIP graph and replayed.
#pragma once
#include "matvec_op.hpp"

#include <vector>

namespace dmrg {

struct MatvecGroup {
    std::vector<GroupedGemm> legs[3];
    std::vector<DevBuf>      ws;     // one intermediate workspace per operator
    hipGraph_t               graph = nullptr;
    hipGraphExec_t           exec  = nullptr;
    int                      k     = 0;
    double                   flop  = 0; // algorithmic FLOP of one run (8 per complex MAC)
    MatvecGroup()                    = default;
    MatvecGroup(const MatvecGroup &) = delete;
    MatvecGroup &operator=(const MatvecGroup &) = delete;
    ~MatvecGroup();
    void run(hipStream_t s);
};

// Bytes of the intermediate workspace one operator's matvec needs in a group.
size_t matvec_group_ws_bytes(const MatVecOp &op, bool fp32);

// Caller-owned storage that lets a group be built without allocating: wss[i] of matvec_group_ws_bytes(ops[i])
// bytes per operator and three descriptor stores (one per leg) of grouped_store_bytes(k) bytes each.
struct GroupStorage {
    void *const *wss            = nullptr;
    void        *stores[3]      = {nullptr, nullptr, nullptr};
    size_t       store_bytes[3] = {0, 0, 0};
};

// y_i = H_i x_i for the k operators; xs/ys are device vectors of each operator's size, in its dtype (or the
// single-precision dtype with fp32). The group keeps pointers into the operators' buffers and into xs/ys:
// they must stay as they are while the group is used. With storage, nothing is allocated (nor freed when the
// group goes).
MatvecGroup *matvec_group_create(MatVecOp *const *ops, int k, const void *const *xs, void *const *ys, bool fp32, bool capture,
                                 const GroupStorage *storage = nullptr);

} // namespace dmrg

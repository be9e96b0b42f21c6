// Allocation-free H_eff matvec for callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc). The two matvec intermediates live in a file-scope
// device array of fixed size instead of the library's grow-only workspace; no leg is split over K (the
// split-K partial slabs would need more room), and GEMM plans are never timed here (a plan tuned earlier for
// the same signature is used, else the size heuristic's).
//
// Each device holds its own copy of the array from the first call on it until the process ends. Every call
// on one device shares it: calls that overlap on one GPU (on different streams or host threads) race on it.
// The intermediates are fully overwritten by every call, so nothing has to be cleared between calls.
#include "contract.hpp"

#include <mutex>
#include <vector>

namespace dmrg {

constexpr size_t NOALLOC_SCRATCH_BYTES = static_cast<size_t>(DMRG_NOALLOC_SCRATCH_BYTES);

// The matvec intermediates of dmrg_matrix_vector_product_noalloc (one copy per device).
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex          g_scratch_mu;
std::vector<void *> g_scratch_ptr; // per device: the address of that device's copy of scratch

// The device address of scratch on the current device, looked up once per device. The array's name is a host
// stand-in, never a device address: kernels get this pointer.
void *device_scratch() {
    int dev = 0;
    DMRG_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if(static_cast<int>(g_scratch_ptr.size()) <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if(!g_scratch_ptr[dev]) {
        void *p = nullptr;
        DMRG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        DMRG_REQUIRE(p, DMRG_ERR_HIP, "hipGetSymbolAddress returned no address for the matvec scratch array");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

} // namespace
} // namespace dmrg

using namespace dmrg;

extern "C" {

dmrg_status dmrg_matrix_vector_product_noalloc(dmrg_dtype dtype, void *res, const void *mps, const int64_t mps_dims[3],
                                               const void *mpo, const int64_t mpo_dims[4], const void *envL,
                                               const int64_t envL_dims[3], const void *envR, const int64_t envR_dims[3],
                                               void *stream) {
    try {
        DMRG_REQUIRE(res && mps && mpo && envL && envR && mps_dims && mpo_dims && envL_dims && envR_dims, DMRG_ERR_ARG,
                     "null argument");
        const Dims   D    = check_matvec_dims(mps_dims, mpo_dims, envL_dims, envR_dims);
        const size_t need = static_cast<size_t>(matvec_ws_elems_nopart(D)) * dtype_size(dtype);
        DMRG_REQUIRE(need <= NOALLOC_SCRATCH_BYTES, DMRG_ERR_ARG,
                     "dmrg_matrix_vector_product_noalloc: the intermediates need " + std::to_string(need) +
                         " bytes, more than the " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                         " bytes of its fixed scratch (use dmrg_matrix_vector_product)");
        void       *ws = device_scratch();
        NoTuneScope no_tune;
        matvec(dtype, D, mps, D.n(), res, D.n(), 1, mpo, envL, envR, ws, static_cast<hipStream_t>(stream), nullptr,
               /*partials=*/false);
        return DMRG_OK;
    } catch(const Error &e) {
        set_last_error(e.what());
        return e.status;
    } catch(const std::exception &e) {
        set_last_error(e.what());
        return DMRG_ERR_ARG;
    } catch(...) {
        set_last_error("unknown error");
        return DMRG_ERR_ARG;
    }
}

int64_t dmrg_matrix_vector_product_noalloc_max_bytes(void) { return static_cast<int64_t>(NOALLOC_SCRATCH_BYTES); }

} // extern "C"

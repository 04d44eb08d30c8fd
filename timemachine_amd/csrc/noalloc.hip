// Allocation-free nonbonded execute for callers that cannot allocate device
// memory (code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc). Every intermediate array of the neighborlist build and
// the tile kernel lies in one file-scope device array of fixed size; the
// caller passes inputs and outputs only.
//
// Each device holds its own copy of `scratch` (at its own address) until the
// process ends. Its address is taken once per device with
// hipGetSymbolAddress while that device is current, and only that address is
// handed to kernels and copies: the array's name in host code is a host-side
// stand-in, not a device address.
//
// Every call on a device shares that device's array and leaves its data
// there; the next call rewrites the descriptor, clears the counters and
// rebuilds the neighborlist before reading anything. Calls that overlap on
// one device (two streams, two host threads) race on the array.
#include <mutex>

#include "potentials.h"

namespace tmamd {

namespace {

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up on the first
// call made with that device current).
void *scratch_address() {
    int dev = 0;
    TM_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        throw Error("execute_device_noalloc(): device index " + std::to_string(dev) + " out of range");
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        TM_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) {
            throw HipError("hipGetSymbolAddress(scratch) returned a null address");
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

} // namespace

void execute_device_noalloc(Potential &pot, int N, int P, const double *d_x, const double *d_p, const double *d_box,
                            double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u,
                            hipStream_t stream) {
    auto *nb = dynamic_cast<NonbondedAllPairs *>(&pot);
    if (nb == nullptr) {
        throw Error("execute_device_noalloc(): the potential is not a NonbondedAllPairs");
    }
    nb->execute_device_in(scratch_address(), NOALLOC_SCRATCH_BYTES, N, P, d_x, d_p, d_box, lambda, d_du_dx, d_du_dp,
                          d_du_dl, d_u, stream);
}

} // namespace tmamd

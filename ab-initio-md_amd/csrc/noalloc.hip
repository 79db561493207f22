// Fixed-size device scratch of the allocation-free entry point
// (aimdhf_grad_one_noalloc): one file-scope array per device, its address
// resolved once per device with that device current. Host code uses only the
// resolved address, never the array's name (which on the host is a stand-in,
// not a device address).

#include "aimdhf_internal.h"

#include <mutex>

namespace aimdhf {

namespace {
// symmetrised D and -W (2 nao^2 doubles) and the fixed-point gradient (6 natm
// int64) of a one-electron gradient with nao <= 1024
constexpr size_t NOALLOC_BYTES = (size_t(16) << 20) + (size_t(64) << 10);
constexpr int MAX_DEVICES = 64;

std::mutex g_mu;
void* g_addr[MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

void* noalloc_scratch(size_t* bytes) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_addr[dev] = p;
    }
    if (bytes) *bytes = NOALLOC_BYTES;
    return g_addr[dev];
}

}  // namespace aimdhf

// grt_noalloc.hip -- the fixed device scratch of grt_trace_device_noalloc,
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The scratch is one file-scope __device__ array of fixed size: work lists,
// per-photon RNG and deferral state and the driver counters for at most
// GRT_NOALLOC_MAX_PHOTONS photons (about 13 MiB; the launch constants travel
// in the kernel arguments).  Every GPU the
// library runs on holds its own copy until the process ends.  In host code
// the array's name is only a stand-in for its device address, so the address
// of each device's copy is looked up once (hipGetSymbolAddress, with that
// device current) and only that address reaches kernels and copies.
//
// Every no-allocation trace on one GPU shares that GPU's array: calls that
// overlap on one GPU (several host threads, or streams that run
// concurrently) race.  Each call clears what must start at zero.
#include <hip/hip_runtime.h>

#include <mutex>

#include "grt_internal.hpp"

namespace grt {
namespace {

constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

constexpr size_t kMaxN = GRT_NOALLOC_MAX_PHOTONS;
constexpr size_t kOffLists = 0;
constexpr size_t kOffRng = align_up(kOffLists + 2 * kMaxN * sizeof(uint32_t));
constexpr size_t kOffFpend = align_up(kOffRng + kMaxN * sizeof(uint32_t));
constexpr size_t kOffCounters = align_up(kOffFpend + kMaxN * sizeof(uint8_t));
constexpr size_t kScratchBytes = align_up(kOffCounters + kReadbackBytes);

}  // namespace
}  // namespace grt

static __device__ __align__(256) unsigned char scratch[grt::kScratchBytes];

namespace grt {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_ptr[kMaxDevices] = {};

// Grid-stride clear of `words` 8-byte words.
__global__ void zero_words(uint64_t* w, size_t words)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) w[i] = 0;
}

// Clears bytes [0, n), n < 8, one thread each.
__global__ void zero_tail(unsigned char* p, unsigned n)
{
    if (threadIdx.x < n) p[threadIdx.x] = 0;
}

}  // namespace

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = static_cast<unsigned char*>(p);
    return hipSuccess;
}

hipError_t noalloc_scratch(int device, TraceScratch* s)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        base = g_ptr[device];
    }
    if (!base) return hipErrorNotInitialized;
    s->lists = reinterpret_cast<uint32_t*>(base + kOffLists);
    s->list_cap = kMaxN;
    s->rng = reinterpret_cast<uint32_t*>(base + kOffRng);
    s->fpend = base + kOffFpend;
    s->counters = reinterpret_cast<uint32_t*>(base + kOffCounters);
    s->fill_kernel = true;
    return hipSuccess;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t stream)
{
    if (bytes == 0) return hipSuccess;
    if (reinterpret_cast<uintptr_t>(p) % 8 != 0) return hipErrorInvalidValue;
    unsigned char* b = static_cast<unsigned char*>(p);
    const size_t words = bytes / 8;
    const unsigned tail = (unsigned)(bytes % 8);
    if (words > 0) {
        const int block = 256;
        const size_t want = (words + block - 1) / block;
        const int blocks = (int)(want < 2048 ? want : 2048);
        hipLaunchKernelGGL(zero_words, dim3(blocks), dim3(block), 0, stream,
                           reinterpret_cast<uint64_t*>(b), words);
    }
    if (tail) hipLaunchKernelGGL(zero_tail, dim3(1), dim3(64), 0, stream, b + 8 * words, tail);
    return hipGetLastError();
}

}  // namespace grt

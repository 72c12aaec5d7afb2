// Allocation-free stable x-plane binning: hy_xplane_order_static.
//
// Same result as hy_pm_xplane_order (the reference's stable counting sort of
// particle indices by wrapped base x-cell, src/mesh/mass_assignment_workspace.inc:
// 264-374), for callers that cannot allocate device memory.  No context: the
// geometry comes in the arguments and every device word the kernels need
// beyond the caller's inputs and outputs lives in one file-scope device array.
//
// Algorithm (deterministic, stable, no per-particle scratch):
//   1. the particles are cut into C contiguous chunks; one workgroup per chunk
//      counts its particles per x-plane in LDS and stores the chunk's column
//      of a plane-major counter matrix cnt[plane][chunk];
//   2. an in-place exclusive scan of the matrix in plane-major order gives,
//      for every (plane, chunk), the first output slot of that chunk's
//      particles in that plane;
//   3. one wavefront per chunk walks its particles in index order, 64 at a
//      time; lanes with equal planes are ranked by lane order with ballots, so
//      equal keys keep their index order (stability) without sorting;
//   4. offsets[plane] = scanned cnt[plane][0], offsets[N] = the grand total.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "hy_device.hpp"
#include "hyowon_pm.h"

namespace {

constexpr uint64_t kCounterWords = 1ull << 22;               // plane x chunk counters (16 MiB)
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;     // counters per scan workgroup
constexpr uint64_t kScanBlocks = kCounterWords / kScanTile;  // 2048 workgroup sums at most
constexpr uint64_t kErrWord = kCounterWords + kScanBlocks;    // error bits
constexpr uint64_t kTotalWord = kErrWord + 1;                 // grand total of the scan
constexpr uint64_t kScratchWords = kTotalWord + 63;           // padded to 256-byte multiple
constexpr uint32_t kMinChunk = 2048;                          // particles per chunk at least
constexpr int kMaxMesh = 1024;                                // LDS histogram of N words
constexpr int kMaxDevices = 64;

}  // namespace

// The one scratch array (per device, held until the process ends).  Host code
// never uses this name as an address: it resolves the device address once
// per device with hipGetSymbolAddress and passes that pointer.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

thread_local std::string t_error;
std::mutex g_mu;
uint32_t* g_ptr[kMaxDevices] = {};

struct StaticError {
    int code;
    std::string what;
};

[[noreturn]] void fail(int code, const std::string& what) { throw StaticError{code, what}; }

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(HY_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Device address of `scratch` on `device` (the current device), resolved once.
uint32_t* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) fail(HY_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_ptr[device] = static_cast<uint32_t*>(p);
    }
    return g_ptr[device];
}

// mesh sizes of hy_pm_create: 2^a (8..1024) or 3*2^a (12..384)
bool supported_mesh(uint64_t N) {
    for (uint64_t p = 8; p <= 1024; p *= 2)
        if (N == p || (N == 3 * p / 2 && N >= 12 && N <= 384)) return true;
    return false;
}

__global__ void k_xp_clear(uint32_t* __restrict__ s) {
    if (threadIdx.x < 2) s[kErrWord + threadIdx.x] = 0u;
}

// 1. per-chunk plane counts; cnt[plane * C + chunk]
__global__ void __launch_bounds__(256) k_xp_count(uint64_t n, const double* __restrict__ x, hy::Geom g, uint64_t chunk,
                                                  uint32_t C, uint32_t* __restrict__ s) {
    __shared__ uint32_t hist[kMaxMesh];
    for (int i = threadIdx.x; i < g.N; i += blockDim.x) hist[i] = 0u;
    __syncthreads();
    const uint64_t b = (uint64_t)blockIdx.x * chunk;
    const uint64_t e = n < b + chunk ? n : b + chunk;
    bool bad = false;
    for (uint64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
        int bx;
        double f;
        if (hy::cic_axis(x[p], g, bx, f)) atomicAdd(&hist[bx], 1u);
        else bad = true;
    }
    if (bad) atomicOr(&s[kErrWord], hy::ERR_POSITION_NONFINITE);
    __syncthreads();
    for (int i = threadIdx.x; i < g.N; i += blockDim.x) s[(uint64_t)i * C + blockIdx.x] = hist[i];
}

__device__ __forceinline__ uint32_t wg_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < (int)kScanThreads; off <<= 1) {
        const uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[kScanThreads - 1];
    const uint32_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

// 2a. workgroup sums of the counter matrix
__global__ void __launch_bounds__(kScanThreads) k_xp_scan_reduce(uint64_t m, uint32_t* __restrict__ s) {
    __shared__ uint32_t sh[kScanThreads];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t sum = 0;
    for (uint32_t i = 0; i < kScanItems; ++i) {
        const uint64_t j = base + (uint64_t)i * kScanThreads + threadIdx.x;
        if (j < m) sum += s[j];
    }
    uint32_t total;
    (void)wg_exclusive_scan(sum, sh, total);
    if (threadIdx.x == 0) s[kCounterWords + blockIdx.x] = total;
}

// 2b. exclusive scan of the workgroup sums (one workgroup) and the grand total
__global__ void __launch_bounds__(kScanThreads) k_xp_scan_sums(uint32_t nblocks, uint32_t* __restrict__ s) {
    __shared__ uint32_t sh[kScanThreads];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblocks; base += kScanThreads) {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < nblocks ? s[kCounterWords + j] : 0u;
        uint32_t total;
        const uint32_t ex = wg_exclusive_scan(v, sh, total);
        if (j < nblocks) s[kCounterWords + j] = ex + carry;
        carry += total;
    }
    if (threadIdx.x == 0) s[kTotalWord] = carry;
}

// 2c. in-place exclusive scan of the matrix: every thread reads its items
// into registers before the workgroup scan and writes them back after it
__global__ void __launch_bounds__(kScanThreads) k_xp_scan_apply(uint64_t m, uint32_t* s) {
    __shared__ uint32_t sh[kScanThreads];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
    for (uint32_t i = 0; i < kScanItems; ++i) {
        v[i] = base + i < m ? s[base + i] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t run = wg_exclusive_scan(sum, sh, total) + s[kCounterWords + blockIdx.x];
    for (uint32_t i = 0; i < kScanItems; ++i) {
        if (base + i < m) s[base + i] = run;
        run += v[i];
    }
}

// 3. stable scatter: one wavefront per chunk, particles in index order
__global__ void __launch_bounds__(64) k_xp_scatter(uint64_t n, const double* __restrict__ x, hy::Geom g,
                                                   uint64_t chunk, uint32_t C, const uint32_t* __restrict__ s,
                                                   uint32_t* __restrict__ order) {
    __shared__ uint32_t next[kMaxMesh];
    for (int i = threadIdx.x; i < g.N; i += 64) next[i] = s[(uint64_t)i * C + blockIdx.x];
    __syncthreads();
    const uint32_t lane = hy::lane_id();
    const uint64_t b = (uint64_t)blockIdx.x * chunk;
    const uint64_t e = n < b + chunk ? n : b + chunk;
    for (uint64_t p0 = b; p0 < e; p0 += 64) {
        const uint64_t p = p0 + lane;
        int bx = -1;
        double f;
        const bool ok = p < e && hy::cic_axis(x[p], g, bx, f);
        // lanes of one plane, ranked by lane order
        uint64_t remaining = __ballot(ok);
        uint32_t slot = 0;
        while (remaining) {
            const int leader = __ffsll((unsigned long long)remaining) - 1;
            const int key = __builtin_amdgcn_readlane(bx, leader);
            const uint64_t mask = __ballot(ok && bx == key) & remaining;
            if (ok && bx == key) slot = next[key] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            __syncthreads();
            if ((int)lane == leader) next[key] += (uint32_t)__popcll(mask);
            __syncthreads();
            remaining &= ~mask;
        }
        if (ok) order[slot] = (uint32_t)p;
    }
}

// 4. plane offsets and the status check
__global__ void k_xp_offsets(int N, uint32_t C, uint64_t n, const uint32_t* __restrict__ s,
                             uint32_t* __restrict__ offsets, uint32_t* __restrict__ err) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= N; i += gridDim.x * blockDim.x) {
        if (i < N) offsets[i] = s[(uint64_t)i * C];
        else {
            offsets[N] = s[kTotalWord];
            if ((uint64_t)s[kTotalWord] != n) atomicOr(err, hy::ERR_ROUTING);
        }
    }
}

}  // namespace

extern "C" int hy_xplane_order_static(double box_size, uint64_t mesh_size, uint64_t n, const double* x,
                                      uint32_t* order, uint32_t* offsets, int device, void* stream) {
    try {
        if (!offsets || (n > 0 && (!x || !order))) fail(HY_EINVAL, "hy_xplane_order_static: null argument");
        if (!std::isfinite(box_size) || box_size <= 0.0) fail(HY_EINVAL, "MeshGeometry box size must be finite and positive");
        if (!supported_mesh(mesh_size)) fail(HY_EINVAL, "mesh size must be 2^a (8..1024) or 3*2^a (12..384)");
        if (n > 0xffffffffull) fail(HY_EOVERFLOW, "particle count exceeds the 32-bit index range");
        const int N = (int)mesh_size;
        hy::Geom g{};
        g.L = box_size;
        g.dx = box_size / (double)N;
        if (!std::isfinite(g.dx) || g.dx <= 0.0) fail(HY_EOVERFLOW, "MeshGeometry cell size is not representable");
        int e2 = 0;
        g.inv_exact = (std::frexp(g.dx, &e2) == 0.5 && std::isfinite(1.0 / g.dx)) ? 1 : 0;
        g.inv_dx = 1.0 / g.dx;
        g.upper = std::nextafter((double)N, 0.0);
        g.N = N;
        g.T = 1;
        g.nT = N;
        g.nbox = 1;
        g.nper = (uint32_t)n;

        // chunks: at least kMinChunk particles each, at most kCounterWords / N of them
        const uint64_t cmax = kCounterWords / (uint64_t)N;
        uint64_t C = std::max<uint64_t>(1, std::min<uint64_t>((n + kMinChunk - 1) / kMinChunk, cmax));
        uint64_t chunk = (n + C - 1) / C;
        chunk = std::max<uint64_t>(64, (chunk + 63) / 64 * 64);
        C = std::max<uint64_t>(1, (n + chunk - 1) / chunk);
        const uint64_t m = (uint64_t)N * C;  // counters in use (<= kCounterWords)
        if (m > kCounterWords) fail(HY_ELOGIC, "x-plane counter matrix exceeds the static scratch");
        const uint64_t nblk = (m + kScanTile - 1) / kScanTile;

        hip_ok(hipSetDevice(device), "hipSetDevice");
        uint32_t* s = scratch_on(device);
        hipStream_t st = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_xp_clear, dim3(1), dim3(64), 0, st, s);
        hipLaunchKernelGGL(k_xp_count, dim3((uint32_t)C), dim3(256), 0, st, n, x, g, chunk, (uint32_t)C, s);
        hipLaunchKernelGGL(k_xp_scan_reduce, dim3((uint32_t)nblk), dim3(kScanThreads), 0, st, m, s);
        hipLaunchKernelGGL(k_xp_scan_sums, dim3(1), dim3(kScanThreads), 0, st, (uint32_t)nblk, s);
        hipLaunchKernelGGL(k_xp_scan_apply, dim3((uint32_t)nblk), dim3(kScanThreads), 0, st, m, s);
        if (n > 0)
            hipLaunchKernelGGL(k_xp_scatter, dim3((uint32_t)C), dim3(64), 0, st, n, x, g, chunk, (uint32_t)C, s,
                               order);
        hipLaunchKernelGGL(k_xp_offsets, dim3((uint32_t)((N + 256) / 256)), dim3(256), 0, st, N, (uint32_t)C, n, s,
                           offsets, s + kErrWord);
        hip_ok(hipGetLastError(), "hy_xplane_order_static launch");
        uint32_t err = 0;
        hip_ok(hipMemcpyAsync(&err, s + kErrWord, sizeof(err), hipMemcpyDeviceToHost, st), "status copy");
        hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (err & hy::ERR_POSITION_NONFINITE) fail(HY_EINVAL, "CIC positions must be finite");
        if (err) fail(HY_ELOGIC, "CIC x-plane binning does not cover the complete particle population");
        return HY_OK;
    } catch (const StaticError& e) {
        t_error = e.what;
        return e.code;
    } catch (const std::exception& e) {
        t_error = e.what();
        return HY_ERUNTIME;
    }
}

extern "C" const char* hy_static_last_error(void) { return t_error.c_str(); }

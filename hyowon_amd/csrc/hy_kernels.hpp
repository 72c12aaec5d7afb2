// HIP kernels of the MI355X particle-mesh gravity step (gfx950, wave64).
//
// Stage map (reference hot path, file:line relative to the reference tree):
//   k_bucket_key / k_scan_* / k_scatter  stable bucketing of particles that
//        feeds the deterministic deposit (reference x-plane counting sort,
//        src/mesh/mass_assignment_workspace.inc:264-374, generalised to
//        tile buckets; determinism restored by k_deposit's index sort).
//   k_deposit    CIC deposit + density source, summation order per cell equal
//        to src/mesh/mass_assignment_deposit.inc:193-240 (bin X with weight
//        1-f, then bin X-1 with weight f, particles in index order, oy outer,
//        oz inner), density contrast as src/gravity/pm_solver.cpp:700-759.
//   k_fft_*      (hy_fft_passes.hpp) r2c / Green / c2r Poisson solve (src/mesh/fft_backend.cpp,
//        src/mesh/green_function.cpp:165-241).
//   k_gather_kdk fused centered-gradient CIC gather
//        (src/mesh/mass_assignment_gradient.cpp:117-206) with the K2 kick,
//        the next K1 kick and the periodic drift
//        (include/cosmo_nbody/time/particle_update.hpp:20-116).
#pragma once

#include "hy_device.hpp"
#include <type_traits>
#include "hy_fft.hpp"

namespace hy {

// ===========================================================================
// Bucketing
// ===========================================================================

// Deposit bucket of a base cell: tile (edge T) and boundary class
// (bit2/1/0: last cell of the tile along x/y/z).
__device__ __forceinline__ uint32_t bucket_of(int bx, int by, int bz, const Geom& g) {
    // T = 16 (every mesh divisible by 16) takes shifts instead of divisions
    const int tx = g.T == 16 ? bx >> 4 : bx / g.T;
    const int ty = g.T == 16 ? by >> 4 : by / g.T;
    const int tz = g.T == 16 ? bz >> 4 : bz / g.T;
    const uint32_t cls = ((uint32_t)(bx - tx * g.T == g.T - 1) << 2)
                       | ((uint32_t)(by - ty * g.T == g.T - 1) << 1)
                       | (uint32_t)(bz - tz * g.T == g.T - 1);
    return (((uint32_t)tx * g.nT + ty) * g.nT + tz) * 8u + cls;
}

// Cell of a deposit record inside its tile: (lx*T + ly)*T + lz.
__device__ __forceinline__ uint32_t record_cell(int bx, int by, int bz, const Geom& g) {
    const int T = g.T;
    const int lx = T == 16 ? bx & 15 : bx % T, ly = T == 16 ? by & 15 : by % T, lz = T == 16 ? bz & 15 : bz % T;
    return (uint32_t)((lx * T + ly) * T + lz);
}

// Wave-aggregated integer atomic: lanes with equal keys reserve consecutive
// slots.  Groups are found with ballots first; then every group leader issues
// its atomic in the same instruction, so a wave pays one atomic round trip.
// Returns this lane's rank (slot) within its bucket.
__device__ __forceinline__ uint32_t aggregated_rank(uint32_t* counts, uint32_t key, bool valid) {
    const uint32_t lane = lane_id();
    uint64_t remaining = __ballot(valid);
    int my_leader = (int)lane;
    uint64_t my_mask = 0;
    while (remaining) {
        const int leader = __ffsll((unsigned long long)remaining) - 1;
        const uint32_t lead_key = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);  // uniform lane: a scalar read
        const uint64_t mask = __ballot(valid && key == lead_key) & remaining;
        if ((mask >> lane) & 1ull) {
            my_leader = leader;
            my_mask = mask;
        }
        remaining &= ~mask;
    }
    uint32_t old = 0;
    if (valid && (int)lane == my_leader) old = atomicAdd(&counts[key], (uint32_t)__popcll(my_mask));
    const uint32_t base = __shfl(old, my_leader);
    return base + (uint32_t)__popcll(my_mask & ((1ull << lane) - 1ull));
}

// mode 0: deposit buckets; mode 1: reference x-plane bins (key = base x).
// shift: positions are routed as x - shift (the interlacing partner of the
// power-spectrum estimator; 0 otherwise, where x - 0 is x exactly).
__global__ void k_bucket_key(uint64_t n, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, Geom g, int mode,
                             uint32_t* __restrict__ keys, uint32_t* __restrict__ ranks,
                             uint32_t* __restrict__ counts, uint32_t* __restrict__ err, double shift) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = p < n;
    uint32_t key = 0;
    bool ok = false;
    if (in) {
        int bx, by = 0, bz = 0;
        double f;
        ok = cic_axis(x[p] - shift, g, bx, f);
        if (mode == 0) ok = ok && cic_axis(y[p] - shift, g, by, f) && cic_axis(z[p] - shift, g, bz, f);
        if (ok) key = mode == 0 ? box_of(p, g) * buckets_per_box(g) + bucket_of(bx, by, bz, g) : (uint32_t)bx;
        else atomicOr(err, ERR_POSITION_NONFINITE);
    }
    const uint32_t rank = aggregated_rank(counts, key, in && ok);
    if (in && !ok) key = 0xffffffffu;
    if (in) {
        keys[p] = key;
        ranks[p] = rank;
    }
}

// Exclusive scan of counts[0..nb) into offsets[0..nb], three launches.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    // Hillis-Steele over 256 threads in LDS (deterministic integer sums).
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        const uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[SCAN_THREADS - 1];
    const uint32_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ void k_scan_reduce(const uint32_t* __restrict__ counts, uint64_t nb,
                              uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t sh[SCAN_THREADS];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint64_t j = base + (uint64_t)i * SCAN_THREADS + threadIdx.x;
        if (j < nb) s += counts[j];
    }
    uint32_t total;
    (void)block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ void k_scan_blocks(uint32_t* __restrict__ block_sums, uint32_t nblocks,
                              uint32_t* __restrict__ grand_total) {
    __shared__ uint32_t sh[SCAN_THREADS];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblocks; base += SCAN_THREADS) {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < nblocks ? block_sums[j] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, sh, total);
        if (j < nblocks) block_sums[j] = ex + carry;
        carry += total;
    }
    if (threadIdx.x == 0) *grand_total = carry;
}

__global__ void k_scan_apply(const uint32_t* __restrict__ counts, uint64_t nb,
                             const uint32_t* __restrict__ block_sums,
                             const uint32_t* __restrict__ grand_total,
                             uint32_t* __restrict__ offsets) {
    __shared__ uint32_t sh[SCAN_THREADS];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < nb) ? counts[base + i] : 0u;
        s += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, sh, total) + block_sums[blockIdx.x];
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < nb) offsets[base + i] = run;
        run += v[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) offsets[nb] = *grand_total;
}

__global__ void k_scatter(uint64_t n, const uint32_t* __restrict__ keys,
                          const uint32_t* __restrict__ ranks, const uint32_t* __restrict__ offsets,
                          uint32_t* __restrict__ order) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        const uint32_t k = keys[p];
        if (k != 0xffffffffu) order[offsets[k] + ranks[p]] = (uint32_t)p;
    }
}

// Deposit record in bucket order: tile-local base cell (lx*T+ly)*T+lz and the
// three CIC fractions (computed in binary64, stored in binary32).
struct DepRecord {
    uint32_t cell;
    float fx, fy, fz;
};

// ---------------------------------------------------------------------------
// Capacity layout of carried deposit records (binary32, uniform mass).
//
// A fused step that drifts the particles writes each particle's deposit
// record for the next step straight into a slot of its bucket: bucket b owns
// cap(c_b) = c_b + c_b/8 + CAP_SLACK slots, c_b its count at this step (the
// particles move little per step), and a record whose rank within its bucket
// reaches the capacity goes to an overflow list instead.  Before the next
// deposit one three-channel scan over the final counts gives
//   ch 0: the exclusive prefix of the counts (dense offsets; its total is the
//         routing identity's grand total),
//   ch 1: the prefix of the overflow counts max(0, c_b - cap_b) of the layout
//         the records were written in (overflow records are placed after the
//         capacity slots, bucket by bucket),
//   ch 2: the prefix of cap(c_b), the layout the next step's records go to.
// The binary32 deposit sums integers, so the order of a bucket's records
// (the ranks come from atomics) leaves every cell sum unchanged.
// ---------------------------------------------------------------------------
constexpr uint32_t CAP_SLACK = 4;

__device__ __forceinline__ uint32_t capacity_of(uint32_t count) { return count + (count >> 3) + CAP_SLACK; }

struct CapScan {
    const uint32_t* counts;  // records per bucket
    uint64_t nb;             // buckets over all boxes
    const uint32_t* cur;     // nb+1 prefix of the layout the records were written in (null: no overflow)
    uint32_t* dense;         // out: nb+1 prefix of the counts (or null)
    uint32_t* ovf_start;     // out: nb+1 prefix of the overflow counts (when cur)
    uint32_t* capn;          // out: nb+1 prefix of the next layout's capacities (or null)
    uint32_t* block_sums;    // 3 per scan block
    uint32_t* grand_total;   // out: total count
};

__device__ __forceinline__ void cap_channels(const CapScan& s, uint64_t b, uint32_t v[3]) {
    const uint32_t c = s.counts[b];
    v[0] = c;
    if (s.cur) {
        const uint32_t cap = s.cur[b + 1] - s.cur[b];
        v[1] = c > cap ? c - cap : 0u;
    } else {
        v[1] = 0u;
    }
    v[2] = capacity_of(c);
}

__global__ void k_cap_reduce(CapScan s) {
    __shared__ uint32_t sh[SCAN_THREADS];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t acc[3] = {0u, 0u, 0u};
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint64_t j = base + (uint64_t)i * SCAN_THREADS + threadIdx.x;
        if (j < s.nb) {
            uint32_t v[3];
            cap_channels(s, j, v);
            for (int ch = 0; ch < 3; ++ch) acc[ch] += v[ch];
        }
    }
    for (int ch = 0; ch < 3; ++ch) {
        uint32_t total;
        (void)block_exclusive_scan(acc[ch], sh, total);
        if (threadIdx.x == 0) s.block_sums[3 * blockIdx.x + ch] = total;
    }
}

// One workgroup: exclusive scan of each channel's block sums; the totals go
// to index nb of each output and the count total to grand_total.
__global__ void k_cap_blocks(CapScan s, uint32_t nblocks) {
    __shared__ uint32_t sh[SCAN_THREADS];
    for (int ch = 0; ch < 3; ++ch) {
        uint32_t carry = 0;
        for (uint32_t base = 0; base < nblocks; base += SCAN_THREADS) {
            const uint32_t j = base + threadIdx.x;
            const uint32_t v = j < nblocks ? s.block_sums[3 * j + ch] : 0u;
            uint32_t total;
            const uint32_t ex = block_exclusive_scan(v, sh, total);
            if (j < nblocks) s.block_sums[3 * j + ch] = ex + carry;
            carry += total;
        }
        if (threadIdx.x == 0) {
            if (ch == 0) {
                *s.grand_total = carry;
                if (s.dense) s.dense[s.nb] = carry;
            } else if (ch == 1 && s.cur) {
                s.ovf_start[s.nb] = carry;
            } else if (ch == 2 && s.capn) {
                s.capn[s.nb] = carry;
            }
        }
    }
}

__global__ void k_cap_apply(CapScan s) {
    __shared__ uint32_t sh[SCAN_THREADS];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS][3];
    uint32_t acc[3] = {0u, 0u, 0u};
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < s.nb) {
            cap_channels(s, base + i, v[i]);
        } else {
            v[i][0] = v[i][1] = v[i][2] = 0u;
        }
        for (int ch = 0; ch < 3; ++ch) acc[ch] += v[i][ch];
    }
    uint32_t run[3];
    for (int ch = 0; ch < 3; ++ch) {
        uint32_t total;
        run[ch] = block_exclusive_scan(acc[ch], sh, total) + s.block_sums[3 * blockIdx.x + ch];
    }
    uint32_t* out[3] = {s.dense, s.cur ? s.ovf_start : nullptr, s.capn};
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < s.nb) {
            for (int ch = 0; ch < 3; ++ch)
                if (out[ch]) out[ch][base + i] = run[ch];
        }
        for (int ch = 0; ch < 3; ++ch) run[ch] += v[i][ch];
    }
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Overflow records of the carried layout cur: entry i holds its bucket, its
// rank and its record; it goes to cur[nb] + ovf_start[b] + (rank - cap_b), so each
// bucket's overflow records follow the capacity slots contiguously.  Grid
// stride up to the device count the gather left.
__global__ void k_place_overflow(const uint32_t* __restrict__ ovf_count, const uint32_t* __restrict__ ovf_key,
                                 const uint32_t* __restrict__ ovf_rank, const DepRecord* __restrict__ ovf_rec,
                                 const uint32_t* __restrict__ cur, const uint32_t* __restrict__ ovf_start,
                                 uint64_t nb, DepRecord* __restrict__ rec) {
    const uint32_t total = *ovf_count;
    const uint32_t base = cur[nb];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const uint32_t b = ovf_key[i];
        const uint32_t cap = cur[b + 1] - cur[b];
        rec[base + ovf_start[b] + (ovf_rank[i] - cap)] = ovf_rec[i];
    }
}

// Scatter with keys made by an earlier call: each key is checked against the
// bucket of the current position first (ERR_STALE_KEYS on a mismatch).
__global__ void k_scatter_checked(uint64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                  const double* __restrict__ z, Geom g, const uint32_t* __restrict__ keys,
                                  const uint32_t* __restrict__ ranks, const uint32_t* __restrict__ offsets,
                                  uint32_t* __restrict__ order, uint32_t* __restrict__ err) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = keys[p];
    int bx, by, bz;
    double f;
    const bool ok = cic_axis(x[p], g, bx, f) && cic_axis(y[p], g, by, f) && cic_axis(z[p], g, bz, f);
    if (!ok) {
        atomicOr(err, ERR_POSITION_NONFINITE);
        return;
    }
    if (k == 0xffffffffu || box_of(p, g) * buckets_per_box(g) + bucket_of(bx, by, bz, g) != k) {
        atomicOr(err, ERR_STALE_KEYS);
        return;
    }
    order[offsets[k] + ranks[p]] = (uint32_t)p;
}

// Binary64 deposit record: tile-local base cell, the particle index (the
// reference's per-cell summation order is by index) and the binary64
// fractions.
struct DepRecord64 {
    uint32_t cell;
    uint32_t index;
    double fx, fy, fz;
};

// Scatter into bucket order, writing the particle index and its deposit
// record.  The bucket is recomputed from the current position and must equal
// the stored key (keys carried over from the previous step are checked here).
// R = DepRecord (binary32 fractions, masses / mass unit) or DepRecord64
// (binary64 fractions and particle index, masses as given).
template <typename R, typename MT>
__global__ void k_scatter_records(uint64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                  const double* __restrict__ z, const double* __restrict__ masses,
                                  double inv_mass_unit, Geom g, const uint32_t* __restrict__ keys,
                                  const uint32_t* __restrict__ ranks, const uint32_t* __restrict__ offsets,
                                  uint32_t* __restrict__ order, R* __restrict__ rec,
                                  MT* __restrict__ rec_mass, uint32_t* __restrict__ err, double shift) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = keys[p];
    int bx, by, bz;
    double fx, fy, fz;
    // The coordinate loads stay behind the short-circuit checks: issued
    // together (cic_axes), the three streams measured 8% slower here.
    const bool ok = cic_axis(x[p] - shift, g, bx, fx) && cic_axis(y[p] - shift, g, by, fy) &&
                    cic_axis(z[p] - shift, g, bz, fz);
    if (!ok) {
        atomicOr(err, ERR_POSITION_NONFINITE);
        return;
    }
    if (k == 0xffffffffu || box_of(p, g) * buckets_per_box(g) + bucket_of(bx, by, bz, g) != k) {
        atomicOr(err, ERR_STALE_KEYS);
        return;
    }
    const uint32_t dst = offsets[k] + ranks[p];
    if (order) order[dst] = (uint32_t)p;
    R r;
    r.cell = record_cell(bx, by, bz, g);
    if constexpr (sizeof(R) == sizeof(DepRecord64)) {
        r.index = (uint32_t)p;
        r.fx = fx;
        r.fy = fy;
        r.fz = fz;
        rec[dst] = r;
        if (masses) rec_mass[dst] = masses[p];
    } else {
        r.fx = (float)fx;
        r.fy = (float)fy;
        r.fz = (float)fz;
        rec[dst] = r;
        if (masses) rec_mass[dst] = (float)(masses[p] * inv_mass_unit);
    }
}

// ===========================================================================
// Workgroup sort (ascending u64), arbitrary n, any address space.
// Bitonic network in the "flip" form: every comparator is ascending, so
// virtual +inf padding past n never has to move into range.
// ===========================================================================
template <typename P>
__device__ __forceinline__ void wg_sort_u64(P keys, uint32_t n) {
    if (n < 2) return;
    uint32_t np2 = 1;
    while (np2 < n) np2 <<= 1;
    const uint32_t half = np2 >> 1;
    for (uint32_t k = 2; k <= np2; k <<= 1) {
        // flip step: compare i with its mirror inside the block of size k
        for (uint32_t c = threadIdx.x; c < half; c += blockDim.x) {
            const uint32_t blk = c / (k >> 1);
            const uint32_t off = c % (k >> 1);
            const uint32_t i = blk * k + off;
            const uint32_t j = blk * k + k - 1 - off;
            if (j < n) {
                const uint64_t a = keys[i], b = keys[j];
                if (a > b) { keys[i] = b; keys[j] = a; }
            }
        }
        __threadfence_block();
        __syncthreads();
        for (uint32_t d = k >> 2; d >= 1; d >>= 1) {
            for (uint32_t c = threadIdx.x; c < half; c += blockDim.x) {
                const uint32_t i = (c / d) * 2 * d + (c % d);
                const uint32_t j = i + d;
                if (j < n) {
                    const uint64_t a = keys[i], b = keys[j];
                    if (a > b) { keys[i] = b; keys[j] = a; }
                }
            }
            __threadfence_block();
            __syncthreads();
        }
    }
}

// Sort each bucket's particle indices ascending (canonical stable order).
// Small buckets sort in LDS, large ones in place in global memory.
constexpr int CANON_THREADS = 256;
constexpr int CANON_LDS = 4096;

__global__ void __launch_bounds__(CANON_THREADS)
k_canonicalize(uint64_t nb, const uint32_t* __restrict__ offsets, uint32_t* __restrict__ order,
               uint64_t* __restrict__ scratch, uint64_t scratch_cap,
               unsigned long long* __restrict__ scratch_cursor, uint32_t* __restrict__ err) {
    __shared__ uint64_t sh[CANON_LDS];
    for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint32_t s = offsets[b], e = offsets[b + 1];
        const uint32_t n = e - s;
        if (n < 2) continue;
        if (n <= CANON_LDS) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) sh[i] = order[s + i];
            __syncthreads();
            wg_sort_u64(sh, n);
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) order[s + i] = (uint32_t)sh[i];
            __syncthreads();
        } else {
            __shared__ uint64_t base_sh;
            if (threadIdx.x == 0) base_sh = atomicAdd(scratch_cursor, (unsigned long long)n);
            __syncthreads();
            const uint64_t base = base_sh;
            __syncthreads();
            if (base + n > scratch_cap) {
                if (threadIdx.x == 0) atomicOr(err, ERR_SCRATCH);
                continue;
            }
            uint64_t* g = scratch + base;
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) g[i] = order[s + i];
            __threadfence_block();
            __syncthreads();
            wg_sort_u64(g, n);
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) order[s + i] = (uint32_t)g[i];
            __syncthreads();
        }
    }
}

// ===========================================================================
// Deterministic CIC deposit + density source
// ===========================================================================
constexpr int DEP_THREADS = 512;
constexpr int DEP_MAX_T = 16;
constexpr int DEP_EXT = (DEP_MAX_T + 1) * (DEP_MAX_T + 1) * (DEP_MAX_T + 1);  // 4913
constexpr int DEP_NRANGES = 20;

template <typename T>
struct DepCfg;
template <>
struct DepCfg<float> { static constexpr int CAP = 2048; };
template <>
struct DepCfg<double> { static constexpr int CAP = 1024; };

// ---------------------------------------------------------------------------
// Exact deposited mass: the reference sums every cell's mass with
// ExactPositiveDoubleSum (src/gravity/pm_solver.cpp:132-180).  Here each box
// has SA_DIGITS counters, one per 32-bit digit of the binary positions
// 2^-1074 .. 2^1230 (every binary64 bit), each holding its digit's sum with
// the carries in its upper bits, plus a word flagging an input that is not a
// finite non-negative number.  The counters are integers, so the total is
// exact whatever the order of the additions; k_mass_exact rounds it once to
// nearest (ties to even), as the host's exact_sum does.
// ---------------------------------------------------------------------------
constexpr int SA_DIGITS = 72;
constexpr int SA_WORDS = SA_DIGITS + 1;
// Each wave of a binary64 deposit workgroup stores its sum as a window of
// SA_WAVE_WIN counters from a base digit (SA_WAVE_WORDS words: base, window;
// base -1: empty); digits outside the window go to the box's counters
// directly.  k_mass_reduce adds the windows up per box.  (Binary32 waves store
// the two digit sums of their fixed-point cells; k_mass_reduce_fixed.)
constexpr int SA_WAVE_WIN = 8;
constexpr int SA_WAVE_WORDS = SA_WAVE_WIN + 1;
constexpr int SA_WAVES_PER_TILE = 8;  // deposit workgroups of 512 threads

// Wave-wide integer sum and minimum in the VALU pipe (no LDS permutes): DPP
// row shifts by 1, 2, 4 and 8 lanes leave each 16-lane row's total (minimum)
// in its last lane, and the four rows' last lanes are read as scalars.  Every
// lane of the wave must be active; the result is the same in every lane.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v, uint32_t fill) {
    // lanes whose source lies outside their row keep fill
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, CTRL, 0xf, 0xf, false);
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    return ((uint64_t)dpp_u32<CTRL>((uint32_t)(v >> 32), 0u) << 32) | dpp_u32<CTRL>((uint32_t)v, 0u);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    v += dpp_u64<0x111>(v);  // row_shr:1
    v += dpp_u64<0x112>(v);  // row_shr:2
    v += dpp_u64<0x114>(v);  // row_shr:4
    v += dpp_u64<0x118>(v);  // row_shr:8
    uint64_t s = 0;
#pragma unroll
    for (int r = 15; r < 64; r += 16)
        s += ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), r) << 32) |
             (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, r);
    return s;
}

// Inclusive prefix sum over the 64 lanes of a wave (every lane active): the
// DPP row shifts give each lane its 16-lane row's prefix, then the totals of
// the rows below its own (read as scalars) are added.
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    v += dpp_u32<0x111>(v, 0u);  // row_shr:1
    v += dpp_u32<0x112>(v, 0u);  // row_shr:2
    v += dpp_u32<0x114>(v, 0u);  // row_shr:4
    v += dpp_u32<0x118>(v, 0u);  // row_shr:8
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
    const uint32_t row = lane_id() >> 4;
    return v + (row > 0 ? r0 : 0u) + (row > 1 ? r1 : 0u) + (row > 2 ? r2 : 0u);
}

__device__ __forceinline__ int wave_min_i32(int v) {
    constexpr uint32_t big = 0x7fffffffu;
    v = min(v, (int)dpp_u32<0x111>((uint32_t)v, big));
    v = min(v, (int)dpp_u32<0x112>((uint32_t)v, big));
    v = min(v, (int)dpp_u32<0x114>((uint32_t)v, big));
    v = min(v, (int)dpp_u32<0x118>((uint32_t)v, big));
    int m = __builtin_amdgcn_readlane(v, 15);
#pragma unroll
    for (int r = 31; r < 64; r += 16) m = min(m, __builtin_amdgcn_readlane(v, r));
    return m;
}

// A lane's running sum of cell masses: a 192-bit integer (a2:a1:a0) in
// units of the binary position 32*base (base < 0: empty), base one digit
// below the first mass it took.  A mass m 2^s (s its bit position, 53-bit m)
// is added exactly when 0 <= s - 32 base <= 127; others go to the counters d
// directly.  sa_acc_flush reduces the wave's sums (DPP wave sums) into the
// wave's digit window.
struct SaAcc {
    int base = -1;
    uint64_t a0 = 0, a1 = 0, a2 = 0;
};

__device__ __forceinline__ void sa_acc_add(SaAcc& a, unsigned long long* d, double x) {
    if (x == 0.0) return;
    if (!(x > 0.0) || !isfinite(x)) {
        atomicOr(&d[SA_DIGITS], 1ull);
        return;
    }
    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    const int e = (int)((bits >> 52) & 0x7ff);
    const uint64_t frac = bits & ((1ull << 52) - 1ull);
    const uint64_t mant = e ? (frac | (1ull << 52)) : frac;
    const int shift = e ? e - 1 : 0;
    if (a.base < 0) a.base = (shift >> 5) > 0 ? (shift >> 5) - 1 : 0;
    const int s = shift - 32 * a.base;
    if (s >= 0 && s <= 127) {
        const int off = s & 63;
        const uint64_t lo = mant << off;
        const uint64_t hi = off ? mant >> (64 - off) : 0ull;
        if (s < 64) {
            a.a0 += lo;
            const uint64_t c0 = a.a0 < lo ? 1ull : 0ull;
            const uint64_t t = a.a1 + hi;
            const uint64_t c1 = t < hi ? 1ull : 0ull;
            a.a1 = t + c0;
            a.a2 += c1 + (a.a1 < c0 ? 1ull : 0ull);
        } else {
            a.a1 += lo;
            a.a2 += hi + (a.a1 < lo ? 1ull : 0ull);
        }
    } else {
        const int limb = shift >> 5, o = shift & 31;
        const uint64_t l = mant << o, h = o ? mant >> (64 - o) : 0ull;
        if ((uint32_t)l) atomicAdd(&d[limb], (unsigned long long)(uint32_t)l);
        if ((uint32_t)(l >> 32) && limb + 1 < SA_DIGITS) atomicAdd(&d[limb + 1], (unsigned long long)(uint32_t)(l >> 32));
        if (h && limb + 2 < SA_DIGITS) atomicAdd(&d[limb + 2], (unsigned long long)h);
    }
}

// Every lane of the wave calls this (the wave reductions need all lanes): the
// wave's window (base, SA_WAVE_WIN digits) goes to out from lanes
// 0..SA_WAVE_WIN; lanes whose sums start more than 2 digits above the wave's
// lowest add theirs to d directly.
__device__ __forceinline__ void sa_acc_flush(const SaAcc& a, unsigned long long* d, unsigned long long* out) {
    const int b0 = wave_min_i32(a.base < 0 ? 0x7fffffff : a.base);
    const uint32_t lane = lane_id();
    if (b0 == 0x7fffffff) {  // no lane holds a mass
        if (lane == 0) out[0] = ~0ull;
        return;
    }
    const uint32_t w[6] = {(uint32_t)a.a0, (uint32_t)(a.a0 >> 32), (uint32_t)a.a1, (uint32_t)(a.a1 >> 32),
                           (uint32_t)a.a2, (uint32_t)(a.a2 >> 32)};
    // lanes whose base is within 2 digits of the wave's lowest share an
    // 8-digit window; the others add their own digits directly
    const int k = a.base - b0;
    const bool shared = a.base >= 0 && k <= 2;
    if (a.base >= 0 && !shared) {
        for (int j = 0; j < 6; ++j)
            if (w[j] && a.base + j < SA_DIGITS) atomicAdd(&d[a.base + j], (unsigned long long)w[j]);
    }
    static_assert(SA_WAVE_WIN == 8, "window of the 6-digit lane sums at offsets 0..2");
    uint64_t word = (uint64_t)b0;
#pragma unroll
    for (int j = 0; j < SA_WAVE_WIN; ++j) {
        uint64_t v = 0;
        if (shared) {
#pragma unroll
            for (int t = 0; t < 6; ++t)
                if (t + k == j) v = w[t];
        }
        v = wave_sum_u64(v);
        if (lane == (uint32_t)j + 1) word = b0 + j < SA_DIGITS ? v : 0ull;
    }
    if (lane <= (uint32_t)SA_WAVE_WIN) out[lane] = word;
}

// Q * unit * 2^-32 exactly as SA_WAVE_WIN 32-bit digits from digit limb, Q =
// (q1 2^64 + q0) an exact integer sum of 32.32 fixed-point cells (q1 < 2^32).
// False when unit is not a normal binary64 of at least 2^-1042 (the sum is
// then flagged invalid).
__device__ __forceinline__ bool sa_fixed_digits(uint64_t q0, uint64_t q1, double unit, int& limb,
                                                uint32_t (&dig)[SA_WAVE_WIN]) {
    const uint64_t bits = (uint64_t)__double_as_longlong(unit);
    const int e = (int)((bits >> 52) & 0x7ff);
    const int shift = e - 1 - 32;
    if (e == 0 || e == 0x7ff || shift < 0) return false;
    const uint64_t m = (bits & ((1ull << 52) - 1ull)) | (1ull << 52);
    // P = Q * m as three 64-bit words
    const uint64_t a_lo = q0 * m, a_hi = __umul64hi(q0, m);
    const uint64_t b_lo = q1 * m, b_hi = __umul64hi(q1, m);
    const uint64_t p1 = a_hi + b_lo;
    const uint64_t p2 = b_hi + (p1 < a_hi ? 1ull : 0ull);
    const uint32_t v[6] = {(uint32_t)a_lo, (uint32_t)(a_lo >> 32), (uint32_t)p1, (uint32_t)(p1 >> 32),
                           (uint32_t)p2, (uint32_t)(p2 >> 32)};
    // the product shifted by `shift` bits: seven 32-bit digits from shift / 32
    limb = shift >> 5;
    const int off = shift & 31;
#pragma unroll
    for (int k = 0; k < SA_WAVE_WIN; ++k) {
        const uint64_t lo = k < 6 ? (uint64_t)v[k] << off : 0ull;
        const uint64_t hi = (k > 0 && k <= 6 && off) ? (uint64_t)v[k - 1] >> (32 - off) : 0ull;
        dig[k] = (uint32_t)((lo | hi) & 0xffffffffull);
    }
    return true;
}

struct DepParams {
    Geom g;
    const double* x;
    const double* y;
    const double* z;
    const double* masses;   // nullptr -> uniform
    double uniform_mass;
    double pos_shift;         // positions are deposited as x - pos_shift (0: as given)
    const uint32_t* offsets;  // bucket offsets (nb+1)
    const uint32_t* order;    // bucketed particle indices
    const DepRecord64* rec64; // binary64 deposit records in bucket order
    const double* rec_mass64; // binary64 masses in bucket order (explicit masses, or signed adjoint values)
    uint64_t n;               // particle count (bound of every order entry)
    double mean_mass;         // M / N_cells
    double inv_volume;        // 1 / dx^3
    double total_mass;        // M (exact input mass)
    double cells_real;        // N_cells as double
    double centred_scale;     // inv_volume / N_cells (normalize path, unused)
    uint64_t norm_mant;       // 1/(N_cells dx^3) as a 64-bit mantissa ...
    int norm_exp;             // ... times 2^norm_exp (the reference holds it in long double)
    int direct;               // 1: (m - mean) * inv_volume ; 0: fma-centred path
    int raw;                  // 1: write the deposited sum itself (adjoint deposits of signed values)
    unsigned long long* mass_digits;  // per box SA_WORDS counters of the exact deposited mass
    unsigned long long* mass_waves;   // per tile and wave: binary64 its digit window (SA_WAVE_WORDS),
                                      // binary32 the two digit sums of its cells
    uint64_t* scratch;        // global fallback keys
    uint64_t scratch_cap;
    unsigned long long* scratch_cursor;
    uint32_t* err;
};

// Key layout: ext cell (13 bits) << 51 | particle index (32 bits) << 19 | slot (19 bits)
__device__ __forceinline__ uint32_t key_cell(uint64_t k) { return (uint32_t)(k >> 51); }
__device__ __forceinline__ uint32_t key_index(uint64_t k) { return (uint32_t)((k >> 19) & 0xffffffffull); }
__device__ __forceinline__ uint32_t key_slot(uint64_t k) { return (uint32_t)(k & 0x7ffffull); }

// |c| * S with S = s_mant * 2^s_exp (64-bit mantissa), rounded first to a
// 64-bit significand and then to binary64: the x87 long-double product and cast
// of the reference's scaled density normalisation, emulated exactly with a
// 128-bit integer product.
__device__ __forceinline__ double scaled_product_ld(double c, uint64_t s_mant, int s_exp) {
    int ea = 0;
    const double fa = frexp(fabs(c), &ea);
    const uint64_t ma = (uint64_t)ldexp(fa, 53);
    const unsigned __int128 q = (unsigned __int128)ma * (unsigned __int128)s_mant;
    const uint64_t qhi = (uint64_t)(q >> 64);
    const int nb = qhi ? 128 - __clzll(qhi) : 64 - __clzll((uint64_t)q);
    int sh = nb - 64;
    uint64_t m64 = (uint64_t)(q >> sh);
    if (sh > 0) {
        const unsigned __int128 rem = q & ((((unsigned __int128)1) << sh) - 1);
        const unsigned __int128 half = ((unsigned __int128)1) << (sh - 1);
        if (rem > half || (rem == half && (m64 & 1ull))) {
            m64 += 1ull;
            if (m64 == 0ull) { m64 = 1ull << 63; sh += 1; }
        }
    }
    uint64_t m53 = m64 >> 11;
    const uint64_t r2 = m64 & 0x7ffull;
    if (r2 > 0x400ull || (r2 == 0x400ull && (m53 & 1ull))) m53 += 1ull;
    return copysign(ldexp((double)m53, (ea - 53) + s_exp + sh + 11), c);
}

template <typename T>
__device__ __forceinline__ T contrast_of(T sum, const DepParams& P) {
    const double m = (double)sum;
    if (P.raw) return (T)m;
    double v;
    if (P.direct) {
        v = (m - P.mean_mass) * P.inv_volume;
    } else {
        const double c = fma(m, P.cells_real, -P.total_mass);
        v = c == 0.0 ? 0.0 : scaled_product_ld(c, P.norm_mant, P.norm_exp);
    }
    return (T)v;
}

// Binary32 deposit: (m - M/N_c) / dx^3 in binary64, rounded to binary32.  It
// differs from the reference's fma-centred form (taken by contrast_of when
// M/N_c is not exact) by about one binary64 ulp, far below the binary32
// rounding of the cell, and avoids that form's 64-bit mantissa arithmetic.
__device__ __forceinline__ float contrast_fp32(double m, const DepParams& P) {
    if (P.raw) return (float)m;
    return (float)((m - P.mean_mass) * P.inv_volume);
}

template <typename T, bool LDS_PAYLOAD, int TE, typename KP, typename FP>
__device__ __forceinline__ void deposit_cells(const DepParams& P, T* __restrict__ mesh, int tx,
                                              int ty, int tz, KP keys, uint32_t M,
                                              const uint32_t* start, FP fx, FP fy, FP fz, FP fm,
                                              unsigned long long* sa, unsigned long long* wave_out) {
    const Geom& g = P.g;
    const int TT = TE > 0 ? TE : g.T;
    const int Te = TT + 1;
    const int cells = TT * TT * TT;
    SaAcc win;
    for (int c = threadIdx.x; c < cells; c += blockDim.x) {
        const int lz = c % TT, ly = (c / TT) % TT, lx = c / (TT * TT);
        T sum = (T)0;
        for (int phase = 0; phase < 2; ++phase) {
            const int ex = lx + 1 - phase;  // phase 0: base x == cell x; phase 1: base x == cell x - 1
            uint32_t cur[4], end[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int oy = q >> 1, oz = q & 1;
                const int cell = (ex * Te + (ly + 1 - oy)) * Te + (lz + 1 - oz);
                cur[q] = start[cell];
                end[q] = start[cell + 1];
            }
            if constexpr (LDS_PAYLOAD) {
                // at most one record per source cell (the common case): order the
                // up to four contributions by particle index with a sorting network
                // and add them in that order
                bool simple = true;
#pragma unroll
                for (int q = 0; q < 4; ++q) simple = simple && end[q] - cur[q] <= 1u;
                if (simple) {
                    uint32_t idx[4], slot[4];
                    int src[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const bool have = cur[q] < end[q];
                        const uint64_t k = have ? keys[cur[q]] : 0ull;
                        idx[q] = have ? key_index(k) : 0xffffffffu;  // empty lists sort last
                        slot[q] = key_slot(k);
                        src[q] = q;
                    }
                    auto cswap = [&](int i, int j) {
                        if (idx[j] < idx[i]) {
                            const uint32_t ti = idx[i], ts = slot[i];
                            const int tq = src[i];
                            idx[i] = idx[j]; slot[i] = slot[j]; src[i] = src[j];
                            idx[j] = ti; slot[j] = ts; src[j] = tq;
                        }
                    };
                    cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if (idx[i] == 0xffffffffu) break;  // the rest are empty too
                        const uint32_t sl = slot[i];
                        const T ffx = fx[sl], ffy = fy[sl], ffz = fz[sl];
                        const T mm = P.masses ? fm[sl] : (T)P.uniform_mass;
                        const int oy = src[i] >> 1, oz = src[i] & 1;
                        const T xw = phase == 0 ? (T)1 - ffx : ffx;
                        const T wy = oy ? ffy : (T)1 - ffy;
                        const T wz = oz ? ffz : (T)1 - ffz;
                        sum += mm * xw * wy * wz;
                    }
                    continue;
                }
            }
            // merge the four source cells by particle index
            while (true) {
                int best = -1;
                uint32_t best_idx = 0xffffffffu;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (cur[q] < end[q]) {
                        const uint32_t idx = key_index(keys[cur[q]]);
                        if (best < 0 || idx < best_idx) { best = q; best_idx = idx; }
                    }
                }
                if (best < 0) break;
                const uint64_t k = keys[cur[best]];
                ++cur[best];
                T ffx, ffy, ffz, mm;
                if constexpr (LDS_PAYLOAD) {
                    const uint32_t s = key_slot(k);
                    ffx = fx[s]; ffy = fy[s]; ffz = fz[s];
                    mm = P.masses ? fm[s] : (T)P.uniform_mass;
                } else {
                    const uint32_t pidx = key_index(k);
                    int b; double f;
                    cic_axis(P.x[pidx] - P.pos_shift, g, b, f); ffx = (T)f;
                    cic_axis(P.y[pidx] - P.pos_shift, g, b, f); ffy = (T)f;
                    cic_axis(P.z[pidx] - P.pos_shift, g, b, f); ffz = (T)f;
                    mm = P.masses ? (T)P.masses[pidx] : (T)P.uniform_mass;
                }
                const int oy = best >> 1, oz = best & 1;
                const T xw = phase == 0 ? (T)1 - ffx : ffx;
                const T wy = oy ? ffy : (T)1 - ffy;
                const T wz = oz ? ffz : (T)1 - ffz;
                sum += mm * xw * wy * wz;
            }
        }
        if (!P.raw) sa_acc_add(win, sa, (double)sum);
        const uint64_t gx = (uint64_t)(tx * g.T + lx), gy = (uint64_t)(ty * g.T + ly),
                       gz = (uint64_t)(tz * g.T + lz);
        mesh[(gx * g.N + gy) * g.N + gz] = contrast_of(sum, P);
    }
    sa_acc_flush(win, sa, wave_out);
}

// Build start[0..Te^3] (prefix positions per ext cell) from sorted keys.
template <typename KP>
__device__ __forceinline__ void build_cell_starts(KP keys, uint32_t M, uint32_t* start, int ncell) {
    for (uint32_t i = threadIdx.x; i <= M; i += blockDim.x) {
        const int c = i < M ? (int)key_cell(keys[i]) : ncell;
        const int prev = i == 0 ? -1 : (int)key_cell(keys[i - 1]);
        for (int q = prev + 1; q <= c; ++q) start[q] = i;
    }
    __syncthreads();
}

// Source ranges of a destination tile: range 0 is the own tile (all eight
// classes, contiguous); ranges 1..19 are the boundary classes c of the seven
// lower neighbour tiles at offset d in {0,1}^3 with (c & d) == d, in the order
//   (0,0) (1,1) (1,3) (1,5) (1,7) (2,2) (2,3) (2,6) (2,7) (3,3) (3,7)
//   (4,4) (4,5) (4,6) (4,7) (5,5) (5,7) (6,6) (6,7) (7,7).
// The d and c of range t are 3-bit fields t of two immediates (a per-lane
// table lookup would cost each tile a memory round trip before its offsets).
constexpr uint64_t kRangeD = 0xfb5b248da491248ull;
constexpr uint64_t kRangeC = 0xff7bf59dfcd7ac8ull;
struct RangeSlot { uint8_t d, c; };
__device__ __forceinline__ RangeSlot range_slot(int t) {
    return RangeSlot{(uint8_t)((kRangeD >> (3 * t)) & 7u), (uint8_t)((kRangeC >> (3 * t)) & 7u)};
}

// Fills rs (start in order[]), rc (prefix of counts, rc[20] = total), rd (d),
// rcls (class) with one global round trip per range, in parallel.
__device__ __forceinline__ void load_source_ranges(const uint32_t* __restrict__ offsets, int tx, int ty,
                                                   int tz, int nT, uint32_t* rs, uint32_t* rc,
                                                   uint8_t* rd, uint8_t* rcls) {
    const int t = threadIdx.x;
    if (t < 64) {  // the first wave: one lane per range, then a wave scan of the counts
        uint32_t cnt = 0;
        if (t < 20) {
            const RangeSlot sl = range_slot(t);
            const int d = sl.d;
            const int sx = (tx - ((d >> 2) & 1) + nT) % nT, sy = (ty - ((d >> 1) & 1) + nT) % nT,
                      sz = (tz - (d & 1) + nT) % nT;
            const uint32_t t8 = (((uint32_t)sx * nT + sy) * nT + sz) * 8u;
            const uint32_t lo = d == 0 ? t8 : t8 + sl.c;
            const uint32_t hi = d == 0 ? t8 + 8 : t8 + sl.c + 1;
            const uint32_t a = offsets[lo], b = offsets[hi];
            rs[t] = a;
            cnt = b - a;
            rd[t] = (uint8_t)d;
            rcls[t] = sl.c;
        }
        const uint32_t incl = wave_incl_scan_u32(cnt);
        if (t < 20) rc[t + 1] = incl;
        if (t == 0) rc[0] = 0;
    }
    __syncthreads();
}

// Exclusive scan of a[0..count) in LDS, in place, by all threads of the
// workgroup (each thread owns a contiguous chunk): a scan inside each
// wavefront, then the wavefront totals (part: one word per wavefront).
__device__ __forceinline__ void wg_exclusive_scan_lds(uint32_t* a, int count, uint32_t* part) {
    const int nt = (int)blockDim.x;
    const int per = (count + nt - 1) / nt;
    const int b = (int)threadIdx.x * per;
    const int e = b + per < count ? b + per : count;
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += a[i];
    const int lane = (int)lane_id(), wave = (int)threadIdx.x / 64;
    const uint32_t incl = wave_incl_scan_u32(s);
    if (lane == 63) part[wave] = incl;
    __syncthreads();
    uint32_t run = incl - s;
    for (int w = 0; w < wave; ++w) run += part[w];
    for (int i = b; i < e; ++i) {
        const uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
    __syncthreads();
}

template <typename T, int TE = 0>
__global__ void __launch_bounds__(DEP_THREADS) k_deposit(DepParams P, T* __restrict__ mesh) {
    using Cfg = DepCfg<T>;
    constexpr int CAP = Cfg::CAP;
    constexpr int PER = CAP / DEP_THREADS;  // records per thread on the LDS path
    static_assert(CAP % DEP_THREADS == 0, "LDS capacity must be a multiple of the workgroup size");
    __shared__ uint64_t sh_keys[CAP];
    __shared__ T sh_fx[CAP], sh_fy[CAP], sh_fz[CAP], sh_m[CAP];
    __shared__ uint32_t sh_start[DEP_EXT + 1];
    __shared__ uint32_t sh_rs[DEP_NRANGES], sh_rc[DEP_NRANGES + 1];
    __shared__ uint8_t sh_rd[DEP_NRANGES];
    __shared__ uint64_t sh_gbase;
    __shared__ uint32_t sh_part[DEP_THREADS / 64];

    const Geom& g = P.g;
    const int nT = g.nT;
    const uint32_t ntiles = (uint32_t)nT * nT * nT;
    const uint32_t gtile = xcd_block(blockIdx.x, gridDim.x);  // box-major tile index
    if (gtile >= ntiles * g.nbox) return;
    const uint32_t box = gtile / ntiles, tile = gtile % ntiles;
    const int tz = tile % nT, ty = (tile / nT) % nT, tx = tile / (nT * nT);
    const int TT = TE > 0 ? TE : g.T;
    const int Te = TT + 1;
    const int ncell = Te * Te * Te;
    mesh += (uint64_t)box * g.N * g.N * g.N;
    static_assert(DEP_THREADS / 64 == SA_WAVES_PER_TILE, "one mass window per wave");
    // exact cell-mass sum: each wave's window, plus the box's counters for the rest
    unsigned long long* box_digits = P.mass_digits + (uint64_t)box * SA_WORDS;
    unsigned long long* wave_out =
        P.mass_waves + ((uint64_t)gtile * SA_WAVES_PER_TILE + threadIdx.x / 64) * SA_WAVE_WORDS;

    __shared__ uint8_t sh_rcls[DEP_NRANGES];
    load_source_ranges(P.offsets + (uint64_t)box * buckets_per_box(g), tx, ty, tz, nT, sh_rs, sh_rc, sh_rd,
                       sh_rcls);
    const uint32_t M = sh_rc[DEP_NRANGES];
    const bool use_lds = M <= (uint32_t)CAP;

    uint64_t* gkeys = nullptr;
    if (!use_lds) {
        if (threadIdx.x == 0) sh_gbase = atomicAdd(P.scratch_cursor, (unsigned long long)M);
        __syncthreads();
        if (sh_gbase + M > P.scratch_cap) {
            if (threadIdx.x == 0) atomicOr(P.err, ERR_SCRATCH);
            if (lane_id() == 0) wave_out[0] = ~0ull;  // this tile's waves hold no mass
            return;
        }
        gkeys = P.scratch + sh_gbase;
    } else {
        for (int i = threadIdx.x; i <= ncell; i += blockDim.x) sh_start[i] = 0u;
        __syncthreads();
    }

    // Load the records (read contiguously per source range): ext cell relative
    // to this tile, particle index, binary64 fractions.  A neighbour record
    // sits on the last cell of its tile along each shifted axis.  On the LDS
    // path each thread keeps its keys in registers and counts them per cell.
    uint64_t mykey[PER];
    uint32_t myrank[PER];
    auto load = [&](uint32_t v) -> uint64_t {
        int r = 0;
        while (sh_rc[r + 1] <= v) ++r;
        const uint32_t slot = sh_rs[r] + (v - sh_rc[r]);
        const DepRecord64 rc = P.rec64[slot];
        const int d = sh_rd[r];
        const int TL = g.T;
        const uint32_t pidx = rc.index < P.n ? rc.index : 0u;  // guards a slot a failed scatter left unwritten
        const int cz = (int)(rc.cell % (uint32_t)TL), cy = (int)((rc.cell / (uint32_t)TL) % (uint32_t)TL),
                  cx = (int)(rc.cell / (uint32_t)(TL * TL));
        int ex = (d & 4) ? 0 : cx + 1;
        int ey = (d & 2) ? 0 : cy + 1;
        int ez = (d & 1) ? 0 : cz + 1;
        if (rc.index >= P.n || rc.cell >= (uint32_t)(TL * TL * TL) || ((d & 4) && cx != TL - 1) ||
            ((d & 2) && cy != TL - 1) || ((d & 1) && cz != TL - 1)) {
            atomicOr(P.err, ERR_COORDINATE);
            ex = ex > TL ? TL : ex;
            ey = ey > TL ? TL : ey;
            ez = ez > TL ? TL : ez;
        }
        const uint32_t cell = (uint32_t)((ex * Te + ey) * Te + ez);
        if (use_lds) {
            sh_fx[v] = (T)rc.fx; sh_fy[v] = (T)rc.fy; sh_fz[v] = (T)rc.fz;
            if (P.masses) sh_m[v] = (T)P.rec_mass64[slot];
        }
        return ((uint64_t)cell << 51) | ((uint64_t)pidx << 19) | (use_lds ? v : 0u);
    };
    if (use_lds) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t v = threadIdx.x + (uint32_t)j * DEP_THREADS;
            mykey[j] = 0;
            myrank[j] = 0;
            if (v < M) {
                mykey[j] = load(v);
                myrank[j] = atomicAdd(&sh_start[key_cell(mykey[j])], 1u);
            }
        }
    } else {
        for (uint32_t v = threadIdx.x; v < M; v += blockDim.x) gkeys[v] = load(v);
    }
    __threadfence_block();
    __syncthreads();

    if (use_lds) {
        // counting sort by cell, then each cell's few records by particle index
        wg_exclusive_scan_lds(sh_start, ncell + 1, sh_part);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t v = threadIdx.x + (uint32_t)j * DEP_THREADS;
            if (v < M) sh_keys[sh_start[key_cell(mykey[j])] + myrank[j]] = mykey[j];
        }
        __syncthreads();
        for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
            const uint32_t b = sh_start[c], e = sh_start[c + 1];
            for (uint32_t i = b + 1; i < e; ++i) {
                const uint64_t k = sh_keys[i];
                uint32_t j = i;
                while (j > b && sh_keys[j - 1] > k) {
                    sh_keys[j] = sh_keys[j - 1];
                    --j;
                }
                sh_keys[j] = k;
            }
        }
        __syncthreads();
        deposit_cells<T, true, TE>(P, mesh, tx, ty, tz, sh_keys, M, sh_start, sh_fx, sh_fy, sh_fz,
                               sh_m, box_digits, wave_out);
    } else {
        wg_sort_u64(gkeys, M);
        build_cell_starts(gkeys, M, sh_start, ncell);
        deposit_cells<T, false, TE>(P, mesh, tx, ty, tz, gkeys, M, sh_start, sh_fx, sh_fy, sh_fz,
                                sh_m, box_digits, wave_out);
    }

}

template __global__ void k_deposit<double, 0>(DepParams, double*);
template __global__ void k_deposit<double, 16>(DepParams, double*);

// ---------------------------------------------------------------------------
// Binary32 deposit: fixed-point accumulation in LDS.
//
// Each CIC contribution m*wx*wy*wz is converted to an unsigned integer in
// units of 2^-32 of the mass unit (the uniform particle mass, or the largest
// explicit mass) and added with an integer LDS atomic.  Integer addition is
// associative, so every cell sum is bit-identical whatever the scheduling: the
// deposit is deterministic without a per-cell order and without float
// atomics.  The resolution (2^-32 of a particle mass per contribution) is far
// below binary32 rounding; a cell holds up to 2^31 particle masses.
// ---------------------------------------------------------------------------
constexpr int DEPFX_THREADS = 512;
constexpr int DEPFX_MAX_CELLS = DEP_MAX_T * DEP_MAX_T * DEP_MAX_T;

struct DepFxParams {
    DepParams base;
    double mass_unit;               // uniform mass, or max explicit mass
    double fixed_to_mass;           // mass_unit * 2^-32
    const DepRecord* rec;           // deposit records by bucket
    const float* rec_mass;          // explicit masses / mass_unit, as rec (or null)
    const uint32_t* start;          // nb+1: first slot of each bucket (dense offsets or a capacity layout)
    const uint32_t* counts;         // records per bucket
    const uint32_t* ovf_start;      // nb+1 prefix of the overflow records (or null: none)
    uint64_t nb;                    // buckets over all boxes (start[nb] = first overflow slot)
};

// The 27 buckets that feed a tile: its own eight classes, then the 19
// neighbour (d, c) slots 1..19 of range_slot.  Each gives a main range (its
// slots, at most its capacity) and an overflow range.
constexpr int DEPFX_BUCKETS = 27;
constexpr int DEPFX_NRANGES = 2 * DEPFX_BUCKETS;
static_assert(DEPFX_NRANGES <= 64, "one lane per range");

// The non-empty ranges in order: rs their first slot, rc the prefix of their
// counts (rc[DEPFX_NRANGES] = total), rd the neighbour offset d, *nranges
// their number; one global round trip, the first wave.
__device__ __forceinline__ void load_fx_ranges(const DepFxParams& Q, uint32_t box_base, int tx, int ty, int tz,
                                               int nT, uint32_t* rs, uint32_t* rc, uint8_t* rd,
                                               uint32_t* nranges) {
    const int t = threadIdx.x;
    if (t < 64) {
        uint32_t cnt = 0, rs_lane = 0;
        uint8_t rd_lane = 0;
        if (t < DEPFX_NRANGES) {
            const int j = t % DEPFX_BUCKETS;
            const bool ovf = t >= DEPFX_BUCKETS;
            int d = 0, cls = j;
            if (j >= 8) {
                const RangeSlot sl = range_slot(j - 7);
                d = sl.d;
                cls = sl.c;
            }
            const int sx = (tx - ((d >> 2) & 1) + nT) % nT, sy = (ty - ((d >> 1) & 1) + nT) % nT,
                      sz = (tz - (d & 1) + nT) % nT;
            const uint32_t b = box_base + (((uint32_t)sx * nT + sy) * nT + sz) * 8u + (uint32_t)cls;
            uint32_t first = 0;
            if (!ovf) {
                const uint32_t a = Q.start[b], e = Q.start[b + 1], c = Q.counts[b];
                first = a;
                cnt = c < e - a ? c : e - a;
            } else if (Q.ovf_start) {
                const uint32_t a = Q.ovf_start[b], e = Q.ovf_start[b + 1];
                first = Q.start[Q.nb] + a;
                cnt = e - a;
            }
            rs_lane = first;
            rd_lane = (uint8_t)d;
        }
        // the non-empty ranges, compacted in order: rc[0..R] is their count
        // prefix and rc[R..DEPFX_NRANGES] repeats the total
        const uint64_t live = __ballot(cnt != 0u);
        const uint32_t at = (uint32_t)__popcll(live & ((1ull << t) - 1ull));
        const uint32_t R = (uint32_t)__popcll(live);
        const uint32_t incl = wave_incl_scan_u32(cnt);
        if (cnt != 0u) {
            rs[at] = rs_lane;
            rd[at] = rd_lane;
            rc[at + 1] = incl;
        }
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        if (t == 0) rc[0] = 0;
        if ((uint32_t)t > R && t <= DEPFX_NRANGES) rc[t] = total;
        if (t == 0) *nranges = R;
    }
    __syncthreads();
}

// The range of record v: the last r < R with rc[r] <= v (binary search).
__device__ __forceinline__ int fx_range_of(const uint32_t* rc, uint32_t R, uint32_t v) {
    int lo = 0, hi = (int)R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rc[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// TE: compile-time tile edge (16, the edge of every mesh divisible by 16:
// cell index arithmetic becomes shifts), or 0 to read it from the geometry.
template <int TE>
__global__ void __launch_bounds__(DEPFX_THREADS) k_deposit_fixed(DepFxParams Q, float* __restrict__ mesh) {
    const DepParams& P = Q.base;
    __shared__ unsigned long long acc[DEPFX_MAX_CELLS];
    __shared__ uint32_t sh_rs[DEPFX_NRANGES], sh_rc[DEPFX_NRANGES + 1];
    __shared__ uint8_t sh_rd[DEPFX_NRANGES];
    __shared__ uint32_t sh_nr;

    const Geom& g = P.g;
    const int nT = g.nT;
    const int T = TE > 0 ? TE : g.T;
    const uint32_t ntiles = (uint32_t)nT * nT * nT;
    const uint32_t gtile = xcd_block(blockIdx.x, gridDim.x);  // box-major tile index
    if (gtile >= ntiles * g.nbox) return;
    const uint32_t box = gtile / ntiles, tile = gtile % ntiles;
    const int tz = tile % nT, ty = (tile / nT) % nT, tx = tile / (nT * nT);
    const int cells = T * T * T;
    mesh += (uint64_t)box * g.N * g.N * g.N;

    for (int c = threadIdx.x; c < cells; c += blockDim.x) acc[c] = 0ull;
    load_fx_ranges(Q, box * buckets_per_box(g), tx, ty, tz, nT, sh_rs, sh_rc, sh_rd, &sh_nr);
    const uint32_t M = sh_rc[DEPFX_NRANGES];
    const uint32_t nr = sh_nr;
    int r = threadIdx.x < M ? fx_range_of(sh_rc, nr, threadIdx.x) : 0;
    bool bad = false;
    // records in pairs: both loads are issued before the first atomic, so a
    // thread with two records waits for one memory round trip, not two
    for (uint32_t v0 = threadIdx.x; v0 < M; v0 += 2 * blockDim.x) {
        DepRecord recs[2];
        float ms[2];
        int ds[2];
        bool have[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t v = v0 + (uint32_t)h * blockDim.x;
            have[h] = v < M;
            if (have[h]) {
                while (sh_rc[r + 1] <= v) ++r;
                const uint32_t slot = sh_rs[r] + (v - sh_rc[r]);
                recs[h] = Q.rec[slot];
                ms[h] = Q.rec_mass ? Q.rec_mass[slot] : 1.0f;
                ds[h] = sh_rd[r];
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (!have[h]) continue;
            const DepRecord rec = recs[h];
            const int d = ds[h];
            const int c = (int)rec.cell;
            const int cz_ = c % T, cy_ = (c / T) % T, cx_ = c / (T * T);
            // a neighbour record sits just below this tile along its shifted axes
            const int lx = (d & 4) ? -1 : cx_;
            const int ly = (d & 2) ? -1 : cy_;
            const int lz = (d & 1) ? -1 : cz_;
            if (((d & 4) && cx_ != T - 1) || ((d & 2) && cy_ != T - 1) || ((d & 1) && cz_ != T - 1)) {
                bad = true;
                continue;
            }
            const float m = ms[h];
            const float wx[2] = {1.0f - rec.fx, rec.fx};
            const float wy[2] = {1.0f - rec.fy, rec.fy};
            const float wz[2] = {1.0f - rec.fz, rec.fz};
#pragma unroll
            for (int ox = 0; ox < 2; ++ox) {
                const int cx = lx + ox;
                if (cx < 0 || cx >= T) continue;
#pragma unroll
                for (int oy = 0; oy < 2; ++oy) {
                    const int cy = ly + oy;
                    if (cy < 0 || cy >= T) continue;
                    const float mxy = m * wx[ox] * wy[oy];
#pragma unroll
                    for (int oz = 0; oz < 2; ++oz) {
                        const int cz = lz + oz;
                        if (cz < 0 || cz >= T) continue;
                        const float w = mxy * wz[oz];
                        // signed 32.32 fixed point (two's complement sums wrap consistently)
                        const unsigned long long q = (unsigned long long)__float2ll_rn(w * 4294967296.0f);
                        atomicAdd(&acc[(cx * T + cy) * T + cz], q);
                    }
                }
            }
        }
    }
    if (bad) atomicOr(P.err, ERR_COORDINATE);
    __syncthreads();

    // the tile's exact mass: its cells' 32.32 integers summed as two 32-bit digit sums
    uint64_t dig0 = 0, dig1 = 0;
    if constexpr (TE > 0 && TE % 4 == 0) {
        // four z-adjacent cells per lane, one 16-byte store (a tile's z-runs
        // start at multiples of TE cells)
        typedef float F4 __attribute__((ext_vector_type(4)));
        for (int c4 = threadIdx.x; c4 < cells / 4; c4 += blockDim.x) {
            const int c = 4 * c4;
            const int lz = c % T, ly = (c / T) % T, lx = c / (T * T);
            float o[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long q = (long long)acc[c + j];
                dig0 += (uint64_t)q & 0xffffffffull;
                dig1 += (uint64_t)q >> 32;
                o[j] = contrast_fp32((double)q * Q.fixed_to_mass, P);
            }
            const uint64_t gx = (uint64_t)(tx * T + lx), gy = (uint64_t)(ty * T + ly), gz = (uint64_t)(tz * T + lz);
            *reinterpret_cast<F4*>(&mesh[(gx * g.N + gy) * g.N + gz]) = F4{o[0], o[1], o[2], o[3]};
        }
    } else {
        for (int c = threadIdx.x; c < cells; c += blockDim.x) {
            const long long q = (long long)acc[c];
            dig0 += (uint64_t)q & 0xffffffffull;
            dig1 += (uint64_t)q >> 32;
            const double mcell = (double)q * Q.fixed_to_mass;
            const int lz = c % T, ly = (c / T) % T, lx = c / (T * T);
            const uint64_t gx = (uint64_t)(tx * T + lx), gy = (uint64_t)(ty * T + ly), gz = (uint64_t)(tz * T + lz);
            mesh[(gx * g.N + gy) * g.N + gz] = contrast_fp32(mcell, P);
        }
    }
    // each wave's two digit sums (below 2^41), for k_mass_reduce_fixed
    static_assert(DEPFX_THREADS / 64 == SA_WAVES_PER_TILE, "one mass pair per wave");
    if (P.raw) dig0 = dig1 = 0;  // signed cells: no mass identity
    const uint64_t s0 = wave_sum_u64(dig0), s1 = wave_sum_u64(dig1);
    if (lane_id() == 0) {
        typedef unsigned long long U2 __attribute__((ext_vector_type(2)));
        *reinterpret_cast<U2*>(P.mass_waves + ((uint64_t)gtile * SA_WAVES_PER_TILE + threadIdx.x / 64) * 2) =
            U2{s0, s1};
    }
}

// Zero fill (used instead of memset nodes in code that callers may capture
// into a HIP graph).
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Step factors written on the stream ahead of a graph replay.
__global__ void k_set_factors(double* __restrict__ f, double k_prev, double k_next, double d_next) {
    if (threadIdx.x == 0) {
        f[0] = k_prev;
        f[1] = k_next;
        f[2] = d_next;
    }
}

// Per box (grid row): the waves' digit windows added up, windows of nearby
// bases in a thread's registers first, then LDS counters, then one atomic per
// non-zero counter and workgroup into the box's counters.  Integer sums: the
// total is exact whatever the order.
__global__ void __launch_bounds__(256) k_mass_reduce(const unsigned long long* __restrict__ waves,
                                                     uint64_t waves_per_box,
                                                     unsigned long long* __restrict__ digits) {
    __shared__ unsigned long long cnt[SA_DIGITS];
    const uint64_t box = blockIdx.y;
    for (int w = threadIdx.x; w < SA_DIGITS; w += blockDim.x) cnt[w] = 0ull;
    __syncthreads();
    constexpr int AW = SA_WAVE_WIN + 2;
    long long tb = -1;
    uint64_t acc[AW];
#pragma unroll
    for (int j = 0; j < AW; ++j) acc[j] = 0ull;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < waves_per_box;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long* ww = waves + (box * waves_per_box + t) * SA_WAVE_WORDS;
        const long long b = (long long)ww[0];
        if (b < 0) continue;
        if (tb < 0) tb = b;
        const long long k = b - tb;
        if (k >= 0 && k <= 2) {
#pragma unroll
            for (int j = 0; j < SA_WAVE_WIN; ++j) {
                const uint64_t v = ww[1 + j];
#pragma unroll
                for (int o = 0; o <= 2; ++o)
                    if (k == o) acc[j + o] += v;
            }
        } else {
            for (int j = 0; j < SA_WAVE_WIN; ++j) {
                const unsigned long long v = ww[1 + j];
                if (v && b + j < SA_DIGITS) atomicAdd(&cnt[b + j], v);
            }
        }
    }
    if (tb >= 0) {
#pragma unroll
        for (int j = 0; j < AW; ++j)
            if (acc[j] && tb + j < SA_DIGITS) atomicAdd(&cnt[tb + j], (unsigned long long)acc[j]);
    }
    __syncthreads();
    for (int w = threadIdx.x; w < SA_DIGITS; w += blockDim.x)
        if (cnt[w]) atomicAdd(&digits[box * SA_WORDS + w], cnt[w]);
}

// Binary32 deposit: per box (grid row), the waves' digit-sum pairs (s0, s1)
// added up per workgroup (below 2^54 each: at most 2^13 waves of under 2^41
// per thread at the grid's 2^16 threads per box and a mesh of 4096^3), and
// the workgroup's exact total (s0 + s1 2^32) * unit * 2^-32 added to the box's
// counters as digits.
__global__ void __launch_bounds__(256) k_mass_reduce_fixed(const unsigned long long* __restrict__ waves,
                                                           uint64_t waves_per_box, double unit,
                                                           unsigned long long* __restrict__ digits) {
    __shared__ uint64_t part[2][4];
    const uint64_t box = blockIdx.y;
    typedef unsigned long long U2 __attribute__((ext_vector_type(2)));
    uint64_t s0 = 0, s1 = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < waves_per_box;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const U2 v = *reinterpret_cast<const U2*>(waves + (box * waves_per_box + t) * 2);
        s0 += v.x;
        s1 += v.y;
    }
    s0 = wave_sum_u64(s0);
    s1 = wave_sum_u64(s1);
    if (lane_id() == 0) {
        part[0][threadIdx.x / 64] = s0;
        part[1][threadIdx.x / 64] = s1;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint64_t a0 = 0, a1 = 0;
    for (int w = 0; w < (int)blockDim.x / 64; ++w) {
        a0 += part[0][w];
        a1 += part[1][w];
    }
    const uint64_t q0 = a0 + (a1 << 32);
    const uint64_t q1 = (a1 >> 32) + (q0 < a0 ? 1ull : 0ull);
    unsigned long long* d = digits + box * SA_WORDS;
    int limb = 0;
    uint32_t dig[SA_WAVE_WIN];
    if (!sa_fixed_digits(q0, q1, unit, limb, dig)) {
        atomicOr(&d[SA_DIGITS], 1ull);
        return;
    }
    for (int k = 0; k < SA_WAVE_WIN; ++k)
        if (dig[k] && limb + k < SA_DIGITS) atomicAdd(&d[limb + k], (unsigned long long)dig[k]);
}

// Per box: the exact deposited mass from its counters, rounded once to nearest
// with ties to even (NaN when a cell was not a finite non-negative number).
// One thread per box.
__global__ void k_mass_exact(const unsigned long long* __restrict__ digits, uint32_t nbox,
                             double* __restrict__ out) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const unsigned long long* d = digits + (uint64_t)b * SA_WORDS;
    if (d[SA_DIGITS]) {
        out[b] = __builtin_nan("");
        return;
    }
    constexpr int ND = SA_DIGITS + 2;
    uint32_t v[ND];
    uint64_t carry = 0;
    for (int i = 0; i < SA_DIGITS; ++i) {
        const uint64_t sum = d[i] + carry;  // d[i] < 2^62, carry < 2^32
        v[i] = (uint32_t)sum;
        carry = sum >> 32;
    }
    v[SA_DIGITS] = (uint32_t)carry;
    v[SA_DIGITS + 1] = (uint32_t)(carry >> 32);
    int top = -1;
    for (int i = ND - 1; i >= 0 && top < 0; --i)
        if (v[i]) top = i * 32 + (31 - __clz(v[i]));
    if (top < 0) {
        out[b] = 0.0;
        return;
    }
    // bits [lo, lo + 64) of the sum
    auto bits64 = [&](int lo) -> uint64_t {
        uint64_t r = 0;
        for (int k = 0; k < 3; ++k) {
            const int i = (lo >> 5) + k;
            if (i < 0 || i >= ND) continue;
            const int sh = k * 32 - (lo & 31);
            const uint64_t w = v[i];
            r |= sh >= 0 ? (sh < 64 ? w << sh : 0ull) : w >> (-sh);
        }
        return r;
    };
    if (top < 53) {
        out[b] = ldexp((double)(bits64(0) & ((1ull << 53) - 1ull)), -1074);
        return;
    }
    const int lo = top - 52;
    uint64_t m = bits64(lo) & ((1ull << 53) - 1ull);
    const int r = lo - 1;  // the rounding bit
    const bool round = (v[r >> 5] >> (r & 31)) & 1u;
    bool sticky = (v[r >> 5] & ((1u << (r & 31)) - 1u)) != 0u;
    for (int i = 0; i < (r >> 5) && !sticky; ++i) sticky = v[i] != 0u;
    if (round && (sticky || (m & 1ull))) ++m;
    out[b] = ldexp((double)m, lo - 1074);
}

// Deposit status on the device, so that calls which do not synchronise stay
// fail-closed: the routing-count identity of the x-plane bins
// (src/mesh/mass_assignment_deposit.inc:269-284: every particle routed exactly
// once) and the mass-conservation bound of src/math/cic_mass_conservation.cpp:
// 25-63 per box, |deposited - input| <= tol_factor * max(input, deposited)
// with a finite positive deposited mass.  One thread per box.
__global__ void k_deposit_status(const uint32_t* __restrict__ grand_total, uint64_t n,
                                 const double* __restrict__ mass_total, uint32_t nbox, double input_per_box,
                                 double tol_factor, uint32_t* __restrict__ err) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    uint32_t bits = 0;
    if (b == 0 && (uint64_t)*grand_total != n) bits |= ERR_ROUTING;
    const double dep = mass_total[b];
    const double scale = dep > input_per_box ? dep : input_per_box;
    if (!isfinite(dep) || !(dep > 0.0) || fabs(dep - input_per_box) > tol_factor * scale)
        bits |= ERR_MASS_CONSERVATION;
    if (bits) atomicOr(err, bits);
}

// ===========================================================================
// Fused centered-gradient gather + kick/drift
// ===========================================================================
// Block schedule for particle arrays stored in Lagrangian lattice order
// p = (ix*S + iy)*S + iz (the order of lattice initial conditions, which
// particles keep for the whole run).  Logical block L covers 256-particle
// chunks of z-rows; consecutive L walk a TX x TY column of lattice rows
// (all z) before the next column, so the blocks resident on one XCD touch a
// compact region of the potential that stays in its L2.  S == 0: identity.
struct LatticeTiling {
    uint32_t S, TX, TY, ZC;  // lattice side, column tile, 256-chunks per z-row
    int lS, lTX, lTY, lZC;   // log2 of each when all four are powers of two, else lS < 0
};

__device__ __forceinline__ uint32_t lattice_block(uint32_t L, const LatticeTiling& t) {
    if (t.S == 0) return L;
    if (t.lS >= 0) {  // powers of two: shifts and masks
        const int lper = t.lTX + t.lTY + t.lZC;
        const uint32_t tile = L >> lper, w = L & ((1u << lper) - 1u);
        const int ltiles_y = t.lS - t.lTY;
        const uint32_t tx = tile >> ltiles_y, ty = tile & ((1u << ltiles_y) - 1u);
        const uint32_t zc = w & ((1u << t.lZC) - 1u), dy = (w >> t.lZC) & ((1u << t.lTY) - 1u);
        const uint32_t dx = w >> (t.lZC + t.lTY);
        const uint32_t ix = (tx << t.lTX) + dx, iy = (ty << t.lTY) + dy;
        return (((ix << t.lS) + iy) << t.lZC) + zc;
    }
    // every quantity is below S^2 * ZC = n / 256 < 2^32
    const uint32_t per_tile = t.TX * t.TY * t.ZC;
    const uint32_t tile = L / per_tile, w = L % per_tile;
    const uint32_t tiles_y = t.S / t.TY;
    const uint32_t tx = tile / tiles_y, ty = tile % tiles_y;
    const uint32_t zc = w % t.ZC, dy = (w / t.ZC) % t.TY, dx = w / (t.ZC * t.TY);
    const uint32_t ix = tx * t.TX + dx, iy = ty * t.TY + dy;
    return (ix * t.S + iy) * t.ZC + zc;
}

template <typename T, typename PT>
struct GatherArgs {
    Geom g;
    uint64_t n;
    const T* phi;               // R after the c2r pass
    const double* x_in;
    const double* y_in;
    const double* z_in;
    double* x_out;
    double* y_out;
    double* z_out;
    const PT* px_in;
    const PT* py_in;
    const PT* pz_in;
    PT* px_out;
    PT* py_out;
    PT* pz_out;
    double* acc_x;              // optional: acc += g (force mode)
    double* acc_y;
    double* acc_z;
    T* g_x;                     // optional: g written (transactional force mode)
    T* g_y;
    T* g_z;
    LatticeTiling lat;          // block schedule over the Lagrangian lattice (or identity)
    uint32_t* next_counts;      // optional: bucket counts/ranks of the drifted positions
    uint32_t* next_keys;        // keys and ranks per particle, or with rec_out the overflow
    uint32_t* next_ranks;       // list's bucket and rank per entry
    const uint32_t* rec_start;  // (binary32) the drifted positions' deposit records are written into
    DepRecord* rec_out;         // the capacity layout rec_start (nb+1 prefix); those past a bucket's
    uint32_t* ovf_count;        // capacity are appended to the overflow list (ovf_rec, next_keys,
    DepRecord* ovf_rec;         // next_ranks)
    const double* dfac;         // optional: k_prev, k_next, d_next read from device memory
                                // (graph replays take new factors without re-capture)
    double k_prev;              // K2 of the previous step
    double k_next;              // K1 of the next step
    double d_next;              // drift factor of the next step
    int has_prev;
    int has_next;
    uint32_t* err;
};

// O32: every wave's particles lie in one box (one box, or a particle count
// per box that is a multiple of 64), whose mesh spans at most 2^32 bytes; the
// stencil then uses 32-bit byte offsets from the box's mesh base (one scalar
// base, no per-lane 64-bit address arithmetic).  REC: the drifted positions' deposit records are
// written into the capacity layout (binary32); otherwise keys and ranks.
// OUT: the force outputs acc and g may be set; without them the kernel keeps
// twelve fewer scalar registers live (8 waves per SIMD instead of 7).
template <typename T, typename PT, bool O32 = false, bool REC = false, bool OUT = true>
__global__ void __launch_bounds__(256) k_gather_kdk(GatherArgs<T, PT> a) {
    // each XCD walks a contiguous particle range (spatially compact for
    // lattice-ordered particles), so potential lines are reused in its L2
    const uint64_t p = (uint64_t)lattice_block(xcd_block(blockIdx.x, gridDim.x), a.lat) * blockDim.x + threadIdx.x;
    const Geom& g = a.g;
    const int N = g.N;
    uint32_t fail = 0;
    bool live = p < a.n;
    uint32_t next_key = 0;
    bool key_ok = false;
    DepRecord next_rec{0u, 0.0f, 0.0f, 0.0f};
    if (live) {
        int ix = 0, iy = 0, iz = 0;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        // the whole particle state is loaded up front: one memory round trip
        // for it, then one for the potential stencil
        const double xp = __builtin_nontemporal_load(&a.x_in[p]), yp = __builtin_nontemporal_load(&a.y_in[p]),
                     zp = __builtin_nontemporal_load(&a.z_in[p]);
        const bool kick = a.has_prev || a.has_next;
        PT qx0 = 0, qy0 = 0, qz0 = 0;
        if (kick) {
            qx0 = __builtin_nontemporal_load(&a.px_in[p]);
            qy0 = __builtin_nontemporal_load(&a.py_in[p]);
            qz0 = __builtin_nontemporal_load(&a.pz_in[p]);
        }
        if (!(cic_axis(xp, g, ix, fx) && cic_axis(yp, g, iy, fy) && cic_axis(zp, g, iz, fz))) {
            fail = ERR_POSITION_NONFINITE;
        } else {
            const T wx[2] = {(T)(1.0 - fx), (T)fx};
            const T wy[2] = {(T)(1.0 - fy), (T)fy};
            const T wz[2] = {(T)(1.0 - fz), (T)fz};
            // The 8 corners' centred differences touch 32 distinct potential
            // values: the x stencil X[k][b][c] at (ix-1+k, iy+b, iz+c), and the
            // y / z stencil ends Y[j][a][c] at (ix+a, iy-1 / iy+2, iz+c) and
            // Z[j][a][b] at (ix+a, iy+b, iz-1 / iz+2).  All 32 loads are
            // issued before any arithmetic.
            using Off = typename std::conditional<O32, uint32_t, uint64_t>::type;
            const Off NN = (Off)N * (Off)N;
            const Off es = O32 ? (Off)sizeof(T) : (Off)1;  // offsets in bytes (O32) or elements
            // O32: the wave's box, read from its first active lane, gives a scalar base
            const uint64_t box = O32 ? (uint64_t)__builtin_amdgcn_readfirstlane(box_of(p, g)) : (uint64_t)box_of(p, g);
            const T* phi = a.phi + box * (uint64_t)N * (uint64_t)N * (uint64_t)N;
            auto ld = [&](Off o) -> T {
                if constexpr (O32) return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(phi) + o);
                else return phi[o];
            };
            Off xo[4], yo[4], zo[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                xo[k] = (Off)wrap_cell(ix - 1 + k, N) * NN * es;
                yo[k] = (Off)wrap_cell(iy - 1 + k, N) * (Off)N * es;
                zo[k] = (Off)wrap_cell(iz - 1 + k, N) * es;
            }
            T X[4][2][2], Y[2][2][2], Z[2][2][2];
            // The stencil as 12 z-contiguous runs: the (ix+a, iy+b) lines hold
            // iz-1..iz+2 (x and z stencils), the x-end and y-end lines
            // iz..iz+1.  A run that does not wrap in z is one vector load.
            typedef T V4 __attribute__((ext_vector_type(4), aligned(sizeof(T))));
            typedef T V2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
            auto run = [&](Off xy, int z0, T* dst, int cnt) {
                if (z0 >= 0 && z0 + cnt <= N) {
                    const Off o = xy + (Off)z0 * es;
                    const char* src = O32 ? reinterpret_cast<const char*>(phi) + o
                                          : reinterpret_cast<const char*>(phi + o);
                    if (cnt == 4) {
                        const V4 v = *reinterpret_cast<const V4*>(src);
                        dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; dst[3] = v.w;
                    } else {
                        const V2 v = *reinterpret_cast<const V2*>(src);
                        dst[0] = v.x; dst[1] = v.y;
                    }
                } else {
                    for (int c = 0; c < cnt; ++c) dst[c] = ld(xy + (Off)wrap_cell(z0 + c, N) * es);
                }
            };
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    T q[4];
                    run(xo[u + 1] + yo[b + 1], iz - 1, q, 4);
                    Z[0][u][b] = q[0];
                    X[u + 1][b][0] = q[1];
                    X[u + 1][b][1] = q[2];
                    Z[1][u][b] = q[3];
                }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                run(xo[0] + yo[b + 1], iz, X[0][b], 2);
                run(xo[3] + yo[b + 1], iz, X[3][b], 2);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int u = 0; u < 2; ++u) run(xo[u + 1] + yo[j * 3], iz, Y[j][u], 2);
            // -(f+ - f-) / (2 dx); when dx is a power of two the quotient equals
            // the product with the exact reciprocal bit for bit
            const T two_dx = (T)(2.0 * g.dx);
            const T inv_two_dx = (T)(0.5 * g.inv_dx);
            const bool exact = g.inv_exact != 0;
            auto cdiff = [&](T d) -> T { return exact ? (-d) * inv_two_dx : -d / two_dx; };
            T gx = 0, gy = 0, gz = 0;
            // Lost non-zero products (src/mesh/mass_assignment_gradient.cpp:157-196):
            // a weight whose three factors are non-zero, a gradient of a non-zero
            // difference, or a contribution of a non-zero weight and gradient,
            // that rounds to zero.  The reference's corner condition
            //   ((w != 0 || w_lost) && g_lost) || (w_lost && g_nonzero) || (w != 0 && c_lost)
            // equals, for the finite weights here, "all three factors non-zero and
            // some axis with d != 0 and (c == 0 or w == 0)": with w == 0 it asks
            // for a non-zero difference (g_lost || g_nonzero), with w != 0 for a
            // non-zero difference whose contribution w * g is zero (g lost, or
            // g != 0 and the product lost).
            const bool nzx[2] = {wx[0] != (T)0, wx[1] != (T)0};
            const bool nzy[2] = {wy[0] != (T)0, wy[1] != (T)0};
            const bool nzz[2] = {wz[0] != (T)0, wz[1] != (T)0};
            bool lost = false;
#pragma unroll
            for (int ox = 0; ox < 2; ++ox) {
#pragma unroll
                for (int oy = 0; oy < 2; ++oy) {
#pragma unroll
                    for (int oz = 0; oz < 2; ++oz) {
                        const T w = wx[ox] * wy[oy] * wz[oz];
                        const T ypl = oy == 0 ? X[ox + 1][1][oz] : Y[1][ox][oz];
                        const T ymi = oy == 0 ? Y[0][ox][oz] : X[ox + 1][0][oz];
                        const T zpl = oz == 0 ? X[ox + 1][oy][1] : Z[1][ox][oy];
                        const T zmi = oz == 0 ? Z[0][ox][oy] : X[ox + 1][oy][0];
                        const T dxv = X[ox + 2][oy][oz] - X[ox][oy][oz];
                        const T dyv = ypl - ymi;
                        const T dzv = zpl - zmi;
                        const T cx = w * cdiff(dxv), cy = w * cdiff(dyv), cz = w * cdiff(dzv);
                        const bool wz0 = w == (T)0;
                        const bool any = (dxv != (T)0 && (cx == (T)0 || wz0)) || (dyv != (T)0 && (cy == (T)0 || wz0)) ||
                                         (dzv != (T)0 && (cz == (T)0 || wz0));
                        lost |= nzx[ox] && nzy[oy] && nzz[oz] && any;
                        gx += cx;
                        gy += cy;
                        gz += cz;
                    }
                }
            }
            if (lost) {
                fail = ERR_GATHER_UNDERFLOW;
            } else if (!isfinite(gx) || !isfinite(gy) || !isfinite(gz)) {
                fail = ERR_FORCE_NONFINITE;
            } else {
                if constexpr (OUT) {
                    if (a.acc_x) {
                        a.acc_x[p] += (double)gx;
                        a.acc_y[p] += (double)gy;
                        a.acc_z[p] += (double)gz;
                    }
                    if (a.g_x) {
                        a.g_x[p] = gx;
                        a.g_y[p] = gy;
                        a.g_z[p] = gz;
                    }
                }
                if (kick) {
                    PT qx = qx0, qy = qy0, qz = qz0;
                    if (!isfinite(qx) || !isfinite(qy) || !isfinite(qz)) {
                        fail = ERR_MOMENTUM_NONFINITE;
                    } else {
                        if (a.has_prev) {
                            const PT k = (PT)(a.dfac ? a.dfac[0] : a.k_prev);
                            qx = qx + (PT)gx * k;
                            qy = qy + (PT)gy * k;
                            qz = qz + (PT)gz * k;
                        }
                        double nx = xp, ny = yp, nz = zp;
                        if (a.has_next) {
                            const PT k = (PT)(a.dfac ? a.dfac[1] : a.k_next);
                            qx = qx + (PT)gx * k;
                            qy = qy + (PT)gy * k;
                            qz = qz + (PT)gz * k;
                            const double D = a.dfac ? a.dfac[2] : a.d_next;
                            const double L = g.L;
                            const double ddx = (double)qx * D, ddy = (double)qy * D, ddz = (double)qz * D;
                            nx = wrap_pos(wrap_pos(xp, L) + minimum_image(ddx, L), L);
                            ny = wrap_pos(wrap_pos(yp, L) + minimum_image(ddy, L), L);
                            nz = wrap_pos(wrap_pos(zp, L) + minimum_image(ddz, L), L);
                        }
                        if (!isfinite(qx) || !isfinite(qy) || !isfinite(qz) || !isfinite(nx)
                            || !isfinite(ny) || !isfinite(nz)) {
                            fail = ERR_UPDATE_NONFINITE;
                        } else {
                            // the particle streams are read and written once per step:
                            // non-temporal, so they do not evict the potential stencil
                            __builtin_nontemporal_store(qx, &a.px_out[p]);
                            __builtin_nontemporal_store(qy, &a.py_out[p]);
                            __builtin_nontemporal_store(qz, &a.pz_out[p]);
                            __builtin_nontemporal_store(nx, &a.x_out[p]);
                            __builtin_nontemporal_store(ny, &a.y_out[p]);
                            __builtin_nontemporal_store(nz, &a.z_out[p]);
                            if (a.next_counts) {
                                // bucket of the drifted position for the next deposit
                                int bx, by, bz;
                                double fxn, fyn, fzn;
                                key_ok = cic_axis(nx, g, bx, fxn) && cic_axis(ny, g, by, fyn)
                                         && cic_axis(nz, g, bz, fzn);
                                if (key_ok) {
                                    next_key = box_of(p, g) * buckets_per_box(g) + bucket_of(bx, by, bz, g);
                                    if constexpr (REC) {  // the record k_scatter_records would write
                                        next_rec.cell = record_cell(bx, by, bz, g);
                                        next_rec.fx = (float)fxn;
                                        next_rec.fy = (float)fyn;
                                        next_rec.fz = (float)fzn;
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
        if (fail) atomicOr(a.err, fail);
    }
    if (a.next_counts) {  // uniform across the grid: every lane reaches the ballots
        const bool have = live && key_ok;
        if constexpr (REC) {
            // the bucket's capacity slots are loaded ahead of the rank atomic, so
            // the two round trips overlap
            uint32_t s0 = 0, s1 = 0;
            if (have) {
                s0 = a.rec_start[next_key];
                s1 = a.rec_start[next_key + 1];
            }
            const uint32_t rank = aggregated_rank(a.next_counts, next_key, have);
            uint32_t slot = 0;
            bool over = false;
            if (have) {
                if (rank < s1 - s0) slot = s0 + rank;
                else over = true;
            }
            const uint64_t om = __ballot(over);
            uint32_t obase = 0;
            if (om) {  // one atomic per wave for its overflow records
                const int leader = __ffsll((unsigned long long)om) - 1;
                if ((int)lane_id() == leader) obase = atomicAdd(a.ovf_count, (uint32_t)__popcll(om));
                obase = (uint32_t)__builtin_amdgcn_readlane((int)obase, leader);
            }
            if (have) {
                if (!over) {
                    a.rec_out[slot] = next_rec;
                } else {
                    const uint32_t o = obase + (uint32_t)__popcll(om & ((1ull << lane_id()) - 1ull));
                    a.ovf_rec[o] = next_rec;
                    a.next_keys[o] = next_key;
                    a.next_ranks[o] = rank;
                }
            }
        } else {
            const uint32_t rank = aggregated_rank(a.next_counts, next_key, have);
            if (live) {
                __builtin_nontemporal_store(key_ok ? next_key : 0xffffffffu, &a.next_keys[p]);
                __builtin_nontemporal_store(rank, &a.next_ranks[p]);
            }
        }
    }
}

// CIC interpolation of up to three meshes of one box
// (src/mesh/mass_assignment_gather.inc:95-181,402-451): v = sum over the x, y,
// z corner offsets, in that order, of f * ((wx * wy) * wz).  A product that
// rounds to zero although its factors are non-zero raises ERR_INTERP_UNDERFLOW,
// as the reference's underflow_error.  out[m] receives v (force mode, a
// non-finite v raises ERR_FORCE_NONFINITE); add[m] receives += v.
template <typename T>
struct InterpArgs {
    Geom g;
    uint64_t n;
    const double* x;
    const double* y;
    const double* z;
    const T* f[3];
    T* out[3];
    double* add[3];
    int meshes;
    uint32_t* err;
};

template <typename T, bool O32>
__global__ void __launch_bounds__(256) k_interp(InterpArgs<T> a) {
    const uint64_t p = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const Geom& g = a.g;
    const int N = g.N;
    int ix, iy, iz;
    double fx, fy, fz;
    if (!cic_axes(a.x[p], a.y[p], a.z[p], g, ix, iy, iz, fx, fy, fz)) {
        atomicOr(a.err, ERR_POSITION_NONFINITE);
        return;
    }
    using Off = typename std::conditional<O32, uint32_t, uint64_t>::type;
    const Off es = O32 ? (Off)sizeof(T) : (Off)1;
    const Off NN = (Off)N * (Off)N;
    const T wx[2] = {(T)(1.0 - fx), (T)fx};
    const T wy[2] = {(T)(1.0 - fy), (T)fy};
    const T wz[2] = {(T)(1.0 - fz), (T)fz};
    const Off xo[2] = {(Off)ix * NN * es, (Off)wrap_cell(ix + 1, N) * NN * es};
    const Off yo[2] = {(Off)iy * (Off)N * es, (Off)wrap_cell(iy + 1, N) * (Off)N * es};
    const bool zrun = iz + 1 < N;
    const Off z0 = (Off)iz * es, z1 = (Off)wrap_cell(iz + 1, N) * es;
    typedef T V2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
    bool lost = false, bad = false;
    for (int m = 0; m < a.meshes; ++m) {
        const T* f = a.f[m];
        auto at = [&](Off o) -> const char* {
            return O32 ? reinterpret_cast<const char*>(f) + o : reinterpret_cast<const char*>(f + o);
        };
        T F[2][2][2];
#pragma unroll
        for (int ox = 0; ox < 2; ++ox)
#pragma unroll
            for (int oy = 0; oy < 2; ++oy) {
                const Off o = xo[ox] + yo[oy];
                if (zrun) {
                    const V2 v = *reinterpret_cast<const V2*>(at(o + z0));
                    F[ox][oy][0] = v.x;
                    F[ox][oy][1] = v.y;
                } else {
                    F[ox][oy][0] = *reinterpret_cast<const T*>(at(o + z0));
                    F[ox][oy][1] = *reinterpret_cast<const T*>(at(o + z1));
                }
            }
        T v = (T)0;
#pragma unroll
        for (int ox = 0; ox < 2; ++ox)
#pragma unroll
            for (int oy = 0; oy < 2; ++oy)
#pragma unroll
                for (int oz = 0; oz < 2; ++oz) {
                    const T xy = wx[ox] * wy[oy];
                    const T w = xy * wz[oz];
                    const T fv = F[ox][oy][oz];
                    if (wx[ox] != (T)0 && wy[oy] != (T)0 && wz[oz] != (T)0 && w == (T)0 && fv != (T)0) lost = true;
                    const T c = fv * w;
                    if (fv != (T)0 && w != (T)0 && c == (T)0) lost = true;
                    v += c;
                }
        if (a.out[m]) {
            if (!isfinite(v)) bad = true;
            a.out[m][p] = v;
        }
        if (a.add[m]) a.add[m][p] += (double)v;
    }
    if (lost) atomicOr(a.err, ERR_INTERP_UNDERFLOW);
    if (bad) atomicOr(a.err, ERR_FORCE_NONFINITE);
}

// Potential-energy diagnostic: sum_p m_p phi(x_p) with phi CIC-interpolated
// (CICMassAssignment::interpolate_mass_weighted_sum_full_mesh,
// src/mesh/mass_assignment_gather.inc:283-399; the caller halves it as
// src/gravity/pm_solver.cpp:817-841 does).  Per particle, in binary64 as the
// reference (its accumulation type is binary64): value = sum over corners
// (x outer, z inner) of phi_c * ((wx wy) wz), then m * value, with the
// reference's lost-product checks (ERR_INTERP_UNDERFLOW).  The summation is the
// reference's deterministic blocked sum (include/cosmo_nbody/math/
// deterministic_sum.hpp:48-104): workgroup b forms the terms of particles
// [b*4096, (b+1)*4096) in LDS, then one lane adds them in index order with
// Neumaier compensation; k_potential_total adds the block sums in order.
constexpr int ENERGY_THREADS = 256;
constexpr int ENERGY_BLOCK = 4096;

__device__ __forceinline__ void neumaier_add(double v, double& sum, double& comp) {
    const double t = sum + v;
    if (fabs(sum) >= fabs(v)) comp += (sum - t) + v;
    else comp += (v - t) + sum;
    sum = t;
}

template <typename T>
__global__ void __launch_bounds__(ENERGY_THREADS) k_potential_partials(Geom g, uint64_t n, const double* __restrict__ x,
                                                                      const double* __restrict__ y,
                                                                      const double* __restrict__ z,
                                                                      const double* __restrict__ masses,
                                                                      double uniform_mass, const T* __restrict__ phi,
                                                                      double* __restrict__ partials,
                                                                      uint32_t* __restrict__ err) {
    __shared__ double terms[ENERGY_BLOCK];
    const uint64_t b0 = (uint64_t)blockIdx.x * ENERGY_BLOCK;
    const uint64_t b1 = n < b0 + ENERGY_BLOCK ? n : b0 + ENERGY_BLOCK;
    const int N = g.N;
    const uint64_t NN = (uint64_t)N * (uint64_t)N;
    bool lost = false, bad = false;
    for (uint64_t p = b0 + threadIdx.x; p < b1; p += ENERGY_THREADS) {
        int ix, iy, iz;
        double fx, fy, fz;
        double weighted = 0.0;
        if (!cic_axes(x[p], y[p], z[p], g, ix, iy, iz, fx, fy, fz)) {
            bad = true;
        } else {
            const double m = masses ? masses[p] : uniform_mass;
            const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
            const uint64_t xs[2] = {(uint64_t)ix, (uint64_t)wrap_cell(ix + 1, N)};
            const uint64_t ys[2] = {(uint64_t)iy, (uint64_t)wrap_cell(iy + 1, N)};
            const uint64_t zs[2] = {(uint64_t)iz, (uint64_t)wrap_cell(iz + 1, N)};
            double value = 0.0;
#pragma unroll
            for (int ox = 0; ox < 2; ++ox)
#pragma unroll
                for (int oy = 0; oy < 2; ++oy)
#pragma unroll
                    for (int oz = 0; oz < 2; ++oz) {
                        const double fv = (double)phi[xs[ox] * NN + ys[oy] * (uint64_t)N + zs[oz]];
                        const double xy = wx[ox] * wy[oy];
                        const double w = xy * wz[oz];
                        if (wx[ox] != 0.0 && wy[oy] != 0.0 && wz[oz] != 0.0 && w == 0.0 && fv != 0.0 && m != 0.0)
                            lost = true;
                        const double c = fv * w;
                        if (fv != 0.0 && w != 0.0 && c == 0.0 && m != 0.0) lost = true;
                        value += c;
                    }
            weighted = m * value;
            if (m != 0.0 && value != 0.0 && weighted == 0.0) lost = true;
        }
        terms[p - b0] = weighted;
    }
    if (bad) atomicOr(err, ERR_POSITION_NONFINITE);
    if (lost) atomicOr(err, ERR_INTERP_UNDERFLOW);
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0, c = 0.0;
        const uint32_t count = (uint32_t)(b1 - b0);
        for (uint32_t i = 0; i < count; ++i) neumaier_add(terms[i], s, c);
        partials[blockIdx.x] = s + c;
    }
}

// 0.5 * (Neumaier sum of the workgroup partials, in order) -> out[0]
__global__ void k_potential_total(const double* __restrict__ partials, uint64_t count, double* __restrict__ out,
                                  uint32_t* __restrict__ err) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0, c = 0.0;
    for (uint64_t i = 0; i < count; ++i) neumaier_add(partials[i], s, c);
    const double total = s + c;
    const double half = 0.5 * total;
    if (!isfinite(total) || (total != 0.0 && half == 0.0)) atomicOr(err, ERR_INTERP_UNDERFLOW);
    out[0] = half;
}

// 1LPT displacement spectrum of one axis (src/ic/lpt_displacement.cpp:306-340):
// psi_k = (-Im d, Re d) * k_d / k^2 with k_d = 0 on the Nyquist index of the
// axis and psi_k = 0 where k^2 = 0, times 1/N^3 (the inverse transform's
// normalisation), written into B [kz][ky][kx] from the r2c half spectrum
// dk [x][y][kz] (binary64).
__global__ void __launch_bounds__(256) k_lpt_spectrum(const double2* __restrict__ dk, int axis, int N,
                                                      const double* __restrict__ kc, const double* __restrict__ kd,
                                                      double norm, double2* __restrict__ B) {
    const int NH = N / 2 + 1;
    const uint64_t total = (uint64_t)NH * N * N;
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int kx = (int)(e % (uint64_t)N), ky = (int)((e / (uint64_t)N) % (uint64_t)N), kz = (int)(e / ((uint64_t)N * N));
    const double qx = kc[kx], qy = kc[ky], qz = kc[kz];
    const double k2 = qx * qx + qy * qy + qz * qz;
    double2 out = {0.0, 0.0};
    if (k2 != 0.0) {
        const double dkv = axis == 0 ? kd[kx] : (axis == 1 ? kd[ky] : kd[kz]);
        const double f = dkv / k2;
        const double2 d = dk[((uint64_t)kx * N + ky) * NH + kz];
        out = {-d.y * f * norm, d.x * f * norm};
    }
    B[e] = out;
}

// 1LPT particle update of one axis (src/ic/lpt_displacement.cpp:99-129,
// 341-384; lattice src/ic/initial_conditions.cpp:296-312): particle
// (i, j, k) of the n^3 lattice takes psi at mesh cell (i s, j s, k s),
// position wrap(q + psi) with q = (i, j or k) * dx, momentum p_factor * psi.
__global__ void __launch_bounds__(256) k_lpt_update(const double* __restrict__ psi, int axis, int N, int n, int s,
                                                    double dx, double L, double p_factor, double* __restrict__ pos,
                                                    double* __restrict__ mom, uint32_t* __restrict__ err) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)n * n * n;
    if (p >= total) return;
    const uint64_t i = p / ((uint64_t)n * n), j = (p / (uint64_t)n) % (uint64_t)n, k = p % (uint64_t)n;
    const double v = psi[((i * s) * (uint64_t)N + j * s) * (uint64_t)N + k * s];
    const uint64_t q = axis == 0 ? i : (axis == 1 ? j : k);
    const double disp = 1.0 * v;
    const double inc = p_factor * v;
    const double cand = (double)q * dx + disp;
    const double mcand = 0.0 + inc;
    const double w = wrap_pos(cand, L);
    if (!isfinite(v) || !isfinite(disp) || !isfinite(inc) || !isfinite(cand) || !isfinite(mcand) || !isfinite(w)) {
        atomicOr(err, ERR_UPDATE_NONFINITE);
        return;
    }
    pos[p] = w;
    mom[p] = mcand;
}

// Transactional commit for the reference-shaped force entry: acc += g.
template <typename T>
__global__ void k_commit_force(uint64_t n, const T* __restrict__ gx, const T* __restrict__ gy,
                               const T* __restrict__ gz, double* __restrict__ ax,
                               double* __restrict__ ay, double* __restrict__ az) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    ax[p] += (double)gx[p];
    ay[p] += (double)gy[p];
    az[p] += (double)gz[p];
}

// Finite checks used by the force entry before anything is published.
__global__ void k_check_finite(uint64_t n, const double* __restrict__ a, const double* __restrict__ b,
                               const double* __restrict__ c, uint32_t bit, uint32_t* __restrict__ err) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (p < n) bad = !isfinite(a[p]) || !isfinite(b[p]) || !isfinite(c[p]);
    if (__ballot(bad) && lane_id() == 0) atomicOr(err, bit);
}

__global__ void k_check_mass(uint64_t n, const double* __restrict__ m, uint32_t* __restrict__ err) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (p < n) bad = !isfinite(m[p]) || m[p] < 0.0;
    if (__ballot(bad) && lane_id() == 0) atomicOr(err, ERR_MASS_INVALID);
}

// ===========================================================================
// Adjoint (vector-Jacobian products) of the fused step.  The forward step is
//   p' = p + (k_prev + k_next) g(x),   x' = wrap(x + d_next p'),
//   g_j(x) = sum_c W(x_j - c) (-D phi)(c),  phi = G s,  s_c = (m_c - M/N_c)/dx^3,
//   m_c = sum_k m W(x_k - c),   D_a f(c) = (f(c + e_a) - f(c - e_a)) / (2 dx).
// For an adjoint vector u (one 3-vector per particle) the product
// J_g(x)^T u has two parts per particle k:
//   (a) through particle k's own gather weights:
//       sum_c dW(x_k - c)/dx_k (u_k . (-D phi)(c)),
//   (b) through the potential: with A_a = sum_j u_j,a W(x_j - c) deposited,
//       chi = G (sum_a D_a A_a) (G symmetric, D_a antisymmetric),
//       (m / dx^3) sum_c chi(c) dW(x_k - c)/dx_k.
// ===========================================================================

// u = lp + d_next lx (the momentum adjoint entering the kick), written over lp.
__global__ void k_adj_kick(uint64_t n, const double* __restrict__ lx, const double* __restrict__ ly,
                           const double* __restrict__ lz, double* __restrict__ lpx, double* __restrict__ lpy,
                           double* __restrict__ lpz, double d_next, int has_next) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !has_next) return;
    lpx[p] = lpx[p] + d_next * lx[p];
    lpy[p] = lpy[p] + d_next * ly[p];
    lpz[p] = lpz[p] + d_next * lz[p];
}

// max |q| over particles (per block, then over blocks): the signed deposit unit.
__global__ void k_absmax(uint64_t n, const double* __restrict__ q, double* __restrict__ out) {
    __shared__ double sh[256];
    double m = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const double v = fabs(q[i]);
        m = v > m ? v : m;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Per-particle values q, scaled by inv_unit, into bucket order for the
// binary32 record deposit (keys/ranks/offsets of the same positions).
template <typename MT>
__global__ void k_scatter_values(uint64_t n, const double* __restrict__ q, double inv_unit,
                                 const uint32_t* __restrict__ keys, const uint32_t* __restrict__ ranks,
                                 const uint32_t* __restrict__ offsets, MT* __restrict__ out) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = keys[p];
    if (k == 0xffffffffu) return;
    if constexpr (sizeof(MT) == sizeof(double)) out[offsets[k] + ranks[p]] = q[p];
    else out[offsets[k] + ranks[p]] = (MT)(q[p] * inv_unit);
}

// out = sum_a D_a A_a (periodic centred differences), one box.
template <typename T>
__global__ void k_div_stencil(const T* __restrict__ ax, const T* __restrict__ ay, const T* __restrict__ az,
                              T* __restrict__ out, int N, double inv_2dx) {
    const uint64_t NN = (uint64_t)N * N;
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= NN * N) return;
    const int iz = (int)(c % N), iy = (int)((c / N) % N), ix = (int)(c / NN);
    const uint64_t xp = (uint64_t)wrap_cell(ix + 1, N) * NN + (uint64_t)iy * N + iz;
    const uint64_t xm = (uint64_t)wrap_cell(ix - 1, N) * NN + (uint64_t)iy * N + iz;
    const uint64_t yp = (uint64_t)ix * NN + (uint64_t)wrap_cell(iy + 1, N) * N + iz;
    const uint64_t ym = (uint64_t)ix * NN + (uint64_t)wrap_cell(iy - 1, N) * N + iz;
    const uint64_t zp = (uint64_t)ix * NN + (uint64_t)iy * N + wrap_cell(iz + 1, N);
    const uint64_t zm = (uint64_t)ix * NN + (uint64_t)iy * N + wrap_cell(iz - 1, N);
    const double d = ((double)ax[xp] - (double)ax[xm]) + ((double)ay[yp] - (double)ay[ym]) +
                     ((double)az[zp] - (double)az[zm]);
    out[c] = (T)(d * inv_2dx);
}

template <typename T>
struct AdjGatherArgs {
    Geom g;
    uint64_t n;
    const double* x;
    const double* y;
    const double* z;
    const double* ux;    // adjoint vector per particle (term (a)); null: no term (a)
    const double* uy;
    const double* uz;
    const T* phi;        // forward potential (term (a))
    const T* chi;        // scalar field of term (b)
    double coef_a;       // lambda += coef_a * (a) + coef_b * (b)
    double coef_b;
    double* lx;
    double* ly;
    double* lz;
    uint32_t* err;
};

template <typename T>
__global__ void __launch_bounds__(256) k_adj_gather(AdjGatherArgs<T> a) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const Geom& g = a.g;
    const int N = g.N;
    int ix, iy, iz;
    double fx, fy, fz;
    if (!cic_axes(a.x[p], a.y[p], a.z[p], g, ix, iy, iz, fx, fy, fz)) {
        atomicOr(a.err, ERR_POSITION_NONFINITE);
        return;
    }
    const double inv_dx = 1.0 / g.dx;
    const double w[3][2] = {{1.0 - fx, fx}, {1.0 - fy, fy}, {1.0 - fz, fz}};
    const double dw[2] = {-inv_dx, inv_dx};
    const uint64_t NN = (uint64_t)N * N;
    const uint64_t boff = (uint64_t)box_of(p, g) * NN * N;
    double ta[3] = {0.0, 0.0, 0.0}, tb[3] = {0.0, 0.0, 0.0};
    const bool with_a = a.ux != nullptr;
    const double u[3] = {with_a ? a.ux[p] : 0.0, with_a ? a.uy[p] : 0.0, with_a ? a.uz[p] : 0.0};
    const double two_dx = 2.0 * g.dx;
#pragma unroll
    for (int ox = 0; ox < 2; ++ox) {
        const int cx = wrap_cell(ix + ox, N);
#pragma unroll
        for (int oy = 0; oy < 2; ++oy) {
            const int cy = wrap_cell(iy + oy, N);
#pragma unroll
            for (int oz = 0; oz < 2; ++oz) {
                const int cz = wrap_cell(iz + oz, N);
                const uint64_t c = boff + (uint64_t)cx * NN + (uint64_t)cy * N + cz;
                // derivative of the corner weight along each axis
                const double dwc[3] = {dw[ox] * w[1][oy] * w[2][oz], w[0][ox] * dw[oy] * w[2][oz],
                                       w[0][ox] * w[1][oy] * dw[oz]};
                const double xv = (double)a.chi[c];
                double s = 0.0;
                if (with_a) {
                    const T* phi = a.phi + boff;
                    const double gx = -((double)phi[(uint64_t)wrap_cell(cx + 1, N) * NN + (uint64_t)cy * N + cz] -
                                        (double)phi[(uint64_t)wrap_cell(cx - 1, N) * NN + (uint64_t)cy * N + cz]) / two_dx;
                    const double gy = -((double)phi[(uint64_t)cx * NN + (uint64_t)wrap_cell(cy + 1, N) * N + cz] -
                                        (double)phi[(uint64_t)cx * NN + (uint64_t)wrap_cell(cy - 1, N) * N + cz]) / two_dx;
                    const double gz = -((double)phi[(uint64_t)cx * NN + (uint64_t)cy * N + wrap_cell(cz + 1, N)] -
                                        (double)phi[(uint64_t)cx * NN + (uint64_t)cy * N + wrap_cell(cz - 1, N)]) / two_dx;
                    s = u[0] * gx + u[1] * gy + u[2] * gz;
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    ta[d] += dwc[d] * s;
                    tb[d] += dwc[d] * xv;
                }
            }
        }
    }
    a.lx[p] += a.coef_a * ta[0] + a.coef_b * tb[0];
    a.ly[p] += a.coef_a * ta[1] + a.coef_b * tb[1];
    a.lz[p] += a.coef_a * ta[2] + a.coef_b * tb[2];
}

}  // namespace hy

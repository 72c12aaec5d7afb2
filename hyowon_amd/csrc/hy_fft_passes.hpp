// The five pencil passes of the PM Poisson solve (gfx950): r2c along z,
// forward y, forward x + Green + backward x, backward y, c2r along z.  Each is
// templated on the mesh size N and a launch variant V (rows per workgroup).
#pragma once

#include "hy_device.hpp"
#include "hy_fft.hpp"

namespace hy {

// ===========================================================================
// FFT passes.  Layouts (NH = N/2+1):
//   R  [x][y][z]   real, N^3            (density source in, potential out)
//   A  [x][kz][y]  complex, N*NH*N      (after the z pass)
//   B  [kz][ky][x] complex, NH*N*N      (after the y pass; x pass in place)
// ===========================================================================
template <typename T>
struct FftArgs {
    T* R;
    cplx<T>* A;
    cplx<T>* B;
    const cplx<T>* tw;   // exp(-2 pi i m / N), m < N
    const T* green_lap;  // discrete Laplacian symbol per index (binary64 derived)
    const T* green_win;  // CIC window per index (deconvolution)
    double green_num;    // 4 pi G
    double norm;         // 1 / N^3
    int deconvolve;
    int kz0;             // first kz plane of this launch (y/x passes run over plane groups)
    uint32_t* err;
};

// Staged copies with compile-time trip counts: every thread issues all of its
// loads before its first dependent store, so each wave keeps PER requests in
// flight instead of one.
template <int COUNT, int NT, typename C, typename Load, typename Store>
__device__ __forceinline__ void staged_copy(int tid, Load load, Store store) {
    constexpr int PER = (COUNT + NT - 1) / NT;
    C v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int l = tid + i * NT;
        if (COUNT % NT == 0 || l < COUNT) v[i] = load(l);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int l = tid + i * NT;
        if (COUNT % NT == 0 || l < COUNT) store(l, v[i]);
    }
}

// z pass: real rows -> half spectrum, written transposed into A.
template <typename T, int N, int V>
__global__ void __launch_bounds__((RowsCfg<T, N / 2, N, V>::THREADS)) k_fft_z_r2c(FftArgs<T> a) {
    using C = cplx<T>;
    constexpr int LEN = N / 2;
    constexpr int NH = N / 2 + 1;
    using RC = RowsCfg<T, LEN, N, V>;
    constexpr int ROWS = RC::ROWS, RS = RC::RS, NT = RC::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int yb = b % (N / ROWS), x = b / (N / ROWS);
    const int y0 = yb * ROWS;
    const int tid = threadIdx.x;
    const C* src = reinterpret_cast<const C*>(a.R + ((uint64_t)x * N + y0) * N);
    staged_copy<ROWS * LEN, NT, C>(
        tid, [&](int l) { return src[(l / LEN) * LEN + (l % LEN)]; },
        [&](int l, C v) { lds[(l / LEN) * RS + lds_slot(l % LEN)] = v; });
    __syncthreads();
    fft_rows<C, LEN, ROWS, RS, -1>(lds, a.tw, 2, tid);
    // untangle: X[k] = E + W^k O, E = (Z[k]+conj Z[L-k])/2, O = -i/2 (Z[k]-conj Z[L-k])
    C* dst = a.A + (uint64_t)x * NH * N + y0;
    staged_copy<ROWS * NH, NT, C>(
        tid,
        [&](int l) {
            const int r = l % ROWS, k = l / ROWS;
            const C zk = lds[r * RS + lds_slot(k % LEN)];
            const C zc = cconj(lds[r * RS + lds_slot((LEN - k) % LEN)]);
            const C e = {(T)0.5 * (zk.x + zc.x), (T)0.5 * (zk.y + zc.y)};
            const C d = csub(zk, zc);
            const C o = {(T)0.5 * d.y, (T)-0.5 * d.x};
            return cadd(e, cmul(a.tw[k], o));
        },
        [&](int l, C v) { dst[(uint64_t)(l / ROWS) * N + (l % ROWS)] = v; });
}

// y pass forward: A rows over y -> B[kz][ky][x]
template <typename T, int N, int V>
__global__ void __launch_bounds__((RowsCfg<T, N, N, V>::THREADS)) k_fft_y_fwd(FftArgs<T> a) {
    using C = cplx<T>;
    constexpr int NH = N / 2 + 1;
    using RC = RowsCfg<T, N, N, V>;
    constexpr int ROWS = RC::ROWS, RS = RC::RS, NT = RC::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int xb = b % (N / ROWS), kz = a.kz0 + (int)(b / (N / ROWS));
    const int x0 = xb * ROWS;
    const int tid = threadIdx.x;
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return a.A[((uint64_t)(x0 + l / N) * NH + kz) * N + (l % N)]; },
        [&](int l, C v) { lds[(l / N) * RS + lds_slot(l % N)] = v; });
    __syncthreads();
    fft_rows<C, N, ROWS, RS, -1>(lds, a.tw, 1, tid);
    C* dst = a.B + (uint64_t)kz * N * N + x0;
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return lds[(l % ROWS) * RS + lds_slot(l / ROWS)]; },
        [&](int l, C v) { dst[(uint64_t)(l / ROWS) * N + (l % ROWS)] = v; });
}

// x pass: forward x FFT, Green multiply (with 1/N^3), backward x FFT, in place on B.
template <typename T, int N, int V>
__global__ void __launch_bounds__((RowsCfg<T, N, (N / 2 + 1) * N, V>::THREADS)) k_fft_x_green(FftArgs<T> a) {
    using C = cplx<T>;
    constexpr int NH = N / 2 + 1;
    using RC = RowsCfg<T, N, NH * N, V>;
    constexpr int ROWS = RC::ROWS, RS = RC::RS, NT = RC::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t q0 = (uint64_t)a.kz0 * N + (uint64_t)b * ROWS;  // row index q = kz*N + ky
    const int tid = threadIdx.x;
    C* rows = a.B + q0 * N;
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return rows[l]; }, [&](int l, C v) { lds[(l / N) * RS + lds_slot(l % N)] = v; });
    __syncthreads();
    fft_rows<C, N, ROWS, RS, -1>(lds, a.tw, 1, tid);
    constexpr int PER = ROWS * N / NT;
    static_assert(ROWS * N % NT == 0, "x pass covers whole rows");
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int l = tid + i * NT;
        const int r = l / N, kx = l % N;
        const uint64_t q = q0 + r;
        const int kz = (int)(q / N), ky = (int)(q % N);
        T f;
        if (kx == 0 && ky == 0 && kz == 0) {
            f = (T)0;
        } else {
            const T D = a.green_lap[kx] + a.green_lap[ky] + a.green_lap[kz];
            T ff = (T)a.green_num / D;
            if (a.deconvolve) {
                const T w = a.green_win[kx] * a.green_win[ky] * a.green_win[kz];
                ff /= (w * w);
            }
            f = ff * (T)a.norm;
        }
        C v = lds[r * RS + lds_slot(kx)];
        v.x *= f;
        v.y *= f;
        lds[r * RS + lds_slot(kx)] = v;
    }
    __syncthreads();
    fft_rows<C, N, ROWS, RS, +1>(lds, a.tw, 1, tid);
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return lds[(l / N) * RS + lds_slot(l % N)]; }, [&](int l, C v) { rows[l] = v; });
}

// y pass backward: B[kz][ky][x] -> A rows over y
template <typename T, int N, int V>
__global__ void __launch_bounds__((RowsCfg<T, N, N, V>::THREADS)) k_fft_y_bwd(FftArgs<T> a) {
    using C = cplx<T>;
    constexpr int NH = N / 2 + 1;
    using RC = RowsCfg<T, N, N, V>;
    constexpr int ROWS = RC::ROWS, RS = RC::RS, NT = RC::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int xb = b % (N / ROWS), kz = a.kz0 + (int)(b / (N / ROWS));
    const int x0 = xb * ROWS;
    const int tid = threadIdx.x;
    const C* src = a.B + (uint64_t)kz * N * N + x0;
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return src[(uint64_t)(l / ROWS) * N + (l % ROWS)]; },
        [&](int l, C v) { lds[(l % ROWS) * RS + lds_slot(l / ROWS)] = v; });
    __syncthreads();
    fft_rows<C, N, ROWS, RS, +1>(lds, a.tw, 1, tid);
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return lds[(l / N) * RS + lds_slot(l % N)]; },
        [&](int l, C v) { a.A[((uint64_t)(x0 + l / N) * NH + kz) * N + (l % N)] = v; });
}

// z pass backward (c2r): A[x][kz][y] -> R rows; flags a non-finite potential.
template <typename T, int N, int V>
__global__ void __launch_bounds__((RowsCfg<T, N / 2, N, V>::THREADS)) k_fft_z_c2r(FftArgs<T> a) {
    using C = cplx<T>;
    constexpr int LEN = N / 2;
    constexpr int NH = N / 2 + 1;
    using RC = RowsCfg<T, LEN, N, V>;
    constexpr int ROWS = RC::ROWS, RS = RC::RS, NT = RC::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int yb = b % (N / ROWS), x = b / (N / ROWS);
    const int y0 = yb * ROWS;
    const int tid = threadIdx.x;
    const C* src = a.A + (uint64_t)x * NH * N + y0;
    staged_copy<ROWS * NH, NT, C>(
        tid, [&](int l) { return src[(uint64_t)(l / ROWS) * N + (l % ROWS)]; },
        [&](int l, C v) {
            const int r = l % ROWS, k = l / ROWS;
            if (k == 0 || k == LEN) v.y = (T)0;  // c2r reads only the independent half
            lds[r * RS + lds_slot(k)] = v;
        });
    __syncthreads();
    // Z'[k] = (X[k] + conj X[L-k]) + i (X[k] - conj X[L-k]) W^-k
    for (int l = tid; l < ROWS * (LEN / 2 + 1); l += NT) {
        const int r = l % ROWS, k = l / ROWS;
        const C xk = lds[r * RS + lds_slot(k)];
        const C xm = lds[r * RS + lds_slot(LEN - k)];
        const C wk = cconj(a.tw[k]);
        const C s1 = cadd(xk, cconj(xm));
        const C d1 = cmul(csub(xk, cconj(xm)), wk);
        const C zk = {s1.x - d1.y, s1.y + d1.x};
        C zm = {(T)0, (T)0};
        const bool pair = k > 0 && (LEN - k) != k;
        if (pair) {
            const C wm = cconj(a.tw[LEN - k]);
            const C s2 = cadd(xm, cconj(xk));
            const C d2 = cmul(csub(xm, cconj(xk)), wm);
            zm = {s2.x - d2.y, s2.y + d2.x};
        }
        lds[r * RS + lds_slot(k)] = zk;
        if (pair) lds[r * RS + lds_slot(LEN - k)] = zm;
    }
    __syncthreads();
    fft_rows<C, LEN, ROWS, RS, +1>(lds, a.tw, 2, tid);
    int mybad = 0;
    C* dst = reinterpret_cast<C*>(a.R + ((uint64_t)x * N + y0) * N);
    staged_copy<ROWS * LEN, NT, C>(
        tid, [&](int l) { return lds[(l / LEN) * RS + lds_slot(l % LEN)]; },
        [&](int l, C v) {
            if (!isfinite(v.x) || !isfinite(v.y)) mybad = 1;
            dst[l] = v;
        });
    if (__ballot(mybad) && lane_id() == 0) atomicOr(a.err, ERR_FIELD_NONFINITE);
}

}  // namespace hy

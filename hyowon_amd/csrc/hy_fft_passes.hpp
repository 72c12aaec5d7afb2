// The five pencil passes of the PM Poisson solve (gfx950): r2c along z,
// forward y, forward x + Green + backward x, backward y, c2r along z.  Each is
// templated on the mesh size N and a launch variant V (rows per workgroup).
// The x, y^-1 and z^-1 passes also come as SG = 1 kernels for the TreePM
// long-range force: continuum Green with the Gaussian split and a spectral
// -i k derivative applied in the pass of its axis.
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
    uint64_t box_real;   // elements per box of R (N^3) and of A, B (N^2 (N/2+1))
    uint64_t box_cplx;
    uint32_t* err;
    // spectral-gradient passes (TreePM long range), read only by the SG = 1 kernels
    const T* kc;         // k component per index
    const T* kd;         // k component per index, 0 on the Nyquist index (derivative axis)
    cplx<T>* xout;       // x pass output (nullptr: in place on B)
    double rs;           // Gaussian split scale r_s
    int deriv;           // 0, or 1 / 2 / 3: multiply by -i k_x / k_y / k_z in the x / y^-1 / z^-1 pass
    int mode_limit;      // spectral estimators: modes with |m_y|, m_z <= mode_limit only (0: all).
                         // The pruned two-level z r2c pass stores the planes kz <= mode_limit;
                         // the forward-only x pass skips workgroups whose rows all have
                         // |m_y| > mode_limit.
};

// Forward-only x pass pruning: true when every row ky0 .. ky0 + rows - 1 of a
// workgroup has |m_y| > limit (ky index in (limit, N - limit)).
__device__ __forceinline__ bool x_rows_pruned(int limit, int N, int ky0, int rows) {
    return limit > 0 && ky0 > limit && ky0 + rows - 1 < N - limit;
}

// Streaming store of a complex value (non-temporal).  Used for the passes'
// contiguous row outputs, which nothing re-reads before the next pass; the
// transposing stores keep the default policy (their partial lines merge in L2).
template <typename C>
__device__ __forceinline__ void st_stream(C* p, C v) {
    using E = decltype(v.x);
    typedef E V2 __attribute__((ext_vector_type(2)));
    const V2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<V2*>(p));
}

// Streaming load of a complex value (non-temporal), for the x passes' row
// inputs (on the z and y passes it measured slower).
template <typename C>
__device__ __forceinline__ C ld_stream(const C* p) {
    using E = decltype(p->x);
    typedef E V2 __attribute__((ext_vector_type(2)));
    const V2 w = __builtin_nontemporal_load(reinterpret_cast<const V2*>(p));
    return C{w.x, w.y};
}

// -i k v
template <typename C, typename T>
__device__ __forceinline__ C mul_mi(C v, T k) {
    return C{v.y * k, -v.x * k};
}

// a / b: binary64 exactly; binary32 as a product with the hardware
// reciprocal (within 1 ulp; the binary32 solve carries no bit-for-bit
// contract), two instructions instead of the correctly rounded division.
template <typename T>
__device__ __forceinline__ T green_div(T a, T b) {
    if constexpr (sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b);
    else return a / b;
}

// Green factor from per-axis table values: p = the discrete Laplacian
// eigenvalue per axis (pure PM, src/mesh/green_function.cpp) or the wave
// number k per axis (SG: TreePM long range, src/mesh/green_function.cpp:
// 165-203: continuum 4 pi G / (-k^2), the Gaussian split exp(-k^2 r_s^2), CIC
// deconvolution), w = the per-axis CIC window.  The caller handles k = 0.
template <typename T, int SG>
__device__ __forceinline__ T green_axes(const FftArgs<T>& a, T px, T py, T pz, T wx, T wy, T wz) {
    if constexpr (SG) {
        const T k2 = px * px + py * py + pz * pz;
        T ff = green_div((T)a.green_num, -k2);
        ff *= exp(-k2 * (T)a.rs * (T)a.rs);
        const T w = wx * wy * wz;
        ff = green_div(ff, w * w);
        return ff * (T)a.norm;
    } else {
        T ff = green_div((T)a.green_num, px + py + pz);
        if (a.deconvolve) {
            const T w = wx * wy * wz;
            ff = green_div(ff, w * w);
        }
        return ff * (T)a.norm;
    }
}

template <typename T>
__device__ __forceinline__ T green_factor_treepm(const FftArgs<T>& a, int kx, int ky, int kz) {
    if (kx == 0 && ky == 0 && kz == 0) return (T)0;
    return green_axes<T, 1>(a, a.kc[kx], a.kc[ky], a.kc[kz], a.green_win[kx], a.green_win[ky], a.green_win[kz]);
}

template <typename T>
__device__ __forceinline__ T green_factor(const FftArgs<T>& a, int kx, int ky, int kz) {
    if (kx == 0 && ky == 0 && kz == 0) return (T)0;
    return green_axes<T, 0>(a, a.green_lap[kx], a.green_lap[ky], a.green_lap[kz], a.green_win[kx],
                            a.green_win[ky], a.green_win[kz]);
}

// Boxes of one geometry are laid out box-major; grid row blockIdx.y is the box.
template <typename T>
__device__ __forceinline__ void fft_box(FftArgs<T>& a) {
    const uint64_t b = blockIdx.y;
    a.R += b * a.box_real;
    a.A += b * a.box_cplx;
    a.B += b * a.box_cplx;
}

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

// Binary32 complex pair (16 bytes) for the transposed reads of the inverse
// passes: elements l and l + 1 of a tile row are adjacent in memory and
// 16-byte aligned (l even, rows start at multiples of ROWS), so one lane loads
// both.
typedef float F4 __attribute__((ext_vector_type(4)));

// z pass: real rows -> half spectrum, written transposed into A.
template <typename T, int N, int V>
__global__ void __launch_bounds__((RowsCfg<T, N / 2, N, V>::THREADS)) k_fft_z_r2c(FftArgs<T> a) {
    fft_box(a);
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
    fft_box(a);
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
// SG = 1: TreePM Green, optional -i k_x, output to xout when given.
template <typename T, int N, int V, int SG = 0>
__global__ void __launch_bounds__((RowsCfg<T, N, (N / 2 + 1) * N, V>::THREADS)) k_fft_x_green(FftArgs<T> a) {
    fft_box(a);
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
    constexpr int PER = ROWS * N / NT;
    static_assert(ROWS * N % NT == 0, "x pass covers whole rows");
    if constexpr (!SG && N % NT == 0) {
        // each thread owns the columns kx = tid + j NT of all ROWS rows: the
        // rows' (ky, kz) table values are uniform over the workgroup and its
        // few kx values are loaded once, all ahead of the row fetch
        constexpr int PX = N / NT;
        T lxv[PX], wxv[PX], lyv[ROWS], lzv[ROWS], wyv[ROWS], wzv[ROWS];
        int kyr[ROWS], kzr[ROWS];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            lxv[j] = a.green_lap[tid + j * NT];
            wxv[j] = a.green_win[tid + j * NT];
        }
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const uint64_t q = q0 + r;
            kzr[r] = (int)(q / N);
            kyr[r] = (int)(q % N);
            lyv[r] = a.green_lap[kyr[r]];
            lzv[r] = a.green_lap[kzr[r]];
            wyv[r] = a.green_win[kyr[r]];
            wzv[r] = a.green_win[kzr[r]];
        }
        staged_copy<ROWS * N, NT, C>(
            tid, [&](int l) { return rows[l]; }, [&](int l, C v) { lds[(l / N) * RS + lds_slot(l % N)] = v; });
        __syncthreads();
        fft_rows<C, N, ROWS, RS, -1>(lds, a.tw, 1, tid);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
#pragma unroll
            for (int j = 0; j < PX; ++j) {
                const int kx = tid + j * NT;
                const T f = (kx == 0 && kyr[r] == 0 && kzr[r] == 0)
                                ? (T)0
                                : green_axes<T, 0>(a, lxv[j], lyv[r], lzv[r], wxv[j], wyv[r], wzv[r]);
                C v = lds[r * RS + lds_slot(kx)];
                v.x *= f;
                v.y *= f;
                lds[r * RS + lds_slot(kx)] = v;
            }
        }
    } else {
        staged_copy<ROWS * N, NT, C>(
            tid, [&](int l) { return rows[l]; }, [&](int l, C v) { lds[(l / N) * RS + lds_slot(l % N)] = v; });
        __syncthreads();
        fft_rows<C, N, ROWS, RS, -1>(lds, a.tw, 1, tid);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int l = tid + i * NT;
            const int r = l / N, kx = l % N;
            const uint64_t q = q0 + r;
            const int kz = (int)(q / N), ky = (int)(q % N);
            T f;
            if constexpr (SG) {
                f = green_factor_treepm(a, kx, ky, kz);
            } else if (kx == 0 && ky == 0 && kz == 0) {
                f = (T)0;
            } else {
                f = green_factor(a, kx, ky, kz);
            }
            C v = lds[r * RS + lds_slot(kx)];
            v.x *= f;
            v.y *= f;
            if (SG && a.deriv == 1) v = mul_mi(v, a.kd[kx]);
            lds[r * RS + lds_slot(kx)] = v;
        }
    }
    __syncthreads();
    fft_rows<C, N, ROWS, RS, +1>(lds, a.tw, 1, tid);
    C* orow = (SG && a.xout) ? a.xout + q0 * N : rows;
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return lds[(l / N) * RS + lds_slot(l % N)]; }, [&](int l, C v) { orow[l] = v; });
}

// x pass without a Green multiply: forward (S = -1, spectral analysis) or
// inverse (S = +1, IC displacements) transform, in place on B, unnormalised.
template <typename T, int N, int V, int S>
__global__ void __launch_bounds__((RowsCfg<T, N, (N / 2 + 1) * N, V>::THREADS)) k_fft_x_dft(FftArgs<T> a) {
    fft_box(a);
    using C = cplx<T>;
    constexpr int NH = N / 2 + 1;
    using RC = RowsCfg<T, N, NH * N, V>;
    constexpr int ROWS = RC::ROWS, RS = RC::RS, NT = RC::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t q0 = (uint64_t)a.kz0 * N + (uint64_t)b * ROWS;
    if (x_rows_pruned(a.mode_limit, N, (int)(q0 % N), ROWS)) return;
    const int tid = threadIdx.x;
    C* rows = a.B + q0 * N;
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return rows[l]; }, [&](int l, C v) { lds[(l / N) * RS + lds_slot(l % N)] = v; });
    __syncthreads();
    fft_rows<C, N, ROWS, RS, S>(lds, a.tw, 1, tid);
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return lds[(l / N) * RS + lds_slot(l % N)]; }, [&](int l, C v) { rows[l] = v; });
}

// y pass backward: B[kz][ky][x] -> A rows over y (SG = 1: optional -i k_y on load)
template <typename T, int N, int V, int SG = 0>
__global__ void __launch_bounds__((RowsCfg<T, N, N, V>::THREADS)) k_fft_y_bwd(FftArgs<T> a) {
    fft_box(a);
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
        tid,
        [&](int l) {
            C v = src[(uint64_t)(l / ROWS) * N + (l % ROWS)];
            if (SG && a.deriv == 2) v = mul_mi(v, a.kd[l / ROWS]);
            return v;
        },
        [&](int l, C v) { lds[(l % ROWS) * RS + lds_slot(l / ROWS)] = v; });
    __syncthreads();
    fft_rows<C, N, ROWS, RS, +1>(lds, a.tw, 1, tid);
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return lds[(l / N) * RS + lds_slot(l % N)]; },
        [&](int l, C v) { a.A[((uint64_t)(x0 + l / N) * NH + kz) * N + (l % N)] = v; });
}

// z pass backward (c2r): A[x][kz][y] -> R rows; flags a non-finite potential.
// SG = 1: optional -i k_z on load.
template <typename T, int N, int V, int SG = 0>
__global__ void __launch_bounds__((RowsCfg<T, N / 2, N, V>::THREADS)) k_fft_z_c2r(FftArgs<T> a) {
    fft_box(a);
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
        tid,
        [&](int l) {
            C v = src[(uint64_t)(l / ROWS) * N + (l % ROWS)];
            if (SG && a.deriv == 3) v = mul_mi(v, a.kd[l / ROWS]);
            return v;
        },
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

// ===========================================================================
// Two-level passes (power-of-two meshes).  Each row transform is a length-A
// DFT per stride-B column read straight from HBM into registers, one LDS
// exchange, and a length-B DFT per row written straight back (or through a
// transposing LDS tile where the pass changes the memory layout).
// ===========================================================================

// z pass: real rows (as LEN = N/2 complex) -> half spectrum, transposed into A.
template <typename T, int N, int V, bool PRUNE = false>
__global__ void __launch_bounds__((TLCfg<T, N / 2, N, V>::THREADS)) k_tl_z_r2c(FftArgs<T> a) {
    fft_box(a);
    using C = cplx<T>;
    constexpr int LEN = N / 2, NH = N / 2 + 1;
    using Cfg = TLCfg<T, LEN, N, V>;
    using TL = TwoLevel<LEN>;
    constexpr int A = TL::A, B = TL::B, XS = TL::XS, ROWS = Cfg::ROWS, TPR = Cfg::TPR, NT = Cfg::THREADS;
    constexpr int NAT = Cfg::NAT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, r = tid / TPR, t = tid % TPR;
    C* xr = lds + r * XS;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int yb = b % (N / ROWS), x = b / (N / ROWS);
    const int y0 = yb * ROWS;
    const C* zrow = reinterpret_cast<const C*>(a.R + ((uint64_t)x * N + y0 + r) * N);
    C v[TL::EMAX], w[A];
    if (t < B) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) v[n1] = zrow[t + B * n1];
        tl_twiddle_load<A, LEN>(w, t, a.tw, 2);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (t < B) {
        dft_reg<A, -1>(v);
        tl_twiddle_apply<A, -1>(v, w);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) xr[k1 * (B + 1) + t] = v[k1];
    }
    __syncthreads();
    if (t < A) {
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) v[n2] = xr[t * (B + 1) + n2];
    }
    __syncthreads();
    if (t < A) {
        dft_reg<B, -1>(v);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) lds[r * NAT + t + A * k2] = v[k2];
    }
    // untangle: X[k] = E + W^k O, E = (Z[k]+conj Z[L-k])/2, O = -i/2 (Z[k]-conj Z[L-k]);
    // thread tid forms the outputs l = tid + i NT (k = l / ROWS), its twiddles
    // W^k loaded ahead of the barrier
    const auto untangle = [&](int l, C w) {
        const int rr = l % ROWS, k = l / ROWS;
        const C zk = lds[rr * NAT + (k % LEN)];
        const C zc = cconj(lds[rr * NAT + (LEN - k) % LEN]);
        const C e = {(T)0.5 * (zk.x + zc.x), (T)0.5 * (zk.y + zc.y)};
        const C d = csub(zk, zc);
        const C o = {(T)0.5 * d.y, (T)-0.5 * d.x};
        return cadd(e, cmul(w, o));
    };
    C* dst = a.A + (uint64_t)x * NH * N + y0;
    // PRUNE (spectral estimators): only the planes kz <= mode_limit are stored
    const int kz_cap = PRUNE && a.mode_limit > 0 && a.mode_limit < NH ? a.mode_limit + 1 : NH;
    if constexpr (sizeof(T) == 4 && ROWS % 2 == 0) {
        // pairs (k, rr) and (k, rr + 1), rr even: one twiddle, one 16-byte store
        constexpr int PAIRS = ROWS * NH / 2, PER = (PAIRS + NT - 1) / NT;
        const int pair_cap = kz_cap * (ROWS / 2);
        C wk[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int m = tid + i * NT;
            if ((PRUNE || PAIRS % NT == 0 || m < PAIRS) && (!PRUNE || m < pair_cap)) wk[i] = a.tw[(2 * m) / ROWS];
        }
        __syncthreads();
        F4 q[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int m = tid + i * NT, l = 2 * m;
            if ((PRUNE || PAIRS % NT == 0 || m < PAIRS) && (!PRUNE || m < pair_cap)) {
                const C c0 = untangle(l, wk[i]), c1 = untangle(l + 1, wk[i]);
                q[i] = F4{c0.x, c0.y, c1.x, c1.y};
            }
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int m = tid + i * NT, l = 2 * m;
            if ((PRUNE || PAIRS % NT == 0 || m < PAIRS) && (!PRUNE || m < pair_cap))
                *reinterpret_cast<F4*>(&dst[(uint64_t)(l / ROWS) * N + (l % ROWS)]) = q[i];
        }
    } else {
        constexpr int COUNT = ROWS * NH, PER = (COUNT + NT - 1) / NT;
        const int cap = kz_cap * ROWS;
        C wk[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int l = tid + i * NT;
            if ((PRUNE || COUNT % NT == 0 || l < COUNT) && (!PRUNE || l < cap)) wk[i] = a.tw[l / ROWS];
        }
        __syncthreads();
        C out[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int l = tid + i * NT;
            if ((PRUNE || COUNT % NT == 0 || l < COUNT) && (!PRUNE || l < cap)) out[i] = untangle(l, wk[i]);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int l = tid + i * NT;
            if ((PRUNE || COUNT % NT == 0 || l < COUNT) && (!PRUNE || l < cap)) dst[(uint64_t)(l / ROWS) * N + (l % ROWS)] = out[i];
        }
    }
}

// y pass forward: A rows over y -> B[kz][ky][x] through a transposing tile.
template <typename T, int N, int V>
__global__ void __launch_bounds__((TLCfg<T, N, N, V>::THREADS)) k_tl_y_fwd(FftArgs<T> a) {
    fft_box(a);
    using C = cplx<T>;
    constexpr int LEN = N, NH = N / 2 + 1;
    using Cfg = TLCfg<T, LEN, N, V>;
    using TL = TwoLevel<LEN>;
    constexpr int A = TL::A, B = TL::B, XS = TL::XS, ROWS = Cfg::ROWS, TPR = Cfg::TPR, NT = Cfg::THREADS;
    constexpr int TS = Cfg::TS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, r = tid / TPR, t = tid % TPR;
    C* xr = lds + r * XS;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int xb = b % (N / ROWS), kz = a.kz0 + (int)(b / (N / ROWS));
    const int x0 = xb * ROWS;
    const C* src = a.A + ((uint64_t)(x0 + r) * NH + kz) * N;
    C v[TL::EMAX], w[A];
    if (t < B) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) v[n1] = src[t + B * n1];
        tl_twiddle_load<A, LEN>(w, t, a.tw, 1);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (t < B) {
        dft_reg<A, -1>(v);
        tl_twiddle_apply<A, -1>(v, w);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) xr[k1 * (B + 1) + t] = v[k1];
    }
    __syncthreads();
    if (t < A) {
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) v[n2] = xr[t * (B + 1) + n2];
    }
    __syncthreads();
    if (t < A) {
        dft_reg<B, -1>(v);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) lds[(t + A * k2) * TS + r] = v[k2];
    }
    __syncthreads();
    C* dst = a.B + (uint64_t)kz * N * N + x0;
    staged_copy<ROWS * N, NT, C>(
        tid, [&](int l) { return lds[(l / ROWS) * TS + (l % ROWS)]; },
        [&](int l, C val) { dst[(uint64_t)(l / ROWS) * N + (l % ROWS)] = val; });
}

// x pass: forward x FFT, Green multiply (with 1/N^3), backward x FFT, in place on B.
// SG = 1: TreePM Green, optional -i k_x, output to xout when given.
// SG = 2: TreePM Green, two outputs from one read of the row: the inverse of
// the -i k_x spectrum to xout, the inverse of the plain spectrum in place.
template <typename T, int N, int V, int SG>
__device__ __forceinline__ void tl_x_green_body(FftArgs<T> a) {
    fft_box(a);
    using C = cplx<T>;
    constexpr int LEN = N, NH = N / 2 + 1;
    using Cfg = TLCfg<T, LEN, NH * N, V>;
    using TL = TwoLevel<LEN>;
    constexpr int A = TL::A, B = TL::B, XS = TL::XS, ROWS = Cfg::ROWS, TPR = Cfg::TPR;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, r = tid / TPR, t = tid % TPR;
    C* xr = lds + r * XS;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t q = (uint64_t)a.kz0 * N + (uint64_t)b * ROWS + r;  // row q = kz*N + ky
    const int kz = (int)(q / N), ky = (int)(q % N);
    C* row = a.B + q * N;
    // The Green factors of this thread's spectrum column X[t + A k2] depend on
    // the row alone: their table loads go out first, the row loads right
    // behind them, so the table latency hides under the row's.  (The TreePM
    // factor, with its exponential, is formed in place after the forward
    // transform instead: held across it, it would halve the occupancy.)
    const T* ptab = SG ? a.kc : a.green_lap;
    T px[B], wx[B], py = (T)0, pz = (T)0, wy = (T)0, wz = (T)0;
    if (!SG && t < A) {
        py = ptab[ky];
        pz = ptab[kz];
        wy = a.green_win[ky];
        wz = a.green_win[kz];
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
            px[k2] = ptab[t + A * k2];
            wx[k2] = a.green_win[t + A * k2];
        }
    }
    C v[TL::EMAX];
    if (t < B) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) v[n1] = ld_stream(&row[t + B * n1]);
    }
    __builtin_amdgcn_sched_barrier(0);
    T f[B];
    if (!SG && t < A) {
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
            const int kx = t + A * k2;
            f[k2] = (kx == 0 && ky == 0 && kz == 0) ? (T)0 : green_axes<T, 0>(a, px[k2], py, pz, wx[k2], wy, wz);
        }
    }
    if (t < B) {
        dft_reg<A, -1>(v);
        tl_twiddle<A, LEN, -1>(v, t, a.tw, 1);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) xr[k1 * (B + 1) + t] = v[k1];
    }
    __syncthreads();
    if (t < A) {
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) v[n2] = xr[t * (B + 1) + n2];
    }
    __syncthreads();
    // level-1 twiddles of the backward transform of the held stride-A column
    // (A and B exchanged)
    C w2l[TwSplit<B>::L], w2h[TwSplit<B>::H];
    if (t < A) {
        dft_reg<B, -1>(v);  // v[k2] = X[t + A k2]
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
            if constexpr (SG) f[k2] = green_factor_treepm(a, t + A * k2, ky, kz);
            v[k2].x *= f[k2];
            v[k2].y *= f[k2];
            if (SG == 1 && a.deriv == 1) v[k2] = mul_mi(v[k2], a.kd[t + A * k2]);
        }
        tl_twiddle_load_split<B, LEN>(w2l, w2h, t, a.tw, 1);
    }
    if constexpr (SG == 2) {
        // the -i k_x spectrum's inverse to xout; v keeps the plain spectrum
        // for the in-place inverse below
        C u[TL::EMAX];
        if (t < A) {
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) u[k2] = mul_mi(v[k2], a.kd[t + A * k2]);
            dft_reg<B, +1>(u);
            tl_twiddle_apply_split<B, +1>(u, w2l, w2h);
#pragma unroll
            for (int k1 = 0; k1 < B; ++k1) xr[k1 * (A + 1) + t] = u[k1];
        }
        __syncthreads();
        if (t < B) {
#pragma unroll
            for (int n = 0; n < A; ++n) u[n] = xr[t * (A + 1) + n];
            dft_reg<A, +1>(u);
            C* drow = a.xout + q * N;
#pragma unroll
            for (int k2 = 0; k2 < A; ++k2) st_stream(&drow[t + B * k2], u[k2]);
        }
        __syncthreads();
    }
    if (t < A) {
        // the twiddle loads above complete while the length-B DFT runs
        __builtin_amdgcn_sched_barrier(0);
        dft_reg<B, +1>(v);
        tl_twiddle_apply_split<B, +1>(v, w2l, w2h);
#pragma unroll
        for (int k1 = 0; k1 < B; ++k1) xr[k1 * (A + 1) + t] = v[k1];
    }
    __syncthreads();
    if (t < B) {
#pragma unroll
        for (int n = 0; n < A; ++n) v[n] = xr[t * (A + 1) + n];
        dft_reg<A, +1>(v);
        C* orow = (SG == 1 && a.xout) ? a.xout + q * N : row;
#pragma unroll
        for (int k2 = 0; k2 < A; ++k2) st_stream(&orow[t + B * k2], v[k2]);
    }
}

template <typename T, int N, int V, int SG = 0>
__global__ void __launch_bounds__((TLCfg<T, N, (N / 2 + 1) * N, V>::THREADS)) k_tl_x_green(FftArgs<T> a) {
    tl_x_green_body<T, N, V, SG>(a);
}

// The SG = 2 pass holds two spectral columns per thread.  The LDS allows two
// waves per SIMD; asked for them, the compiler fits the kernel in 256
// registers with a few spilled (left alone it takes 272 and runs one wave per
// SIMD, 1.3x slower at N = 1024).
template <typename T, int N, int V>
__global__ void __launch_bounds__((TLCfg<T, N, (N / 2 + 1) * N, V>::THREADS)) __attribute__((amdgpu_waves_per_eu(2)))
k_tl_x_green_dual(FftArgs<T> a) {
    tl_x_green_body<T, N, V, 2>(a);
}

// x pass without a Green multiply: forward (S = -1, spectral analysis) or
// inverse (S = +1, IC displacements) transform, in place on B, unnormalised.
template <typename T, int N, int V, int S>
__global__ void __launch_bounds__((TLCfg<T, N, (N / 2 + 1) * N, V>::THREADS)) k_tl_x_dft(FftArgs<T> a) {
    fft_box(a);
    using C = cplx<T>;
    constexpr int LEN = N, NH = N / 2 + 1;
    using Cfg = TLCfg<T, LEN, NH * N, V>;
    using TL = TwoLevel<LEN>;
    constexpr int A = TL::A, B = TL::B, XS = TL::XS, ROWS = Cfg::ROWS, TPR = Cfg::TPR;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, r = tid / TPR, t = tid % TPR;
    C* xr = lds + r * XS;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t q = (uint64_t)a.kz0 * N + (uint64_t)b * ROWS + r;
    if (x_rows_pruned(a.mode_limit, N, (int)((q - r) % N), ROWS)) return;
    C* row = a.B + q * N;
    C v[TL::EMAX], w[A];
    if (t < B) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) v[n1] = ld_stream(&row[t + B * n1]);
        tl_twiddle_load<A, LEN>(w, t, a.tw, 1);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (t < B) {
        dft_reg<A, S>(v);
        tl_twiddle_apply<A, S>(v, w);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) xr[k1 * (B + 1) + t] = v[k1];
    }
    __syncthreads();
    if (t < A) {
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) v[n2] = xr[t * (B + 1) + n2];
        dft_reg<B, S>(v);  // v[k2] = X[t + A k2]
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) st_stream(&row[t + A * k2], v[k2]);
    }
}

// y pass backward: B[kz][ky][x] -> A rows over y through a transposing tile
// (SG = 1: optional -i k_y on load; SG = 2: two outputs from one read of B,
// the inverse of the -i k_y column to A and -i k_z times the inverse of the
// plain column to xout, in the layout of A; k_z is uniform over a workgroup).
template <typename T, int N, int V, int SG = 0>
__global__ void __launch_bounds__((TLCfg<T, N, N, V>::THREADS)) k_tl_y_bwd(FftArgs<T> a) {
    fft_box(a);
    using C = cplx<T>;
    constexpr int LEN = N, NH = N / 2 + 1;
    using Cfg = TLCfg<T, LEN, N, V>;
    using TL = TwoLevel<LEN>;
    constexpr int A = TL::A, B = TL::B, XS = TL::XS, ROWS = Cfg::ROWS, TPR = Cfg::TPR, NT = Cfg::THREADS;
    constexpr int TS = Cfg::TS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, r = tid / TPR, t = tid % TPR;
    C* xr = lds + r * XS;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int xb = b % (N / ROWS), kz = a.kz0 + (int)(b / (N / ROWS));
    const int x0 = xb * ROWS;
    const C* src = a.B + (uint64_t)kz * N * N + x0;
    C wl[TwSplit<A>::L], wh[TwSplit<A>::H];
    if (t < B) tl_twiddle_load_split<A, LEN>(wl, wh, t, a.tw, 1);
    if constexpr (sizeof(T) == 4 && ROWS % 2 == 0 && (ROWS * N / 2) % NT == 0) {
        // 16-byte loads of two adjacent x of one ky row (ROWS even: a pair
        // never straddles rows, so one -i k_y factor serves both)
        constexpr int PER = ROWS * N / 2 / NT;
        F4 q[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int l = 2 * (tid + i * NT);
            q[i] = *reinterpret_cast<const F4*>(&src[(uint64_t)(l / ROWS) * N + (l % ROWS)]);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int l = 2 * (tid + i * NT);
            C c0{q[i].x, q[i].y}, c1{q[i].z, q[i].w};
            if (SG == 1 && a.deriv == 2) {
                const T kk = a.kd[l / ROWS];
                c0 = mul_mi(c0, kk);
                c1 = mul_mi(c1, kk);
            }
            lds[(l / ROWS) * TS + (l % ROWS)] = c0;
            lds[(l / ROWS) * TS + (l % ROWS) + 1] = c1;
        }
    } else {
        staged_copy<ROWS * N, NT, C>(
            tid,
            [&](int l) {
                C val = src[(uint64_t)(l / ROWS) * N + (l % ROWS)];
                if (SG == 1 && a.deriv == 2) val = mul_mi(val, a.kd[l / ROWS]);
                return val;
            },
            [&](int l, C val) { lds[(l / ROWS) * TS + (l % ROWS)] = val; });
    }
    __syncthreads();
    C v[TL::EMAX];
    if (t < B) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) v[n1] = lds[(t + B * n1) * TS + r];
    }
    __syncthreads();
    const uint64_t orow = ((uint64_t)(x0 + r) * NH + kz) * N;
    if constexpr (SG == 2) {
        // the -i k_y column's inverse to A; v keeps the plain column for the
        // inverse to xout below (LDS holds the pass to two waves per SIMD, so
        // the second held column costs no occupancy)
        C u[TL::EMAX];
        if (t < B) {
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) u[n1] = mul_mi(v[n1], a.kd[t + B * n1]);
            dft_reg<A, +1>(u);
            tl_twiddle_apply_split<A, +1>(u, wl, wh);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) xr[k1 * (B + 1) + t] = u[k1];
        }
        __syncthreads();
        if (t < A) {
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) u[n2] = xr[t * (B + 1) + n2];
            dft_reg<B, +1>(u);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) st_stream(&a.A[orow + t + A * k2], u[k2]);
        }
        __syncthreads();
    }
    if (t < B) {
        dft_reg<A, +1>(v);
        tl_twiddle_apply_split<A, +1>(v, wl, wh);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) xr[k1 * (B + 1) + t] = v[k1];
    }
    __syncthreads();
    if (t < A) {
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) v[n2] = xr[t * (B + 1) + n2];
        dft_reg<B, +1>(v);
        if constexpr (SG == 2) {
            const T kzf = a.kd[kz];
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) v[k2] = mul_mi(v[k2], kzf);
        }
        C* dst = (SG == 2 ? a.xout : a.A) + orow;
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) st_stream(&dst[t + A * k2], v[k2]);
    }
}

// z pass backward (c2r): A[x][kz][y] -> R rows; flags a non-finite potential.
// SG = 1: optional -i k_z on load.
template <typename T, int N, int V, int SG = 0>
__global__ void __launch_bounds__((TLCfg<T, N / 2, N, V>::THREADS)) k_tl_z_c2r(FftArgs<T> a) {
    fft_box(a);
    using C = cplx<T>;
    constexpr int LEN = N / 2, NH = N / 2 + 1;
    using Cfg = TLCfg<T, LEN, N, V>;
    using TL = TwoLevel<LEN>;
    constexpr int A = TL::A, B = TL::B, XS = TL::XS, ROWS = Cfg::ROWS, TPR = Cfg::TPR, NT = Cfg::THREADS;
    constexpr int NAT = Cfg::NAT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, r = tid / TPR, t = tid % TPR;
    C* xr = lds + r * XS;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int yb = b % (N / ROWS), x = b / (N / ROWS);
    const int y0 = yb * ROWS;
    const C* src = a.A + (uint64_t)x * NH * N + y0;
    // the untangle twiddles W^k of this thread's column (k = t + B n1) and its
    // level-1 twiddles, in flight with the row loads
    C wu[A], wl[TwSplit<A>::L], wh[TwSplit<A>::H];
    if (t < B) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) wu[n1] = a.tw[t + B * n1];
        tl_twiddle_load_split<A, LEN>(wl, wh, t, a.tw, 2);
    }
    const auto put = [&](int l, C val) {
        const int rr = l % ROWS, k = l / ROWS;
        if (k == 0 || k == LEN) val.y = (T)0;  // c2r reads only the independent half
        lds[rr * NAT + k] = val;
    };
    if constexpr (sizeof(T) == 4 && ROWS % 2 == 0) {
        // 16-byte loads of two adjacent y of one kz row (ROWS even: one -i k_z
        // factor serves the pair)
        constexpr int PAIRS = ROWS * NH / 2, PER = (PAIRS + NT - 1) / NT;
        F4 q[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int m = tid + i * NT;
            const int l = 2 * m;
            if (PAIRS % NT == 0 || m < PAIRS) q[i] = *reinterpret_cast<const F4*>(&src[(uint64_t)(l / ROWS) * N + (l % ROWS)]);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int m = tid + i * NT;
            const int l = 2 * m;
            if (PAIRS % NT == 0 || m < PAIRS) {
                C c0{q[i].x, q[i].y}, c1{q[i].z, q[i].w};
                if (SG && a.deriv == 3) {
                    const T kk = a.kd[l / ROWS];
                    c0 = mul_mi(c0, kk);
                    c1 = mul_mi(c1, kk);
                }
                put(l, c0);
                put(l + 1, c1);
            }
        }
    } else {
        staged_copy<ROWS * NH, NT, C>(
            tid,
            [&](int l) {
                C val = src[(uint64_t)(l / ROWS) * N + (l % ROWS)];
                if (SG && a.deriv == 3) val = mul_mi(val, a.kd[l / ROWS]);
                return val;
            },
            put);
    }
    __syncthreads();
    C v[TL::EMAX];
    if (t < B) {
        // Z'[k] = (X[k] + conj X[L-k]) + i (X[k] - conj X[L-k]) W^-k
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) {
            const int k = t + B * n1;
            const C xk = lds[r * NAT + k];
            const C xm = lds[r * NAT + LEN - k];
            const C s1 = cadd(xk, cconj(xm));
            const C d1 = cmul(csub(xk, cconj(xm)), cconj(wu[n1]));
            v[n1] = C{s1.x - d1.y, s1.y + d1.x};
        }
    }
    __syncthreads();
    if (t < B) {
        dft_reg<A, +1>(v);
        tl_twiddle_apply_split<A, +1>(v, wl, wh);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) xr[k1 * (B + 1) + t] = v[k1];
    }
    __syncthreads();
    int mybad = 0;
    if (t < A) {
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) v[n2] = xr[t * (B + 1) + n2];
        dft_reg<B, +1>(v);
        C* dst = reinterpret_cast<C*>(a.R + ((uint64_t)x * N + y0 + r) * N);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
            if (!isfinite(v[k2].x) || !isfinite(v[k2].y)) mybad = 1;
            st_stream(&dst[t + A * k2], v[k2]);
        }
    }
    if (__ballot(mybad) && lane_id() == 0) atomicOr(a.err, ERR_FIELD_NONFINITE);
}

}  // namespace hy

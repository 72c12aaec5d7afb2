// LDS-staged pencil FFT building blocks for the PM Poisson solve (gfx950).
//
// A workgroup holds ROWS rows of LEN complex values in LDS (row stride
// LEN+1) and runs an in-place Stockham autosort transform over all rows at
// once: every stage each thread loads E/R radix-R butterflies into registers,
// twiddles and transforms them, then (after a barrier) stores them at their
// autosorted positions.  LEN = 2^a or 3*2^a; radices 3, 8, 4, 2.
//
// Twiddles come from one table tw[m] = exp(-2*pi*i*m/NT) (NT = mesh size),
// evaluated in binary64 on the host, read with a stride NT/LEN.
//
// Conventions follow the reference FFT backend (src/mesh/fft_backend.cpp:
// 1088-1145): forward transforms are unnormalized, the inverse is the
// unnormalized backward DFT; the 1/N^3 factor is folded into the Green
// multiply of the x pass.
#pragma once

#include "hy_device.hpp"

namespace hy {

// ---------------------------------------------------------------------------
// Compile-time plan
// ---------------------------------------------------------------------------
template <int LEN>
struct FftPlan {
    static constexpr bool has3 = (LEN % 3) == 0;
    static constexpr int E = has3 ? (LEN < 24 ? LEN : 24) : (LEN < 16 ? LEN : 16);
    static constexpr int TPR = LEN / E;  // threads per row
    static constexpr int radix(int s) {
        int rem = LEN;
        int r = 1;
        for (int i = 0; i <= s; ++i) {
            if (rem == 1) return 1;
            if (i == 0 && rem % 3 == 0) r = 3;
            else if (rem % 8 == 0) r = 8;
            else if (rem % 4 == 0) r = 4;
            else r = 2;
            rem /= r;
        }
        return r;
    }
    static constexpr int stages() {
        int s = 0;
        while (radix(s) != 1) ++s;
        return s;
    }
    static constexpr int prod_before(int s) {
        int p = 1;
        for (int i = 0; i < s; ++i) p *= radix(i);
        return p;
    }
};

// ---------------------------------------------------------------------------
// Complex helpers
// ---------------------------------------------------------------------------
template <typename C>
HY_HD C cadd(C a, C b) { return {a.x + b.x, a.y + b.y}; }
template <typename C>
HY_HD C csub(C a, C b) { return {a.x - b.x, a.y - b.y}; }
template <typename C>
HY_HD C cmul(C a, C b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C>
HY_HD C cconj(C a) { return {a.x, -a.y}; }
// multiply by s*i, s = +1 or -1
template <int S, typename C>
HY_HD C cmuli(C a) {
    if constexpr (S > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}

// In-register DFTs of radix 2, 3, 4, 8.  S = -1 forward, +1 backward.
template <int S, typename C>
HY_HD void dft2(C* u) {
    const C a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
}

template <int S, typename C>
HY_HD void dft3(C* u) {
    using T = decltype(u[0].x);
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    const C t1 = cadd(u[1], u[2]);
    const C t2 = {u[0].x - (T)0.5 * t1.x, u[0].y - (T)0.5 * t1.y};
    const C d0 = csub(u[1], u[2]);
    const C d = {(T)S * h * d0.x, (T)S * h * d0.y};
    const C id = {-d.y, d.x};
    u[0] = cadd(u[0], t1);
    u[1] = cadd(t2, id);
    u[2] = csub(t2, id);
}

template <int S, typename C>
HY_HD void dft4(C* u) {
    const C t0 = cadd(u[0], u[2]);
    const C t1 = csub(u[0], u[2]);
    const C t2 = cadd(u[1], u[3]);
    const C t3 = cmuli<S>(csub(u[1], u[3]));
    u[0] = cadd(t0, t2);
    u[2] = csub(t0, t2);
    u[1] = cadd(t1, t3);
    u[3] = csub(t1, t3);
}

template <int S, typename C>
HY_HD void dft8(C* u) {
    using T = decltype(u[0].x);
    C e[4] = {u[0], u[2], u[4], u[6]};
    C o[4] = {u[1], u[3], u[5], u[7]};
    dft4<S>(e);
    dft4<S>(o);
    const T r = (T)0.70710678118654752440;
    // w8^q for q = 1, 2, 3 with w8 = exp(S*2*pi*i/8)
    const C o1 = {r * (o[1].x - (T)S * o[1].y), r * ((T)S * o[1].x + o[1].y)};
    const C o2 = cmuli<S>(o[2]);
    const C o3 = {r * (-o[3].x - (T)S * o[3].y), r * ((T)S * o[3].x - o[3].y)};
    u[0] = cadd(e[0], o[0]);
    u[4] = csub(e[0], o[0]);
    u[1] = cadd(e[1], o1);
    u[5] = csub(e[1], o1);
    u[2] = cadd(e[2], o2);
    u[6] = csub(e[2], o2);
    u[3] = cadd(e[3], o3);
    u[7] = csub(e[3], o3);
}

template <int R, int S, typename C>
HY_HD void dftR(C* u) {
    if constexpr (R == 2) dft2<S>(u);
    else if constexpr (R == 3) dft3<S>(u);
    else if constexpr (R == 4) dft4<S>(u);
    else dft8<S>(u);
}

// ---------------------------------------------------------------------------
// Register DFTs of length 2..32 (natural order in and out) and the two-level
// row FFT built on them.
// ---------------------------------------------------------------------------
// exp(2 pi i m / 32), m = 0..31 (cos, sin), evaluated in binary64.
constexpr double kW32[32][2] = {
    {                     1.0,                      0.0},
    {      0.9807852804032304,      0.19509032201612825},
    {      0.9238795325112867,       0.3826834323650898},
    {      0.8314696123025452,       0.5555702330196022},
    {      0.7071067811865476,       0.7071067811865475},
    {      0.5555702330196023,       0.8314696123025452},
    {     0.38268343236508984,       0.9238795325112867},
    {     0.19509032201612833,       0.9807852804032304},
    {   6.123233995736766e-17,                      1.0},
    {     -0.1950903220161282,       0.9807852804032304},
    {     -0.3826834323650897,       0.9238795325112867},
    {      -0.555570233019602,       0.8314696123025455},
    {     -0.7071067811865475,       0.7071067811865476},
    {     -0.8314696123025453,       0.5555702330196022},
    {     -0.9238795325112867,       0.3826834323650899},
    {     -0.9807852804032304,       0.1950903220161286},
    {                    -1.0,   1.2246467991473532e-16},
    {     -0.9807852804032304,     -0.19509032201612836},
    {     -0.9238795325112868,     -0.38268343236508967},
    {     -0.8314696123025455,       -0.555570233019602},
    {     -0.7071067811865477,      -0.7071067811865475},
    {     -0.5555702330196022,      -0.8314696123025452},
    {    -0.38268343236509034,      -0.9238795325112865},
    {    -0.19509032201612866,      -0.9807852804032303},
    { -1.8369701987210297e-16,                     -1.0},
    {      0.1950903220161283,      -0.9807852804032304},
    {        0.38268343236509,      -0.9238795325112866},
    {      0.5555702330196018,      -0.8314696123025455},
    {      0.7071067811865474,      -0.7071067811865477},
    {      0.8314696123025452,      -0.5555702330196022},
    {      0.9238795325112865,      -0.3826834323650904},
    {      0.9807852804032303,     -0.19509032201612872},
};

// exp(S 2 pi i m / R) for R | 32, as a complex constant of type C.
template <int R, int S, typename C>
HY_HD C wconst(int m) {
    using T = decltype(C{}.x);
    const int q = ((m % R) + R) % R * (32 / R);
    return C{(T)kW32[q][0], (T)(S * kW32[q][1])};
}

// In-register DFT of length R (power of two <= 32), natural order in place:
// R = A * B with a length-A DFT over each stride-B column, the twiddle
// exp(S 2 pi i n2 k1 / R), then a length-B DFT over each row.
template <int R, int S, typename C>
HY_HD void dft_reg(C* u) {
    if constexpr (R <= 8) {
        dftR<R, S>(u);
    } else {
        constexpr int A = 4, B = R / 4;
        C y[R];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) {
            C col[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) col[n1] = u[B * n1 + n2];
            dft4<S>(col);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1)
                y[k1 * B + n2] = (k1 * n2 == 0) ? col[k1] : cmul(col[k1], wconst<R, S, C>(k1 * n2));
        }
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            C row[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) row[n2] = y[k1 * B + n2];
            dft_reg<B, S>(row);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) u[k1 + A * k2] = row[k2];
        }
    }
}

// Two-level row FFT of a power-of-two length LEN = A * B (A, B <= 32):
//   level 1: thread n2 < B holds the stride-B column x[n2 + B n1] (n1 < A),
//            transforms it (length A) and multiplies by W_LEN^(n2 k1);
//   exchange: Y[k1][n2] through LDS (row stride B + 1);
//   level 2: thread k1 < A holds row Y[k1][*], transforms it (length B) and
//            then holds X[k1 + A k2], k2 < B.
// The inverse of a spectrum held as in level 2 is the same procedure with
// A and B exchanged.
template <int LEN>
struct TwoLevel {
    static constexpr int lg(int v) { return v <= 1 ? 0 : 1 + lg(v / 2); }
    static constexpr int A = 1 << (lg(LEN) / 2);
    static constexpr int B = LEN / A;
    static constexpr int TPR = A > B ? A : B;
    static constexpr int EMAX = A > B ? A : B;
    // exchange region per row (complex): A rows of B+1 or B rows of A+1
    static constexpr int XS0 = (A * (B + 1) > B * (A + 1) ? A * (B + 1) : B * (A + 1));
    static constexpr int XS = XS0 % 2 == 0 ? XS0 + 1 : XS0;
    static_assert(A * B == LEN && A <= 32 && B <= 32 && A >= 2, "two-level FFT length");
};

// Level-1 twiddle: v[k1] *= W_LEN^(S col k1), twiddles from the table
// tw[m * stride] = exp(-2 pi i m / LEN).
// Level-1 twiddles split into the loads, issued with a pass's row loads so
// their latency hides under the row fetch, and the multiplies, which match
// tl_twiddle's arithmetic.
template <int A, int LEN, typename C>
HY_HD void tl_twiddle_load(C* w, int col, const C* tw, int stride) {
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) w[k1] = tw[((col * k1) % LEN) * stride];
}

template <int A, int S, typename C>
HY_HD void tl_twiddle_apply(C* v, const C* w) {
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) {
        C x = w[k1];
        if constexpr (S > 0) x.y = -x.y;
        v[k1] = cmul(v[k1], x);
    }
}

// Level-1 twiddles W^(col k1), k1 = 1 .. A-1, as the products
// W^(col j) W^(col L h) (k1 = j + L h) of L - 1 + H - 1 table entries instead
// of A - 1: for A = 32 a row holds 10 complex factors in registers rather
// than 31 (the registers that bound the passes' occupancy), at one rounding
// more per twiddle.
template <int A>
struct TwSplit {
    static constexpr int L = A >= 16 ? 8 : A;
    static constexpr int H = (A + L - 1) / L;
};

template <int A, int LEN, typename C>
HY_HD void tl_twiddle_load_split(C* lo, C* hi, int col, const C* tw, int stride) {
    constexpr int L = TwSplit<A>::L, H = TwSplit<A>::H;
#pragma unroll
    for (int j = 1; j < L; ++j) lo[j] = tw[((col * j) % LEN) * stride];
#pragma unroll
    for (int h = 1; h < H; ++h) hi[h] = tw[((col * L * h) % LEN) * stride];
}

template <int A, int S, typename C>
HY_HD void tl_twiddle_apply_split(C* v, const C* lo, const C* hi) {
    constexpr int L = TwSplit<A>::L;
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) {
        const int j = k1 % L, h = k1 / L;
        C x = h == 0 ? lo[j] : (j == 0 ? hi[h] : cmul(lo[j], hi[h]));
        if constexpr (S > 0) x.y = -x.y;
        v[k1] = cmul(v[k1], x);
    }
}

// The table loads go out in batches of TB ahead of their multiplies (a
// scheduling barrier keeps the compiler from interleaving them with the
// uses), so a row pays (A - 1) / TB memory round trips for its twiddles rather
// than up to one per twiddle.
template <int A, int LEN, int S, int TB = 16, typename C>
HY_HD void tl_twiddle(C* v, int col, const C* tw, int stride) {
#pragma unroll
    for (int k0 = 1; k0 < A; k0 += TB) {
        C w[TB];
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            if (k0 + j < A) w[j] = tw[((col * (k0 + j)) % LEN) * stride];
        }
#ifdef __HIP_DEVICE_COMPILE__
        __builtin_amdgcn_sched_barrier(0);
#endif
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            if (k0 + j < A) {
                C x = w[j];
                if constexpr (S > 0) x.y = -x.y;
                v[k0 + j] = cmul(v[k0 + j], x);
            }
        }
    }
}

// LDS slot of logical element j of a row: one padding slot every 32 entries,
// so the stride-R stores of the first Stockham stage and the row-strided
// transposes spread over all banks.
__device__ __forceinline__ constexpr int lds_slot(int j) { return j + (j >> 5); }

template <int LEN>
struct LdsRow {
    static constexpr int SPAN = LEN + (LEN >> 5) + 1;          // LEN+1 logical entries
    static constexpr int RS = (SPAN % 2 == 0) ? SPAN + 1 : SPAN;  // odd stride in complex units
};

// One Stockham stage over all rows held in LDS.
template <typename C, int LEN, int ROWS, int RS, int STAGE, int S>
__device__ __forceinline__ void fft_stage(C* lds, const C* __restrict__ tw, int tw_stride,
                                          int tid) {
    using P = FftPlan<LEN>;
    constexpr int R = P::radix(STAGE);
    constexpr int p = P::prod_before(STAGE);
    constexpr int M = LEN / R;
    constexpr int NB = P::E / R;  // butterflies per thread
    constexpr int TPR = P::TPR;
    const int row = tid / TPR;
    const int t = tid % TPR;
    C* x = lds + row * RS;
    C v[P::E];
    if (row < ROWS) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int i = t + j * TPR;
            const int k = i % p;
            C u[R];
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = x[lds_slot(i + r * M)];
            if (k != 0) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const int m = (r * k * (LEN / (p * R))) % LEN;
                    C w = tw[m * tw_stride];
                    if constexpr (S > 0) w.y = -w.y;
                    u[r] = cmul(u[r], w);
                }
            }
            dftR<R, S>(u);
#pragma unroll
            for (int r = 0; r < R; ++r) v[j * R + r] = u[r];
        }
    }
    __syncthreads();
    if (row < ROWS) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int i = t + j * TPR;
            const int k = i % p;
            const int base = (i - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) x[lds_slot(base + r * p)] = v[j * R + r];
        }
    }
    __syncthreads();
}

template <typename C, int LEN, int ROWS, int RS, int S, int STAGE = 0>
__device__ __forceinline__ void fft_rows(C* lds, const C* __restrict__ tw, int tw_stride,
                                         int tid) {
    if constexpr (STAGE < FftPlan<LEN>::stages()) {
        fft_stage<C, LEN, ROWS, RS, STAGE, S>(lds, tw, tw_stride, tid);
        fft_rows<C, LEN, ROWS, RS, S, STAGE + 1>(lds, tw, tw_stride, tid);
    }
}

// ---------------------------------------------------------------------------
// Per-pass launch geometry
// ---------------------------------------------------------------------------
// Launch variants of every pass (chosen per pass by the autotuner): an LDS
// budget per workgroup, which sets the rows per workgroup and therefore the
// workgroups resident per CU (160 KB of LDS each) and the length of the
// contiguous runs of the transposing stores.
constexpr int kFftVariants = 4;
constexpr int fft_lds_budget(int v) {
    return v == 0 ? 80 * 1024 : v == 1 ? 40 * 1024 : v == 2 ? 136 * 1024 : 20 * 1024;
}
constexpr int fft_thread_cap(int v) { return v == 2 ? 1024 : 512; }

constexpr int pow2_floor(int v) {
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}

// Rows per workgroup for a transform of length LEN whose rows are blocked from
// a dimension of extent DIM (ROWS must divide DIM), launch variant V.
template <typename T, int LEN, int DIM, int V = 0>
struct RowsCfg {
    using C = cplx<T>;
    static constexpr int RS = LdsRow<LEN>::RS;
    static constexpr int by_threads = fft_thread_cap(V) / FftPlan<LEN>::TPR;
    static constexpr int by_lds = fft_lds_budget(V) / (RS * (int)sizeof(C));
    static constexpr int cap = by_threads < by_lds ? by_threads : by_lds;
    static constexpr int dim_pow2 = (DIM & (-DIM));  // largest power of two dividing DIM
    static constexpr int ROWS0 = pow2_floor(cap < 1 ? 1 : cap);
    static constexpr int ROWS = ROWS0 < dim_pow2 ? ROWS0 : dim_pow2;
    static constexpr int THREADS = ROWS * FftPlan<LEN>::TPR;
    static constexpr int LDS_BYTES = ROWS * RS * (int)sizeof(C);
    static_assert(LDS_BYTES <= 160 * 1024, "LDS per workgroup");
};

// Launch geometry of the two-level passes: ROWS rows per workgroup, TPR
// threads per row; the LDS holds either the per-row exchange (XS complex per
// row), a natural-order copy of the rows (LEN + 1 per row) or the transposed
// tile of the y passes (LEN x (ROWS + 1)).
template <typename T, int LEN, int DIM, int V = 0>
struct TLCfg {
    using C = cplx<T>;
    using TL = TwoLevel<LEN>;
    static constexpr int TPR = TL::TPR;
    static constexpr int PER_ROW = (TL::XS > LEN + 1 ? TL::XS : LEN + 1) + LEN / 8 + 1;
    static constexpr int by_threads = fft_thread_cap(V) / TPR;
    static constexpr int by_lds = fft_lds_budget(V) / (PER_ROW * (int)sizeof(C));
    static constexpr int cap = by_threads < by_lds ? by_threads : by_lds;
    static constexpr int dim_pow2 = (DIM & (-DIM));
    static constexpr int ROWS0 = pow2_floor(cap < 1 ? 1 : cap);
    static constexpr int ROWS = ROWS0 < dim_pow2 ? ROWS0 : dim_pow2;
    static constexpr int THREADS = ROWS * TPR;
    static constexpr int NAT = LEN + 1;  // natural-order row stride (odd)
    static constexpr int TS = ROWS + 1;  // transposed tile row stride
    static constexpr int ELEMS = (ROWS * TL::XS > ROWS * NAT ? ROWS * TL::XS : ROWS * NAT) > LEN * TS
                                     ? (ROWS * TL::XS > ROWS * NAT ? ROWS * TL::XS : ROWS * NAT)
                                     : LEN * TS;
    static constexpr int LDS_BYTES = ELEMS * (int)sizeof(C);
    // the x passes use only the per-row exchange regions (no transposed tile)
    static constexpr int X_LDS_BYTES = ROWS * TL::XS * (int)sizeof(C);
    static_assert(LDS_BYTES <= 160 * 1024, "LDS per workgroup");
};

}  // namespace hy

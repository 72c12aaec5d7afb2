// Matter power-spectrum estimator kernels (gfx950, wave64): shell binning of
// the r2c spectrum of the CIC density contrast.
//
// The per-mode arithmetic follows MatterPowerSpectrum::compute_pk
// (src/analysis/matter_power_spectrum.cpp:127-206): |k| as
// hypot(hypot(kx, ky), kz) (include/cosmo_nbody/core/types.hpp:42-44), the
// logarithmic shell of analysis/spectral_numeric.hpp:124-141, raw power
// L^3 |delta_k|^2, the CIC power deconvolution 1 / W(k)^2, the aliased shot
// noise S prod_i (1 - 2/3 sin^2(k_i dx / 2)), multiplicity 2 off the kz = 0
// and Nyquist planes.  Shell sums are reduced deterministically: one wave owns
// each row and its own LDS accumulator, waves and then workgroups are summed
// in a fixed order.
#pragma once

#include "hy_device.hpp"

namespace hy {

// hypot with the correction step of C. F. Borges, "An Improved Algorithm for
// hypot(a,b)" (arXiv:1904.09481, 2019), the variant without fused
// multiply-add: the result of glibc's hypot, on which the reference's |k|
// rests, for operands whose magnitudes lie in [2^-450, 2^450] or are zero.
// Built with -ffp-contract=off (no contraction of the correction terms).
HY_HD double hypot_corrected(double x, double y) {
    x = fabs(x);
    y = fabs(y);
    const double ax = x < y ? y : x;
    const double ay = x < y ? x : y;
    if (ay <= ax * 0x1p-54) return ax + ay;
    double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        const double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        const double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}

// Shell of |k| in edges[0..nb]: [e_i, e_i+1), the last shell closed; -1 outside.
HY_HD int shell_index(const double* edges, int nb, double v) {
    if (v < edges[0] || v > edges[nb]) return -1;
    if (v == edges[nb]) return nb - 1;
    int lo = 0, hi = nb + 1;  // first edge above v
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (edges[mid] > v) hi = mid;
        else lo = mid + 1;
    }
    return lo - 1;
}

constexpr int PK_Q = 5;  // per shell: sum k, sum p_raw, sum p_corrected, sum deconvolution, modes
constexpr int PK_BLOCK = 256;

struct PkArgs {
    const void* m1;       // spectrum [kz][ky][kx] (the x pass output), binary32 or binary64 complex
    const void* m2;       // interlaced partner spectrum, or nullptr
    const double* kc;     // k component per mesh index (N)
    const double* win;    // per-axis CIC window sinc^2(k dx / 2) (N)
    const double* ali;    // per-axis alias factor 1 - 2/3 sin^2(k dx / 2) (N)
    const double* edges;  // nb + 1 shell edges
    double* partials;     // [gridDim.x][nb][PK_Q]
    uint32_t* err;
    int N;
    int nb;
    int waves;            // waves per workgroup owning an LDS accumulator each
    int subtract;         // subtract the aliased shot noise
    double kmax;          // largest binned |k|
    double kf;            // fundamental 2 pi / L
    double perp_skip;     // rows with hypot(ky, kz) above this hold no binned mode
    double amp;           // transform value -> delta_k (L^3 / (M N^3) for a density-source mesh)
    double volume;        // L^3
    double shot;          // shot-noise reference
    double shift;         // interlacing shift (dx / 2)
    uint64_t row_step;    // visiting order q = (k row_step) mod rows, row_step coprime to rows
};

template <typename T>
__global__ void __launch_bounds__(PK_BLOCK) k_pk_bin(PkArgs a) {
    extern __shared__ __align__(16) double pk_lds[];  // [waves][nb][PK_Q], then nb + 1 edges
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    const int slab = a.nb * PK_Q;
    double* edges = pk_lds + (size_t)a.waves * slab;
    for (int i = threadIdx.x; i < a.waves * slab; i += blockDim.x) pk_lds[i] = 0.0;
    for (int i = threadIdx.x; i <= a.nb; i += blockDim.x) edges[i] = a.edges[i];
    __syncthreads();
    using C = cplx<T>;
    const C* m1 = static_cast<const C*>(a.m1);
    const C* m2 = static_cast<const C*>(a.m2);
    const int N = a.N, NH = N / 2 + 1;
    const bool even = (N & 1) == 0;
    const uint64_t rows = (uint64_t)NH * (uint64_t)N;
    const uint64_t all_waves = (uint64_t)gridDim.x * (uint64_t)a.waves;
    bool bad = false;
    if (wave < a.waves) {
        double* acc = pk_lds + (size_t)wave * slab;
        // The binned rows form a disk of small |k_perp|.  Strided over the
        // rows in storage order (a stride that is a multiple of N), a wave
        // would see one ky index in every row it owns: the waves on the disk's
        // ky would bin every mode and the rest none.  Each wave takes a
        // contiguous block of the permuted order q = (k row_step) mod rows
        // instead, whose consecutive rows lie row_step (about rows / golden
        // ratio) apart: every wave samples the whole half spectrum.
        const uint64_t per = (rows + all_waves - 1) / all_waves;
        const uint64_t k0 = ((uint64_t)blockIdx.x * a.waves + wave) * per;
        const uint64_t k1 = k0 + per < rows ? k0 + per : rows;
        uint64_t q = k0 < rows ? (k0 * a.row_step) % rows : 0;
        for (uint64_t k = k0; k < k1; ++k, q = q + a.row_step >= rows ? q + a.row_step - rows : q + a.row_step) {
            const int iz = (int)(q / (uint64_t)N), iy = (int)(q % (uint64_t)N);
            const double ky = a.kc[iy], kz = a.kc[iz];
            if (hypot_corrected(ky, kz) > a.perp_skip) continue;  // uniform over the wave
            const double mult = (iz == 0 || (even && iz == N / 2)) ? 1.0 : 2.0;
            // |m_x| <= sqrt(kmax^2 - k_perp^2) / k_f (+1 for rounding): the
            // x indices [0, mxm] and [N - mxm, N) hold every binned mode of the row
            const double rem = a.kmax * a.kmax - (ky * ky + kz * kz);
            const int mxm = min(N / 2, (int)(sqrt(fmax(rem, 0.0)) / a.kf) + 1);
            const int lo2 = max(N - mxm, N / 2 + 1);
            const int span1 = mxm + 1, span = span1 + (N - lo2);
            for (int c0 = 0; c0 < span; c0 += 64) {
                const int c = c0 + lane;
                const int ix = c < span1 ? c : lo2 + (c - span1);
                int shell = -1;
                double v[PK_Q];
                if (c < span) {
                    const double kx = a.kc[ix];
                    const double kmag = hypot_corrected(hypot_corrected(kx, ky), kz);
                    if (kmag > 0.0 && !(kmag > a.kmax)) shell = shell_index(edges, a.nb, kmag);
                    if (shell >= 0) {
                        const C u = m1[q * (uint64_t)N + ix];
                        double re = (double)u.x * a.amp, im = (double)u.y * a.amp;
                        if (m2) {
                            // 0.5 (m1 + m2 e^{i phase}), phase = -(kx + ky + kz) shift
                            // (src/analysis/fourier_density_field.cpp:314-326)
                            const C w = m2[q * (uint64_t)N + ix];
                            const double r2 = (double)w.x * a.amp, i2 = (double)w.y * a.amp;
                            const double ph = -(kx + ky + kz) * a.shift;
                            double sn, cs;
                            sincos(ph, &sn, &cs);
                            re = 0.5 * (re + (r2 * cs - i2 * sn));
                            im = 0.5 * (im + (r2 * sn + i2 * cs));
                        }
                        const double praw = (re * re + im * im) * a.volume;
                        const double win = a.win[ix] * a.win[iy] * a.win[iz];
                        const double dec = 1.0 / (win * win);
                        double pc = praw;
                        if (a.subtract) pc = praw - a.shot * a.ali[ix] * a.ali[iy] * a.ali[iz];
                        pc *= dec;
                        if (!isfinite(praw) || !isfinite(pc) || !isfinite(re) || !isfinite(im)) {
                            bad = true;
                            shell = -1;
                        } else {
                            v[0] = kmag * mult;
                            v[1] = praw * mult;
                            v[2] = pc * mult;
                            v[3] = dec * mult;
                            v[4] = mult;
                        }
                    }
                }
                // one fixed-order butterfly sum per distinct shell in this chunk
                unsigned long long pending = __ballot(shell >= 0);
                while (pending) {
                    const int leader = __ffsll((long long)pending) - 1;
                    const int s = __builtin_amdgcn_readlane(shell, leader);
                    const bool mine = shell == s;
                    double t[PK_Q];
#pragma unroll
                    for (int k = 0; k < PK_Q; ++k) t[k] = mine ? v[k] : 0.0;
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
                        for (int k = 0; k < PK_Q; ++k) t[k] += __shfl_xor(t[k], off);
                    }
                    if (lane == 0) {
#pragma unroll
                        for (int k = 0; k < PK_Q; ++k) acc[s * PK_Q + k] += t[k];
                    }
                    pending &= ~__ballot(mine);
                }
            }
        }
    }
    if (bad) atomicOr(a.err, ERR_SPECTRUM_NONFINITE);
    __syncthreads();
    for (int i = threadIdx.x; i < slab; i += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < a.waves; ++w) s += pk_lds[(size_t)w * slab + i];
        a.partials[(size_t)blockIdx.x * slab + i] = s;
    }
}

// Fixed-order sum of the workgroup partials, one workgroup per output:
// out[i] = sum_g partials[g][i] (strided per-thread sums, then a fixed tree).
__global__ void __launch_bounds__(256) k_pk_reduce(const double* __restrict__ partials, int groups, int slab,
                                                   double* __restrict__ out) {
    __shared__ double red[256];
    const int i = blockIdx.x;
    double s = 0.0;
    for (int g = threadIdx.x; g < groups; g += blockDim.x) s += partials[(size_t)g * slab + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = red[0];
}

}  // namespace hy

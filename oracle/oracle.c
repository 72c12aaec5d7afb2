/*
 * CPU oracle for the HYOWON particle-mesh gravity step (binary64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never part of it.
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  The restated
 * arithmetic keeps the reference's operation order; reference citations are
 * file:line relative to the reference source tree.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;

void or_set_threads(int threads) { g_threads = threads; }

int or_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

#define OR_NT or_get_threads()

static double g_stage_s[OR_STAGES];

static double wall_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

void or_stage_reset(void) { memset(g_stage_s, 0, sizeof g_stage_s); }

void or_stage_times(double* out) { memcpy(out, g_stage_s, sizeof g_stage_s); }

/* ------------------------------------------------------------------------- */
/* Constants: include/cosmo_nbody/cosmology/units.hpp:14-30                   */
/* ------------------------------------------------------------------------- */
static const double OR_PI = 3.14159265358979323846;

double or_gravitational_constant(void) {
    const double g_si = 6.67430e-11;
    const double mpc_m = 3.085677581491367e22;
    const double msun_kg = 1.98847e30;
    const double mass_unit_kg = 1e10 * msun_kg;
    const double velocity_unit = 1000.0;
    return g_si * mass_unit_kg / (mpc_m * velocity_unit * velocity_unit);
}

double or_rho_crit0(void) {
    const double h0 = 100.0;
    return 3.0 * h0 * h0 / (8.0 * OR_PI * or_gravitational_constant());
}

/* ------------------------------------------------------------------------- */
/* Periodic geometry: include/cosmo_nbody/math/periodic_box.hpp:27-61         */
/* ------------------------------------------------------------------------- */
double or_wrap(double x, double L) {
    if (!isfinite(x) || !isfinite(L) || L <= 0.0) return NAN;
    if (x >= 0.0 && x < L) return x;
    double w = fmod(x, L);
    if (!isfinite(w)) return NAN;
    if (w < 0.0) w += L;
    if (w >= L) w = 0.0;
    return w == 0.0 ? 0.0 : w;
}

double or_minimum_image(double d, double L) {
    if (!isfinite(d) || !isfinite(L) || L <= 0.0) return NAN;
    const double half = 0.5 * L;
    if (!isfinite(half) || half <= 0.0) return NAN;
    if (d >= -half && d < half) return d;
    double w = remainder(d, L);
    if (!isfinite(w)) return NAN;
    if (w >= half) w -= L;
    else if (w < -half) w += L;
    return w == 0.0 ? 0.0 : w;
}

/* ------------------------------------------------------------------------- */
/* CIC coordinate: src/mesh/cic_coordinate.hpp:95-121                         */
/* ------------------------------------------------------------------------- */
int or_cic_coordinate(double position, double L, double dx, uint64_t N,
                      int64_t* base, double* fraction) {
    const double upper = nextafter((double)N, 0.0);
    const double w = or_wrap(position, L);
    if (!isfinite(w) || w < 0.0 || !(w < L)) return OR_ELOGIC;
    const double unbounded = w / dx;
    if (!isfinite(unbounded)) return OR_ELOGIC;
    double s = unbounded;
    if (s < 0.0) s = 0.0;
    if (s > upper) s = upper;
    const int64_t b = (int64_t)floor(s);
    if (b < 0 || (uint64_t)b >= N) return OR_ELOGIC;
    const double f = s - (double)b;
    if (!(f >= 0.0 && f < 1.0)) return OR_ELOGIC;
    *base = b;
    *fraction = f;
    return OR_OK;
}

static inline uint64_t wrap_index(int64_t i, uint64_t N) {
    int64_t w = i % (int64_t)N;
    if (w < 0) w += (int64_t)N;
    return (uint64_t)w;
}

/* ------------------------------------------------------------------------- */
/* Stable x-plane counting sort: mass_assignment_workspace.inc:264-374         */
/* ------------------------------------------------------------------------- */
int or_xplane_order(uint64_t n, const double* x, double L, uint64_t N,
                    uint64_t* offsets, uint32_t* order) {
    const double dx = L / (double)N;
    uint32_t* bin = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    if (!bin) return OR_ERUNTIME;
    int err = OR_OK;
    for (uint64_t i = 0; i <= N; ++i) offsets[i] = 0;
    for (uint64_t p = 0; p < n; ++p) {
        int64_t b; double f;
        if (or_cic_coordinate(x[p], L, dx, N, &b, &f) != OR_OK) { err = OR_ELOGIC; break; }
        bin[p] = (uint32_t)b;
        offsets[b + 1]++;
    }
    if (err == OR_OK) {
        for (uint64_t i = 0; i < N; ++i) offsets[i + 1] += offsets[i];
        uint64_t* cursor = (uint64_t*)malloc(N * sizeof(uint64_t));
        if (!cursor) { free(bin); return OR_ERUNTIME; }
        memcpy(cursor, offsets, N * sizeof(uint64_t));
        for (uint64_t p = 0; p < n; ++p) order[cursor[bin[p]]++] = (uint32_t)p;
        free(cursor);
    }
    free(bin);
    return err;
}

/* ------------------------------------------------------------------------- */
/* Deposit bucket key (HIP deposit contract; no reference counterpart)        */
/* ------------------------------------------------------------------------- */
uint64_t or_bucket_tile_edge(uint64_t N) {
    for (uint64_t t = 16; t >= 1; --t)
        if (N % t == 0) return t;
    return 1;
}

int or_bucket_order(uint64_t n, const double* x, const double* y, const double* z,
                    double L, uint64_t N, uint64_t T,
                    uint64_t* offsets, uint32_t* order) {
    if (T == 0 || N % T != 0) return OR_EINVAL;
    const uint64_t nT = N / T;
    const uint64_t nb = nT * nT * nT * 8;
    const double dx = L / (double)N;
    uint64_t* key = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    if (!key) return OR_ERUNTIME;
    for (uint64_t i = 0; i <= nb; ++i) offsets[i] = 0;
    for (uint64_t p = 0; p < n; ++p) {
        int64_t bx, by, bz; double f;
        if (or_cic_coordinate(x[p], L, dx, N, &bx, &f) != OR_OK
            || or_cic_coordinate(y[p], L, dx, N, &by, &f) != OR_OK
            || or_cic_coordinate(z[p], L, dx, N, &bz, &f) != OR_OK) {
            free(key);
            return OR_ELOGIC;
        }
        const uint64_t tx = (uint64_t)bx / T, ty = (uint64_t)by / T, tz = (uint64_t)bz / T;
        const uint64_t cls = ((uint64_t)((uint64_t)bx % T == T - 1) << 2)
                           | ((uint64_t)((uint64_t)by % T == T - 1) << 1)
                           | (uint64_t)((uint64_t)bz % T == T - 1);
        key[p] = ((tx * nT + ty) * nT + tz) * 8 + cls;
        offsets[key[p] + 1]++;
    }
    for (uint64_t i = 0; i < nb; ++i) offsets[i + 1] += offsets[i];
    uint64_t* cursor = (uint64_t*)malloc(nb * sizeof(uint64_t));
    if (!cursor) { free(key); return OR_ERUNTIME; }
    memcpy(cursor, offsets, nb * sizeof(uint64_t));
    for (uint64_t p = 0; p < n; ++p) order[cursor[key[p]]++] = (uint32_t)p;
    free(cursor);
    free(key);
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* Deposit: src/mesh/mass_assignment_deposit.inc:15-334                        */
/* Plane task for destination plane X visits bin X with x-weight (1-f), then   */
/* bin X-1 with f; each particle adds ((m*wx)*wy)*wz for oy outer, oz inner.   */
/* ------------------------------------------------------------------------- */
int or_deposit(uint64_t n, const double* x, const double* y, const double* z,
               const double* masses, double uniform_mass, double L, uint64_t N,
               double* mesh) {
    const double dx = L / (double)N;
    for (uint64_t p = 0; p < n; ++p)
        if (!isfinite(x[p]) || !isfinite(y[p]) || !isfinite(z[p])) return OR_EINVAL;
    if (masses) {
        for (uint64_t p = 0; p < n; ++p)
            if (!isfinite(masses[p]) || masses[p] < 0.0) return OR_EINVAL;
    } else if (!isfinite(uniform_mass) || uniform_mass < 0.0) {
        return OR_EINVAL;
    }
    const uint64_t N2 = N * N;
    memset(mesh, 0, N2 * N * sizeof(double));
    if (n == 0) return OR_OK;

    uint64_t* offsets = (uint64_t*)malloc((N + 1) * sizeof(uint64_t));
    uint32_t* order = (uint32_t*)malloc(n * sizeof(uint32_t));
    if (!offsets || !order) { free(offsets); free(order); return OR_ERUNTIME; }
    double t0 = wall_s();
    int err = or_xplane_order(n, x, L, N, offsets, order);
    if (err != OR_OK) { free(offsets); free(order); return err; }
    g_stage_s[OR_STAGE_SORT] += wall_s() - t0;
    t0 = wall_s();

    int failed = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(OR_NT) reduction(|: failed)
    for (int64_t X = 0; X < (int64_t)N; ++X) {
        double* plane = mesh + (uint64_t)X * N2;
        for (int pass = 0; pass < 2; ++pass) {
            const uint64_t bin = pass == 0 ? (uint64_t)X : wrap_index(X - 1, N);
            for (uint64_t c = offsets[bin]; c < offsets[bin + 1]; ++c) {
                const uint32_t p = order[c];
                int64_t ix, iy, iz; double tx, ty, tz;
                if (or_cic_coordinate(x[p], L, dx, N, &ix, &tx) != OR_OK
                    || or_cic_coordinate(y[p], L, dx, N, &iy, &ty) != OR_OK
                    || or_cic_coordinate(z[p], L, dx, N, &iz, &tz) != OR_OK) {
                    failed = 1;
                    continue;
                }
                (void)ix;
                const double xw = pass == 1 ? tx : (1.0 - tx);
                const double wy[2] = {1.0 - ty, ty};
                const double wz[2] = {1.0 - tz, tz};
                const double m = masses ? masses[p] : uniform_mass;
                for (int oy = 0; oy < 2; ++oy) {
                    const uint64_t gy = wrap_index(iy + oy, N);
                    for (int oz = 0; oz < 2; ++oz) {
                        const uint64_t gz = wrap_index(iz + oz, N);
                        plane[gy * N + gz] += m * xw * wy[oy] * wz[oz];
                    }
                }
            }
        }
    }
    free(offsets);
    free(order);
    g_stage_s[OR_STAGE_DEPOSIT] += wall_s() - t0;
    return failed ? OR_ELOGIC : OR_OK;
}

/* ------------------------------------------------------------------------- */
/* Exact non-negative sum (contract of math/exact_positive_sum.hpp):          */
/* integer multiples of 2^-1074 in 32-bit digits, one final rounding.         */
/* ------------------------------------------------------------------------- */
#define OR_LIMBS 72
typedef struct { uint64_t d[OR_LIMBS]; uint64_t pending; } or_superacc;

static void sa_norm(or_superacc* s) {
    uint64_t carry = 0;
    for (int i = 0; i < OR_LIMBS; ++i) {
        const uint64_t v = s->d[i] + carry;
        s->d[i] = v & 0xffffffffu;
        carry = v >> 32;
    }
    s->pending = 0;
}

static void sa_add_wide(or_superacc* s, unsigned __int128 v, uint32_t shift) {
    uint32_t limb = shift / 32, off = shift % 32;
    /* v << off fits in 128+31 bits: split into pieces of 32 bits. */
    unsigned __int128 lo = v << off;
    const unsigned __int128 hi = off ? (v >> (128 - off)) : 0;
    for (int k = 0; k < 4 && limb + k < OR_LIMBS; ++k) {
        s->d[limb + k] += (uint64_t)(lo & 0xffffffffu);
        lo >>= 32;
    }
    if (limb + 4 < OR_LIMBS) s->d[limb + 4] += (uint64_t)hi;
    if (++s->pending >= (1u << 30)) sa_norm(s);
}

static int sa_add(or_superacc* s, double v, uint64_t reps) {
    if (!isfinite(v) || v < 0.0) return -1;
    if (v == 0.0 || reps == 0) return 0;
    uint64_t bits;
    memcpy(&bits, &v, 8);
    const uint32_t e = (uint32_t)((bits >> 52) & 0x7ff);
    const uint64_t frac = bits & ((1ull << 52) - 1);
    const uint64_t mant = e ? (frac | (1ull << 52)) : frac;
    const uint32_t shift = e ? e - 1 : 0;
    sa_add_wide(s, (unsigned __int128)mant * reps, shift);
    return 0;
}

static uint64_t sa_bits(const or_superacc* s, int64_t lo, int count) {
    /* bits [lo, lo+count) as an integer (count <= 64), lo may be negative. */
    uint64_t r = 0;
    for (int k = 0; k < count; ++k) {
        const int64_t b = lo + k;
        if (b < 0) continue;
        const uint64_t limb = (uint64_t)b / 32;
        if (limb >= OR_LIMBS) break;
        if ((s->d[limb] >> (b % 32)) & 1u) r |= 1ull << k;
    }
    return r;
}

static int sa_any_below(const or_superacc* s, int64_t bit) {
    for (int64_t b = 0; b < bit; ++b) {
        const uint64_t limb = (uint64_t)b / 32;
        if ((s->d[limb] >> (b % 32)) & 1u) return 1;
    }
    return 0;
}

static double sa_value(or_superacc* s) {
    sa_norm(s);
    int64_t top = -1;
    for (int i = OR_LIMBS - 1; i >= 0 && top < 0; --i)
        if (s->d[i]) {
            for (int b = 31; b >= 0; --b)
                if ((s->d[i] >> b) & 1u) { top = (int64_t)i * 32 + b; break; }
        }
    if (top < 0) return 0.0;
    if (top < 53) return ldexp((double)sa_bits(s, 0, 53), -1074);
    const int64_t lo = top - 52;
    uint64_t mant = sa_bits(s, lo, 53);
    const int round = (int)sa_bits(s, lo - 1, 1);
    const int sticky = sa_any_below(s, lo - 1);
    if (round && (sticky || (mant & 1u))) mant += 1;
    return ldexp((double)mant, (int)(lo - 1074));
}

double or_exact_sum(const double* values, uint64_t count) {
    or_superacc* s = (or_superacc*)calloc(1, sizeof(or_superacc));
    if (!s) return NAN;
    for (uint64_t i = 0; i < count; ++i)
        if (sa_add(s, values[i], 1) != 0) { free(s); return NAN; }
    const double v = sa_value(s);
    free(s);
    return v;
}

double or_exact_repeated(double value, uint64_t count) {
    or_superacc* s = (or_superacc*)calloc(1, sizeof(or_superacc));
    if (!s) return NAN;
    if (sa_add(s, value, count) != 0) { free(s); return NAN; }
    const double v = sa_value(s);
    free(s);
    return v;
}

/* src/math/cic_mass_conservation.cpp:25-63 */
int or_require_cic_mass_conservation(double particle_mass, double deposited_mass,
                                     uint64_t particle_count) {
    if (!isfinite(particle_mass) || particle_mass <= 0.0
        || !isfinite(deposited_mass) || deposited_mass <= 0.0) return OR_ERUNTIME;
    if (particle_count == 0) return OR_EINVAL;
    const long double eps = (long double)DBL_EPSILON;
    const long double ops = 32.0L * (long double)particle_count + 1.0L;
    const long double nu = ops * eps;
    if (!isfinite((double)nu) || nu >= 0.25L) return OR_EOVERFLOW;
    const long double gamma = nu / (1.0L - nu);
    const long double scale = particle_mass > deposited_mass
        ? (long double)particle_mass : (long double)deposited_mass;
    const long double tol = (gamma + 64.0L * eps) * scale;
    const long double diff = fabsl((long double)deposited_mass - (long double)particle_mass);
    return diff > tol ? OR_ERUNTIME : OR_OK;
}

/* ------------------------------------------------------------------------- */
/* Density source: src/gravity/pm_solver.cpp:700-759 and                      */
/* include/cosmo_nbody/math/density_normalization.hpp:32-120                  */
/* ------------------------------------------------------------------------- */
int or_density_contrast(double* mesh, uint64_t N, double dx, double total_mass) {
    const uint64_t cells = N * N * N;
    const double nc = (double)cells;
    const double inv_nc = 1.0 / nc;
    const double mean = total_mass / nc;
    const double area = dx * dx;
    const double volume = area * dx;
    const double inv_volume = 1.0 / volume;
    const int exact_mean = isfinite(mean) && mean > 0.0 && fma(mean, nc, -total_mass) == 0.0;
    if (exact_mean && isfinite(inv_volume) && inv_volume > 0.0) {
#pragma omp parallel for schedule(static) num_threads(OR_NT)
        for (int64_t i = 0; i < (int64_t)cells; ++i)
            mesh[i] = (mesh[i] - mean) * inv_volume;
        return OR_OK;
    }
    /* normalize() without an exact mean quotient: the fma-centred numerator
     * N_c*m - M times 1/(N_c*dx^3), that factor held in long double
     * (((1/N_c)/dx)/dx)/dx, the product rounded once to binary64
     * (density_normalization.hpp:136-142 with the scaled product of
     * scaled_positive_product.hpp:66-78,425-458). */
    (void)inv_nc;
    long double scale = 1.0L;
    scale /= (long double)nc;
    scale /= (long double)dx;
    scale /= (long double)dx;
    scale /= (long double)dx;
    int failed = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: failed)
    for (int64_t i = 0; i < (int64_t)cells; ++i) {
        const double m = mesh[i];
        if (!isfinite(m) || m < 0.0) { failed = 1; continue; }
        const double centred = fma(m, nc, -total_mass);
        if (!isfinite(centred)) { failed = 1; continue; }
        if (centred == 0.0) { mesh[i] = 0.0; continue; }
        const double v = (double)(scale * (long double)fabs(centred));
        if (!isfinite(v) || v == 0.0) { failed = 1; continue; }
        mesh[i] = copysign(v, centred);
    }
    return failed ? OR_EOVERFLOW : OR_OK;
}

/* ------------------------------------------------------------------------- */
/* FFT: mixed-radix (2,3,4,5) Stockham per line, naive DFT for other primes.  */
/* Twiddles exp(sign*2*pi*i*k/n) evaluated in long double.                     */
/* ------------------------------------------------------------------------- */
typedef struct { double re, im; } cplx;

static void dft_naive(uint64_t n, cplx* a, cplx* tmp, int sign) {
    for (uint64_t k = 0; k < n; ++k) {
        long double sr = 0, si = 0;
        for (uint64_t j = 0; j < n; ++j) {
            const long double ang = (long double)sign * 2.0L * 3.141592653589793238462643383279502884L
                * (long double)((k * j) % n) / (long double)n;
            const long double c = cosl(ang), s = sinl(ang);
            sr += a[j].re * c - a[j].im * s;
            si += a[j].re * s + a[j].im * c;
        }
        tmp[k].re = (double)sr;
        tmp[k].im = (double)si;
    }
    memcpy(a, tmp, n * sizeof(cplx));
}

typedef struct {
    uint64_t n;
    int nfac;
    int fac[64];
    int naive;
    cplx* tw; /* exp(sign*2 pi i k/n), k < n, for sign = -1 */
} fft_plan;

static int plan_make(fft_plan* p, uint64_t n) {
    p->n = n; p->nfac = 0; p->naive = 0;
    uint64_t m = n;
    while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
    while (m % 3 == 0) { p->fac[p->nfac++] = 3; m /= 3; }
    while (m % 5 == 0) { p->fac[p->nfac++] = 5; m /= 5; }
    if (m != 1) p->naive = 1;
    p->tw = (cplx*)malloc(n * sizeof(cplx));
    if (!p->tw) return -1;
    for (uint64_t k = 0; k < n; ++k) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        p->tw[k].re = (double)cosl(ang);
        p->tw[k].im = (double)sinl(ang);
    }
    return 0;
}

static void plan_free(fft_plan* p) { free(p->tw); p->tw = NULL; }

static inline cplx cmul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

/* Stockham autosort: stage with radix R, p = product of earlier radices.
 * in[i + r*n/R] -> twiddle by w^(r*k), k = i mod p, DFT_R -> out[(i-k)*R + k + r*p] */
static void fft_exec(const fft_plan* p, cplx* a, cplx* work, int sign) {
    const uint64_t n = p->n;
    if (n <= 1) return;
    if (p->naive) { dft_naive(n, a, work, sign); return; }
    cplx* src = a;
    cplx* dst = work;
    uint64_t prod = 1;
    for (int s = 0; s < p->nfac; ++s) {
        const int R = p->fac[s];
        const uint64_t m = n / (uint64_t)R;
        const uint64_t step = n / (prod * (uint64_t)R);
        for (uint64_t i = 0; i < m; ++i) {
            const uint64_t k = i % prod;
            cplx u[5];
            for (int r = 0; r < R; ++r) {
                u[r] = src[i + (uint64_t)r * m];
                if (r && k) {
                    cplx w = p->tw[((uint64_t)r * k * step) % n];
                    if (sign > 0) w.im = -w.im;
                    u[r] = cmul(u[r], w);
                }
            }
            cplx v[5];
            if (R == 2) {
                v[0].re = u[0].re + u[1].re; v[0].im = u[0].im + u[1].im;
                v[1].re = u[0].re - u[1].re; v[1].im = u[0].im - u[1].im;
            } else if (R == 4) {
                const cplx t0 = {u[0].re + u[2].re, u[0].im + u[2].im};
                const cplx t1 = {u[0].re - u[2].re, u[0].im - u[2].im};
                const cplx t2 = {u[1].re + u[3].re, u[1].im + u[3].im};
                cplx t3 = {u[1].re - u[3].re, u[1].im - u[3].im};
                /* multiply t3 by sign*i */
                const cplx t3r = sign < 0 ? (cplx){t3.im, -t3.re} : (cplx){-t3.im, t3.re};
                v[0].re = t0.re + t2.re; v[0].im = t0.im + t2.im;
                v[2].re = t0.re - t2.re; v[2].im = t0.im - t2.im;
                v[1].re = t1.re + t3r.re; v[1].im = t1.im + t3r.im;
                v[3].re = t1.re - t3r.re; v[3].im = t1.im - t3r.im;
            } else {
                for (int q = 0; q < R; ++q) {
                    long double sr = 0, si = 0;
                    for (int r = 0; r < R; ++r) {
                        cplx w = p->tw[((uint64_t)((q * r) % R) * (n / (uint64_t)R)) % n];
                        if (sign > 0) w.im = -w.im;
                        sr += (long double)u[r].re * w.re - (long double)u[r].im * w.im;
                        si += (long double)u[r].re * w.im + (long double)u[r].im * w.re;
                    }
                    v[q].re = (double)sr; v[q].im = (double)si;
                }
            }
            const uint64_t j = (i - k) * (uint64_t)R + k;
            for (int r = 0; r < R; ++r) dst[j + (uint64_t)r * prod] = v[r];
        }
        prod *= (uint64_t)R;
        cplx* t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(cplx));
}

int or_fft_1d(uint64_t n, double* data, int sign) {
    fft_plan p;
    if (plan_make(&p, n) != 0) return OR_ERUNTIME;
    cplx* work = (cplx*)malloc(n * sizeof(cplx));
    if (!work) { plan_free(&p); return OR_ERUNTIME; }
    fft_exec(&p, (cplx*)data, work, sign);
    free(work);
    plan_free(&p);
    return OR_OK;
}

/* Apply 1-D FFTs along one axis of an array of `count` lines with `stride`
 * between consecutive elements; line start offsets from index->offset map. */
static int fft_lines(const fft_plan* p, cplx* base, uint64_t lines,
                     uint64_t (*line_offset)(uint64_t, const uint64_t*),
                     const uint64_t* geo, uint64_t stride, int sign) {
    const uint64_t n = p->n;
    int failed = 0;
#pragma omp parallel num_threads(OR_NT) reduction(|: failed)
    {
        cplx* line = (cplx*)malloc(n * sizeof(cplx));
        cplx* work = (cplx*)malloc(n * sizeof(cplx));
        if (!line || !work) failed = 1;
#pragma omp for schedule(static)
        for (int64_t l = 0; l < (int64_t)lines; ++l) {
            if (!line || !work) continue;
            cplx* start = base + line_offset((uint64_t)l, geo);
            for (uint64_t i = 0; i < n; ++i) line[i] = start[i * stride];
            fft_exec(p, line, work, sign);
            for (uint64_t i = 0; i < n; ++i) start[i * stride] = line[i];
        }
        free(line);
        free(work);
    }
    return failed ? OR_ERUNTIME : OR_OK;
}

/* geo = {N, NH} */
static uint64_t off_y(uint64_t l, const uint64_t* g) { /* l = ix*NH + iz */
    const uint64_t ix = l / g[1], iz = l % g[1];
    return iz + g[1] * (g[0] * ix);
}
static uint64_t off_x(uint64_t l, const uint64_t* g) { /* l = iy*NH + iz */
    const uint64_t iy = l / g[1], iz = l % g[1];
    return iz + g[1] * iy;
}

int or_fft_r2c_3d(uint64_t N, const double* real, double* out) {
    const uint64_t NH = N / 2 + 1;
    fft_plan p;
    if (plan_make(&p, N) != 0) return OR_ERUNTIME;
    cplx* c = (cplx*)out;
    int failed = 0;
#pragma omp parallel num_threads(OR_NT) reduction(|: failed)
    {
        cplx* line = (cplx*)malloc(N * sizeof(cplx));
        cplx* work = (cplx*)malloc(N * sizeof(cplx));
        if (!line || !work) failed = 1;
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)(N * N); ++r) {
            if (!line || !work) continue;
            const double* src = real + (uint64_t)r * N;
            for (uint64_t i = 0; i < N; ++i) { line[i].re = src[i]; line[i].im = 0.0; }
            fft_exec(&p, line, work, -1);
            memcpy(c + (uint64_t)r * NH, line, NH * sizeof(cplx));
        }
        free(line);
        free(work);
    }
    const uint64_t geo[2] = {N, NH};
    if (!failed) failed = fft_lines(&p, c, N * NH, off_y, geo, NH, -1);
    if (!failed) failed = fft_lines(&p, c, N * NH, off_x, geo, N * NH, -1);
    plan_free(&p);
    return failed ? OR_ERUNTIME : OR_OK;
}

int or_fft_c2r_3d(uint64_t N, const double* in, double* real) {
    const uint64_t NH = N / 2 + 1;
    fft_plan p;
    if (plan_make(&p, N) != 0) return OR_ERUNTIME;
    cplx* c = (cplx*)malloc(N * N * NH * sizeof(cplx));
    if (!c) { plan_free(&p); return OR_ERUNTIME; }
    memcpy(c, in, N * N * NH * sizeof(cplx));
    const uint64_t geo[2] = {N, NH};
    int failed = fft_lines(&p, c, N * NH, off_x, geo, N * NH, +1);
    if (!failed) failed = fft_lines(&p, c, N * NH, off_y, geo, NH, +1);
    if (!failed) {
#pragma omp parallel num_threads(OR_NT) reduction(|: failed)
        {
            cplx* line = (cplx*)malloc(N * sizeof(cplx));
            cplx* work = (cplx*)malloc(N * sizeof(cplx));
            if (!line || !work) failed = 1;
#pragma omp for schedule(static)
            for (int64_t r = 0; r < (int64_t)(N * N); ++r) {
                if (!line || !work) continue;
                const cplx* h = c + (uint64_t)r * NH;
                for (uint64_t k = 0; k < NH; ++k) line[k] = h[k];
                for (uint64_t k = NH; k < N; ++k) {
                    line[k].re = h[N - k].re;
                    line[k].im = -h[N - k].im;
                }
                /* A c2r transform reads only the independent half: imaginary
                 * parts of the self-conjugate z modes are discarded. */
                line[0].im = 0.0;
                if (N % 2 == 0) line[N / 2].im = 0.0;
                fft_exec(&p, line, work, +1);
                double* dst = real + (uint64_t)r * N;
                for (uint64_t i = 0; i < N; ++i) dst[i] = line[i].re;
            }
            free(line);
            free(work);
        }
    }
    free(c);
    plan_free(&p);
    if (failed) return OR_ERUNTIME;
    const double norm = 1.0 / ((double)N * (double)N * (double)N);
    const uint64_t cells = N * N * N;
#pragma omp parallel for schedule(static) num_threads(OR_NT)
    for (int64_t i = 0; i < (int64_t)cells; ++i) real[i] *= norm;
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* Green function: src/mesh/green_function.cpp:15-241                          */
/* ------------------------------------------------------------------------- */
static int64_t mode_number(uint64_t i, uint64_t N) {
    return (int64_t)i <= (int64_t)(N / 2) ? (int64_t)i : (int64_t)i - (int64_t)N;
}

static double window_1d(uint64_t i, uint64_t N) {
    const long double arg = 3.141592653589793238462643383279502884L
        * (long double)mode_number(i, N) / (long double)N;
    if (arg == 0.0L) return 1.0;
    long double s = sinl(arg) / arg;
    s = fabsl(s);
    if (s > 1.0L) s = 1.0L;
    return (double)(s * s);
}

typedef struct { double* lap; double* win; double G; } green_tables;

static int green_tables_make(green_tables* t, uint64_t N, double L) {
    t->lap = (double*)malloc(N * sizeof(double));
    t->win = (double*)malloc(N * sizeof(double));
    if (!t->lap || !t->win) { free(t->lap); free(t->win); return OR_ERUNTIME; }
    const double n = (double)N;
    const double dx = L / (double)N;
    const double inv = 1.0 / dx;
    const double inv2 = inv * inv;
    for (uint64_t i = 0; i < N; ++i) {
        const double h = OR_PI * (double)mode_number(i, N) / n;
        const double s = sin(h);
        t->lap[i] = -4.0 * inv2 * s * s;
        t->win[i] = window_1d(i, N);
    }
    t->G = or_gravitational_constant();
    return OR_OK;
}

static void green_tables_free(green_tables* t) { free(t->lap); free(t->win); }

static double green_eval(const green_tables* t, int deconvolve,
                         uint64_t ix, uint64_t iy, uint64_t iz) {
    if (ix == 0 && iy == 0 && iz == 0) return 0.0;
    const double D = t->lap[ix] + t->lap[iy] + t->lap[iz];
    double f = 4.0 * OR_PI * t->G / D;
    if (deconvolve) {
        const double w = t->win[ix] * t->win[iy] * t->win[iz];
        f /= (w * w);
    }
    return f;
}

double or_green_kernel(uint64_t N, double L, int deconvolve_cic,
                       uint64_t ix, uint64_t iy, uint64_t iz) {
    green_tables t;
    if (green_tables_make(&t, N, L) != OR_OK) return NAN;
    const double v = green_eval(&t, deconvolve_cic, ix, iy, iz);
    green_tables_free(&t);
    return v;
}

int or_green_apply(uint64_t N, double L, int deconvolve_cic, double* out) {
    green_tables t;
    if (green_tables_make(&t, N, L) != OR_OK) return OR_ERUNTIME;
    const uint64_t NH = N / 2 + 1;
    cplx* c = (cplx*)out;
    int failed = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: failed)
    for (int64_t ix = 0; ix < (int64_t)N; ++ix)
        for (uint64_t iy = 0; iy < N; ++iy)
            for (uint64_t iz = 0; iz < NH; ++iz) {
                const double f = green_eval(&t, deconvolve_cic, (uint64_t)ix, iy, iz);
                if (!isfinite(f)) failed = 1;
                cplx* v = &c[iz + NH * (iy + N * (uint64_t)ix)];
                v->re *= f;
                v->im *= f;
            }
    green_tables_free(&t);
    return failed ? OR_EOVERFLOW : OR_OK;
}

/* ------------------------------------------------------------------------- */
/* Centered-gradient gather: src/mesh/mass_assignment_gradient.cpp:46-226      */
/* ------------------------------------------------------------------------- */
int or_gather_centered_gradient_add(uint64_t n, const double* phi,
                                    const double* x, const double* y, const double* z,
                                    double L, uint64_t N,
                                    double* ax, double* ay, double* az) {
    const double dx = L / (double)N;
    const double two_dx = 2.0 * dx;
    for (uint64_t p = 0; p < n; ++p) {
        if (!isfinite(x[p]) || !isfinite(y[p]) || !isfinite(z[p])) return OR_EINVAL;
        if (!isfinite(ax[p]) || !isfinite(ay[p]) || !isfinite(az[p])) return OR_EINVAL;
    }
    int failed = 0, lost = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: failed, lost)
    for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
        const uint64_t p = (uint64_t)pi;
        int64_t ix, iy, iz; double fx, fy, fz;
        if (or_cic_coordinate(x[p], L, dx, N, &ix, &fx) != OR_OK
            || or_cic_coordinate(y[p], L, dx, N, &iy, &fy) != OR_OK
            || or_cic_coordinate(z[p], L, dx, N, &iz, &fz) != OR_OK) {
            failed = 1;
            continue;
        }
        const double wx[2] = {1.0 - fx, fx};
        const double wy[2] = {1.0 - fy, fy};
        const double wz[2] = {1.0 - fz, fz};
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int ox = 0; ox < 2; ++ox) {
            const uint64_t cx = wrap_index(ix + ox, N);
            const uint64_t cxp = wrap_index((int64_t)cx + 1, N), cxm = wrap_index((int64_t)cx - 1, N);
            for (int oy = 0; oy < 2; ++oy) {
                const uint64_t cy = wrap_index(iy + oy, N);
                const uint64_t cyp = wrap_index((int64_t)cy + 1, N), cym = wrap_index((int64_t)cy - 1, N);
                for (int oz = 0; oz < 2; ++oz) {
                    const uint64_t cz = wrap_index(iz + oz, N);
                    const uint64_t czp = wrap_index((int64_t)cz + 1, N), czm = wrap_index((int64_t)cz - 1, N);
                    const double w = wx[ox] * wy[oy] * wz[oz];
                    /* cic_weight_3d (src/mesh/cic_coordinate.hpp:26-38) */
                    const int w_lost = wx[ox] != 0.0 && wy[oy] != 0.0 && wz[oz] != 0.0 && w == 0.0;
                    const double dxv = phi[(cxp * N + cy) * N + cz] - phi[(cxm * N + cy) * N + cz];
                    const double dyv = phi[(cx * N + cyp) * N + cz] - phi[(cx * N + cym) * N + cz];
                    const double dzv = phi[(cx * N + cy) * N + czp] - phi[(cx * N + cy) * N + czm];
                    const double g_x = -dxv / two_dx;
                    const double g_y = -dyv / two_dx;
                    const double g_z = -dzv / two_dx;
                    /* lost non-zero products: mass_assignment_gradient.cpp:157-196 */
                    if ((w != 0.0 || w_lost)
                        && ((dxv != 0.0 && g_x == 0.0) || (dyv != 0.0 && g_y == 0.0) || (dzv != 0.0 && g_z == 0.0)))
                        lost = 1;
                    if (w_lost && (g_x != 0.0 || g_y != 0.0 || g_z != 0.0)) lost = 1;
                    const double c_x = w * g_x, c_y = w * g_y, c_z = w * g_z;
                    if (w != 0.0
                        && ((g_x != 0.0 && c_x == 0.0) || (g_y != 0.0 && c_y == 0.0) || (g_z != 0.0 && c_z == 0.0)))
                        lost = 1;
                    gx += c_x;
                    gy += c_y;
                    gz += c_z;
                }
            }
        }
        ax[p] += gx;
        ay[p] += gy;
        az[p] += gz;
        if (!isfinite(ax[p]) || !isfinite(ay[p]) || !isfinite(az[p])) failed |= 2;
    }
    if (failed & 1) return OR_ELOGIC;
    if (lost) return OR_EUNDERFLOW;
    if (failed & 2) return OR_EOVERFLOW;
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* PM force refresh: src/gravity/pm_solver.cpp:569-852                         */
/* ------------------------------------------------------------------------- */
int or_pm_force(uint64_t n, const double* x, const double* y, const double* z,
                const double* masses, double uniform_mass,
                double L, uint64_t N, int deconvolve_cic,
                double* ax, double* ay, double* az, double* phi_out) {
    const uint64_t cells = N * N * N;
    const uint64_t NH = N / 2 + 1;
    if (!masses && (!isfinite(uniform_mass) || uniform_mass <= 0.0)) return OR_EINVAL;
    double* mesh = phi_out ? phi_out : (double*)malloc(cells * sizeof(double));
    double* spec = (double*)malloc(N * N * NH * 2 * sizeof(double));
    if (!mesh || !spec) {
        if (!phi_out) free(mesh);
        free(spec);
        return OR_ERUNTIME;
    }
    int err = or_deposit(n, x, y, z, masses, uniform_mass, L, N, mesh);
    double input_mass = 0.0;
    double t0 = wall_s();
    if (err == OR_OK) {
        input_mass = masses ? or_exact_sum(masses, n) : or_exact_repeated(uniform_mass, n);
        if (!isfinite(input_mass) || input_mass <= 0.0 || n == 0) err = OR_ERUNTIME;
    }
    if (err == OR_OK) {
        const double deposited = or_exact_sum(mesh, cells);
        if (!isfinite(deposited) || deposited <= 0.0) err = OR_ERUNTIME;
        else err = or_require_cic_mass_conservation(input_mass, deposited, n);
    }
    if (err == OR_OK) err = or_density_contrast(mesh, N, L / (double)N, input_mass);
    double t1 = wall_s();
    g_stage_s[OR_STAGE_MASS_CONTRAST] += t1 - t0;
    if (err == OR_OK) err = or_fft_r2c_3d(N, mesh, spec);
    t0 = wall_s();
    g_stage_s[OR_STAGE_FFT_FORWARD] += t0 - t1;
    if (err == OR_OK) err = or_green_apply(N, L, deconvolve_cic, spec);
    t1 = wall_s();
    g_stage_s[OR_STAGE_GREEN] += t1 - t0;
    if (err == OR_OK) err = or_fft_c2r_3d(N, spec, mesh);
    t0 = wall_s();
    g_stage_s[OR_STAGE_FFT_INVERSE] += t0 - t1;
    if (err == OR_OK) err = or_gather_centered_gradient_add(n, mesh, x, y, z, L, N, ax, ay, az);
    g_stage_s[OR_STAGE_GATHER] += wall_s() - t0;
    free(spec);
    if (!phi_out) free(mesh);
    return err;
}

/* ------------------------------------------------------------------------- */
/* TreePM long-range PM: continuum Green with the Gaussian split, spectral     */
/* gradient, CIC interpolation (src/mesh/green_function.cpp:165-203,           */
/* src/gravity/pm_solver.cpp:768-810, src/mesh/mass_assignment_gather.inc:     */
/* 95-181,402-451).                                                            */
/* ------------------------------------------------------------------------- */
static double k_component(uint64_t i, uint64_t N, double L) {
    return 2.0 * OR_PI * (double)mode_number(i, N) / L;
}

int or_green_apply_treepm(uint64_t N, double L, double split_scale, double* out) {
    if (!isfinite(split_scale) || split_scale <= 0.0) return OR_EINVAL;
    green_tables t;
    if (green_tables_make(&t, N, L) != OR_OK) return OR_ERUNTIME;
    const uint64_t NH = N / 2 + 1;
    cplx* c = (cplx*)out;
    int failed = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: failed)
    for (int64_t ix = 0; ix < (int64_t)N; ++ix)
        for (uint64_t iy = 0; iy < N; ++iy)
            for (uint64_t iz = 0; iz < NH; ++iz) {
                double f = 0.0;
                if (!(ix == 0 && iy == 0 && iz == 0)) {
                    const double kx = k_component((uint64_t)ix, N, L), ky = k_component(iy, N, L),
                                 kz = k_component(iz, N, L);
                    const double k2 = kx * kx + ky * ky + kz * kz;
                    f = 4.0 * OR_PI * t.G / (-k2);
                    f *= exp(-k2 * split_scale * split_scale);
                    const double w = t.win[ix] * t.win[iy] * t.win[iz];
                    f /= (w * w);
                }
                if (!isfinite(f)) failed = 1;
                cplx* v = &c[iz + NH * (iy + N * (uint64_t)ix)];
                v->re *= f;
                v->im *= f;
            }
    green_tables_free(&t);
    return failed ? OR_EOVERFLOW : OR_OK;
}

/* out += CIC interpolation of field: sum over (x, y, z) corner offsets of
 * field * ((wx * wy) * wz) in that order.  Returns OR_EUNDERFLOW when a
 * non-zero weight or field-weight product rounds to zero. */
int or_interpolate_add(uint64_t n, const double* field, const double* x, const double* y,
                       const double* z, double L, uint64_t N, double* out) {
    const double dx = L / (double)N;
    for (uint64_t p = 0; p < n; ++p)
        if (!isfinite(x[p]) || !isfinite(y[p]) || !isfinite(z[p])) return OR_EINVAL;
    int lost = 0, bad = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: lost, bad)
    for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
        const uint64_t p = (uint64_t)pi;
        int64_t ix, iy, iz;
        double fx, fy, fz;
        if (or_cic_coordinate(x[p], L, dx, N, &ix, &fx) != OR_OK
            || or_cic_coordinate(y[p], L, dx, N, &iy, &fy) != OR_OK
            || or_cic_coordinate(z[p], L, dx, N, &iz, &fz) != OR_OK) {
            bad = 1;
            continue;
        }
        const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
        double v = 0.0;
        for (int ox = 0; ox < 2; ++ox) {
            const uint64_t gx = wrap_index(ix + ox, N);
            for (int oy = 0; oy < 2; ++oy) {
                const uint64_t gy = wrap_index(iy + oy, N);
                for (int oz = 0; oz < 2; ++oz) {
                    const uint64_t gz = wrap_index(iz + oz, N);
                    const double xy = wx[ox] * wy[oy];
                    const double w = xy * wz[oz];
                    const double f = field[(gx * N + gy) * N + gz];
                    if (wx[ox] != 0.0 && wy[oy] != 0.0 && wz[oz] != 0.0 && w == 0.0 && f != 0.0) lost = 1;
                    const double c = f * w;
                    if (f != 0.0 && w != 0.0 && c == 0.0) lost = 1;
                    v += c;
                }
            }
        }
        out[p] += v;
    }
    if (bad) return OR_ELOGIC;
    if (lost) return OR_EUNDERFLOW;
    return OR_OK;
}

/* Potential-energy diagnostic sum_p m_p phi(x_p) (core::Accum is binary64):
 * CICMassAssignment::interpolate_mass_weighted_sum_full_mesh
 * (src/mesh/mass_assignment_gather.inc:283-399) over the fixed-block Neumaier sum
 * of include/cosmo_nbody/math/deterministic_sum.hpp:48-104 (blocks of 4096);
 * masses NULL selects the uniform mass.  *out = 0.5 * sum (pm_solver.cpp:817-841). */
static void neumaier(double v, double* sum, double* comp) {
    const double t = *sum + v;
    if (fabs(*sum) >= fabs(v)) *comp += (*sum - t) + v;
    else *comp += (v - t) + *sum;
    *sum = t;
}

int or_potential_energy(uint64_t n, const double* field, const double* x, const double* y, const double* z,
                        const double* masses, double uniform_mass, double L, uint64_t N, double* out) {
    const double dx = L / (double)N;
    for (uint64_t p = 0; p < n; ++p)
        if (!isfinite(x[p]) || !isfinite(y[p]) || !isfinite(z[p])) return OR_EINVAL;
    const uint64_t block = 4096;
    const uint64_t nb = n ? 1 + (n - 1) / block : 0;
    double* part = (double*)calloc(nb ? nb : 1, sizeof(double));
    if (!part) return OR_ERUNTIME;
    int lost = 0, bad = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: lost, bad)
    for (int64_t bi = 0; bi < (int64_t)nb; ++bi) {
        const uint64_t b0 = (uint64_t)bi * block, b1 = b0 + block < n ? b0 + block : n;
        double sum = 0.0, comp = 0.0;
        for (uint64_t p = b0; p < b1; ++p) {
            int64_t ix, iy, iz;
            double fx, fy, fz;
            if (or_cic_coordinate(x[p], L, dx, N, &ix, &fx) != OR_OK
                || or_cic_coordinate(y[p], L, dx, N, &iy, &fy) != OR_OK
                || or_cic_coordinate(z[p], L, dx, N, &iz, &fz) != OR_OK) {
                bad = 1;
                continue;
            }
            const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
            const double m = masses ? masses[p] : uniform_mass;
            double v = 0.0;
            int lost_p = 0;
            for (int ox = 0; ox < 2; ++ox) {
                const uint64_t gx = wrap_index(ix + ox, N);
                for (int oy = 0; oy < 2; ++oy) {
                    const uint64_t gy = wrap_index(iy + oy, N);
                    for (int oz = 0; oz < 2; ++oz) {
                        const uint64_t gz = wrap_index(iz + oz, N);
                        const double f = field[(gx * N + gy) * N + gz];
                        const double xy = wx[ox] * wy[oy];
                        const double w = xy * wz[oz];
                        if (wx[ox] != 0.0 && wy[oy] != 0.0 && wz[oz] != 0.0 && w == 0.0 && f != 0.0 && m != 0.0)
                            lost_p = 1;
                        const double c = f * w;
                        if (f != 0.0 && w != 0.0 && c == 0.0 && m != 0.0) lost_p = 1;
                        v += c;
                    }
                }
            }
            const double weighted = m * v;
            if (m != 0.0 && v != 0.0 && weighted == 0.0) lost_p = 1;
            if (lost_p) {
                lost = 1;
                continue;
            }
            neumaier(weighted, &sum, &comp);
        }
        part[bi] = sum + comp;
    }
    double total = 0.0, comp = 0.0;
    for (uint64_t b = 0; b < nb; ++b) neumaier(part[b], &total, &comp);
    free(part);
    if (bad) return OR_ELOGIC;
    if (lost) return OR_EUNDERFLOW;
    const double sum = total + comp;
    if (!isfinite(sum)) return OR_EUNDERFLOW;
    const double half = 0.5 * sum;
    if (sum != 0.0 && half == 0.0) return OR_EUNDERFLOW;
    *out = half;
    return OR_OK;
}

/* The long-range PM force of TreePM (PMForceMethod::treepm_long_range):
 * deposit, density source, r2c, continuum Green x exp(-k^2 r_s^2) / W^2, then
 * per axis -i k_axis phi (k_axis = 0 on the Nyquist index), c2r with 1/N^3 and
 * CIC interpolation added into a_axis.  mesh_out (optional, 3 N^3) receives
 * the three force meshes. */
int or_pm_force_treepm(uint64_t n, const double* x, const double* y, const double* z,
                       const double* masses, double uniform_mass, double L, uint64_t N,
                       double split_scale, double* ax, double* ay, double* az, double* mesh_out) {
    const uint64_t cells = N * N * N, NH = N / 2 + 1, cm = N * N * NH;
    if (!masses && (!isfinite(uniform_mass) || uniform_mass <= 0.0)) return OR_EINVAL;
    double* mesh = (double*)malloc(cells * sizeof(double));
    double* spec = (double*)malloc(cm * 2 * sizeof(double));
    double* tmp = (double*)malloc(cm * 2 * sizeof(double));
    if (!mesh || !spec || !tmp) { free(mesh); free(spec); free(tmp); return OR_ERUNTIME; }
    int err = or_deposit(n, x, y, z, masses, uniform_mass, L, N, mesh);
    double input_mass = 0.0;
    if (err == OR_OK) {
        input_mass = masses ? or_exact_sum(masses, n) : or_exact_repeated(uniform_mass, n);
        if (!isfinite(input_mass) || input_mass <= 0.0 || n == 0) err = OR_ERUNTIME;
    }
    if (err == OR_OK) {
        const double deposited = or_exact_sum(mesh, cells);
        if (!isfinite(deposited) || deposited <= 0.0) err = OR_ERUNTIME;
        else err = or_require_cic_mass_conservation(input_mass, deposited, n);
    }
    if (err == OR_OK) err = or_density_contrast(mesh, N, L / (double)N, input_mass);
    if (err == OR_OK) err = or_fft_r2c_3d(N, mesh, spec);
    if (err == OR_OK) err = or_green_apply_treepm(N, L, split_scale, spec);
    double* acc[3] = {ax, ay, az};
    for (int axis = 0; axis < 3 && err == OR_OK; ++axis) {
        for (uint64_t ix = 0; ix < N; ++ix) {
            double kx = k_component(ix, N, L);
            if (N % 2 == 0 && ix == N / 2) kx = 0.0;
            for (uint64_t iy = 0; iy < N; ++iy) {
                double ky = k_component(iy, N, L);
                if (N % 2 == 0 && iy == N / 2) ky = 0.0;
                for (uint64_t iz = 0; iz < NH; ++iz) {
                    double kz = k_component(iz, N, L);
                    if (N % 2 == 0 && iz == N / 2) kz = 0.0;
                    const double k = axis == 0 ? kx : (axis == 1 ? ky : kz);
                    const uint64_t i = 2 * (iz + NH * (iy + N * ix));
                    tmp[i] = spec[i + 1] * k;
                    tmp[i + 1] = -spec[i] * k;
                }
            }
        }
        err = or_fft_c2r_3d(N, tmp, mesh);
        if (err == OR_OK && mesh_out) memcpy(mesh_out + (uint64_t)axis * cells, mesh, cells * sizeof(double));
        if (err == OR_OK) err = or_interpolate_add(n, mesh, x, y, z, L, N, acc[axis]);
    }
    free(mesh);
    free(spec);
    free(tmp);
    return err;
}

/* ------------------------------------------------------------------------- */
/* KDK particle update: include/cosmo_nbody/time/particle_update.hpp:20-116    */
/* ------------------------------------------------------------------------- */
static double add_periodic_displacement(double pos, double disp, double L) {
    if (!isfinite(pos) || !isfinite(disp) || !isfinite(L) || L <= 0.0) return NAN;
    const double wp = or_wrap(pos, L);
    const double rd = or_minimum_image(disp, L);
    if (!isfinite(wp) || !isfinite(rd)) return NAN;
    return or_wrap(wp + rd, L);
}

int or_kick_drift(uint64_t n, double* x, double* y, double* z,
                  double* px, double* py, double* pz,
                  const double* ax, const double* ay, const double* az,
                  double K, double D, double L) {
    /* leapfrog_integrator.cpp:170-225: preflight every particle, then publish. */
    int invalid = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: invalid)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (!isfinite(x[i]) || !isfinite(y[i]) || !isfinite(z[i]) || !isfinite(D)) { invalid = 1; continue; }
        if (!isfinite(px[i]) || !isfinite(py[i]) || !isfinite(pz[i])
            || !isfinite(ax[i]) || !isfinite(ay[i]) || !isfinite(az[i]) || !isfinite(K)) { invalid = 1; continue; }
        const double qx = px[i] + ax[i] * K, qy = py[i] + ay[i] * K, qz = pz[i] + az[i] * K;
        if (!isfinite(qx) || !isfinite(qy) || !isfinite(qz)) { invalid = 1; continue; }
        const double dx = qx * D, dy = qy * D, dz = qz * D;
        if (!isfinite(dx) || !isfinite(dy) || !isfinite(dz)) { invalid = 1; continue; }
        if (!isfinite(add_periodic_displacement(x[i], dx, L))
            || !isfinite(add_periodic_displacement(y[i], dy, L))
            || !isfinite(add_periodic_displacement(z[i], dz, L))) invalid = 1;
    }
    if (invalid) return OR_EOVERFLOW;
#pragma omp parallel for schedule(static) num_threads(OR_NT)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const double qx = px[i] + ax[i] * K, qy = py[i] + ay[i] * K, qz = pz[i] + az[i] * K;
        px[i] = qx; py[i] = qy; pz[i] = qz;
        x[i] = add_periodic_displacement(x[i], qx * D, L);
        y[i] = add_periodic_displacement(y[i], qy * D, L);
        z[i] = add_periodic_displacement(z[i], qz * D, L);
    }
    return OR_OK;
}

int or_kick(uint64_t n, double* px, double* py, double* pz,
            const double* ax, const double* ay, const double* az, double K) {
    int invalid = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: invalid)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (!isfinite(px[i]) || !isfinite(py[i]) || !isfinite(pz[i])
            || !isfinite(ax[i]) || !isfinite(ay[i]) || !isfinite(az[i]) || !isfinite(K)) { invalid = 1; continue; }
        const double qx = px[i] + ax[i] * K, qy = py[i] + ay[i] * K, qz = pz[i] + az[i] * K;
        if (!isfinite(qx) || !isfinite(qy) || !isfinite(qz)) invalid = 1;
    }
    if (invalid) return OR_EOVERFLOW;
#pragma omp parallel for schedule(static) num_threads(OR_NT)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        px[i] = px[i] + ax[i] * K;
        py[i] = py[i] + ay[i] * K;
        pz[i] = pz[i] + az[i] * K;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* Background expansion: src/cosmology/cosmology_model.cpp:25-57,167-178       */
/* ------------------------------------------------------------------------- */
double or_hubble(double a, double omega_m, double omega_lambda) {
    if (!isfinite(a) || a <= 0.0) return NAN;
    const double sqrt_a = sqrt(a);
    const double sqrt_m = sqrt(omega_m);
    const double sqrt_l = sqrt(omega_lambda);
    double matter;
    if (a <= 1.0) {
        long double m = (long double)sqrt_m;
        m = m / (long double)a;
        m = m / (long double)sqrt_a;
        matter = (double)m;
    } else {
        const double inv = 1.0 / sqrt_a;
        matter = sqrt_m * (inv / a);
    }
    const double e = hypot(matter, sqrt_l);
    return 100.0 * e;
}

/* Gauss-Legendre nodes/weights on [-1,1] (positive half), computed by Newton
 * iteration on P_n in long double and rounded to binary64. */
static int gl_ready = 0;
static double gl16_x[8], gl16_w[8], gl8_x[4], gl8_w[4];

static void gl_nodes(int n, double* xs, double* ws) {
    const long double pi = 3.141592653589793238462643383279502884L;
    const int half = n / 2;
    for (int i = 0; i < half; ++i) {
        /* root index counted from the largest; store ascending */
        long double x = cosl(pi * ((long double)(i + 1) - 0.25L) / ((long double)n + 0.5L));
        long double dp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = x;
            for (int k = 2; k <= n; ++k) {
                const long double pk = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1; p1 = pk;
            }
            dp = (long double)n * (x * p1 - p0) / (x * x - 1.0L);
            const long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-30L) break;
        }
        {
            long double p0 = 1.0L, p1 = x;
            for (int k = 2; k <= n; ++k) {
                const long double pk = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1; p1 = pk;
            }
            dp = (long double)n * (x * p1 - p0) / (x * x - 1.0L);
        }
        const long double w = 2.0L / ((1.0L - x * x) * dp * dp);
        xs[half - 1 - i] = (double)x;
        ws[half - 1 - i] = (double)w;
    }
}

static void gl_init(void) {
    if (gl_ready) return;
    gl_nodes(16, gl16_x, gl16_w);
    gl_nodes(8, gl8_x, gl8_w);
    gl_ready = 1;
}

static void compensated_add(long double term, long double* sum, long double* comp) {
    const long double upd = *sum + term;
    if (fabsl(*sum) >= fabsl(term)) *comp += (*sum - upd) + term;
    else *comp += (term - upd) + *sum;
    *sum = upd;
}

typedef double (*or_integrand)(double u, const double* ctx);

static int gl_composite(or_integrand f, const double* ctx, double lower, double upper,
                        uint64_t cells, long double* e8, long double* e16) {
    const double width = upper - lower;
    const double cw = width / (double)cells;
    const double hw = 0.5 * cw;
    if (!(width > 0.0) || !(cw > 0.0) || !(hw > 0.0)) return OR_EUNDERFLOW;
    long double t8 = 0, c8 = 0, t16 = 0, c16 = 0;
    for (uint64_t c = 0; c < cells; ++c) {
        const double centre = lower + ((double)c + 0.5) * cw;
        long double l8 = 0, l8c = 0, l16 = 0, l16c = 0;
        for (int i = 0; i < 4; ++i) {
            const double off = hw * gl8_x[i];
            const double a = centre - off, b = centre + off;
            if (!(lower < a && a < b && b < upper)) return OR_EUNDERFLOW;
            const double fa = f(a, ctx), fb = f(b, ctx);
            if (!isfinite(fa) || !isfinite(fb)) return OR_EOVERFLOW;
            compensated_add((long double)gl8_w[i] * ((long double)fa + (long double)fb), &l8, &l8c);
        }
        for (int i = 0; i < 8; ++i) {
            const double off = hw * gl16_x[i];
            const double a = centre - off, b = centre + off;
            if (!(lower < a && a < b && b < upper)) return OR_EUNDERFLOW;
            const double fa = f(a, ctx), fb = f(b, ctx);
            if (!isfinite(fa) || !isfinite(fb)) return OR_EOVERFLOW;
            compensated_add((long double)gl16_w[i] * ((long double)fa + (long double)fb), &l16, &l16c);
        }
        compensated_add((long double)hw * (l8 + l8c), &t8, &c8);
        compensated_add((long double)hw * (l16 + l16c), &t16, &c16);
    }
    *e8 = t8 + c8;
    *e16 = t16 + c16;
    return OR_OK;
}

static long double binary64_spacing(double v) {
    const double up = nextafter(v, INFINITY), dn = nextafter(v, -INFINITY);
    long double s = 0.0L;
    if (isfinite(up)) { const long double d = fabsl((long double)up - (long double)v); if (d > s) s = d; }
    if (isfinite(dn)) { const long double d = fabsl((long double)v - (long double)dn); if (d > s) s = d; }
    return s;
}

static int gl_resolved(or_integrand f, const double* ctx, double lower, double upper, double* out) {
    gl_init();
    long double p8, p16, c8, c16;
    int err = gl_composite(f, ctx, lower, upper, 1, &p8, &p16);
    if (err) return err;
    for (uint64_t cells = 2; cells < (1ull << 40); cells *= 2) {
        err = gl_composite(f, ctx, lower, upper, cells, &c8, &c16);
        if (err) return err;
        const double rep = (double)c16;
        const long double sp = binary64_spacing(rep);
        if (fabsl(c16 - c8) <= sp && fabsl(c16 - p16) <= sp) { *out = rep; return OR_OK; }
        p16 = c16;
        (void)p8;
    }
    return OR_EOVERFLOW;
}

/* ctx = {base, omega_m, omega_lambda, kind}; kind 0 drift 1/(a*aH), 1 kick 1/(aH) */
static double dk_integrand(double u, const double* ctx) {
    const double base = ctx[0];
    const double ratio = exp(u);
    double a = base * ratio;
    if (!isfinite(ratio) || !isfinite(a) || a <= 0.0) a = exp(log(base) + u);
    const double ah = a * or_hubble(a, ctx[1], ctx[2]);
    return ctx[3] == 0.0 ? 1.0 / (a * ah) : 1.0 / ah;
}

static int dk_factor(double a1, double a2, double om, double ol, int kind, double* out) {
    if (!isfinite(a1) || !isfinite(a2) || a1 <= 0.0 || a2 <= 0.0) return OR_EINVAL;
    if (a1 == a2) { *out = 0.0; return OR_OK; }
    const int fwd = a2 > a1;
    const double base = fwd ? a1 : a2, upper = fwd ? a2 : a1;
    const double rel = (upper - base) / base;
    const double w = isfinite(rel) ? log1p(rel) : log(upper) - log(base);
    if (!isfinite(w)) return OR_EOVERFLOW;
    if (w <= 0.0) return OR_EUNDERFLOW;
    const double ctx[4] = {base, om, ol, (double)kind};
    double v;
    const int err = gl_resolved(dk_integrand, ctx, 0.0, w, &v);
    if (err) return err;
    *out = fwd ? v : -v;
    return OR_OK;
}

int or_drift_factor(double a1, double a2, double om, double ol, double* out) {
    return dk_factor(a1, a2, om, ol, 0, out);
}

int or_kick_factor(double a1, double a2, double om, double ol, double* out) {
    return dk_factor(a1, a2, om, ol, 1, out);
}

/* Correctly rounded sqrt(lhs*rhs): candidate from long double, then exact
 * comparison of squared midpoints in binary128 (products are exact there). */
double or_geometric_mean(double lhs, double rhs) {
    if (!(lhs > 0.0) || !(rhs > 0.0) || !isfinite(lhs) || !isfinite(rhs)) return NAN;
    const __float128 prod = (__float128)lhs * (__float128)rhs;
    double r = (double)sqrtl((long double)lhs * (long double)rhs);
    for (int it = 0; it < 4; ++it) {
        const double up = nextafter(r, INFINITY), dn = nextafter(r, 0.0);
        const __float128 mid_up = ((__float128)r + (__float128)up) / 2;
        const __float128 mid_dn = ((__float128)r + (__float128)dn) / 2;
        if (mid_up * mid_up < prod) { r = up; continue; }
        if (mid_dn * mid_dn > prod) { r = dn; continue; }
        if (mid_up * mid_up == prod) { /* tie: even significand */
            uint64_t b; memcpy(&b, &r, 8);
            if (b & 1u) r = up;
        } else if (mid_dn * mid_dn == prod) {
            uint64_t b; memcpy(&b, &r, 8);
            if (b & 1u) r = dn;
        }
        break;
    }
    return r;
}

/* ------------------------------------------------------------------------- */
/* Boundary table: src/time/time_stepper.cpp:16-64,194-276                     */
/* ------------------------------------------------------------------------- */
static int ts_log_ratio(double upper, double lower, double* out) {
    if (!isfinite(upper) || !isfinite(lower) || !(upper > lower) || lower <= 0.0) return OR_EINVAL;
    const double rel = (upper - lower) / lower;
    const double v = isfinite(rel) ? log1p(rel) : log(upper) - log(lower);
    if (!isfinite(v) || v <= 0.0) return OR_EOVERFLOW;
    *out = v;
    return OR_OK;
}

static int ts_count(double span, double dlna, double* out) {
    const double ratio = span / dlna;
    if (!isfinite(ratio) || ratio <= 0.0) return OR_EOVERFLOW;
    const double c = ceil(nextafter(ratio, INFINITY));
    if (!isfinite(c)) return OR_EOVERFLOW;
    *out = c < 1.0 ? 1.0 : c;
    return OR_OK;
}

static int ts_advance(double cur, double dlna, double* out) {
    const double inc = expm1(dlna);
    const double next = isfinite(inc) ? fma(cur, inc, cur) : exp(log(cur) + dlna);
    if (!isfinite(next)) return OR_EOVERFLOW;
    *out = next;
    return OR_OK;
}

static int ts_push(double* out, uint64_t cap, uint64_t* count, double next) {
    if (*count == 0) return OR_ELOGIC;
    const double prev = out[*count - 1];
    if (!(next > prev) || !isfinite(next)) return OR_EUNDERFLOW;
    const double h = or_geometric_mean(prev, next);
    if (!(h > prev) || !(h < next)) return OR_EUNDERFLOW;
    if (*count >= cap) return OR_EOVERFLOW;
    out[(*count)++] = next;
    return OR_OK;
}

int or_time_boundaries(double a_start, double a_final, double dlna,
                       const double* stops_in, uint64_t n_stops,
                       double* out, uint64_t cap, uint64_t* count) {
    if (!isfinite(a_start) || !isfinite(a_final) || !isfinite(dlna)) return OR_EINVAL;
    if (a_start <= 0.0 || a_final <= 0.0 || dlna <= 0.0 || a_start > a_final) return OR_EINVAL;
    if (cap == 0) return OR_EOVERFLOW;
    *count = 0;
    out[(*count)++] = a_start;
    if (a_start == a_final) return OR_OK;
    double* stops = (double*)malloc((n_stops + 1) * sizeof(double));
    if (!stops) return OR_ERUNTIME;
    uint64_t ns = 0;
    double prev_stop = a_start;
    for (uint64_t i = 0; i < n_stops; ++i) {
        const double s = stops_in[i];
        if (!isfinite(s) || s <= a_start || s > a_final || !(s > prev_stop)) { free(stops); return OR_EINVAL; }
        prev_stop = s;
        stops[ns++] = s;
    }
    if (ns == 0 || stops[ns - 1] != a_final) stops[ns++] = a_final;
    int err = OR_OK;
    for (uint64_t si = 0; si < ns && err == OR_OK; ++si) {
        const double stop = stops[si];
        const double anchor = out[*count - 1];
        double span, creal;
        err = ts_log_ratio(stop, anchor, &span);
        if (!err) err = ts_count(span, dlna, &creal);
        if (err) break;
        const uint64_t cnt = (uint64_t)creal;
        int appended = 0;
        for (uint64_t k = 1; k < cnt && err == OR_OK; ++k) {
            const double off = dlna * (double)k;
            double next;
            err = ts_advance(anchor, off, &next);
            if (err) break;
            if (next >= stop) { err = ts_push(out, cap, count, stop); appended = 1; break; }
            const double th = or_geometric_mean(next, stop);
            if (!(th > next && th < stop)) { err = ts_push(out, cap, count, stop); appended = 1; break; }
            err = ts_push(out, cap, count, next);
        }
        if (!err && !appended) err = ts_push(out, cap, count, stop);
    }
    free(stops);
    return err;
}

/* ------------------------------------------------------------------------- */
/* Full run: step loop of src/runtime/simulation_runner.cpp:1348-1527 over     */
/* src/time/leapfrog_integrator.cpp:143-279 (force reused across steps).       */
/* ------------------------------------------------------------------------- */
static void zero3(uint64_t n, double* a, double* b, double* c) {
#pragma omp parallel for schedule(static) num_threads(OR_NT)
    for (int64_t i = 0; i < (int64_t)n; ++i) { a[i] = 0.0; b[i] = 0.0; c[i] = 0.0; }
}

int or_run(uint64_t n, double* x, double* y, double* z,
           double* px, double* py, double* pz, double* ax, double* ay, double* az,
           const double* masses, double uniform_mass, double L, uint64_t N,
           int deconvolve_cic, double om, double ol,
           const double* b, uint64_t nb) {
    if (nb < 1) return OR_EINVAL;
    zero3(n, ax, ay, az);
    int err = or_pm_force(n, x, y, z, masses, uniform_mass, L, N, deconvolve_cic, ax, ay, az, NULL);
    for (uint64_t s = 0; s + 1 < nb && err == OR_OK; ++s) {
        const double a0 = b[s], a1 = b[s + 1];
        const double ah = or_geometric_mean(a0, a1);
        double D, K1, K2;
        err = or_drift_factor(a0, a1, om, ol, &D);
        if (!err) err = or_kick_factor(a0, ah, om, ol, &K1);
        if (!err) err = or_kick_factor(ah, a1, om, ol, &K2);
        double t0 = wall_s();
        if (!err) err = or_kick_drift(n, x, y, z, px, py, pz, ax, ay, az, K1, D, L);
        if (!err) zero3(n, ax, ay, az);
        g_stage_s[OR_STAGE_KDK] += wall_s() - t0;
        if (!err) err = or_pm_force(n, x, y, z, masses, uniform_mass, L, N, deconvolve_cic, ax, ay, az, NULL);
        t0 = wall_s();
        if (!err) err = or_kick(n, px, py, pz, ax, ay, az, K2);
        g_stage_s[OR_STAGE_KDK] += wall_s() - t0;
    }
    return err;
}


/* ------------------------------------------------------------------------ */
/* Mode-keyed Gaussian initial field (src/ic/philox_rng.cpp:1-90,
 * src/ic/random_field.cpp:70-330). */
void or_philox4x32_10(const uint32_t* counter, const uint32_t* key, uint32_t* out) {
    uint32_t c0 = counter[0], c1 = counter[1], c2 = counter[2], c3 = counter[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        if (round != 9) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static double open53(uint32_t hi, uint32_t lo) {
    const uint64_t bits = ((uint64_t)hi << 21) | ((uint64_t)lo >> 11);
    if (bits == 0) return 0x1.0p-54;
    return (double)bits * 0x1.0p-53;
}

void or_uniform_pair(uint64_t seed, int32_t kx, int32_t ky, int32_t kz, uint32_t stream, double* out) {
    const uint32_t ctr[4] = {(uint32_t)kx, (uint32_t)ky, (uint32_t)kz, stream};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    or_philox4x32_10(ctr, key, w);
    out[0] = open53(w[0], w[1]);
    out[1] = open53(w[2], w[3]);
}

void or_normal_pair(uint64_t seed, int32_t kx, int32_t ky, int32_t kz, uint32_t stream, double* out) {
    double u[2];
    or_uniform_pair(seed, kx, ky, kz, stream, u);
    const double r = sqrt(-2.0 * log(u[0]));
    const double th = 2.0 * 3.14159265358979323846 * u[1];
    out[0] = r * cos(th);
    out[1] = r * sin(th);
}

static long long or_mode(long long i, long long n) { return i <= n / 2 ? i : i - n; }

int or_gaussian_modes(uint64_t N, uint64_t seed, uint64_t max_mode, const double* var_m2, uint32_t flags,
                      double* out) {
    const long long n = (long long)N, nh = n / 2 + 1;
    const int fixed = (flags & 1u) != 0, reverse = (flags & 2u) != 0;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        for (long long j = 0; j < n; ++j) {
            for (long long k = 0; k < nh; ++k) {
                double* o = out + 2 * ((i * n + j) * nh + k);
                const long long mx = or_mode(i, n), my = or_mode(j, n), mz = k;
                o[0] = 0.0;
                o[1] = 0.0;
                if (mx == 0 && my == 0 && mz == 0) continue;
                if (llabs(mx) > (long long)max_mode || llabs(my) > (long long)max_mode || mz > (long long)max_mode)
                    continue;
                const int plane = (k == 0) || (n % 2 == 0 && k == n / 2);
                long long ci = i, cj = j;
                int conj = 0, self = 0;
                if (plane) {
                    const long long ti = (n - i) % n, tj = (n - j) % n;
                    self = ti == i && tj == j;
                    if (!self && (ti < i || (ti == i && tj < j))) {
                        ci = ti;
                        cj = tj;
                        conj = 1;
                    }
                }
                const int32_t cx = (int32_t)or_mode(ci, n), cy = (int32_t)or_mode(cj, n), cz = (int32_t)mz;
                const double var = var_m2[mx * mx + my * my + mz * mz];
                double re, im;
                if (fixed) {
                    double u[2];
                    or_uniform_pair(seed, cx, cy, cz, 1u, u);
                    const double amp = sqrt(var);
                    if (self) {
                        re = (u[0] < 0.5 ? -1.0 : 1.0) * amp;
                        im = 0.0;
                    } else {
                        const double ph = 2.0 * 3.14159265358979323846 * u[0];
                        re = amp * cos(ph);
                        im = amp * sin(ph);
                        if (conj) im = -im;
                    }
                } else {
                    double g[2];
                    or_normal_pair(seed, cx, cy, cz, 0u, g);
                    if (self) {
                        re = sqrt(var) * g[0];
                        im = 0.0;
                    } else {
                        const double sd = sqrt(var / 2.0);
                        re = sd * g[0];
                        im = sd * g[1];
                        if (conj) im = -im;
                    }
                }
                if (reverse) {
                    re = -re;
                    im = -im;
                }
                o[0] = re;
                o[1] = im;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Matter power spectrum: MatterPowerSpectrum::compute_pk                     */
/* (src/analysis/matter_power_spectrum.cpp:57-241) over                       */
/* FourierDensityBuilder::build (src/analysis/fourier_density_field.cpp:      */
/* 112-343), with the helpers of analysis/spectral_numeric.hpp and            */
/* analysis/cic_spectral_window.hpp.                                          */
/* ------------------------------------------------------------------------- */

/* log(upper/lower) through log1p of the relative gap (spectral_numeric.hpp:28-47) */
/* ------------------------------------------------------------------------- */
/* First-order LPT: src/ic/lpt_displacement.cpp:89-129 (Nyquist rule, particle */
/* update), 202-221 (factors), 301-384 (per-axis spectrum, c2r, lattice         */
/* sampling); lattice src/ic/initial_conditions.cpp:296-312.                    */
/* ------------------------------------------------------------------------- */
int or_lpt_first_order(uint64_t N, const double* delta_k, uint64_t n_side, double L, double p_factor,
                       double* x, double* y, double* z, double* px, double* py, double* pz) {
    if (N < 2 || n_side == 0 || N % n_side != 0 || !(L > 0.0)) return OR_EINVAL;
    const uint64_t NH = N / 2 + 1;
    const uint64_t s = N / n_side;
    const double dx = L / (double)n_side;
    double* pos[3] = {x, y, z};
    double* mom[3] = {px, py, pz};
    const uint64_t np = n_side * n_side * n_side;
    /* lattice q = index * dx, zero momenta */
    for (uint64_t i = 0; i < n_side; ++i)
        for (uint64_t j = 0; j < n_side; ++j)
            for (uint64_t k = 0; k < n_side; ++k) {
                const uint64_t p = (i * n_side + j) * n_side + k;
                x[p] = (double)i * dx;
                y[p] = (double)j * dx;
                z[p] = (double)k * dx;
                px[p] = py[p] = pz[p] = 0.0;
            }
    double* psi_k = (double*)malloc(N * N * NH * 2 * sizeof(double));
    double* psi = (double*)malloc(N * N * N * sizeof(double));
    if (!psi_k || !psi) { free(psi_k); free(psi); return OR_ERUNTIME; }
    int bad = 0;
    for (int dim = 0; dim < 3 && !bad; ++dim) {
#pragma omp parallel for schedule(static) num_threads(OR_NT)
        for (int64_t ix = 0; ix < (int64_t)N; ++ix)
            for (uint64_t iy = 0; iy < N; ++iy) {
                const double kx = k_component((uint64_t)ix, N, L);
                const double ky = k_component(iy, N, L);
                for (uint64_t iz = 0; iz < NH; ++iz) {
                    const double kz = k_component(iz, N, L);
                    const double full[3] = {kx, ky, kz};
                    const uint64_t idx[3] = {(uint64_t)ix, iy, iz};
                    const double k2 = kx * kx + ky * ky + kz * kz;
                    const uint64_t c = ((uint64_t)ix * N + iy) * NH + iz;
                    if (k2 == 0.0) {
                        psi_k[2 * c] = 0.0;
                        psi_k[2 * c + 1] = 0.0;
                        continue;
                    }
                    const int nyq = N % 2 == 0 && idx[dim] == N / 2;
                    const double f = (nyq ? 0.0 : full[dim]) / k2;
                    const double re = delta_k[2 * c], im = delta_k[2 * c + 1];
                    psi_k[2 * c] = -im * f;
                    psi_k[2 * c + 1] = re * f;
                }
            }
        if (or_fft_c2r_3d(N, psi_k, psi) != OR_OK) { bad = 1; break; }
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: bad)
        for (int64_t p = 0; p < (int64_t)np; ++p) {
            const uint64_t i = (uint64_t)p / (n_side * n_side), j = ((uint64_t)p / n_side) % n_side,
                           k = (uint64_t)p % n_side;
            const double v = psi[((i * s) * N + j * s) * N + k * s];
            const double disp = 1.0 * v;
            const double inc = p_factor * v;
            const double cand = pos[dim][p] + disp;
            const double mc = mom[dim][p] + inc;
            const double w = or_wrap(cand, L);
            if (!isfinite(v) || !isfinite(disp) || !isfinite(inc) || !isfinite(cand) || !isfinite(mc) || !isfinite(w)) {
                bad = 1;
                continue;
            }
            pos[dim][p] = w;
            mom[dim][p] = mc;
        }
    }
    free(psi_k);
    free(psi);
    return bad ? OR_ERUNTIME : OR_OK;
}

static int pk_log_ratio(double upper, double lower, double* out) {
    if (!isfinite(upper) || !isfinite(lower) || lower <= 0.0 || upper < lower) return OR_EINVAL;
    if (upper == lower) { *out = 0.0; return OR_OK; }
    const double rel = (upper - lower) / lower;
    const double v = isfinite(rel) ? log1p(rel) : log(upper) - log(lower);
    if (!isfinite(v) || v <= 0.0) return OR_EOVERFLOW;
    *out = v;
    return OR_OK;
}

/* lower * (upper/lower)^f (spectral_numeric.hpp:49-87) */
static int pk_log_interp(double lower, double upper, double f, double* out) {
    if (!isfinite(lower) || !isfinite(upper) || lower <= 0.0 || upper < lower || !isfinite(f) || f < 0.0 ||
        f > 1.0)
        return OR_EINVAL;
    if (f == 0.0 || lower == upper) { *out = lower; return OR_OK; }
    if (f == 1.0) { *out = upper; return OR_OK; }
    double span = 0.0;
    const int e = pk_log_ratio(upper, lower, &span);
    if (e != OR_OK) return e;
    double r;
    if (span <= 1.0) {
        r = lower * exp(f * span);
    } else {
        const long double lv = logl((long double)lower) + (long double)f * (long double)span;
        const long double w = expl(lv);
        if (!isfinite(w) || w <= 0.0L || w > (long double)DBL_MAX) return OR_EOVERFLOW;
        r = (double)w;
    }
    if (!isfinite(r) || r <= 0.0 || r < lower || r > upper) return OR_EOVERFLOW;
    *out = r;
    return OR_OK;
}

int or_pk_edges(double minimum, double maximum, int bins, double* edges) {
    if (!isfinite(minimum) || !isfinite(maximum) || minimum <= 0.0 || maximum <= minimum || bins <= 0)
        return OR_EINVAL;
    for (int i = 0; i <= bins; ++i) {
        const int e = pk_log_interp(minimum, maximum, (double)i / (double)bins, &edges[i]);
        if (e != OR_OK) return e;
        if (i != 0 && !(edges[i] > edges[i - 1])) return OR_EOVERFLOW;
    }
    return OR_OK;
}

/* [edge_i, edge_i+1), last bin closed; -1 outside (spectral_numeric.hpp:124-141) */
int or_pk_bin_index(const double* edges, int bins, double v) {
    if (v < edges[0] || v > edges[bins]) return -1;
    if (v == edges[bins]) return bins - 1;
    int lo = 0, hi = bins + 1; /* first edge > v */
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (edges[mid] > v) hi = mid; else lo = mid + 1;
    }
    return lo - 1;
}

/* sinc^2(k dx / 2) in long double, |sinc| clamped to 1 (cic_spectral_window.hpp:13-37) */
long double or_pk_window_axis(double k, double dx) {
    const long double arg = 0.5L * (long double)k * (long double)dx;
    if (arg == 0.0L) return 1.0L;
    const long double s = sinl(arg) / arg;
    long double m = fabsl(s);
    if (m > 1.0L) m = 1.0L;
    return m * m;
}

/* value * L^3 in a shared exponent domain (spectral_numeric.hpp:170-205) */
static double pk_volume_times(double L, long double value) {
    if (value == 0.0L) return 0.0;
    int be = 0, ve = 0, ce = 0;
    const long double bf = frexpl((long double)L, &be);
    const long double vf = frexpl(value, &ve);
    long double c = vf;
    for (int i = 0; i < 3; ++i) c *= bf;
    c = frexpl(c, &ce);
    return (double)scalbnl(c, 3 * be + ve + ce);
}

/* |v|^2 L^3 (spectral_numeric.hpp:304-321 with :240-271) */
double or_pk_raw_power(double L, double re, double im) {
    const double scale = fmax(fabs(re), fabs(im));
    if (scale == 0.0) return 0.0;
    const long double rs = (long double)re / (long double)scale;
    const long double is = (long double)im / (long double)scale;
    const long double shape = rs * rs + is * is;
    int be = 0, se = 0, ce = 0;
    const long double bf = frexpl((long double)L, &be);
    const long double sf = frexpl((long double)scale, &se);
    long double c = shape * sf * sf * bf * bf * bf;
    c = frexpl(c, &ce);
    return (double)scalbnl(c, se + se + 3 * be + ce);
}

/* Exact sum of products of binary64 factors, rounded once to nearest-even:
 * the contract of math/exact_binary64_product_sum.hpp.  Products of up to five
 * factors are formed exactly as 320-bit integers times a power of two and
 * added into a two's-complement fixed-point accumulator spanning
 * [2^-5376, 2^5376). */
#define PK_ACC_LIMBS 170
#define PK_ACC_BASE (-5376)
typedef struct { uint64_t d[PK_ACC_LIMBS]; } pk_bigacc;

static void pk_acc_add_shifted(pk_bigacc* a, const uint64_t* v, int nv, int64_t bit, int negate) {
    /* a += (negate ? -1 : 1) * v * 2^bit, bit >= 0 relative to the accumulator */
    uint64_t tmp[PK_ACC_LIMBS];
    memset(tmp, 0, sizeof(tmp));
    const int64_t limb = bit / 64, sh = bit % 64;
    for (int i = 0; i < nv; ++i) {
        const int64_t j = limb + i;
        if (j < PK_ACC_LIMBS) tmp[j] |= v[i] << sh;
        if (sh && j + 1 < PK_ACC_LIMBS) tmp[j + 1] |= v[i] >> (64 - sh);
    }
    if (negate) {
        unsigned __int128 c = 1;
        for (int i = 0; i < PK_ACC_LIMBS; ++i) {
            c += (unsigned __int128)(~tmp[i]);
            tmp[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    unsigned __int128 c = 0;
    for (int i = 0; i < PK_ACC_LIMBS; ++i) {
        c += (unsigned __int128)a->d[i] + tmp[i];
        a->d[i] = (uint64_t)c;
        c >>= 64;
    }
}

static double pk_exact_product_sum(const double* const* terms, const int* counts, int nterms) {
    pk_bigacc acc;
    memset(&acc, 0, sizeof(acc));
    for (int t = 0; t < nterms; ++t) {
        uint64_t v[6] = {1, 0, 0, 0, 0, 0};
        int64_t e = 0;
        int neg = 0, zero = 0;
        for (int f = 0; f < counts[t]; ++f) {
            const double x = terms[t][f];
            if (x == 0.0) { zero = 1; break; }
            if (x < 0.0) neg ^= 1;
            int ex = 0;
            const double fr = frexp(fabs(x), &ex);
            const uint64_t m = (uint64_t)ldexp(fr, 53);
            e += ex - 53;
            unsigned __int128 c = 0;
            for (int i = 0; i < 6; ++i) {
                c += (unsigned __int128)v[i] * m;
                v[i] = (uint64_t)c;
                c >>= 64;
            }
        }
        if (zero) continue;
        pk_acc_add_shifted(&acc, v, 6, e - PK_ACC_BASE, neg);
    }
    /* sign and magnitude */
    const int negative = (acc.d[PK_ACC_LIMBS - 1] >> 63) != 0;
    if (negative) {
        unsigned __int128 c = 1;
        for (int i = 0; i < PK_ACC_LIMBS; ++i) {
            c += (unsigned __int128)(~acc.d[i]);
            acc.d[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    int top = -1;
    for (int i = PK_ACC_LIMBS - 1; i >= 0 && top < 0; --i)
        if (acc.d[i]) top = i * 64 + 63 - __builtin_clzll(acc.d[i]);
    if (top < 0) return 0.0;
    /* keep 53 significant bits (fewer below the normal range) */
    const int64_t vexp = (int64_t)top + PK_ACC_BASE; /* value in [2^vexp, 2^(vexp+1)) */
    int64_t lsb = top - 52;
    if (vexp - 52 < -1074) lsb = -1074 - PK_ACC_BASE;
    uint64_t mant = 0;
    for (int64_t b = top; b >= lsb; --b)
        mant = (mant << 1) | ((acc.d[b / 64] >> (b % 64)) & 1u);
    int round = 0, sticky = 0;
    if (lsb - 1 >= 0) {
        round = (int)((acc.d[(lsb - 1) / 64] >> ((lsb - 1) % 64)) & 1u);
        for (int64_t b = lsb - 2; b >= 0 && !sticky; --b)
            if ((acc.d[b / 64] >> (b % 64)) & 1u) sticky = 1;
    }
    if (round && (sticky || (mant & 1u))) mant += 1;
    const double r = ldexp((double)mant, (int)(lsb + PK_ACC_BASE));
    return negative ? -r : r;
}

/* |v|^2 L^3 - s (spectral_numeric.hpp:326-391).  Severe cancellation is
 * resolved by an exact sum of the binary64 products, rounded once. */
double or_pk_raw_power_minus(double L, double re, double im, double s) {
    const double scale = fmax(fabs(re), fabs(im));
    if (scale == 0.0) return s == 0.0 ? 0.0 : -s;
    const long double rs = (long double)re / (long double)scale;
    const long double is = (long double)im / (long double)scale;
    const long double shape = rs * rs + is * is;
    int be = 0, se = 0, re_ = 0;
    const long double bf = frexpl((long double)L, &be);
    const long double sf = frexpl((long double)scale, &se);
    long double rc = shape * sf * sf * bf * bf * bf;
    rc = frexpl(rc, &re_);
    const int rexp = 2 * se + 3 * be + re_;
    if (s == 0.0) return (double)scalbnl(rc, rexp);
    int sexp = 0;
    const long double sc = frexpl((long double)s, &sexp);
    const int common = rexp > sexp ? rexp : sexp;
    const long double ar = scalbnl(rc, rexp - common);
    const long double as = scalbnl(sc, sexp - common);
    const long double diff = ar - as;
    const long double env = 32.0L * LDBL_EPSILON * (fabsl(ar) + fabsl(as));
    if (fabsl(ar) + fabsl(as) > 0.0L && fabsl(diff) <= env) {
        const double t1[5] = {re, re, L, L, L}, t2[5] = {im, im, L, L, L}, t3[1] = {-s};
        const double* terms[3] = {t1, t2, t3};
        const int counts[3] = {5, 5, 1};
        return pk_exact_product_sum(terms, counts, 3);
    }
    return (double)scalbnl(diff, common);
}

typedef struct { long double k, raw, corr, dec; uint64_t count; } pk_acc;

/* Neumaier-compensated non-negative sum (fourier_density_field.cpp:27-59) */
static void pk_comp_add(long double v, long double* sum, long double* comp) {
    if (v == 0.0L) return;
    const long double u = *sum + v;
    *comp += fabsl(*sum) >= fabsl(v) ? (*sum - u) + v : (v - u) + *sum;
    *sum = u;
}

/* One contrast transform: deposit, density / mean - 1, r2c, times 1/N^3
 * (fourier_density_field.cpp:230-295). */
static int pk_build_one(uint64_t n, const double* x, const double* y, const double* z, const double* w,
                        double L, uint64_t N, double mean, double* mesh, double* modes) {
    int e = or_deposit(n, x, y, z, w, 1.0, L, N, mesh);
    if (e != OR_OK) return e;
    const uint64_t cells = N * N * N;
    for (uint64_t i = 0; i < cells; ++i) {
        const double c = mesh[i] / mean - 1.0;
        if (!isfinite(c)) return OR_ERUNTIME;
        mesh[i] = c;
    }
    e = or_fft_r2c_3d(N, mesh, modes);
    if (e != OR_OK) return e;
    const double norm = 1.0 / (double)cells;
    const uint64_t m = N * N * (N / 2 + 1) * 2;
    for (uint64_t i = 0; i < m; ++i) modes[i] *= norm;
    return OR_OK;
}

int or_power_spectrum(uint64_t n, const double* x, const double* y, const double* z,
                      const double* masses, double uniform_mass, double L, uint64_t N, int num_bins,
                      int subtract_shot_noise, int interlaced, double max_fraction, double* out) {
    if (num_bins < 1) return OR_EINVAL;
    if (!isfinite(max_fraction) || max_fraction <= 0.0 || max_fraction > 1.0) return OR_EINVAL;
    if (interlaced && subtract_shot_noise) return OR_EINVAL;
    if (N < 2 || n == 0) return OR_EINVAL;
    for (uint64_t p = 0; p < n; ++p)
        if (!isfinite(x[p]) || !isfinite(y[p]) || !isfinite(z[p])) return OR_EINVAL;

    /* mass moments in units of one positive mass (fourier_density_field.cpp:160-225) */
    double wsum, msf;
    double* w = NULL;
    if (!masses) {
        if (!isfinite(uniform_mass) || uniform_mass <= 0.0) return OR_EINVAL;
        wsum = (double)n;
        msf = (double)(1.0L / (long double)n);
    } else {
        double unit = 0.0;
        for (uint64_t p = 0; p < n; ++p) {
            if (!isfinite(masses[p]) || masses[p] <= 0.0) return OR_EINVAL;
            unit = fmax(unit, masses[p]);
        }
        w = (double*)malloc(n * sizeof(double));
        if (!w) return OR_ERUNTIME;
        long double s1 = 0.0L, c1 = 0.0L, s2 = 0.0L, c2 = 0.0L;
        for (uint64_t p = 0; p < n; ++p) {
            w[p] = masses[p] / unit;
            const long double v = (long double)w[p];
            pk_comp_add(v, &s1, &c1);
            pk_comp_add(v * v, &s2, &c2);
        }
        const long double sw = s1 + c1, sw2 = s2 + c2;
        wsum = (double)sw;
        const double r = (double)(sw2 / (sw * sw));
        if (r > 1.0 + 64.0 * DBL_EPSILON) { free(w); return OR_ELOGIC; }
        msf = fmin(1.0, r);
    }
    const uint64_t NH = N / 2 + 1, cells = N * N * N, cm = N * N * NH;
    const double mean = wsum / (double)cells;
    double* mesh = (double*)malloc(cells * sizeof(double));
    double* modes = (double*)malloc(cm * 2 * sizeof(double));
    double* m2 = interlaced ? (double*)malloc(cm * 2 * sizeof(double)) : NULL;
    double* sx = interlaced ? (double*)malloc(3 * n * sizeof(double)) : NULL;
    double* edges = (double*)malloc((size_t)(num_bins + 1) * sizeof(double));
    pk_acc* slabs = (pk_acc*)calloc(N * (uint64_t)num_bins, sizeof(pk_acc));
    int err = OR_OK;
    if (!mesh || !modes || !edges || !slabs || (interlaced && (!m2 || !sx))) { err = OR_ERUNTIME; goto done; }
    err = pk_build_one(n, x, y, z, w, L, N, mean, mesh, modes);
    if (err != OR_OK) goto done;
    const double dx = L / (double)N;
    if (interlaced) {
        /* a second deposit shifted by -dx/2 and a phase-corrected average
         * (fourier_density_field.cpp:300-336) */
        const double shift = 0.5 * dx;
        for (uint64_t p = 0; p < n; ++p) {
            sx[p] = or_wrap(x[p] - shift, L);
            sx[n + p] = or_wrap(y[p] - shift, L);
            sx[2 * n + p] = or_wrap(z[p] - shift, L);
        }
        err = pk_build_one(n, sx, sx + n, sx + 2 * n, w, L, N, mean, mesh, m2);
        if (err != OR_OK) goto done;
        for (uint64_t ix = 0; ix < N; ++ix) {
            const double kx = 2.0 * OR_PI * (double)mode_number(ix, N) / L;
            for (uint64_t iy = 0; iy < N; ++iy) {
                const double ky = 2.0 * OR_PI * (double)mode_number(iy, N) / L;
                for (uint64_t iz = 0; iz < NH; ++iz) {
                    const double kz = 2.0 * OR_PI * (double)mode_number(iz, N) / L;
                    const uint64_t i = 2 * (iz + NH * (iy + N * ix));
                    const double ph = -(kx + ky + kz) * shift;
                    const double c = cos(ph), s = sin(ph);
                    const double pr = m2[i] * c - m2[i + 1] * s;
                    const double pi = m2[i] * s + m2[i + 1] * c;
                    modes[i] = 0.5 * (modes[i] + pr);
                    modes[i + 1] = 0.5 * (modes[i + 1] + pi);
                }
            }
        }
    }
    modes[0] = 0.0;
    modes[1] = 0.0;

    {
        const double knyq = OR_PI * (double)N / L;
        const double kmax = max_fraction * knyq;
        const double kf = 2.0 * OR_PI / L;
        if (!(kmax > kf)) { err = OR_EINVAL; goto done; }
        err = or_pk_edges(kf, kmax, num_bins, edges);
        if (err != OR_OK) goto done;
        const double shot = pk_volume_times(L, (long double)msf);
        const int even = (N % 2) == 0;
        int failed = 0;
#pragma omp parallel for schedule(static) num_threads(OR_NT) reduction(|: failed)
        for (int64_t ix = 0; ix < (int64_t)N; ++ix) {
            pk_acc* local = slabs + (uint64_t)ix * (uint64_t)num_bins;
            const double kx = 2.0 * OR_PI * (double)mode_number((uint64_t)ix, N) / L;
            const double sxa = 0.5 * kx * dx;
            const long double cx = or_pk_window_axis(kx, dx);
            for (uint64_t iy = 0; iy < N; ++iy) {
                const double ky = 2.0 * OR_PI * (double)mode_number(iy, N) / L;
                const double sya = 0.5 * ky * dx;
                const long double cy = or_pk_window_axis(ky, dx);
                for (uint64_t iz = 0; iz < NH; ++iz) {
                    const double kz = 2.0 * OR_PI * (double)mode_number(iz, N) / L;
                    const double sza = 0.5 * kz * dx;
                    const long double cz = or_pk_window_axis(kz, dx);
                    const double kmag = hypot(hypot(kx, ky), kz);
                    if (!(kmag > 0.0) || kmag > kmax) continue;
                    const uint64_t i = 2 * (iz + NH * (iy + N * (uint64_t)ix));
                    const double praw = or_pk_raw_power(L, modes[i], modes[i + 1]);
                    const long double win = cx * cy * cz;
                    if (!(win > 0.0L) || win > 1.0L) { failed = 1; continue; }
                    const long double decw = 1.0L / (win * win);
                    const double dec = (double)decw;
                    const double s2x = sin(sxa) * sin(sxa);
                    const double s2y = sin(sya) * sin(sya);
                    const double s2z = sin(sza) * sin(sza);
                    const double alias = shot * (1.0 - 2.0 / 3.0 * s2x) * (1.0 - 2.0 / 3.0 * s2y) *
                                         (1.0 - 2.0 / 3.0 * s2z);
                    double pc = praw;
                    if (subtract_shot_noise) pc = or_pk_raw_power_minus(L, modes[i], modes[i + 1], alias);
                    pc *= dec;
                    if (!isfinite(pc)) { failed = 1; continue; }
                    const uint64_t mult = (iz == 0 || (even && iz == N / 2)) ? 1u : 2u;
                    const int b = or_pk_bin_index(edges, num_bins, kmag);
                    if (b < 0) continue;
                    const long double lm = (long double)mult;
                    local[b].k += (long double)kmag * lm;
                    local[b].raw += (long double)praw * lm;
                    local[b].corr += (long double)pc * lm;
                    local[b].dec += (long double)dec * lm;
                    local[b].count += mult;
                }
            }
        }
        if (failed) { err = OR_EOVERFLOW; goto done; }
        for (int b = 0; b < num_bins; ++b) {
            pk_acc t = {0.0L, 0.0L, 0.0L, 0.0L, 0};
            for (uint64_t ix = 0; ix < N; ++ix) {
                const pk_acc* s = slabs + ix * (uint64_t)num_bins + b;
                t.k += s->k;
                t.raw += s->raw;
                t.corr += s->corr;
                t.dec += s->dec;
                t.count += s->count;
            }
            double* o = out + 8 * b;
            o[0] = edges[b];
            o[2] = edges[b + 1];
            o[6] = shot;
            o[1] = o[3] = o[4] = o[5] = o[7] = 0.0;
            if (t.count == 0) continue;
            const long double cnt = (long double)t.count;
            o[1] = (double)(t.k / cnt);
            o[3] = (double)(t.raw / cnt);
            o[4] = (double)(t.corr / cnt);
            o[5] = (double)(t.dec / cnt);
            o[7] = (double)t.count;
        }
    }
done:
    free(w);
    free(mesh);
    free(modes);
    free(m2);
    free(sx);
    free(edges);
    free(slabs);
    return err;
}

// Device-side building blocks shared by the PM kernels (gfx950, wave64).
//
// Periodic wrap, minimum image and the per-axis CIC coordinate restate the
// reference arithmetic exactly (binary64, no contraction: the library is built
// with -ffp-contract=off), so cell keys and fractions are bit-identical to the
// CPU path:
//   wrap / minimum_image : include/cosmo_nbody/math/periodic_box.hpp:27-61
//   CIC coordinate       : src/mesh/cic_coordinate.hpp:95-121
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// helpers usable from host code too (CPU checks of the index arithmetic)
#define HY_HD __host__ __device__ __forceinline__

namespace hy {

// Error bits raised by kernels into a device word; the host maps them to the
// reference exception classes.
enum : uint32_t {
    ERR_POSITION_NONFINITE = 1u << 0,  // invalid_argument "CIC positions must be finite"
    ERR_COORDINATE = 1u << 1,          // logic_error from CICCoordinateAxis
    ERR_MASS_INVALID = 1u << 2,        // invalid_argument (negative / non-finite mass)
    ERR_SCRATCH = 1u << 3,             // runtime_error: deposit scratch exhausted
    ERR_FIELD_NONFINITE = 1u << 4,     // invalid_argument: potential not finite
    ERR_FORCE_NONFINITE = 1u << 5,     // overflow_error: gathered force not finite
    ERR_UPDATE_NONFINITE = 1u << 6,    // overflow_error: kick/drift not finite
    ERR_MOMENTUM_NONFINITE = 1u << 7,  // overflow_error: kick input not finite
    ERR_STALE_KEYS = 1u << 8,          // reused bucket keys do not match the positions
    ERR_SPECTRUM_NONFINITE = 1u << 9,  // runtime_error: a Fourier mode or shell power is not finite
    ERR_INTERP_UNDERFLOW = 1u << 10,   // underflow_error: a CIC interpolation product rounded to zero
    ERR_MASS_CONSERVATION = 1u << 11,  // runtime_error: deposited mass outside the forward-error bound
    ERR_ROUTING = 1u << 12,            // logic_error: bucket routing did not cover each particle once
    ERR_GATHER_UNDERFLOW = 1u << 13,   // underflow_error: a centered-gradient gather product rounded to zero
};

struct Geom {
    double L;       // box size
    double dx;      // L / N
    double inv_dx;  // 1 / dx, used only when dx is a power of two (then w*inv_dx == w/dx)
    int inv_exact;  // dx is a power of two
    double upper;   // nextafter((double)N, 0)
    int N;          // mesh cells per axis
    int T;          // deposit tile edge (divides N, <= 16)
    int nT;         // N / T
    uint32_t nbox;  // boxes of this geometry held by the context (>= 1)
    uint32_t nper;  // particles per box in the current call (arrays are box-major)
};

// Box of global particle p (particle arrays concatenate the boxes).
__device__ __forceinline__ uint32_t box_of(uint64_t p, const Geom& g) {
    return g.nbox > 1 ? (uint32_t)p / g.nper : 0u;
}

// Deposit buckets per box: (N/T)^3 tiles x 8 boundary classes.
__device__ __forceinline__ uint32_t buckets_per_box(const Geom& g) {
    return (uint32_t)g.nT * (uint32_t)g.nT * (uint32_t)g.nT * 8u;
}

__device__ __forceinline__ double wrap_pos(double x, double L) {
    if (!isfinite(x)) return __builtin_nan("");
    if (x >= 0.0 && x < L) return x;
    double w = fmod(x, L);
    if (w < 0.0) w += L;
    if (w >= L) w = 0.0;
    return w == 0.0 ? 0.0 : w;
}

__device__ __forceinline__ double minimum_image(double d, double L) {
    if (!isfinite(d)) return __builtin_nan("");
    const double half = 0.5 * L;
    if (d >= -half && d < half) return d;
    double w = remainder(d, L);
    if (w >= half) w -= L;
    else if (w < -half) w += L;
    return w == 0.0 ? 0.0 : w;
}

// Returns false when the coordinate is not admissible (non-finite input).
__device__ __forceinline__ bool cic_axis(double pos, const Geom& g, int& base, double& frac) {
    const double w = wrap_pos(pos, g.L);
    if (!(w >= 0.0 && w < g.L)) return false;
    double s = g.inv_exact ? w * g.inv_dx : w / g.dx;
    s = s < 0.0 ? 0.0 : s;
    s = s > g.upper ? g.upper : s;
    const double f = floor(s);
    base = (int)f;
    frac = s - f;
    return true;
}

// cic_axis over the three axes of a position already loaded, all three
// evaluated (no short circuit), so a caller's three coordinate loads go out
// together instead of each waiting on the previous axis's check.
__device__ __forceinline__ bool cic_axes(double px, double py, double pz, const Geom& g, int& bx, int& by, int& bz,
                                         double& fx, double& fy, double& fz) {
    const bool ox = cic_axis(px, g, bx, fx);
    const bool oy = cic_axis(py, g, by, fy);
    const bool oz = cic_axis(pz, g, bz, fz);
    return ox && oy && oz;
}

__device__ __forceinline__ int wrap_cell(int i, int N) {
    return i < 0 ? i + N : (i >= N ? i - N : i);
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2.  Remap so that each XCD receives a contiguous range of
// logical blocks (neighbouring tiles / rows then share that L2).
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t XCDS = 8;
    const uint32_t per = nblocks / XCDS;
    const uint32_t full = per * XCDS;
    if (bid >= full) return bid;  // ragged tail keeps identity mapping
    return (bid % XCDS) * per + bid / XCDS;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

template <typename T>
struct cplx_t;
template <>
struct cplx_t<float> { using type = float2; };
template <>
struct cplx_t<double> { using type = double2; };
template <typename T>
using cplx = typename cplx_t<T>::type;

}  // namespace hy

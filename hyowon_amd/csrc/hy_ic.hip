// Mode-keyed Gaussian initial field on the GPU (gfx950).
//
// Every r2c mode (i, j, k) of an N^3 IC mesh draws its own Philox4x32-10 block
// keyed by the seed with counter (m_x, m_y, m_z, stream) of its canonical
// (integer) mode, so realisations do not depend on the mesh size or on the
// order of generation (src/ic/random_field.cpp:70-330, src/ic/philox_rng.cpp:
// 1-90).  One thread per mode; the variance of each mode comes from a table
// indexed by m_x^2 + m_y^2 + m_z^2 (the power spectrum is a function of |m|).
#include "hyowon_pm.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

namespace {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// 53 addressed bits mapped into (0,1): j > 0 -> j 2^-53, 0 -> 2^-54.
__device__ __forceinline__ double open53(uint32_t hi, uint32_t lo) {
    const uint64_t bits = ((uint64_t)hi << 21) | ((uint64_t)lo >> 11);
    return bits == 0 ? 0x1.0p-54 : (double)bits * 0x1.0p-53;
}

__device__ __forceinline__ long long mode_of(long long i, long long n) { return i <= n / 2 ? i : i - n; }

__global__ void k_gaussian_modes(uint64_t N, uint64_t seed, uint64_t max_mode, const double* __restrict__ var,
                                 uint32_t flags, double2* __restrict__ out) {
    const long long n = (long long)N, nh = n / 2 + 1;
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)(n * n * nh)) return;
    const long long k = (long long)(idx % nh), j = (long long)((idx / nh) % n), i = (long long)(idx / (nh * n));
    const long long mx = mode_of(i, n), my = mode_of(j, n), mz = k;
    double2 v = {0.0, 0.0};
    const long long km = (long long)max_mode;
    if (!(mx == 0 && my == 0 && mz == 0) && llabs(mx) <= km && llabs(my) <= km && mz <= km) {
        const bool plane = k == 0 || (n % 2 == 0 && k == n / 2);
        long long ci = i, cj = j;
        bool conj = false, self = false;
        if (plane) {
            const long long ti = (n - i) % n, tj = (n - j) % n;
            self = ti == i && tj == j;
            if (!self && (ti < i || (ti == i && tj < j))) {
                ci = ti;
                cj = tj;
                conj = true;
            }
        }
        const bool fixed = (flags & 1u) != 0;
        uint32_t c[4] = {(uint32_t)(int32_t)mode_of(ci, n), (uint32_t)(int32_t)mode_of(cj, n), (uint32_t)(int32_t)mz,
                         fixed ? 1u : 0u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u1 = open53(c[0], c[1]), u2 = open53(c[2], c[3]);
        const double vr = var[mx * mx + my * my + mz * mz];
        const double pi = 3.14159265358979323846;
        double re, im;
        if (fixed) {
            const double amp = sqrt(vr);
            if (self) {
                re = (u1 < 0.5 ? -1.0 : 1.0) * amp;
                im = 0.0;
            } else {
                const double ph = 2.0 * pi * u1;
                re = amp * cos(ph);
                im = amp * sin(ph);
                if (conj) im = -im;
            }
        } else {
            const double r = sqrt(-2.0 * log(u1));
            const double th = 2.0 * pi * u2;
            const double g1 = r * cos(th), g2 = r * sin(th);
            if (self) {
                re = sqrt(vr) * g1;
                im = 0.0;
            } else {
                const double sd = sqrt(vr / 2.0);
                re = sd * g1;
                im = sd * g2;
                if (conj) im = -im;
            }
        }
        if (flags & 2u) {
            re = -re;
            im = -im;
        }
        v = {re, im};
    }
    out[idx] = v;
}

thread_local std::string g_ic_error;

}  // namespace

extern "C" {

int hy_ic_gaussian_modes(uint64_t N, uint64_t seed, uint64_t max_mode, const double* variance_by_m2, void* out,
                         uint32_t flags, int device, void* stream) {
    if (N < 2 || N > 4096 || !variance_by_m2 || !out) {
        g_ic_error = "hy_ic_gaussian_modes: mesh size must be 2..4096 and pointers non-null";
        return HY_EINVAL;
    }
    if (max_mode > N / 2) {
        g_ic_error = "hy_ic_gaussian_modes: max_mode exceeds N/2";
        return HY_EINVAL;
    }
    if (hipSetDevice(device) != hipSuccess) {
        g_ic_error = "hipSetDevice failed";
        return HY_EDEVICE;
    }
    const uint64_t count = N * N * (N / 2 + 1);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_gaussian_modes, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, N, seed, max_mode,
                       variance_by_m2, flags, static_cast<double2*>(out));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_ic_error = std::string("k_gaussian_modes: ") + hipGetErrorString(e);
        return HY_EDEVICE;
    }
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e2 != hipSuccess) {
        g_ic_error = std::string("hipStreamSynchronize: ") + hipGetErrorString(e2);
        return HY_EDEVICE;
    }
    return HY_OK;
}

const char* hy_ic_last_error(void) { return g_ic_error.c_str(); }

}  // extern "C"

// Table of the compiled FFT pass kernels: per mesh size, per pass, per launch
// variant.  The instantiations live in hy_fft_f32.hip and hy_fft_f64.hip so
// the two precisions compile in parallel.
#pragma once

#include "hy_fft_passes.hpp"

// Supported mesh sizes (compile-time FFT plans).
#define HY_MESH_SIZES(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(1024)

namespace hy {

enum { FFT_Z_R2C = 0, FFT_Y_FWD, FFT_X_GREEN, FFT_Y_BWD, FFT_Z_C2R, FFT_PASSES };

template <typename T>
struct FftPass {
    void (*fn)(FftArgs<T>) = nullptr;
    int rows = 0, threads = 0, lds = 0;
};

template <typename T>
struct FftKernels {
    FftPass<T> pass[FFT_PASSES][kFftVariants];
    FftPass<T> x_fwd;  // forward-only x pass of the spectral estimators (in place on B)
    FftPass<T> z_fwd;  // z r2c pass of the spectral estimators (stores kz <= mode_limit only)
    FftPass<T> x_inv;  // inverse-only x pass of the IC displacements (in place on B)
    FftPass<T> sg[3];  // spectral-gradient x, y^-1 and z^-1 passes (TreePM long range; variant 0)
    bool sg_x_dual = false;  // sg[0] writes the -i k_x and the plain inverse from one read of B
    FftPass<T> sg_y_dual;    // y^-1 pass writing the -i k_y and the plain inverse from one read of B (or none)
    int variants = 0;  // launch variants compiled for this size (variant 0 is the default)
};

// Sizes below this have one launch variant (their passes are latency-bound and
// the variants collapse onto the same row counts).
constexpr int kFftTunedFrom = 192;

// Defined in hy_fft_f32.hip / hy_fft_f64.hip; false for an unsupported size.
bool lookup_fft(int N, FftKernels<float>& out);
bool lookup_fft(int N, FftKernels<double>& out);

// Power-of-two meshes from 128 up use the two-level register passes; the
// others (3*2^a, small meshes) use the LDS Stockham passes.
constexpr bool fft_two_level(int N) { return N >= 128 && (N & (N - 1)) == 0; }

#ifdef HY_FFT_INSTANTIATE
template <typename T, int N, int V>
void fill_variant(FftKernels<T>& k) {
    if constexpr (fft_two_level(N)) {
        using RZ = TLCfg<T, N / 2, N, V>;
        using RY = TLCfg<T, N, N, V>;
        using RX = TLCfg<T, N, (N / 2 + 1) * N, V>;
        k.pass[FFT_Z_R2C][V] = {k_tl_z_r2c<T, N, V>, RZ::ROWS, RZ::THREADS, RZ::LDS_BYTES};
        k.pass[FFT_Y_FWD][V] = {k_tl_y_fwd<T, N, V>, RY::ROWS, RY::THREADS, RY::LDS_BYTES};
        k.pass[FFT_X_GREEN][V] = {k_tl_x_green<T, N, V>, RX::ROWS, RX::THREADS, RX::X_LDS_BYTES};
        k.pass[FFT_Y_BWD][V] = {k_tl_y_bwd<T, N, V>, RY::ROWS, RY::THREADS, RY::LDS_BYTES};
        k.pass[FFT_Z_C2R][V] = {k_tl_z_c2r<T, N, V>, RZ::ROWS, RZ::THREADS, RZ::LDS_BYTES};
        if constexpr (V == 0) {
            k.x_fwd = {k_tl_x_dft<T, N, V, -1>, RX::ROWS, RX::THREADS, RX::X_LDS_BYTES};
            k.z_fwd = {k_tl_z_r2c<T, N, V, true>, RZ::ROWS, RZ::THREADS, RZ::LDS_BYTES};
            k.x_inv = {k_tl_x_dft<T, N, V, +1>, RX::ROWS, RX::THREADS, RX::X_LDS_BYTES};
            if constexpr (sizeof(T) == 4) {
                // binary32: the second held column fits the register file
                k.sg[0] = {k_tl_x_green_dual<T, N, V>, RX::ROWS, RX::THREADS, RX::X_LDS_BYTES};
                k.sg_x_dual = true;
                k.sg_y_dual = {k_tl_y_bwd<T, N, V, 2>, RY::ROWS, RY::THREADS, RY::LDS_BYTES};
            } else {
                k.sg[0] = {k_tl_x_green<T, N, V, 1>, RX::ROWS, RX::THREADS, RX::X_LDS_BYTES};
            }
            k.sg[1] = {k_tl_y_bwd<T, N, V, 1>, RY::ROWS, RY::THREADS, RY::LDS_BYTES};
            k.sg[2] = {k_tl_z_c2r<T, N, V, 1>, RZ::ROWS, RZ::THREADS, RZ::LDS_BYTES};
        }
    } else {
        using RZ = RowsCfg<T, N / 2, N, V>;
        using RY = RowsCfg<T, N, N, V>;
        using RX = RowsCfg<T, N, (N / 2 + 1) * N, V>;
        k.pass[FFT_Z_R2C][V] = {k_fft_z_r2c<T, N, V>, RZ::ROWS, RZ::THREADS, RZ::LDS_BYTES};
        k.pass[FFT_Y_FWD][V] = {k_fft_y_fwd<T, N, V>, RY::ROWS, RY::THREADS, RY::LDS_BYTES};
        k.pass[FFT_X_GREEN][V] = {k_fft_x_green<T, N, V>, RX::ROWS, RX::THREADS, RX::LDS_BYTES};
        k.pass[FFT_Y_BWD][V] = {k_fft_y_bwd<T, N, V>, RY::ROWS, RY::THREADS, RY::LDS_BYTES};
        k.pass[FFT_Z_C2R][V] = {k_fft_z_c2r<T, N, V>, RZ::ROWS, RZ::THREADS, RZ::LDS_BYTES};
        if constexpr (V == 0) {
            k.x_fwd = {k_fft_x_dft<T, N, V, -1>, RX::ROWS, RX::THREADS, RX::LDS_BYTES};
            k.z_fwd = k.pass[FFT_Z_R2C][V];  // stores every plane
            k.x_inv = {k_fft_x_dft<T, N, V, +1>, RX::ROWS, RX::THREADS, RX::LDS_BYTES};
            k.sg[0] = {k_fft_x_green<T, N, V, 1>, RX::ROWS, RX::THREADS, RX::LDS_BYTES};
            k.sg[1] = {k_fft_y_bwd<T, N, V, 1>, RY::ROWS, RY::THREADS, RY::LDS_BYTES};
            k.sg[2] = {k_fft_z_c2r<T, N, V, 1>, RZ::ROWS, RZ::THREADS, RZ::LDS_BYTES};
        }
    }
}

template <typename T, int N>
FftKernels<T> make_fft() {
    FftKernels<T> k{};
    fill_variant<T, N, 0>(k);
    k.variants = 1;
    if constexpr (N >= kFftTunedFrom) {
        fill_variant<T, N, 1>(k);
        fill_variant<T, N, 2>(k);
        fill_variant<T, N, 3>(k);
        k.variants = kFftVariants;
    }
    return k;
}

template <typename T>
bool lookup_fft_impl(int N, FftKernels<T>& out) {
    switch (N) {
#define HY_CASE(n) \
    case n: out = make_fft<T, n>(); return true;
        HY_MESH_SIZES(HY_CASE)
#undef HY_CASE
        default: return false;
    }
}
#endif

}  // namespace hy

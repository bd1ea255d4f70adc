// Device building blocks for the Born-series kernels on gfx950 (CDNA4, 64-lane wavefronts).
//
// Complex arithmetic, radix-2/3/4/5/8 and radix-7 DFT butterflies, and the self-sorting
// Stockham pass that runs over a tile of pencils staged in LDS. The pass schedule (twiddle index
// j*k*N/(p*R), output index (i-k)*R + k + j*p) is modelled on the CPU by the oracle's stockham_fft.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mmx_plan.hpp"

namespace mmx {


template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T r, i;
};

template <typename T> __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <typename T> __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.r - b.r, a.i - b.i}; }
template <typename T> __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <typename T> __device__ __forceinline__ cx<T> scal(cx<T> a, T s) { return {a.r * s, a.i * s}; }
template <typename T> __device__ __forceinline__ cx<T> conjg(cx<T> a) { return {a.r, -a.i}; }
template <typename T> __device__ __forceinline__ T norm2(cx<T> a) { return a.r * a.r + a.i * a.i; }
// |a|^2 evaluated in FP64 (norm accumulation)
template <typename T> __device__ __forceinline__ double norm2_f64(cx<T> a) {
    return double(a.r) * double(a.r) + double(a.i) * double(a.i);
}
// a * (DIR * i): DIR = -1 for the forward transform, +1 for the inverse.
template <typename T, int DIR> __device__ __forceinline__ cx<T> mul_dir_i(cx<T> a) {
    return DIR > 0 ? cx<T>{-a.i, a.r} : cx<T>{a.i, -a.r};
}
// a / b for complex b (used for the Green's function, once per voxel)
template <typename T> __device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
    T den = b.r * b.r + b.i * b.i;
    T inv = T(1) / den;
    return {(a.r * b.r + a.i * b.i) * inv, (a.i * b.r - a.r * b.i) * inv};
}

// ----------------------------------------------------------------------------------------------
// Butterflies: y_j = sum_m a_m exp(DIR 2 pi i j m / R), in place.
// ----------------------------------------------------------------------------------------------
template <typename T, int DIR> __device__ __forceinline__ void dft2(cx<T>* a) {
    cx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
}

template <typename T, int DIR> __device__ __forceinline__ void dft4(cx<T>* a) {
    cx<T> t0 = a[0] + a[2], t1 = a[0] - a[2];
    cx<T> t2 = a[1] + a[3], t3 = mul_dir_i<T, DIR>(a[1] - a[3]);
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + t3;
    a[3] = t1 - t3;
}

template <typename T, int DIR> __device__ __forceinline__ void dft8(cx<T>* a) {
    cx<T> e[4] = {a[0], a[2], a[4], a[6]};
    cx<T> o[4] = {a[1], a[3], a[5], a[7]};
    dft4<T, DIR>(e);
    dft4<T, DIR>(o);
    const T h = T(0.70710678118654752440084436210485);
    // o1 *= (h, DIR h); o2 *= DIR i; o3 *= (-h, DIR h)
    o[1] = cx<T>{h * (o[1].r - T(DIR) * o[1].i), h * (o[1].i + T(DIR) * o[1].r)};
    o[2] = mul_dir_i<T, DIR>(o[2]);
    o[3] = cx<T>{-h * (o[3].r + T(DIR) * o[3].i), h * (T(DIR) * o[3].r - o[3].i)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = e[k] + o[k];
        a[k + 4] = e[k] - o[k];
    }
}

template <typename T, int DIR> __device__ __forceinline__ void dft3(cx<T>* a) {
    const T c = T(-0.5);
    const T s = T(DIR) * T(0.86602540378443864676372317075294);
    cx<T> t = a[1] + a[2], u = a[1] - a[2];
    cx<T> m = a[0] + scal(t, c);
    cx<T> iu = cx<T>{-u.i * s, u.r * s};  // i s u
    a[0] = a[0] + t;
    a[1] = m + iu;
    a[2] = m - iu;
}

template <typename T, int DIR> __device__ __forceinline__ void dft5(cx<T>* a) {
    const T c1 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
    const T s1 = T(DIR) * T(0.95105651629515357211643933337938);  // sin(2pi/5)
    const T s2 = T(DIR) * T(0.58778525229247312916870595463907);  // sin(4pi/5)
    cx<T> t1 = a[1] + a[4], t2 = a[2] + a[3];
    cx<T> u1 = a[1] - a[4], u2 = a[2] - a[3];
    cx<T> m1 = a[0] + scal(t1, c1) + scal(t2, c2);
    cx<T> m2 = a[0] + scal(t2, c1) + scal(t1, c2);
    cx<T> v1 = scal(u1, s1) + scal(u2, s2);
    cx<T> v2 = scal(u1, s2) - scal(u2, s1);
    cx<T> iv1 = cx<T>{-v1.i, v1.r}, iv2 = cx<T>{-v2.i, v2.r};
    a[0] = a[0] + t1 + t2;
    a[1] = m1 + iv1;
    a[4] = m1 - iv1;
    a[2] = m2 + iv2;
    a[3] = m2 - iv2;
}

// Radix 7: direct sum over the symmetric pairs (a_m + a_{7-m}, a_m - a_{7-m}).
template <typename T, int DIR> __device__ __forceinline__ void dft7(cx<T>* a) {
    const T c[4] = {T(1), T(0.62348980185873353052500488400424), T(-0.22252093395631440428890256449679),
                    T(-0.90096886790241912623610231950745)};                       // cos(2 pi q / 7)
    const T s[4] = {T(0), T(DIR) * T(0.78183148246802980870844452667406), T(DIR) * T(0.97492791218182360701813168299393),
                    T(DIR) * T(0.43388373911755812047576833284836)};              // DIR sin(2 pi q / 7)
    cx<T> t[4], u[4];
#pragma unroll
    for (int m = 1; m <= 3; ++m) { t[m] = a[m] + a[7 - m]; u[m] = a[m] - a[7 - m]; }
    cx<T> y0 = a[0] + t[1] + t[2] + t[3];
#pragma unroll
    for (int j = 1; j <= 3; ++j) {
        cx<T> re = a[0], im{T(0), T(0)};
#pragma unroll
        for (int m = 1; m <= 3; ++m) {
            const int q = (j * m) % 7;
            const T cq = q <= 3 ? c[q] : c[7 - q];
            const T sq = q <= 3 ? s[q] : -s[7 - q];
            re = re + scal(t[m], cq);
            im = im + scal(u[m], sq);
        }
        const cx<T> iim{-im.i, im.r};
        a[j] = re + iim;
        a[7 - j] = re - iim;
    }
    a[0] = y0;
}

template <typename T, int R, int DIR> __device__ __forceinline__ void dft(cx<T>* a) {
    if constexpr (R == 2) dft2<T, DIR>(a);
    else if constexpr (R == 3) dft3<T, DIR>(a);
    else if constexpr (R == 4) dft4<T, DIR>(a);
    else if constexpr (R == 5) dft5<T, DIR>(a);
    else if constexpr (R == 7) dft7<T, DIR>(a);
    else dft8<T, DIR>(a);
}

// ----------------------------------------------------------------------------------------------
// Tile geometry in LDS: element j of pencil m lives at j*SJ + m*SM (complex units).
//   strided tiles ("columns"): SJ = M, SM = 1, butterflies enumerated with m fastest;
//   row tiles: SJ = 1, SM = N + pad, butterflies enumerated with i fastest.
// ----------------------------------------------------------------------------------------------
struct TileGeom {
    int N;       // FFT length
    int M;       // pencils in the tile
    int SJ, SM;  // LDS strides
    int m_fast;  // 1: consecutive threads take consecutive pencils
    FDiv fM;     // division by M
};

// Per-radix-pass constants of one FFT length (host-computed): butterflies per pencil and stride product.
struct RadixPlan {
    int nrad;
    int rad[kMaxRadices];
    FDiv per[kMaxRadices];  // N / R
    FDiv p[kMaxRadices];    // product of the radices before this pass
};

// One Stockham pass of radix R over all pencils of the tile. Each thread owns at most
// Q = ceil(E/R) butterflies (the host sizes the block so that M*N <= E*blockDim).
template <typename T, int R, int DIR, int E>
__device__ __forceinline__ void stockham_pass(cx<T>* lds, const cx<T>* tw, const TileGeom& g, const FDiv& fper,
                                              const FDiv& fp, int tid, int nth) {
    constexpr int Q = (E + R - 1) / R;
    const int per = fper.d;         // butterflies per pencil
    const int p = fp.d;
    const int total = g.M * per;
    const int twstride = g.N / (p * R);
    cx<T> v[Q][R];
    int obase[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = tid + q * nth;
        obase[q] = -1;
        if (b < total) {
            int m, i;
            if (g.m_fast) { i = fdiv(b, g.fM); m = b - i * g.M; }
            else { m = fdiv(b, fper); i = b - m * per; }
            const int k = i - fdiv(i, fp) * p;
            const int base = m * g.SM;
#pragma unroll
            for (int j = 0; j < R; ++j) v[q][j] = lds[(i + j * per) * g.SJ + base];
            if (p > 1) {
#pragma unroll
                for (int j = 1; j < R; ++j) {
                    cx<T> t = tw[j * k * twstride];
                    if (DIR > 0) t.i = -t.i;
                    v[q][j] = v[q][j] * t;
                }
            }
            dft<T, R, DIR>(v[q]);
            obase[q] = ((i - k) * R + k) * g.SJ + base;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if (obase[q] >= 0) {
#pragma unroll
            for (int j = 0; j < R; ++j) lds[obase[q] + j * p * g.SJ] = v[q][j];
        }
    }
    __syncthreads();
}

// Full FFT of every pencil of the tile (natural order in, natural order out, unnormalised).
template <typename T, int DIR, int E>
__device__ void tile_fft(cx<T>* lds, const cx<T>* tw, const TileGeom& g, const RadixPlan& rp, int tid, int nth) {
    for (int r = 0; r < rp.nrad; ++r) {
        switch (rp.rad[r]) {
            case 8: stockham_pass<T, 8, DIR, E>(lds, tw, g, rp.per[r], rp.p[r], tid, nth); break;
            case 4: stockham_pass<T, 4, DIR, E>(lds, tw, g, rp.per[r], rp.p[r], tid, nth); break;
            case 2: stockham_pass<T, 2, DIR, E>(lds, tw, g, rp.per[r], rp.p[r], tid, nth); break;
            case 3: stockham_pass<T, 3, DIR, E>(lds, tw, g, rp.per[r], rp.p[r], tid, nth); break;
            case 5: stockham_pass<T, 5, DIR, E>(lds, tw, g, rp.per[r], rp.p[r], tid, nth); break;
            case 7: stockham_pass<T, 7, DIR, E>(lds, tw, g, rp.per[r], rp.p[r], tid, nth); break;
            default: break;  // rejected by the host plan
        }
    }
}

// Wavefront-shuffle sum of a double across the block; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* red /* >= 16 */) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) s += red[w];
    }
    __syncthreads();
    return s;
}

}  // namespace mmx

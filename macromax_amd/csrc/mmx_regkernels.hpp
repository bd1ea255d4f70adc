// Pencil passes for power-of-two axis lengths with the register-resident FFT (mmx_regfft.hpp).
//
// Same pass kinds and semantics as pencil_pass_kernel (mmx_kernels.hpp); the tile's data lives in registers:
// thread (pencil p, position t) holds points t + m*TP (m < E) of each of its NC components.
//   row tiles:     lanes run along a row (t fastest) -> each row is read and written contiguously;
//   strided tiles: lanes run across the P columns (p fastest) -> P*8-byte segments per row of the tile.
#pragma once
#include "mmx_kernels.hpp"
#include "mmx_regfft.hpp"

namespace mmx {

template <typename T, int N, int E, int KIND, int NC>
__global__ void __launch_bounds__(512) reg_pass_kernel(PassArgs<T> a) {
    constexpr bool PRO = (KIND == kProFwd || KIND == kAll);
    constexpr bool GREEN = (KIND == kGreen || KIND == kAll);
    constexpr bool EPI = (KIND == kInvEpi || KIND == kAll);
    constexpr int TP = N / E;
    constexpr int LTP = ilog2(TP);
    extern __shared__ unsigned char smem_raw[];
    __shared__ double red[16];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* xbuf = tw + N;
    const int tid = threadIdx.x;
    const bool rows = (a.inner == 1);
    const int P = a.P;
    const int LP = 31 - __clz(P);  // P is a power of two

    int64_t blk = blockIdx.x;
    if (a.xcd_remap) {
        const int64_t n = gridDim.x, q = n >> 3, rem = n & 7, r = blk & 7;
        blk = r * q + (r < rem ? r : rem) + (blk >> 3);
    }
    const int64_t b = fdiv64(blk, a.ftpp);
    const int64_t tl = blk - b * a.tiles_per_problem;
    const Coef<T> cf = (a.coef != nullptr) ? a.coef[b] : Coef<T>{T(0), T(1), T(1), T(0), T(0), T(0)};

    int p, t;
    cx<T>* xb;
    int ps;
    if (rows) { t = tid & (TP - 1); p = tid >> LTP; xb = xbuf + p * padded_len(N); ps = 1; }
    else { p = tid & (P - 1); t = tid >> LP; xb = xbuf + p; ps = P; }

    int64_t r0 = 0, o = 0, ib = 0;
    if (rows) r0 = tl * P;
    else { o = fdiv64(tl, a.fntile); ib = (tl - o * a.ntile_inner) * P; }
    const bool valid = rows ? (r0 + p < a.nouter) : (ib + p < a.inner);
    // offset of point j of this thread's pencil within one component of one problem
    const int64_t pbase = rows ? (r0 + p) * N : o * (int64_t)N * a.inner + ib + p;
    const int64_t jstride = rows ? 1 : a.inner;
    const int64_t pb = b * a.bstride;

    for (int q = tid; q < N; q += blockDim.x) tw[q] = a.tw[q];
    __syncthreads();

    cx<T> v[NC][E];
    const cx<T> zero{T(0), T(0)};
    if constexpr (PRO) {
        const cx<T>* Eb = a.E + b * a.E_bstride;
        const cx<T>* sb = a.s + b * a.s_bstride;
        const cx<T>* chib = a.chi + b * a.chi_bstride;
        const int64_t V = a.comp_stride;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            cx<T> x[3] = {zero, zero, zero}, y[3] = {zero, zero, zero};
            if (valid) {
                const int64_t vox = pbase + (int64_t)(t + m * TP) * jstride;
#pragma unroll
                for (int c = 0; c < NC; ++c) x[c] = Eb[c * V + vox];
                chi_apply<T>(a, chib, vox, x, y, cf);
#pragma unroll
                for (int c = 0; c < NC; ++c) y[c] = y[c] + scal(sb[c * V + vox], cf.a);
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c][m] = y[c];
        }
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int64_t off = pbase + (int64_t)(t + m * TP) * jstride;
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c][m] = valid ? a.src[pb + c * a.comp_stride + off] : zero;
        }
    }

    if constexpr (GREEN) {
        reg_fft<T, N, E, -1, NC>(v, t, tw, xb, ps);
        // k components off the pass axis are fixed for this thread's pencil
        int64_t col = rows ? 0 : ib + p;
        int64_t row = rows ? r0 + p : o;
        int64_t i0 = 0, i1 = 0, i2 = 0;
        if (a.axis_orig == 0) { i1 = fdiv64(col, a.fn2); i2 = col - i1 * a.n2; }
        else if (a.axis_orig == 1) { i0 = row; i2 = col; }
        else { i0 = fdiv64(row, a.fn1); i1 = row - i0 * a.n1; }
        T kf[3];
        kf[0] = (a.axis_orig != 0) ? a.kvec[i0] : T(0);
        kf[1] = (a.axis_orig != 1 && a.n1 > 1) ? a.kvec[a.k1ofs + i1] : T(0);
        kf[2] = (a.axis_orig != 2 && a.n2 > 1) ? a.kvec[a.k2ofs + i2] : T(0);
        const bool off_dc = (i0 | i1 | i2) != 0;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int j = t + m * TP;
            const T kj = a.kvec[a.kofs_axis + j];
            const T k0 = a.axis_orig == 0 ? kj : kf[0];
            const T k1 = a.axis_orig == 1 ? kj : kf[1];
            const T k2 = a.axis_orig == 2 ? kj : kf[2];
            cx<T> f[3] = {zero, zero, zero};
#pragma unroll
            for (int c = 0; c < NC; ++c) f[c] = v[c][m];
            green_apply<T>(f, NC, k0, k1, k2, (j == 0) && !off_dc, a, cf);
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c][m] = f[c];
        }
        reg_fft<T, N, E, +1, NC>(v, t, tw, xb, ps);
    } else if constexpr (KIND == kFwd || KIND == kProFwd) {
        reg_fft<T, N, E, -1, NC>(v, t, tw, xb, ps);
    } else {
        reg_fft<T, N, E, +1, NC>(v, t, tw, xb, ps);
    }

    if constexpr (EPI) {
        const cx<T>* Eb = a.E + b * a.E_bstride;
        const cx<T>* chib = a.chi + b * a.chi_bstride;
        cx<T>* Eo = a.Eout + b * a.E_bstride;
        const int64_t V = a.comp_stride;
        double dd = 0.0, ee = 0.0;
        if (valid) {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int64_t vox = pbase + (int64_t)(t + m * TP) * jstride;
                cx<T> e[3] = {zero, zero, zero}, w[3] = {zero, zero, zero}, d[3] = {zero, zero, zero};
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    e[c] = Eb[c * V + vox];
                    w[c] = v[c][m] - e[c];
                }
                chi_apply<T>(a, chib, vox, w, d, cf);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const cx<T> en = e[c] + d[c];
                    Eo[c * V + vox] = en;
                    dd += norm2_f64(d[c]);
                    ee += norm2_f64(en);
                }
            }
        }
        dd = block_sum(dd, red);
        ee = block_sum(ee, red);
        if (tid == 0) {
            a.partials[2 * blk + 0] = dd;
            a.partials[2 * blk + 1] = ee;
        }
    } else if (valid) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int64_t off = pbase + (int64_t)(t + m * TP) * jstride;
#pragma unroll
            for (int c = 0; c < NC; ++c) a.dst[pb + c * a.comp_stride + off] = scal(v[c][m], a.scale);
        }
    }
}

}  // namespace mmx

// Fused pencil passes of the magnetic iteration (complex64, register-resident FFT lengths 64-1024 on the first and
// last grid axes). The iteration of mmx_magnetic.hpp,
//
//     E^ = F E, curl E;  [Z1; Y1] = [chi_EE E + chi_HE curl E + s; chi_EH E + chi_HH curl E];
//     B^ = G (Z1^ + i k x Y1^) - E^;  [Z2; Y2] = [chi_EE B + chi_HE curl B; chi_EH B + chi_HH curl B];
//     d = Z2 + curl Y2,  E' = E + d,
//
// is evaluated with the k-space products inside the pass along the first axis ("mid" passes: forward FFT, product,
// inverse FFT) and the real-space products inside the pass along the last, contiguous axis ("pw" passes: inverse FFT
// of the transform that ends there, the voxel-wise products, forward FFT of the transform that starts there). With
// plain passes along the middle axis of 3D grids in between, one iteration is
//
//     fwd C (E) | [B fwd] | mid 1: fwd A, E^ -> w1, i k x, inv A | [B inv] | pw 1: inv C, [Z1; Y1], fwd C |
//     [B fwd] | mid 2: fwd A, B^, i k x B^, inv A | [B inv] | pw 2: inv C, Z2 -> w1, Y2, fwd C | [B fwd] |
//     mid 3: fwd A, i k x, inv A | [B inv] | final: inv C, d, E', norms
//
// (A the first axis, C the last), eight kernels on a 2D grid instead of nineteen. The inverse transforms' 1/N is
// applied by the pw passes.
#pragma once
#include "mmx_magnetic.hpp"
#include "mmx_regkernels.hpp"

namespace mmx {

// Pencil geometry of a register-path tile (as in reg_pass_kernel): pencil p, point t + m TP of it at byte offset
// toff + m mstep from the tile's first element tb; `valid` false for pencils past the edge of the last tile.
struct MagTile {
    int64_t tb;
    int64_t row, col;  // outer index (rows tiles: the row; strided: the block before the axis) and column
    uint32_t toff, mstep;
    bool valid;
};

template <typename T, int N, int E>
__device__ __forceinline__ MagTile mag_tile(const PassArgs<T>& g, int p, int t) {
    constexpr int TP = N / E;
    constexpr uint32_t ES = sizeof(cx<T>);
    int64_t blk = blockIdx.x;
    if (g.xcd_remap) {
        const int64_t n = gridDim.x, q = n >> 3, rem = n & 7, r = blk & 7;
        blk = r * q + (r < rem ? r : rem) + (blk >> 3);
    }
    MagTile m;
    if (g.inner == 1) {
        const int64_t r0 = blk * g.P;
        m.valid = r0 + p < g.nouter;
        m.row = r0 + p;
        m.col = 0;
        m.tb = r0 * N;
        m.toff = (uint32_t)(p * N + t) * ES;
        m.mstep = (uint32_t)TP * ES;
    } else {
        const int64_t o = fdiv64(blk, g.fntile);
        const int64_t ib = (blk - o * g.ntile_inner) * g.P;
        m.valid = ib + p < g.inner;
        m.row = o;
        m.col = ib + p;
        m.tb = o * (int64_t)N * g.inner + ib;
        m.toff = ((uint32_t)p + (uint32_t)t * (uint32_t)g.inner) * ES;
        m.mstep = (uint32_t)TP * (uint32_t)g.inner * ES;
    }
    return m;
}

template <typename T, int N, int E, int NC>
__device__ __forceinline__ void mag_load(cx<T> (&v)[NC][E], const cx<T>* base, const MagTile& tl, int64_t V) {
    const cx<T> zero{T(0), T(0)};
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c][m] = zero;
    if (!tl.valid) return;
    const cx<T>* b = base + tl.tb;
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c][m] = ldg<true>(b + c * V, tl.toff + m * tl.mstep);
}

template <typename T, int N, int E, int NC>
__device__ __forceinline__ void mag_store(const cx<T> (&v)[NC][E], cx<T>* base, const MagTile& tl, int64_t V) {
    if (!tl.valid) return;
    cx<T>* b = base + tl.tb;
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
        for (int c = 0; c < NC; ++c) stg<true>(b + c * V, tl.toff + m * tl.mstep, v[c][m]);
}

// Mid passes along the first grid axis (g.axis_orig == 0): forward FFT of NC = 3 components of `src`, a k-space
// product, inverse FFT, result to `dst` (unscaled).
//   OP 0: E^ to w1; i k x E^
//   OP 2: i k x
// (OP 1, [Z1^; Y1^] -> B^ = G (Z1^ + i k x Y1^) - E^ and [B^; i k x B^], runs in mag_mid_split_kernel.)
template <typename T, int N, int E, int NC, int OP>
__global__ void __launch_bounds__(512) mag_mid_kernel(PassArgs<T> g, MagArgs<T> a, const cx<T>* src, cx<T>* dst) {
    constexpr int TP = N / E;
    constexpr int LTP = ilog2(TP);
    extern __shared__ unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* xbuf = tw + N;
    const int tid = threadIdx.x;
    const bool rows = g.inner == 1;
    const int P = g.P;
    const int LP = 31 - __clz(P);
    int p, t, ps;
    cx<T>* xb;
    if (rows) { t = tid & (TP - 1); p = tid >> LTP; xb = xbuf + p * padded_len(N, E); ps = 1; }
    else { p = tid & (P - 1); t = tid >> LP; xb = xbuf + p; ps = P; }
    const MagTile tl = mag_tile<T, N, E>(g, p, t);
    const int64_t V = a.V;
    load_twiddles<T, N, E>(tw, g.tw, tid, blockDim.x);
    cx<T> v[NC][E];
    mag_load<T, N, E, NC>(v, src, tl, V);
    // k off the pass axis is fixed for the pencil; along it, k of point t + m TP (pencils past the edge: index 0)
    const int64_t col = tl.valid ? tl.col : 0;
    const int64_t i1 = fdiv64(col, g.fn2), i2 = col - i1 * g.n2;
    const T k1 = g.n1 > 1 ? g.kvec[g.k1ofs + i1] : T(0);
    const T k2 = g.n2 > 1 ? g.kvec[g.k2ofs + i2] : T(0);
    const bool dc_pencil = (i1 | i2) == 0;
    T kpre[E];
#pragma unroll
    for (int m = 0; m < E; ++m) kpre[m] = g.kvec[t + m * TP];
    static_assert(NC == 3 && (OP == 0 || OP == 2), "OP 1 runs in mag_mid_split_kernel");
    __syncthreads();
    reg_fft<T, N, E, -1, NC>(v, t, tw, xb, ps);
    const cx<T> zero{T(0), T(0)};
    if constexpr (OP == 0) {
        mag_store<T, N, E, NC>(v, a.w1, tl, V);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const cx<T> f[3] = {v[0][m], v[1][m], v[2][m]};
        cx<T> o[3];
        cross_ik(kpre[m], k1, k2, f, o);
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c][m] = o[c];
    }
    reg_fft<T, N, E, +1, NC>(v, t, tw, xb, ps);
    mag_store<T, N, E, NC>(v, dst, tl, V);
}

// Mid pass OP 1 with its six components in two thread groups of P pencils each (blockDim = 2 P TP): group 0
// transforms Z1 and writes B, group 1 transforms Y1 and writes curl B, each three components per thread (half the
// serial work of the six-component kernel, whose passes at 0.26 M voxels are bound by one workgroup's chain). The
// coupling goes through LDS after the forward FFTs: group 1 leaves i k x Y1^, group 0 forms
// B^ = G (Z1^ + i k x Y1^) - E^ in place of it, group 1 takes B^ for i k x B^. Dynamic LDS: twiddles, one exchange
// buffer per group, and the coupling buffer of 3 P N values ([component][pencil][point]).
template <typename T, int N, int E>
__global__ void __launch_bounds__(512) mag_mid_split_kernel(PassArgs<T> g, MagArgs<T> a, const cx<T>* src,
                                                            cx<T>* dst) {
    constexpr int TP = N / E;
    constexpr int LTP = ilog2(TP);
    extern __shared__ unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    const int P = g.P;
    const int LP = 31 - __clz(P);
    const int gsz = P * TP;
    const int grp = (int)threadIdx.x >= gsz ? 1 : 0;
    const int gtid = (int)threadIdx.x - grp * gsz;
    const int plen = P * padded_len(N, E);
    cx<T>* xbuf = tw + N + grp * plen;
    cx<T>* xfer = tw + N + 2 * plen;
    const bool rows = g.inner == 1;
    int p, t, ps;
    cx<T>* xb;
    if (rows) { t = gtid & (TP - 1); p = gtid >> LTP; xb = xbuf + p * padded_len(N, E); ps = 1; }
    else { p = gtid & (P - 1); t = gtid >> LP; xb = xbuf + p; ps = P; }
    const MagTile tl = mag_tile<T, N, E>(g, p, t);
    const int64_t V = a.V;
    load_twiddles<T, N, E>(tw, g.tw, threadIdx.x, blockDim.x);
    cx<T> v[3][E];
    mag_load<T, N, E, 3>(v, src + grp * 3 * V, tl, V);
    const int64_t col = tl.valid ? tl.col : 0;
    const int64_t i1 = fdiv64(col, g.fn2), i2 = col - i1 * g.n2;
    const T k1 = g.n1 > 1 ? g.kvec[g.k1ofs + i1] : T(0);
    const T k2 = g.n2 > 1 ? g.kvec[g.k2ofs + i2] : T(0);
    const bool dc_pencil = (i1 | i2) == 0;
    T kpre[E];
#pragma unroll
    for (int m = 0; m < E; ++m) kpre[m] = g.kvec[t + m * TP];
    cx<T> eh[3][E];  // E^ (group 0), requested before the forward FFT so that it arrives during it
    if (grp == 0) mag_load<T, N, E, 3>(eh, a.w1, tl, V);
    __syncthreads();
    reg_fft<T, N, E, -1, 3>(v, t, tw, xb, ps);
    cx<T>* xf = xfer + p * N + t;  // component c, point t + m TP of this pencil: xf[c P N + m TP]
    if (grp == 1) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const cx<T> y[3] = {v[0][m], v[1][m], v[2][m]};
            cx<T> o[3];
            cross_ik(kpre[m], k1, k2, y, o);
#pragma unroll
            for (int c = 0; c < 3; ++c) xf[c * P * N + m * TP] = o[c];
        }
    }
    __syncthreads();
    if (grp == 0) {
        const T ar = a.alpha_r, inv_ai = T(1) / a.alpha_i;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const T k0 = kpre[m];
            cx<T> f[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) f[c] = v[c][m] + xf[c * P * N + m * TP];
            const T kk = k0 * k0 + k1 * k1 + k2 * k2;
            const cx<T> gs = green_scalar(kk, ar, inv_ai, T(1));  // -i alpha_i / (k^2 - alpha)
            const cx<T> gl = green_longitudinal(ar, inv_ai, T(1));  // i alpha_i / alpha
            cx<T> b[3];
            if (dc_pencil && t + m * TP == 0) {  // k = 0: the first nd components count as longitudinal
#pragma unroll
                for (int c = 0; c < 3; ++c) b[c] = (c < a.nd ? gl : gs) * f[c];
            } else {
                const cx<T> proj = scal(f[0], k0) + scal(f[1], k1) + scal(f[2], k2);
                const cx<T> q = (gl - gs) * scal(proj, T(1) / kk);
                b[0] = gs * f[0] + scal(q, k0);
                b[1] = gs * f[1] + scal(q, k1);
                b[2] = gs * f[2] + scal(q, k2);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                b[c] = b[c] - eh[c][m];
                v[c][m] = b[c];
                xf[c * P * N + m * TP] = b[c];
            }
        }
    }
    __syncthreads();
    if (grp == 1) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const cx<T> b[3] = {xf[m * TP], xf[P * N + m * TP], xf[2 * P * N + m * TP]};
            cx<T> cb[3];
            cross_ik(kpre[m], k1, k2, b, cb);
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c][m] = cb[c];
        }
    }
    reg_fft<T, N, E, +1, 3>(v, t, tw, xb, ps);
    mag_store<T, N, E, 3>(v, dst + grp * 3 * V, tl, V);
}

// pw passes along the last (contiguous) grid axis: inverse FFT of the components of `src` (the last inverse axis:
// scaled by 1/V), voxel-wise products, then
//   STAGE 1 (curl E): [Z1; Y1] = [chi_EE E + chi_HE curl E + s; chi_EH E + chi_HH curl E], forward FFT, to dst
//                     (6 components; mag_pw1_split_kernel);
//   STAGE 2 ([B; curl B]): Z2 = chi_EE B + chi_HE curl B to w1 (real space), Y2 = chi_EH B + chi_HH curl B, forward
//                     FFT, to dst (3 components; mag_pw2_split_kernel);
//   STAGE 3 (curl Y2): d = Z2 + curl Y2, E' = E + d to Eout, FP64 partials of |d|^2, |E'|^2 per tile (mag_pw_kernel).
//
// pw pass STAGE 1 in two thread groups of P rows each (blockDim = 2 P TP): both take the inverse FFT of curl E^
// (three components), then group 0 forms Z1 = chi_EE E + chi_HE curl E + s and group 1 Y1 = chi_EH E + chi_HH curl E,
// and each transforms and writes its three components (dst + group 3 V). Dynamic LDS: twiddles and one exchange
// buffer per group.
template <typename T, int N, int E>
__global__ void __launch_bounds__(512) mag_pw1_split_kernel(PassArgs<T> g, MagArgs<T> a, const cx<T>* src,
                                                            cx<T>* dst) {
    constexpr int TP = N / E;
    constexpr int LTP = ilog2(TP);
    extern __shared__ unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    const int gsz = g.P * TP;
    const int grp = (int)threadIdx.x >= gsz ? 1 : 0;
    const int gtid = (int)threadIdx.x - grp * gsz;
    const int t = gtid & (TP - 1);
    const int p = gtid >> LTP;
    cx<T>* xb = tw + N + grp * g.P * padded_len(N, E) + p * padded_len(N, E);
    const MagTile tl = mag_tile<T, N, E>(g, p, t);
    const int64_t V = a.V;
    const cx<T> zero{T(0), T(0)};
    load_twiddles<T, N, E>(tw, g.tw, threadIdx.x, blockDim.x);
    cx<T> x[3][E];
    mag_load<T, N, E, 3>(x, src, tl, V);
    // E, and s for group 0, requested before the inverse FFT so that they arrive during it
    cx<T> e[3][E], sv[3][E];
    mag_load<T, N, E, 3>(e, a.E, tl, V);
    if (grp == 0) mag_load<T, N, E, 3>(sv, a.s, tl, V);
    else {
#pragma unroll
        for (int m = 0; m < E; ++m)
#pragma unroll
            for (int c = 0; c < 3; ++c) sv[c][m] = zero;
    }
    __syncthreads();
    reg_fft<T, N, E, +1, 3>(x, t, tw, xb, 1);
    const T sc = g.scale;
    const MagChi<T> ce = grp == 0 ? a.ee : a.eh;  // on E
    const MagChi<T> cc = grp == 0 ? a.he : a.hh;  // on curl E
    cx<T> y[3][E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const cx<T> ev[3] = {e[0][m], e[1][m], e[2][m]};
        cx<T> acc[3] = {sv[0][m], sv[1][m], sv[2][m]};
        const cx<T> c[3] = {scal(x[0][m], sc), scal(x[1][m], sc), scal(x[2][m], sc)};
        if (tl.valid) {
            const int64_t vox = tl.tb + tl.toff / sizeof(cx<T>) + (int64_t)m * TP;
            mag_mul_add(ce, vox, V, ev, acc);
            mag_mul_add(cc, vox, V, c, acc);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) y[i][m] = acc[i];
    }
    reg_fft<T, N, E, -1, 3>(y, t, tw, xb, 1);
    mag_store<T, N, E, 3>(y, dst + grp * 3 * V, tl, V);
}

// pw pass STAGE 2 in two thread groups of P rows each: group 0 takes the inverse FFT of B^, group 1 that of
// i k x B^; both leave their three components in LDS and read the other group's; group 0 writes
// Z2 = chi_EE B + chi_HE curl B to w1, group 1 transforms Y2 = chi_EH B + chi_HH curl B and writes it to dst. Group 0
// runs the same forward transform on its Z2 (result unused), so that both groups pass the same barriers. Dynamic
// LDS: twiddles, one exchange buffer per group, and 6 P N values ([component][row][point]).
template <typename T, int N, int E>
__global__ void __launch_bounds__(512) mag_pw2_split_kernel(PassArgs<T> g, MagArgs<T> a, const cx<T>* src,
                                                            cx<T>* dst) {
    constexpr int TP = N / E;
    constexpr int LTP = ilog2(TP);
    extern __shared__ unsigned char smem_raw[];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    const int P = g.P;
    const int gsz = P * TP;
    const int grp = (int)threadIdx.x >= gsz ? 1 : 0;
    const int gtid = (int)threadIdx.x - grp * gsz;
    const int t = gtid & (TP - 1);
    const int p = gtid >> LTP;
    cx<T>* xb = tw + N + grp * P * padded_len(N, E) + p * padded_len(N, E);
    cx<T>* xf = tw + N + 2 * P * padded_len(N, E) + p * N + t;  // component c, point t + m TP: xf[c P N + m TP]
    const MagTile tl = mag_tile<T, N, E>(g, p, t);
    const int64_t V = a.V;
    const cx<T> zero{T(0), T(0)};
    load_twiddles<T, N, E>(tw, g.tw, threadIdx.x, blockDim.x);
    cx<T> x[3][E];
    mag_load<T, N, E, 3>(x, src + grp * 3 * V, tl, V);
    __syncthreads();
    reg_fft<T, N, E, +1, 3>(x, t, tw, xb, 1);
    const T sc = g.scale;
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
        for (int c = 0; c < 3; ++c) xf[(grp * 3 + c) * P * N + m * TP] = scal(x[c][m], sc);
    __syncthreads();
    const MagChi<T> cb_chi = grp == 0 ? a.ee : a.eh;  // on B
    const MagChi<T> cc_chi = grp == 0 ? a.he : a.hh;  // on curl B
    cx<T> y[3][E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const cx<T> b[3] = {xf[m * TP], xf[P * N + m * TP], xf[2 * P * N + m * TP]};
        const cx<T> cb[3] = {xf[3 * P * N + m * TP], xf[4 * P * N + m * TP], xf[5 * P * N + m * TP]};
        cx<T> acc[3] = {zero, zero, zero};
        if (tl.valid) {
            const int64_t vox = tl.tb + tl.toff / sizeof(cx<T>) + (int64_t)m * TP;
            mag_mul_add(cb_chi, vox, V, b, acc);
            mag_mul_add(cc_chi, vox, V, cb, acc);
            if (grp == 0) {
                cx<T>* zt = a.w1 + tl.tb;
#pragma unroll
                for (int i = 0; i < 3; ++i) stg<true>(zt + i * V, tl.toff + m * tl.mstep, acc[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) y[i][m] = acc[i];
    }
    reg_fft<T, N, E, -1, 3>(y, t, tw, xb, 1);
    if (grp == 1) mag_store<T, N, E, 3>(y, dst, tl, V);
}

template <typename T, int N, int E, int STAGE>
__global__ void __launch_bounds__(512) mag_pw_kernel(PassArgs<T> g, MagArgs<T> a, const cx<T>* src, cx<T>* dst) {
    static_assert(STAGE == 3, "STAGE 1 and 2 run in the two-group pw kernels");
    constexpr int TP = N / E;
    constexpr int LTP = ilog2(TP);
    extern __shared__ unsigned char smem_raw[];
    __shared__ double red[16];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* xbuf = tw + N;
    const int tid = threadIdx.x;
    const int t = tid & (TP - 1);
    const int p = tid >> LTP;
    cx<T>* xb = xbuf + p * padded_len(N, E);
    const MagTile tl = mag_tile<T, N, E>(g, p, t);
    const int64_t V = a.V;
    load_twiddles<T, N, E>(tw, g.tw, tid, blockDim.x);
    cx<T> x[3][E];
    mag_load<T, N, E, 3>(x, src, tl, V);
    // Z2 and E, requested before the inverse FFT so that they arrive during it
    cx<T> f1[3][E], f2[3][E];
    mag_load<T, N, E, 3>(f1, a.w1, tl, V);
    mag_load<T, N, E, 3>(f2, a.E, tl, V);
    __syncthreads();
    reg_fft<T, N, E, +1, 3>(x, t, tw, xb, 1);
    const T sc = g.scale;
    {
        double dd = 0.0, ee = 0.0;
        if (tl.valid) {
            cx<T>* Eo = a.Eout + tl.tb;
#pragma unroll
            for (int m = 0; m < E; ++m) {
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const uint32_t bo = tl.toff + m * tl.mstep;
                    const cx<T> d = f1[i][m] + scal(x[i][m], sc);
                    const cx<T> en = f2[i][m] + d;
                    stg<true>(Eo + i * V, bo, en);
                    dd += norm2_f64(d);
                    ee += norm2_f64(en);
                }
            }
        }
        dd = block_sum(dd, red);
        ee = block_sum(ee, red);
        if (threadIdx.x == 0) {
            a.partials[2 * blockIdx.x] = dd;
            a.partials[2 * blockIdx.x + 1] = ee;
        }
    }
}

}  // namespace mmx

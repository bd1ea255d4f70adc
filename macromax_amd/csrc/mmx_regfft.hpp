// Register-resident Stockham FFT for axis lengths q * 2^k with q = 1, 3 or 5 (compile-time N).
//
// A pencil of N points is owned by TP = N/E threads; thread t holds the E points at positions t + m*TP
// (m < E) in registers. Every stage of radix R (R divides E) runs E/R butterflies per thread on those registers;
// butterfly q of thread t is Stockham butterfly i = t + q*TP with inputs at i + j*N/R, which are exactly the
// registers m = q + j*E/R. Between stages the outputs go through LDS once (write at their Stockham output
// position, read back the new stride-TP set); after the last stage register m holds output t + m*TP, so the
// first stage reads and the last stage writes HBM directly, with no LDS staging.
#pragma once
#include "mmx_device.hpp"

namespace mmx {

template <typename T, int DIR> __device__ __forceinline__ void dft16(cx<T>* a) {
    // 16 = 4 x 4: DFT4 over n1 (stride 4), twiddle W16^(n2 k1), DFT4 over n2; output k = k1 + 4 k2
    cx<T> b[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cx<T> c[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
        dft4<T, DIR>(c);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) b[4 * k1 + n2] = c[k1];
    }
    const T h = T(0.70710678118654752440084436210485);
    const T c1 = T(0.92387953251128675612818318939679), s1 = T(0.38268343236508977172845998403040);
    // b[4*k1 + n2] *= W16^(n2*k1), W16^m = cos(2 pi m/16) + DIR i sin(2 pi m/16)
    auto tw = [&](cx<T> x, T cr, T si) { return cx<T>{x.r * cr - T(DIR) * x.i * si, x.i * cr + T(DIR) * x.r * si}; };
    b[5] = tw(b[5], c1, s1);      // m = 1
    b[6] = tw(b[6], h, h);        // m = 2
    b[7] = tw(b[7], s1, c1);      // m = 3
    b[9] = tw(b[9], h, h);        // m = 2
    b[10] = mul_dir_i<T, DIR>(b[10]);  // m = 4
    b[11] = tw(b[11], -h, h);     // m = 6
    b[13] = tw(b[13], s1, c1);    // m = 3
    b[14] = tw(b[14], -h, h);     // m = 6
    b[15] = tw(b[15], -c1, -s1);  // m = 9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cx<T> c[4] = {b[4 * k1], b[4 * k1 + 1], b[4 * k1 + 2], b[4 * k1 + 3]};
        dft4<T, DIR>(c);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = c[k2];
    }
}

template <typename T, int R, int DIR> __device__ __forceinline__ void dft_reg(cx<T>* a) {
    if constexpr (R == 2) dft2<T, DIR>(a);
    else if constexpr (R == 3) dft3<T, DIR>(a);
    else if constexpr (R == 4) dft4<T, DIR>(a);
    else if constexpr (R == 5) dft5<T, DIR>(a);
    else if constexpr (R == 8) dft8<T, DIR>(a);
    else dft16<T, DIR>(a);
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }
constexpr int pow2_part(int x) { return x & -x; }

// Radix of stage s for N = q * 2^n points with an odd factor q (3 or 5 here) and E registers per thread (E a
// multiple of q with an even power-of-two part): the power-of-two stages come first, each the largest power of two
// <= min(pow2 part of E, 16) that still divides what is left of N's power-of-two part; q is the last stage. The
// product of the radices before every stage is then a power of two, so a butterfly's index within its group is
// i & (P - 1).
constexpr int stage_radix(int N, int E, int s) {
    const int cap = pow2_part(E) < 16 ? pow2_part(E) : 16;
    int rem = pow2_part(N), r = 1;
    for (int q = 0; q <= s; ++q) {
        if (rem == 1) return N / pow2_part(N);
        r = cap;
        while (r > rem) r >>= 1;
        rem /= r;
    }
    return r;
}
constexpr int stage_prod(int N, int E, int s) {  // product of the radices before stage s
    int p = 1;
    for (int q = 0; q < s; ++q) p *= stage_radix(N, E, q);
    return p;
}
constexpr int num_stages(int N, int E) {
    int s = 0, p = 1;
    while (p < N) { p *= stage_radix(N, E, s); ++s; }
    return s;
}
// Twiddle table in LDS, laid out per stage: stage s >= 1 (radix R, P = product of earlier radices) keeps
// w^(j k N/(P R)) for j = 1..R-1, k = 0..P-1 at tw_stage_off(s) + (j-1) P + k, so that the lanes of a wave, which
// differ in k, read consecutive entries. The stages together hold at most N - 1 entries.
constexpr int tw_stage_off(int N, int E, int s) {
    int off = 0;
    for (int q = 1; q < s; ++q) off += (stage_radix(N, E, q) - 1) * stage_prod(N, E, q);
    return off;
}

// Exponent (in units of 2 pi / N) of staged-table entry e, for the stages from S on.
template <int N, int E, int S>
__device__ __forceinline__ int tw_stage_exponent(int e) {
    if constexpr (S >= num_stages(N, E)) {
        return 0;
    } else {
        constexpr int P = stage_prod(N, E, S), R = stage_radix(N, E, S), off = tw_stage_off(N, E, S);
        if (e < off + (R - 1) * P) return (1 + (e - off) / P) * ((e - off) & (P - 1)) * (N / (P * R));
        return tw_stage_exponent<N, E, S + 1>(e);
    }
}

// Fill the LDS twiddle table (N entries) from the natural-order table g[i] = exp(-2 pi i I / N) in global memory.
template <typename T, int N, int E>
__device__ __forceinline__ void load_twiddles(cx<T>* tw, const cx<T>* g, int tid, int nth) {
    for (int e = tid; e < N - 1; e += nth) tw[e] = g[tw_stage_exponent<N, E, 1>(e)];
}

// LDS position of exchange entry `pos`: one pad slot every 2^pad_shift(E) entries, which breaks the power-of-two
// strides of the exchange. LDS has 64 four-byte banks; an 8-byte read serves 32 lanes per cycle and an 8-byte write
// 16. With E = 8 (three radix-8/4 stages) a slot every 8 entries makes the strided-tile exchange and the rows-tile
// writes conflict-free (rows-tile reads then pair up 3 of 32 lanes); with E = 16 a slot every 16 is conflict-free.
template <int E> __device__ __forceinline__ int padpos(int pos) { return pos + (pos >> pad_shift(E)); }


template <int N, int E> constexpr bool reg_schedule_ok() {
    for (int q = 0; q < num_stages(N, E); ++q) {
        const int R = stage_radix(N, E, q), P = stage_prod(N, E, q);
        if (E % R != 0 || (P & (P - 1)) != 0) return false;
    }
    return N % E == 0 && pow2_part(N / E) == N / E;
}

// One stage on the registers v[NC][E] of all NC components.
template <typename T, int N, int E, int S, int DIR, int NC>
__device__ __forceinline__ void reg_stage(cx<T> (&v)[NC][E], int t, const cx<T>* tw) {
    constexpr int R = stage_radix(N, E, S);
    constexpr int P = stage_prod(N, E, S);
    constexpr int TP = N / E;
    constexpr int BQ = E / R;
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
        const int i = t + q * TP;
        const int k = i & (P - 1);
        cx<T> w[R];
        if constexpr (P > 1) {
#pragma unroll
            for (int j = 1; j < R; ++j) {
                w[j] = tw[tw_stage_off(N, E, S) + (j - 1) * P + k];
                if (DIR > 0) w[j].i = -w[j].i;
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            cx<T> a[R];
#pragma unroll
            for (int j = 0; j < R; ++j) a[j] = v[c][q + j * BQ];
            if constexpr (P > 1) {
#pragma unroll
                for (int j = 1; j < R; ++j) a[j] = a[j] * w[j];
            }
            dft_reg<T, R, DIR>(a);
#pragma unroll
            for (int j = 0; j < R; ++j) v[c][q + j * BQ] = a[j];
        }
    }
}

// Exchange after stage S: outputs to their Stockham positions in LDS, then read back positions t + m*TP.
// `xb` is this pencil's exchange buffer (padded_len(N, E) entries, with stride `ps` between entries: 1 for
// pencil-major buffers, the pencil count for interleaved ones). All NC components go through the buffer one
// after the other (one buffer per pencil keeps the LDS per block small enough for three blocks per CU; an NC-buffer
// exchange with one barrier pair per stage measured 3 % slower at 256^3).
template <typename T, int N, int E, int S, int NC>
__device__ __forceinline__ void reg_exchange(cx<T> (&v)[NC][E], int t, cx<T>* xb, int ps) {
    constexpr int R = stage_radix(N, E, S);
    constexpr int P = stage_prod(N, E, S);
    constexpr int TP = N / E;
    constexpr int BQ = E / R;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int i = t + q * TP;
            const int k = i & (P - 1);
            const int base = (i - k) * R + k;
#pragma unroll
            for (int j = 0; j < R; ++j) xb[padpos<E>(base + j * P) * ps] = v[c][q + j * BQ];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[c][m] = xb[padpos<E>(t + m * TP) * ps];
        __syncthreads();
    }
}

// Whole FFT (unnormalised) of the NC pencils held in v, natural order in and out.
template <typename T, int N, int E, int DIR, int NC, int S = 0>
__device__ __forceinline__ void reg_fft(cx<T> (&v)[NC][E], int t, const cx<T>* tw, cx<T>* xb, int ps) {
    static_assert(reg_schedule_ok<N, E>(), "every radix must divide E and the threads per pencil be a power of 2");
    reg_stage<T, N, E, S, DIR, NC>(v, t, tw);
    if constexpr (S + 1 < num_stages(N, E)) {
        reg_exchange<T, N, E, S, NC>(v, t, xb, ps);
        reg_fft<T, N, E, DIR, NC, S + 1>(v, t, tw, xb, ps);
    }
}

}  // namespace mmx

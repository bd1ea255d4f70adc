// Voxel-wise kernels of the unfused Born iteration, and the transposition of the four-step FFT.
//
// Grids whose axes cannot be transformed within one pencil pass (an axis longer than one LDS tile, or whose length has
// a prime factor above 7 and so goes through Bluestein's algorithm) run the iteration of MacroMax solver.py:1070-1085
// unfused: x = chi_eff E + s_eff (born_pro_kernel), the forward FFT over all axes, the Green's operator in k-space
// (green_pointwise_kernel), the inverse FFT, then d = chi_eff (x - E), E' = E + d and the FP64 norm partials
// (born_epi_kernel). Each is one pass over the fields; the same per-problem state words as the fused passes let a
// batch skip finished problems and pick each problem's current field buffer.
#pragma once
#include "mmx_kernels.hpp"

namespace mmx {

// Problem b of a launch: blockIdx.y; returns false when the problem is finished (or, for the prologue of a batch,
// when its work buffer is not stale), and selects the current and next field buffers by the parity bit.
template <typename T>
__device__ __forceinline__ bool problem_active(const PassArgs<T>& a, int64_t b, const cx<T>*& Ecur, cx<T>*& Enext) {
    Ecur = a.E;
    Enext = a.Eout;
    if (a.state == nullptr) return true;
    const unsigned st = a.state[b];
    if ((st & kStateDone) || (a.only_stale && !(st & kStateStale))) return false;
    if (st & kStateParity) {
        Ecur = a.Eout;
        Enext = const_cast<cx<T>*>(a.E);
    }
    return true;
}

template <typename T>
__device__ __forceinline__ Coef<T> problem_coef(const PassArgs<T>& a, int64_t b) {
    if (a.coef != nullptr) return a.coef[b];
    Coef<T> cf;
    cf.alpha_r = T(0); cf.alpha_i = T(1); cf.a = T(1); cf.c_r = T(0); cf.c_i = T(0); cf.pad_ = T(0);
    return cf;
}

// dst[b] = chi_eff E_cur[b] + s_eff[b] (a.src unused). src_from_out: read E from Eout instead (the next iteration's
// prologue from the field an iteration just wrote, assuming it is accepted).
template <typename T>
__global__ void __launch_bounds__(256) born_pro_kernel(PassArgs<T> a, int src_from_out) {
    const int64_t b = blockIdx.y;
    const cx<T>* Ecur;
    cx<T>* Enext;
    if (!problem_active(a, b, Ecur, Enext)) return;
    const cx<T>* Eb = (src_from_out ? Enext : Ecur) + b * a.E_bstride;
    const cx<T>* sb = a.s + b * a.s_bstride;
    const cx<T>* chib = a.chi + b * a.chi_bstride;
    cx<T>* Wb = a.dst + b * a.bstride;
    const Coef<T> cf = problem_coef(a, b);
    const int64_t V = a.comp_stride;
    const cx<T> z{T(0), T(0)};
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        cx<T> x[3] = {z, z, z}, y[3] = {z, z, z};
#pragma unroll
        for (int c = 0; c < 3; ++c) if (c < a.C) x[c] = Eb[c * V + v];
        chi_apply<T>(a, chib, v, x, y, cf);
#pragma unroll
        for (int c = 0; c < 3; ++c) if (c < a.C) Wb[c * V + v] = y[c] + scal(sb[c * V + v], cf.a);
    }
}

// W[b] <- G W[b] in k-space (natural fftfreq order on every axis), scaled by a.inv_ntotal.
template <typename T>
__global__ void __launch_bounds__(256) green_pointwise_kernel(PassArgs<T> a) {
    const int64_t b = blockIdx.y;
    const cx<T>* Ecur;
    cx<T>* Enext;
    if (!problem_active(a, b, Ecur, Enext)) return;
    cx<T>* Wb = a.dst + b * a.bstride;
    const Coef<T> cf = problem_coef(a, b);
    const int64_t V = a.comp_stride;
    const cx<T> z{T(0), T(0)};
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        int64_t i0, i1, i2;
        unravel3(v, a, i0, i1, i2);
        const T k0 = a.kvec[i0];
        const T k1 = (a.n1 > 1) ? a.kvec[a.k1ofs + i1] : T(0);
        const T k2 = (a.n2 > 1) ? a.kvec[a.k2ofs + i2] : T(0);
        cx<T> f[3] = {z, z, z};
#pragma unroll
        for (int c = 0; c < 3; ++c) if (c < a.C) f[c] = Wb[c * V + v];
        green_apply<T>(f, a.C, k0, k1, k2, v == 0, a, cf);
#pragma unroll
        for (int c = 0; c < 3; ++c) if (c < a.C) Wb[c * V + v] = f[c];
    }
}

// d = chi_eff (W - E), E' = E + d into the next field buffer, FP64 partial sums of |d|^2 and |E'|^2 per block into
// partials[2 (b gridDim.x + blockIdx.x) + {0, 1}].
template <typename T>
__global__ void __launch_bounds__(256) born_epi_kernel(PassArgs<T> a) {
    __shared__ double red[16];
    const int64_t b = blockIdx.y;
    const cx<T>* Ecur;
    cx<T>* Enext;
    if (!problem_active(a, b, Ecur, Enext)) return;
    const cx<T>* Eb = Ecur + b * a.E_bstride;
    cx<T>* Eo = Enext + b * a.E_bstride;
    const cx<T>* Wb = a.src + b * a.bstride;
    const cx<T>* chib = a.chi + b * a.chi_bstride;
    const Coef<T> cf = problem_coef(a, b);
    const int64_t V = a.comp_stride;
    const cx<T> z{T(0), T(0)};
    double dd = 0.0, ee = 0.0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        cx<T> e[3] = {z, z, z}, w[3] = {z, z, z}, d[3] = {z, z, z};
#pragma unroll
        for (int c = 0; c < 3; ++c) if (c < a.C) {
            e[c] = Eb[c * V + v];
            w[c] = Wb[c * V + v] - e[c];
        }
        chi_apply<T>(a, chib, v, w, d, cf);
#pragma unroll
        for (int c = 0; c < 3; ++c) if (c < a.C) {
            const cx<T> en = e[c] + d[c];
            Eo[c * V + v] = en;
            dd += norm2_f64(d[c]);
            ee += norm2_f64(en);
        }
    }
    dd = block_sum(dd, red);
    ee = block_sum(ee, red);
    if (threadIdx.x == 0) {
        const int64_t slot = b * gridDim.x + blockIdx.x;
        a.partials[2 * slot + 0] = dd;
        a.partials[2 * slot + 1] = ee;
    }
}

// Four-step FFT of an axis of N = N1 N2 points (index n = n1 N2 + n2): after the length-N1 FFTs along n1 and the
// twiddled length-N2 FFTs along n2, element (k1, k2) holds output k = k1 + N1 k2; this kernel writes it to its natural
// position. in: [outer][N1][N2][inner], out: [outer][N2][N1][inner] (out != in).
template <typename T>
__global__ void __launch_bounds__(256) axis_transpose_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                                                             int64_t outer, int64_t N1, int64_t N2, int64_t inner) {
    if (inner >= 64) {  // runs of `inner` contiguous values: copy them as they are
        const int64_t total = outer * N1 * N2 * inner;
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
             q += (int64_t)gridDim.x * blockDim.x) {
            const int64_t i = q % inner, r = q / inner;  // r = (o, k2, k1) of the output
            const int64_t k1 = r % N1, r2 = r / N1, k2 = r2 % N2, o = r2 / N2;
            out[q] = in[((o * N1 + k1) * N2 + k2) * inner + i];
        }
        return;
    }
    // 32 x 32 tiles of the (k1, k2) plane through LDS, one (outer, inner) plane per blockIdx.z-range
    __shared__ cx<T> tile[32][33];
    const int64_t t1 = (N1 + 31) / 32, t2 = (N2 + 31) / 32;
    const int64_t planes = outer * inner;
    for (int64_t job = blockIdx.x; job < planes * t1 * t2; job += gridDim.x) {
        const int64_t pl = job / (t1 * t2), tt = job % (t1 * t2);
        const int64_t b1 = (tt / t2) * 32, b2 = (tt % t2) * 32;
        const int64_t o = pl / inner, i = pl % inner;
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
        for (int r = ty; r < 32; r += 8) {
            const int64_t k1 = b1 + r, k2 = b2 + tx;
            if (k1 < N1 && k2 < N2) tile[r][tx] = in[((o * N1 + k1) * N2 + k2) * inner + i];
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            const int64_t k2 = b2 + r, k1 = b1 + tx;
            if (k1 < N1 && k2 < N2) out[((o * N2 + k2) * N1 + k1) * inner + i] = tile[tx][r];
        }
        __syncthreads();
    }
}

}  // namespace mmx

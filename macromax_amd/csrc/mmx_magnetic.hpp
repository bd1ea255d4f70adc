// Voxel-wise kernels of the Born iteration for magnetic and bi-(an)isotropic media (MacroMax solver.py:531-552,
// 564-584, 1070-1085), where the susceptibility operator has curl terms:
//
//     chi X = chi_EE X + chi_HE curl X + curl(chi_EH X + chi_HH curl X),   curl X = F^-1 (i k x F X).
//
// One iteration d = chi (G (chi E + s) - E) is evaluated with the FFTs of 24 field components (eight 3-component
// transforms) and five voxel-wise kernels, keeping in k-space what only k-space needs:
//   1. E^ = F E;  W2 = F^-1 (i k x E^) = curl E
//   2. [Z1; Y1] = [chi_EE E + chi_HE curl E + s; chi_EH E + chi_HH curl E]            (magnetic_real_kernel)
//   3. [Z1^; Y1^] = F [Z1; Y1]
//   4. B^ = G (Z1^ + i k x Y1^) - E^;  [B^; i k x B^]                                  (magnetic_green_kernel)
//   5. [B; curl B] = F^-1 [B^; i k x B^]
//   6. [Z2; Y2] = [chi_EE B + chi_HE curl B; chi_EH B + chi_HH curl B]                 (magnetic_real_kernel)
//   7. Y2 <- F^-1 (i k x F Y2) = curl Y2                                                (cross_ik_kernel between)
//   8. d = Z2 + curl Y2, E' = E + d, FP64 norm partials                                 (magnetic_final_kernel)
// (chi E + s and G (chi E + s) are never formed in real space.)
#pragma once
#include "mmx_kernels.hpp"

namespace mmx {

// A susceptibility tensor of the magnetic operator: kind 1 a scalar field, kind 2 a full 3x3 field ([9][V] row-major);
// vstride 0 for a constant (one value, or one 3x3 matrix, for every voxel).
template <typename T>
struct MagChi {
    const cx<T>* p;
    int kind;
    int64_t vstride;
};

template <typename T>
struct MagArgs {
    MagChi<T> ee, he, eh, hh;
    const cx<T>* E;     // [3][V]
    cx<T>* Eout;        // [3][V]
    const cx<T>* s;     // [3][V]
    cx<T>* w1;          // [3][V]  E^
    cx<T>* w2;          // [3][V]  curl E
    cx<T>* w3;          // [6][V]
    double* partials;   // 2 per block of the final kernel
    int64_t V;
    int64_t n1, n2;
    FDiv64 fn1, fn2;
    const T* kvec;
    int k1ofs, k2ofs;
    int nd;             // grid dimensions (components >= nd count as transverse at k = 0)
    T alpha_r, alpha_i;
};

template <typename T>
__device__ __forceinline__ void mag_mul_add(const MagChi<T>& m, int64_t v, int64_t V, const cx<T>* x, cx<T>* y) {
    const int64_t o = v * m.vstride;
    if (m.kind == 1) {
        const cx<T> c = m.p[o];
#pragma unroll
        for (int i = 0; i < 3; ++i) y[i] = y[i] + c * x[i];
    } else {
        const int64_t P = m.vstride ? V : 1;  // plane stride of the 9 entries
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) y[i] = y[i] + m.p[(3 * i + j) * P + o] * x[j];
    }
}

template <typename T>
__device__ __forceinline__ void mag_k(const MagArgs<T>& a, int64_t v, T& k0, T& k1, T& k2) {
    const int64_t t = fdiv64(v, a.fn2);
    const int64_t i2 = v - t * a.n2;
    const int64_t i0 = fdiv64(t, a.fn1);
    const int64_t i1 = t - i0 * a.n1;
    k0 = a.kvec[i0];
    k1 = a.n1 > 1 ? a.kvec[a.k1ofs + i1] : T(0);
    k2 = a.n2 > 1 ? a.kvec[a.k2ofs + i2] : T(0);
}

// i k x f
template <typename T>
__device__ __forceinline__ void cross_ik(T k0, T k1, T k2, const cx<T>* f, cx<T>* out) {
    const cx<T> c0 = scal(f[2], k1) - scal(f[1], k2);
    const cx<T> c1 = scal(f[0], k2) - scal(f[2], k0);
    const cx<T> c2 = scal(f[1], k0) - scal(f[0], k1);
    out[0] = cx<T>{-c0.i, c0.r};
    out[1] = cx<T>{-c1.i, c1.r};
    out[2] = cx<T>{-c2.i, c2.r};
}

// dst = i k x src (3 components; dst may equal src)
template <typename T>
__global__ void __launch_bounds__(256) cross_ik_kernel(MagArgs<T> a, const cx<T>* src, cx<T>* dst) {
    const int64_t V = a.V;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        T k0, k1, k2;
        mag_k(a, v, k0, k1, k2);
        const cx<T> f[3] = {src[v], src[V + v], src[2 * V + v]};
        cx<T> o[3];
        cross_ik(k0, k1, k2, f, o);
        dst[v] = o[0];
        dst[V + v] = o[1];
        dst[2 * V + v] = o[2];
    }
}

// w3 = [chi_EE x + chi_HE cx (+ s); chi_EH x + chi_HH cx] for x, cx (= curl x) given as [3][V] each.
template <typename T>
__global__ void __launch_bounds__(256) magnetic_real_kernel(MagArgs<T> a, const cx<T>* x_, const cx<T>* cx_, int add_s) {
    const int64_t V = a.V;
    const cx<T> z{T(0), T(0)};
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const cx<T> x[3] = {x_[v], x_[V + v], x_[2 * V + v]};
        const cx<T> c[3] = {cx_[v], cx_[V + v], cx_[2 * V + v]};
        cx<T> zz[3] = {z, z, z}, yy[3] = {z, z, z};
        if (add_s) {
            zz[0] = a.s[v];
            zz[1] = a.s[V + v];
            zz[2] = a.s[2 * V + v];
        }
        mag_mul_add(a.ee, v, V, x, zz);
        mag_mul_add(a.he, v, V, c, zz);
        mag_mul_add(a.eh, v, V, x, yy);
        mag_mul_add(a.hh, v, V, c, yy);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            a.w3[i * V + v] = zz[i];
            a.w3[(3 + i) * V + v] = yy[i];
        }
    }
}

// k-space: A = Z1^ + i k x Y1^, B^ = G A - E^ with the dyadic Green's operator (g (A - P_L A) + (i alpha_i / alpha)
// P_L A, g = -i alpha_i / (k^2 - alpha); P_L = I on the first nd components at k = 0); w3 = [B^; i k x B^].
template <typename T>
__global__ void __launch_bounds__(256) magnetic_green_kernel(MagArgs<T> a) {
    const int64_t V = a.V;
    const T ar = a.alpha_r, inv_ai = T(1) / a.alpha_i;
    const cx<T> gl = green_longitudinal(ar, inv_ai, T(1));  // i alpha_i / alpha
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        T k0, k1, k2;
        mag_k(a, v, k0, k1, k2);
        const cx<T> y[3] = {a.w3[3 * V + v], a.w3[4 * V + v], a.w3[5 * V + v]};
        cx<T> f[3];
        cross_ik(k0, k1, k2, y, f);
#pragma unroll
        for (int i = 0; i < 3; ++i) f[i] = f[i] + a.w3[i * V + v];
        const T kk = k0 * k0 + k1 * k1 + k2 * k2;
        const cx<T> g = green_scalar(kk, ar, inv_ai, T(1));  // -i alpha_i / (k^2 - alpha)
        cx<T> b[3];
        if (v == 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) b[i] = (i < a.nd ? gl : g) * f[i];
        } else {
            const cx<T> proj = scal(f[0], k0) + scal(f[1], k1) + scal(f[2], k2);
            const cx<T> q = (gl - g) * scal(proj, T(1) / kk);
            b[0] = g * f[0] + scal(q, k0);
            b[1] = g * f[1] + scal(q, k1);
            b[2] = g * f[2] + scal(q, k2);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) b[i] = b[i] - a.w1[i * V + v];
        cx<T> cb[3];
        cross_ik(k0, k1, k2, b, cb);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            a.w3[i * V + v] = b[i];
            a.w3[(3 + i) * V + v] = cb[i];
        }
    }
}

// d = Z2 + curl Y2 (w3 = [Z2; curl Y2]), E' = E + d, FP64 partials of |d|^2 and |E'|^2 per block.
template <typename T>
__global__ void __launch_bounds__(256) magnetic_final_kernel(MagArgs<T> a) {
    __shared__ double red[16];
    const int64_t V = a.V;
    double dd = 0.0, ee = 0.0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const cx<T> d = a.w3[i * V + v] + a.w3[(3 + i) * V + v];
            const cx<T> en = a.E[i * V + v] + d;
            a.Eout[i * V + v] = en;
            dd += norm2_f64(d);
            ee += norm2_f64(en);
        }
    }
    dd = block_sum(dd, red);
    ee = block_sum(ee, red);
    if (threadIdx.x == 0) {
        a.partials[2 * blockIdx.x] = dd;
        a.partials[2 * blockIdx.x + 1] = ee;
    }
}

}  // namespace mmx

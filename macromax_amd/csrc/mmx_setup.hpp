// Preconditioner set-up reductions over the permittivity (FP64), MacroMax solver.py:313-401:
//   sigma: max over voxels of the largest eigenvalue of (eps - shift I)^H (eps - shift I)   (calc_DeltaEE_beta2,
//          calc_sigmaEE, solver.py:392-401, for a real shift alpha_r beta),
//   gain:  min and max over voxels of the eigenvalues of -i/2 (eps - eps^H)                  (has_gain / has_loss,
//          solver.py:322-333).
// eps is complex128, [1, 1, V] (isotropic) or [3, 3, V] (row-major 3x3 per voxel). Eigenvalues of the Hermitian
// 3x3 matrices use the closed trigonometric form (the largest one is accurate to a few ulp of the matrix norm).
#pragma once
#include "mmx_device.hpp"

namespace mmx {

struct Herm3 {
    double a00, a11, a22;   // real diagonal
    cx<double> a01, a02, a12;  // upper triangle (a10 = conj(a01), ...)
};

// Eigenvalues l0 <= l1 <= l2 of a Hermitian 3x3 matrix.
__device__ __forceinline__ void herm3_eigvals(const Herm3& h, double& l0, double& l1, double& l2) {
    const double p1 = norm2_f64(h.a01) + norm2_f64(h.a02) + norm2_f64(h.a12);
    const double q = (h.a00 + h.a11 + h.a22) / 3.0;
    const double d0 = h.a00 - q, d1 = h.a11 - q, d2 = h.a22 - q;
    const double p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * p1;
    if (p2 <= 0.0) {  // a multiple of the identity
        l0 = l1 = l2 = q;
        return;
    }
    const double p = sqrt(p2 / 6.0);
    const double ip = 1.0 / p;
    // B = (A - q I) / p; r = det(B) / 2 (real for a Hermitian B)
    const double b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip;
    const cx<double> b01 = scal(h.a01, ip), b02 = scal(h.a02, ip), b12 = scal(h.a12, ip);
    // det = b00 (b11 b22 - |b12|^2) - |b01|^2 b22 - |b02|^2 b11 + 2 Re(b01 b12 conj(b02))
    const cx<double> t = b01 * b12;
    const double re = t.r * b02.r + t.i * b02.i;
    double r = 0.5 * (b00 * (b11 * b22 - norm2_f64(b12)) - norm2_f64(b01) * b22 - norm2_f64(b02) * b11 + 2.0 * re);
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    const double phi = acos(r) / 3.0;
    l2 = q + 2.0 * p * cos(phi);
    l0 = q + 2.0 * p * cos(phi + 2.0943951023931954923084289221863);  // + 2 pi / 3
    l1 = 3.0 * q - l0 - l2;
}

__device__ __forceinline__ cx<double> eps_at(const cx<double>* eps, int64_t V, int i, int j, int64_t v) {
    return eps[(int64_t)(3 * i + j) * V + v];
}

// Per-block extreme values -> partials[2 blk + {0, 1}]. MODE 0: {max sigma^2, unused}; MODE 1: {min, max} gain
// eigenvalue.
template <int MODE>
__global__ void __launch_bounds__(256) setup_extreme_kernel(const cx<double>* eps, int64_t V, int full3, double shift,
                                                            double* partials) {
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        if (!full3) {
            const cx<double> e = eps[v];
            if (MODE == 0) {
                const cx<double> d{e.r - shift, e.i};
                hi = fmax(hi, norm2_f64(d));
            } else {
                lo = fmin(lo, e.i);  // -i/2 (e - conj(e)) = Im(e)
                hi = fmax(hi, e.i);
            }
            continue;
        }
        cx<double> m[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) m[i][j] = eps_at(eps, V, i, j, v);
        Herm3 h;
        if (MODE == 0) {
            // H = D^H D, D = eps - shift I
#pragma unroll
            for (int i = 0; i < 3; ++i) m[i][i].r -= shift;
            auto col_dot = [&](int a, int b) {  // sum_k conj(D_ka) D_kb
                cx<double> s{0.0, 0.0};
#pragma unroll
                for (int k = 0; k < 3; ++k) s = s + conjg(m[k][a]) * m[k][b];
                return s;
            };
            h.a00 = col_dot(0, 0).r;
            h.a11 = col_dot(1, 1).r;
            h.a22 = col_dot(2, 2).r;
            h.a01 = col_dot(0, 1);
            h.a02 = col_dot(0, 2);
            h.a12 = col_dot(1, 2);
        } else {
            // H = -i/2 (eps - eps^H): H_ij = -i/2 (e_ij - conj(e_ji))
            auto entry = [&](int i, int j) {
                const cx<double> w = m[i][j] - conjg(m[j][i]);
                return cx<double>{0.5 * w.i, -0.5 * w.r};
            };
            h.a00 = entry(0, 0).r;
            h.a11 = entry(1, 1).r;
            h.a22 = entry(2, 2).r;
            h.a01 = entry(0, 1);
            h.a02 = entry(0, 2);
            h.a12 = entry(1, 2);
        }
        double l0, l1, l2;
        herm3_eigvals(h, l0, l1, l2);
        lo = fmin(lo, l0);
        hi = fmax(hi, l2);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = MODE == 0 ? smax[0] : smin[0];
        partials[2 * blockIdx.x + 1] = smax[0];
    }
}

__global__ void __launch_bounds__(256) setup_extreme_final_kernel(const double* partials, int n, double* out) {
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        lo = fmin(lo, partials[2 * q]);
        hi = fmax(hi, partials[2 * q + 1]);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smin[0];
        out[1] = smax[0];
    }
}

}  // namespace mmx

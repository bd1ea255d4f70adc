// Kernels of the Born-series iteration on gfx950.
//
// One iteration d = chi [G (chi E + s) - E], E' = E + d (MacroMax solver.py:1070-1085) runs as
// 2*ndim - 1 "pencil passes" over the field, each reading and writing the field once:
//
//   pass 1         innermost axis, forward FFT, prologue  x = chi E + s            (rows tile)
//   passes 2..d-1  the middle axes, forward FFT                                     (strided tile)
//   pass d         outermost axis: forward FFT -> Green's multiply -> inverse FFT   (strided tile)
//   passes ..      the middle axes, inverse FFT                                     (strided tile)
//   pass 2d-1      innermost axis, inverse FFT, epilogue d = chi (W - E), E' = E + d,
//                  FP64 partial sums of |d|^2 and |E'|^2 per workgroup              (rows tile)
//
// A tile holds P pencils of every component it needs in LDS; all FFT stages of that axis run in LDS.
// The 1/N of the inverse transform is folded into the Green's multiply.
#pragma once
#include "mmx_device.hpp"

namespace mmx {

// Per-problem preconditioner constants. chi_eff = a chi + c I, s_eff = a s, alpha for G.
// After a rejected update the host (or the device) changes alpha, a and c instead of rewriting chi
// (MacroMax solver.py:514-528 rescales chi and s in memory; the affine form is the same operator).
template <typename T>
struct Coef {
    T alpha_r, alpha_i;
    T a;
    T c_r, c_i;
    T pad_;
};

// Per-problem flags of batched solves (one 32-bit word per problem, owned by the device-side controller).
constexpr unsigned kStateParity = 1u;  // the current field lives in the second field buffer (Eout)
constexpr unsigned kStateDone = 2u;    // finished (residue or iteration limit): no pass touches the problem again
constexpr unsigned kStateStale = 4u;   // the work buffer does not hold the first pass of the current field

template <typename T>
struct PassArgs {
    // axis geometry (per component, per problem)
    int N;
    RadixPlan rp;             // radices and their division constants
    FDiv fN, fP, fncomp;      // division by N, P, ncomp
    FDiv64 fn1, fn2, fntile;  // unravel of voxel indices; tile index -> (outer, inner tile)
    FDiv64 ftpp;              // division by tiles_per_problem
    int64_t inner;            // element stride of the axis (1: rows)
    int64_t nouter;           // outer blocks per component per problem (rows: number of rows)
    int P;                    // pencils per tile and component
    int ntile_inner;          // strided tiles along the inner extent
    int64_t tiles_per_problem;
    int xcd_remap;            // 1: consecutive tiles run on the same XCD (blocks are dealt to 8 XCDs round-robin)
    int ncomp;                // components held per tile (1 when components are folded into nouter)
    int64_t comp_stride;      // V, elements between components
    int64_t bstride;          // elements between problems (C*V) for src/dst in fused passes
    int pad;                  // rows layout padding (complex elements)
    const cx<T>* tw;          // forward twiddles exp(-2 pi i t / N), t < N
    const cx<T>* src;
    cx<T>* dst;
    int dir;                  // plain passes: -1 forward, +1 inverse
    T scale;                  // plain passes: output scale
    int prologue;             // 0: load src; 1: x = chi_eff E + s_eff
    int green;                // 1: forward FFT, Green's multiply, inverse FFT
    int epilogue;             // 0: store dst; 1: Born update into Eout + norms
    // Born operands
    int C;                    // field components (1 or 3)
    int chi_kind;             // 1: [1,1,*g]; 2: [3,3,*g]; 3: symmetric [6,*g]
    const cx<T>* E;
    const cx<T>* chi;
    const cx<T>* s;
    int64_t E_bstride, chi_bstride, s_bstride;
    cx<T>* Eout;
    double* partials;         // 2 per (problem, tile)
    double* norms;            // per problem: {sum |d|^2, sum |Eout|^2}
    const Coef<T>* coef;      // one per problem
    // batched solves (state == nullptr otherwise): per-problem flags; passes skip finished problems, the first and
    // last passes read the current field from E or Eout by the parity bit and write the next one to the other
    const unsigned* state;
    int only_stale;           // first-pass launch: only the problems whose work buffer is stale
    // support of the source: s is zero outside [s_lo, s_hi) on every (padded) grid axis; s_restricted = 0 when the
    // box is the whole grid. Passes that add s skip its loads outside the box.
    int s_restricted;
    int s_lo[3], s_hi[3];
    FDiv64 frpp;              // single-component passes: outer index o -> problem (o / (C * outer extent))
    // transforms that are not one plain Stockham FFT of the pencil (plain passes only)
    int NT;                   // transform length held in LDS per pencil: N, or the Bluestein length
    int blue;                 // 1: Bluestein: chirp, FFT of length NT, product with bhat, inverse FFT, chirp
    const cx<T>* chirp;       // N values exp(-i pi n^2 / N), conjugated in the kernel for the inverse transform
    const cx<T>* bhat;        // NT values: FFT of the wrapped conjugate chirp of this pass's direction, / NT
    int tw4;                  // 1: four-step sub-pass: point n of a pencil whose outer index is r is multiplied by
    const cx<T>* tw4tab;      //    w^(DIR k1 n) before the FFT, k1 = r mod tw4_N1, w = tw4tab[1] = exp(-2 pi i / N4)
    FDiv64 ftw4;              //    division by tw4_N1
    // Green's function
    int64_t n0, n1, n2;       // grid shape padded with trailing 1s
    const T* kvec;            // k_0 at [0, n0), k_1 at [k1ofs, +n1), k_2 at [k2ofs, +n2)
    int k1ofs, k2ofs;
    int nd_data;              // grid ndim
    int vectorial;
    T inv_ntotal;
    int axis_orig;            // original grid axis (0..2) of this pass
    int kofs_axis;            // offset of that axis' k-vector in kvec
};

// Green's operator at one k for the C components held in f[] (MacroMax solver.py:566-581 with
// backend/__init__.py:606-649): g = -i alpha_i / (k^2 - alpha);
// vectorial: F <- g (F - Pi_L F) + (i alpha_i / alpha) Pi_L F, Pi_L = k k^T / k^2, Pi_L(0) = I on the
// first nd_data components and 0 on the others. The result is scaled by 1/N_total.
// 1/x: the hardware reciprocal (1 ulp) in single precision, a correctly rounded division in double.
template <typename T> __device__ __forceinline__ T recip(T x) { return T(1) / x; }
template <> __device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// g = -i alpha_i / (k^2 - alpha) = (1 - i u) / (1 + u^2) with u = (k^2 - alpha_r) / alpha_i, and the longitudinal
// i alpha_i / alpha = (1 + i v) / (1 + v^2) with v = alpha_r / alpha_i: no alpha_i^2 is formed, which overflows
// single precision for the alpha_i of 1e19 and more that the preconditioner search can give (a magnetic medium on a
// fine grid, where beta tends to 0).
template <typename T>
__device__ __forceinline__ cx<T> green_scalar(T kk, T alpha_r, T inv_ai, T scale) {
    const T u = (kk - alpha_r) * inv_ai;
    const T r = scale * recip(T(1) + u * u);
    return cx<T>{r, -u * r};
}
template <typename T>
__device__ __forceinline__ cx<T> green_longitudinal(T alpha_r, T inv_ai, T scale) {
    const T v = alpha_r * inv_ai;
    const T r = scale / (T(1) + v * v);
    return cx<T>{r, v * r};
}

template <typename T>
__device__ __forceinline__ void green_apply(cx<T>* f, int C, T k0, T k1, T k2, bool dc, const PassArgs<T>& a,
                                            const Coef<T>& cf) {
    const T kk = k0 * k0 + k1 * k1 + k2 * k2;
    const T inv_ai = T(1) / cf.alpha_i;  // uniform: computed once per tile
    const cx<T> g = green_scalar(kk, cf.alpha_r, inv_ai, a.inv_ntotal);
    if (!a.vectorial || C == 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) if (c < C) f[c] = g * f[c];
        return;
    }
    const cx<T> gl = green_longitudinal(cf.alpha_r, inv_ai, a.inv_ntotal);
    if (dc) {  // k = 0: the first nd_data components count as longitudinal
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] = (c < a.nd_data ? gl : g) * f[c];
        return;
    }
    // g (F - P_L F) + gl P_L F = g F + (gl - g) k (k . F) / k^2
    const cx<T> proj = scal(f[0], k0) + scal(f[1], k1) + scal(f[2], k2);
    const cx<T> q = (gl - g) * scal(proj, recip(kk));
    f[0] = g * f[0] + scal(q, k0);
    f[1] = g * f[1] + scal(q, k1);
    f[2] = g * f[2] + scal(q, k2);
}

// Is the pencil with outer index `row` (the axes before the pass axis, C order, one component of one problem)
// inside the source box on those axes? Passes that add s run along the innermost axis.
template <typename T>
__device__ __forceinline__ bool src_row_in_box(const PassArgs<T>& a, int64_t row) {
    if (!a.s_restricted) return true;
    if (a.axis_orig == 2) {
        const int64_t i0 = fdiv64(row, a.fn1);
        const int64_t i1 = row - i0 * a.n1;
        return i0 >= a.s_lo[0] && i0 < a.s_hi[0] && i1 >= a.s_lo[1] && i1 < a.s_hi[1];
    }
    if (a.axis_orig == 1) return row >= a.s_lo[0] && row < a.s_hi[0];
    return true;
}

template <typename T>
__device__ __forceinline__ void unravel3(int64_t v, const PassArgs<T>& a, int64_t& i0, int64_t& i1, int64_t& i2) {
    const int64_t t = fdiv64(v, a.fn2);
    i2 = v - t * a.n2;
    i0 = fdiv64(t, a.fn1);
    i1 = t - i0 * a.n1;
}

// chi at one voxel, loaded once and applied as often as needed. chi_kind 1: one value per voxel; 2: 3x3
// row-major [9, V]; 3: complex-symmetric 3x3 packed as [xx, xy, xz, yy, yz, zz] ([6, V]); 4: uniaxial
// chi = u I + w n n^T with a real unit axis n, stored as planes [u, w, (n_x, n_y), n_z] ([4, V] complex with the
// real n_z in the first V reals of the last plane): 28 bytes per voxel in complex64 instead of 48.
template <typename T>
__device__ __forceinline__ T uniaxial_nz(const cx<T>* chi_b, int64_t V, int64_t v) {
    return reinterpret_cast<const T*>(chi_b + 3 * V)[v];
}

template <typename T>
struct ChiVox {
    cx<T> m0, m1, m2, m3, m4, m5, m6, m7, m8;  // named (not an array): every access stays in registers
};

template <typename T>
__device__ __forceinline__ void chi_load(const PassArgs<T>& a, const cx<T>* chi_b, int64_t v, ChiVox<T>& q) {
    const int64_t V = a.comp_stride;
    q.m0 = chi_b[v];
    if (a.chi_kind == 4) {
        q.m1 = chi_b[V + v];
        q.m2 = chi_b[2 * V + v];
        q.m3 = cx<T>{uniaxial_nz(chi_b, V, v), T(0)};
        return;
    }
    if (a.chi_kind >= 2) {
        q.m1 = chi_b[V + v];
        q.m2 = chi_b[2 * V + v];
        q.m3 = chi_b[3 * V + v];
        q.m4 = chi_b[4 * V + v];
        q.m5 = chi_b[5 * V + v];
        if (a.chi_kind == 2) {
            q.m6 = chi_b[6 * V + v];
            q.m7 = chi_b[7 * V + v];
            q.m8 = chi_b[8 * V + v];
        }
    }
}

// y = chi_eff x = a chi x + c x
template <typename T>
__device__ __forceinline__ void chi_mulv(const PassArgs<T>& a, const ChiVox<T>& q, const cx<T>* x, cx<T>* y,
                                        const Coef<T>& cf) {
    const cx<T> cc{cf.c_r, cf.c_i};
    if (a.chi_kind == 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) if (c < a.C) y[c] = scal(q.m0 * x[c], cf.a) + cc * x[c];
        return;
    }
    cx<T> y0, y1, y2;
    if (a.chi_kind == 4) {  // u x + w n (n . x)
        const cx<T> t = q.m1 * (scal(x[0], q.m2.r) + scal(x[1], q.m2.i) + scal(x[2], q.m3.r));
        y0 = q.m0 * x[0] + scal(t, q.m2.r);
        y1 = q.m0 * x[1] + scal(t, q.m2.i);
        y2 = q.m0 * x[2] + scal(t, q.m3.r);
    } else if (a.chi_kind == 2) {
        y0 = q.m0 * x[0] + q.m1 * x[1] + q.m2 * x[2];
        y1 = q.m3 * x[0] + q.m4 * x[1] + q.m5 * x[2];
        y2 = q.m6 * x[0] + q.m7 * x[1] + q.m8 * x[2];
    } else {  // packed symmetric {xx, xy, xz, yy, yz, zz}
        y0 = q.m0 * x[0] + q.m1 * x[1] + q.m2 * x[2];
        y1 = q.m1 * x[0] + q.m3 * x[1] + q.m4 * x[2];
        y2 = q.m2 * x[0] + q.m4 * x[1] + q.m5 * x[2];
    }
    y[0] = scal(y0, cf.a) + cc * x[0];
    y[1] = scal(y1, cf.a) + cc * x[1];
    y[2] = scal(y2, cf.a) + cc * x[2];
}

template <typename T>
__device__ __forceinline__ void chi_apply(const PassArgs<T>& a, const cx<T>* chi_b, int64_t v, const cx<T>* x,
                                          cx<T>* y, const Coef<T>& cf) {
    ChiVox<T> q;
    chi_load<T>(a, chi_b, v, q);
    chi_mulv<T>(a, q, x, y, cf);
}

// Per-tile FP64 partial sums of |d|^2 and |Eout|^2 -> partials[2 blk + {0, 1}]; reduce_partials_kernel adds them
// in tile order after the pass (deterministic). A last-block-done reduction inside the pass would need an
// agent-scope release per block, which writes back the whole L2 of the XCD on this hardware.
template <typename T>
__device__ __forceinline__ void write_partials(double dd, double ee, const PassArgs<T>& a, int64_t blk, double* red) {
    dd = block_sum(dd, red);
    ee = block_sum(ee, red);
    if (threadIdx.x == 0) {
        a.partials[2 * blk + 0] = dd;
        a.partials[2 * blk + 1] = ee;
    }
}


template <int KIND>
struct PassFlags {
    static constexpr bool LOAD_PRO = (KIND == kProFwd || KIND == kAll);
    static constexpr bool FWD_FIRST = (KIND == kFwd || KIND == kProFwd || KIND == kGreen || KIND == kAll);
    static constexpr bool GREEN = (KIND == kGreen || KIND == kAll || KIND == kGreenEpiPro);
    static constexpr bool INV = (KIND == kInv || KIND == kGreen || KIND == kInvEpi || KIND == kAll ||
                                 KIND == kEpiPro || KIND == kGreenEpiPro);
    static constexpr bool EPI = (KIND == kInvEpi || KIND == kAll || KIND == kEpiPro || KIND == kGreenEpiPro);
    static constexpr bool PRO_AFTER = (KIND == kEpiPro || KIND == kGreenEpiPro);
    static constexpr bool STORE = !(KIND == kInvEpi || KIND == kAll);
};

// Position of point j of pencil p (one component) in a tile.
__device__ __forceinline__ int tile_pos(const TileGeom& g, int p, int j) { return j * g.SJ + p * g.SM; }

// The transform of a plain pass (one component per pencil) in the tile: the four-step twiddles (if any), then either
// the Stockham FFT of length NT = N or Bluestein's algorithm (X_k = c_k sum_n (x_n c_n) conj(c_(k-n)) with the chirp
// c_n = exp(-i pi n^2 / N), the circular convolution of length NT >= 2N - 1 as FFT, product with bhat, inverse FFT;
// conjugated chirps for the inverse transform). Points N..NT-1 of each pencil are scratch.
template <typename T, int DIR, int E>
__device__ void plain_transform(cx<T>* tile, const cx<T>* tw, const TileGeom& g, const PassArgs<T>& a, int tid, int nth,
                                int64_t r0, int64_t o, bool rows) {
    const int N = a.N, P = a.P;
    const int nv = P * N;
    if (a.tw4) {
        for (int idx = tid; idx < nv; idx += nth) {
            int p, j;
            if (rows) { p = fdiv(idx, a.fN); j = idx - p * N; }
            else { j = fdiv(idx, a.fP); p = idx - j * P; }
            const int64_t r = rows ? r0 + p : o;
            const int64_t k1 = r - fdiv64(r, a.ftw4) * a.ftw4.d;
            cx<T> w = a.tw4tab[k1 * j];
            if (DIR > 0) w.i = -w.i;
            const int q = tile_pos(g, p, j);
            tile[q] = tile[q] * w;
        }
        __syncthreads();
    }
    if (!a.blue) {
        tile_fft<T, DIR, E>(tile, tw, g, a.rp, tid, nth);
        return;
    }
    const int NT = a.NT;
    const FDiv fNT = make_fdiv(NT);
    for (int idx = tid; idx < P * NT; idx += nth) {
        int p, j;
        if (rows) { p = fdiv(idx, fNT); j = idx - p * NT; }
        else { j = fdiv(idx, a.fP); p = idx - j * P; }
        const int q = tile_pos(g, p, j);
        if (j < N) {
            cx<T> c = a.chirp[j];
            if (DIR > 0) c.i = -c.i;
            tile[q] = tile[q] * c;
        } else {
            tile[q] = cx<T>{T(0), T(0)};
        }
    }
    __syncthreads();
    tile_fft<T, -1, E>(tile, tw, g, a.rp, tid, nth);
    for (int idx = tid; idx < P * NT; idx += nth) {
        int p, j;
        if (rows) { p = fdiv(idx, fNT); j = idx - p * NT; }
        else { j = fdiv(idx, a.fP); p = idx - j * P; }
        const int q = tile_pos(g, p, j);
        tile[q] = tile[q] * a.bhat[j];
    }
    __syncthreads();
    tile_fft<T, +1, E>(tile, tw, g, a.rp, tid, nth);
    for (int idx = tid; idx < nv; idx += nth) {
        int p, j;
        if (rows) { p = fdiv(idx, a.fN); j = idx - p * N; }
        else { j = fdiv(idx, a.fP); p = idx - j * P; }
        cx<T> c = a.chirp[j];
        if (DIR > 0) c.i = -c.i;
        const int q = tile_pos(g, p, j);
        tile[q] = tile[q] * c;
    }
    __syncthreads();
}

template <typename T, int E, int KIND>
__global__ void __launch_bounds__(1024) pencil_pass_kernel(PassArgs<T> a) {
    using F = PassFlags<KIND>;
    extern __shared__ unsigned char smem_raw[];
    __shared__ double red[16];
    cx<T>* tw = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* tile = tw + a.NT;
    const int tid = threadIdx.x;
    const int nth = blockDim.x;
    const int N = a.N;
    const bool rows = (a.inner == 1);

    // which problem and which tile
    int64_t blk = blockIdx.x;
    if (a.xcd_remap) {
        // block i runs on XCD i % 8; give each XCD a contiguous range of tiles so neighbouring tiles share an L2
        const int64_t n = gridDim.x, q = n >> 3, rem = n & 7, r = blk & 7;
        blk = r * q + (r < rem ? r : rem) + (blk >> 3);
    }
    const int64_t b = fdiv64(blk, a.ftpp);
    const int64_t t = blk - b * a.tiles_per_problem;
    Coef<T> cf;
    if (a.coef != nullptr) {
        cf = a.coef[b];
    } else {
        cf.alpha_r = T(0); cf.alpha_i = T(1); cf.a = T(1); cf.c_r = T(0); cf.c_i = T(0); cf.pad_ = T(0);
    }

    TileGeom g;
    g.N = a.NT;
    g.M = a.ncomp * a.P;
    if (rows) { g.SJ = 1; g.SM = a.NT + a.pad; g.m_fast = 0; }
    else { g.SJ = g.M; g.SM = 1; g.m_fast = 1; }
    g.fM.d = g.M;
    g.fM.inv = 1.0f / float(g.M);


    // tile origin: rows -> first row r0; strided -> (outer block o, first inner index i0)
    int64_t r0 = 0, o = 0, ib = 0;
    if (rows) r0 = t * a.P;
    else { o = fdiv64(t, a.fntile); ib = (t - o * a.ntile_inner) * a.P; }
    const int64_t pb = b * a.bstride;  // problem offset in src/dst (fused passes; 0 when folded)
    const cx<T>* Ecur = a.E;
    cx<T>* Enext = a.Eout;
    if (a.state != nullptr) {
        // single-component passes hold pencils of one problem per tile (strided tiles never span problems)
        constexpr bool single = (KIND == kFwd || KIND == kInv);
        const unsigned st = a.state[single ? fdiv64(rows ? r0 : o, a.frpp) : b];
        if ((st & kStateDone) || (a.only_stale && !(st & kStateStale))) return;
        if (st & kStateParity) {
            Ecur = a.Eout;
            Enext = const_cast<cx<T>*>(a.E);
        }
    }

    for (int q = tid; q < a.NT; q += nth) tw[q] = a.tw[q];

    // ---------------- load ----------------
    if constexpr (F::LOAD_PRO) {
        // rows tile with all C components; voxel (row, j)
        const cx<T>* Eb = Ecur + b * a.E_bstride;
        const cx<T>* sb = a.s + b * a.s_bstride;
        const cx<T>* chib = a.chi + b * a.chi_bstride;
        const int64_t V = a.comp_stride;
        const int nvox = a.P * N;
        for (int idx = tid; idx < nvox; idx += nth) {
            const int p = fdiv(idx, a.fN);
            const int j = idx - p * N;
            const int64_t row = r0 + p;
            cx<T> y[3] = {{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
            if (row < a.nouter) {
                const int64_t v = row * N + j;
                cx<T> x[3] = {{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
#pragma unroll
                for (int c = 0; c < 3; ++c) if (c < a.C) x[c] = Eb[c * V + v];
                chi_apply<T>(a, chib, v, x, y, cf);
                const int ax = a.axis_orig;
                if (!a.s_restricted || (j >= a.s_lo[ax] && j < a.s_hi[ax] && src_row_in_box(a, row))) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) if (c < a.C) y[c] = y[c] + scal(sb[c * V + v], cf.a);
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) if (c < a.C) tile[(c * a.P + p) * g.SM + j] = y[c];
        }
    } else {
      if (rows) {
        const int total = a.ncomp * a.P * N;
        for (int idx = tid; idx < total; idx += nth) {
            const int rest = fdiv(idx, a.fN);
            const int j = idx - rest * N;
            const int c = fdiv(rest, a.fP);
            const int p = rest - c * a.P;
            const int64_t row = r0 + p;
            cx<T> val{T(0), T(0)};
            if (row < a.nouter) val = a.src[pb + c * a.comp_stride + row * N + j];
            tile[(c * a.P + p) * g.SM + j] = val;
        }
    } else {
        const int total = a.ncomp * a.P * N;
        const int64_t obase = pb + o * (int64_t)N * a.inner + ib;
        for (int idx = tid; idx < total; idx += nth) {
            const int rest = fdiv(idx, a.fP);
            const int p = idx - rest * a.P;
            const int j = fdiv(rest, a.fncomp);
            const int c = rest - j * a.ncomp;
            cx<T> val{T(0), T(0)};
            if (ib + p < a.inner) val = a.src[obase + c * a.comp_stride + j * a.inner + p];
            tile[idx] = val;  // == j*M + c*P + p
        }
      }
    }
    __syncthreads();

    // ---------------- FFT (and Green) ----------------
    if constexpr (KIND == kFwd) plain_transform<T, -1, E>(tile, tw, g, a, tid, nth, r0, o, rows);
    else if constexpr (F::FWD_FIRST) tile_fft<T, -1, E>(tile, tw, g, a.rp, tid, nth);
    if constexpr (F::GREEN) {
        const int nvox = a.P * N;
        for (int idx = tid; idx < nvox; idx += nth) {
            int p, j;
            int64_t v;
            if (rows) { p = fdiv(idx, a.fN); j = idx - p * N; v = (r0 + p) * N + j; if (r0 + p >= a.nouter) continue; }
            else { j = fdiv(idx, a.fP); p = idx - j * a.P; if (ib + p >= a.inner) continue; v = (o * N + j) * a.inner + ib + p; }
            int64_t i0, i1, i2;
            unravel3(v, a, i0, i1, i2);
            const T k0 = a.kvec[i0];
            const T k1 = (a.n1 > 1) ? a.kvec[a.k1ofs + i1] : T(0);
            const T k2 = (a.n2 > 1) ? a.kvec[a.k2ofs + i2] : T(0);
            cx<T> f[3] = {{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
#pragma unroll
            for (int c = 0; c < 3; ++c) if (c < a.ncomp) f[c] = tile[j * g.SJ + (c * a.P + p) * g.SM];
            green_apply<T>(f, a.ncomp, k0, k1, k2, v == 0, a, cf);
#pragma unroll
            for (int c = 0; c < 3; ++c) if (c < a.ncomp) tile[j * g.SJ + (c * a.P + p) * g.SM] = f[c];
        }
        __syncthreads();
    }
    if constexpr (KIND == kInv) plain_transform<T, +1, E>(tile, tw, g, a, tid, nth, r0, o, rows);
    else if constexpr (F::INV) tile_fft<T, +1, E>(tile, tw, g, a.rp, tid, nth);

    // ---------------- store ----------------
    if constexpr (F::EPI) {
        // rows tile with all C components: d = chi_eff (W - E), E' = E + d (then, for the fused kinds,
        // the next iteration's x = chi_eff E' + s_eff goes back into the tile)
        const cx<T>* Eb = Ecur + b * a.E_bstride;
        const cx<T>* chib = a.chi + b * a.chi_bstride;
        cx<T>* Eo = Enext + b * a.E_bstride;
        const int64_t V = a.comp_stride;
        double dd = 0.0, ee = 0.0;
        const int nvox = a.P * N;
        for (int idx = tid; idx < nvox; idx += nth) {
            const int p = fdiv(idx, a.fN);
            const int j = idx - p * N;
            const int64_t row = r0 + p;
            if (row >= a.nouter) continue;
            const int64_t v = row * N + j;
            const cx<T> z{T(0), T(0)};
            cx<T> e[3] = {z, z, z}, w[3] = {z, z, z}, d[3] = {z, z, z};
#pragma unroll
            for (int c = 0; c < 3; ++c) if (c < a.C) {
                e[c] = Eb[c * V + v];
                w[c] = tile[(c * a.P + p) * g.SM + j] - e[c];
            }
            ChiVox<T> q;
            chi_load<T>(a, chib, v, q);
            chi_mulv<T>(a, q, w, d, cf);
            cx<T> en[3] = {z, z, z};
#pragma unroll
            for (int c = 0; c < 3; ++c) if (c < a.C) {
                en[c] = e[c] + d[c];
                Eo[c * V + v] = en[c];
                dd += norm2_f64(d[c]);
                ee += norm2_f64(en[c]);
            }
            if constexpr (F::PRO_AFTER) {
                const cx<T>* sb = a.s + b * a.s_bstride;
                cx<T> x[3] = {z, z, z};
                chi_mulv<T>(a, q, en, x, cf);
                const int ax = a.axis_orig;
                if (!a.s_restricted || (j >= a.s_lo[ax] && j < a.s_hi[ax] && src_row_in_box(a, row))) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) if (c < a.C) x[c] = x[c] + scal(sb[c * V + v], cf.a);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) if (c < a.C) tile[(c * a.P + p) * g.SM + j] = x[c];
            }
        }
        write_partials(dd, ee, a, blk, red);
        if constexpr (F::PRO_AFTER) {
            __syncthreads();
            tile_fft<T, -1, E>(tile, tw, g, a.rp, tid, nth);
        }
    }
    if constexpr (F::STORE) {
      if (rows) {
        const int total = a.ncomp * a.P * N;
        for (int idx = tid; idx < total; idx += nth) {
            const int rest = fdiv(idx, a.fN);
            const int j = idx - rest * N;
            const int c = fdiv(rest, a.fP);
            const int p = rest - c * a.P;
            const int64_t row = r0 + p;
            if (row < a.nouter) a.dst[pb + c * a.comp_stride + row * N + j] = scal(tile[(c * a.P + p) * g.SM + j], a.scale);
        }
    } else {
        const int total = a.ncomp * a.P * N;
        const int64_t obase = pb + o * (int64_t)N * a.inner + ib;
        for (int idx = tid; idx < total; idx += nth) {
            const int rest = fdiv(idx, a.fP);
            const int p = idx - rest * a.P;
            const int j = fdiv(rest, a.fncomp);
            const int c = rest - j * a.ncomp;
            if (ib + p < a.inner) a.dst[obase + c * a.comp_stride + j * a.inner + p] = scal(tile[idx], a.scale);
        }
      }
    }
}

// Sum of per-tile partials -> out[2*b + {0,1}], fixed order (deterministic). One block per problem.
__global__ void __launch_bounds__(1024) reduce_partials_kernel(const double* partials, int64_t tiles, double* out) {
    __shared__ double red[16];
    const int64_t b = blockIdx.x;
    const double2* pp = reinterpret_cast<const double2*>(partials + 2 * b * tiles);
    const int64_t nt = blockDim.x;
    // eight independent partial sums per thread so that all its loads are in flight together; the order of the
    // additions depends only on the tile count and the block size
    double2 acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = make_double2(0.0, 0.0);
    int64_t q = threadIdx.x;
    for (; q + 7 * nt < tiles; q += 8 * nt) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = pp[q + u * nt];
#pragma unroll
        for (int u = 0; u < 8; ++u) { acc[u].x += v[u].x; acc[u].y += v[u].y; }
    }
    for (int u = 0; q < tiles; q += nt, ++u) {
        const double2 v = pp[q];
        acc[u & 7].x += v.x;
        acc[u & 7].y += v.y;
    }
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) { s0 += acc[u].x; s1 += acc[u].y; }
    s0 = block_sum(s0, red);
    s1 = block_sum(s1, red);
    if (threadIdx.x == 0) {
        out[2 * b] = s0;
        out[2 * b + 1] = s1;
    }
}

// Per-problem record of a batched solve (doubles; index constants are mirrored in include/macromax_hip.h).
constexpr int kTrPrev = 0;       // ||d|| of the last accepted update (+inf before the first)
constexpr int kTrFieldNorm = 1;  // ||E|| after the last accepted update
constexpr int kTrResidue = 2;    // ||d_prev|| / ||E||, 0 while ||d_prev|| = 0
constexpr int kTrIter = 3;       // iterations done, accepted or not
constexpr int kTrAlphaR = 4;     // current alpha and the affine chi/s coefficients, in FP64 (coef holds them in T)
constexpr int kTrAlphaI = 5;
constexpr int kTrA = 6;
constexpr int kTrCr = 7;
constexpr int kTrCi = 8;
constexpr int kTrDoneClock = 9;  // device wall clock when the problem finished
constexpr int kTrRejected = 10;  // rejected updates
constexpr int kTrLen = 12;

// Accept / reject and stop test of every unfinished problem after one batched iteration; the same rules as
// Solution.__iter__ and Solution.solve on the host (MacroMax solver.py:1078-1096, 1100-1123): accept when
// ||d|| < ||d_prev|| (the first update always), else alpha_i <- 1.5 alpha_i with chi_eff = a chi + c I rescaled
// as in solver.py:514-528 and the work buffer marked stale; stop when !(residue > tol && iterations < max_iter).
template <typename T>
__global__ void __launch_bounds__(256) batch_control_kernel(int64_t nbatch, const double* norms, unsigned* state,
                                                            Coef<T>* coef, double* track, double tol,
                                                            double max_iter) {
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nbatch; b += (int64_t)gridDim.x * blockDim.x) {
        unsigned st = state[b];
        if (st & kStateDone) continue;
        double* tr = track + b * kTrLen;
        st &= ~kStateStale;  // this iteration's last pass computed the next first pass from the field it wrote
        const double it = tr[kTrIter] + 1.0;
        tr[kTrIter] = it;
        const double d = sqrt(norms[2 * b]);
        if (d < tr[kTrPrev]) {
            st ^= kStateParity;
            tr[kTrPrev] = d;
            tr[kTrFieldNorm] = sqrt(norms[2 * b + 1]);
        } else {
            const double ai = tr[kTrAlphaI];
            const double ai_new = 1.5 * ai;
            const double A = ai / ai_new;
            const double B = (ai_new - ai) / ai_new;  // i (alpha_old - alpha_new) / alpha_i,new is real here
            tr[kTrA] *= A;
            tr[kTrCr] = tr[kTrCr] * A + B;
            tr[kTrCi] *= A;
            tr[kTrAlphaI] = ai_new;
            tr[kTrRejected] += 1.0;
            Coef<T> c;
            c.alpha_r = T(tr[kTrAlphaR]);
            c.alpha_i = T(ai_new);
            c.a = T(tr[kTrA]);
            c.c_r = T(tr[kTrCr]);
            c.c_i = T(tr[kTrCi]);
            c.pad_ = T(0);
            coef[b] = c;
            st |= kStateStale;
        }
        const double prev = tr[kTrPrev];
        const double res = prev > 0.0 ? prev / tr[kTrFieldNorm] : 0.0;
        tr[kTrResidue] = res;
        if (!(res > tol && it < max_iter)) {
            st |= kStateDone;
            tr[kTrDoneClock] = double(wall_clock64());
        }
        state[b] = st;
    }
}

__global__ void wall_clock_kernel(double* out) { out[0] = double(wall_clock64()); }

// sum |x|^2 in FP64: per-block partials (pass 1), then reduce_partials-style sum (pass 2).
template <typename T>
__global__ void __launch_bounds__(256) norm2_partial_kernel(const cx<T>* x, int64_t n, double* partials) {
    __shared__ double red[16];
    double s = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        s += norm2_f64(x[q]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) sum_kernel(const double* partials, int64_t n, double* out) {
    __shared__ double red[16];
    double s = 0.0;
    for (int64_t q = threadIdx.x; q < n; q += blockDim.x) s += partials[q];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = s;
}

// y = chi x per voxel (chi [1,1,V], [3,3,V], packed symmetric [6,V] or uniaxial [4,V]; x, y [C,V] with batch
// strides).
template <typename T>
__global__ void __launch_bounds__(256) chi_mul_kernel(const cx<T>* chi, int chi_kind, const cx<T>* x, cx<T>* y,
                                                      int C, int64_t V, int64_t nbatch) {
    const int64_t total = V * nbatch;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = q / V;
        const int64_t v = q - b * V;
        const cx<T>* xb = x + b * C * V;
        cx<T>* yb = y + b * C * V;
        if (chi_kind == 1) {
            const cx<T> c0 = chi[v];
            for (int c = 0; c < C; ++c) yb[c * V + v] = c0 * xb[c * V + v];
        } else if (chi_kind == 4) {
            const cx<T> u = chi[v], w = chi[V + v], nxy = chi[2 * V + v];
            const T nz = uniaxial_nz(chi, V, v);
            const cx<T> x0 = xb[v], x1 = xb[V + v], x2 = xb[2 * V + v];
            const cx<T> t = w * (scal(x0, nxy.r) + scal(x1, nxy.i) + scal(x2, nz));
            yb[v] = u * x0 + scal(t, nxy.r);
            yb[V + v] = u * x1 + scal(t, nxy.i);
            yb[2 * V + v] = u * x2 + scal(t, nz);
        } else {
            const cx<T> xv[3] = {xb[v], xb[V + v], xb[2 * V + v]};
            cx<T> m[9];
            if (chi_kind == 2) {
#pragma unroll
                for (int e = 0; e < 9; ++e) m[e] = chi[e * V + v];
            } else {
                const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
#pragma unroll
                for (int e = 0; e < 9; ++e) m[e] = chi[map[e] * V + v];
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) yb[r * V + v] = m[3 * r] * xv[0] + m[3 * r + 1] * xv[1] + m[3 * r + 2] * xv[2];
        }
    }
}

// Pi_L F at every k (MacroMax backend/__init__.py:606-649). F, out: [3, V] (out may alias F).
template <typename T>
__global__ void __launch_bounds__(256) longitudinal_projection_kernel(const cx<T>* F, cx<T>* out, int64_t n0,
                                                                      int64_t n1, int64_t n2, const T* kvec,
                                                                      int k1ofs, int k2ofs, int nd_data) {
    const int64_t V = n0 * n1 * n2;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i2 = v % n2, t = v / n2, i1 = t % n1, i0 = t / n1;
        const T k0 = kvec[i0];
        const T k1 = (n1 > 1) ? kvec[k1ofs + i1] : T(0);
        const T k2 = (n2 > 1) ? kvec[k2ofs + i2] : T(0);
        cx<T> f[3] = {F[v], F[V + v], F[2 * V + v]};
        cx<T> pl[3];
        if (v == 0) {
            for (int c = 0; c < 3; ++c) pl[c] = (c < nd_data) ? f[c] : cx<T>{T(0), T(0)};
        } else {
            const T kk = k0 * k0 + k1 * k1 + k2 * k2;
            cx<T> proj = scal(f[0], k0) + scal(f[1], k1) + scal(f[2], k2);
            proj = scal(proj, T(1) / kk);
            pl[0] = scal(proj, k0);
            pl[1] = scal(proj, k1);
            pl[2] = scal(proj, k2);
        }
        for (int c = 0; c < 3; ++c) out[c * V + v] = pl[c];
    }
}

}  // namespace mmx

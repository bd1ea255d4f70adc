// Pencil passes with the register-resident FFT (mmx_regfft.hpp) for the axis lengths it is instantiated for.
//
// Same pass kinds and semantics as pencil_pass_kernel (mmx_kernels.hpp); the tile's data lives in registers:
// thread (pencil p, position t) holds points t + m*TP (m < E) of each of its NC components.
//   row tiles:     lanes run along a row (t fastest) -> each row is read and written contiguously;
//   strided tiles: lanes run across the P columns (p fastest) -> P*8-byte segments per row of the tile.
#pragma once
#include "mmx_kernels.hpp"
#include "mmx_regfft.hpp"

namespace mmx {

// Element at a 32-bit byte offset from a wave-uniform base: the compiler keeps the base in SGPRs and issues
// global loads/stores with a single 32-bit VGPR offset (no 64-bit address arithmetic per access).
template <typename T> __device__ __forceinline__ const cx<T>& at(const cx<T>* base, uint32_t boff) {
    return *reinterpret_cast<const cx<T>*>(reinterpret_cast<const char*>(base) + boff);
}
template <typename T> __device__ __forceinline__ cx<T>& at(cx<T>* base, uint32_t boff) {
    return *reinterpret_cast<cx<T>*>(reinterpret_cast<char*>(base) + boff);
}

// chi of one voxel for a compile-time chi kind CK (1: scalar, 2: full 3x3 row-major, 3: packed symmetric,
// 4: uniaxial u I + w n n^T).
// chi_t is the chi of the tile (chi + tile offset tb, in complex elements), boff the byte offset of the voxel from it.
template <typename T, int CK>
__device__ __forceinline__ void chi_load_ck(const cx<T>* chi_t, int64_t V, uint32_t boff, ChiVox<T>& q, int64_t tb) {
    q.m0 = at(chi_t, boff);
    if constexpr (CK == 4) {
        q.m1 = at(chi_t + V, boff);
        q.m2 = at(chi_t + 2 * V, boff);
        // n_z: the reals of the last plane, element tb + boff / sizeof(cx) of that real array
        const T* nz_t = reinterpret_cast<const T*>(chi_t - tb + 3 * V) + tb;
        q.m3.r = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(nz_t) + boff / 2);
        q.m3.i = T(0);
        return;
    }
    if constexpr (CK >= 2) {
        q.m1 = at(chi_t + V, boff);
        q.m2 = at(chi_t + 2 * V, boff);
        q.m3 = at(chi_t + 3 * V, boff);
        q.m4 = at(chi_t + 4 * V, boff);
        q.m5 = at(chi_t + 5 * V, boff);
    }
    if constexpr (CK == 2) {
        q.m6 = at(chi_t + 6 * V, boff);
        q.m7 = at(chi_t + 7 * V, boff);
        q.m8 = at(chi_t + 8 * V, boff);
    }
}

// y = chi_eff x = a chi x + c x for NC components
template <typename T, int CK, int NC>
__device__ __forceinline__ void chi_mul_ck(const ChiVox<T>& q, const cx<T>* x, cx<T>* y, const Coef<T>& cf) {
    const cx<T> cc{cf.c_r, cf.c_i};
    if constexpr (CK == 1) {
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] = scal(q.m0 * x[c], cf.a) + cc * x[c];
    } else {
        cx<T> y0, y1, y2;
        if constexpr (CK == 4) {  // u x + w n (n . x)
            const cx<T> t = q.m1 * (scal(x[0], q.m2.r) + scal(x[1], q.m2.i) + scal(x[2], q.m3.r));
            y0 = q.m0 * x[0] + scal(t, q.m2.r);
            y1 = q.m0 * x[1] + scal(t, q.m2.i);
            y2 = q.m0 * x[2] + scal(t, q.m3.r);
        } else if constexpr (CK == 2) {
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
}

// Inputs of one voxel for the update (and the fused next prologue): E, chi and optionally s.
template <typename T, int NC, bool WITH_S> struct EpiIn {
    cx<T> e[NC];
    ChiVox<T> q;
    cx<T> s[WITH_S ? NC : 1];
};

template <typename T, int NC, int CK, bool WITH_S>
__device__ __forceinline__ void epi_load(EpiIn<T, NC, WITH_S>& in, const cx<T>* Et, const cx<T>* chit,
                                         const cx<T>* st, int64_t V, uint32_t bo, int64_t tb) {
#pragma unroll
    for (int c = 0; c < NC; ++c) in.e[c] = at(Et + c * V, bo);
    chi_load_ck<T, CK>(chit, V, bo, in.q, tb);
    if constexpr (WITH_S) {
#pragma unroll
        for (int c = 0; c < NC; ++c) in.s[c] = at(st + c * V, bo);
    }
}

// The host launches this kernel only when (N * inner + P) * sizeof(cx<T>) < 2^32, so every offset within a tile
// fits the 32-bit byte offsets used below. CK is the chi kind of the passes that apply chi (0 for the others).
// SB = 1 (update passes that also compute the next first pass, with a source confined to a box): s is added after
// the update loop, only inside the box, instead of being loaded with every voxel's E and chi.
template <typename T, int N, int E, int KIND, int NC, int CK, int SB>
__global__ void __launch_bounds__(512) reg_pass_kernel(PassArgs<T> a) {
    using F = PassFlags<KIND>;
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
    Coef<T> cf;
    if (a.coef != nullptr) {
        cf = a.coef[b];
    } else {
        cf.alpha_r = T(0); cf.alpha_i = T(1); cf.a = T(1); cf.c_r = T(0); cf.c_i = T(0); cf.pad_ = T(0);
    }

    int p, t;
    cx<T>* xb;
    int ps;
    if (rows) { t = tid & (TP - 1); p = tid >> LTP; xb = xbuf + p * padded_len(N, E); ps = 1; }
    else { p = tid & (P - 1); t = tid >> LP; xb = xbuf + p; ps = P; }

    int64_t r0 = 0, o = 0, ib = 0;
    if (rows) r0 = tl * P;
    else { o = fdiv64(tl, a.fntile); ib = (tl - o * a.ntile_inner) * P; }
    const bool valid = rows ? (r0 + p < a.nouter) : (ib + p < a.inner);
    // element offset of the tile within one component of one problem (uniform), and the byte offset of this
    // thread's register m from it: toff + m * mstep
    constexpr uint32_t ES = sizeof(cx<T>);
    const int64_t tb = rows ? r0 * N : o * (int64_t)N * a.inner + ib;
    const uint32_t toff = rows ? (uint32_t)(p * N + t) * ES : ((uint32_t)p + (uint32_t)t * (uint32_t)a.inner) * ES;
    const uint32_t mstep = rows ? (uint32_t)TP * ES : (uint32_t)TP * (uint32_t)a.inner * ES;
    const int64_t pb = b * a.bstride;
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

    load_twiddles<T, N, E>(tw, a.tw, tid, blockDim.x);
    __syncthreads();

    cx<T> v[NC][E];
    const cx<T> zero{T(0), T(0)};
    // source support of this pencil: rows inside the box, points j in [s_jlo, s_jhi) (passes that add s are row
    // passes along the innermost axis)
    bool s_row = true;
    int s_jlo = 0, s_jhi = N;
    if constexpr (F::LOAD_PRO || F::PRO_AFTER) {
        if (a.s_restricted) {
            s_row = src_row_in_box(a, r0 + p);
            s_jlo = a.s_lo[a.axis_orig];
            s_jhi = a.s_hi[a.axis_orig];
        }
    }
    if constexpr (F::LOAD_PRO) {
        const int64_t V = a.comp_stride;
        const cx<T>* Et = Ecur + b * a.E_bstride + tb;
        const cx<T>* st = a.s + b * a.s_bstride + tb;
        const cx<T>* chit = a.chi + b * a.chi_bstride + tb;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            cx<T> x[3] = {zero, zero, zero}, y[3] = {zero, zero, zero};
            if (valid) {
                const uint32_t bo = toff + m * mstep;
#pragma unroll
                for (int c = 0; c < NC; ++c) x[c] = at(Et + c * V, bo);
                ChiVox<T> qv;
                chi_load_ck<T, CK>(chit, V, bo, qv, tb);
                chi_mul_ck<T, CK, NC>(qv, x, y, cf);
                const int j = t + m * TP;
                if (s_row && j >= s_jlo && j < s_jhi) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) y[c] = y[c] + scal(at(st + c * V, bo), cf.a);
                }
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c][m] = y[c];
            if (NC == 3 && (m & 1)) asm volatile("" ::: "memory");  // bound the loads in flight per thread
        }
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m)
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c][m] = zero;
        if (valid) {
            const cx<T>* srct = a.src + pb + tb;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const uint32_t bo = toff + m * mstep;
#pragma unroll
                for (int c = 0; c < NC; ++c) v[c][m] = at(srct + c * a.comp_stride, bo);
            }
        }
    }

    // k components for the Green pass, loaded before the forward FFT so that the loads complete during it: the
    // components off the pass axis are fixed for this thread's pencil, the one along it varies with the point
    T kf[3], kpre[F::GREEN ? E : 1];
    bool off_dc = true;
    if constexpr (F::GREEN) {
        // pencils past the edge of a partial tile use index 0 (their results are not stored)
        int64_t col = (rows || !valid) ? 0 : ib + p;
        int64_t row = !valid ? 0 : (rows ? r0 + p : o);
        int64_t i0 = 0, i1 = 0, i2 = 0;
        if (a.axis_orig == 0) { i1 = fdiv64(col, a.fn2); i2 = col - i1 * a.n2; }
        else if (a.axis_orig == 1) { i0 = row; i2 = col; }
        else { i0 = fdiv64(row, a.fn1); i1 = row - i0 * a.n1; }
        kf[0] = (a.axis_orig != 0) ? a.kvec[i0] : T(0);
        kf[1] = (a.axis_orig != 1 && a.n1 > 1) ? a.kvec[a.k1ofs + i1] : T(0);
        kf[2] = (a.axis_orig != 2 && a.n2 > 1) ? a.kvec[a.k2ofs + i2] : T(0);
        off_dc = (i0 | i1 | i2) != 0;
#pragma unroll
        for (int m = 0; m < E; ++m) kpre[m] = a.kvec[a.kofs_axis + t + m * TP];
    }

    if constexpr (F::FWD_FIRST) reg_fft<T, N, E, -1, NC>(v, t, tw, xb, ps);
    if constexpr (F::GREEN) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int j = t + m * TP;
            const T kj = kpre[m];
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
    }
    // the update's inputs for the first voxel are requested before the inverse FFT, so that they arrive during it
    constexpr bool S_INLINE = F::PRO_AFTER && SB == 0;
    EpiIn<T, NC, S_INLINE> cur;
    if constexpr (F::EPI) {
        if (valid) {
            epi_load<T, NC, CK, S_INLINE>(cur, Ecur + b * a.E_bstride + tb, a.chi + b * a.chi_bstride + tb,
                                          a.s + b * a.s_bstride + tb, a.comp_stride, toff, tb);
        }
    }
    if constexpr (F::INV) reg_fft<T, N, E, +1, NC>(v, t, tw, xb, ps);

    if constexpr (F::EPI) {
        const int64_t V = a.comp_stride;
        const cx<T>* Et = Ecur + b * a.E_bstride + tb;
        const cx<T>* chit = a.chi + b * a.chi_bstride + tb;
        const cx<T>* st = a.s + b * a.s_bstride + tb;
        cx<T>* Eot = Enext + b * a.E_bstride + tb;
        double dd = 0.0, ee = 0.0;
        if (valid) {
            // Loads and stores retire in issue order on this hardware, so every input of voxel m + 1 (E, chi and,
            // for the fused next prologue, s) is requested before voxel m's result is stored (two voxels ahead
            // measured no faster).
            EpiIn<T, NC, S_INLINE> nxt;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                if (m + 1 < E) epi_load<T, NC, CK, S_INLINE>(nxt, Et, chit, st, V, toff + (m + 1) * mstep, tb);
                const uint32_t bo = toff + m * mstep;
                cx<T> w[3] = {zero, zero, zero}, d[3] = {zero, zero, zero}, en[3] = {zero, zero, zero};
#pragma unroll
                for (int c = 0; c < NC; ++c) w[c] = v[c][m] - cur.e[c];
                chi_mul_ck<T, CK, NC>(cur.q, w, d, cf);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    en[c] = cur.e[c] + d[c];
                    at(Eot + c * V, bo) = en[c];
                    dd += norm2_f64(d[c]);
                    ee += norm2_f64(en[c]);
                }
                if constexpr (F::PRO_AFTER) {
                    // the next iteration's x = chi_eff E' + s_eff, with chi already in registers
                    cx<T> x[3] = {zero, zero, zero};
                    chi_mul_ck<T, CK, NC>(cur.q, en, x, cf);
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        if constexpr (S_INLINE) v[c][m] = x[c] + scal(cur.s[c], cf.a);
                        else v[c][m] = x[c];
                    }
                }
                if (m + 1 < E) cur = nxt;
            }
            if constexpr (F::PRO_AFTER && SB != 0) {
                if (s_row) {
#pragma unroll
                    for (int m = 0; m < E; ++m) {
                        const int j = t + m * TP;
                        if (j >= s_jlo && j < s_jhi) {
#pragma unroll
                            for (int c = 0; c < NC; ++c)
                                v[c][m] = v[c][m] + scal(at(st + c * V, toff + m * mstep), cf.a);
                        }
                    }
                }
            }
        }
        write_partials(dd, ee, a, blk, red);
        if constexpr (F::PRO_AFTER) reg_fft<T, N, E, -1, NC>(v, t, tw, xb, ps);
    }
    if constexpr (F::STORE) {
      if (valid) {
        cx<T>* dstt = a.dst + pb + tb;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t bo = toff + m * mstep;
#pragma unroll
            for (int c = 0; c < NC; ++c) at(dstt + c * a.comp_stride, bo) = scal(v[c][m], a.scale);
        }
      }
    }
}

}  // namespace mmx

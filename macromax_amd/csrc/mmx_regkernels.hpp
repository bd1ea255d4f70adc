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

// Field traffic of the passes is streamed: every value is read once and written once per pass, and a field
// (hundreds of MiB) does not fit the caches, so loads and stores carry the non-temporal hint (the "nt" cache policy).
// Measured on MI355X with 7 read and 6 written planes of 128 MiB: 4.8 TB/s with default loads and stores, 5.5 TB/s
// non-temporal; 3 + 3 planes: 4.9 -> 5.7 TB/s.
// The fused Green pass is the exception: its strided tiles move 64-byte row segments, which neighbouring tiles on
// one XCD complete to whole cache lines in L2; with the hint it took 227 instead of 173 us at 256^3.
#ifndef MMX_NT
#define MMX_NT 1
#endif
template <int KIND> constexpr bool nt_kind() { return MMX_NT && KIND != kGreen; }
template <bool NT, typename T> __device__ __forceinline__ cx<T> ldg(const cx<T>* base, uint32_t boff) {
    const cx<T>& r = at(base, boff);
    if constexpr (NT) return cx<T>{__builtin_nontemporal_load(&r.r), __builtin_nontemporal_load(&r.i)};
    else return r;
}
template <bool NT, typename T> __device__ __forceinline__ T ldg_real(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T> __device__ __forceinline__ void stg(cx<T>* base, uint32_t boff, cx<T> v) {
    cx<T>& r = at(base, boff);
    if constexpr (NT) {
        __builtin_nontemporal_store(v.r, &r.r);
        __builtin_nontemporal_store(v.i, &r.i);
    } else {
        r = v;
    }
}

// chi of one voxel for a compile-time chi kind CK (1: scalar, 2: full 3x3 row-major, 3: packed symmetric,
// 4: uniaxial u I + w n n^T).
// chi_t is the chi of the tile (chi + tile offset tb, in complex elements), boff the byte offset of the voxel from it.
template <bool NT, typename T, int CK>
__device__ __forceinline__ void chi_load_ck(const cx<T>* chi_t, int64_t V, uint32_t boff, ChiVox<T>& q, int64_t tb) {
    q.m0 = ldg<NT>(chi_t, boff);
    if constexpr (CK == 4) {
        q.m1 = ldg<NT>(chi_t + V, boff);
        q.m2 = ldg<NT>(chi_t + 2 * V, boff);
        // n_z: the reals of the last plane, element tb + boff / sizeof(cx) of that real array
        const T* nz_t = reinterpret_cast<const T*>(chi_t - tb + 3 * V) + tb;
        q.m3.r = ldg_real<NT>(reinterpret_cast<const T*>(reinterpret_cast<const char*>(nz_t) + boff / 2));
        q.m3.i = T(0);
        return;
    }
    if constexpr (CK >= 2) {
        q.m1 = ldg<NT>(chi_t + V, boff);
        q.m2 = ldg<NT>(chi_t + 2 * V, boff);
        q.m3 = ldg<NT>(chi_t + 3 * V, boff);
        q.m4 = ldg<NT>(chi_t + 4 * V, boff);
        q.m5 = ldg<NT>(chi_t + 5 * V, boff);
    }
    if constexpr (CK == 2) {
        q.m6 = ldg<NT>(chi_t + 6 * V, boff);
        q.m7 = ldg<NT>(chi_t + 7 * V, boff);
        q.m8 = ldg<NT>(chi_t + 8 * V, boff);
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

// s_in: whether this voxel lies inside the source's support (only read when the source is restricted to a box;
// outside it s is zero and is not loaded)
template <bool NT, typename T, int NC, int CK, bool WITH_S>
__device__ __forceinline__ void epi_load(EpiIn<T, NC, WITH_S>& in, const cx<T>* Et, const cx<T>* chit,
                                         const cx<T>* st, int64_t V, uint32_t bo, int64_t tb, bool s_in = true) {
#pragma unroll
    for (int c = 0; c < NC; ++c) in.e[c] = ldg<NT>(Et + c * V, bo);
    chi_load_ck<NT, T, CK>(chit, V, bo, in.q, tb);
    if constexpr (WITH_S) {
#pragma unroll
        for (int c = 0; c < NC; ++c) in.s[c] = cx<T>{T(0), T(0)};
        if (s_in) {
#pragma unroll
            for (int c = 0; c < NC; ++c) in.s[c] = ldg<NT>(st + c * V, bo);
        }
    }
}

// Voxels of the update whose inputs (E, chi and the source) are requested ahead of their use, as a ring: the first
// PRE voxels' loads are issued before the inverse FFT, and voxel m + PRE's when voxel m is taken out. Loads and stores
// retire in issue order on this hardware, so with a distance of one voxel the update loop waited for one memory round
// trip per voxel. Measured at 256^3 (three components, uniaxial chi, non-temporal traffic): 0.907 ms per iteration
// with PRE = 1, 0.843 with 2, 0.835 with 4, 0.846 with 8 (all 8 voxels up front: more registers, fewer waves).
#ifndef MMX_EPI_PREFETCH
#define MMX_EPI_PREFETCH 0
#endif
template <int NC, int CK, int E>
constexpr int epi_prefetch() {
    if (MMX_EPI_PREFETCH > 0) return MMX_EPI_PREFETCH < E ? MMX_EPI_PREFETCH : E;
    if (NC == 1) return E;
    return (CK == 1 || CK == 4) ? (E < 4 ? E : 4) : 2;
}

// The host launches this kernel only when (N * inner + P) * sizeof(cx<T>) < 2^32, so every offset within a tile
// fits the 32-bit byte offsets used below. CK is the chi kind of the passes that apply chi (0 for the others).
// SB = 1 (update passes that also compute the next first pass, with a source confined to a box): s is loaded only
// inside the box (masked loads), instead of with every voxel's E and chi.
template <typename T, int N, int E, int KIND, int NC, int CK, int SB>
__global__ void __launch_bounds__(512) reg_pass_kernel(PassArgs<T> a) {
    using F = PassFlags<KIND>;
    constexpr bool NT = nt_kind<KIND>();
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

    // the fused Green pass loads its twiddle table first (measured 173 against 180 us at 256^3); the other passes
    // request it after their pencil loads, so that both arrive in one memory round trip (the table's LDS store waits
    // for every load issued before it)
    constexpr bool TW_FIRST = (KIND == kGreen);
    if constexpr (TW_FIRST) {
        load_twiddles<T, N, E>(tw, a.tw, tid, blockDim.x);
        __syncthreads();
    }
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
                for (int c = 0; c < NC; ++c) x[c] = ldg<NT>(Et + c * V, bo);
                ChiVox<T> qv;
                chi_load_ck<NT, T, CK>(chit, V, bo, qv, tb);
                chi_mul_ck<T, CK, NC>(qv, x, y, cf);
                const int j = t + m * TP;
                if (s_row && j >= s_jlo && j < s_jhi) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) y[c] = y[c] + scal(ldg<NT>(st + c * V, bo), cf.a);
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
                for (int c = 0; c < NC; ++c) v[c][m] = ldg<NT>(srct + c * a.comp_stride, bo);
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

    if constexpr (F::FWD_FIRST) {
        if constexpr (!TW_FIRST) {
            load_twiddles<T, N, E>(tw, a.tw, tid, blockDim.x);
            __syncthreads();
        }
        reg_fft<T, N, E, -1, NC>(v, t, tw, xb, ps);
    }
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
    // the update's inputs for the first PRE voxels are requested before the inverse FFT, so that they arrive during
    // it; a restricted source is read only inside its box (masked loads), a dense one with every voxel
    constexpr bool S_IN = F::PRO_AFTER;
    constexpr int PRE = F::EPI ? epi_prefetch<NC, CK, E>() : 1;
    EpiIn<T, NC, S_IN> slot[PRE];
    const int64_t V = a.comp_stride;
    const cx<T>* Et = Ecur + b * a.E_bstride + tb;
    const cx<T>* chit = a.chi + b * a.chi_bstride + tb;
    const cx<T>* st = a.s + b * a.s_bstride + tb;
    auto s_inside = [&](int m) {
        if constexpr (SB == 0) return true;
        const int j = t + m * TP;
        return s_row && j >= s_jlo && j < s_jhi;
    };
    if constexpr (F::EPI) {
        if (valid) {
#pragma unroll
            for (int q = 0; q < PRE; ++q)
                epi_load<NT, T, NC, CK, S_IN>(slot[q], Et, chit, st, V, toff + q * mstep, tb, s_inside(q));
        }
    }
    if constexpr (!F::FWD_FIRST && !TW_FIRST) {
        load_twiddles<T, N, E>(tw, a.tw, tid, blockDim.x);
        __syncthreads();
    }
    if constexpr (F::INV) reg_fft<T, N, E, +1, NC>(v, t, tw, xb, ps);

    if constexpr (F::EPI) {
        cx<T>* Eot = Enext + b * a.E_bstride + tb;
        double dd = 0.0, ee = 0.0;
        if (valid) {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const EpiIn<T, NC, S_IN> cur = slot[m % PRE];
                // refill the slot with voxel m + PRE before voxel m's stores are issued
                if (m + PRE < E)
                    epi_load<NT, T, NC, CK, S_IN>(slot[m % PRE], Et, chit, st, V, toff + (m + PRE) * mstep, tb,
                                              s_inside(m + PRE));
                const uint32_t bo = toff + m * mstep;
                cx<T> w[3] = {zero, zero, zero}, d[3] = {zero, zero, zero}, en[3] = {zero, zero, zero};
#pragma unroll
                for (int c = 0; c < NC; ++c) w[c] = v[c][m] - cur.e[c];
                chi_mul_ck<T, CK, NC>(cur.q, w, d, cf);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    en[c] = cur.e[c] + d[c];
                    stg<NT>(Eot + c * V, bo, en[c]);
                    dd += norm2_f64(d[c]);
                    ee += norm2_f64(en[c]);
                }
                if constexpr (F::PRO_AFTER) {
                    // the next iteration's x = chi_eff E' + s_eff, with chi already in registers
                    cx<T> x[3] = {zero, zero, zero};
                    chi_mul_ck<T, CK, NC>(cur.q, en, x, cf);
#pragma unroll
                    for (int c = 0; c < NC; ++c) v[c][m] = x[c] + scal(cur.s[c], cf.a);
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
            for (int c = 0; c < NC; ++c) stg<NT>(dstt + c * a.comp_stride, bo, scal(v[c][m], a.scale));
        }
      }
    }
}

}  // namespace mmx

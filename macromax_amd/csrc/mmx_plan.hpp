// Host-side planning of the FFT passes: axis factorisation (direct, Bluestein or four-step), tile shapes of the
// LDS-staged and register-resident pencil passes, and the k-vectors. Plain C++ without the HIP runtime, so that the
// CPU tests can build it with address and undefined-behaviour sanitizers (tests/host_planning_test.cpp).
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/macromax_hip.h"

namespace mmx {

constexpr int kMaxRadices = 12;

// Pass kinds: which prologue / transform / epilogue a pencil pass runs.
enum PassKind : int {
    kFwd = 0,     // load, forward FFT, store
    kInv = 1,     // load, inverse FFT, store * scale
    kProFwd = 2,  // Born prologue (chi E + s), forward FFT, store
    kGreen = 3,   // load, forward FFT, Green's operator, inverse FFT, store
    kInvEpi = 4,  // load, inverse FFT, Born epilogue (update + norms)
    kAll = 5,     // Born prologue, forward, Green, inverse, Born epilogue (single-axis grids)
    kEpiPro = 6,  // load, inverse FFT, Born epilogue, then the next iteration's prologue and forward FFT, store
    kGreenEpiPro = 7,  // single-axis grids: load (k-space), Green, inverse, epilogue, prologue, forward, store
};

// LDS position of exchange entry `pos` of the register-resident FFT: one pad slot every 2^pad_shift(E) entries
// (mmx_regfft.hpp), and the padded length of one pencil's exchange buffer.
constexpr int pad_shift(int E) { return E == 8 ? 3 : 4; }  // 12 and 24 points per thread: a slot every 16
constexpr int padded_len(int N, int E) { return N + (N >> pad_shift(E)); }


// Division by a launch-constant divisor without the integer-divide sequence: a float reciprocal gives the quotient to
// within one, the remainder test corrects it. Exact for 0 <= x < 2^22 and 1 <= d < 2^22 (x and d are exact in
// float, the estimate x * (1/d) carries a relative error below 2^-23, so |estimate - x/d| < 1/2). constexpr makes the
// functions callable from kernels and from host code alike; (float)x rounds to nearest and (int) truncates, the
// conversions a kernel would spell __int2float_rn and __float2int_rz.
struct FDiv {
    int d;
    float inv;
};

[[gnu::always_inline]] constexpr FDiv make_fdiv(int d) { return FDiv{d, 1.0f / float(d)}; }

[[gnu::always_inline]] constexpr int fdiv(int x, const FDiv& f) {
    const int q = int(float(x) * f.inv);
    const int r = x - q * f.d;
    return q + (r >= f.d) - (r < 0);
}

// 64-bit variant through a double reciprocal (exact for 0 <= x < 2^50).
struct FDiv64 {
    int64_t d;
    double inv;
};

[[gnu::always_inline]] constexpr FDiv64 make_fdiv64(int64_t d) { return FDiv64{d, 1.0 / double(d)}; }

[[gnu::always_inline]] constexpr int64_t fdiv64(int64_t x, const FDiv64& f) {
    const int64_t q = int64_t(double(x) * f.inv);
    const int64_t r = x - q * f.d;
    return q + (r >= f.d) - (r < 0);
}

// Text of the last error on this thread, and the error return of the planning and launch code.
inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

constexpr int kMaxN = 4096;          // longest axis a single pass holds in LDS
constexpr size_t kLdsBudget = 65536; // preferred dynamic LDS per workgroup (more only when one pencil needs it)
constexpr int kElemsF = 8;           // complex elements per thread per Stockham pass (complex64)
constexpr int kElemsD = 4;           // (complex128)
// Larger per-thread budgets, used only when a tile of one long pencil would need more than 1024 threads.
constexpr int kElemsFBig = 16;
constexpr int kElemsDBig = 8;

struct AxisPlan {
    int orig;        // original grid axis (0..2)
    int N;           // points along the axis
    int NT;          // length of the FFTs that transform it in one pass: N, or the Bluestein length (> 2N - 2)
    int nrad;
    int rad[kMaxRadices];  // radices of NT
    int64_t inner;   // stride of the axis
    int64_t nouter;  // blocks before the axis (per field)
    void* tw;        // device twiddles (forward), NT complex
    int blue;        // 1: Bluestein's algorithm (N has a prime factor above 7)
    void* chirp;     // N complex exp(-i pi n^2 / N)
    void* bhat_f;    // NT complex: FFT of the wrapped conjugate chirp / NT (forward transforms)
    void* bhat_i;    // the same for the inverse transform (the chirp itself)
    int split;       // 1: four-step: the axis is N1 x N2 sub-axes (n = n1 N2 + n2), each transformed in its own pass
    int N1, N2;
    void* tw4;       // N complex exp(-2 pi i t / N), the four-step twiddles w^(k1 n2)
};


// Radix orders of the LDS-staged Stockham transforms (plan parameter MMX_PARAM_RADIX_ORDER): each takes the radices
// greedily in its order of preference, and its stages run in that order. 0: radix 8 first (fewest stages);
// 1: radix 4 first (more stages, fewer registers per butterfly, no idle lanes at four points per thread);
// 2: the odd radices first.
constexpr int kRadixOrders = 3;
inline bool factorize(int n, int* rad, int& nrad, int order = 0) {
    static const int prefs[kRadixOrders][6] = {{8, 4, 2, 3, 5, 7}, {4, 2, 8, 3, 5, 7}, {7, 5, 3, 8, 4, 2}};
    nrad = 0;
    if (order < 0 || order >= kRadixOrders) return false;
    int m = n;
    for (int r : prefs[order]) {
        while (m % r == 0 && m > 1) {
            if (nrad >= kMaxRadices) return false;
            rad[nrad++] = r;
            m /= r;
        }
    }
    return m == 1;
}

inline bool smooth7(int64_t n) {
    if (n < 1) return false;
    for (int r : {2, 3, 5, 7})
        while (n % r == 0) n /= r;
    return n == 1;
}

inline int64_t next_smooth7(int64_t n) {
    while (!smooth7(n)) ++n;
    return n;
}

// One pass can transform an axis of n points directly (Stockham, n 7-smooth) or by Bluestein's algorithm (a 7-smooth
// transform length >= 2n - 1), both at most kMaxN points in LDS.
inline bool direct_ok(int64_t n) { return n <= kMaxN && smooth7(n); }
inline bool blue_ok(int64_t n) { return n >= 2 && next_smooth7(2 * n - 1) <= kMaxN; }

// Four-step split n = n1 n2 with both sub-lengths transformable in one pass: both 7-smooth if possible, then the
// most balanced. Returns false when none exists (a prime factor above kMaxN / 2, or n beyond kMaxN^2).
inline bool choose_split(int64_t n, int& n1, int& n2) {
    int best_score = 1 << 30;
    for (int64_t d = 2; d * d <= n; ++d) {
        if (n % d) continue;
        const int64_t e = n / d;
        if (e > kMaxN) continue;
        const bool ok_d = direct_ok(d) || blue_ok(d), ok_e = direct_ok(e) || blue_ok(e);
        if (!ok_d || !ok_e) continue;
        const int score = (direct_ok(d) ? 0 : 1 << 20) + (direct_ok(e) ? 0 : 1 << 20) + (int)(e - d);
        if (score < best_score) {
            best_score = score;
            n1 = (int)d;
            n2 = (int)e;
        }
    }
    return best_score < (1 << 30);
}

// k-vectors 2 pi fftfreq(n, d) in the order of numpy.fft.fftfreq (utils/ft/grid.py:426-439). For the transposed
// operator the Nyquist entry -pi/d of every even axis becomes +pi/d: A = F^-1 D(k) F has A^T = F^-1 D(k[-n]) F, and
// k[-n] = -k[n] everywhere except at n = N/2, where the dyadic Green's cross terms change sign.
inline std::vector<double> k_vectors(const int64_t shape3[3], const double step3[3], int transposed) {
    std::vector<double> k;
    for (int a = 0; a < 3; ++a) {
        const int64_t n = shape3[a];
        const double val = 1.0 / (double(n) * step3[a]);
        const int64_t npos = (n - 1) / 2 + 1;
        for (int64_t i = 0; i < n; ++i) {
            double f = (i < npos ? double(i) : double(i - n)) * val;
            if (transposed && n % 2 == 0 && n > 1 && i == n / 2) f = -f;
            k.push_back(2.0 * M_PI * f);
        }
    }
    return k;
}

inline int pow2floor(int64_t x) {
    int r = 1;
    while ((int64_t)r * 2 <= x) r *= 2;
    return r;
}

// Tile shape for one pass: pencils per tile P, threads, dynamic LDS bytes, tiles per problem.
struct TileCfg {
    int P, nth, pad;
    int big;                    // 1: the kernel variant with the larger per-thread element budget
    size_t lds;
    int64_t nouter;             // outer blocks the pass iterates (folded passes include batch and components)
    int64_t ntile_inner;
    int64_t tiles_per_problem;
    int64_t nblocks;
};

// fused: tiles hold `ncomp` components of one problem, grid = tiles_per_problem * nbatch.
// plain: (batch, component, outer) folded into nouter, ncomp = 1, `nbatch` counts fields.
template <typename T>
inline int tile_config(const AxisPlan& ax, int ncomp, bool fused, int64_t nbatch, size_t lds_max, TileCfg& t,
                int p_request = 0, size_t lds_budget = kLdsBudget, bool want_big = false) {
    int E = sizeof(T) == 4 ? kElemsF : kElemsD;
    t.big = 0;
    if (want_big || (int64_t)(fused ? ncomp : 1) * ax.NT > (int64_t)1024 * E) {
        E = sizeof(T) == 4 ? kElemsFBig : kElemsDBig;
        t.big = 1;
    }
    const size_t el = 2 * sizeof(T);
    const int N = ax.NT;  // points per pencil in LDS
    const bool rows = ax.inner == 1;
    if (!fused) ncomp = 1;
    t.nouter = fused ? ax.nouter : ax.nouter * nbatch;
    const int64_t mult = fused ? nbatch : 1;  // independent tile groups besides nouter
    t.pad = rows ? 1 : 0;
    const size_t per = (size_t)ncomp * (size_t)(N + t.pad) * el;
    size_t budget = lds_budget > lds_max ? lds_max : lds_budget;
    if (N * el + per > budget) budget = lds_max;
    if (N * el + per > budget) return fail(MMX_EUNSUP, "FFT length too large for one LDS tile");
    int64_t pm = (int64_t)((budget - N * el) / per);
    const int64_t pmax_thr = (int64_t)1024 * E / ((int64_t)ncomp * N);
    if (pmax_thr < 1) return fail(MMX_EUNSUP, "FFT length too large for one workgroup");
    if (pm > pmax_thr) pm = pmax_thr;
    if (rows) {
        const int64_t total_rows = t.nouter * mult;
        int64_t target = 8;
        const int64_t by_tiles = total_rows / 1024;  // keep >= ~1024 tiles when the field allows it
        if (by_tiles < target) target = by_tiles < 1 ? 1 : by_tiles;
        if (p_request > 0) target = p_request;
        if (pm > target) pm = target;
        if (pm > t.nouter) pm = t.nouter;
        t.P = p_request > 0 ? (int)(pm < 1 ? 1 : pm) : pow2floor(pm < 1 ? 1 : pm);
        t.ntile_inner = 1;
        t.tiles_per_problem = (t.nouter + t.P - 1) / t.P;
    } else {
        const int64_t target = p_request > 0 ? p_request : 16;
        if (pm > target) pm = target;
        t.P = ax.inner < pm ? (int)ax.inner : (p_request > 0 ? (int)(pm < 1 ? 1 : pm) : pow2floor(pm < 1 ? 1 : pm));
        // few tiles in total: trade segment length for parallelism
        if (p_request <= 0)
            while (t.P > 2 && t.nouter * mult * ((ax.inner + t.P - 1) / t.P) < 512) t.P /= 2;
        t.ntile_inner = (ax.inner + t.P - 1) / t.P;
        t.tiles_per_problem = t.nouter * t.ntile_inner;
    }
    const int64_t need = ((int64_t)ncomp * t.P * N + E - 1) / E;
    t.nth = (int)(((need + 63) / 64) * 64);
    if (t.nth < 64) t.nth = 64;
    if (t.nth > 1024) return fail(MMX_EUNSUP, "internal: tile needs more than 1024 threads");
    t.lds = (size_t)N * el + (size_t)ncomp * t.P * (N + t.pad) * el;
    t.nblocks = t.tiles_per_problem * mult;
    return MMX_OK;
}

// ---- register-resident kernels (complex64, N = 2^k, 3 * 2^k or 5 * 2^k, 64 <= N <= 2048) ----
// points per thread: fewer for the fused last pass, which also holds E, chi and s of its voxels; lengths with a
// factor 3 use 6 points per thread for three components (radix-2 stages, then radix 3; 132-172 VGPRs, measured 2-6 %
// faster than 12 points with 192-225 VGPRs) and 24 for one; lengths with a factor 5 use 10 and 20
#ifndef MMX_E_FUSED3
#define MMX_E_FUSED3 8
#endif
#ifndef MMX_E_GREEN3
#define MMX_E_GREEN3 8
#endif
#ifndef MMX_E_RADIX3_C3
#define MMX_E_RADIX3_C3 6
#endif
constexpr int reg_elems(int N, int NC, int KIND) {
    return (N % 5 == 0)                              ? (NC == 3 ? 10 : 20)
           : (N % 3 == 0)                              ? (NC == 3 ? MMX_E_RADIX3_C3 : 24)
           : (KIND == kEpiPro || KIND == kGreenEpiPro) ? (NC == 3 ? MMX_E_FUSED3 : 8)
           : (KIND == kGreen && NC == 3)             ? MMX_E_GREEN3
           : (NC == 3 ? 8 : (N >= 256 ? 16 : 8));
}

// Lengths the register kernels are instantiated for; their in-tile byte offsets are 32-bit, so the strided extent
// N * inner (plus at most 512 pencils) must stay below 2^32 bytes.
inline bool reg_supported(const AxisPlan& ax) {
    if (ax.blue || ax.split || ax.NT != ax.N) return false;
    const bool len = ax.N == 64 || ax.N == 128 || ax.N == 256 || ax.N == 512 || ax.N == 96 || ax.N == 192 ||
                     ax.N == 384 || ax.N == 80 || ax.N == 160 || ax.N == 320 || ax.N == 640 || ax.N == 768 ||
                     ax.N == 1024 || ax.N == 2048;
    return len && ((int64_t)ax.N * ax.inner + 512) * (int64_t)(2 * sizeof(float)) < (int64_t(1) << 32);
}

inline int reg_threads_per_pencil(int N, int NC, int kind) { return N / reg_elems(N, NC, kind); }

// Tile shape of a register-path pass: P pencils (a power of two) of TP threads each, at most 512 threads.
inline int reg_tile_config(const AxisPlan& ax, int ncomp, bool fused, int64_t nbatch, int p_request, int kind, TileCfg& t) {
    const int NC = fused ? ncomp : 1;
    const int TP = reg_threads_per_pencil(ax.N, NC, kind);
    const bool rows = ax.inner == 1;
    t.nouter = fused ? ax.nouter : ax.nouter * nbatch;
    const int64_t mult = fused ? nbatch : 1;
    t.pad = 0;
    t.big = 0;
    int pmax = 512 / TP;
    if (pmax < 1) return fail(MMX_EUNSUP, "internal: register FFT pencil too long");
    // defaults from the tile sweep at 128^3 and 256^3: 8 rows of 3 components / 16 rows of 1; 32 columns for
    // single-component strided passes; 8 columns of 3 components (16 of 1) for the Green pass
    int want = p_request > 0 ? p_request : (rows ? (NC == 3 ? 8 : 16) : (fused ? (NC == 3 ? 8 : 16) : 32));
    if (want > pmax) want = pmax;
    // Small grids: smaller tiles give more CUs work, down to 4 rows per tile, and to 16 columns (128-byte row
    // segments) for single-component strided passes or 4 for fused ones. A sweep of 128^2-512^2 and 64^3 grids
    // measured smaller tiles slower (a 512^2 3-component FFT: 28 us with 1-row / 2-column tiles, 15 us with 8 / 16;
    // a 256^2 scalar Born iteration: 22.4 us with 1 / 2, 19.3 us with 4 / 4; a 64^3 vector one: 46.6 us with 4-row
    // tiles, 42.4 us with 8).
    const int rows_min = 4, cols_min = fused ? 4 : 16;
    if (rows) {
        const int64_t total_rows = t.nouter * mult;
        if (p_request <= 0) {
            while (want > rows_min && total_rows / want < 256) want >>= 1;  // keep >= ~256 tiles when possible
            while (want > 1 && want > t.nouter) want >>= 1;
        }
        t.P = pow2floor(want);
        t.ntile_inner = 1;
        t.tiles_per_problem = (t.nouter + t.P - 1) / t.P;
    } else {
        int P = pow2floor(want);
        while (P > 1 && P > ax.inner) P >>= 1;
        if (p_request <= 0)
            while (P > cols_min && t.nouter * mult * ((ax.inner + P - 1) / P) < 512) P >>= 1;
        t.P = P;
        t.ntile_inner = (ax.inner + P - 1) / P;
        t.tiles_per_problem = t.nouter * t.ntile_inner;
    }
    t.nth = t.P * TP;
    const int E = ax.N / TP;
    t.lds = (size_t)(ax.N + (size_t)t.P * padded_len(ax.N, E)) * (2 * sizeof(float));
    t.nblocks = t.tiles_per_problem * mult;
    return MMX_OK;
}

}  // namespace mmx

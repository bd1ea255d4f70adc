// CPU test of the host-side FFT planning (macromax_amd/csrc/mmx_plan.hpp), built with g++ under the address and
// undefined-behaviour sanitizers by tests/test_host_planning.py. Each check compares the planning code with an
// independent statement of what it must satisfy: a sieve for 7-smoothness, the sorted 7-smooth (Hamming) sequence for
// next_smooth7, brute-force divisor search for the four-step split, exhaustive division for the fast divisors, the
// numpy ifftshift form of fftfreq for the k-vectors, and the launch limits of gfx950 (1024 threads and 160 KiB of
// LDS per workgroup) for the tile shapes. Prints one line per group and exits non-zero on the first failure.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <random>
#include <set>
#include <vector>

#include "../macromax_amd/csrc/mmx_plan.hpp"

using namespace mmx;

static int g_failures = 0;

#define CHECK(cond, ...)                                              \
    do {                                                              \
        if (!(cond)) {                                                \
            fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
            fprintf(stderr, __VA_ARGS__);                             \
            fprintf(stderr, "\n");                                    \
            if (++g_failures > 20) exit(1);                           \
        }                                                             \
    } while (0)

constexpr size_t kLdsMax = 160 * 1024;  // LDS per workgroup on gfx950

// Largest prime factor of every n < lim (sieve).
static std::vector<int> largest_prime_factor(int lim) {
    std::vector<int> lpf(lim, 0);
    for (int p = 2; p < lim; ++p)
        if (lpf[p] == 0)
            for (int m = p; m < lim; m += p) lpf[m] = p;
    if (lim > 1) lpf[1] = 1;
    return lpf;
}

// All 7-smooth numbers up to lim, sorted (products 2^a 3^b 5^c 7^d).
static std::vector<int64_t> hamming7(int64_t lim) {
    std::vector<int64_t> out;
    for (int64_t a = 1; a <= lim; a *= 2)
        for (int64_t b = a; b <= lim; b *= 3)
            for (int64_t c = b; c <= lim; c *= 5)
                for (int64_t d = c; d <= lim; d *= 7) out.push_back(d);
    std::sort(out.begin(), out.end());
    return out;
}

static void test_smooth_and_factorize() {
    const int lim = 20000;
    const std::vector<int> lpf = largest_prime_factor(lim);
    for (int n = 1; n < lim; ++n) {
        const bool smooth = lpf[n] <= 7;
        CHECK(smooth7(n) == smooth, "n=%d", n);
        for (int order = 0; order < kRadixOrders; ++order) {
            int rad[kMaxRadices], nrad = -1;
            const bool ok = factorize(n, rad, nrad, order);
            if (n > kMaxN) continue;
            CHECK(ok == smooth, "factorize(%d, order %d)", n, order);
            if (!ok) continue;
            int64_t prod = 1;
            int n8 = 0, n4 = 0, n2 = 0;
            for (int i = 0; i < nrad; ++i) {
                CHECK(rad[i] == 2 || rad[i] == 3 || rad[i] == 4 || rad[i] == 5 || rad[i] == 7 || rad[i] == 8,
                      "n=%d radix %d", n, rad[i]);
                prod *= rad[i];
                n8 += rad[i] == 8;
                n4 += rad[i] == 4;
                n2 += rad[i] == 2;
            }
            CHECK(prod == n, "n=%d order %d product %lld", n, order, (long long)prod);
            CHECK(nrad <= kMaxRadices, "n=%d nrad=%d", n, nrad);
            // the power-of-two part 2^e: order 1 takes radix 4 while it divides, so no 8 and at most one 2;
            // the others take 8 first, then at most one 4 or one 2
            int e = 0;
            for (int m = n; m % 2 == 0; m /= 2) ++e;
            if (order == 1) CHECK(n8 == 0 && n4 == e / 2 && n2 == e % 2, "n=%d order 1", n);
            else CHECK(n8 == e / 3 && n4 == (e % 3 == 2) && n2 == (e % 3 == 1), "n=%d order %d", n, order);
            // order 2 runs every odd radix before the power-of-two ones, the others after them
            bool seen_even = false, seen_odd = false, odd_after_even = false, even_after_odd = false;
            for (int i = 0; i < nrad; ++i) {
                const bool odd = rad[i] % 2 == 1;
                if (odd && seen_even) odd_after_even = true;
                if (!odd && seen_odd) even_after_odd = true;
                seen_even |= !odd;
                seen_odd |= odd;
            }
            CHECK(order == 2 ? !odd_after_even : !even_after_odd, "n=%d order %d stage order", n, order);
        }
        {
            int rad[kMaxRadices], nrad = 0;
            CHECK(!factorize(n, rad, nrad, kRadixOrders), "order %d accepted", kRadixOrders);
        }
    }
    CHECK(!smooth7(0) && !smooth7(-4), "non-positive lengths are not 7-smooth");
    // every length a Bluestein plan can ask for (2n - 1 for n <= kMaxN), then sparse samples up to 2^22
    const std::vector<int64_t> h = hamming7(int64_t(1) << 23);
    for (int64_t n = 1; n <= 2 * kMaxN + 1; ++n)
        CHECK(next_smooth7(n) == *std::lower_bound(h.begin(), h.end(), n), "next_smooth7(%lld)", (long long)n);
    std::mt19937_64 rng(7);
    for (int t = 0; t < 2000; ++t) {
        const int64_t n = 1 + int64_t(rng() % ((int64_t(1) << 22)));
        CHECK(next_smooth7(n) == *std::lower_bound(h.begin(), h.end(), n), "next_smooth7(%lld)", (long long)n);
    }
    printf("smooth7, next_smooth7, factorize: ok\n");
}

static bool one_pass_ok(int64_t n) { return direct_ok(n) || blue_ok(n); }

static void test_axis_choice() {
    const std::vector<int> lpf = largest_prime_factor(kMaxN + 1);
    for (int n = 1; n <= kMaxN; ++n) {
        CHECK(direct_ok(n) == (lpf[n] <= 7), "direct_ok(%d)", n);
        // Bluestein: a 7-smooth transform length of at least 2n - 1 that fits one LDS tile
        if (n >= 2) {
            const int64_t nt = next_smooth7(2 * n - 1);
            CHECK(blue_ok(n) == (nt <= kMaxN), "blue_ok(%d)", n);
            if (blue_ok(n)) CHECK(nt >= 2 * n - 1 && smooth7(nt), "n=%d NT=%lld", n, (long long)nt);
        }
    }
    CHECK(blue_ok(2039) && !blue_ok(2053), "Bluestein reaches primes up to 2048");
    // four-step split: brute force over all divisor pairs
    std::vector<int64_t> cases;
    for (int64_t n = 4097; n < 4097 + 3000; ++n) cases.push_back(n);
    for (int64_t n : {int64_t(8192), int64_t(12288), int64_t(8198), int64_t(4096) * 4096, int64_t(4096) * 4096 + 1,
                      int64_t(1) << 22, int64_t(2039) * 2029, int64_t(4099) * 2, int64_t(3) * 4096 * 5})
        cases.push_back(n);
    for (int64_t n : cases) {
        bool any = false;
        for (int64_t d = 2; d * d <= n; ++d)
            if (n % d == 0 && n / d <= kMaxN && one_pass_ok(d) && one_pass_ok(n / d)) any = true;
        int n1 = -1, n2 = -1;
        const bool got = choose_split(n, n1, n2);
        CHECK(got == any, "choose_split(%lld) = %d, brute force %d", (long long)n, got, any);
        if (got) {
            CHECK((int64_t)n1 * n2 == n, "n=%lld split %d x %d", (long long)n, n1, n2);
            CHECK(n1 <= n2 && n2 <= kMaxN && one_pass_ok(n1) && one_pass_ok(n2), "n=%lld split %d x %d", (long long)n, n1, n2);
            // a split of two directly transformable lengths is preferred whenever one exists
            bool direct_pair = false;
            for (int64_t d = 2; d * d <= n; ++d)
                if (n % d == 0 && n / d <= kMaxN && direct_ok(d) && direct_ok(n / d)) direct_pair = true;
            if (direct_pair) CHECK(direct_ok(n1) && direct_ok(n2), "n=%lld split %d x %d", (long long)n, n1, n2);
        }
    }
    int n1 = 0, n2 = 0;
    CHECK(choose_split(8192, n1, n2) && n1 == 64 && n2 == 128, "8192 -> %d x %d (most balanced)", n1, n2);
    CHECK(!choose_split(8191, n1, n2), "8191 is prime");
    printf("direct_ok, blue_ok, choose_split: ok\n");
}

static void test_fdiv() {
    // exhaustive over x in [0, 2^22) for small, smooth, prime and large divisors
    std::vector<int> ds;
    for (int d = 1; d <= 64; ++d) ds.push_back(d);
    for (int d : {96, 100, 127, 128, 255, 256, 257, 1000, 1023, 1025, 2039, 3072, 4095, 4096, 4097, 65535, 65537,
                  (1 << 21) - 1, (1 << 21) + 1, (1 << 22) - 1})
        ds.push_back(d);
    for (int d : ds) {
        const FDiv f = make_fdiv(d);
        int bad = 0;
        for (int x = 0; x < (1 << 22); ++x)
            if (fdiv(x, f) != x / d) ++bad;
        CHECK(bad == 0, "fdiv by %d: %d wrong quotients", d, bad);
    }
    std::mt19937_64 rng(11);
    for (int t = 0; t < 2000000; ++t) {
        const int d = 1 + int(rng() % ((1 << 22) - 1));
        const int x = int(rng() % (1 << 22));
        CHECK(fdiv(x, make_fdiv(d)) == x / d, "fdiv(%d, %d)", x, d);
    }
    // 64-bit: random, and around exact multiples where the estimate is most likely off by one
    for (int t = 0; t < 2000000; ++t) {
        const int64_t d = 1 + int64_t(rng() % ((t & 1) ? (int64_t(1) << 40) : 5000));
        int64_t x = int64_t(rng() % (int64_t(1) << 50));
        if (t % 3 == 0) {
            const int64_t q = x / d;
            x = q * d + int64_t(t % 5) - 2;
            if (x < 0 || x >= (int64_t(1) << 50)) continue;
        }
        CHECK(fdiv64(x, make_fdiv64(d)) == x / d, "fdiv64(%lld, %lld)", (long long)x, (long long)d);
    }
    printf("fdiv, fdiv64: ok\n");
}

static void test_k_vectors() {
    const int64_t shapes[][3] = {{1, 1, 1}, {4, 1, 1}, {5, 7, 2}, {256, 256, 256}, {11, 8, 40}, {3, 2, 1}};
    const double steps[][3] = {{1.0, 1.0, 1.0}, {0.1, 1.0, 1.0}, {0.25, 0.5, 2.0}, {1.0 / 16, 1.0 / 16, 1.0 / 16},
                               {0.3, 0.7, 0.05}, {1.0, 2.0, 3.0}};
    for (int c = 0; c < 6; ++c)
        for (int tr = 0; tr < 2; ++tr) {
            const std::vector<double> k = k_vectors(shapes[c], steps[c], tr);
            size_t ofs = 0;
            for (int a = 0; a < 3; ++a) {
                const int64_t n = shapes[c][a];
                const double d = steps[c][a];
                double sum = 0.0;
                for (int64_t i = 0; i < n; ++i) {
                    // fftfreq as the shifted index ((i + floor(n/2)) mod n) - floor(n/2)
                    const double m = double((i + n / 2) % n - n / 2);
                    double want = 2.0 * M_PI * m / (double(n) * d);
                    if (tr && n % 2 == 0 && n > 1 && i == n / 2) want = -want;
                    CHECK(fabs(k[ofs + i] - want) <= 1e-12 * fabs(want) + 1e-300, "shape %d axis %d i=%lld", c, a, (long long)i);
                    sum += k[ofs + i];
                }
                // the spectrum is symmetric about 0 except for the Nyquist term -pi/d (+pi/d transposed)
                const double nyq = (n % 2 == 0 && n > 1) ? (tr ? 1.0 : -1.0) * M_PI / d : 0.0;
                CHECK(fabs(sum - nyq) < 1e-9 * (1.0 + fabs(nyq)), "shape %d axis %d sum %g", c, a, sum);
                CHECK(k[ofs] == 0.0, "shape %d axis %d: k[0] = %g", c, a, k[ofs]);
                ofs += n;
            }
            CHECK(ofs == k.size(), "shape %d: %zu k-values", c, k.size());
        }
    printf("k_vectors: ok\n");
}

static AxisPlan make_axis(int N, int64_t inner, int64_t nouter) {
    AxisPlan ax{};
    ax.N = N;
    ax.NT = N;
    if (!direct_ok(N)) {
        ax.blue = 1;
        ax.NT = (int)next_smooth7(2 * (int64_t)N - 1);
    }
    factorize(ax.NT, ax.rad, ax.nrad);
    ax.inner = inner;
    ax.nouter = nouter;
    return ax;
}

template <typename T>
static void check_tile(const AxisPlan& ax, int ncomp, bool fused, int64_t nbatch, int p_request, size_t budget,
                       bool want_big, int& n_ok) {
    TileCfg t{};
    const int rc = tile_config<T>(ax, ncomp, fused, nbatch, kLdsMax, t, p_request, budget, want_big);
    const int nc = fused ? ncomp : 1;
    const size_t el = 2 * sizeof(T);
    if (rc != MMX_OK) {
        // refused only when one pencil (plus the twiddles) cannot fit the LDS or one workgroup
        const int E = sizeof(T) == 4 ? kElemsFBig : kElemsDBig;
        const bool lds_short = (size_t)ax.NT * el + (size_t)nc * (ax.NT + (ax.inner == 1)) * el > kLdsMax;
        const bool thr_short = (int64_t)nc * ax.NT > (int64_t)1024 * E;
        CHECK(lds_short || thr_short, "N=%d NT=%d inner=%lld nc=%d: refused", ax.N, ax.NT, (long long)ax.inner, nc);
        return;
    }
    ++n_ok;
    const int E = t.big ? (sizeof(T) == 4 ? kElemsFBig : kElemsDBig) : (sizeof(T) == 4 ? kElemsF : kElemsD);
    CHECK(t.P >= 1 && t.nth >= 64 && t.nth <= 1024 && t.nth % 64 == 0, "N=%d P=%d nth=%d", ax.N, t.P, t.nth);
    CHECK((int64_t)t.nth * E >= (int64_t)nc * t.P * ax.NT, "N=%d P=%d nth=%d E=%d", ax.N, t.P, t.nth, E);
    CHECK(t.lds <= kLdsMax, "N=%d P=%d lds=%zu", ax.N, t.P, t.lds);
    CHECK(t.lds == (size_t)ax.NT * el + (size_t)nc * t.P * (ax.NT + t.pad) * el, "N=%d lds=%zu", ax.N, t.lds);
    // in-tile LDS indices stay within the exact range of the fast division
    CHECK((int64_t)nc * t.P * (ax.NT + t.pad) < (int64_t(1) << 22), "N=%d P=%d", ax.N, t.P);
    const int64_t mult = fused ? nbatch : 1;
    CHECK(t.nouter == (fused ? ax.nouter : ax.nouter * nbatch), "nouter");
    if (ax.inner == 1) {
        CHECK(t.P <= std::max<int64_t>(1, t.nouter), "N=%d rows P=%d nouter=%lld", ax.N, t.P, (long long)t.nouter);
        CHECK(t.tiles_per_problem * t.P >= t.nouter && (t.tiles_per_problem - 1) * t.P < t.nouter, "row tiles");
    } else {
        CHECK(t.P <= ax.inner, "N=%d cols P=%d inner=%lld", ax.N, t.P, (long long)ax.inner);
        CHECK(t.ntile_inner * t.P >= ax.inner && (t.ntile_inner - 1) * t.P < ax.inner, "column tiles");
        CHECK(t.tiles_per_problem == t.nouter * t.ntile_inner, "tiles per problem");
    }
    CHECK(t.nblocks == t.tiles_per_problem * mult, "nblocks");
    if (p_request > 0) CHECK(t.P <= p_request, "P=%d above the request %d", t.P, p_request);
}

static void test_tiles() {
    std::vector<int> Ns;
    for (int n = 2; n <= kMaxN; ++n)
        if (direct_ok(n) && (n < 64 || n % 8 == 0 || n % 7 == 0 || n % 9 == 0)) Ns.push_back(n);
    for (int n : {11, 13, 101, 127, 1021, 2039, 2047}) Ns.push_back(n);
    int n_ok = 0;
    for (int N : Ns)
        for (int64_t inner : {int64_t(1), int64_t(3), int64_t(64), int64_t(4096)})
            for (int64_t nouter : {int64_t(1), int64_t(5), int64_t(4096)})
                for (int nc : {1, 3}) {
                    const AxisPlan ax = make_axis(N, inner, nouter);
                    for (bool fused : {false, true})
                        for (int preq : {0, 1, 8, 48}) {
                            check_tile<float>(ax, nc, fused, fused ? 2 : 6, preq, kLdsBudget, false, n_ok);
                            check_tile<double>(ax, nc, fused, fused ? 2 : 6, preq, kLdsBudget, false, n_ok);
                        }
                    check_tile<float>(ax, nc, true, 1, 0, kLdsMax, true, n_ok);
                    check_tile<double>(ax, nc, true, 1, 0, kLdsMax, true, n_ok);
                }
    // every single-component pass of an axis up to kMaxN plans, in both precisions
    for (int N : Ns)
        for (int64_t inner : {int64_t(1), int64_t(256)}) {
            const AxisPlan ax = make_axis(N, inner, 16);
            TileCfg t{};
            CHECK(tile_config<float>(ax, 1, false, 1, kLdsMax, t) == MMX_OK, "float N=%d inner=%lld", N, (long long)inner);
            CHECK(tile_config<double>(ax, 1, false, 1, kLdsMax, t) == MMX_OK, "double N=%d inner=%lld", N, (long long)inner);
        }
    printf("tile_config: ok (%d tile shapes checked)\n", n_ok);
}

static void test_reg_tiles() {
    int n_ok = 0, n_sup = 0;
    for (int N = 2; N <= kMaxN; ++N) {
        const AxisPlan ax = make_axis(N, 1, 64);
        if (!reg_supported(ax)) continue;
        ++n_sup;
        for (int kind : {(int)kFwd, (int)kInv, (int)kProFwd, (int)kGreen, (int)kEpiPro})
            for (int nc : {1, 3})
                for (int64_t inner : {int64_t(1), int64_t(2), int64_t(64), int64_t(N)})
                    for (int64_t nouter : {int64_t(1), int64_t(3), int64_t(N)})
                        for (int preq : {0, 1, 4, 16, 64}) {
                            const AxisPlan a = make_axis(N, inner, nouter);
                            const bool fused = kind != kFwd && kind != kInv;
                            const int NC = fused ? nc : 1;
                            const int E = reg_elems(N, NC, kind);
                            CHECK(N % E == 0, "N=%d E=%d", N, E);
                            TileCfg t{};
                            const int rc = reg_tile_config(a, nc, fused, 2, preq, kind, t);
                            CHECK(rc == MMX_OK, "N=%d kind=%d", N, kind);
                            if (rc != MMX_OK) continue;
                            ++n_ok;
                            const int TP = N / E;
                            CHECK(t.P >= 1 && (t.P & (t.P - 1)) == 0, "N=%d P=%d", N, t.P);
                            CHECK(t.nth == t.P * TP && t.nth <= 512, "N=%d P=%d nth=%d", N, t.P, t.nth);
                            CHECK(t.lds <= kLdsMax, "N=%d lds=%zu", N, t.lds);
                            CHECK(t.lds == (size_t)(N + (size_t)t.P * padded_len(N, E)) * 8, "N=%d lds", N);
                            if (inner == 1)
                                CHECK(t.tiles_per_problem * t.P >= t.nouter, "N=%d row tiles", N);
                            else
                                CHECK(t.P <= inner && t.ntile_inner * t.P >= inner, "N=%d P=%d inner=%lld", N, t.P,
                                      (long long)inner);
                        }
    }
    CHECK(n_sup == 14, "%d register-path lengths", n_sup);
    // the 32-bit in-tile offsets of the register kernels
    CHECK(!reg_supported(make_axis(2048, int64_t(1) << 18, 1)), "2048 x 2^18 strided extent exceeds 2^32 bytes");
    CHECK(reg_supported(make_axis(2048, int64_t(1) << 17, 1)), "2048 x 2^17");
    CHECK(!reg_supported(make_axis(101, 1, 1)) && !reg_supported(make_axis(4096, 1, 1)), "outside the register set");
    printf("reg_tile_config: ok (%d tile shapes over %d lengths)\n", n_ok, n_sup);
}

int main() {
    setvbuf(stdout, nullptr, _IOLBF, 0);
    test_smooth_and_factorize();
    test_axis_choice();
    test_fdiv();
    test_k_vectors();
    test_tiles();
    test_reg_tiles();
    if (g_failures) {
        fprintf(stderr, "%d failures\n", g_failures);
        return 1;
    }
    printf("host planning: all checks passed\n");
    return 0;
}

// C ABI of libmacromax_hip.so: plans, pass scheduling and launches (declarations in include/macromax_hip.h).
#include <atomic>
#include <complex>
#include <mutex>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

#include "../../include/macromax_hip.h"
#include "mmx_kernels.hpp"
#include "mmx_plan.hpp"
#include "mmx_magnetic.hpp"
#include "mmx_pointwise.hpp"
#include "mmx_regkernels.hpp"
#include "mmx_magfused.hpp"
#include "mmx_setup.hpp"

using namespace mmx;

// Norm partials of mmx_born_step_noalloc: one array per device, held until the process ends. Host code never uses
// this name as an address (it is a host-side stand-in); it uses the device address that hipGetSymbolAddress
// returns for the current device (noalloc_scratch below).
static __device__ __align__(256) double scratch[MMX_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

#define MMX_HIP_CHECK(expr)                                                                    \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(MMX_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct mmx_plan_s {
    int dtype;
    int ndim;
    int64_t shape3[3];
    int64_t V;
    int naxes;
    AxisPlan ax[3];  // non-trivial axes, outermost first
    AxisPlan sub[3][2];  // sub-axes of split axes: [0] along n1 (stride N2 inner), [1] along n2 (stride inner)
    void* fft_scratch;   // one field batch: the four-step transposition's target
    size_t fft_scratch_bytes;
    int fused_ok[2];     // per component count (1, 3): -1 unknown, 1 the fused pass schedule fits, 0 unfused
    void* kvec;      // device k-vectors, n0 + n1 + n2 reals
    double step3[3]; // grid step per axis (1 for padded axes)
    int transposed;  // 1: k-vectors of the transposed operator (Nyquist entries negated)
    int64_t s_lo[3], s_hi[3];  // support box of the sources of later Born calls (default: the whole grid)
    int k1ofs, k2ofs;
    void* scratch;   // FP64 partial sums
    size_t scratch_bytes;
    int64_t scratch_gen;  // number of (re)allocations of scratch: work captured in a graph holds its address
    // mmx_born_iterate_split: two partial buffers, each with the event that marks its last pass on the launch stream
    void* slot_partials[2];
    size_t slot_bytes[2];
    hipEvent_t slot_done[2];
    int device;
    size_t lds_max;  // device limit of shared memory per block
    // tuning (0 = automatic)
    int g_axis;
    int p_rows, p_strided, p_green;
    int xcd_remap;
    int big_elems;   // bit k set: pass kind k uses the larger per-thread element budget
    int reg_fft;     // 1: power-of-two axes use the register-resident FFT kernels
    int radix_order; // radix order of the LDS-staged transforms (factorize)
    size_t lds_budget;
};

namespace {

template <typename T>
int upload_complex(const std::vector<std::complex<double>>& h, void** dev) {
    std::vector<cx<T>> tmp(h.size());
    for (size_t i = 0; i < h.size(); ++i) tmp[i] = cx<T>{T(h[i].real()), T(h[i].imag())};
    MMX_HIP_CHECK(hipMalloc(dev, tmp.size() * sizeof(cx<T>)));
    MMX_HIP_CHECK(hipMemcpy(*dev, tmp.data(), tmp.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
    return MMX_OK;
}

// Bluestein tables of an axis of N points with transform length NT (double precision on the host): the chirp
// c_n = exp(-i pi n^2 / N) (n^2 reduced mod 2N first), and bhat = DFT_NT(b) / NT for the convolution kernel
// b_m = conj(c_m) at m and NT - m (forward transforms) or c_m (inverse transforms).
template <typename T>
int make_bluestein(AxisPlan& ax) {
    const int N = ax.N, NT = ax.NT;
    std::vector<std::complex<double>> c(N), bf(NT, 0.0), bi(NT, 0.0), hf(NT), hi(NT), w(NT);
    for (int n = 0; n < N; ++n) {
        const int64_t q = ((int64_t)n * n) % (2 * (int64_t)N);
        const double ang = -M_PI * double(q) / double(N);
        c[n] = std::complex<double>(cos(ang), sin(ang));
    }
    for (int m = 0; m < N; ++m) {
        bf[m] = std::conj(c[m]);
        bi[m] = c[m];
        if (m > 0) {
            bf[NT - m] = std::conj(c[m]);
            bi[NT - m] = c[m];
        }
    }
    for (int t = 0; t < NT; ++t) {
        const double ang = -2.0 * M_PI * double(t) / double(NT);
        w[t] = std::complex<double>(cos(ang), sin(ang));
    }
    for (int k = 0; k < NT; ++k) {
        std::complex<double> sf = 0.0, si = 0.0;
        for (int m = 0; m < NT; ++m) {
            const std::complex<double> e = w[(int64_t)m * k % NT];
            sf += bf[m] * e;
            si += bi[m] * e;
        }
        hf[k] = sf / double(NT);
        hi[k] = si / double(NT);
    }
    int rc = upload_complex<T>(c, &ax.chirp);
    if (!rc) rc = upload_complex<T>(hf, &ax.bhat_f);
    if (!rc) rc = upload_complex<T>(hi, &ax.bhat_i);
    return rc;
}

template <typename T>
int upload_real(const std::vector<double>& h, void** dev) {
    std::vector<T> tmp(h.begin(), h.end());
    MMX_HIP_CHECK(hipMalloc(dev, tmp.size() * sizeof(T)));
    MMX_HIP_CHECK(hipMemcpy(*dev, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return MMX_OK;
}

template <typename T>
int reupload_real(const std::vector<double>& h, void* dev) {
    std::vector<T> tmp(h.begin(), h.end());
    MMX_HIP_CHECK(hipDeviceSynchronize());
    MMX_HIP_CHECK(hipMemcpy(dev, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return MMX_OK;
}

template <typename T>
int make_twiddles(int N, void** dev) {
    std::vector<cx<T>> tw(N);
    for (int t = 0; t < N; ++t) {
        // exp(-2 pi i t / N) evaluated in double; sin/cos of the reduced angle keep symmetry exact
        double ang = -2.0 * M_PI * double(t) / double(N);
        tw[t] = cx<T>{T(cos(ang)), T(sin(ang))};
    }
    MMX_HIP_CHECK(hipMalloc(dev, N * sizeof(cx<T>)));
    MMX_HIP_CHECK(hipMemcpy(*dev, tw.data(), N * sizeof(cx<T>), hipMemcpyHostToDevice));
    return MMX_OK;
}

// One pass's plan for an axis of n points: the Stockham radices and twiddles of the transform length, and the
// Bluestein tables when n has a prime factor above 7.
int init_axis(mmx_plan_s* p, AxisPlan& ax, int64_t n, int64_t inner, int64_t nouter) {
    ax.N = (int)n;
    ax.inner = inner;
    ax.nouter = nouter;
    ax.blue = 0;
    ax.split = 0;
    if (direct_ok(n)) {
        ax.NT = (int)n;
    } else if (blue_ok(n)) {
        ax.NT = (int)next_smooth7(2 * n - 1);
        ax.blue = 1;
    } else {
        return fail(MMX_EUNSUP, "internal: axis of " + std::to_string(n) + " points in one pass");
    }
    if (!factorize(ax.NT, ax.rad, ax.nrad, p->radix_order)) return fail(MMX_EUNSUP, "internal: transform length radices");
    int rc = p->dtype == MMX_C64 ? make_twiddles<float>(ax.NT, &ax.tw) : make_twiddles<double>(ax.NT, &ax.tw);
    if (!rc && ax.blue) rc = p->dtype == MMX_C64 ? make_bluestein<float>(ax) : make_bluestein<double>(ax);
    return rc;
}

int ensure_scratch(mmx_plan_s* p, size_t bytes) {
    if (p->scratch_bytes >= bytes) return MMX_OK;
    if (p->scratch) MMX_HIP_CHECK(hipFree(p->scratch));
    p->scratch = nullptr;
    p->scratch_bytes = 0;
    MMX_HIP_CHECK(hipMalloc(&p->scratch, bytes));
    p->scratch_bytes = bytes;
    ++p->scratch_gen;
    return MMX_OK;
}

template <typename T>
void fill_geometry(const mmx_plan_s* p, const AxisPlan& ax, PassArgs<T>& a) {
    a.N = ax.N;
    a.NT = ax.NT;
    a.blue = ax.blue;
    a.chirp = reinterpret_cast<const cx<T>*>(ax.chirp);
    a.rp.nrad = ax.nrad;
    int prod = 1;
    for (int r = 0; r < kMaxRadices; ++r) {
        const int R = r < ax.nrad ? ax.rad[r] : 1;
        a.rp.rad[r] = R;
        a.rp.per[r] = make_fdiv(ax.NT / R);
        a.rp.p[r] = make_fdiv(prod);
        prod *= R;
    }
    a.fN = make_fdiv(ax.N);
    a.fn1 = make_fdiv64(p->shape3[1]);
    a.fn2 = make_fdiv64(p->shape3[2]);
    a.inner = ax.inner;
    a.tw = reinterpret_cast<const cx<T>*>(ax.tw);
    a.n0 = p->shape3[0];
    a.n1 = p->shape3[1];
    a.n2 = p->shape3[2];
    a.kvec = reinterpret_cast<const T*>(p->kvec);
    a.k1ofs = p->k1ofs;
    a.k2ofs = p->k2ofs;
    a.nd_data = p->ndim;
    a.inv_ntotal = T(1.0 / double(p->V));
    a.comp_stride = p->V;
    a.axis_orig = ax.orig;
    a.kofs_axis = ax.orig == 0 ? 0 : (ax.orig == 1 ? p->k1ofs : p->k2ofs);
}

// Dynamic LDS beyond 64 KiB has to be enabled per kernel (up to the 160 KiB of a gfx950 CU).
template <auto KERNEL>
void allow_lds(size_t lds) {
    if (lds <= 65536) return;
    static std::atomic<size_t> raised{0};  // per kernel: the largest dynamic LDS size enabled so far
    if (lds > raised.load()) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(KERNEL), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        raised.store(lds);
    }
}

template <int N, int KIND, int NC, int CK>
void launch_reg(const TileCfg& t, const PassArgs<float>& a, hipStream_t stream) {
    constexpr int E = reg_elems(N, NC, KIND);
    if constexpr (KIND == kEpiPro || KIND == kGreenEpiPro) {
        if (a.s_restricted) {
            allow_lds<&reg_pass_kernel<float, N, E, KIND, NC, CK, 1>>(t.lds);
            hipLaunchKernelGGL((reg_pass_kernel<float, N, E, KIND, NC, CK, 1>), dim3((unsigned)t.nblocks),
                               dim3(t.nth), t.lds, stream, a);
            return;
        }
    }
    allow_lds<&reg_pass_kernel<float, N, E, KIND, NC, CK, 0>>(t.lds);
    hipLaunchKernelGGL((reg_pass_kernel<float, N, E, KIND, NC, CK, 0>), dim3((unsigned)t.nblocks), dim3(t.nth),
                       t.lds, stream, a);
}

template <int KIND, int NC, int CK>
int launch_reg_n(int N, const TileCfg& t, const PassArgs<float>& a, hipStream_t stream) {
    switch (N) {
        case 64: launch_reg<64, KIND, NC, CK>(t, a, stream); break;
        case 128: launch_reg<128, KIND, NC, CK>(t, a, stream); break;
        case 256: launch_reg<256, KIND, NC, CK>(t, a, stream); break;
        case 512: launch_reg<512, KIND, NC, CK>(t, a, stream); break;
        case 96: launch_reg<96, KIND, NC, CK>(t, a, stream); break;
        case 192: launch_reg<192, KIND, NC, CK>(t, a, stream); break;
        case 384: launch_reg<384, KIND, NC, CK>(t, a, stream); break;
        case 80: launch_reg<80, KIND, NC, CK>(t, a, stream); break;
        case 160: launch_reg<160, KIND, NC, CK>(t, a, stream); break;
        case 320: launch_reg<320, KIND, NC, CK>(t, a, stream); break;
        case 640: launch_reg<640, KIND, NC, CK>(t, a, stream); break;
        case 768: launch_reg<768, KIND, NC, CK>(t, a, stream); break;
        case 1024: launch_reg<1024, KIND, NC, CK>(t, a, stream); break;
        case 2048: launch_reg<2048, KIND, NC, CK>(t, a, stream); break;
        default: return fail(MMX_EINVAL, "internal: register FFT length");
    }
    return MMX_OK;
}

constexpr bool kind_applies_chi(int kind) {
    return kind == kProFwd || kind == kInvEpi || kind == kAll || kind == kEpiPro || kind == kGreenEpiPro;
}

// Instantiation for the component count and (for passes that apply chi) the chi kind of the launch.
template <int KIND>
int launch_reg_kind(int N, bool three, int chi_kind, const TileCfg& t, const PassArgs<float>& a, hipStream_t stream) {
    if constexpr (KIND == kFwd || KIND == kInv) {  // single-component passes
        if (three) return fail(MMX_EINVAL, "internal: plain pass with three components");
        return launch_reg_n<KIND, 1, 0>(N, t, a, stream);
    } else if constexpr (!kind_applies_chi(KIND)) {
        return three ? launch_reg_n<KIND, 3, 0>(N, t, a, stream) : launch_reg_n<KIND, 1, 0>(N, t, a, stream);
    } else {
        if (!three) {
            if (chi_kind != 1) return fail(MMX_EINVAL, "internal: 3x3 chi with one component");
            return launch_reg_n<KIND, 1, 1>(N, t, a, stream);
        }
        switch (chi_kind) {
            case 1: return launch_reg_n<KIND, 3, 1>(N, t, a, stream);
            case 2: return launch_reg_n<KIND, 3, 2>(N, t, a, stream);
            case 3: return launch_reg_n<KIND, 3, 3>(N, t, a, stream);
            case 4: return launch_reg_n<KIND, 3, 4>(N, t, a, stream);
            default: return fail(MMX_EINVAL, "internal: chi kind");
        }
    }
}

// Launch one pencil pass (see tile_config for `fused`).
template <typename T, int KIND, int E>
void launch_kind_e(const TileCfg& t, const PassArgs<T>& a, hipStream_t stream) {
    if (t.lds > 65536) {
        static size_t raised = 0;  // per instantiation: the largest dynamic LDS size requested so far
        if (t.lds > raised) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&pencil_pass_kernel<T, E, KIND>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)t.lds);
            raised = t.lds;
        }
    }
    hipLaunchKernelGGL((pencil_pass_kernel<T, E, KIND>), dim3((unsigned)t.nblocks), dim3(t.nth), t.lds, stream, a);
}

template <typename T, int KIND>
void launch_kind(const TileCfg& t, const PassArgs<T>& a, hipStream_t stream) {
    constexpr int E = sizeof(T) == 4 ? kElemsF : kElemsD;
    constexpr int EB = sizeof(T) == 4 ? kElemsFBig : kElemsDBig;
    if (t.big) launch_kind_e<T, KIND, EB>(t, a, stream);
    else launch_kind_e<T, KIND, E>(t, a, stream);
}

int launch_pass_reg(const mmx_plan_s* p, const AxisPlan& ax, PassArgs<float> a, int64_t nbatch, bool fused,
                    int kind, hipStream_t stream) {
    TileCfg t;
    const bool rows = ax.inner == 1;
    const int preq = rows ? p->p_rows : (fused ? p->p_green : p->p_strided);
    if (!fused) a.ncomp = 1;
    int rc = reg_tile_config(ax, a.ncomp, fused, nbatch, preq, kind, t);
    if (rc) return rc;
    a.xcd_remap = (p->xcd_remap >> kind) & 1;
    if (!fused) a.bstride = 0;
    a.nouter = t.nouter;
    a.P = t.P;
    a.pad = 0;
    a.ntile_inner = (int)t.ntile_inner;
    a.tiles_per_problem = t.tiles_per_problem;
    a.fntile = make_fdiv64(t.ntile_inner);
    a.ftpp = make_fdiv64(t.tiles_per_problem);
    if (t.nblocks <= 0) return MMX_OK;
    if (t.nblocks > 0x7fffffffLL) return fail(MMX_EINVAL, "grid too large");
    const bool three = fused && a.ncomp == 3;
    const int ck = a.chi_kind;
    switch (kind) {
        case kFwd: rc = launch_reg_kind<kFwd>(ax.N, three, ck, t, a, stream); break;
        case kInv: rc = launch_reg_kind<kInv>(ax.N, three, ck, t, a, stream); break;
        case kProFwd: rc = launch_reg_kind<kProFwd>(ax.N, three, ck, t, a, stream); break;
        case kGreen: rc = launch_reg_kind<kGreen>(ax.N, three, ck, t, a, stream); break;
        case kInvEpi: rc = launch_reg_kind<kInvEpi>(ax.N, three, ck, t, a, stream); break;
        case kAll: rc = launch_reg_kind<kAll>(ax.N, three, ck, t, a, stream); break;
        case kEpiPro: rc = launch_reg_kind<kEpiPro>(ax.N, three, ck, t, a, stream); break;
        case kGreenEpiPro: rc = launch_reg_kind<kGreenEpiPro>(ax.N, three, ck, t, a, stream); break;
        default: return fail(MMX_EINVAL, "internal: unknown pass kind");
    }
    if (rc) return rc;
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

template <typename T>
int launch_pass(const mmx_plan_s* p, const AxisPlan& ax, PassArgs<T> a, int64_t nbatch, bool fused, int kind,
                hipStream_t stream) {
    if (ax.split) return fail(MMX_EINVAL, "internal: a split axis has no single pass");
    if ((ax.blue || a.tw4) && kind != kFwd && kind != kInv)
        return fail(MMX_EINVAL, "internal: Bluestein and four-step transforms run in plain passes only");
    a.bhat = reinterpret_cast<const cx<T>*>(kind == kInv ? ax.bhat_i : ax.bhat_f);
    if constexpr (sizeof(T) == 4) {
        if (p->reg_fft && reg_supported(ax) && !a.tw4) return launch_pass_reg(p, ax, a, nbatch, fused, kind, stream);
    }
    TileCfg t;
    const bool rows = ax.inner == 1;
    const int preq = rows ? p->p_rows : (fused ? p->p_green : p->p_strided);
    int rc = tile_config<T>(ax, a.ncomp, fused, nbatch, p->lds_max, t, preq, p->lds_budget,
                            (p->big_elems >> kind) & 1);
    if (rc) return rc;
    a.xcd_remap = (p->xcd_remap >> kind) & 1;
    if (!fused) {
        a.ncomp = 1;
        a.bstride = 0;
    }
    a.nouter = t.nouter;
    a.P = t.P;
    a.pad = t.pad;
    a.ntile_inner = (int)t.ntile_inner;
    a.tiles_per_problem = t.tiles_per_problem;
    a.fP = make_fdiv(t.P);
    a.fncomp = make_fdiv(a.ncomp > 0 ? a.ncomp : 1);
    a.fntile = make_fdiv64(t.ntile_inner);
    a.ftpp = make_fdiv64(t.tiles_per_problem);
    if ((int64_t)a.ncomp * t.P * (ax.NT + t.pad) >= (1 << 22)) return fail(MMX_EUNSUP, "internal: tile index range");
    if (t.nblocks <= 0) return MMX_OK;
    if (t.nblocks > 0x7fffffffLL) return fail(MMX_EINVAL, "grid too large");
    switch (kind) {
        case kFwd: launch_kind<T, kFwd>(t, a, stream); break;
        case kInv: launch_kind<T, kInv>(t, a, stream); break;
        case kProFwd: launch_kind<T, kProFwd>(t, a, stream); break;
        case kGreen: launch_kind<T, kGreen>(t, a, stream); break;
        case kInvEpi: launch_kind<T, kInvEpi>(t, a, stream); break;
        case kAll: launch_kind<T, kAll>(t, a, stream); break;
        case kEpiPro: launch_kind<T, kEpiPro>(t, a, stream); break;
        case kGreenEpiPro: launch_kind<T, kGreenEpiPro>(t, a, stream); break;
        default: return fail(MMX_EINVAL, "internal: unknown pass kind");
    }
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

template <typename T>
PassArgs<T> blank_args() {
    PassArgs<T> a;
    memset(&a, 0, sizeof(a));
    a.scale = T(1);
    a.dir = -1;
    return a;
}

int ensure_fft_scratch(mmx_plan_s* p, size_t bytes) {
    if (p->fft_scratch_bytes >= bytes) return MMX_OK;
    if (p->fft_scratch) MMX_HIP_CHECK(hipFree(p->fft_scratch));
    p->fft_scratch = nullptr;
    p->fft_scratch_bytes = 0;
    MMX_HIP_CHECK(hipMalloc(&p->fft_scratch, bytes));
    p->fft_scratch_bytes = bytes;
    ++p->scratch_gen;
    return MMX_OK;
}

bool has_split(const mmx_plan_s* p) {
    for (int q = 0; q < p->naxes; ++q)
        if (p->ax[q].split) return true;
    return false;
}

// FFT of `nfields` consecutive fields over every axis, innermost first. scale_inverse: the inverse transform's 1/V
// goes on the last pass (the unfused Born iteration folds it into its Green's operator instead).
template <typename T>
int fft_passes(mmx_plan_s* p, const void* in, void* out, int64_t nfields, int direction, bool scale_inverse,
               hipStream_t stream) {
    if (nfields <= 0) return MMX_OK;
    if (p->naxes == 0) {
        if (in != out)
            MMX_HIP_CHECK(hipMemcpyAsync(out, in, nfields * p->V * sizeof(cx<T>), hipMemcpyDeviceToDevice, stream));
        return MMX_OK;
    }
    if (has_split(p)) {
        const int rc = ensure_fft_scratch(p, (size_t)nfields * p->V * sizeof(cx<T>));
        if (rc) return rc;
    }
    const int kind = direction < 0 ? kFwd : kInv;
    const void* src = in;
    for (int q = p->naxes - 1; q >= 0; --q) {
        const AxisPlan& ax = p->ax[q];
        const T scale = (direction > 0 && q == 0 && scale_inverse) ? T(1.0 / double(p->V)) : T(1);
        if (!ax.split) {
            PassArgs<T> a = blank_args<T>();
            fill_geometry<T>(p, ax, a);
            a.src = reinterpret_cast<const cx<T>*>(src);
            a.dst = reinterpret_cast<cx<T>*>(out);
            a.dir = direction < 0 ? -1 : 1;
            a.scale = scale;
            int rc = launch_pass<T>(p, ax, a, nfields, false, kind, stream);
            if (rc) return rc;
        } else {
            // four-step: length-N1 FFTs along n1, then the twiddled length-N2 FFTs along n2 into the scratch buffer,
            // then the transposition of (k1, k2) to the natural order k = k1 + N1 k2
            cx<T>* scr = reinterpret_cast<cx<T>*>(p->fft_scratch);
            PassArgs<T> a = blank_args<T>();
            fill_geometry<T>(p, p->sub[q][0], a);
            a.src = reinterpret_cast<const cx<T>*>(src);
            a.dst = reinterpret_cast<cx<T>*>(out);
            a.dir = direction < 0 ? -1 : 1;
            int rc = launch_pass<T>(p, p->sub[q][0], a, nfields, false, kind, stream);
            if (rc) return rc;
            PassArgs<T> b = blank_args<T>();
            fill_geometry<T>(p, p->sub[q][1], b);
            b.src = reinterpret_cast<const cx<T>*>(out);
            b.dst = scr;
            b.dir = a.dir;
            b.scale = scale;
            b.tw4 = 1;
            b.tw4tab = reinterpret_cast<const cx<T>*>(ax.tw4);
            b.ftw4 = make_fdiv64(ax.N1);
            rc = launch_pass<T>(p, p->sub[q][1], b, nfields, false, kind, stream);
            if (rc) return rc;
            const int64_t outer = nfields * ax.nouter;
            int64_t nblk = ax.inner >= 64 ? (outer * ax.N * ax.inner + 255) / 256
                                          : outer * ax.inner * ((ax.N1 + 31) / 32) * ((ax.N2 + 31) / 32);
            if (nblk > 65536) nblk = 65536;
            if (nblk < 1) nblk = 1;
            hipLaunchKernelGGL((axis_transpose_kernel<T>), dim3((unsigned)nblk), dim3(256), 0, stream, scr,
                               reinterpret_cast<cx<T>*>(out), outer, (int64_t)ax.N1, (int64_t)ax.N2, ax.inner);
            MMX_HIP_CHECK(hipGetLastError());
        }
        src = out;
    }
    return MMX_OK;
}

template <typename T>
int fft_impl(mmx_plan_s* p, const void* in, void* out, int64_t nbatch, int direction, hipStream_t stream) {
    return fft_passes<T>(p, in, out, nbatch, direction, true, stream);
}

// Tile configuration of the pass that writes the norm partials (same choice as launch_pass makes).
template <typename T>
int last_pass_tiles(const mmx_plan_s* p, const AxisPlan& inner, int C, int64_t nbatch, int kind, TileCfg& tl) {
    if (sizeof(T) == 4 && p->reg_fft && reg_supported(inner))
        return reg_tile_config(inner, C, true, nbatch, p->p_rows, kind, tl);
    return tile_config<T>(inner, C, true, nbatch, p->lds_max, tl, p->p_rows, p->lds_budget, (p->big_elems >> kind) & 1);
}

// norms[2b + {0,1}] = sums of the last pass's per-tile partials of problem b, in tile order.
int reduce_norms(const double* partials, int64_t tiles, int64_t nbatch, double* norms, hipStream_t stream) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)nbatch), dim3(1024), 0, stream, partials, tiles, norms);
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

template <typename T>
int born_check(const mmx_plan_s* p, int C, int chi_kind) {
    if (C != 1 && C != 3) return fail(MMX_EINVAL, "ncomp must be 1 or 3");
    if (chi_kind < 1 || chi_kind > 4) return fail(MMX_EINVAL, "chi_kind must be 1, 2, 3 or 4");
    if (chi_kind >= 2 && C != 3) return fail(MMX_EINVAL, "3x3 chi needs 3 components");
    if (p->naxes == 0) return fail(MMX_EUNSUP, "grid without any axis longer than 1");
    return MMX_OK;
}

template <typename T>
PassArgs<T> born_base(const mmx_plan_s* p, const void* E, void* Eout, const void* s, const void* chi, int C,
                      int chi_kind, const void* coef, int64_t s_bstride, int64_t chi_bstride, double* partials,
                      double* norms = nullptr) {
    PassArgs<T> base = blank_args<T>();
    base.norms = norms;
    base.C = C;
    base.chi_kind = chi_kind;
    base.E = reinterpret_cast<const cx<T>*>(E);
    base.chi = reinterpret_cast<const cx<T>*>(chi);
    base.s = reinterpret_cast<const cx<T>*>(s);
    base.E_bstride = (int64_t)C * p->V;
    base.s_bstride = s_bstride;
    base.chi_bstride = chi_bstride;
    base.Eout = reinterpret_cast<cx<T>*>(Eout);
    base.partials = partials;
    base.coef = reinterpret_cast<const Coef<T>*>(coef);
    base.vectorial = (C == 3);
    base.s_restricted = 0;
    for (int a = 0; a < 3; ++a) {
        base.s_lo[a] = (int)p->s_lo[a];
        base.s_hi[a] = (int)p->s_hi[a];
        if (p->s_lo[a] != 0 || p->s_hi[a] != p->shape3[a]) base.s_restricted = 1;
    }
    return base;
}

// ---- the unfused iteration: voxel-wise kernels around whole-grid FFTs (axes that need Bluestein's algorithm or the
// four-step split, or fused tiles that do not fit in LDS) ----

// Whether the fused pass schedule serves `C` components on this plan (cached per plan; tuning parameters reset it).
template <typename T>
bool fused_schedule_fits(mmx_plan_s* p, int C) {
    int& f = p->fused_ok[C == 3 ? 1 : 0];
    if (f >= 0) return f == 1;
    bool ok = true;
    for (int q = 0; q < p->naxes; ++q)
        if (p->ax[q].blue || p->ax[q].split) ok = false;
    if (ok && p->naxes > 0) {
        const std::string saved = g_last_error;
        const int d = p->naxes;
        TileCfg t;
        const int kinds1[] = {kAll, kGreenEpiPro};
        const int kinds3[] = {kProFwd, kInvEpi, kEpiPro};
        const int* kinds = d == 1 ? kinds1 : kinds3;
        const int nk = d == 1 ? 2 : 3;
        for (int i = 0; i < nk && ok; ++i)
            if (last_pass_tiles<T>(p, p->ax[d - 1], C, 1, kinds[i], t)) ok = false;
        if (ok && d >= 2) {
            int g = p->g_axis < 0 ? 0 : p->g_axis;
            if (g > d - 2) g = d - 2;
            const AxisPlan& ax = p->ax[g];
            const int rc = (sizeof(T) == 4 && p->reg_fft && reg_supported(ax))
                               ? reg_tile_config(ax, C, true, 1, p->p_green, kGreen, t)
                               : tile_config<T>(ax, C, true, 1, p->lds_max, t, p->p_green, p->lds_budget,
                                                (p->big_elems >> kGreen) & 1);
            if (rc) ok = false;
        }
        g_last_error = saved;
    }
    f = ok ? 1 : 0;
    return ok;
}

int pointwise_blocks(int64_t V) {
    int64_t g = (V + 255) / 256;
    if (g > 1024) g = 1024;
    return (int)(g < 1 ? 1 : g);
}

// dst = chi_eff E + s_eff per voxel (E from Eout when src_from_out: the speculative next prologue).
template <typename T>
int unfused_prologue(mmx_plan_s* p, PassArgs<T> a, void* dst, int64_t nbatch, int src_from_out, hipStream_t stream) {
    if (nbatch > 65535) return fail(MMX_EINVAL, "more than 65535 problems per call on a grid of this shape");
    fill_geometry<T>(p, p->ax[0], a);
    a.dst = reinterpret_cast<cx<T>*>(dst);
    a.bstride = (int64_t)a.C * p->V;
    a.comp_stride = p->V;
    hipLaunchKernelGGL((born_pro_kernel<T>), dim3(pointwise_blocks(p->V), (unsigned)nbatch), dim3(256), 0, stream, a,
                       src_from_out);
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

// From W = chi_eff E + s_eff: W <- G W (forward FFT, Green's operator, inverse FFT), then E' = E + chi_eff (W - E)
// with the norm partials (pointwise_blocks(V) per problem).
template <typename T>
int unfused_core(mmx_plan_s* p, PassArgs<T> a, void* W, int64_t nbatch, hipStream_t stream) {
    if (nbatch > 65535) return fail(MMX_EINVAL, "more than 65535 problems per call on a grid of this shape");
    const int64_t nf = nbatch * a.C;
    int rc = fft_passes<T>(p, W, W, nf, -1, false, stream);
    if (rc) return rc;
    fill_geometry<T>(p, p->ax[0], a);
    a.bstride = (int64_t)a.C * p->V;
    a.comp_stride = p->V;
    a.vectorial = (a.C == 3);
    a.src = reinterpret_cast<const cx<T>*>(W);
    a.dst = reinterpret_cast<cx<T>*>(W);
    const dim3 grid(pointwise_blocks(p->V), (unsigned)nbatch);
    hipLaunchKernelGGL((green_pointwise_kernel<T>), grid, dim3(256), 0, stream, a);
    MMX_HIP_CHECK(hipGetLastError());
    rc = fft_passes<T>(p, W, W, nf, +1, false, stream);
    if (rc) return rc;
    hipLaunchKernelGGL((born_epi_kernel<T>), grid, dim3(256), 0, stream, a);
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

// W = forward FFT along the innermost axis of chi_eff E + s_eff (the first pass of an iteration).
template <typename T>
int prologue_impl(mmx_plan_s* p, const void* E, void* W, const void* s, const void* chi, int C, int chi_kind,
                  const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride, hipStream_t stream,
                  const unsigned* state = nullptr, const void* E_alt = nullptr) {
    int rc = born_check<T>(p, C, chi_kind);
    if (rc) return rc;
    if (nbatch <= 0) return MMX_OK;
    const AxisPlan& inner = p->ax[p->naxes - 1];
    PassArgs<T> a = born_base<T>(p, E, const_cast<void*>(E_alt), s, chi, C, chi_kind, coef, s_bstride, chi_bstride,
                                 nullptr);
    a.state = state;
    a.only_stale = state != nullptr;
    if (!fused_schedule_fits<T>(p, C)) return unfused_prologue<T>(p, a, W, nbatch, 0, stream);  // W = chi E + s
    fill_geometry<T>(p, inner, a);
    a.ncomp = C;
    a.bstride = (int64_t)C * p->V;
    a.src = reinterpret_cast<cx<T>*>(W);
    a.dst = reinterpret_cast<cx<T>*>(W);
    return launch_pass<T>(p, inner, a, nbatch, true, kProFwd, stream);
}

// One iteration from W (first pass already applied to E): remaining forward passes, Green, inverse passes, then
// the fused last pass: Eout = E + chi_eff (W - E), norms, and W <- first pass of the next iteration on Eout.
// Partial-sum buffer `slot` (0 or 1) of mmx_born_iterate_split, grown to `bytes`; its event is created with it.
int ensure_slot(mmx_plan_s* p, int slot, size_t bytes) {
    if (!p->slot_done[slot]) MMX_HIP_CHECK(hipEventCreateWithFlags(&p->slot_done[slot], hipEventDisableTiming));
    if (p->slot_bytes[slot] >= bytes) return MMX_OK;
    if (p->slot_partials[slot]) MMX_HIP_CHECK(hipFree(p->slot_partials[slot]));
    p->slot_partials[slot] = nullptr;
    p->slot_bytes[slot] = 0;
    MMX_HIP_CHECK(hipMalloc(&p->slot_partials[slot], bytes));
    p->slot_bytes[slot] = bytes;
    ++p->scratch_gen;
    return MMX_OK;
}

// Where an iteration's norm partials go and where they are reduced: the plan's scratch and the launch stream, or
// (norm_stream set) partial slot `slot` and norm_stream, once the launch stream has finished the passes.
struct NormRoute {
    hipStream_t norm_stream = nullptr;
    int slot = 0;
    double* norms_host = nullptr;  // pinned host copy of the norms (norm_stream only), or none
};

int partials_for(mmx_plan_s* p, const NormRoute& r, size_t bytes, double** partials) {
    int rc = r.norm_stream ? ensure_slot(p, r.slot, bytes) : ensure_scratch(p, bytes);
    if (rc) return rc;
    *partials = reinterpret_cast<double*>(r.norm_stream ? p->slot_partials[r.slot] : p->scratch);
    return MMX_OK;
}

int reduce_routed(mmx_plan_s* p, const NormRoute& r, const double* partials, int64_t tiles, int64_t nbatch,
                  double* norms, hipStream_t stream) {
    if (!r.norm_stream) return reduce_norms(partials, tiles, nbatch, norms, stream);
    MMX_HIP_CHECK(hipEventRecord(p->slot_done[r.slot], stream));
    MMX_HIP_CHECK(hipStreamWaitEvent(r.norm_stream, p->slot_done[r.slot], 0));
    const int rc = reduce_norms(partials, tiles, nbatch, norms, r.norm_stream);
    if (rc || !r.norms_host) return rc;
    MMX_HIP_CHECK(hipMemcpyAsync(r.norms_host, norms, (size_t)(2 * nbatch) * sizeof(double), hipMemcpyDeviceToHost,
                                 r.norm_stream));
    return MMX_OK;
}

template <typename T>
int iterate_impl(mmx_plan_s* p, const void* E, void* Eout, void* Wv, const void* s, const void* chi, int C,
                 int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride,
                 double* norms, hipStream_t stream, const unsigned* state = nullptr, NormRoute route = NormRoute()) {
    int rc = born_check<T>(p, C, chi_kind);
    if (rc) return rc;
    if (nbatch <= 0) return MMX_OK;
    if (!fused_schedule_fits<T>(p, C)) {
        const int64_t tiles = pointwise_blocks(p->V);
        double* partials = nullptr;
        rc = partials_for(p, route, (size_t)(2 * tiles * nbatch) * sizeof(double), &partials);
        if (rc) return rc;
        PassArgs<T> a = born_base<T>(p, E, Eout, s, chi, C, chi_kind, coef, s_bstride, chi_bstride, partials, norms);
        a.state = state;
        rc = unfused_core<T>(p, a, Wv, nbatch, stream);
        if (!rc) rc = unfused_prologue<T>(p, a, Wv, nbatch, 1, stream);
        if (!rc) rc = reduce_routed(p, route, partials, tiles, nbatch, norms, stream);
        return rc;
    }
    const int d = p->naxes;
    const AxisPlan& inner = p->ax[d - 1];
    const int last_kind = d == 1 ? kGreenEpiPro : kEpiPro;
    TileCfg tl;
    rc = last_pass_tiles<T>(p, inner, C, nbatch, last_kind, tl);
    if (rc) return rc;
    const int64_t tiles_last = tl.tiles_per_problem;
    double* partials = nullptr;
    rc = partials_for(p, route, (size_t)(2 * tiles_last * nbatch) * sizeof(double), &partials);
    if (rc) return rc;
    PassArgs<T> base = born_base<T>(p, E, Eout, s, chi, C, chi_kind, coef, s_bstride, chi_bstride, partials, norms);
    base.state = state;
    cx<T>* W = reinterpret_cast<cx<T>*>(Wv);
    if (d >= 2) {
        int g = p->g_axis;
        if (g < 0) g = 0;
        if (g > d - 2) g = d - 2;
        int others[3], nothers = 0;
        for (int q = d - 2; q >= 0; --q)
            if (q != g) others[nothers++] = q;
        for (int q = 0; q < nothers; ++q) {
            PassArgs<T> a = base;
            fill_geometry<T>(p, p->ax[others[q]], a);
            a.frpp = make_fdiv64((int64_t)C * p->ax[others[q]].nouter);
            a.src = W;
            a.dst = W;
            rc = launch_pass<T>(p, p->ax[others[q]], a, nbatch * C, false, kFwd, stream);
            if (rc) return rc;
        }
        {
            PassArgs<T> a = base;
            fill_geometry<T>(p, p->ax[g], a);
            a.ncomp = C;
            a.bstride = (int64_t)C * p->V;
            a.src = W;
            a.dst = W;
            rc = launch_pass<T>(p, p->ax[g], a, nbatch, true, kGreen, stream);
            if (rc) return rc;
        }
        for (int q = nothers - 1; q >= 0; --q) {
            PassArgs<T> a = base;
            fill_geometry<T>(p, p->ax[others[q]], a);
            a.frpp = make_fdiv64((int64_t)C * p->ax[others[q]].nouter);
            a.src = W;
            a.dst = W;
            rc = launch_pass<T>(p, p->ax[others[q]], a, nbatch * C, false, kInv, stream);
            if (rc) return rc;
        }
    }
    {
        PassArgs<T> a = base;
        fill_geometry<T>(p, inner, a);
        a.ncomp = C;
        a.bstride = (int64_t)C * p->V;
        a.src = W;
        a.dst = W;
        rc = launch_pass<T>(p, inner, a, nbatch, true, last_kind, stream);
        if (rc) return rc;
    }
    return reduce_routed(p, route, partials, tiles_last, nbatch, norms, stream);
}

// fixed_partials != nullptr: the per-tile partial sums go there (at most fixed_cap bytes) instead of the plan's
// scratch buffer, and nothing is allocated.
template <typename T>
int born_impl(mmx_plan_s* p, const void* E, void* Eout, const void* s, const void* chi, int C, int chi_kind,
              const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride, double* norms,
              hipStream_t stream, double* fixed_partials = nullptr, size_t fixed_cap = 0) {
    if (C != 1 && C != 3) return fail(MMX_EINVAL, "ncomp must be 1 or 3");
    if (chi_kind < 1 || chi_kind > 4) return fail(MMX_EINVAL, "chi_kind must be 1, 2, 3 or 4");
    if (chi_kind >= 2 && C != 3) return fail(MMX_EINVAL, "3x3 chi needs 3 components");
    if (nbatch <= 0) return MMX_OK;
    if (p->naxes == 0) return fail(MMX_EUNSUP, "grid without any axis longer than 1");
    const int d = p->naxes;
    const AxisPlan& inner = p->ax[d - 1];
    const bool fused = fused_schedule_fits<T>(p, C);
    if (!fused && fixed_partials != nullptr && has_split(p))
        return fail(MMX_EINVAL, "this grid's four-step FFT needs a scratch buffer: use mmx_born_step");
    // partial sums: one pair per tile of the last pass
    TileCfg tl;
    int rc = MMX_OK;
    if (fused) rc = last_pass_tiles<T>(p, inner, C, nbatch, d == 1 ? kAll : kInvEpi, tl);
    else tl.tiles_per_problem = pointwise_blocks(p->V);
    if (rc) return rc;
    const int64_t tiles_last = tl.tiles_per_problem;
    const size_t partial_bytes = (size_t)(2 * tiles_last * nbatch) * sizeof(double);
    double* partials;
    if (fixed_partials != nullptr) {
        if (partial_bytes > fixed_cap)
            return fail(MMX_EINVAL, "launch needs " + std::to_string(partial_bytes) +
                                        " bytes of norm partials, more than the " + std::to_string(fixed_cap) +
                                        " bytes of the static scratch array (fewer problems per call, or "
                                        "mmx_born_step)");
        partials = fixed_partials;
    } else {
        rc = ensure_scratch(p, partial_bytes);
        if (rc) return rc;
        partials = reinterpret_cast<double*>(p->scratch);
    }
    PassArgs<T> base = born_base<T>(p, E, Eout, s, chi, C, chi_kind, coef, s_bstride, chi_bstride, partials, norms);
    cx<T>* W = reinterpret_cast<cx<T>*>(Eout);
    if (!fused) {  // Eout is the work buffer: x = chi E + s, G x, then E' = E + chi (G x - E) voxel by voxel
        rc = unfused_prologue<T>(p, base, W, nbatch, 0, stream);
        if (!rc) rc = unfused_core<T>(p, base, W, nbatch, stream);
        if (!rc) rc = reduce_norms(partials, tiles_last, nbatch, norms, stream);
        return rc;
    }

    if (d == 1) {
        PassArgs<T> a = base;
        fill_geometry<T>(p, inner, a);
        a.prologue = 1;
        a.green = 1;
        a.epilogue = 1;
        a.ncomp = C;
        a.bstride = (int64_t)C * p->V;
        a.src = W;
        a.dst = W;
        rc = launch_pass<T>(p, inner, a, nbatch, true, kAll, stream);
        if (rc) return rc;
    } else {
        // the Green pass runs on axis g (0 = outermost); the innermost axis carries prologue and epilogue
        int g = p->g_axis;
        if (g < 0) g = 0;
        if (g > d - 2) g = d - 2;
        int others[3], nothers = 0;
        for (int q = d - 2; q >= 0; --q)
            if (q != g) others[nothers++] = q;
        {   // pass 1: prologue + forward FFT along the innermost axis
            PassArgs<T> a = base;
            fill_geometry<T>(p, inner, a);
            a.ncomp = C;
            a.bstride = (int64_t)C * p->V;
            a.src = W;
            a.dst = W;
            rc = launch_pass<T>(p, inner, a, nbatch, true, kProFwd, stream);
            if (rc) return rc;
        }
        for (int q = 0; q < nothers; ++q) {  // remaining axes, forward
            PassArgs<T> a = base;
            fill_geometry<T>(p, p->ax[others[q]], a);
            a.src = W;
            a.dst = W;
            rc = launch_pass<T>(p, p->ax[others[q]], a, nbatch * C, false, kFwd, stream);
            if (rc) return rc;
        }
        {   // axis g: forward FFT, Green's operator, inverse FFT
            PassArgs<T> a = base;
            fill_geometry<T>(p, p->ax[g], a);
            a.ncomp = C;
            a.bstride = (int64_t)C * p->V;
            a.src = W;
            a.dst = W;
            rc = launch_pass<T>(p, p->ax[g], a, nbatch, true, kGreen, stream);
            if (rc) return rc;
        }
        for (int q = nothers - 1; q >= 0; --q) {  // remaining axes, inverse (1/N folded into G)
            PassArgs<T> a = base;
            fill_geometry<T>(p, p->ax[others[q]], a);
            a.src = W;
            a.dst = W;
            rc = launch_pass<T>(p, p->ax[others[q]], a, nbatch * C, false, kInv, stream);
            if (rc) return rc;
        }
        {   // last pass: inverse FFT along the innermost axis + Born update + norms
            PassArgs<T> a = base;
            fill_geometry<T>(p, inner, a);
            a.ncomp = C;
            a.bstride = (int64_t)C * p->V;
            a.src = W;
            a.dst = W;
            rc = launch_pass<T>(p, inner, a, nbatch, true, kInvEpi, stream);
            if (rc) return rc;
        }
    }
    return reduce_norms(partials, tiles_last, nbatch, norms, stream);
}

// One Born iteration of every unfinished problem of a batch, then the accept/reject decision and the stop test on
// the device. EA and EB are the two field buffers (problem b's current field is in EB when its parity bit is set).
template <typename T>
int batch_step_impl(mmx_plan_s* p, void* EA, void* EB, void* W, const void* s, const void* chi, int C, int chi_kind,
                    void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride, unsigned* state,
                    double* track, double* norms, double tol, int64_t max_iter, hipStream_t stream) {
    if (nbatch <= 0) return MMX_OK;
    int rc = prologue_impl<T>(p, EA, W, s, chi, C, chi_kind, coef, nbatch, s_bstride, chi_bstride, stream, state, EB);
    if (rc) return rc;
    rc = iterate_impl<T>(p, EA, EB, W, s, chi, C, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms, stream,
                         state);
    if (rc) return rc;
    const int nblk = (int)((nbatch + 255) / 256 < 1024 ? (nbatch + 255) / 256 : 1024);
    hipLaunchKernelGGL((batch_control_kernel<T>), dim3(nblk), dim3(256), 0, stream, nbatch, norms, state,
                       reinterpret_cast<Coef<T>*>(coef), track, tol, double(max_iter));
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

// One magnetic Born iteration (see mmx_magnetic.hpp for the sequence).
// ---- fused magnetic iteration (mmx_magfused.hpp): complex64, first and last axes of 64..1024 points (powers of two)
#ifndef MMX_MAG_FUSED
#define MMX_MAG_FUSED 1
#endif
inline bool mag_fused_length(const AxisPlan& ax) {
    const int n = ax.N;
    return (n == 64 || n == 128 || n == 256 || n == 512 || n == 1024) && reg_supported(ax);
}

// Dynamic LDS of the two-group OP 1 mid pass: twiddles, an exchange buffer per group, the coupling buffer.
constexpr size_t mag_split_lds(int N, int E, int P) {
    return (size_t)(N + 2 * (size_t)P * padded_len(N, E) + 3 * (size_t)P * N) * sizeof(cx<float>);
}
// ... and of the two-group pw pass STAGE 2, whose coupling buffer holds six components.
constexpr size_t mag_pw2_lds(int N, int E, int P) {
    return (size_t)(N + 2 * (size_t)P * padded_len(N, E) + 6 * (size_t)P * N) * sizeof(cx<float>);
}

template <int N, int OP>
void launch_mag_mid(const TileCfg& t, const PassArgs<float>& g, const MagArgs<float>& a, const cx<float>* src,
                    cx<float>* dst, hipStream_t st) {
    constexpr int E = reg_elems(N, 3, kGreen);
    if constexpr (OP == 1) {
        const size_t lds = mag_split_lds(N, E, t.P);
        allow_lds<&mag_mid_split_kernel<float, N, E>>(lds);
        hipLaunchKernelGGL((mag_mid_split_kernel<float, N, E>), dim3((unsigned)t.nblocks), dim3(2 * t.nth), lds, st,
                           g, a, src, dst);
    } else {
        allow_lds<&mag_mid_kernel<float, N, E, 3, OP>>(t.lds);
        hipLaunchKernelGGL((mag_mid_kernel<float, N, E, 3, OP>), dim3((unsigned)t.nblocks), dim3(t.nth), t.lds, st, g,
                           a, src, dst);
    }
}

template <int N, int STAGE>
void launch_mag_pw(const TileCfg& t, const PassArgs<float>& g, const MagArgs<float>& a, const cx<float>* src,
                   cx<float>* dst, hipStream_t st) {
    constexpr int E = reg_elems(N, 3, kEpiPro);
    if constexpr (STAGE == 1) {
        const size_t lds = (size_t)(N + 2 * (size_t)t.P * padded_len(N, E)) * sizeof(cx<float>);
        allow_lds<&mag_pw1_split_kernel<float, N, E>>(lds);
        hipLaunchKernelGGL((mag_pw1_split_kernel<float, N, E>), dim3((unsigned)t.nblocks), dim3(2 * t.nth), lds, st,
                           g, a, src, dst);
    } else if constexpr (STAGE == 2) {
        const size_t lds = mag_pw2_lds(N, E, t.P);
        allow_lds<&mag_pw2_split_kernel<float, N, E>>(lds);
        hipLaunchKernelGGL((mag_pw2_split_kernel<float, N, E>), dim3((unsigned)t.nblocks), dim3(2 * t.nth), lds, st,
                           g, a, src, dst);
    } else {
        allow_lds<&mag_pw_kernel<float, N, E, STAGE>>(t.lds);
        hipLaunchKernelGGL((mag_pw_kernel<float, N, E, STAGE>), dim3((unsigned)t.nblocks), dim3(t.nth), t.lds, st, g,
                           a, src, dst);
    }
}

// step 0..2: mid pass with k-space product OP = step; 3..5: pw pass STAGE = step - 2
template <int N>
void launch_mag_step(int step, const TileCfg& t, const PassArgs<float>& g, const MagArgs<float>& a,
                     const cx<float>* src, cx<float>* dst, hipStream_t st) {
    switch (step) {
        case 0: launch_mag_mid<N, 0>(t, g, a, src, dst, st); break;
        case 1: launch_mag_mid<N, 1>(t, g, a, src, dst, st); break;
        case 2: launch_mag_mid<N, 2>(t, g, a, src, dst, st); break;
        case 3: launch_mag_pw<N, 1>(t, g, a, src, dst, st); break;
        case 4: launch_mag_pw<N, 2>(t, g, a, src, dst, st); break;
        default: launch_mag_pw<N, 3>(t, g, a, src, dst, st); break;
    }
}

// Geometry and tiles of a fused magnetic pass along `ax` (mid passes: the first axis, strided tiles of three
// components as the Green pass; pw passes: the last axis, row tiles as the update pass).
int mag_geometry(const mmx_plan_s* p, const AxisPlan& ax, bool mid, PassArgs<float>& g, TileCfg& t,
                 bool split = false) {
    g = blank_args<float>();
    fill_geometry<float>(p, ax, g);
    const int kind = mid ? kGreen : kEpiPro;
    int rc = reg_tile_config(ax, 3, true, 1, mid ? p->p_green : p->p_rows, kind, t);
    if (rc) return rc;
    if (split) {  // two thread groups per tile: at most 512 threads and a CU's LDS
        const int TP = reg_threads_per_pencil(ax.N, 3, kind);
        int P = t.P;
        while (P > 1 && (2 * P * TP > 512 || mag_pw2_lds(ax.N, ax.N / TP, P) > p->lds_max)) P >>= 1;
        if (P != t.P && (rc = reg_tile_config(ax, 3, true, 1, P, kind, t))) return rc;
        if (2 * t.nth > 512) return fail(MMX_EINVAL, "internal: magnetic tile too wide");
    }
    g.nouter = t.nouter;
    g.P = t.P;
    g.ntile_inner = (int)t.ntile_inner;
    g.tiles_per_problem = t.tiles_per_problem;
    g.fntile = make_fdiv64(t.ntile_inner);
    g.ftpp = make_fdiv64(t.tiles_per_problem);
    g.xcd_remap = (p->xcd_remap >> kind) & 1;
    g.scale = mid ? 1.0f : float(1.0 / double(p->V));
    if (t.nblocks <= 0 || t.nblocks > 0x7fffffffLL) return fail(MMX_EINVAL, "internal: magnetic pass grid");
    return MMX_OK;
}

int mag_launch(const mmx_plan_s* p, int step, const MagArgs<float>& a, const cx<float>* src, cx<float>* dst,
               hipStream_t st) {
    const bool mid = step < 3;
    const AxisPlan& ax = mid ? p->ax[0] : p->ax[p->naxes - 1];
    PassArgs<float> g;
    TileCfg t;
    int rc = mag_geometry(p, ax, mid, g, t, step == 1 || step == 3 || step == 4);
    if (rc) return rc;
    switch (ax.N) {
        case 64: launch_mag_step<64>(step, t, g, a, src, dst, st); break;
        case 128: launch_mag_step<128>(step, t, g, a, src, dst, st); break;
        case 256: launch_mag_step<256>(step, t, g, a, src, dst, st); break;
        case 512: launch_mag_step<512>(step, t, g, a, src, dst, st); break;
        case 1024: launch_mag_step<1024>(step, t, g, a, src, dst, st); break;
        default: return fail(MMX_EINVAL, "internal: fused magnetic pass length");
    }
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

// One plain pass along the middle axis of a 3D grid (nfields fields, unscaled).
int mag_middle_pass(const mmx_plan_s* p, const cx<float>* src, cx<float>* dst, int64_t nfields, int direction,
                    hipStream_t st) {
    PassArgs<float> a = blank_args<float>();
    fill_geometry<float>(p, p->ax[1], a);
    a.src = src;
    a.dst = dst;
    a.dir = direction;
    return launch_pass<float>(p, p->ax[1], a, nfields, false, direction < 0 ? kFwd : kInv, st);
}

bool mag_fused_ok(const mmx_plan_s* p) {
    if (!MMX_MAG_FUSED || p->dtype != MMX_C64 || !p->reg_fft || p->naxes < 2 || p->naxes > 3) return false;
    if (!mag_fused_length(p->ax[0]) || !mag_fused_length(p->ax[p->naxes - 1])) return false;
    return p->naxes == 2 || !p->ax[1].split;
}

// The iteration with the fused passes (see mmx_magfused.hpp for the schedule).
int magnetic_fused(mmx_plan_s* p, MagArgs<float>& a, double* norms, hipStream_t st, const NormRoute& route) {
    const int64_t V = p->V;
    const bool three = p->naxes == 3;
    PassArgs<float> gf;
    TileCfg tf;
    int rc = mag_geometry(p, p->ax[p->naxes - 1], false, gf, tf);
    if (rc) return rc;
    rc = partials_for(p, route, (size_t)(2 * tf.nblocks) * sizeof(double), &a.partials);
    if (rc) return rc;
    cx<float>* w1 = a.w1;
    cx<float>* w2 = a.w2;
    cx<float>* w3 = a.w3;
    // forward transform of E along the last axis (and the middle one): plain passes, three fields
    {
        PassArgs<float> b = blank_args<float>();
        fill_geometry<float>(p, p->ax[p->naxes - 1], b);
        b.src = a.E;
        b.dst = w2;
        b.dir = -1;
        if ((rc = launch_pass<float>(p, p->ax[p->naxes - 1], b, 3, false, kFwd, st))) return rc;
    }
    if (three && (rc = mag_middle_pass(p, w2, w2, 3, -1, st))) return rc;
    if ((rc = mag_launch(p, 0, a, w2, w2, st))) return rc;       // E^ -> w1, i k x E^
    if (three && (rc = mag_middle_pass(p, w2, w2, 3, +1, st))) return rc;
    if ((rc = mag_launch(p, 3, a, w2, w3, st))) return rc;       // curl E; [Z1; Y1], forward along the last axis
    if (three && (rc = mag_middle_pass(p, w3, w3, 6, -1, st))) return rc;
    if ((rc = mag_launch(p, 1, a, w3, w3, st))) return rc;       // [B^; i k x B^]
    if (three && (rc = mag_middle_pass(p, w3, w3, 6, +1, st))) return rc;
    if ((rc = mag_launch(p, 4, a, w3, w2, st))) return rc;       // [B; curl B]; Z2 -> w1; Y2 forward
    if (three && (rc = mag_middle_pass(p, w2, w2, 3, -1, st))) return rc;
    if ((rc = mag_launch(p, 2, a, w2, w2, st))) return rc;       // i k x Y2^
    if (three && (rc = mag_middle_pass(p, w2, w2, 3, +1, st))) return rc;
    if ((rc = mag_launch(p, 5, a, w2, w1, st))) return rc;       // curl Y2; d, E', partials
    return reduce_routed(p, route, a.partials, tf.nblocks, 1, norms, st);
}

template <typename T>
int magnetic_impl(mmx_plan_s* p, const void* E, void* Eout, const void* s, const void* const* chi, const int* kinds,
                  const int64_t* vstride, const double* alpha, void* work, double* norms, hipStream_t st,
                  const NormRoute& route = NormRoute()) {
    if (p->naxes == 0) return fail(MMX_EUNSUP, "grid without any axis longer than 1");
    MagArgs<T> a;
    memset(&a, 0, sizeof(a));
    MagChi<T>* m[4] = {&a.ee, &a.he, &a.eh, &a.hh};
    for (int i = 0; i < 4; ++i) {
        if (!chi[i]) return fail(MMX_EINVAL, "null susceptibility tensor");
        if (kinds[i] != 1 && kinds[i] != 2) return fail(MMX_EINVAL, "susceptibility kinds are 1 (scalar) or 2 (3x3)");
        if (vstride[i] != 0 && vstride[i] != 1) return fail(MMX_EINVAL, "voxel strides are 0 or 1");
        m[i]->p = reinterpret_cast<const cx<T>*>(chi[i]);
        m[i]->kind = kinds[i];
        m[i]->vstride = vstride[i];
    }
    const int64_t V = p->V;
    cx<T>* w = reinterpret_cast<cx<T>*>(work);
    a.E = reinterpret_cast<const cx<T>*>(E);
    a.Eout = reinterpret_cast<cx<T>*>(Eout);
    a.s = reinterpret_cast<const cx<T>*>(s);
    a.w1 = w;
    a.w2 = w + 3 * V;
    a.w3 = w + 6 * V;
    a.V = V;
    a.n1 = p->shape3[1];
    a.n2 = p->shape3[2];
    a.fn1 = make_fdiv64(p->shape3[1]);
    a.fn2 = make_fdiv64(p->shape3[2]);
    a.kvec = reinterpret_cast<const T*>(p->kvec);
    a.k1ofs = p->k1ofs;
    a.k2ofs = p->k2ofs;
    a.nd = p->ndim;
    a.alpha_r = T(alpha[0]);
    a.alpha_i = T(alpha[1]);
    if constexpr (sizeof(T) == 4) {
        if (mag_fused_ok(p)) return magnetic_fused(p, a, norms, st, route);
    }
    const int nblk = pointwise_blocks(V);
    double* partials = nullptr;
    int rc = partials_for(p, route, (size_t)(2 * nblk) * sizeof(double), &partials);
    if (rc) return rc;
    a.partials = partials;
    const dim3 grid(nblk), blk(256);
    auto check = [&]() -> int {
        MMX_HIP_CHECK(hipGetLastError());
        return MMX_OK;
    };
    // E^ and curl E
    if ((rc = fft_passes<T>(p, E, a.w1, 3, -1, true, st))) return rc;
    hipLaunchKernelGGL((cross_ik_kernel<T>), grid, blk, 0, st, a, a.w1, a.w2);
    if ((rc = check())) return rc;
    if ((rc = fft_passes<T>(p, a.w2, a.w2, 3, +1, true, st))) return rc;
    // [Z1; Y1], their transforms, B^ and i k x B^, back to [B; curl B]
    hipLaunchKernelGGL((magnetic_real_kernel<T>), grid, blk, 0, st, a, a.E, a.w2, 1);
    if ((rc = check())) return rc;
    if ((rc = fft_passes<T>(p, a.w3, a.w3, 6, -1, true, st))) return rc;
    hipLaunchKernelGGL((magnetic_green_kernel<T>), grid, blk, 0, st, a);
    if ((rc = check())) return rc;
    if ((rc = fft_passes<T>(p, a.w3, a.w3, 6, +1, true, st))) return rc;
    // [Z2; Y2] and curl Y2
    hipLaunchKernelGGL((magnetic_real_kernel<T>), grid, blk, 0, st, a, a.w3, a.w3 + 3 * V, 0);
    if ((rc = check())) return rc;
    cx<T>* y2 = a.w3 + 3 * V;
    if ((rc = fft_passes<T>(p, y2, y2, 3, -1, true, st))) return rc;
    hipLaunchKernelGGL((cross_ik_kernel<T>), grid, blk, 0, st, a, y2, y2);
    if ((rc = check())) return rc;
    if ((rc = fft_passes<T>(p, y2, y2, 3, +1, true, st))) return rc;
    // d = Z2 + curl Y2, E' and the norms
    hipLaunchKernelGGL((magnetic_final_kernel<T>), grid, blk, 0, st, a);
    if ((rc = check())) return rc;
    return reduce_routed(p, route, a.partials, nblk, 1, norms, st);
}

// Device address of `scratch` on the plan's device, looked up once per device (with that device current) and kept.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch(const mmx_plan_s* p, double** out) {
    int cur = -1;
    MMX_HIP_CHECK(hipGetDevice(&cur));
    if (cur != p->device) return fail(MMX_EINVAL, "the plan's device is not the current device");
    if (cur < 0 || cur >= kMaxDevices) return fail(MMX_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_ptr[cur] == nullptr) {
        void* ptr = nullptr;
        MMX_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr) return fail(MMX_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_noalloc_ptr[cur] = ptr;
    }
    *out = reinterpret_cast<double*>(g_noalloc_ptr[cur]);
    return MMX_OK;
}

template <typename T>
int norm2_impl(mmx_plan_s* p, const void* x, int64_t n, double* out, hipStream_t stream) {
    const int nblk = 1024;
    int rc = ensure_scratch(p, nblk * sizeof(double));
    if (rc) return rc;
    double* partials = reinterpret_cast<double*>(p->scratch);
    hipLaunchKernelGGL((norm2_partial_kernel<T>), dim3(nblk), dim3(256), 0, stream,
                       reinterpret_cast<const cx<T>*>(x), n, partials);
    MMX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, stream, partials, (int64_t)nblk, out);
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

int grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

template <typename T>
int chi_mul_impl(mmx_plan_s* p, const void* chi, int chi_kind, const void* x, void* y, int C, int64_t nbatch,
                 hipStream_t stream) {
    if (chi_kind >= 2 && C != 3) return fail(MMX_EINVAL, "3x3 chi needs 3 components");
    if (chi_kind < 1 || chi_kind > 4) return fail(MMX_EINVAL, "chi_kind must be 1, 2, 3 or 4");
    hipLaunchKernelGGL((chi_mul_kernel<T>), dim3(grid_for(p->V * nbatch)), dim3(256), 0, stream,
                       reinterpret_cast<const cx<T>*>(chi), chi_kind, reinterpret_cast<const cx<T>*>(x),
                       reinterpret_cast<cx<T>*>(y), C, p->V, nbatch);
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

template <typename T>
int lproj_impl(mmx_plan_s* p, const void* F, void* out, hipStream_t stream) {
    hipLaunchKernelGGL((longitudinal_projection_kernel<T>), dim3(grid_for(p->V)), dim3(256), 0, stream,
                       reinterpret_cast<const cx<T>*>(F), reinterpret_cast<cx<T>*>(out), p->shape3[0],
                       p->shape3[1], p->shape3[2], reinterpret_cast<const T*>(p->kvec), p->k1ofs, p->k2ofs,
                       p->ndim);
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

}  // namespace

extern "C" {

int mmx_version(void) { return 100; }

int mmx_register_fft_length(int64_t n) {
    if (n < 1 || n > kMaxN) return 0;
    AxisPlan ax{};
    ax.N = ax.NT = (int)n;
    ax.inner = 1;
    return reg_supported(ax) ? 1 : 0;
}

const char* mmx_last_error(void) { return g_last_error.c_str(); }

int mmx_plan_create(mmx_plan_t* out, int ndim, const int64_t* shape, const double* step, int dtype) {
    g_last_error.clear();
    if (!out || !shape || !step) return fail(MMX_EINVAL, "null argument");
    *out = nullptr;
    if (ndim < 1 || ndim > 3) return fail(MMX_EINVAL, "grid ndim must be 1, 2 or 3");
    if (dtype != MMX_C64 && dtype != MMX_C128) return fail(MMX_EINVAL, "dtype must be MMX_C64 or MMX_C128");
    mmx_plan_s* p = new mmx_plan_s();
    memset(p, 0, sizeof(*p));
    p->dtype = dtype;
    p->ndim = ndim;
    for (int a = 0; a < 3; ++a) p->shape3[a] = a < ndim ? shape[a] : 1;
    p->V = p->shape3[0] * p->shape3[1] * p->shape3[2];
    for (int a = 0; a < 3; ++a) {
        if (p->shape3[a] < 1) { delete p; return fail(MMX_EINVAL, "shape entries must be >= 1"); }
    }
    (void)hipGetDevice(&p->device);
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device) != hipSuccess || v <= 0)
            v = 65536;
        p->lds_max = (size_t)v;
    }
    p->g_axis = 1;  // Green pass on the middle axis of 3-D grids (clamped to the outermost axis otherwise)
    // measured at 256^3: the strided Green pass and the update passes gain from running neighbouring tiles on one
    // XCD (their 64-byte row segments share L2 lines); the plain passes do not
    p->xcd_remap = (1 << kGreen) | (1 << kInvEpi) | (1 << kEpiPro) | (1 << kGreenEpiPro);
    p->reg_fft = 1;
    p->radix_order = 0;
    p->lds_budget = kLdsBudget;
    p->fused_ok[0] = p->fused_ok[1] = -1;
    // non-trivial axes, outermost first
    p->naxes = 0;
    for (int a = 0; a < 3; ++a) {
        const int64_t n = p->shape3[a];
        if (n == 1) continue;
        const int q = p->naxes++;
        AxisPlan& ax = p->ax[q];
        ax.orig = a;
        int64_t inner = 1, outer = 1;
        for (int b = a + 1; b < 3; ++b) inner *= p->shape3[b];
        for (int b = 0; b < a; ++b) outer *= p->shape3[b];
        int rc;
        if (direct_ok(n) || blue_ok(n)) {
            rc = init_axis(p, ax, n, inner, outer);
        } else {
            int n1 = 0, n2 = 0;
            if (n > (int64_t)kMaxN * kMaxN || !choose_split(n, n1, n2)) {
                mmx_plan_destroy(p);  // frees the tables of the axes before this one
                return fail(MMX_EUNSUP, "FFT length " + std::to_string(n) +
                                            " is not supported (a prime factor above 2048, or more than 4096^2 points)");
            }
            ax.N = (int)n;
            ax.NT = (int)n;
            ax.inner = inner;
            ax.nouter = outer;
            ax.split = 1;
            ax.N1 = n1;
            ax.N2 = n2;
            p->sub[q][0].orig = p->sub[q][1].orig = a;
            rc = init_axis(p, p->sub[q][0], n1, (int64_t)n2 * inner, outer);
            if (!rc) rc = init_axis(p, p->sub[q][1], n2, inner, outer * n1);
            if (!rc) rc = dtype == MMX_C64 ? make_twiddles<float>((int)n, &ax.tw4) : make_twiddles<double>((int)n, &ax.tw4);
        }
        if (rc) { mmx_plan_destroy(p); return rc; }
    }
    for (int a = 0; a < 3; ++a) p->step3[a] = a < ndim ? step[a] : 1.0;
    p->transposed = 0;
    for (int a = 0; a < 3; ++a) { p->s_lo[a] = 0; p->s_hi[a] = p->shape3[a]; }
    p->k1ofs = (int)p->shape3[0];
    p->k2ofs = (int)(p->shape3[0] + p->shape3[1]);
    const std::vector<double> k = k_vectors(p->shape3, p->step3, p->transposed);
    int rc = dtype == MMX_C64 ? upload_real<float>(k, &p->kvec) : upload_real<double>(k, &p->kvec);
    if (rc) { mmx_plan_destroy(p); return rc; }
    *out = p;
    return MMX_OK;
}

int mmx_plan_destroy(mmx_plan_t p) {
    if (!p) return MMX_OK;
    auto free_axis = [](AxisPlan& ax) {
        for (void* d : {ax.tw, ax.chirp, ax.bhat_f, ax.bhat_i, ax.tw4})
            if (d) (void)hipFree(d);
    };
    for (int a = 0; a < p->naxes; ++a) {
        free_axis(p->ax[a]);
        free_axis(p->sub[a][0]);
        free_axis(p->sub[a][1]);
    }
    if (p->kvec) (void)hipFree(p->kvec);
    if (p->scratch) (void)hipFree(p->scratch);
    if (p->fft_scratch) (void)hipFree(p->fft_scratch);
    for (int q = 0; q < 2; ++q) {
        if (p->slot_partials[q]) (void)hipFree(p->slot_partials[q]);
        if (p->slot_done[q]) (void)hipEventDestroy(p->slot_done[q]);
    }
    delete p;
    return MMX_OK;
}

int mmx_plan_set_param(mmx_plan_t p, int param, int64_t value) {
    g_last_error.clear();
    if (!p) return fail(MMX_EINVAL, "null plan");
    if (value < 0) return fail(MMX_EINVAL, "parameter values are >= 0");
    p->fused_ok[0] = p->fused_ok[1] = -1;
    switch (param) {
        case MMX_PARAM_G_AXIS: p->g_axis = (int)value; break;
        case MMX_PARAM_P_ROWS: p->p_rows = (int)value; break;
        case MMX_PARAM_P_STRIDED: p->p_strided = (int)value; break;
        case MMX_PARAM_P_GREEN: p->p_green = (int)value; break;
        case MMX_PARAM_XCD_REMAP:
            if (value < 0 || value > 255) return fail(MMX_EINVAL, "XCD remap mask must be in [0, 255]");
            p->xcd_remap = (int)value;
            break;
        case MMX_PARAM_LDS_BUDGET: p->lds_budget = value ? (size_t)value : kLdsBudget; break;
        case MMX_PARAM_BIG_ELEMS: p->big_elems = (int)value; break;
        case MMX_PARAM_REG_FFT: p->reg_fft = value ? 1 : 0; break;
        case MMX_PARAM_TRANSPOSED: {
            const int t = value ? 1 : 0;
            if (t != p->transposed) {
                p->transposed = t;
                const std::vector<double> k = k_vectors(p->shape3, p->step3, p->transposed);
                const int rc = p->dtype == MMX_C64 ? reupload_real<float>(k, p->kvec)
                                                   : reupload_real<double>(k, p->kvec);
                if (rc) return rc;
            }
            break;
        }
        case MMX_PARAM_RADIX_ORDER: {
            if (value >= kRadixOrders) return fail(MMX_EINVAL, "radix order must be 0, 1 or 2");
            // every transform length is 7-smooth, so each order factorises it; only the stage list changes
            // (a four-step axis is transformed by the passes of its two sub-axes)
            for (int q = 0; q < p->naxes; ++q) {
                AxisPlan* axs[2] = {&p->sub[q][0], &p->sub[q][1]};
                if (!p->ax[q].split) axs[0] = &p->ax[q];
                for (int i = 0; i < (p->ax[q].split ? 2 : 1); ++i)
                    if (!factorize(axs[i]->NT, axs[i]->rad, axs[i]->nrad, (int)value))
                        return fail(MMX_EUNSUP, "internal: transform length radices");
            }
            p->radix_order = (int)value;
            break;
        }
        case MMX_PARAM_SCRATCH_GEN: return fail(MMX_EINVAL, "MMX_PARAM_SCRATCH_GEN is read-only");
        default: return fail(MMX_EINVAL, "unknown parameter");
    }
    return MMX_OK;
}

int mmx_plan_get_param(mmx_plan_t p, int param, int64_t* value) {
    g_last_error.clear();
    if (!p || !value) return fail(MMX_EINVAL, "null argument");
    switch (param) {
        case MMX_PARAM_G_AXIS: *value = p->g_axis; break;
        case MMX_PARAM_P_ROWS: *value = p->p_rows; break;
        case MMX_PARAM_P_STRIDED: *value = p->p_strided; break;
        case MMX_PARAM_P_GREEN: *value = p->p_green; break;
        case MMX_PARAM_XCD_REMAP: *value = p->xcd_remap; break;
        case MMX_PARAM_LDS_BUDGET: *value = (int64_t)p->lds_budget; break;
        case MMX_PARAM_BIG_ELEMS: *value = p->big_elems; break;
        case MMX_PARAM_REG_FFT: *value = p->reg_fft; break;
        case MMX_PARAM_TRANSPOSED: *value = p->transposed; break;
        case MMX_PARAM_SCRATCH_GEN: *value = p->scratch_gen; break;
        case MMX_PARAM_RADIX_ORDER: *value = p->radix_order; break;
        default: return fail(MMX_EINVAL, "unknown parameter");
    }
    return MMX_OK;
}

int mmx_plan_set_source_support(mmx_plan_t p, const int64_t* lo, const int64_t* hi) {
    g_last_error.clear();
    if (!p) return fail(MMX_EINVAL, "null plan");
    if ((lo == nullptr) != (hi == nullptr)) return fail(MMX_EINVAL, "give both lo and hi, or neither");
    int64_t l[3], h[3];
    for (int a = 0; a < 3; ++a) {
        l[a] = (lo && a < p->ndim) ? lo[a] : 0;
        h[a] = (hi && a < p->ndim) ? hi[a] : p->shape3[a];
        if (l[a] < 0 || h[a] > p->shape3[a] || l[a] > h[a]) return fail(MMX_EINVAL, "source box outside the grid");
        if (h[a] >= (int64_t(1) << 31)) return fail(MMX_EINVAL, "source box too large");
    }
    for (int a = 0; a < 3; ++a) { p->s_lo[a] = l[a]; p->s_hi[a] = h[a]; }
    return MMX_OK;
}

int mmx_fft(mmx_plan_t p, const void* in, void* out, int64_t nbatch, int direction, void* stream) {
    g_last_error.clear();
    if (!p || !in || !out) return fail(MMX_EINVAL, "null argument");
    if (direction != -1 && direction != 1) return fail(MMX_EINVAL, "direction must be -1 or +1");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    return p->dtype == MMX_C64 ? fft_impl<float>(p, in, out, nbatch, direction, s)
                               : fft_impl<double>(p, in, out, nbatch, direction, s);
}

int mmx_born_step(mmx_plan_t p, const void* E, void* Eout, const void* s, const void* chi, int ncomp, int chi_kind,
                  const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride, double* norms,
                  void* stream) {
    g_last_error.clear();
    if (!p || !E || !Eout || !s || !chi || !coef || !norms) return fail(MMX_EINVAL, "null argument");
    if (E == Eout) return fail(MMX_EINVAL, "E and Eout must be different buffers");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    return p->dtype == MMX_C64
               ? born_impl<float>(p, E, Eout, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms, st)
               : born_impl<double>(p, E, Eout, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms, st);
}

int mmx_born_step_noalloc(mmx_plan_t p, const void* E, void* Eout, const void* s, const void* chi, int ncomp,
                          int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride,
                          double* norms, void* stream) {
    g_last_error.clear();
    if (!p || !E || !Eout || !s || !chi || !coef || !norms) return fail(MMX_EINVAL, "null argument");
    if (E == Eout) return fail(MMX_EINVAL, "E and Eout must be different buffers");
    double* partials = nullptr;
    int rc = noalloc_scratch(p, &partials);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const size_t cap = MMX_NOALLOC_SCRATCH_BYTES;
    return p->dtype == MMX_C64
               ? born_impl<float>(p, E, Eout, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms, st,
                                  partials, cap)
               : born_impl<double>(p, E, Eout, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms,
                                   st, partials, cap);
}

int mmx_magnetic_iterate(mmx_plan_t p, const void* E, void* Eout, const void* s, const void* const* chi,
                         const int* chi_kind, const int64_t* chi_vstride, const double* alpha, void* work,
                         double* norms, void* stream) {
    g_last_error.clear();
    if (!p || !E || !Eout || !s || !chi || !chi_kind || !chi_vstride || !alpha || !work || !norms)
        return fail(MMX_EINVAL, "null argument");
    if (E == Eout || E == work || Eout == work) return fail(MMX_EINVAL, "E, Eout and work must be different buffers");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    return p->dtype == MMX_C64
               ? magnetic_impl<float>(p, E, Eout, s, chi, chi_kind, chi_vstride, alpha, work, norms, st)
               : magnetic_impl<double>(p, E, Eout, s, chi, chi_kind, chi_vstride, alpha, work, norms, st);
}

int mmx_magnetic_iterate_split(mmx_plan_t p, const void* E, void* Eout, const void* s, const void* const* chi,
                               const int* chi_kind, const int64_t* chi_vstride, const double* alpha, void* work,
                               double* norms, double* norms_host, int slot, void* stream, void* norm_stream) {
    g_last_error.clear();
    if (!p || !E || !Eout || !s || !chi || !chi_kind || !chi_vstride || !alpha || !work || !norms || !norm_stream)
        return fail(MMX_EINVAL, "null argument");
    if (E == Eout || E == work || Eout == work) return fail(MMX_EINVAL, "E, Eout and work must be different buffers");
    if (slot != 0 && slot != 1) return fail(MMX_EINVAL, "slot must be 0 or 1");
    if (norm_stream == stream) return fail(MMX_EINVAL, "norm_stream must differ from stream");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    NormRoute route;
    route.norm_stream = reinterpret_cast<hipStream_t>(norm_stream);
    route.slot = slot;
    route.norms_host = norms_host;
    return p->dtype == MMX_C64
               ? magnetic_impl<float>(p, E, Eout, s, chi, chi_kind, chi_vstride, alpha, work, norms, st, route)
               : magnetic_impl<double>(p, E, Eout, s, chi, chi_kind, chi_vstride, alpha, work, norms, st, route);
}

int mmx_born_prologue(mmx_plan_t p, const void* E, void* W, const void* s, const void* chi, int ncomp, int chi_kind,
                      const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride, void* stream) {
    g_last_error.clear();
    if (!p || !E || !W || !s || !chi || !coef) return fail(MMX_EINVAL, "null argument");
    if (E == W) return fail(MMX_EINVAL, "E and W must be different buffers");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    return p->dtype == MMX_C64
               ? prologue_impl<float>(p, E, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, st)
               : prologue_impl<double>(p, E, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, st);
}

int mmx_born_iterate(mmx_plan_t p, const void* E, void* Eout, void* W, const void* s, const void* chi, int ncomp,
                     int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride,
                     double* norms, void* stream) {
    g_last_error.clear();
    if (!p || !E || !Eout || !W || !s || !chi || !coef || !norms) return fail(MMX_EINVAL, "null argument");
    if (E == Eout || E == W || Eout == W) return fail(MMX_EINVAL, "E, Eout and W must be three different buffers");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    return p->dtype == MMX_C64
               ? iterate_impl<float>(p, E, Eout, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms, st)
               : iterate_impl<double>(p, E, Eout, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms, st);
}

int mmx_born_iterate_split(mmx_plan_t p, const void* E, void* Eout, void* W, const void* s, const void* chi,
                           int ncomp, int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride,
                           int64_t chi_bstride, double* norms, double* norms_host, int slot, void* stream,
                           void* norm_stream) {
    g_last_error.clear();
    if (!p || !E || !Eout || !W || !s || !chi || !coef || !norms || !norm_stream) return fail(MMX_EINVAL, "null argument");
    if (E == Eout || E == W || Eout == W) return fail(MMX_EINVAL, "E, Eout and W must be three different buffers");
    if (slot != 0 && slot != 1) return fail(MMX_EINVAL, "slot must be 0 or 1");
    if (norm_stream == stream) return fail(MMX_EINVAL, "norm_stream must differ from stream");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    NormRoute route;
    route.norm_stream = reinterpret_cast<hipStream_t>(norm_stream);
    route.slot = slot;
    route.norms_host = norms_host;
    return p->dtype == MMX_C64
               ? iterate_impl<float>(p, E, Eout, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms,
                                     st, nullptr, route)
               : iterate_impl<double>(p, E, Eout, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride,
                                      norms, st, nullptr, route);
}

int mmx_batch_step(mmx_plan_t p, void* EA, void* EB, void* W, const void* s, const void* chi, int ncomp,
                   int chi_kind, void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride, uint32_t* state,
                   double* track, double* norms, double tol, int64_t max_iter, void* stream) {
    g_last_error.clear();
    if (!p || !EA || !EB || !W || !s || !chi || !coef || !state || !track || !norms)
        return fail(MMX_EINVAL, "null argument");
    if (EA == EB || EA == W || EB == W) return fail(MMX_EINVAL, "EA, EB and W must be different buffers");
    if (max_iter < 0) return fail(MMX_EINVAL, "max_iter must be >= 0");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    unsigned* stt = reinterpret_cast<unsigned*>(state);
    return p->dtype == MMX_C64
               ? batch_step_impl<float>(p, EA, EB, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride,
                                        stt, track, norms, tol, max_iter, st)
               : batch_step_impl<double>(p, EA, EB, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride,
                                         stt, track, norms, tol, max_iter, st);
}

int mmx_wall_clock(double* out, void* stream) {
    g_last_error.clear();
    if (!out) return fail(MMX_EINVAL, "null argument");
    hipLaunchKernelGGL(wall_clock_kernel, dim3(1), dim3(1), 0, reinterpret_cast<hipStream_t>(stream), out);
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

int mmx_wall_clock_rate_khz(int64_t* khz) {
    g_last_error.clear();
    if (!khz) return fail(MMX_EINVAL, "null argument");
    int dev = 0, rate = 0;
    MMX_HIP_CHECK(hipGetDevice(&dev));
    MMX_HIP_CHECK(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev));
    *khz = rate;
    return MMX_OK;
}

// Set-up extremes over a complex128 permittivity (mode 0: sigma^2 with a shift; 1: gain eigenvalue range).
int setup_extreme(mmx_plan_s* p, const void* eps, int full3, double shift, int mode, double* out, hipStream_t st) {
    const int64_t V = p->V;
    int nblk = (int)((V + 255) / 256);
    if (nblk > 2048) nblk = 2048;
    if (nblk < 1) nblk = 1;
    const int rc = ensure_scratch(p, (size_t)2 * nblk * sizeof(double));
    if (rc) return rc;
    double* partials = reinterpret_cast<double*>(p->scratch);
    const cx<double>* e = reinterpret_cast<const cx<double>*>(eps);
    if (mode == 0)
        hipLaunchKernelGGL((setup_extreme_kernel<0>), dim3(nblk), dim3(256), 0, st, e, V, full3, shift, partials);
    else
        hipLaunchKernelGGL((setup_extreme_kernel<1>), dim3(nblk), dim3(256), 0, st, e, V, full3, shift, partials);
    MMX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(setup_extreme_final_kernel, dim3(1), dim3(256), 0, st, partials, nblk, out);
    MMX_HIP_CHECK(hipGetLastError());
    return MMX_OK;
}

int mmx_setup_sigma(mmx_plan_t p, const void* eps, int eps_kind, double shift, double* out, void* stream) {
    g_last_error.clear();
    if (!p || !eps || !out) return fail(MMX_EINVAL, "null argument");
    if (eps_kind != 1 && eps_kind != 2) return fail(MMX_EINVAL, "eps_kind must be 1 (isotropic) or 2 (3x3)");
    return setup_extreme(p, eps, eps_kind == 2, shift, 0, out, reinterpret_cast<hipStream_t>(stream));
}

int mmx_setup_gain(mmx_plan_t p, const void* eps, int eps_kind, double* out, void* stream) {
    g_last_error.clear();
    if (!p || !eps || !out) return fail(MMX_EINVAL, "null argument");
    if (eps_kind != 1 && eps_kind != 2) return fail(MMX_EINVAL, "eps_kind must be 1 (isotropic) or 2 (3x3)");
    return setup_extreme(p, eps, eps_kind == 2, 0.0, 1, out, reinterpret_cast<hipStream_t>(stream));
}

int mmx_norm2(mmx_plan_t p, const void* x, int64_t n, double* out, void* stream) {
    g_last_error.clear();
    if (!p || !x || !out) return fail(MMX_EINVAL, "null argument");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    return p->dtype == MMX_C64 ? norm2_impl<float>(p, x, n, out, s) : norm2_impl<double>(p, x, n, out, s);
}

int mmx_chi_mul(mmx_plan_t p, const void* chi, int chi_kind, const void* x, void* y, int ncomp, int64_t nbatch,
                void* stream) {
    g_last_error.clear();
    if (!p || !chi || !x || !y) return fail(MMX_EINVAL, "null argument");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    return p->dtype == MMX_C64 ? chi_mul_impl<float>(p, chi, chi_kind, x, y, ncomp, nbatch, s)
                               : chi_mul_impl<double>(p, chi, chi_kind, x, y, ncomp, nbatch, s);
}

int mmx_longitudinal_projection_ft(mmx_plan_t p, const void* F, void* out, void* stream) {
    g_last_error.clear();
    if (!p || !F || !out) return fail(MMX_EINVAL, "null argument");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    return p->dtype == MMX_C64 ? lproj_impl<float>(p, F, out, s) : lproj_impl<double>(p, F, out, s);
}

}  // extern "C"

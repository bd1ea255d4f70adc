// Launchers of the register-resident pass kernels (mmx_regkernels.hpp), one translation unit per axis length
// (mmx_reg_<N>.hip) so that the kernels of different lengths compile in parallel; mmx_capi.hip dispatches to them.
#pragma once
#include <atomic>
#include <hip/hip_runtime.h>

#include "../../include/macromax_hip.h"
#include "mmx_regkernels.hpp"

namespace mmx {

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

// Points per thread of a register-kernel pass: fewer for the fused last pass, which also holds E, chi and s of
// its voxels; lengths with a factor 3 use 6 points per thread for three components (radix-2 stages, then radix 3;
// 132-172 VGPRs, measured 2-6 % faster than 12 points with 192-225 VGPRs) and 24 for one.
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
    return (N % 3 == 0)                              ? (NC == 3 ? MMX_E_RADIX3_C3 : 24)
           : (KIND == kEpiPro || KIND == kGreenEpiPro) ? (NC == 3 ? MMX_E_FUSED3 : 8)
           : (KIND == kGreen && NC == 3)             ? MMX_E_GREEN3
           : (NC == 3 ? 8 : (N >= 256 ? 16 : 8));
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

constexpr bool kind_applies_chi(int kind) {
    return kind == kProFwd || kind == kInvEpi || kind == kAll || kind == kEpiPro || kind == kGreenEpiPro;
}

// Instantiation for the component count and (for passes that apply chi) the chi kind of the launch. Returns
// MMX_EINVAL for a combination no pass uses.
template <int N, int KIND>
int launch_reg_kind(bool three, int chi_kind, const TileCfg& t, const PassArgs<float>& a, hipStream_t stream) {
    if constexpr (KIND == kFwd || KIND == kInv) {  // single-component passes
        if (three) return MMX_EINVAL;
        launch_reg<N, KIND, 1, 0>(t, a, stream);
    } else if constexpr (!kind_applies_chi(KIND)) {
        if (three) launch_reg<N, KIND, 3, 0>(t, a, stream);
        else launch_reg<N, KIND, 1, 0>(t, a, stream);
    } else {
        if (!three) {
            if (chi_kind != 1) return MMX_EINVAL;
            launch_reg<N, KIND, 1, 1>(t, a, stream);
            return MMX_OK;
        }
        switch (chi_kind) {
            case 1: launch_reg<N, KIND, 3, 1>(t, a, stream); break;
            case 2: launch_reg<N, KIND, 3, 2>(t, a, stream); break;
            case 3: launch_reg<N, KIND, 3, 3>(t, a, stream); break;
            case 4: launch_reg<N, KIND, 3, 4>(t, a, stream); break;
            default: return MMX_EINVAL;
        }
    }
    return MMX_OK;
}

template <int N>
int launch_reg_len(int kind, bool three, int chi_kind, const TileCfg& t, const PassArgs<float>& a,
                   hipStream_t stream) {
    switch (kind) {
        case kFwd: return launch_reg_kind<N, kFwd>(three, chi_kind, t, a, stream);
        case kInv: return launch_reg_kind<N, kInv>(three, chi_kind, t, a, stream);
        case kProFwd: return launch_reg_kind<N, kProFwd>(three, chi_kind, t, a, stream);
        case kGreen: return launch_reg_kind<N, kGreen>(three, chi_kind, t, a, stream);
        case kInvEpi: return launch_reg_kind<N, kInvEpi>(three, chi_kind, t, a, stream);
        case kAll: return launch_reg_kind<N, kAll>(three, chi_kind, t, a, stream);
        case kEpiPro: return launch_reg_kind<N, kEpiPro>(three, chi_kind, t, a, stream);
        case kGreenEpiPro: return launch_reg_kind<N, kGreenEpiPro>(three, chi_kind, t, a, stream);
        default: return MMX_EINVAL;
    }
}

// One launcher per length, each defined in its own translation unit (mmx_reg_<N>.hip).
#define MMX_REG_LENGTHS(X) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
#define MMX_DECLARE_REG_LAUNCHER(N)                                                                     \
    int launch_reg_##N(int kind, bool three, int chi_kind, const TileCfg& t, const PassArgs<float>& a, \
                       hipStream_t stream);
MMX_REG_LENGTHS(MMX_DECLARE_REG_LAUNCHER)
#undef MMX_DECLARE_REG_LAUNCHER

}  // namespace mmx

// Streaming ceiling probe (a measurement aid, not part of the solver library): the HBM bandwidth a kernel reaches
// when it only streams the planes one pencil pass touches, with the access pattern of the register-FFT passes but
// no arithmetic. bench.py --stream-ceiling reports it beside the passes' achieved bandwidth.
//
// Each workgroup takes 8 rows of 256 complex64 values (256 threads; thread (p, t) handles points t + 32 m of row p,
// m < 8, as the update pass does), reads NR planes and writes NW planes, all planes V complex values apart, with
// non-temporal loads and stores. Built by macromax_amd/build.py into libmmx_stream_probe.so.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

struct cxf {
    float r, i;
};

template <int NR, int NW>
__global__ void __launch_bounds__(256) stream_rows_kernel(const cxf* __restrict__ in, cxf* __restrict__ out,
                                                          int64_t V) {
    const int t = threadIdx.x & 31, p = threadIdx.x >> 5;
    const int64_t tb = (int64_t)blockIdx.x * 2048;
    cxf v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int64_t o = tb + p * 256 + t + m * 32;
        cxf s{0.f, 0.f};
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const cxf* q = in + k * V + o;
            s.r += __builtin_nontemporal_load(&q->r);
            s.i += __builtin_nontemporal_load(&q->i);
        }
        v[m] = s;
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int64_t o = tb + p * 256 + t + m * 32;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            cxf* q = out + k * V + o;
            __builtin_nontemporal_store(v[m].r + float(k), &q->r);
            __builtin_nontemporal_store(v[m].i, &q->i);
        }
    }
}

template <int NR, int NW>
void launch(const void* in, void* out, int64_t V, hipStream_t st) {
    hipLaunchKernelGGL((stream_rows_kernel<NR, NW>), dim3((unsigned)(V / 2048)), dim3(256), 0, st,
                       reinterpret_cast<const cxf*>(in), reinterpret_cast<cxf*>(out), V);
}

}  // namespace

extern "C" {

// One streaming pass over V complex64 values per plane (V a multiple of 2048): (nread, nwrite) in {(3, 3) the plain
// and Green passes of three components, (6, 6), (10, 6) the update pass with uniaxial chi: W, E, chi (4 planes) read;
// E', W written}. `in` holds nread planes, `out` nwrite planes. Returns 0, or -1 for other shapes.
int mmx_stream_probe(int nread, int nwrite, int64_t V, const void* in, void* out, void* stream) {
    if (V <= 0 || V % 2048 != 0 || !in || !out) return -1;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (nread == 3 && nwrite == 3) launch<3, 3>(in, out, V, st);
    else if (nread == 6 && nwrite == 6) launch<6, 6>(in, out, V, st);
    else if (nread == 10 && nwrite == 6) launch<10, 6>(in, out, V, st);
    else return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"

/*
 * C ABI of the MI355X (gfx950) Born-series library, libmacromax_hip.so.
 *
 * All array arguments are device pointers to contiguous complex arrays (interleaved re/im; complex64 when the
 * plan's dtype is MMX_C64, complex128 for MMX_C128) laid out as MacroMax lays out its fields:
 * [C, 1, n0, n1, n2] in C order, last axis fastest (MacroMax backend/__init__.py:130-144). `stream` is a
 * hipStream_t (NULL = the null stream). Functions return 0 on success and a negative MMX_E* code on error;
 * mmx_last_error() then describes it. No function throws across the ABI. A plan is used by one host thread.
 */
#ifndef MACROMAX_HIP_H
#define MACROMAX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMX_C64 0
#define MMX_C128 1

#define MMX_OK 0
#define MMX_EINVAL -1    /* bad argument (shape, size, kind) */
#define MMX_EUNSUP -2    /* unsupported FFT length (a prime factor above 2048, or more than 4096^2 points) */
#define MMX_EHIP -3      /* HIP runtime error */
#define MMX_ENOMEM -4

typedef struct mmx_plan_s* mmx_plan_t;

/* Library version (major*10000 + minor*100 + patch). */
int mmx_version(void);
/* Text of the last error on this thread (empty if none). */
const char* mmx_last_error(void);
/* 1 if complex64 axes of n points have register-resident FFT kernels (MMX_PARAM_REG_FFT chooses between them and the
 * LDS-staged kernels), else 0. No reference counterpart (its FFTs are library calls); for tuners. */
int mmx_register_fft_length(int64_t n);

/*
 * Plan for a grid of `ndim` (1..3) axes with `shape[ndim]` samples and k0-scaled sample pitch `step[ndim]`.
 * Holds twiddle tables and the k-vectors k_a = 2 pi fftfreq(n_a, step_a) (MacroMax utils/ft/grid.py:426-439)
 * on the current device. Replaces the state that MacroMax BackEnd.__init__ builds
 * (backend/__init__.py:43-67; pyFFTW plans in backend/numpy.py:46-89). Axis lengths: up to 4096 points with prime
 * factors up to 7 in one pass; a length with a larger prime factor (up to 2048 points) by Bluestein's algorithm in one
 * pass; longer axes as a four-step product n1 x n2 of such lengths (two passes and a transposition, with a scratch
 * buffer of one field batch allocated on first use). Grids that need Bluestein or four-step axes, or whose fused tiles
 * do not fit in LDS (e.g. three complex128 components of 4096 points), run the Born iteration unfused: voxel-wise
 * kernels around whole-grid FFTs, with the same results.
 */
int mmx_plan_create(mmx_plan_t* plan, int ndim, const int64_t* shape, const double* step, int dtype);
int mmx_plan_destroy(mmx_plan_t plan);

/*
 * Tuning parameters of a plan (no counterpart in the reference: its FFTs are library calls). 0 restores the
 * automatic choice. Values out of range are clamped to what the device and the tile shape allow.
 *   MMX_PARAM_G_AXIS     which non-trivial axis (0 = outermost) hosts the fused forward/Green/inverse pass;
 *                        the innermost axis always hosts the susceptibility prologue and update epilogue
 *   MMX_PARAM_P_ROWS     pencils per tile in passes along the contiguous axis
 *   MMX_PARAM_P_STRIDED  pencils (contiguous columns) per tile in single-component strided passes
 *   MMX_PARAM_P_GREEN    pencils per tile in the fused Green pass (all components in one tile)
 *   MMX_PARAM_XCD_REMAP  bit mask over pass kinds (bits as for MMX_PARAM_BIG_ELEMS, plus 6 inverse+update+next
 *                        prologue, 7 single-axis with next prologue): those passes remap workgroups so that
 *                        neighbouring tiles run on one XCD (default: Green and update passes)
 *   MMX_PARAM_LDS_BUDGET preferred LDS bytes per workgroup
 *   MMX_PARAM_BIG_ELEMS  bit mask over pass kinds (bit 0 forward, 1 inverse, 2 prologue+forward, 3 Green,
 *                        4 inverse+update, 5 single-axis): those passes give each thread twice the elements
 *   MMX_PARAM_REG_FFT    1 (default): complex64 axes of 64, 80, 96, 128, 160, 192, 256, 320, 384, 512, 640,
 *                        768, 1024 or 2048 points use the register-resident
 *                        FFT kernels; 0: every axis uses the LDS-staged kernels
 *   MMX_PARAM_TRANSPOSED 0 (default): the operator of the reference (k = 2 pi fftfreq, utils/ft/grid.py:426-439);
 *                        1: its transpose, for adjoint solves by reciprocity: the Nyquist entry of every
 *                        even-length axis is +pi/d instead of -pi/d (only the dyadic Green's cross terms
 *                        k_i k_j / k^2 on those planes change). Synchronises the device when it changes.
 *   MMX_PARAM_SCRATCH_GEN read-only (mmx_plan_get_param): how many times the plan's device scratch buffer (FP64
 *                        partial sums) has been allocated. Launches recorded in a graph hold that buffer's address,
 *                        so a caller replaying a graph re-records it when this number has changed since capture.
 *   MMX_PARAM_RADIX_ORDER radix order of the LDS-staged transforms (every axis the register kernels do not take):
 *                        0 (default) radix 8 first, then 4, 2, 3, 5, 7; 1 radix 4 first (4, 2, 8, 3, 5, 7);
 *                        2 the odd radices first (7, 5, 3, 8, 4, 2). Stages run in the order their radices are
 *                        taken. Values above 2 are rejected (MMX_EINVAL).
 */
#define MMX_PARAM_G_AXIS 1
#define MMX_PARAM_P_ROWS 2
#define MMX_PARAM_P_STRIDED 3
#define MMX_PARAM_P_GREEN 4
#define MMX_PARAM_XCD_REMAP 5
#define MMX_PARAM_LDS_BUDGET 6
#define MMX_PARAM_BIG_ELEMS 7
#define MMX_PARAM_REG_FFT 8
#define MMX_PARAM_TRANSPOSED 9
#define MMX_PARAM_SCRATCH_GEN 10
#define MMX_PARAM_RADIX_ORDER 11
int mmx_plan_set_param(mmx_plan_t plan, int param, int64_t value);
/*
 * Support of the sources s of the Born calls that follow on this plan: s is zero outside the box
 * lo[a] <= i_a < hi[a] (a < ndim), so the passes that add s read it only inside. NULL, NULL restores the whole
 * grid (the default). No reference counterpart: the reference stores and reads s densely (solver.py:702-720,
 * 1070). A caller that changes s, or passes a batch with several sources, sets the box (their union) first.
 */
int mmx_plan_set_source_support(mmx_plan_t plan, const int64_t* lo, const int64_t* hi);
int mmx_plan_get_param(mmx_plan_t plan, int param, int64_t* value);

/*
 * FFT over the spatial axes of `nbatch` consecutive fields of the plan's grid.
 * direction -1: unnormalised forward transform; +1: inverse transform with 1/N.
 * in may equal out. Replaces BackEnd.ft / BackEnd.ift (backend/__init__.py:234-257; torch.py:183-205;
 * numpy.py:112-133).
 */
int mmx_fft(mmx_plan_t plan, const void* in, void* out, int64_t nbatch, int direction, void* stream);

/*
 * Per-problem preconditioner coefficients, an array of `nbatch` records of 6 reals of the plan's real type:
 *   {alpha_re, alpha_im, a, c_re, c_im, 0}
 * chi_eff = a chi + c I and s_eff = a s; alpha enters the Green's function. a = 1, c = 0 reproduce MacroMax's
 * chi and s exactly; a rejected update (solver.py:1089-1094) changes alpha, a and c instead of chi and s.
 */

/*
 * One Born iteration for `nbatch` problems on the plan's grid (MacroMax Solution.__iter__,
 * solver.py:1070-1085):
 *     d = chi_eff [ G(chi_eff E + s_eff) - E ],   Eout = E + d,
 * G the scalar (ncomp == 1) or dyadic (ncomp == 3) Green's operator (solver.py:564-584,
 * backend/__init__.py:606-649). chi_kind 1: chi is [1,1,grid] (isotropic); 2: chi is [3,3,grid] (needs ncomp 3);
 * 3: chi is complex-symmetric (chi^T = chi, reciprocal media) packed as [6,grid] = {xx, xy, xz, yy, yz, zz};
 * 4: chi is uniaxial, chi = u I + w n n^T with a real unit axis n per voxel, stored as [4,grid] complex planes
 * {u, w, n_x + i n_y, n_z} where the last plane holds n_z as the first grid-size reals (28 bytes per voxel in
 * complex64 instead of 48; the storage has no reference counterpart, the operator is the reference's chi).
 * E, Eout: [nbatch][ncomp][grid]; s: [nbatch or 1][ncomp][grid] (s_bstride 0 shares one source);
 * chi: [nbatch or 1][chi size] (chi_bstride 0 shares one medium). Eout is also the work buffer; E is only read.
 * norms (device, 2*nbatch doubles): {sum |d|^2, sum |Eout|^2} per problem, accumulated in FP64 with a fixed
 * reduction order. The caller decides acceptance (||d|| < ||d_prev||, solver.py:1083) and swaps E and Eout.
 */
int mmx_born_step(mmx_plan_t plan, const void* E, void* Eout, const void* s, const void* chi, int ncomp,
                  int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride,
                  double* norms, void* stream);

/*
 * mmx_born_step without any device allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as mmx_born_step (which allocates or grows the plan's scratch buffer for its
 * per-tile FP64 norm partials on first use). Here the partials go to a static device array of
 * MMX_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs on until the process ends; a launch that
 * needs more (2 doubles per tile of the last pass per problem: e.g. 128 KiB per problem at 256^3 with three
 * components) fails with MMX_EINVAL before anything is launched. The array needs no clearing: every launch writes
 * each partial it reads. All calls on one device share that device's array, so calls that overlap on one device
 * (other streams, other host threads) race. The plan's device must be current.
 */
#define MMX_NOALLOC_SCRATCH_BYTES (4u << 20)
int mmx_born_step_noalloc(mmx_plan_t plan, const void* E, void* Eout, const void* s, const void* chi, int ncomp,
                          int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride,
                          double* norms, void* stream);

/*
 * The same iteration split so that consecutive iterations share one pass over the field (the last pass of
 * iteration n also computes the first pass of iteration n+1, reading chi, E and s once). W is a work buffer of the
 * size of E, distinct from E and Eout.
 *   mmx_born_prologue: W = FFT along the innermost axis of (chi_eff E + s_eff); call it before the first
 *                      mmx_born_iterate, and again whenever E, s, chi or coef changed outside mmx_born_iterate
 *                      (e.g. after a rejected update).
 *   mmx_born_iterate:  completes the iteration begun in W: Eout = E + d, norms as for mmx_born_step, and
 *                      W = FFT along the innermost axis of (chi_eff Eout + s_eff), the prologue of the next
 *                      iteration assuming Eout is accepted. If the caller rejects Eout it must re-run
 *                      mmx_born_prologue on E (with the new coef) before the next mmx_born_iterate.
 * Replaces Solution.__iter__ (solver.py:1064-1098) for callers that iterate many times.
 */
int mmx_born_prologue(mmx_plan_t plan, const void* E, void* W, const void* s, const void* chi, int ncomp,
                      int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride,
                      void* stream);
int mmx_born_iterate(mmx_plan_t plan, const void* E, void* Eout, void* W, const void* s, const void* chi, int ncomp,
                     int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride,
                     double* norms, void* stream);

/*
 * mmx_born_iterate with its norm reduction on a second stream: the passes run on `stream`; the FP64 sum of their
 * per-tile partials runs on `norm_stream` (a different stream of the same device) once `stream` has finished them,
 * so that work enqueued on `stream` next (the following iteration) does not wait for it. The partials go to one of
 * two buffers of the plan, chosen by `slot` (0 or 1), each with an event that the call records on `stream` and
 * that norm_stream waits for: a call may reuse a slot only once the reduction of the previous call with that slot
 * has run (alternate the slots and wait on norm_stream for call n before issuing call n + 2), since it overwrites
 * the slot's partials and records its event again. Eout, W and norms as for mmx_born_iterate; norms is written on
 * norm_stream, and copied from there to `norms_host` (2 nbatch doubles of pinned host memory; NULL: no copy).
 * Replaces Solution.__iter__ (solver.py:1064-1098), as mmx_born_iterate.
 */
int mmx_born_iterate_split(mmx_plan_t plan, const void* E, void* Eout, void* W, const void* s, const void* chi,
                           int ncomp, int chi_kind, const void* coef, int64_t nbatch, int64_t s_bstride,
                           int64_t chi_bstride, double* norms, double* norms_host, int slot, void* stream,
                           void* norm_stream);

/*
 * One Born iteration for a magnetic or bi-(an)isotropic medium (MacroMax Solution.__iter__ with the curl-term
 * susceptibility operator, solver.py:531-552 and 1064-1085; three components):
 *     chi X = chi_EE X + chi_HE curl X + curl(chi_EH X + chi_HH curl X),  d = chi (G (chi E + s) - E),  Eout = E + d,
 * curls with respect to the k0-scaled grid (i k x in k-space, backend/__init__.py:483-509), G the dyadic Green's
 * operator for alpha = alpha[0] + i alpha[1]. chi[4] = {chi_EE, chi_HE, chi_EH, chi_HH}: chi_kind[i] 1 for a scalar
 * field ([1,1,grid]), 2 for a 3x3 field ([3,3,grid] row-major); chi_vstride[i] 1 for a field, 0 for a constant (one
 * value or one 3x3 matrix). E, Eout, s: [3,1,grid]; work: 12 * grid complex values of scratch (distinct from E and
 * Eout); norms (device, 2 doubles): {sum |d|^2, sum |Eout|^2} in FP64. The caller applies the accept rule
 * (solver.py:1078-1096) as for mmx_born_step. Complex64 plans whose first and last axes have 64, 128, 256, 512 or
 * 1024 points (2D or 3D, register kernels on) run fused pencil passes: the k-space products inside the passes along
 * the first axis, the voxel-wise products inside those along the last axis (eight kernels per 2D iteration); other
 * plans run the transforms and five voxel-wise kernels separately. Both compute the same iteration.
 */
int mmx_magnetic_iterate(mmx_plan_t plan, const void* E, void* Eout, const void* s, const void* const* chi,
                         const int* chi_kind, const int64_t* chi_vstride, const double* alpha, void* work,
                         double* norms, void* stream);
/*
 * mmx_magnetic_iterate with its norm reduction on a second stream, from partial slot `slot`, and the norms copied to
 * `norms_host` (pinned, or NULL), with the rules of mmx_born_iterate_split. Replaces the same reference code as
 * mmx_magnetic_iterate.
 */
int mmx_magnetic_iterate_split(mmx_plan_t plan, const void* E, void* Eout, const void* s, const void* const* chi,
                               const int* chi_kind, const int64_t* chi_vstride, const double* alpha, void* work,
                               double* norms, double* norms_host, int slot, void* stream, void* norm_stream);

/*
 * Batched solves: nbatch problems of one grid shape (many right-hand sides of one medium when chi_bstride = 0, or
 * many media) iterate together with the accept/reject decision and the stop test taken on the device, so that a
 * run of calls can be captured in a graph and replayed without host synchronisation. No single reference
 * function corresponds: it replaces the host loop of Solution.__iter__ / Solution.solve (solver.py:1064-1123)
 * that ScatteringMatrix runs once per right-hand side (matrix.py:789-819), with the same rules per problem.
 *   EA, EB: [nbatch, C, 1, *grid] field buffers; problem b's current field is in EB when bit 0 of state[b] is set.
 *   W:      [nbatch, C, 1, *grid] work buffer (the first pass of the next iteration).
 *   coef:   nbatch records as for mmx_born_step; the device rewrites a record when it raises alpha_i.
 *   state:  nbatch words: bit 0 (MMX_STATE_PARITY) field in EB, bit 1 (MMX_STATE_DONE) finished, bit 2
 *           (MMX_STATE_STALE) W must be recomputed (set it for every problem before the first call).
 *   track:  nbatch records of MMX_TRACK_LEN doubles, indexed by MMX_TRACK_*; set PREV = +inf, ITER = 0, ALPHA_*,
 *           A = 1, C_* = 0 (or the coef values) before the first call.
 *   norms:  2 * nbatch doubles of scratch.
 * One call runs one iteration of every unfinished problem: the first pass where W is stale, the remaining passes,
 * then per problem: accept if ||d|| < ||d_prev|| (swap by the parity bit), else alpha_i <- 1.5 alpha_i
 * (solver.py:1083-1096); residue = ||d_prev|| / ||E||; done when !(residue > tol && iterations < max_iter).
 * The first call with a given plan and nbatch may allocate scratch; later calls allocate nothing and issue only
 * kernel launches on `stream`.
 */
#define MMX_STATE_PARITY 1u
#define MMX_STATE_DONE 2u
#define MMX_STATE_STALE 4u
#define MMX_TRACK_PREV 0
#define MMX_TRACK_FIELD_NORM 1
#define MMX_TRACK_RESIDUE 2
#define MMX_TRACK_ITER 3
#define MMX_TRACK_ALPHA_R 4
#define MMX_TRACK_ALPHA_I 5
#define MMX_TRACK_A 6
#define MMX_TRACK_C_R 7
#define MMX_TRACK_C_I 8
#define MMX_TRACK_DONE_CLOCK 9
#define MMX_TRACK_REJECTED 10
#define MMX_TRACK_LEN 12
int mmx_batch_step(mmx_plan_t plan, void* EA, void* EB, void* W, const void* s, const void* chi, int ncomp,
                   int chi_kind, void* coef, int64_t nbatch, int64_t s_bstride, int64_t chi_bstride, uint32_t* state,
                   double* track, double* norms, double tol, int64_t max_iter, void* stream);

/* Device wall clock (constant rate, see mmx_wall_clock_rate_khz) written to out[0] by a one-thread kernel on
 * `stream`; the batch controller stamps MMX_TRACK_DONE_CLOCK with the same clock. No reference counterpart
 * (the reference times solves on the host, examples/benchmark.py:66-84). */
int mmx_wall_clock(double* out, void* stream);
int mmx_wall_clock_rate_khz(int64_t* khz);

/*
 * Preconditioner set-up reductions over a complex128 permittivity eps ([1,1,grid] for eps_kind 1, [3,3,grid]
 * row-major for eps_kind 2), whatever the plan's dtype; `out` is 2 device doubles.
 *   mmx_setup_sigma: out[1] = max over voxels of the largest eigenvalue of (eps - shift I)^H (eps - shift I)
 *                    (calc_DeltaEE_beta2 / calc_sigmaEE, solver.py:392-401, shift = alpha_r beta real);
 *   mmx_setup_gain:  out[0], out[1] = min and max over voxels of the eigenvalues of -i/2 (eps - eps^H)
 *                    (has_gain / has_loss, solver.py:322-333).
 * Closed-form 3x3 Hermitian eigenvalues in FP64 (the reference's mat3_eigh, backend/__init__.py:692-850).
 */
int mmx_setup_sigma(mmx_plan_t plan, const void* eps, int eps_kind, double shift, double* out, void* stream);
int mmx_setup_gain(mmx_plan_t plan, const void* eps, int eps_kind, double* out, void* stream);

/* out[0] = sum |x|^2 over n complex elements, FP64 accumulation (backend/__init__.py:872-874, torch.py:316-317). */
int mmx_norm2(mmx_plan_t plan, const void* x, int64_t n, double* out, void* stream);

/*
 * y = chi x per voxel for `nbatch` fields x of ncomp components (BackEnd.mul(chi, x), backend/__init__.py:415-444,
 * torch.py:220-249, for chi of shape [1,1,grid], [3,3,grid], packed symmetric [6,grid] or uniaxial [4,grid], chi_kind
 * 1 to 4 as for mmx_born_step). y may alias x.
 */
int mmx_chi_mul(mmx_plan_t plan, const void* chi, int chi_kind, const void* x, void* y, int ncomp,
                int64_t nbatch, void* stream);

/*
 * out = Pi_L F = k (k . F) / k^2 with Pi_L(k=0) = I on the first ndim components and 0 on the others
 * (BackEnd.longitudinal_projection_ft, backend/__init__.py:606-649). F, out: [3][grid]; out may alias F.
 */
int mmx_longitudinal_projection_ft(mmx_plan_t plan, const void* F, void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MACROMAX_HIP_H */

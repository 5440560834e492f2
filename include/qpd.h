/*
 * qpd.h - C ABI of the MI355X kinematic-diffraction library (libqpd.so).
 *
 * One call computes the diffraction intensity of a volume:
 *
 *   A(k) = S(k) sum_a exp(-i (k + G).R_a) / V [FFT(fE_a)(k)/n - i k . FFT(fE_a u_a)(k)/n]
 *   fE_a(r) = sum_p w_p(r) f_pa exp(-i qC . u_ap(r)),   u_ap = u(r) + su_pa . P(r)
 *   I(k) = fftshift(max(|A(k)|^2, 1e-50))
 *
 * which is the computation of src/python/diffraction.py:62-113 and of the
 * source / fft.forward / accumulate / intensity sequence of
 * src/heffte/qpop_diffraction_mgpu.cu:160-186 in the reference.
 *
 * Conventions: every pointer argument named *_dev or documented as device memory
 * is device memory owned by the caller; the library owns only its plan tables.
 * Work is enqueued on the stream passed in (a hipStream_t, NULL = default
 * stream) and is asynchronous: no call synchronises the device except
 * qpd_plan_create / qpd_plan_destroy.  Functions return QPD_OK (0) or a negative
 * code; qpd_last_error() gives the message of the calling thread's last error.
 * A plan may not be used by two host threads at once; distinct plans may run
 * concurrently on distinct streams.
 *
 * Memory layout: a grid is given as three memory dims, dims[0] fastest; axis[d]
 * is the physical axis (0 = x, 1 = y, 2 = z) of memory dim d.  FerroDyn files
 * and src/heffte use x fastest: dims = {nx, ny, nz}, axis = {0, 1, 2}.  The
 * reference's Python arrays are (x, y, z) C order, z fastest: dims = {nz, ny, nx},
 * axis = {2, 1, 0}.  Fields and the intensity share the grid's layout.
 */
#ifndef QPD_H_
#define QPD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QPD_OK 0
#define QPD_ERR_ARG (-1)
#define QPD_ERR_UNSUPPORTED (-2)
#define QPD_ERR_HIP (-3)
#define QPD_ERR_NOMEM (-4)

#define QPD_MAX_ATOMS 64
#define QPD_MAX_PHASES 16

typedef struct qpd_plan qpd_plan;

/* Grid of one volume; see the layout note above.  h[] is the spacing of the
 * physical axes x, y, z in nm (SizeContext.dx/dy/dz, size_context.py:178-180). */
typedef struct qpd_grid {
    int64_t dims[3];
    int axis[3];
    double h[3];
} qpd_grid;

/* The crystal of parameter.atom.in after AtomList.setup (atom_list.py:37-193) and
 * reciprocal_centre (diffraction.py:36-48).  Host pointers, copied at plan creation. */
typedef struct qpd_crystal {
    int n_atom;
    int n_phase;
    int n_struc;
    const double* R;     /* (n_atom, 3) atom positions xAtom . aC, nm           (diffraction.py:94) */
    const double* f_re;  /* (n_phase, n_atom) form factors, real part            (atom_list.py:102-111) */
    const double* f_im;  /* (n_phase, n_atom) imaginary part */
    const double* su;    /* (3, n_struc, n_phase, n_atom) shift per unit P, nm   (diffraction.py:93) */
    double G[3];         /* reciprocal lattice point, nm^-1 */
    double qC[3];        /* strain-corrected centre, nm^-1 */
    double V;            /* cell volume det(aC), nm^3 */
} qpd_crystal;

/* Per-cell inputs of one volume, device memory, in the grid's layout, n cells each.
 * Element type "real" is float for a precision-32 plan and double for precision 64. */
typedef struct qpd_fields {
    const void* P;                 /* (n_struc, n) real: order parameter, C/m^2     (diffraction.py:89) */
    const double* U;               /* (3, n) double: displacement in nm, physical x, y, z components,
                                      u_z reference ramp included                     (cli.py:215-217) */
    const signed char* phase;      /* (n) phase index of the cell, -1 = does not scatter
                                                                                      (kernels.cuh:55-66) */
    const void* fraction;          /* (n_phase, n) real phase fractions, or NULL = one-hot from phase */
    const void* region;            /* (n) real scattering weight, or NULL = 1 */
} qpd_fields;

/* Field ingest: the raw, cropped FerroDyn fields of a run as read from the files (file
 * order, x fastest, each component its own array, in the files' own precision) and the
 * material ids of the crop's columns over the full stack.  qpd_ingest_fields turns them
 * into the plan inputs of the x-fastest layout (dims {nx, ny, nz}, axes {0, 1, 2}):
 *   P_out (3, n) = P / P_div in the plan precision,
 *   U_out (3, n) = u / U_div in double, plus on z the reference ramp
 *                  (cumsum_z dc[mat] - dc[mat] / 2) dz summed over the full stack from z = 0,
 *   phase_out (n) = phase_of[mat],
 * with the operations in the order of the reference CLI (cli.py:189-217; the native
 * `columns` kernel, kernels.cuh:34-50), so the results equal the host preparation bit for
 * bit.  Material ids outside [0, n_material) give dc = 0 and phase -1.  All pointers are
 * device memory; dc and phase_of hold n_material entries. */
typedef struct qpd_ingest {
    int64_t dims[3];               /* crop extents nx, ny, nz */
    int64_t nz_full;               /* z extent of the material field */
    int64_t z0;                    /* first z of the crop in the full stack */
    const void* P[3];              /* cropped P_x, P_y, P_z */
    int P_bits;                    /* 32 or 64: precision of the P arrays */
    double P_div;                  /* order-parameter unit (OS0) */
    const void* U[3];              /* cropped u_x, u_y, u_z in m, or NULL: no displacement file */
    int U_bits;                    /* 32 or 64 */
    double U_div;                  /* metres per nm (L0) */
    const int32_t* material;       /* (nz_full, ny, nx) material id of each cell of the crop's columns */
    int n_material;
    const double* dc;              /* c_m(T) / aC_z - 1 per material (size_context reference strain) */
    const signed char* phase_of;   /* phase index per material, -1 = does not scatter */
    double dz;                     /* nm */
} qpd_ingest;

int qpd_ingest_fields(const qpd_ingest* in, int precision, void* P_out, double* U_out, signed char* phase_out,
                      void* stream);

/* Sums of a float64 array A (dims[0] fastest) over each memory dim, for reciprocal-space
 * maps (rsm.py:88-113: with the default bins on the rectilinear q grid, the `sum` map of
 * two axes is the intensity summed over the third).  S0 (dims[1] x dims[2], dims[1]
 * fastest) sums over dims[0], S1 (dims[0] x dims[2]) over dims[1], S2 (dims[0] x dims[1])
 * over dims[2]; a NULL output is skipped.  Each sum runs in a fixed order. */
int qpd_project(const int64_t dims[3], const double* A, double* S0, double* S1, double* S2, void* stream);

/* Reverse the axis order of n_comp fields of a grid: src holds n_comp arrays of dims[0] x
 * dims[1] x dims[2] elements (dims[0] fastest, elem_bytes 1, 4 or 8 each), dst receives them
 * with dims[2] fastest (element (i0, i1, i2) moves from i0 + dims[0] (i1 + dims[1] i2) to
 * i2 + dims[2] (i1 + dims[1] i0)).  Device memory; src and dst must not overlap.  The CLI uses
 * it to run x-fastest fields z fastest when the x rows are longer than every compile-time
 * source kernel, and to return the intensity to x fastest.  No counterpart in the reference,
 * whose arrays keep one layout (cli.py:189-217). */
int qpd_reverse_axes(const int64_t dims[3], int n_comp, int elem_bytes, const void* src, void* dst, void* stream);

/* Create a plan for one grid, crystal and precision (32: complex FP32 bulk with FP64
 * phase reduction and FP64-generated twiddles; 64: complex FP64 throughout).
 * Replaces the per-call setup of diffraction.py:75-94 and the HeFFTe plan of
 * qpop_diffraction_mgpu.cu:154-157.  Allocates and uploads the plan tables. */
int qpd_plan_create(qpd_plan** plan, const qpd_grid* grid, const qpd_crystal* crystal, int precision);

/* Plan options of qpd_plan_create_opts.
 *
 * Pass order: pass 0 always transforms the contiguous dims[0] (and forms the source);
 * passes 1 and 2 transform the two strided dims, in memory order (dims[1] then dims[2]) or
 * swapped.  QPD_ORDER_AUTO keeps an order in which the cell's atoms form source groups, then
 * takes the order of the smaller error estimate below.
 *
 * Arithmetic of a precision-32 plan (P, fractions and region are FP32 either way):
 *   QPD_ARITH_FP32    complex FP32 transforms and spectra (FP64 phase reduction and
 *                     FP64-generated tables);
 *   QPD_ARITH_MIXED   complex FP64 transforms, FP32 spectra between passes (the memory traffic
 *                     of QPD_ARITH_FP32);
 *   QPD_ARITH_FP64    FP64 transforms and spectra, the FP32 inputs widened on the device first;
 *   QPD_ARITH_AUTO    QPD_ARITH_MIXED when its error estimate is at most 6e-4, else
 *                     QPD_ARITH_FP64.
 * An FP32 transform of a line that holds a strong peak leaves errors of order 1e-7 of the
 * peak amplitude on its other outputs, |dlog10 I| of about 1e-3 at voxels 1e-10 of the peak
 * intensity, and where the lines of a volume are alike they repeat from line to line and the
 * later transforms do not average them down; the transforms therefore run in FP64.  The
 * rounding of FP32 spectra is relative to each stored value; the FP64 transforms after it
 * spread it out: |dlog10 I| of about hypot(5.8e-3 / sqrt(N2), 3.5e-3 / sqrt(N1 N2)) with N2
 * the length of the last pass and N1 that of pass 1 (a term vanishes when its length or
 * plane is one cell).  A precision-64 plan is FP64 throughout (QPD_ARITH_AUTO or
 * QPD_ARITH_FP64).  No counterpart in the reference (complex128 throughout,
 * diffraction.py:97-111). */
#define QPD_ORDER_AUTO 0
#define QPD_ORDER_MEMORY 1
#define QPD_ORDER_SWAPPED 2
#define QPD_ARITH_AUTO 0
#define QPD_ARITH_FP32 1
#define QPD_ARITH_MIXED 2
#define QPD_ARITH_FP64 3

/* qpd_plan_create with explicit pass order and arithmetic (qpd_plan_create is this call with
 * QPD_ORDER_AUTO and QPD_ARITH_AUTO). */
int qpd_plan_create_opts(qpd_plan** plan, const qpd_grid* grid, const qpd_crystal* crystal, int precision, int order,
                         int arith);

/* What a plan chose: the memory dim each pass transforms (pass_dims[0] = 0), the precision
 * of its transforms and of the spectra stored between passes (32 or 64), and the error
 * estimate of that arithmetic (max |dlog10 I| over voxels >= 1e-10 of the peak; 0 when
 * everything is FP64, -1 for QPD_ARITH_FP32, which has no bound).  NULL outputs are skipped. */
int qpd_plan_info(const qpd_plan* plan, int pass_dims[3], int* compute_bits, int* storage_bits,
                  double* dlog10_estimate);

/* The bits of each pass's transforms (32 or 64), in pass order (qpd_plan_info's pass_dims).
 * The automatic arithmetic of an FP32 plan may give FP32 transforms to long strided passes
 * (about 3200 points and more) while its error estimate stays within its bound; its
 * compute_bits above is pass 0's.  No counterpart in the reference. */
int qpd_plan_pass_bits(const qpd_plan* plan, int bits[3]);

/* Release a plan (synchronises the device). */
int qpd_plan_destroy(qpd_plan* plan);

/* Bytes of device workspace qpd_intensity needs (intermediate spectra). */
size_t qpd_workspace_bytes(const qpd_plan* plan);

/* Number of kernel launches one qpd_intensity call enqueues (a split pass launches twice). */
int qpd_launches_per_volume(const qpd_plan* plan);

/* Launch schedules.  Each pass of a plan can run one of several schedules: a
 * compile-time-length kernel of this build (every role in FP32; the source, strided and
 * intensity roles in FP64, named "... f64"), in FP32 a long strided length split into
 * two launches of registered lengths N1 x N2 (N2-point sub-transforms with the twiddle
 * in place, then N1-point transforms with the pass's accumulation), or the generic
 * run-time-radix kernel.  Schedule 0 is the default; an autotuner times the others
 * and selects one.  Selecting is a host-side change that applies to calls enqueued
 * afterwards (a captured graph keeps the launches it recorded).  No counterpart in the
 * reference, which delegates its transforms to cuFFT / HeFFTe (qpop_diffraction_mgpu.cu:
 * 154-176); these entry points replace those libraries' plan-time algorithm choice. */
int qpd_plan_schedules(const qpd_plan* plan, int pass);      /* how many schedules pass has */
int qpd_plan_schedule_name(const qpd_plan* plan, int pass, int index, char* buf, int cap);
                                                             /* name into buf; returns its length */
int qpd_plan_set_schedule(qpd_plan* plan, int pass, int index);
int qpd_plan_get_schedule(const qpd_plan* plan, int pass);   /* the selected index */

/* Lengths with a compile-time FP32 kernel for a role (0 source pass, 1 contiguous transform,
 * 2 strided 3 -> 2 pass, 3 strided pass to the intensity, 4 strided transform, 5 first
 * step of a split length, 6 and 7 roles 2 and 3 as the second step of a split length,
 * 8 and 9 the strided pass over source groups with one or two z-classes, 10 and 11 those
 * as the second step of a split length, 12 the source pass over source groups with one
 * wave per row), ascending:
 * writes up to cap of them and returns how many there are.  No counterpart in the
 * reference. */
int qpd_fast_lengths(int role, int64_t* lengths, int cap);

/* The intensity of one volume: fftshifted max(|A|^2, 1e-50) as double, n values in
 * the grid's layout, written to I_dev.  work_dev holds qpd_workspace_bytes(plan).
 * Replaces diffraction.py:96-112 and qpop_diffraction_mgpu.cu:166-186. */
int qpd_intensity(qpd_plan* plan, const qpd_fields* fields, double* I_dev, void* work_dev, void* stream);

/* Enqueue only pass `pass` (0, 1 or 2) of qpd_intensity: 0 = source + transform along
 * dims[0], 1 and 2 = transforms along the strided dims in the plan's pass order, the last
 * with the intensity (qpd_plan_info gives the order).  Has no
 * counterpart in the reference; it exists so that each kernel can be timed on its own.
 * Running passes 0, 1, 2 in order is exactly qpd_intensity. */
int qpd_intensity_pass(qpd_plan* plan, const qpd_fields* fields, double* I_dev, void* work_dev, int pass,
                       void* stream);

/* qpd_intensity without a workspace argument.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The intermediate spectra go to a static device
 * array of QPD_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
 * process ends; a plan whose qpd_workspace_bytes exceeds it is rejected with
 * QPD_ERR_UNSUPPORTED before anything is enqueued.  The call allocates nothing (no hipMalloc,
 * no hipFree); it uses the current device, which must be the plan's.  All calls on one GPU
 * share that GPU's array, which keeps the last call's contents: calls that overlap on one GPU
 * (enqueued on different streams, or from several threads without ordering) race, so keep
 * them on one stream.  No part of the array has to start at zero: each pass writes every
 * intermediate before the next pass reads it. */
#define QPD_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int qpd_intensity_noalloc(qpd_plan* plan, const qpd_fields* fields, double* I_dev, void* stream);

/* Algorithmic bytes one launch of pass `pass` moves per cell: each input read once and
 * each output written once (fields, intensity and source inputs as the plan uses them). */
double qpd_pass_bytes_per_cell(const qpd_plan* plan, const qpd_fields* fields, int pass);

/* Forward, unnormalised, in-place 3-D DFT (exponent -2 pi i jk/N on every dim) of
 * n_fields complex fields stored one after another, dims[0] fastest; complex
 * interleaved float (precision 32) or double (64).  The transform that
 * xp.fft.fftn performs at diffraction.py:105-106, exposed for testing. */
int qpd_fft3d(const int64_t dims[3], int precision, void* data_dev, int n_fields, void* stream);

/* Whether this build can transform a 1-D length: any length up to 65536 (prime factors up to
 * 17 as butterfly stages, larger ones as direct stages of O(R) work per output).  A plan
 * needs more: a launch shape for the axis the length is on (contiguous rows up to 8192
 * cells; FP32 strided axes of any such length, split in two launches where it factors into
 * lengths with kernels of their own, else up to 8192; FP64 strided axes up to 8192);
 * qpd_plan_create reports QPD_ERR_UNSUPPORTED otherwise.  Replaces the any-length transforms
 * of xp.fft.fftn (diffraction.py:105-106) and HeFFTe (qpop_diffraction_mgpu.cu:154-157). */
int qpd_fft_length_supported(int64_t n);

/* Message of the calling thread's last error ("" when none). */
const char* qpd_last_error(void);

/* ABI version of the library (major * 100 + minor). */
int qpd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* QPD_H_ */

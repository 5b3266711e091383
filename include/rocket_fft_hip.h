/*
 * rocket_fft_hip.h -- C ABI of the MI355X (gfx950) transform engine.
 *
 * The ten numba_* entry points keep the names, argument order and meaning of
 * rocket-fft's own C ABI (rocket_fft/_pocketfft_numba.cpp:25-223), which
 * rocket-fft's Numba intrinsics bind by symbol name with the LLVM prototypes
 * of rocket_fft/pocketfft.py:32-128 (i64 for integers, i1 for bools, double
 * for fct, i8* for arrays).  Arrays are Numba array structs: shape first,
 * then strides in BYTES (numba _arraystruct.h).
 *
 * Memory: data pointers may be device memory (hipMalloc / torch CUDA
 * tensors: computed in place on the GPU) or ordinary host memory (staged
 * through device buffers, then copied back; correctness path).
 *
 * Semantics follow the reference engine exactly: unnormalised transforms
 * times fct, fct applied once (first axis), zero-size arrays are a no-op,
 * in-place (ain == aout) allowed for c2c, dct, dst, r2r_*.  Like the
 * reference, the functions return nothing; a failure (invalid DCT type,
 * HIP error, unsupported geometry) leaves aout unspecified, prints one line
 * to stderr and is reported by rfh_last_error() on the calling thread.
 *
 * By default each call completes before it returns (the reference is
 * synchronous).  rfh_set_stream() lets a thread run its calls
 * asynchronously on its own HIP stream instead.
 */
#ifndef ROCKET_FFT_HIP_H
#define ROCKET_FFT_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Numba's array struct (numba/_arraystruct.h); shape_and_strides holds
 * 2*ndim entries: the shape, then the strides in bytes. */
typedef struct {
    void *meminfo;
    void *parent;
    int64_t nitems;
    int64_t itemsize;
    void *data;
    int64_t shape_and_strides[];
} arystruct_t;

/* ---- drop-in entry points (rocket_fft/_pocketfft_numba.cpp) ---------------- */

/* _pocketfft_numba.cpp:25-29: smallest 2^a3^b5^c7^d11^e >= target (real:
 * 2^a3^b5^c), target itself below 13 (real: below 7).  Bit-exact. */
uint64_t numba_good_size(uint64_t target, bool real);

/* _pocketfft_numba.cpp:31-49: complex -> complex over axes; itemsize 16 of
 * ain selects complex128, else complex64. */
void numba_c2c(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               bool forward, double fct, uint64_t nthreads);

/* _pocketfft_numba.cpp:51-69: DCT type 1..4 over axes; itemsize 8 -> float64. */
void numba_dct(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               uint64_t type, double fct, bool ortho, uint64_t nthreads);

/* _pocketfft_numba.cpp:71-89: DST type 1..4 over axes. */
void numba_dst(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               uint64_t type, double fct, bool ortho, uint64_t nthreads);

/* _pocketfft_numba.cpp:91-109: real -> half complex (last axis n/2+1),
 * shape taken from ain; conjugated output when !forward. */
void numba_r2c(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               bool forward, double fct, uint64_t nthreads);

/* _pocketfft_numba.cpp:111-143: r2c into a full-size complex aout, then the
 * Hermitian mirror fills the other half. */
void numba_c2c_sym(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout,
                   arystruct_t *axes, bool forward, double fct, uint64_t nthreads);

/* _pocketfft_numba.cpp:145-163: half complex -> real; shape taken from aout. */
void numba_c2r(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               bool forward, double fct, uint64_t nthreads);

/* _pocketfft_numba.cpp:165-183: FFTPACK halfcomplex real transforms. */
void numba_r2r_fftpack(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout,
                       arystruct_t *axes, bool real2hermitian, bool forward, double fct,
                       uint64_t nthreads);

/* _pocketfft_numba.cpp:185-203: Hartley transform, axis by axis. */
void numba_r2r_separable_hartley(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout,
                                 arystruct_t *axes, double fct, uint64_t nthreads);

/* _pocketfft_numba.cpp:205-223: true N-D Hartley transform. */
void numba_r2r_genuine_hartley(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout,
                               arystruct_t *axes, double fct, uint64_t nthreads);

/* ---- additions without a reference counterpart ------------------------------ */

/* Last error of the calling thread: 0 = none.  Copies the message (NUL
 * terminated, truncated to len) when buf is non-null.  rfh_clear_error()
 * resets it.  For tests and callers that want to check the void entry
 * points above. */
int rfh_last_error(char *buf, size_t len);
void rfh_clear_error(void);

/* Run this thread's subsequent calls on `stream` (a hipStream_t; NULL = the
 * default stream).  blocking = true (the default) makes every call wait for
 * its work; false returns after enqueueing (device pointers only; host
 * arrays always block).  Returns 0 on success. */
int rfh_set_stream(void *stream, bool blocking);

/* Enables (default) or disables the compile-time specialised line kernels
 * for power-of-two lengths; disabled, every transform runs the generic
 * kernels.  Process-wide; meant for tests that compare the two paths. */
int rfh_set_fast_path(bool enabled);

/* ---- measurement --------------------------------------------------------------
 * No reference counterpart.  rfh_kernel_timing(true) brackets every line-kernel
 * launch with HIP events on the stream it is launched on (not while a stream is
 * being captured into a graph: leave it off then).  rfh_kernel_timing_report
 * waits for the recorded launches, forgets them, and writes one line per kernel
 * label: "family op n=<length> c64|c128 [strided]\t<launches>\t<total ms>\t
 * <algorithmic bytes: each launch's lines read once and written once>".
 * Returns the report's length (the text is truncated to len - 1). */
int rfh_kernel_timing(bool enabled);
int rfh_kernel_timing_report(char *buf, size_t len);

/* ---- shifted transforms -----------------------------------------------------
 * numba_c2c with numpy.roll shifts fused into the transform:
 *   aout = roll(c2c(roll(ain, shift_in, axes)), shift_out, axes),
 * shift_in[i] / shift_out[i] (any sign; NULL: none) rolling along axes[i],
 * each axis at most once.  fftshift is a roll by n // 2, ifftshift by
 * -(n // 2) (overloads.py:1253-1310 compose them with the transforms as
 * separate calls).  The rolls become index rotations in the loads of each
 * axis's first pass and the stores of its last pass, so no extra pass runs
 * and the result equals the composition bit for bit.  Fusable: lengths whose
 * plan is one LDS pass, every Bluestein length, and four-step (multi-pass)
 * lengths n1 x n2 whose input roll is a multiple of n2 and output roll a
 * multiple of n1 (every fftshift / ifftshift of an even length with both
 * factors even); other rolls of those lengths return 5 with nothing
 * written, and the caller
 * runs the rolls and the transform separately.  Complex64 transforms run
 * on the FP32 kernels whatever rfh_set_precision says.  No reference
 * counterpart.  Returns 0, 5 (not fusable), or another error code. */
int rfh_c2c_shifted(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
                    bool forward, double fct, uint64_t nthreads, const int64_t *shift_in,
                    const int64_t *shift_out);

/* ---- plan cache -------------------------------------------------------------
 * Plans (radix schedules, FP64-generated twiddle tables, the Bluestein
 * kernel's spectrum) are built once per (device, length, precision, LDS
 * budget, contiguous / interleaved lines) and kept for the process, as the
 * reference keeps its plans in pocketfft's LRU cache
 * (_pocketfft_hdronly.h:3170-3222, get_plan).  rfh_plan_cache adds an
 * on-disk cache of the host records under `dir`, keyed by transform kind,
 * length, precision, LDS budget, stride class and GPU architecture, so that a
 * later process skips the table generation; dir NULL or "" turns it off.
 * Files are checksummed and written atomically; one that does not check is
 * ignored and rewritten.  clear_memory also drops this process's plans from
 * the cache (the next transform of each shape rebuilds or reloads its plan;
 * the dropped plans' device tables stay allocated until the process ends, so
 * work already enqueued and graphs already captured stay valid).  No
 * reference counterpart for the disk part.  Returns 0 or an error code. */
int rfh_plan_cache(const char *dir, bool clear_memory);
/* Plans loaded from / written to the disk cache since the library loaded. */
int rfh_plan_cache_stats(uint64_t *hits, uint64_t *writes);
/* Host-only (no GPU needed), for tests of the file format: builds the plan
 * record of (n, precision, lds_budget > 0 bytes, interleaved, real), writes
 * it to `path`, reads it back and compares.  0: identical; 1: the file did
 * not read back; 2: a field differs; -1: error (rfh_last_error). */
int rfh_plan_record_roundtrip(uint64_t n, bool double_precision, uint64_t lds_budget,
                              bool interleaved, bool real, const char *path);

/* ---- allocation-free variant --------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): numba_c2c's transform with the same arguments (nthreads
 * dropped), on device (or managed) memory of one GPU only, enqueued on
 * `stream` without synchronising.  Its four-step intermediates live in a
 * 64 MiB static device array of the library (one per GPU, held until the
 * process ends) instead of an allocated workspace; a transform needing more
 * than that is rejected with an error.  Plans and matrix-core tables are
 * taken from the caches only: run numba_c2c once with the same shape (and
 * precision mode) first, else the call fails.  Calls that overlap on one GPU
 * share the array and race.  Returns 0, or the error code of rfh_last_error. */
int rfh_c2c_noalloc(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
                    bool forward, double fct, void *stream);

/* ---- split-precision form ---------------------------------------------------
 * Process-wide arithmetic of complex64 transforms of length 4096 on
 * contiguous lines: 0 = plain FP32 butterflies (default), 1 = radix-16 stages
 * as DFT matrix products on FP16 matrix cores with hi/lo operand splits and
 * FP32 accumulation, 2 = the same on BF16 with hi/mid/lo splits, 3 = BF16 with
 * hi/lo splits only (precision study; about 1e-5).  No reference
 * counterpart.  Returns the previous mode. */
int rfh_set_precision(int mode);
/* Number of matrix-core (split-precision) kernel launches since the library
 * was loaded, in precision mode `mode` (1..3), or in all modes (0); lets a
 * caller or test confirm that a transform actually ran on the matrix cores
 * rather than on the FP32 butterflies.  No reference counterpart. */
uint64_t rfh_mfma_launches(int mode);
/* The DFT-16 operand table the matrix-core kernels of mode 1..3 load for
 * direction -1 (forward) or +1: per split part t, per {A_re, A_im}, per lane
 * l < 64, eight values ([(2 t + part) * 64 + l] * 8 + i), each exact in the
 * operand format, generated from FP64 roots on the host.  Writes at most cap
 * floats to out and returns the table's length (-1: bad mode or direction).
 * Host-only (no GPU needed).  No reference counterpart. */
int rfh_mfma_operands(int mode, int direction, float *out, int cap);
/* One 16x16x32 matrix-core product on device buffers laid out as the
 * hardware operands (a, b: 64 lanes x 8 floats, d: 64 x 4; bf16 != 0 selects
 * BF16, else FP16), for tests of the operand layout. */
int rfh_mfma_probe(int bf16, const float *a, const float *b, float *d, void *stream);

/* ---- ragged batches ---------------------------------------------------------
 * Many independent transforms of uneven lengths and kinds (complex c2c, real
 * r2c / c2r), planned once and executed in one persistent launch (a job
 * queue drained by resident workgroups, largest job first).  No reference
 * counterpart: rocket-fft calls its engine once per array.  Each job is one
 * contiguous line in device memory; fct, forward and the r2c/c2r
 * conjugation rules mean what they mean in numba_c2c / numba_r2c /
 * numba_c2r.  Execution allocates nothing and can be captured in a HIP
 * graph; the buffers named at creation are the ones every execution uses.
 * autotune = true times LDS budgets and grid sizes on the batch and keeps
 * the fastest (remembered per batch signature for the process). */
enum { RFH_JOB_C2C = 0, RFH_JOB_R2C = 1, RFH_JOB_C2R = 2 };
typedef struct {
    int32_t kind;     /* RFH_JOB_* */
    int32_t forward;
    uint64_t n;       /* transform length (the real length for r2c / c2r) */
    const void *in;   /* device memory, one contiguous line */
    void *out;        /* device memory, one contiguous line */
    double fct;
} rfh_job_t;
int rfh_batch_create(const rfh_job_t *jobs, int64_t njobs, bool double_precision,
                     bool autotune, void *stream, void **handle);
int rfh_batch_execute(void *handle, void *stream);
/* The same jobs launched one transform at a time through the per-call
 * engine (the comparison point of the persistent launch). */
int rfh_batch_execute_serial(void *handle, void *stream);
/* autotuned: 0 = not tuned, 1 = timed by this call, 2 = result remembered by
 * this process, 3 = read from the cache directory. */
int rfh_batch_info(void *handle, int *grid, int *lds_bytes, int *autotuned, double *tuned_ms);
/* Directory of the on-disk autotune cache (NULL or "" = none, the default):
 * results are read from and appended to <dir>/ragged_tune_<arch>_<CUs>cu.txt,
 * keyed by a 128-bit hash of the batch signature (precision, kind and length
 * of every job).  clear_memory forgets this process's remembered results.
 * The directory must exist; an unwritable one only disables writing.
 * Process-wide. */
int rfh_tune_cache(const char *dir, bool clear_memory);
int rfh_batch_destroy(void *handle);

/* ---- per-line scaling (FFTLog transforms) -----------------------------------
 * In place on device memory: nlines lines of len elements, line starts
 * `stride` elements apart.  mode 0: real lines, x[l][j] *= w[j] (w: len
 * doubles); mode 1: complex lines, X[l][k] = conj(X[l][k]) * c[k] (w: len
 * complex values as 2*len doubles).  w is host memory, rounded to the data's
 * precision.  Completes before returning.  No reference counterpart: it is
 * the spectrum and bias step of the fht / ifht overloads
 * (rocket_fft/overloads.py:858-902, 1755-1809) on the GPU. */
int rfh_line_scale(void *data, int64_t nlines, int64_t len, int64_t stride, const double *w,
                   int mode, bool double_precision, void *stream);

/* Cyclic shift, numpy.roll over every axis at once: out (C-contiguous, shape
 * `shape`) [i] = in[(i_k - shifts[k]) mod shape[k]], input strides in
 * elements (any sign), elements of 1, 2, 4, 8 or 16 bytes, device memory,
 * enqueued on `stream` (out must not overlap in).  No reference counterpart
 * in the C ABI: it is the roll / fftshift / ifftshift overloads
 * (rocket_fft/overloads.py:762-855, 1253-1310) on the GPU. */
int rfh_roll(const void *in, void *out, int ndim, const int64_t *shape, const int64_t *in_strides,
             const int64_t *shifts, int64_t elem_bytes, void *stream);

/* Plan facts of the reference engine, restated bit-exactly:
 * complex FFTPACK factor list (_pocketfft_hdronly.h:1750-1778), real FFTPACK
 * factor list (:2629-2652), and the FFTPACK-vs-Bluestein decision
 * (:2843-2861 complex, :2885-2903 real) with its Bluestein length n2
 * (:2765-2767).  factors return the count (entries beyond cap dropped);
 * plan_choice returns 1 for Bluestein (n2 written) and 0 for FFTPACK. */
int rfh_cfftp_factors(uint64_t n, uint64_t *out, int cap);
int rfh_rfftp_factors(uint64_t n, uint64_t *out, int cap);
int rfh_plan_choice(uint64_t n, bool real, uint64_t *n2);

/* GPU execution schedule chosen for a complex transform of length n over
 * contiguous lines (double_precision selects complex128): kind 0 = single
 * pass in LDS, 1 = Bluestein in LDS, 2 = two-pass (four-step), 3 = Bluestein
 * two-pass, 4 = multi-pass (a column pass of n1, then the rows of n2 = n / n1
 * by their own two- or multi-pass plan), 5 = Bluestein around multi-pass
 * transforms of its padded length; radices of the (first) LDS pass in
 * execution order; n1, n2 = the split (0 when single pass).  Returns the
 * number of radices. */
int rfh_gpu_schedule(uint64_t n, bool double_precision, int *kind, int *radices, int cap,
                     uint64_t *n1, uint64_t *n2);

#ifdef __cplusplus
}
#endif

#endif /* ROCKET_FFT_HIP_H */

/*
 * gfftx.h -- C ABI of the MI355X (gfx950) FFT engine behind GFFT's plan-and-call surface.
 *
 * The reference (scpp/gfft 0.32) is a header-only C++ library: a transform is planned by
 * instantiating GenerateTransform<...> / Transform<...> and run by the virtual
 * fft(T* data) (in place) or fft(const T* src, T* dst) (out of place).  This header is the
 * thin C layer a new GPU parallelisation policy calls from those fft() members; the C++
 * policy layer in include/gfftx/gfft.hpp is that caller.
 *
 * Conventions (identical to GFFT):
 *   - IDs: value types DOUBLE=0, FLOAT=1, COMPLEX_DOUBLE=2, COMPLEX_FLOAT=3
 *     (src/gfftpolicy.h:79,95,111,127); transform types DFT=0, IDFT=1, RDFT=2, IRDFT=3
 *     (src/gfftpolicy.h:226,245,263,283); places IN_PLACE=0, OUT_OF_PLACE=1
 *     (src/gfftpolicy.h:144,181).
 *   - Data are interleaved (re,im) pairs; std::complex<T> has the same layout.
 *   - DFT computes X_k = sum_n x_n exp(-2 pi i n k / N), unnormalised; IDFT uses exp(+) and
 *     multiplies by 1/N (src/metaroot.h:104-120, src/gfftswap.h:339-374).  Output is in
 *     natural order.
 *   - n is the complex length.  For RDFT/IRDFT n is half the real length: the transform
 *     maps 2n reals to n complex values in Numerical-Recipes packing
 *     [X_0, X_n, Re X_1, Im X_1, ..., Re X_{n-1}, Im X_{n-1}] (src/gfftswap.h:239-287);
 *     GFFTX_FLAG_R2C_BINS selects the n+1-bin layout instead (the buffer then holds
 *     2n+2 scalars on the complex side).
 *
 * Lengths: any n >= 2 (GFFT's out-of-place path takes any N, src/gfftalg.h:459-507).
 * Products of 2, 3, 5 and 7 run as Stockham passes (FP32 up to 2^28, FP64 up to 2^26); any
 * other n runs as a chirp-z convolution of a 2/3/5-smooth length m >= 2n-1 with the same
 * limits on m.  RDFT/IRDFT need an even n (GFFT's Separate assumes it,
 * src/gfftswap.h:255,279).
 *
 * Pointers may be device memory (the call is enqueued on `stream` and returns) or host
 * memory (the call copies through pinned staging buffers in chunks, overlapping the host
 * copies, both DMA directions and the transform of earlier batch items, and returns after
 * the result is back, as GFFT's synchronous fft() requires; GFFTX_STAGE_CHUNK_MB sets the
 * chunk, default 16).  `stream` is a hipStream_t; NULL is the null stream.  Calls on one plan
 * may come from several threads (GFFT's fft() is reentrant): the plan runs them one at a
 * time, a host-pointer call until its result is back, a device-pointer call until its work
 * is enqueued.  Calls on one plan from different streams are safe: when the plan's calls
 * share device arrays (multi-pass scratch, chirp-z work arrays, graphs, ragged batches), a
 * call on another stream than the previous call's first waits until the device has finished
 * its earlier work (a host wait, only at such a switch; calls on one stream cost nothing
 * extra).  A call on a stream the caller is capturing into a graph is not ordered that way;
 * the caller orders the replays of its own graphs.
 * No function throws; each returns a gfftx_status.  gfftx_last_error() gives detail for the
 * calling thread.
 */
#ifndef GFFTX_H
#define GFFTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFFTX_VERSION_MAJOR 0
#define GFFTX_VERSION_MINOR 1

enum gfftx_value_type { GFFTX_DOUBLE = 0, GFFTX_FLOAT = 1, GFFTX_COMPLEX_DOUBLE = 2, GFFTX_COMPLEX_FLOAT = 3 };
enum gfftx_transform_type { GFFTX_DFT = 0, GFFTX_IDFT = 1, GFFTX_RDFT = 2, GFFTX_IRDFT = 3 };
enum gfftx_place { GFFTX_IN_PLACE = 0, GFFTX_OUT_OF_PLACE = 1 };

typedef enum gfftx_status {
    GFFTX_SUCCESS = 0,
    GFFTX_ERR_INVALID_ARG = 1,   /* bad pointer, size or ID */
    GFFTX_ERR_UNSUPPORTED = 2,   /* combination not available (GFFT: "Requested transform is
                                    not compiled!", src/gfftgen.h:133-145) */
    GFFTX_ERR_ALLOC = 3,         /* device or host allocation failed */
    GFFTX_ERR_DEVICE = 4,        /* no usable GPU, or a HIP runtime error */
    GFFTX_ERR_ALIAS = 5,         /* out-of-place call with src == dst */
    GFFTX_ERR_LAUNCH = 6         /* kernel launch failed */
} gfftx_status;

/* Plan flags */
#define GFFTX_FLAG_R2C_BINS 0x1u     /* real transforms: n+1 bins instead of NR packing   */
#define GFFTX_FLAG_SPLIT_MFMA 0x2u   /* FP32 C2C: radix-16 stages as split-precision MFMA */
#define GFFTX_FLAG_NO_AUTOTUNE 0x4u  /* use the built-in schedule, ignore the tuning cache */
#define GFFTX_FLAG_GRAPH 0x8u        /* replay execution through a captured hipGraph      */
#define GFFTX_FLAG_PROFILE 0x10u     /* time every kernel launch with a HIP event pair
                                        (FFT tile and MFMA kernels: events attached to the
                                        dispatch; element-wise kernels: recorded around it) */

typedef struct gfftx_plan_s* gfftx_plan;
typedef struct gfftx_ragged_s* gfftx_ragged;

/* Create a plan for one transform of length n.
 * Replaces: GenerateTransform<...>::CreateTransformObject(n, vtype_id, trans_id, dim,
 * parall_id, place_id) (src/gfftgen.h:222-231) and Transform<...>::Create()
 * (src/gfftgen.h:81-83) for the HIP<device> parallelisation policy. */
int gfftx_plan_create(gfftx_plan* plan, uint64_t n, int vtype, int ttype, int place, int device,
                      uint32_t flags);

/* As gfftx_plan_create for `batch` transforms stored back to back (distance n complex;
 * n+1 for RDFT/IRDFT with GFFTX_FLAG_R2C_BINS on the complex side).  No reference
 * counterpart: GFFT's benchmark loops fft() over disjoint chunks of one buffer
 * (test/gfft_performance.h:110-116); a batched plan runs that loop as one launch. */
int gfftx_plan_create_batched(gfftx_plan* plan, uint64_t n, uint64_t batch, int vtype, int ttype,
                              int place, int device, uint32_t flags);

/* In-place transform.  Replaces: IN_PLACE::Function<...>::fft(T* data)
 * (src/gfftpolicy.h:163-172), i.e. AbstractFFT_inp<T>::fft (src/gfftpolicy.h:46-51). */
int gfftx_exec_inplace(gfftx_plan plan, void* data, void* stream);

/* Out-of-place transform; src is left unchanged, dst must not alias src.  Replaces:
 * OUT_OF_PLACE::Function<...>::fft(const T* src, T* dst) (src/gfftpolicy.h:203-212), i.e.
 * AbstractFFT_oop<T>::fft (src/gfftpolicy.h:53-58). */
int gfftx_exec_oop(gfftx_plan plan, const void* src, void* dst, void* stream);

/* Several in-place transforms in one call: plans[i] transforms data[i], i < count, in that
 * order on `stream` (device pointers: enqueued; host pointers: each returns when done, as
 * gfftx_exec_inplace).  The same plan may appear several times.  Every plan must be an
 * IN_PLACE plan and every pointer non-NULL; otherwise nothing is enqueued and
 * GFFTX_ERR_INVALID_ARG is returned.  A failure part-way returns that call's status, with
 * its index in gfftx_last_error(); the calls before it stay enqueued.  Replaces: a loop of
 * fft(T* data) calls over GFFT objects, e.g. the disjoint-chunk loop of
 * test/gfft_performance.h:110-116 or one object per length from one GenerateTransform
 * factory (src/gfftgen.h:222-231). */
int gfftx_exec_batch(const gfftx_plan* plans, void* const* data, size_t count, void* stream);

/* In-place transform that allocates nothing, for callers that cannot allocate during a call
 * (code under graph capture, tight serving loops, harnesses that forbid allocation).  No
 * counterpart in the reference.  As gfftx_exec_inplace on device pointers, except that the
 * intermediate arrays of a multi-pass plan come from a static device array of the library
 * (64 MiB per GPU, held until the process ends) instead of the plan's own scratch, and that
 * no event, graph or staging buffer is created.  Limits: device pointers only; lengths of
 * the factors 2, 3, 5 and 7 (chirp-z plans return GFFTX_ERR_UNSUPPORTED); a multi-pass
 * plan's batch may hold at most 32 MiB (2^22 FP32 or 2^21 FP64 complex points; larger
 * returns GFFTX_ERR_INVALID_ARG).  All such calls on one GPU share the static array: calls
 * that overlap on one GPU (other streams, other threads) race, and the caller orders them. */
int gfftx_exec_inplace_noalloc(gfftx_plan plan, void* data, void* stream);

/* Wait until the work enqueued on `stream` (a hipStream_t; NULL: the null stream) has
 * finished.  No reference counterpart: GFFT's fft() is synchronous
 * (src/gfftpolicy.h:163-172, :203-212), and the C++ policy layer calls this after each
 * device-pointer call to keep that guarantee. */
int gfftx_stream_synchronize(void* stream);

/* Release a plan.  Replaces: `delete` of the object CreateTransformObject returned
 * (virtual destructors, src/gfftpolicy.h:50,57). */
int gfftx_plan_destroy(gfftx_plan plan);

/* Write a JSON description of the plan (passes, kernels, radices, bytes) into buf.
 * No reference counterpart (GFFT's plan is a compile-time type). */
int gfftx_plan_describe(gfftx_plan plan, char* buf, size_t len);

/* Per-kernel timings of a plan created with GFFTX_FLAG_PROFILE: for each launch slot i
 * (i < max), total_ms[i] is the summed duration of that launch over all executions since
 * the last query and count[i] the number of executions; names[i] (optional, may be NULL)
 * receives the kernel name.  Waits for recorded work to finish and resets the counters.
 * Returns the number of launch slots via *nslots.  No reference counterpart (GFFT times
 * with clock(), test/gfft_performance.h:104-123). */
int gfftx_plan_profile(gfftx_plan plan, double* total_ms, uint64_t* count, const char** names, int max,
                       int* nslots);

/* Ragged batch: `count` transforms of individual lengths stored back to back in one buffer
 * (transform i starts at complex offset sum_{j<i} lengths[j]).  Any length n >= 2 within
 * the plan limits, FP32 or FP64.  In FP32, lengths 2^k, 3*2^k, 5*2^k from 16 to 4096 run
 * in one persistent launch; every other length class (all of them in FP64) runs as batched
 * launches on side streams that overlap it (a class with a prime factor above 5 as a
 * chirp-z plan over its transforms); one call (or one hipGraph with GFFTX_FLAG_GRAPH)
 * covers all.  With autotuning on, plan creation also times the batch with every length as
 * a class launch and keeps that form where it is faster (the choice is cached per set of
 * length classes; GFFTX_RAGGED_PERSISTENT_MAX fixes it).
 * No reference counterpart: the nearest is one GFFT object per length from one
 * GenerateTransform<NList,...> factory (src/gfftgen.h:189-233) called in a loop. */
int gfftx_ragged_create(gfftx_ragged* plan, const uint64_t* lengths, uint64_t count, int vtype,
                        int ttype, int device, uint32_t flags);
int gfftx_ragged_exec(gfftx_ragged plan, void* data, void* stream);
int gfftx_ragged_destroy(gfftx_ragged plan);
/* JSON description of a ragged plan (classes, items, grid, the persistent-kernel limit
 * and how it was chosen).  No reference counterpart. */
int gfftx_ragged_describe(gfftx_ragged plan, char* buf, size_t len);
/* Summed duration and count of ragged launches since the last query (GFFTX_FLAG_PROFILE).
 * No reference counterpart. */
int gfftx_ragged_profile(gfftx_ragged plan, double* total_ms, uint64_t* count);

/* Digit-reversal permutation dst[rev(k)] = src[k] over radix^digits complex elements
 * (device pointers).  Replaces GFFTswap2<radix,digits,T>::apply (src/gfftswap.h:74-140) when
 * an explicitly reordered array is wanted; the transforms themselves need no reordering
 * pass (their indexing is self-sorting). */
int gfftx_digit_reverse(const void* src, void* dst, uint64_t radix, uint32_t digits, int vtype,
                        void* stream);

/* Human-readable text for a status code. */
const char* gfftx_strerror(int status);
/* Detail of the calling thread's most recent failure ("" if none). */
const char* gfftx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GFFTX_H */

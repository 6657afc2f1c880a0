/*
 * mpx -- MI355X-native NNEDI3 / RAVU prescalers: the C ABI.
 *
 * The reference (bjin/mpv-prescalers) is a set of mpv/libplacebo user shaders; its
 * "interface" is the `//!` directive block at the head of every pass, which mpv parses
 * and dispatches.  Each entry point below names the directive lines it replaces.
 * All pointers are device pointers unless stated otherwise; sizes are in pixels,
 * pitches and strides in ELEMENTS (not bytes).  No entry point aborts: every failure is
 * a negative status with a thread-local message from mpx_last_error().
 */
#ifndef MPX_H
#define MPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define MPX_OK 0
#define MPX_SKIPPED 1             /* WHEN is false for these sizes: output = input copy */
#define MPX_ERR_ARG (-1)          /* bad argument (null pointer, bad enum, bad stride)  */
#define MPX_ERR_UNKNOWN_VARIANT (-2)
#define MPX_ERR_UNAVAILABLE (-3)  /* variant known but its weights are not shipped     */
#define MPX_ERR_SHAPE (-4)        /* sizes out of range or batch > plan's max_batch    */
#define MPX_ERR_BUDGET (-5)       /* allocation would exceed the context's HBM budget  */
#define MPX_ERR_HIP (-6)          /* HIP runtime error (message has the HIP error)     */
#define MPX_ERR_IO (-7)           /* weight file missing or corrupt                    */
#define MPX_ERR_UNSUPPORTED (-8)  /* combination not supported (e.g. precision)        */

/* ---- element types of planes ------------------------------------------------------ */
#define MPX_U8 0    /* 8-bit unorm, value / 255 (HOOKED_mul of an 8-bit texture)          */
#define MPX_U16 1   /* 16-bit container, value / (2^bits - 1), bits given at plan time   */
#define MPX_F16 2
#define MPX_F32 3

/* ---- NNEDI3 contraction precision (RAVU always computes in FP32) ------------------ */
#define MPX_PREC_FP32 0   /* plain FP32 VALU kernel                                       */
#define MPX_PREC_FP16 1   /* FP16 MFMA, one product, FP32 accumulation                    */
#define MPX_PREC_BF16 2   /* BF16 MFMA, high/low split (three products), FP32 accumulation */
#define MPX_PREC_FP8 3    /* FP8 e4m3 MFMA, per-neuron power-of-two scales, FP32 accum.   */
#define MPX_PREC_FP16X3 4 /* FP16 MFMA, high/low split (three products), FP32 accumulation */
/* OR'ed into `precision` of mpx_plan_create: store the intermediates (NNEDI3's double_y
 * plane, RAVU's int11 plane) in FP16, round to nearest even, as mpv stores a hook's SAVEd
 * textures and pass outputs in its default rgba16f FBOs (e.g. nnedi3-nns256-win8x6.hook:18,
 * ravu-r3.hook:19).  With MPX_F16 output the plan then reproduces mpv's default storage;
 * without the flag the intermediates stay FP32 (closer to the FP64 shader math). */
#define MPX_STORE_FP16 0x100

/* ---- variant families / channel forms --------------------------------------------- */
#define MPX_FAMILY_NNEDI3 0
#define MPX_FAMILY_RAVU 1
#define MPX_FAMILY_RAVU_LITE 2
#define MPX_FAMILY_RAVU_3X 3
#define MPX_FAMILY_RAVU_ZOOM 4

#define MPX_CH_LUMA 0   /* //!HOOK LUMA, one plane                                        */
#define MPX_CH_YUV 1    /* //!HOOK NATIVE, 3 planes, hash from plane 0                    */
#define MPX_CH_RGB 2    /* //!HOOK MAIN, 3 planes, hash from BT.709 luma                  */

typedef struct mpx_variant_info {
  int32_t family;       /* MPX_FAMILY_*                                                   */
  int32_t channels;     /* MPX_CH_*                                                       */
  int32_t planes;       /* 1 or 3                                                         */
  int32_t radius;       /* RAVU radius r (0 for NNEDI3)                                   */
  int32_t nns;          /* NNEDI3 neurons (0 for RAVU)                                    */
  int32_t win_w, win_h; /* NNEDI3 window (8 x 4 or 8 x 6)                                 */
  int32_t ar;           /* anti-ringing variant                                           */
  int32_t available;    /* weights shipped (0 for ravu-zoom-ar-r3*)                       */
  int32_t applies;      /* WHEN evaluated for the given in/target sizes                   */
  int32_t out_w, out_h; /* size of the hook's output (in size when it does not apply)     */
  int32_t scale_x, scale_y; /* integer scale per axis; 0 for arbitrary-ratio (zoom)       */
  int32_t align;        /* 1 for OFFSET ALIGN (zoom)                                      */
  float offset_x, offset_y; /* accumulated OFFSET of the output, in output pixels         */
  char hook[8];         /* "LUMA" | "NATIVE" | "MAIN"                                     */
} mpx_variant_info;

typedef struct mpx_ctx mpx_ctx;
typedef struct mpx_plan mpx_plan;

/* Full I/O description for mpx_run_ex.  Strides in elements; planes of one frame are
 * cstride apart, frames bstride apart. */
typedef struct mpx_io {
  const void* in;
  int64_t in_pitch, in_cstride, in_bstride;
  void* out;
  int64_t out_pitch, out_cstride, out_bstride;
  int32_t batch;
  int32_t* buckets;     /* optional (may be NULL): per-site LUT row ids, int32, see
                           mpx_plan_bucket_count; for parity reporting               */
} mpx_io;

/* ---- host-only queries (no GPU needed) -------------------------------------------- */

/* Number of shipped variant names and the i-th name (file names without ".hook"). */
int mpx_variant_count(void);
const char* mpx_variant_name(int i);

/* Geometry / gating of `name` for an input of in_w x in_h displayed at target_w x
 * target_h.  Replaces mpv's evaluation of //!WHEN, //!WIDTH, //!HEIGHT, //!OFFSET of the
 * variant's passes (e.g. ravu-lite-r3.hook:20,123-125; nnedi3-nns256-win8x6.hook:19,
 * 349-352; ravu-zoom-r2.hook:18-21; compute/ravu-3x-r3.hook:18-20). */
int mpx_query(const char* name, int in_w, int in_h, int target_w, int target_h,
              mpx_variant_info* info);

/* Check every weight file variant `name` loads from `weights_dir` (NULL = the default
 * directory) without a device: MPXW header, shape against the shape the kernels index
 * with, file size, and for the tables the reference ships the SHA-256 of the data
 * against the value recorded when they were extracted.  MPX_ERR_IO names the first bad
 * file; MPX_ERR_UNAVAILABLE when an unshipped table (ravu-zoom-ar-r3) is not supplied.
 * mpx_plan_create makes the same checks before it uploads anything.  Replaces mpv's
 * parsing of the `//!TEXTURE` block's SIZE / FORMAT lines and its hex data (e.g.
 * ravu-zoom-ar-r2-rgb.hook:226-234: SIZE 18 2592, FORMAT rgba16f). */
int mpx_weights_verify(const char* weights_dir, const char* name);

/* ---- context ---------------------------------------------------------------------- */

/* Bind a HIP device, set the HBM budget (bytes; 0 = no limit) for weights + plan
 * scratch, and the directory holding the shipped *.bin weight files (NULL = the
 * directory compiled in / MPX_WEIGHTS_DIR).  Replaces mpv's loading of the `.hook`
 * file and its //!TEXTURE blocks (e.g. ravu-lite-r3.hook:132-136). */
int mpx_create(int device, uint64_t hbm_budget, const char* weights_dir, mpx_ctx** out);
void mpx_destroy(mpx_ctx* ctx);
/* Bytes of device memory the context holds (weights + all live plans' scratch). */
uint64_t mpx_ctx_bytes(const mpx_ctx* ctx);

/* ---- plans ------------------------------------------------------------------------ */

/* Prepare `name` for fixed sizes and element types: uploads the weights once (kept
 * resident in the context), allocates scratch for max_batch frames, picks the kernels.
 * `bits` is the bit depth of MPX_U16 input (ignored otherwise).  Replaces the pass list
 * of the hook file (//!HOOK, //!BIND, //!SAVE, //!COMPONENTS, //!COMPUTE lines, e.g.
 * ravu-r3.hook:15-20,148-154,282-288,416-425).  Returns MPX_SKIPPED (and a valid plan
 * that copies) when //!WHEN is false for these sizes. */
int mpx_plan_create(mpx_ctx* ctx, const char* name, int in_w, int in_h, int target_w,
                    int target_h, int in_type, int bits, int out_type, int precision,
                    int max_batch, mpx_plan** out);
int mpx_plan_info(const mpx_plan* plan, mpx_variant_info* info);
/* Number of int32 bucket ids one frame writes when mpx_io.buckets is set (0 for NNEDI3). */
int64_t mpx_plan_bucket_count(const mpx_plan* plan);
void mpx_plan_destroy(mpx_plan* plan);

/* Run the plan on `batch` frames, asynchronously on `hip_stream` (a hipStream_t; NULL =
 * default stream).  Planes of a frame are contiguous (cstride = pitch * height).
 * Allocates nothing.  Replaces one frame's dispatch of all the hook's passes (e.g.
 * ravu-lite-r3.hook:15-131 step1 + step2). */
int mpx_run(mpx_plan* plan, const void* in, int64_t in_pitch, int64_t in_bstride, void* out,
            int64_t out_pitch, int64_t out_bstride, int batch, void* hip_stream);
int mpx_run_ex(mpx_plan* plan, const mpx_io* io, void* hip_stream);

/* Run the plan on n (1..MPX_MAX_FRAMES, <= max_batch) frames of the plan's shape that live
 * at independent addresses (ins[i] / outs[i]: host arrays of device pointers; planes of
 * a frame contiguous).  One launch per pass for all of them, no copies and no
 * allocation -- what a serving loop uses to batch frames of separate requests.  Replaces
 * mpv dispatching the hook's passes once per frame. */
#define MPX_MAX_FRAMES 16
int mpx_run_frames(mpx_plan* plan, int n, const void* const* ins, void* const* outs, int64_t in_pitch,
                   int64_t out_pitch, void* hip_stream);

/* Autotune the plan's launch shape on the caller's buffers: runs the plan on `io`
 * (warm-up + `iters` timed runs per candidate, HIP events on hip_stream, synchronising)
 * over the candidate shapes and keeps the fastest in the plan.  NNEDI3 MFMA plans: the
 * persistent grid (blocks per CU) and, for nns = 64 and 128, the neuron split across waves.
 * RAVU plans (lite, 3x, full, zoom): the tile shape of their kernels (read the choice back
 * with mpx_plan_tile); *grid_per_cu and *split are 0.  NNEDI3 FP32 plans have no tunable
 * shape: MPX_OK, nothing run.  Writes the output buffers.  Outputs (may be NULL): the
 * chosen blocks per CU (0 = occupancy maximum) and split, and the best time per run in
 * microseconds.  Replaces the fixed workgroup shape of the compute shaders (e.g.
 * compute/ravu-3x-r3.hook:22 `//!COMPUTE 96 24 32 8`, compute/ravu-zoom-r3.hook:23). */
int mpx_plan_tune(mpx_plan* plan, const mpx_io* io, void* hip_stream, int iters, int* grid_per_cu, int* split,
                  float* best_us);

/* Set the plan's launch shape directly, e.g. one found by mpx_plan_tune in an earlier
 * process (the tuner's result can be cached per GPU model and frame size).  grid_per_cu:
 * 0 (occupancy maximum) or 1..64 blocks per CU (capped at the occupancy maximum); split:
 * 0, or 1 for NNEDI3 MFMA plans with nns = 64 or 128.  Plans without a tunable shape
 * accept only (0, 0).  Enqueues nothing.  No counterpart in the reference (its compute
 * shaders fix the workgroup shape, e.g. compute/ravu-3x-r3.hook:22). */
int mpx_plan_set_shape(mpx_plan* plan, int grid_per_cu, int split);

/* Tile shape of a RAVU plan's kernels: candidate `tile` of the family's list (0 = the
 * default; mpx_plan_tune picks the fastest).  One block of the plan's kernels covers a
 * tile_w x tile_h block of input pixels (ravu-lite, ravu-3x, full RAVU) or output pixels
 * (ravu-zoom), one thread per pixel.  mpx_plan_tile outputs (may be NULL): the current
 * index, the number of candidates (1 for NNEDI3 plans, with width = height = 0) and the
 * current tile's size.  mpx_plan_set_tile enqueues nothing and changes no result, only
 * speed; call it while no other thread runs the plan.  No counterpart in the reference
 * (the compute shaders fix their workgroup, e.g. compute/ravu-lite-r3.hook:22
 * `//!COMPUTE 64 16 32 8`, compute/ravu-zoom-r3.hook:23 `//!COMPUTE 32 8`). */
int mpx_plan_set_tile(mpx_plan* plan, int tile);
int mpx_plan_tile(const mpx_plan* plan, int* tile, int* count, int* tile_w, int* tile_h);

/* ---- the scaling pass that consumes a prescaler's OFFSET ---------------------------- */
#define MPX_FILTER_BILINEAR 0
#define MPX_FILTER_CATMULL_ROM 1
#define MPX_FILTER_MITCHELL 2
#define MPX_FILTER_SPLINE36 3
#define MPX_FILTER_LANCZOS3 4

/* Resample `planes` planes per frame of in_w x in_h (io->batch frames, strides as in
 * mpx_io) to out_w x out_h with a separable filter, sampling the input at
 * (j + 1/2) * in/out + offset per axis: pass the prescaler's mpx_variant_info.offset_x/y
 * to get a display-aligned image.  Replaces mpv's main scaler pass, which applies the
 * accumulated `//!OFFSET` of the hooks before it (e.g. nnedi3-nns256-win8x6.hook:352,
 * ravu-r3.hook:426-428).  Types MPX_U8/U16 (value/(2^bits-1)), F16, F32 in; F16/F32 out.
 * Allocates nothing; asynchronous on hip_stream. */
int mpx_resample(const mpx_io* io, int in_w, int in_h, int in_type, int bits, int out_w, int out_h,
                 int out_type, int planes, int filter, float offset_x, float offset_y, void* hip_stream);

/* Run the plan on n (1..MPX_MAX_FRAMES, <= max_batch) frames of DIFFERENT sizes in one
 * launch per pass: frame i is widths[i] x heights[i] (each at most the plan's input size),
 * rows in_pitches[i] / out_pitches[i] elements apart, planes contiguous; every frame is
 * scaled by the plan's factors (the plan must apply: NNEDI3, ravu, ravu-lite, ravu-3x;
 * not ravu-zoom, whose output size is the plan's target).  The kernels walk one tile
 * queue across all frames.  No allocation.  Replaces mpv running the hook's passes on
 * each frame at its own size. */
int mpx_run_frames_sized(mpx_plan* plan, int n, const void* const* ins, void* const* outs, const int* widths,
                         const int* heights, const int64_t* in_pitches, const int64_t* out_pitches,
                         void* hip_stream);

/* mpx_run_ex with the plan's intermediates (the NNEDI3 double_y plane when both axes
 * double; the RAVU int11 plane) in a device array that the library holds in its code
 * object, MPX_STATIC_SCRATCH_BYTES per device, instead of the plan's scratch.  It calls
 * no allocating routine of HIP or of any library.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  MPX_ERR_SHAPE, before anything is
 * enqueued, when the batch's intermediates exceed MPX_STATIC_SCRATCH_BYTES (e.g. 4
 * frames of 1920x1080 NNEDI3 2x need 4 x 16.6 MB; one 2560x1440 ravu-rN-yuv frame needs
 * 44.2 MB).  Every call on one device shares that device's array: calls whose work
 * overlaps on one device (two streams, or two host threads) race on it and give wrong
 * results; serialise them.  Nothing in the array needs to start at zero (each pass writes
 * every element it is read back from).  Plans without intermediates run as mpx_run_ex. */
#define MPX_STATIC_SCRATCH_BYTES (64u << 20)
int mpx_run_static(mpx_plan* plan, const mpx_io* io, void* hip_stream);

/* ---- request queue: continuous batching for hosts that serve many frames ----------- */
#define MPX_PENDING 2             /* mpx_poll: the request has not finished yet          */
#define MPX_ERR_DEADLINE (-9)     /* the request's deadline passed before it was launched */

/* One frame to upscale.  `in` / `out` are device pointers owned by the caller and must
 * stay valid until mpx_poll reports the request finished; planes are contiguous
 * (in: pitch in_w, out: pitch out_w of mpx_query's answer for these sizes).  ready_stream:
 * the hipStream_t that produces `in` (NULL: `in` is ready now); the launch waits for the
 * work submitted to it before mpx_submit, without the host waiting.  deadline_us: a
 * request still queued that long after mpx_submit fails with MPX_ERR_DEADLINE instead of
 * running (0: no deadline). */
typedef struct mpx_request {
  const char* name;
  const void* in;
  int32_t in_w, in_h, in_type, bits;
  void* out;
  int32_t out_type;
  int32_t target_w, target_h; /* display size; 0, 0 = the variant's 2x (3x for ravu-3x) */
  int32_t precision;          /* MPX_PREC_* (| MPX_STORE_FP16)                          */
  void* ready_stream;
  int64_t deadline_us;
} mpx_request;

typedef struct mpx_queue mpx_queue;

/* A request queue over `ctx` (which it shares: plans and weights stay resident under the
 * context's HBM budget; the queue allocates no frame memory, outputs are the caller's).
 * A scheduler thread admits requests in arrival order and groups them by plan: frames at
 * their variant's default scale whose sizes fit max_w x max_h share launches whatever their
 * sizes (mpx_run_frames_sized), the others group by size (mpx_run_frames); up to
 * max_batch (<= MPX_MAX_FRAMES) frames per launch, launches spread over `streams` HIP
 * streams of its own, at most 2 launches in flight per stream so that arrivals keep
 * joining the next launch.  A device error fails every unfinished request and the queue
 * (submits then return MPX_ERR_HIP).  Replaces mpv dispatching the hook's passes once per
 * frame (e.g. ravu-lite-r3.hook:15-21, 119-126) with batched launches across requests. */
int mpx_queue_create(mpx_ctx* ctx, int streams, int max_batch, int max_w, int max_h, mpx_queue** out);
/* Finish every submitted request (run what is queued), then free the queue. */
void mpx_queue_destroy(mpx_queue* q);
/* Queue one request; *ticket identifies it.  Checks the variant, sizes, types and
 * pointers first (nothing is queued on an error).  Never waits for the GPU. */
int mpx_submit(mpx_queue* q, const mpx_request* r, uint64_t* ticket);
/* MPX_OK: the request finished and its output is complete (visible to any stream);
 * MPX_PENDING: not yet after waiting up to timeout_us (0: do not wait, < 0: wait until
 * it finishes); a negative status: the request failed (message in mpx_last_error).
 * latency_us (may be NULL): on MPX_OK or an error, the microseconds from mpx_submit to the
 * moment the queue saw the request finish.  A ticket is retired once it returns MPX_OK or
 * an error. */
int mpx_poll(mpx_queue* q, uint64_t ticket, int64_t timeout_us, int64_t* latency_us);
/* Launches issued so far (how many requests shared a launch is submitted / launches). */
uint64_t mpx_queue_launches(const mpx_queue* q);
/* hold != 0: the scheduler takes nothing from the queue (submits still queue); hold == 0:
 * it takes everything queued again, grouped as usual -- to submit a set of frames that
 * should share launches.  mpx_queue_destroy releases a hold. */
int mpx_queue_hold(mpx_queue* q, int hold);

/* Message of the last failing call on this thread ("" if none). */
const char* mpx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MPX_H */

/*
 * dfttest2_mi355x -- C ABI of the MI355X DFTTest block pipeline.
 *
 * A plan holds one set of DFTTest plugin parameters (the reference plugin's DFTTest
 * arguments) on one GPU; process calls filter whole planes with the fused gfx950
 * kernel.  All entry points are plain C: opaque handle, plain pointers and sizes,
 * int status (0 = success, negative = error; dft2_last_error() names the error).
 *
 * Reference interfaces replaced (paths relative to the vs-dfttest2 reference tree):
 *   dft2_plan_create    <- DFTTestCreate, hiprtc_source/source.cpp:710-934 (argument
 *                          string hiprtc_source/source.cpp:1025-1046, defaults and
 *                          checks :719-847)
 *   dft2_plan_destroy   <- DFTTestFree, hiprtc_source/source.cpp:694-707
 *   dft2_process_frame  <- DFTTestGetFrame, hiprtc_source/source.cpp:445-692 (one
 *                          output frame of one plane from its T source frames)
 *   dft2_process_frame_keyed
 *                       <- the same on host planes, with the source planes' device
 *                          copies kept between calls (the reference uploads T planes
 *                          per call, hiprtc_source/source.cpp:570-573)
 *   dft2_process_batch  <- the same per-frame work over whole clips resident in device
 *                          memory, many clips/planes per launch (no reference
 *                          counterpart; the reference launches per frame and plane)
 *   dft2_process_batch_sigma, dft2_backward_batch
 *                       <- no reference counterpart: the differentiable form of the
 *                          filter (sigma as a device array) and its backward pass
 *   dft2_batch_*        <- no reference counterpart: a fixed set of clips of uneven
 *                          size set up once (job table, scratch, autotuned tile shape)
 *                          and run as one captured hipGraph launch; replaces the
 *                          reference's per-frame launch pair (hiprtc_source/source.cpp:575-640)
 *   dft2_rdft           <- RDFT, cpu_source/source.cpp:269-340 (registered :390-398)
 *   dft2_version        <- Version, cpu_source/source.cpp:342-354
 */
#ifndef DFTTEST2_MI355X_H
#define DFTTEST2_MI355X_H

#include <stddef.h>

#if defined(__GNUC__)
#define DFT2_API __attribute__((visibility("default")))
#else
#define DFT2_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define DFT2_OK 0
#define DFT2_ERR_ARG (-1)     /* invalid argument (message in dft2_last_error) */
#define DFT2_ERR_HIP (-2)     /* HIP runtime error */
#define DFT2_ERR_NOMEM (-3)   /* device or host allocation failed */
#define DFT2_ERR_STATE (-4)   /* plan in an unusable state */

/* Sample formats (VapourSynth sampleType / bitsPerSample). */
#define DFT2_SAMPLE_INT 0
#define DFT2_SAMPLE_FLOAT 1

/* Precision of the DFT contractions. */
#define DFT2_PRECISION_FP32 0      /* FP32 accuracy: FP32 MFMA; with 8-bit samples the x
                                    * transform runs on the BF16 MFMA over an exact
                                    * three-part split of its basis (exact products) */
#define DFT2_PRECISION_SPLIT_BF16 1 /* hi+lo BF16 operands, three products, FP32 accumulate.
                                    * Coefficient error about 2^-15 of the block's L1 norm
                                    * (FP32: 2^-18).  Filter types 0, 1, 2, 5 and 6 only:
                                    * plans with types 3 and 4, whose gains exceed 1 and (type
                                    * 3) jump at pmin/pmax, are rejected.  Type 2 scales the
                                    * error by sigma.  Type 1 decides psd < sigma within that
                                    * error: near-threshold coefficients may be kept or
                                    * dropped otherwise than in FP64 (as in FP32, more often). */

/* Stream argument naming the null (default) stream: the work is queued there and not
 * waited for (NULL instead selects the plan's own stream and a blocking call).  Same
 * value as HIP's hipStreamLegacy; the library passes stream 0 to HIP for it. */
#define DFT2_DEFAULT_STREAM ((void*)1)

/* Plugin parameters, with the meaning and layout of the reference DFTTest arguments.
 * block_size 16 with radius 0..3 (the reference's fused backends) runs on the fused
 * kernel; other block sizes (1..64) and radii up to 7 (the reference's FFT-library
 * backends) on the generic block path, FP32 only, as far as one block's two complex
 * T x B x (B/2+1) buffers fit a CU's LDS.  B below is block_size. */
typedef struct dft2_params {
    int radius;           /* "radius": 0..7, T = 2*radius+1 */
    int block_size;       /* "block_size": 1..64 */
    int block_step;       /* "block_step": 1..block_size (sbsize - sosize) */
    int filter_type;      /* "filter_type": 0..6 */
    int zero_mean;        /* "zero_mean": 0 or 1 */
    const double* window; /* "window": T*B*B values, [t][y][x] */
    int window_len;
    const double* sigma;  /* "sigma": 1 value (scalar) or T*B*(B/2+1) values, [t][ky][kx] */
    int sigma_len;
    double sigma2;        /* "sigma2" */
    double pmin;          /* "pmin" */
    double pmax;          /* "pmax" */
    const double* window_freq; /* "window_freq": T*B*(B/2+1)*2 values [t][ky][kx][re,im]; needed if zero_mean */
    int window_freq_len;
} dft2_params;

/* One plane of one clip resident in device memory.  Strides and pitches are in bytes.
 * Output frames [frame_begin, frame_end) are written; source frames are taken from
 * [0, n_frames) with the temporal clamp of the reference. */
typedef struct dft2_job {
    const void* src;
    void* dst;
    long long src_frame_pitch;
    long long src_stride;
    long long dst_frame_pitch;
    long long dst_stride;
    int width;
    int height;
    int n_frames;
    int frame_begin;
    int frame_end;
} dft2_job;

/* One plane of one clip for the backward pass (float32 planes in device memory).
 * grad_in receives d loss / d src and is accumulated with +=, so the caller zeroes it. */
typedef struct dft2_grad_job {
    const void* src;       /* forward input planes */
    const void* grad_out;  /* d loss / d output planes */
    void* grad_in;         /* d loss / d input planes (accumulated) */
    long long src_frame_pitch;
    long long src_stride;
    long long go_frame_pitch;
    long long go_stride;
    long long gi_frame_pitch;
    long long gi_stride;
    int width;
    int height;
    int n_frames;
    int frame_begin;       /* output frames [frame_begin, frame_end) whose gradient is given */
    int frame_end;
} dft2_grad_job;

typedef struct dft2_plan dft2_plan;

/* Creates a plan on `device_id`.  sample_type / bits_per_sample give the clip format
 * (integer 8..16 bit, or float 32).  num_streams >= 1 concurrent calls are admitted,
 * each with its own stream and scratch (the reference's num_streams). */
DFT2_API int dft2_plan_create(const dft2_params* params, int sample_type, int bits_per_sample, int device_id,
                     int num_streams, int precision, dft2_plan** out_plan);

/* Releases the plan and its device memory. NULL is accepted. */
DFT2_API int dft2_plan_destroy(dft2_plan* plan);

/* Streams: every `void* stream` below is a hipStream_t.  NULL selects the plan's (or
 * batch's) own stream and makes the call wait for completion.  Any other handle queues
 * the work on that stream and returns at once; DFT2_DEFAULT_STREAM queues it on the
 * null (default) stream, which NULL cannot name here. */

/* Filters one output frame of one plane.  src_frames holds the T = 2*radius+1 source
 * planes in temporal order (already clamped by the caller, as GetFrame gathers them).
 * memory = 0: device pointers, work is queued on `stream` (hipStream_t; NULL = the
 * plan's own stream, and the call waits for completion; that stream does not wait for
 * other streams, so the planes must be complete when the call is made).  Calls from
 * several threads on one plan are safe; at most num_streams run at once, the others
 * wait for a free stream slot.  memory = 1: host pointers;
 * the call stages through device memory and returns when dst is written.  Pageable host
 * planes pass through a pinned buffer of the call's stream slot; pinned ones are copied
 * directly. */
DFT2_API int dft2_process_frame(dft2_plan* plan, const void* const* src_frames, long long src_stride, void* dst,
                       long long dst_stride, int width, int height, int memory, void* stream);

/* dft2_process_frame on host planes (memory = 1) that keeps device copies of the source
 * planes between calls.  src_keys[t] names the content of src_frames[t]: the caller
 * guarantees that a key always names the same width x height samples for the life of the
 * plan (or until dft2_plan_forget_frames), as (frame number, plane) does for a
 * VapourSynth clip, whose frames are immutable.  A plane whose key an earlier call
 * uploaded is not copied again, so consecutive output frames upload one new source plane
 * each instead of T.  The copies are shared by the plan's stream slots and calls from
 * several threads; the plan keeps 3 * (T + 1) * num_streams of them, enough for three
 * planes of consecutive frames per stream slot (least recently used ones are replaced).
 * Returns when dst is written.  Replaces the per-call uploads of the
 * reference's GetFrame (hiprtc_source/source.cpp:570-573), which have no cache. */
DFT2_API int dft2_process_frame_keyed(dft2_plan* plan, const void* const* src_frames,
                                      const unsigned long long* src_keys, long long src_stride, void* dst,
                                      long long dst_stride, int width, int height);

/* Forgets every source plane dft2_process_frame_keyed holds: later calls upload them
 * again whatever their keys (their device memory is kept for reuse).  No reference
 * counterpart. */
DFT2_API int dft2_plan_forget_frames(dft2_plan* plan);

/* Filters every job (device memory) with one fused launch per frame chunk, queued on
 * `stream` (NULL = the plan's own stream, and the call waits for completion).  A chunk
 * holds as many frames as the stream slot's tile scratch allows (environment variable
 * DFT2_SCRATCH_MB at plan creation, default 4096), in whole groups of the cross-frame
 * form's frames per work item.  Plans with a separable window take that form (8 output
 * frames per work item, each source slice transformed once per item) at tbsize 5, and at
 * tbsize 7 for float samples; DFT2_REUSE=K at plan creation overrides K (1 = one frame
 * per item).  Calls whose jobs have one output frame each take the one-frame form. */
DFT2_API int dft2_process_batch(dft2_plan* plan, const dft2_job* jobs, int n_jobs, void* stream);

/* As dft2_process_batch, for callers that cannot allocate (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc): the call allocates nothing.  No
 * reference counterpart.  Its working memory (job table and tile regions) is one 64 MiB
 * array of device memory per GPU, held from library load until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU (different streams or
 * threads) race, so the caller serialises them.  At most 64 jobs per call; frames run in
 * chunks that fit the array, and jobs whose single frame needs more tile scratch than it
 * holds (about 3 frames of 1080p YUV420 fit) are rejected.  Kernel timing
 * (dft2_plan_set_timing) does not record these launches.  NULL stream: the plan's first
 * stream, and the call waits for completion. */
DFT2_API int dft2_process_batch_noalloc(dft2_plan* plan, const dft2_job* jobs, int n_jobs, void* stream);

/* As dft2_process_batch, with sigma read from device memory: sigma_dev holds T*16*9
 * floats [t][ky][kx] and the filter uses sigma_dev * sigma_scale (sigma_scale = the
 * window's sum of squares for ftype 0/1 in user units, as python/dfttest2.py:329-333).
 * Gains at kx = 0, 8 average each coefficient's own and conjugate partner's sigma.
 * No reference counterpart: the differentiable form of the filter. */
DFT2_API int dft2_process_batch_sigma(dft2_plan* plan, const dft2_job* jobs, int n_jobs, const float* sigma_dev,
                             float sigma_scale, void* stream);

/* Parity instrumentation for the hard-threshold filters: as dft2_process_batch for a plan
 * with filter_type 1 (either precision) or 3 (FP32) and a symmetric sigma, and also writes every block's gain decisions to decisions_dev (device memory,
 * decisions_len 64-bit entries).  Layout: for each job in order, each output frame
 * [frame_begin, frame_end), each block row by < pad_num(height) and column bx <
 * pad_num(width) (cpu_source/getframe_impl.cpp:18-20), (radius+1)*4 masks [g][slot];
 * bit `lane` (0..63, lane = 8p + q, slot = 2e + k) is the decision on coefficient
 * (kt, ky, kx): for q != 0, ky = (p or 0 when p = 0) if e = 0, else (16 - p or 8 when
 * p = 0), and (kt, ky, kx) = (g, ky, q) if k = 0, ((T - g) % T, (16 - ky) % 16, q) if k = 1;
 * for q = 0, (g, ky, 0) if k = 0 and (g, ky, 8) if k = 1.  The decision of type 1 is
 * psd < sigma (dropped), of type 3 pmin <= psd <= pmax (sigma applies).  Plans on the
 * generic block path (block_size != 16 or radius > 3) decide every coefficient of the
 * T x block_size x (block_size/2+1) spectrum: per block, ceil(that / 64) masks, bit i of
 * the sequence = coefficient i in [kt][ky][kx] order (any sigma).  No reference
 * counterpart. */
DFT2_API int dft2_process_batch_decisions(dft2_plan* plan, const dft2_job* jobs, int n_jobs,
                                          unsigned long long* decisions_dev, long long decisions_len, void* stream);

/* Backward pass of dft2_process_batch_sigma for the Wiener filter (filter_type 0) on
 * float32 planes: accumulates d loss / d src into each job's grad_in and
 * d loss / d sigma_dev into sigma_grad_dev (T*16*9 floats, device, +=), given grad_out.
 * Block spectra are recomputed, not stored.  grad_in is accumulated with atomic adds, so
 * the order of its sums (not its value beyond rounding) varies from call to call.  With a
 * separable window and radius >= 1 a work item covers 8 consecutive output frames (the
 * environment variables DFT2_BWD_REUSE=K, K = 1 for one frame per item, and
 * DFT2_BWD_TILE_H=H override the frames per item and the tile height in blocks).
 * No reference counterpart. */
DFT2_API int dft2_backward_batch(dft2_plan* plan, const dft2_grad_job* jobs, int n_jobs, const float* sigma_dev,
                        float sigma_scale, float* sigma_grad_dev, void* stream);

/* ---- Kernel timing ----------------------------------------------------------------- */

typedef struct dft2_kernel_timing {
    long long fused_launches;    /* fused kernel launches recorded */
    double fused_ms;             /* their summed duration */
    long long finalize_launches; /* finalize kernel launches recorded */
    double finalize_ms;          /* their summed duration */
    long long blocks;            /* output blocks (block positions x output frames) the
                                    recorded block launches computed */
    long long slice_spectra;     /* spatial slice spectra they computed, counted per block
                                    position: T per output block in the one-frame form,
                                    frames + T - 1 per group of consecutive frames in the
                                    cross-frame form */
} dft2_kernel_timing;

/* enable = 1: from now on the plan's process calls record HIP events around every fused
 * and finalize launch, on the stream the launch uses (three events per launch pair);
 * the totals start from zero.  At most 64 launch pairs stay pending: a process call
 * that would exceed that first waits for the oldest pair and adds it to the totals, so
 * a long run holds a fixed number of events (and, with more than 64 launch pairs in
 * flight, waits for the GPU).  enable = 0 stops recording.  No reference counterpart:
 * instrumentation for throughput and roofline measurement. */
DFT2_API int dft2_plan_set_timing(dft2_plan* plan, int enable);

/* Waits for the recorded events and returns the totals since timing was enabled. */
DFT2_API int dft2_plan_get_timing(dft2_plan* plan, dft2_kernel_timing* out);

/* ---- Batches: many clips, one graph launch ---------------------------------------- */

typedef struct dft2_batch dft2_batch;

typedef struct dft2_batch_options {
    int autotune;      /* 1: time every valid tile shape and launch form on a sample of the
                          batch and keep the fastest (tile_w/tile_h/persistent ignored) */
    int tile_w;        /* blocks per workgroup tile along x (0 = the plan's default) */
    int tile_h;        /* blocks per workgroup tile along y (0 = the plan's default) */
    int persistent;    /* 1: persistent grid of resident workgroups over a work queue */
    int sample_frames; /* autotune: output frames per job in the timing sample (0 = 2) */
} dft2_batch_options;

typedef struct dft2_batch_info {
    int tile_w, tile_h;      /* chosen tile (blocks) */
    int persistent;          /* chosen launch form */
    int workgroups_per_cu;   /* resident fused workgroups per CU at this tile's LDS */
    int compute_units;       /* CUs of the device */
    int lds_bytes;           /* LDS per fused workgroup */
    int n_chunks;            /* fused + finalize launch pairs in the graph */
    long long work_items;    /* tiles x frames over all chunks */
    long long scratch_bytes; /* tile-region scratch held by the batch */
    int n_candidates;        /* candidates the autotuner timed (0 = not tuned) */
    float tuned_ms;          /* autotuner's time of the chosen candidate on the sample */
} dft2_batch_info;

typedef struct dft2_batch_candidate {
    int tile_w, tile_h;
    int persistent;
    int workgroups_per_cu;
    int lds_bytes;
    float ms;                /* average time on the autotuner's sample */
} dft2_batch_candidate;

/* Sets up `jobs` (device memory; the buffers must stay valid and at the same addresses
 * for the batch's lifetime) for repeated runs.  options = NULL means autotune with the
 * defaults.  With autotune the jobs' first sample_frames output frames are written
 * while timing.  The batch holds its own scratch, job table and captured graph. */
DFT2_API int dft2_batch_create(dft2_plan* plan, const dft2_job* jobs, int n_jobs, const dft2_batch_options* options,
                      dft2_batch** out_batch);

/* Runs the whole batch: one hipGraphLaunch on `stream` (NULL = the batch's own stream,
 * and the call waits for completion).  Runs of one batch must not overlap: they share
 * its scratch and work-queue counters. */
DFT2_API int dft2_batch_launch(dft2_batch* batch, void* stream);

/* Chosen configuration and sizes of the batch. */
DFT2_API int dft2_batch_get_info(const dft2_batch* batch, dft2_batch_info* info);

/* The autotuner's measurement for candidate `index` (0 .. n_candidates-1). */
DFT2_API int dft2_batch_get_candidate(const dft2_batch* batch, int index, dft2_batch_candidate* out);

/* Releases the batch (waits for its stream).  NULL is accepted. */
DFT2_API int dft2_batch_destroy(dft2_batch* batch);

/* FP64 real DFT of `data` viewed as `shape` (ndim 1..3), unnormalised, exponent -2*pi*i,
 * real along the last axis.  out receives prod(shape[:-1])*(shape[-1]/2+1) complex values
 * as interleaved (re, im) doubles; out_len is its capacity in doubles. */
DFT2_API int dft2_rdft(const double* data, int data_len, const int* shape, int ndim, double* out, int out_len);

/* Version string of the library. */
DFT2_API const char* dft2_version(void);

/* Message of the last error on the calling thread ("" if none). */
DFT2_API const char* dft2_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* DFTTEST2_MI355X_H */

// Internal definitions shared by the host C ABI (dft2_capi.cpp) and the gfx950
// kernels (dft2_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dft2 {

// Workgroup shape of the fused kernel: 8 wave64 waves sharing one tile accumulator.
constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;

// Layout of the per-plan constant table (floats).  All offsets are multiples of 4.
//   WX [x][n]   forward x transform, real basis n: 0..8 = cos(2 pi n x/16), 9..15 = sin(2 pi (n-8) x/16)
//   WY [m][y]   forward y transform, same basis along y
//   GX [n][x]   inverse x transform, WX^T with weights 1 (n = 0, 8) or 2 (others)
//   GY [y][m]   inverse y transform, WY^T with the same weights
//   WIN [t][y][x] analysis window times the sample scale (float32, as the plugin stores it)
//   SYN [y][x]  synthesis window (window slice t = radius)
//   The temporal transform is indexed from the centre slice, t' = t - r (t' = -r..r): its
//   basis functions are even (cos) or odd (sin) in t', slices r + k and r - k pair up,
//   and at the centre slice every sin row vanishes.  A coefficient differs from the
//   reference's (t = 0..T-1) by the phase exp(2 pi i kt r / T) only, which leaves every
//   gain unchanged; the window spectrum in TAB carries the same phase.
//   WT [j][k]   temporal real basis at t' = k (k = 1..r used): j = 0..r cos(2 pi j k/T),
//               j = r+1..2r sin(2 pi (j-r) k/T)
//   CT [j]      centre-slice inverse weights: 1 (j = 0), 2 (j = 1..r), 0 (sin rows)
//   TAB [g][slot][lane][4]  per-lane gain table of the middle stage: sigma, partner sigma,
//               Re and Im of window_freq at the coefficient the lane owns (see kernel),
//               times exp(2 pi i kt r / T).
constexpr int kOffWX = 0;
constexpr int kOffWY = 256;
constexpr int kOffGX = 512;
constexpr int kOffGY = 768;
constexpr int kOffWIN = 1024;                 // up to 7*256
constexpr int kOffSYN = kOffWIN + 7 * 256;    // 256
constexpr int kOffWT = kOffSYN + 256;         // up to 49, padded to 64
constexpr int kOffCT = kOffWT + 64;           // up to 7, padded to 16
constexpr int kOffTAB = kOffCT + 16;          // up to 4*4*64*4
constexpr int kTabFloats(int radius) { return (radius + 1) * 4 * 64 * 4; }
// Separable window w[t][y][x] = a[t] b[y] c[x] (the form DFTTest builds): the spatial
// factors fold into the bases, so the kernel needs no per-lane window values.
//   SWX [x][n] = WX[x][n] c[x]    (forward x operand; read transposed: inverse x operand)
//   SWY [m][y] = WY[m][y] b[y]    (forward y operand; read transposed: inverse y operand)
//   AS  [t]    = sample scale * a[t]   (per-slice factor of the analysis window; a[t] is
//                symmetric about the centre, a[r + k] = a[r - k], for the separable form)
//   CTS [j]    = CT[j] * a[r]          (centre weights with the synthesis window's a[r])
//   WTS [j][k] = WT[j][k] * a[r + k] / a[r]  (temporal basis with the slice factors
//                relative to the centre slice's; the kernel scales its x analysis operand
//                by AS[r], which every slice then carries; a[r] != 0 for the form)
constexpr int kOffSWX = kOffTAB + 4 * 4 * 64 * 4;
constexpr int kOffSWY = kOffSWX + 256;
constexpr int kOffAS = kOffSWY + 256;         // up to 7, padded to 8
constexpr int kOffCTS = kOffAS + 8;           // up to 7, padded to 8
constexpr int kOffWTS = kOffCTS + 8;          // up to 49, padded to 64
constexpr int kConstFloats = kOffWTS + 64;

// Source frames of one output frame at the largest radius (7, the generic block path).
constexpr int kMaxFramePtrs = 15;

// One plane of one clip in a launch (device copy of dft2_job plus derived geometry).
struct JobDev {
    const void* src;
    void* dst;
    long long src_fp, src_rs;   // frame pitch and row stride in samples
    long long dst_fp, dst_rs;
    int width, height, n_frames;
    int f_begin, f_count;       // output frames of this chunk
    int pad_w, pad_h, nbx, nby, ox, oy;
    int tiles_x, tiles_y;
    long long item_begin;       // first work item (fused kernel) of this job
    long long frame_begin;      // first job-frame index (finalize kernel) of this job
    int use_fptr;               // 1: source frame f is fptr[f] (f < T <= kMaxFramePtrs) instead of src + f*src_fp
    const void* fptr[kMaxFramePtrs];
    // Backward pass only: gradient of the output (same geometry as dst) and gradient of
    // the input (same geometry as src, accumulated with +=), pitches in samples.
    const void* grad_out;
    void* grad_in;
    long long go_fp, go_rs, gi_fp, gi_rs;
    // Gain-decision record (LaunchArgs::decisions): index of block (0, 0) of output frame
    // f_begin of this job in the record.
    long long dec_base;
    // Cross-frame form (LaunchArgs::reuse_k > 1): first work item of this job, one per
    // (group of reuse_k consecutive output frames, tile).
    long long group_begin;
};

struct LaunchArgs {
    const float* consts;
    const JobDev* jobs;
    int n_jobs;
    float* scratch;             // per work item: region_h * region_w floats
    int block_step;
    int step_magic;             // ceil(65536 / block_step): x / block_step by div_small
    int tile_bw, tile_bh;       // tile size in blocks
    int region_w, region_h;     // accumulator region of a full tile
    float inv_wf0;              // 1 / Re window_freq[0][0][0]
    float sigma2, pmin, pmax;
    int filter_type;
    int zero_mean;
    int asym;                   // sigma not symmetric under (kt,ky) -> (-kt,-ky) at kx = 0, 8
    float out_scale;            // sample scale (2^(8-bits) for int, 255 for float)
    int peak;                   // 2^bits - 1
    // Optional device sigma [T][16][9] (times sigma_scale) replacing the plan's sigma in
    // the gain table; used by the differentiable path, where sigma is a tensor.
    const float* sigma_dev;
    float sigma_scale;
    float* sigma_grad;          // backward: d loss / d sigma_dev, accumulated with +=
    // Gain table of this launch in global memory ((radius+1)*4*64 entries, see kOffTAB),
    // built by launch_gain_table from sigma_dev; nullptr = build per workgroup.
    const float4* tab_dev;
    // Persistent launch: work items [0, n_items) are taken from the counter *queue
    // (zeroed before the launch); nullptr = one work item per workgroup.  A launch has
    // fewer than 2^31 work items.
    unsigned* queue;
    int n_items;
    // Gain-decision record of the hard-threshold filters (instantiations with REC only):
    // per block (dec_base + (frame * nby + by) * nbx + bx), (radius+1)*4 64-bit masks
    // [g][slot], bit `lane` = the decision of the coefficient middle-stage lane `lane`
    // owns (filter type 1: psd < sigma, dropped; type 3: pmin <= psd <= pmax).
    unsigned long long* decisions;
    // Cross-frame form of the fused kernel: a work item is a tile over reuse_k consecutive
    // output frames, which share the spatial spectra of their common source slices; each
    // frame's accumulator is a ring of ring_bands bands of block_step rows (ring_bands =
    // ceil(16 / block_step)).  reuse_k = 1: one frame per work item, full-region accumulator.
    int reuse_k;
    int ring_bands;
};

// Generic block path (block sizes other than 16, tbsize above 7): table layout in floats,
// built in FP64 and rounded to float, and the launch arguments.
//   cos/sin(2 pi m / B), m < B; cos/sin(2 pi m / T), m < T;
//   WIN [t][y][x] analysis window times the sample scale; SYN [y][x] window slice r;
//   SIG [t][ky][kx] sigma (kx < B/2 + 1); WF [t][ky][kx][re, im] window spectrum.
struct GenericArgs {
    const float* cos_b;
    const float* sin_b;
    const float* cos_t;
    const float* sin_t;
    const float* win;
    const float* syn;
    const float* sig;
    const float* wf;
    const JobDev* jobs;
    int n_jobs;
    float* blockout;      // per block (work item): its B x B windowed centre-slice output
    int n_items;
    int block_size, half, t_len, radius, block_step;  // B, B/2 + 1, T, r, S
    int filter_type, zero_mean;
    float inv_wf0, sigma2, pmin, pmax;
    float out_scale;
    int peak;
    // Gain-decision record (filter types 1 and 3; nullptr = none): per block, dec_words
    // 64-bit masks, bit i = coefficient i in [kt][ky][kx] order (type 1: psd < sigma;
    // type 3: pmin <= psd <= pmax); block index job.dec_base + its index within the job.
    unsigned long long* decisions;
    int dec_words;
};
// LDS per wave of the generic block kernel: two complex T x B x (B/2+1) buffers.
inline size_t generic_lds_per_wave(int block_size, int t_len) {
    return sizeof(float) * 4 * static_cast<size_t>(t_len) * block_size * (block_size / 2 + 1);
}
hipError_t launch_generic_blocks(const GenericArgs& g, int sample_code, int waves_per_group, hipStream_t stream);
hipError_t launch_generic_gather(const GenericArgs& g, int sample_code, int max_width, int max_height, long long frames,
                                 hipStream_t stream);

// Selects the fused kernel instantiation.
struct FusedKey {
    int radius;
    int sample_code;
    int precision;
    int filter_type;
    int persistent; // 1: persistent grid over the work queue (LaunchArgs::queue)
    int asym;       // LaunchArgs::asym
    int sep;        // 1: separable-window form (constant table SWX/SWY/AS/CTS); the persistent
                    // form exists only with it
    int rec;        // 1: records gain decisions (LaunchArgs::decisions); exists for filter
                    // type 1 with a symmetric sigma and type 3 in FP32, one item per
                    // workgroup
    int reuse;      // 1: cross-frame form (LaunchArgs::reuse_k); separable, radius >= 1,
                    // one item per workgroup, no decision record
};

// Sample type codes.
enum SampleCode { kU8 = 0, kU16 = 1, kF32 = 2 };

// Fused kernel instantiations of one radius (each radius is its own object file).
template <int R>
struct RadiusTag {};
const void* fused_kernel_ptr(const FusedKey& k, RadiusTag<0>);
const void* fused_kernel_ptr(const FusedKey& k, RadiusTag<1>);
const void* fused_kernel_ptr(const FusedKey& k, RadiusTag<2>);
const void* fused_kernel_ptr(const FusedKey& k, RadiusTag<3>);

// Fused kernel over `grid` workgroups: one work item each, or a persistent grid over the
// work queue when a.queue is set.
hipError_t launch_fused(const LaunchArgs& a, const FusedKey& k, unsigned grid, size_t lds_bytes,
                        hipStream_t stream);
// Workgroups of the fused kernel resident per CU with `lds_bytes` of LDS (0 on error).
int fused_workgroups_per_cu(const FusedKey& k, size_t lds_bytes);
// Builds the gain table of a launch with sigma from a.sigma_dev into `out` (device,
// kTabFloats(radius) floats); bias: sigma + 1e-15 as the Wiener forward gain takes it.
hipError_t launch_gain_table(const LaunchArgs& a, int radius, bool bias, float4* out, hipStream_t stream);
// Zeroes n work-queue counters with a kernel (safe inside graph capture).
hipError_t launch_queue_reset(unsigned* queue, int n, hipStream_t stream);
// Finalize over the n_items work items (frame, tile) of a fused launch.
hipError_t launch_finalize(const LaunchArgs& a, int sample_code, long long n_items, hipStream_t stream);
// Backward pass over n_items work items: one per (frame, tile), or (a.reuse_k > 1, separable
// window, radius >= 1) the cross-frame form, one per (group of reuse_k frames, tile).
hipError_t launch_backward(const LaunchArgs& a, int radius, long long n_items, size_t lds_bytes, hipStream_t stream);
size_t fused_lds_bytes(int radius, int region_w, int region_h);
// LDS of the cross-frame form: reuse_k rings of ceil(16 / S) bands of S rows.
size_t fused_reuse_lds_bytes(int radius, int region_w, int block_step, int reuse_k);
size_t backward_lds_bytes(int radius, int region_w, int region_h);
// LDS of the cross-frame backward: reuse_k + 2 radius rings of ceil(16 / S) bands of S rows.
size_t backward_reuse_lds_bytes(int radius, int region_w, int block_step, int reuse_k);
// Workgroups of the backward kernel (reuse: the cross-frame form) resident per CU with
// `lds` bytes of LDS (0 on error or when the form does not exist).
int backward_workgroups_per_cu(int radius, bool reuse, size_t lds);

} // namespace dft2

// Fused DFTTest block pipeline for gfx950 (MI355X).
//
// One workgroup owns a tile of tile_bh x tile_bw blocks of one output frame of one
// plane.  Each wave64 takes blocks of the tile in turn and, per block:
//   1. loads the T 16x16 slices straight from HBM in the MFMA A-operand layout (lane
//      holds row y = lane%16, columns 4*(lane/16)..+3), applying the mirror padding and
//      temporal clamp in the address calculation (no padded copy is ever made) and the
//      analysis window (cpu_source/getframe_impl.cpp:105-156);
//   2. runs the separable forward real DFT as two chained MFMA contractions per slice,
//      D2 = WY * (slice * WX), in a real cos/sin basis (16 real values per axis); the
//      output of the first MFMA is already the B operand of the second, so no transpose
//      is needed.  FP32 MFMA, except that 8-bit samples, exact in BF16, meet the exact
//      three-part BF16 split of WX on the BF16 MFMA (exact products, FP32 sums), and
//      9- to 16-bit samples enter as their two bytes on the 16x16x32 BF16 MFMA;
//   3. forms the temporal real basis in registers, stages the T*256 values in LDS and
//      regroups them so that each lane owns one (y pair, x pair) group: from its 4*T
//      values it forms its 2*T complex coefficients X[kt][ky][kx], applies zero-mean,
//      the gain and the mean re-add (cpu_source/kernel.hpp:1819-1897), maps them back
//      to the real basis and evaluates the inverse over t at the centre slice;
//   4. runs the inverse real DFT (C2R convention, cpu_source/kernel.hpp:1978-1982) as
//      two chained MFMA contractions, multiplies by the synthesis window and stages the
//      block output in LDS; after each pass of 8 blocks the pixels' owning threads add
//      the staged outputs covering them to the tile accumulator in LDS, in a fixed order
//      and without atomics (cpu_source/getframe_impl.cpp:158-175).
// The finished tile region goes to a scratch buffer; the finalize kernel sums the at
// most 4 tile regions covering each pixel and converts to the sample type with the
// reference rounding (cpu_source/getframe_impl.cpp:177-221).
//
// The real-basis formulation is exact: a 16-point real transform has 16 real degrees
// of freedom (Re k = 0..8, Im k = 1..7), so every axis is a 16x16 (spatial) or TxT
// (temporal) real matrix, and the complex 3-D coefficients are sums/differences of the
// eight cos/sin products (see the middle stage below).

// Compiled once per radius (DFT2_FUSED_RADIUS = 0..3), each object providing
// fused_kernel_ptr(FusedKey, RadiusTag<R>) for its radius, so the instantiations build
// in parallel.

#include "dft2_device.h"

#ifndef DFT2_FUSED_RADIUS
#error "DFT2_FUSED_RADIUS must be defined (0..3)"
#endif

namespace dft2 {
namespace {

// Radius 0/1, Wiener or threshold gain, separable window: at most 80 VGPRs, so that six
// waves per SIMD (three workgroups per CU) fit.  The cross-frame form keeps T slice
// spectra live across its frames: at most 128 VGPRs at radius 1 and 2 (two workgroups
// per CU, which its LDS also allows).
// REC: also writes every block's gain decisions to a.decisions (parity instrumentation of
// the hard-threshold filters; separate instantiations, so the others carry no trace of it).
// REUSE: the cross-frame form.  A work item is a tile over K = a.reuse_k consecutive
// output frames; each wave keeps its block's T slice spectra D2 in registers and slides
// them from one frame to the next, so each further frame computes the spatial transform
// of one new slice instead of T (the reference recomputes every slice for every output
// frame, cpu_source/getframe_impl.cpp:330-386).  The separable window makes a slice's D2
// independent of its position in the window (the slice factors sit in the temporal
// basis).  Each frame's accumulator is a ring of ceil(16/S) bands of S rows: band r is
// final once block row r is done, and is then copied to the frame's scratch region and
// cleared, so K accumulators take the LDS of about one 16-row strip each.
template <int R, typename Tin, int FTM, int PREC, bool PERSIST, int ASYM, bool SEP, bool REC = false,
          bool REUSE = false>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(REUSE ? (R <= 2 ? 4 : 1)
                                          : (R <= 1 && FTM != 2 && SEP && !PERSIST && !REC ? 6 : 1))))
void fused_kernel(LaunchArgs a) {
    static_assert(!REUSE || (SEP && !PERSIST && !REC && R >= 1), "cross-frame form: separable, one item, R >= 1");
    constexpr int T = 2 * R + 1;
    // Overlap-add: staged block outputs summed by the pixels' owners after a barrier, in a
    // fixed order (LDS float atomics, the first form, cost more at every radius: 4K tbsize
    // 5, 171.5 against 143.8 ms).
    // Per-wave LDS: the T*256 spectrum staging, reused for V (16 x 20) and for the block
    // output, staged with row pitch kOut at offset wave*S (<= 7*16 floats) for the
    // overlap-add: 16 rows * 20 + 112 <= 448 floats.  Pitch 20 puts the four lane groups'
    // rows 4*(lane/16)+i in different banks.
    constexpr int kOut = 20;
    constexpr int kWbuf = (T * 256 > 448) ? T * 256 : 448;
    extern __shared__ __attribute__((aligned(16))) float smem[]; // float4 accesses

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = sgpr(tid >> 6);
    const int l16 = lane & 15;
    const int lg = lane >> 4;

    // Work items: one per workgroup, or (PERSIST) taken in turn from the work-queue
    // counter a.queue by a persistent grid until a.n_items are done.  The two forms are
    // separate instantiations: the loop costs registers the one-item form does not need.
    // Work items of one launch number fewer than 2^31 (the host splits larger work), so
    // item arithmetic stays 32-bit.
    __shared__ int s_item;
    int item = blockIdx.x;
    if constexpr (PERSIST) {
        if (tid == 0) {
            s_item = static_cast<int>(atomicAdd(a.queue, 1u));
        }
        __syncthreads();
        item = __builtin_amdgcn_readfirstlane(s_item);
    }

    const int S = a.block_step;
    const int rw = a.region_w;
    // Accumulators: the full tile region, or (REUSE) K rings of nbands bands of S rows.
    const int K = REUSE ? sgpr(a.reuse_k) : 1;
    const int nbands = REUSE ? sgpr(a.ring_bands) : 0;
    const int ring_floats = nbands * S * rw;
    float* acc = smem;
    float4* tab = reinterpret_cast<float4*>(smem + (REUSE ? K * ring_floats : a.region_h * rw));
    float* stage = reinterpret_cast<float*>(tab + (R + 1) * 4 * 64);
    float* wbuf = stage + wave * kWbuf;
    fill_tab<R, FTM == 0>(tab, a, tid, kThreads);

    // ---- per-lane constant operands ----
    // SEP: separable window, its spatial factors folded into the bases (kOffSWX), so no
    // per-lane window values are held (fewer registers, more waves per SIMD).
    constexpr int kOffX = SEP ? kOffSWX : kOffWX;
    constexpr int kOffY = SEP ? kOffSWY : kOffWY;
    float wx[4], wy[4], gx[4], gy[4], syn[4];
    float win[SEP ? 1 : T][4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        wx[s] = a.consts[kOffX + (4 * lg + s) * 16 + l16];
        wy[s] = a.consts[kOffY + l16 * 16 + 4 * lg + s];
        // Inverse operands: the unweighted basis (WX, WY read transposed).  The middle
        // stage leaves every coefficient of a non-degenerate pair at twice its value
        // (the 1/2 of the complex <-> real basis maps is not applied), which is exactly
        // the weight 2 of those basis functions in the inverse.
        gx[s] = a.consts[kOffX + l16 * 16 + 4 * lg + s];
        gy[s] = a.consts[kOffY + (4 * lg + s) * 16 + l16];
        if constexpr (!SEP) {
            syn[s] = a.consts[kOffSYN + (4 * lg + s) * 16 + l16];
#pragma unroll
            for (int t = 0; t < T; t++) {
                win[t][s] = a.consts[kOffWIN + (t * 16 + l16) * 16 + 4 * lg + s];
            }
        }
    }
    // SEP: sample scale * temporal window factor of the centre slice, applied to every
    // slice through the x analysis operand; the other slices' factors relative to it are
    // in the temporal basis (WTS).
    if constexpr (SEP) {
        const float as_c = a.consts[kOffAS + R];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            wx[s] *= as_c;
        }
    }
    // Centre-slice weights of the temporal inverse: c0 for the DC pair and 2 c0 for the
    // others (c0 = 1, or the centre slice's temporal synthesis factor when separable).
    // c0 rides in the inverse x operand, the 2 in the middle stage's fma.
    const float c0 = a.consts[SEP ? kOffCTS : kOffCT];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        gx[s] *= c0;
    }
    // The first contraction's form by sample type (separable window).  8-bit samples:
    // FP32, as BF16 (exact) against the exact three-part BF16 split of WX; split BF16, as
    // FP16 (exact, four instructions per four samples) against the two FP16 parts of
    // kBasisH16 * WX, whose scale the y operand takes out.  (FP16 in the FP32 form too:
    // the split form 1.1 % faster, the FP32 form 0.3 % slower, measured on 1080p u8.)
    // 16-bit samples: their two bytes on the 16x16x32 BF16 MFMA.
    constexpr bool kExactIn = SEP && sizeof(Tin) == 1;
    constexpr bool kExactIn16 = SEP && sizeof(Tin) == 2;
    constexpr bool kH16In = kExactIn && PREC == 1;
    const Opnd WXo = make_opnd<PREC>(wx[0], wx[1], wx[2], wx[3]);
    const Opnd3 WX3 = make_opnd3(wx[0], wx[1], wx[2], wx[3]);          // FP32, 8-bit samples
    const OpndH WXh = make_opndh<PREC>(wx[0], wx[1], wx[2], wx[3]);    // split, 8-bit samples
    const OpndW WXw = make_opndw<PREC>(wx[0], wx[1], wx[2], wx[3]);    // 16-bit samples
    constexpr float kYScale = kH16In ? 1.0f / kBasisH16 : 1.0f; // exact (a power of two)
    const Opnd WYo = make_opnd<PREC>(wy[0] * kYScale, wy[1] * kYScale, wy[2] * kYScale, wy[3] * kYScale);
    const Opnd GXo = make_opnd<PREC>(gx[0], gx[1], gx[2], gx[3]);
    const Opnd GYo = make_opnd<PREC>(gy[0], gy[1], gy[2], gy[3]);
    float wt[T][T];
#pragma unroll
    for (int j = 0; j < T; j++) {
#pragma unroll
        for (int t = 0; t < T; t++) {
            wt[j][t] = a.consts[(SEP ? kOffWTS : kOffWT) + j * T + t];
        }
    }

    // Middle-stage ownership: lane = 8*p + q owns y pair p (rows m = p, p+8 of the real
    // basis) and x pair q (columns n = q, q+8).  Pair 0 is (cos 0, cos 8) on both axes,
    // two purely real-axis coefficients; pairs 1..7 are (cos k, sin k).
    const int mp = lane >> 3;
    const int mq = lane & 7;
    const bool ey = mp != 0;
    const bool ex = mq != 0;
    // The same as 1/0 factors: "c ? a + b : a" is fma(cf, b, a), one instruction and the
    // same value for finite b.
    const float eyf = ey ? 1.0f : 0.0f;
    const float exf = ex ? 1.0f : 0.0f;
    const int wq = l16 & 7;
    const int wc = l16 >> 3;

    // Overlap-add of a pass for block steps 4..6 (one-frame form): the pass's columns (at
    // most 7S + 16 <= 64) are one per lane, and column x = lane is covered by at most four
    // of the pass's blocks, k0 .. k0 + 3 with k0 = max(hi - 3, 0).  Which of those four
    // cover it, and where block k0's value sits, depend on the lane and S only, so they are
    // found once here (the masks as lane masks, no per-pass compares or divisions).  The
    // bounds are those of a full pass (kWaves blocks); a pass with fewer blocks reads zeros
    // from the staging buffers of its idle waves, which clear them.  1080p u8 tbsize 3, per
    // half-clip launch on one box: fused 16.71 -> 16.22 ms (FP32), 13.53 -> 13.21 ms (split
    // BF16); tbsize 1 and 7 unchanged.  The path holds a VGPR and four lane masks for the
    // kernel's life, so it is compiled only where that costs no spills and no waves: the
    // one-item, separable, symmetric 8-bit instantiations of radius 0, 1 and 3.  The others
    // keep the general gather: with this path the float ones at radius 1 spill (the
    // differentiable forward, 8 x 256x256x16, 0.64 -> 0.69 ms), 16-bit, asymmetric,
    // persistent and radius-2 ones lose a wave per SIMD, and the cross-frame form's radius-2
    // ones spill more (4K 10-bit tbsize 5, FP32: 94.0 -> 97.2 ms).
    constexpr bool kGatherLane = !REUSE && !PERSIST && !REC && SEP && sizeof(Tin) == 1 && ASYM == 0 && R != 2;
    const bool gfast = kGatherLane && S >= 4 && 7 * S + 16 <= 64;
    const int g_lo = div_small(max(lane + S - 16, 0), a.step_magic);
    const int g_hi = min(div_small(lane, a.step_magic), kWaves - 1);
    const int g_k0 = max(g_hi - 3, 0);
    const bool gm0 = g_k0 >= g_lo && g_k0 <= g_hi;
    const bool gm1 = g_k0 + 1 >= g_lo && g_k0 + 1 <= g_hi;
    const bool gm2 = g_k0 + 2 >= g_lo && g_k0 + 2 <= g_hi;
    const bool gm3 = g_k0 + 3 >= g_lo && g_k0 + 3 <= g_hi;
    const int g_off = g_k0 * kWbuf + lane;

    while (item < a.n_items) {
    // ---- work item -> (job, frame, tile); REUSE: (job, first frame of a group of K, tile) ----
    const JobDev& job = a.jobs[sgpr(find_job<REUSE ? 2 : 0>(a.jobs, a.n_jobs, item))];
    const int tiles_x = sgpr(job.tiles_x);
    const int tiles = tiles_x * sgpr(job.tiles_y);
    const int job_begin = sgpr(static_cast<int>(REUSE ? job.group_begin : job.item_begin));
    if constexpr (!PERSIST) {
        // workgroup -> work item of the same job, each XCD a contiguous part of the job
        const int f_count = sgpr(job.f_count);
        item = job_xcd_item(item, job_begin, (REUSE ? (f_count + K - 1) / K : f_count) * tiles);
    }
    const int local = item - job_begin;
    const int group = local / tiles;
    const int tile = local - group * tiles;
    const int f_local = group * K;
    // frames of this item (K, fewer at the end of the job's frame range)
    const int kcount = REUSE ? min(K, sgpr(job.f_count) - f_local) : 1;
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int n = sgpr(job.f_begin) + f_local;
    const int by0 = ty * a.tile_bh;
    const int bx0 = tx * a.tile_bw;
    const int nby = min(a.tile_bh, sgpr(job.nby) - by0);
    const int nbx = min(a.tile_bw, sgpr(job.nbx) - bx0);
    const int rows_used = S * (nby - 1) + 16;
    const int width = sgpr(job.width);
    const int height = sgpr(job.height);
    const int ox = sgpr(job.ox);
    const int oy = sgpr(job.oy);
    // Row stride in samples; a frame holds fewer than 2^31 samples (checked on the host),
    // so sample offsets within a frame are 32-bit.
    const int rs = sgpr(static_cast<int>(job.src_rs));
    const unsigned lane_off = static_cast<unsigned>(l16 * rs + 4 * lg);

    // The region (rows_used x rw, rw a multiple of 4), or the K rings, zeroed; read out as
    // float4.
    float4* acc4 = reinterpret_cast<float4*>(acc);
    for (int i = tid; i < (REUSE ? K * ring_floats : rows_used * rw) / 4; i += kThreads) {
        acc4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }

    // Source planes of the T slices (temporal clamp, cpu_source/getframe_impl.cpp:294-298).
    // wide_ok: every plane base and the row pitch are multiples of kRaw4Align, so blocks
    // whose first sample is too take one wide load per slice.
    // REUSE: frame n + k (k >= 1) adds the source slice n + k + R, whose descriptor is made
    // when its samples are fetched (slice_rsrc).
    __amdgpu_buffer_rsrc_t frame[T];
    constexpr unsigned kAl = kRaw4Align<Tin>;
    bool wide_ok = (static_cast<unsigned>(rs) * sizeof(Tin)) % kAl == 0;
    const int n_frames = sgpr(job.n_frames);
    const unsigned frame_bytes = static_cast<unsigned>(rs * height) * static_cast<unsigned>(sizeof(Tin));
    auto slice_base = [&](int fi) -> const Tin* {
        fi = min(max(fi, 0), n_frames - 1);
        return sgpr(job.use_fptr) != 0 ? sgpr_ptr(reinterpret_cast<const Tin*>(job.fptr[fi]))
                                       : sgpr_ptr(reinterpret_cast<const Tin*>(job.src)) + static_cast<long long>(fi) *
                                                                                         sgpr64(job.src_fp);
    };
#pragma unroll
    for (int t = 0; t < T; t++) {
        const Tin* base = slice_base(n - R + t);
        frame[t] = frame_rsrc(base, frame_bytes);
        wide_ok = wide_ok && (reinterpret_cast<unsigned long long>(base) % kAl) == 0;
    }
    if constexpr (REUSE) {
        for (int k = 1; k < kcount; k++) {
            wide_ok = wide_ok && (reinterpret_cast<unsigned long long>(slice_base(n + k + R)) % kAl) == 0;
        }
    }
    // Samples of block (fy, fx) of the tile for this lane: row y = lane%16, columns
    // 4*(lane/16)..+3 of every slice, as raw bits.  Blocks that lie inside the plane
    // (most of them) take one offset per lane, and one wide load per slice when aligned;
    // the others apply the mirror padding per sample.
    constexpr unsigned kSz = sizeof(Tin);
    // fetch(fy, fx, fr, dst): slices from the N descriptors fr[] into dst[0..N-1].
    auto fetch = [&](int fy, int fx, const auto& fr, auto& dst) {
        constexpr int N = sizeof(fr) / sizeof(fr[0]);
        fx = opaque_uniform(fx); // a wave's column is fixed: keep its offsets out of the loop's registers
        const int y0 = (by0 + fy) * S - oy;
        const int x0 = (bx0 + fx) * S - ox;
        if (y0 >= 0 && y0 + 15 < height && x0 >= 0 && x0 + 15 < width) {
            const unsigned off = (static_cast<unsigned>(y0 * rs + x0) + lane_off) * kSz;
            if (wide_ok && (static_cast<unsigned>(y0 * rs + x0) * kSz) % kAl == 0) {
#pragma unroll
                for (int t = 0; t < N; t++) {
                    dst[t] = load4_aligned<Tin>(fr[t], off);
                }
            } else {
                const unsigned offs[4] = {off, off + kSz, off + 2 * kSz, off + 3 * kSz};
#pragma unroll
                for (int t = 0; t < N; t++) {
                    dst[t] = load4_each<Tin>(fr[t], offs);
                }
            }
        } else {
            const unsigned roff = static_cast<unsigned>(mirror(y0 + l16, height) * rs);
            unsigned offs[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                offs[s] = (roff + static_cast<unsigned>(mirror(x0 + 4 * lg + s, width))) * kSz;
            }
#pragma unroll
            for (int t = 0; t < N; t++) {
                dst[t] = load4_each<Tin>(fr[t], offs);
            }
        }
    };

    // REUSE: copies band b (region rows b*S .. b*S+S-1, fewer past rows_used) of every frame
    // of the item from ring slot `slot` to the frame's scratch region and zeroes the slot.
    // Scratch region of frame f_local + k = the work item (f_local + k, tile) of the
    // one-frame form, which finalize reads.
    // The K bands are spread over all threads of the workgroup (one float4 each), so no
    // wave carries the copy alone into the next barrier.
    auto flush_band = [&](int b, int slot) {
        const int rows = min(S, rows_used - b * S);
        const int n4 = rows * rw / 4;
        for (int e = tid; e < kcount * n4; e += kThreads) {
            const int k = e / n4;
            const int i = e - k * n4;
            float4* src = reinterpret_cast<float4*>(acc + k * ring_floats + slot * S * rw);
            const long long region_item = sgpr(static_cast<int>(job.item_begin)) + static_cast<long long>(f_local + k) * tiles + tile;
            float4* dst = reinterpret_cast<float4*>(a.scratch + region_item * (a.region_h * rw) + b * S * rw);
            dst[i] = src[i];
            src[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    };

    // Each block row r of the tile is done in passes of kWaves blocks: in pass p wave w
    // computes block (r, p*kWaves + w).  The wave's next block is fetched one block ahead,
    // so its load latency overlaps the current block's transforms.
    const int n_pass = (nbx + kWaves - 1) / kWaves;
    const int my_pass = nbx > wave ? (nbx - wave + kWaves - 1) / kWaves : 0; // passes with a block of this wave
    raw4<Tin> nxt[T];
    if (!REUSE && my_pass > 0) {
        fetch(0, wave, frame, nxt);
    }

    __syncthreads();

    floatx4 d2[T]; // slice spectra; REUSE slides them from one frame of the item to the next
    int r_band = 0; // REUSE: r mod nbands, the ring slot of band r
    for (int r = 0; r < nby; r++) {
    for (int pass = 0; pass < n_pass; pass++) {
    for (int k = 0; k < kcount; k++) {
    if (pass < my_pass) {
        // Slices to transform: all T for the item's first frame, else the one new slice,
        // which the prefetch put in nxt[T - 1].
        const bool all = !REUSE || k == 0;
        if (REUSE && k == 0) {
            // The block's T slices, fetched here rather than one block ahead: a prefetch
            // would keep T raw slices live across the middle stage of all K frames.
            fetch(r, pass * kWaves + wave, frame, nxt);
        }
        raw4<Tin> cur_raw[T]; // the samples as loaded (16-bit samples' exact path)
        float cur[T][4];      // and as floats (REUSE: converted where used, one slice or T)
#pragma unroll
        for (int t = 0; t < T; t++) {
            cur_raw[t] = nxt[t];
            if constexpr (!REUSE && !(SEP && sizeof(Tin) == 1 && PREC == 1)) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    cur[t][s] = raw4_value<Tin>(nxt[t], s);
                }
            }
        }
        {
            if constexpr (REUSE) {
                // the same block at the next frame: only its newest slice
                if (k + 1 < kcount) {
                    const __amdgpu_buffer_rsrc_t fr1[1] = {frame_rsrc(slice_base(n + k + 1 + R), frame_bytes)};
                    raw4<Tin>(&last)[1] = *reinterpret_cast<raw4<Tin>(*)[1]>(&nxt[T - 1]);
                    fetch(r, pass * kWaves + wave, fr1, last);
                }
            } else {
                const bool row_end = pass + 1 == my_pass;
                if (!row_end || r + 1 < nby) {
                    fetch(row_end ? r + 1 : r, (row_end ? 0 : pass + 1) * kWaves + wave, frame, nxt);
                }
            }
        }

        // ---- forward: D2_t = WY * (slice_t * WX) ----
        // Separable window: the slices enter unscaled (their factors are in WX and WTS).
        // 8-bit samples are exact in BF16 and in FP16, so the first contraction takes them
        // as they are: FP32, the BF16 MFMA over the exact three-part split of WX; split
        // BF16, the FP16 MFMA over the two FP16 parts of the scaled WX (every product
        // exact, FP32 accumulation).  16-bit samples enter as their two bytes on the
        // 16x16x32 BF16 MFMA (mm16_bytes).
        if (!all) {
#pragma unroll
            for (int t = 0; t + 1 < T; t++) {
                d2[t] = d2[t + 1];
            }
        }
#pragma unroll
        for (int t = 0; t < T; t++) {
            if (!all && t + 1 < T) {
                continue;
            }
            if constexpr (REUSE && !kH16In) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    cur[t][s] = raw4_value<Tin>(cur_raw[t], s);
                }
            }
            const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
            floatx4 d1;
            if constexpr (kH16In) {
                d1 = mm16_h<PREC>(bytes_h(cur_raw[t].v[0]), WXh, zero);
            } else if constexpr (kExactIn) {
                d1 = mm16_exact3(make_opnd_exact(cur[t][0], cur[t][1], cur[t][2], cur[t][3]), WX3, zero);
            } else if constexpr (kExactIn16) {
                d1 = mm16_bytes<PREC>(bytes_w(cur_raw[t].v[0], cur_raw[t].v[1]), WXw, zero);
            } else {
                float wv[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    wv[q] = SEP ? cur[t][q] : win[t][q] * cur[t][q];
                }
                d1 = mm16<PREC>(make_opnd<PREC>(wv[0], wv[1], wv[2], wv[3]), WXo, zero);
            }
            d2[t] = mm16<PREC>(WYo, make_opnd<PREC>(d1[0], d1[1], d1[2], d1[3]), zero);
        }

        // ---- temporal real basis (centre-indexed), staged as [j][p][q][b*2+c] ----
        // cos(2 pi j k/T) is even and sin odd in k = t - R, so the rows need only the
        // sums and differences of the slice pairs (R + k, R - k).
        floatx4 tsum[R + 1], tdif[R + 1]; // [k] for k = 1..R
#pragma unroll
        for (int k = 1; k <= R; k++) {
            tsum[k] = d2[R + k] + d2[R - k];
            tdif[k] = d2[R + k] - d2[R - k];
        }
        const floatx4 dc = d2[R];
#pragma unroll
        for (int j = 0; j < T; j++) {
            floatx4 u;
            if (j == 0) {
                u = dc;
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    u += SEP ? tsum[k] * wt[0][k] : tsum[k];
                }
            } else if (j <= R) {
                u = dc;
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    u += tsum[k] * wt[j][k];
                }
            } else if constexpr (R > 0) {
                u = tdif[1] * wt[j][1];
#pragma unroll
                for (int k = 2; k <= R; k++) {
                    u += tdif[k] * wt[j][k];
                }
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = 4 * lg + i;
                wbuf[((j * 8 + (m & 7)) * 8 + wq) * 4 + (m >> 3) * 2 + wc] = u[i];
            }
        }
        wave_lds_sync();

        // ---- middle stage: complex coefficients, gain, centre-slice inverse over t ----
        float4 U4[T];
#pragma unroll
        for (int j = 0; j < T; j++) {
            U4[j] = *reinterpret_cast<const float4*>(&wbuf[((j * 8 + mp) * 8 + mq) * 4]);
        }
        const float gf = a.zero_mean ? wbuf[0] * a.inv_wf0 : 0.0f;
        float v00 = 0.0f, v01 = 0.0f, v10 = 0.0f, v11 = 0.0f;
#pragma unroll
        for (int g = 0; g <= R; g++) {
            // c* = cos-t component, s* = sin-t component (zero for g = 0);
            // index [b][c] = (y slot, x slot).
            const float4 uc = U4[g];
            const float4 us = g ? U4[R + g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float c00 = uc.x, c01 = uc.y, c10 = uc.z, c11 = uc.w;
            const float s00 = us.x, s01 = us.y, s10 = us.z, s11 = us.w;

            // Two (t,y) evaluations per x slot c:
            //   proper y pair:  e1 = (+kt, +ky), e2 = (+kt, -ky)
            //   pair 0 (rows cos0, cos8): e1 = row ky=0, e2 = row ky=8
            float z1r[2], z1i[2], z2r[2], z2i[2];
            {
                const float cc0[2] = {c00, c01}, cc1[2] = {c10, c11};
                const float ss0[2] = {s00, s01}, ss1[2] = {s10, s11};
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    z1r[c] = fmaf(-eyf, ss1[c], cc0[c]);
                    z1i[c] = -fmaf(eyf, cc1[c], ss0[c]);
                    z2r[c] = ey ? (cc0[c] + ss1[c]) : cc1[c];
                    z2i[c] = ey ? (cc1[c] - ss0[c]) : -ss1[c];
                }
            }

            float zr[2][2], zi[2][2]; // [e][c] after the gains
            // Decision record of this block at temporal pair g: [g][slot], slot = 2e + k.
            unsigned long long* rec = nullptr;
            if constexpr (REC) {
                const long long blk = sgpr64(job.dec_base) +
                                      (static_cast<long long>(f_local) * sgpr(job.nby) + by0 + r) * sgpr(job.nbx) +
                                      bx0 + pass * kWaves + wave;
                rec = a.decisions + (blk * (R + 1) + g) * 4;
            }
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const float z0r = e ? z2r[0] : z1r[0];
                const float z0i = e ? z2i[0] : z1i[0];
                const float zxr = e ? z2r[1] : z1r[1];
                const float zxi = e ? z2i[1] : z1i[1];
                // proper x pair: A = Z0 - i Z1 (+kx), B = conj(Z0) - i conj(Z1) (conjugate
                // partner side); pair 0: A = column kx=0, B = column kx=8.
                float ar = fmaf(exf, zxi, z0r);
                float ai = fmaf(-exf, zxr, z0i);
                float br = ex ? z0r - zxi : zxr;
                float bi = ex ? -z0i - zxr : zxi;
                const bool da = filter_coef<FTM, ASYM>(ar, ai, tab_entry<ASYM>(&tab[(g * 4 + 2 * e) * 64 + lane]), gf, a);
                const bool db = filter_coef<FTM, ASYM>(br, bi, tab_entry<ASYM>(&tab[(g * 4 + 2 * e + 1) * 64 + lane]), gf, a);
                if constexpr (REC) {
                    const unsigned long long ma = __ballot(da);
                    const unsigned long long mb = __ballot(db);
                    if (lane == 0) {
                        rec[2 * e] = ma;
                        rec[2 * e + 1] = mb;
                    }
                }
                // x slots back from (A, B), doubled for a proper pair (see gx above).
                zr[e][0] = fmaf(exf, br, ar);
                zi[e][0] = fmaf(-exf, bi, ai);
                zr[e][1] = ex ? -(ai + bi) : br;
                zi[e][1] = ex ? ar - br : bi;
            }

            // Back to the real (t, y) basis at the centre slice, where the sin-t rows
            // vanish: only the cos-t components, times the centre weights (1 for g = 0,
            // 2 otherwise; their common factor c0 is in the inverse x operand).
            const float wc_ = g ? 2.0f : 1.0f;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float pr = zr[0][c], pi = zi[0][c], qr = zr[1][c], qi = zi[1][c];
                // doubled for a proper y pair (see gy above)
                const float u_c0 = fmaf(eyf, qr, pr); // cos t, y slot 0 (pr + qr for a proper pair)
                const float u_c1 = ey ? qi - pi : qr; // cos t, y slot 1
                if (c == 0) {
                    v00 = fmaf(wc_, u_c0, v00);
                    v10 = fmaf(wc_, u_c1, v10);
                } else {
                    v01 = fmaf(wc_, u_c0, v01);
                    v11 = fmaf(wc_, u_c1, v11);
                }
            }
        }
        wave_lds_sync();
        // V[m][n], row pitch 20 floats, aliased onto the staging buffer.
        wbuf[mp * 20 + mq] = v00;
        wbuf[mp * 20 + mq + 8] = v01;
        wbuf[(mp + 8) * 20 + mq] = v10;
        wbuf[(mp + 8) * 20 + mq + 8] = v11;
        wave_lds_sync();

        // ---- inverse: out = GY * (V * GX) ----
        const float4 v = *reinterpret_cast<const float4*>(&wbuf[l16 * 20 + 4 * lg]);
        const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
        const floatx4 f1 = mm16<PREC>(make_opnd<PREC>(v.x, v.y, v.z, v.w), GXo, zero);
        const floatx4 o = mm16<PREC>(GYo, make_opnd<PREC>(f1[0], f1[1], f1[2], f1[3]), zero);

        // ---- synthesis window, block output staged in LDS as [y][x] at offset wave*S ----
        wave_lds_sync(); // V is read before the staged output overwrites it
        float* srow = wbuf + wave * S + 4 * lg * kOut + l16;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            srow[i * kOut] = SEP ? o[i] : o[i] * syn[i];
        }
    } else if (gfast) {
        // no block of this wave in this pass: the gather's full-pass bounds read its
        // staging rows (y * kOut + x, y < 16, x < 64) as zeros
        for (int i = lane; i < 16 * kOut + 64; i += 64) {
            wbuf[i] = 0.0f;
        }
    }
        // ---- overlap-add of pass `pass` of block row r (REUSE: of frame k), without atomics ----
        // Every pixel (y, x) of the pass's 16-row band sums the staged outputs of the
        // pass's blocks covering it; the thread that owns the pixel in this phase adds the
        // sum to the accumulator.  With x local to the pass (origin: its first block), the
        // value of the pass's block k for pixel x sits at stage[k*kWbuf + y*kOut + x] (the
        // staging offset k*S cancels the block origin), so a wave's lanes read consecutive
        // banks.
        __syncthreads();
        if (gfast) {
            constexpr int kRows = 16 / kWaves;
            const int nb = min(kWaves, nbx - pass * kWaves);
            if (lane < S * (nb - 1) + 16) {
                float old[kRows], sum[kRows];
#pragma unroll
                for (int i = 0; i < kRows; i++) {
                    const int dy = i * kWaves + wave;
                    old[i] = acc[(r * S + dy) * rw + pass * kWaves * S + lane];
                    // blocks k0 .. k0 + 3 in order, as gather_head sums them; all four read
                    // (every address lies in the staging area), so that the reads pair up
                    // and are not made conditional
                    const float* p = stage + dy * kOut + g_off;
                    float v0 = p[0], v1 = p[kWbuf], v2 = p[2 * kWbuf], v3 = p[3 * kWbuf];
                    asm volatile("" : "+v"(v0), "+v"(v1), "+v"(v2), "+v"(v3));
                    float s = 0.0f;
                    s += gm0 ? v0 : 0.0f;
                    s += gm1 ? v1 : 0.0f;
                    s += gm2 ? v2 : 0.0f;
                    s += gm3 ? v3 : 0.0f;
                    sum[i] = s;
                }
#pragma unroll
                for (int i = 0; i < kRows; i++) {
                    const int dy = i * kWaves + wave;
                    acc[(r * S + dy) * rw + pass * kWaves * S + lane] = old[i] + sum[i];
                }
            }
        } else {
            const int Sg = opaque_uniform(S); // as fx in fetch: per-lane bounds stay in the loop
            const int Sm = opaque_uniform(a.step_magic); // for div_small
            const int nb = min(kWaves, nbx - pass * kWaves);
            const int cols = Sg * (nb - 1) + 16;
            for (int x0 = 0; x0 < cols; x0 += 64) {
                const int x = x0 + lane;
                if (x >= cols) {
                    continue;
                }
                // column x is covered by blocks ceil((x - 15) / S) .. floor(x / S); the
                // bounds and masks serve the wave's rows y = wave, wave + 8
                const int lo = div_small(max(x + Sg - 16, 0), Sm);
                const int hi = min(div_small(x, Sm), nb - 1);
                // All reads of the wave's rows (accumulator and staged values) first, one
                // LDS round trip, then the sums and the writes.
                constexpr int kRows = 16 / kWaves;
                // Accumulator row of the wave's row i (r * S + dy, dy = i * kWaves + wave):
                // at (r * S + dy) * rw, or (REUSE) in frame k's ring, band r + dy / S in slot
                // (r + dy / S) mod nbands.
                float* acol = acc + (r * Sg + wave) * rw + pass * kWaves * Sg + x;
                float* arow[kRows];
#pragma unroll
                for (int i = 0; i < kRows; i++) {
                    if constexpr (REUSE) {
                        const int dy = i * kWaves + wave;
                        const int q = div_small(dy, Sm);
                        int band = r_band + q;
                        band = band >= nbands ? band - nbands : band;
                        arow[i] = acc + k * ring_floats + (band * Sg + dy - q * Sg) * rw + pass * kWaves * Sg + x;
                    } else {
                        arow[i] = acol + i * kWaves * rw;
                    }
                }
                const int k0 = max(hi - 3, 0);
                float old[kRows], sum[kRows];
#pragma unroll
                for (int i = 0; i < kRows; i++) {
                    old[i] = *arow[i];
                    sum[i] = gather_head<kWbuf>(stage + (i * kWaves + wave) * kOut + x, k0, lo, hi);
                }
#pragma unroll
                for (int i = 0; i < kRows; i++) {
                    sum[i] = gather_tail<kWbuf>(stage + (i * kWaves + wave) * kOut + x, k0, lo, sum[i]);
                    *arow[i] = old[i] + sum[i];
                }
            }
        }
        __syncthreads(); // staged outputs consumed before the next pass reuses the buffers
    } // k
    } // pass
    if constexpr (REUSE) {
        // Band r is final in every frame of the item (later block rows start below it):
        // copy it to the frames' scratch regions and clear its ring slot for band r +
        // nbands.  The next writes to the slot follow the next pass's first barrier.
        flush_band(r, r_band);
        r_band = r_band + 1 == nbands ? 0 : r_band + 1;
    }
    } // r

    if constexpr (REUSE) {
        for (int b = nby; b * S < rows_used; b++) {
            flush_band(b, r_band);
            r_band = r_band + 1 == nbands ? 0 : r_band + 1;
        }
    } else {
        float4* out = reinterpret_cast<float4*>(a.scratch + static_cast<long long>(item) * (a.region_h * rw));
        for (int i = tid; i < rows_used * rw / 4; i += kThreads) {
            out[i] = acc4[i];
        }
    }
    if constexpr (!PERSIST) {
        break;
    }
    __syncthreads(); // the accumulator is read out and s_item consumed before reuse
    if (tid == 0) {
        s_item = static_cast<int>(atomicAdd(a.queue, 1u));
    }
    __syncthreads();
    item = __builtin_amdgcn_readfirstlane(s_item);
    }
}

template <int R, typename Tin, int PREC, bool PERSIST, bool SEP>
const void* fused_ptr_ft(int filter_type, bool asym, bool rec, bool reuse) {
    // Gains: 0 Wiener, 1 hard threshold (each with a symmetric and an asymmetric form),
    // 2 the generic form for types 2..6.  Decision records exist for type 1 with a
    // symmetric sigma and for type 3 in FP32, one item per workgroup.  The cross-frame
    // form exists for separable windows at radius >= 1, one item per workgroup.
    if (reuse) {
        if constexpr (!PERSIST && SEP && R >= 1) {
            if (filter_type == 0) {
                return asym ? reinterpret_cast<const void*>(fused_kernel<R, Tin, 0, PREC, false, 1, true, false, true>)
                            : reinterpret_cast<const void*>(fused_kernel<R, Tin, 0, PREC, false, 0, true, false, true>);
            }
            if (filter_type == 1) {
                return asym ? reinterpret_cast<const void*>(fused_kernel<R, Tin, 1, PREC, false, 1, true, false, true>)
                            : reinterpret_cast<const void*>(fused_kernel<R, Tin, 1, PREC, false, 0, true, false, true>);
            }
            return reinterpret_cast<const void*>(fused_kernel<R, Tin, 2, PREC, false, 2, true, false, true>);
        }
        return nullptr;
    }
    if (rec) {
        if constexpr (!PERSIST) {
            if (filter_type == 1 && !asym) {
                return reinterpret_cast<const void*>(fused_kernel<R, Tin, 1, PREC, false, 0, SEP, true>);
            }
            if constexpr (PREC == 0) {
                if (filter_type == 3) {
                    return reinterpret_cast<const void*>(fused_kernel<R, Tin, 2, 0, false, 2, SEP, true>);
                }
            }
        }
        return nullptr;
    }
    if (filter_type == 0) {
        return asym ? reinterpret_cast<const void*>(fused_kernel<R, Tin, 0, PREC, PERSIST, 1, SEP>)
                    : reinterpret_cast<const void*>(fused_kernel<R, Tin, 0, PREC, PERSIST, 0, SEP>);
    }
    if (filter_type == 1) {
        return asym ? reinterpret_cast<const void*>(fused_kernel<R, Tin, 1, PREC, PERSIST, 1, SEP>)
                    : reinterpret_cast<const void*>(fused_kernel<R, Tin, 1, PREC, PERSIST, 0, SEP>);
    }
    return reinterpret_cast<const void*>(fused_kernel<R, Tin, 2, PREC, PERSIST, 2, SEP>);
}

template <int R, typename Tin, bool PERSIST>
const void* fused_ptr_prec(const FusedKey& k) {
    // The persistent form exists for separable windows only; the one-item-per-workgroup
    // form exists for both.
    if constexpr (PERSIST) {
        if (!k.sep) {
            return nullptr;
        }
        return k.precision == 1 ? fused_ptr_ft<R, Tin, 1, true, true>(k.filter_type, k.asym != 0, k.rec != 0, k.reuse != 0)
                                : fused_ptr_ft<R, Tin, 0, true, true>(k.filter_type, k.asym != 0, k.rec != 0, k.reuse != 0);
    } else {
        if (k.sep) {
            return k.precision == 1 ? fused_ptr_ft<R, Tin, 1, false, true>(k.filter_type, k.asym != 0, k.rec != 0, k.reuse != 0)
                                    : fused_ptr_ft<R, Tin, 0, false, true>(k.filter_type, k.asym != 0, k.rec != 0, k.reuse != 0);
        }
        return k.precision == 1 ? fused_ptr_ft<R, Tin, 1, false, false>(k.filter_type, k.asym != 0, k.rec != 0, k.reuse != 0)
                                : fused_ptr_ft<R, Tin, 0, false, false>(k.filter_type, k.asym != 0, k.rec != 0, k.reuse != 0);
    }
}

template <int R, bool PERSIST>
const void* fused_ptr_type(const FusedKey& k) {
    switch (k.sample_code) {
    case kU8:
        return fused_ptr_prec<R, uint8_t, PERSIST>(k);
    case kU16:
        return fused_ptr_prec<R, uint16_t, PERSIST>(k);
    default:
        return fused_ptr_prec<R, float, PERSIST>(k);
    }
}

} // namespace

const void* fused_kernel_ptr(const FusedKey& k, RadiusTag<DFT2_FUSED_RADIUS>) {
    return k.persistent ? fused_ptr_type<DFT2_FUSED_RADIUS, true>(k) : fused_ptr_type<DFT2_FUSED_RADIUS, false>(k);
}

} // namespace dft2

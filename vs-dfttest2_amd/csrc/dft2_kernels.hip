// Fused DFTTest block pipeline for gfx950 (MI355X).
//
// One workgroup owns a tile of tile_bh x tile_bw blocks of one output frame of one
// plane.  Each wave64 takes blocks of the tile in turn and, per block:
//   1. loads the T 16x16 slices straight from HBM in the MFMA A-operand layout (lane
//      holds row y = lane%16, columns 4*(lane/16)..+3), applying the mirror padding and
//      temporal clamp in the address calculation (no padded copy is ever made) and the
//      analysis window (cpu_source/getframe_impl.cpp:105-156);
//   2. runs the separable forward real DFT as two chained FP32 MFMA contractions per
//      slice, D2 = WY * (slice * WX), in a real cos/sin basis (16 real values per axis);
//      the output of the first MFMA is already the B operand of the second, so no
//      transpose is needed;
//   3. forms the temporal real basis in registers, stages the T*256 values in LDS and
//      regroups them so that each lane owns one (y pair, x pair) group: from its 4*T
//      values it forms its 2*T complex coefficients X[kt][ky][kx], applies zero-mean,
//      the gain and the mean re-add (cpu_source/kernel.hpp:1819-1897), maps them back
//      to the real basis and evaluates the inverse over t at the centre slice;
//   4. runs the inverse real DFT (C2R convention, cpu_source/kernel.hpp:1978-1982) as
//      two chained MFMA contractions, multiplies by the synthesis window and overlap-
//      adds into the tile accumulator in LDS (cpu_source/getframe_impl.cpp:158-175).
// The finished tile region goes to a scratch buffer; the finalize kernel sums the at
// most 4 tile regions covering each pixel and converts to the sample type with the
// reference rounding (cpu_source/getframe_impl.cpp:177-221).
//
// The real-basis formulation is exact: a 16-point real transform has 16 real degrees
// of freedom (Re k = 0..8, Im k = 1..7), so every axis is a 16x16 (spatial) or TxT
// (temporal) real matrix, and the complex 3-D coefficients are sums/differences of the
// eight cos/sin products (see the middle stage below).

#include <type_traits>

#include "dft2_internal.h"

namespace dft2 {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short shortx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// One lane's share of a 16-long contraction operand: the value at K slot (lane/16, s)
// for s = 0..3, plus its split into BF16 hi + lo parts for the split-precision form.
struct Opnd {
    float f[4];
    bf16x4 hi, lo;
};

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// BF16 hi part of (a, b) as one packed register, and the FP32 values of its halves (a
// BF16 is the upper half of an FP32 with the same value).
__device__ __forceinline__ unsigned pack_hi(float a, float b, float& ha, float& hb) {
    const unsigned h = __builtin_bit_cast(unsigned, __builtin_convertvector(floatx2{a, b}, bf16x2));
    ha = __builtin_bit_cast(float, h << 16);
    hb = __builtin_bit_cast(float, h & 0xffff0000u);
    return h;
}

__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(floatx2{a, b}, bf16x2));
}

template <int PREC>
__device__ __forceinline__ Opnd make_opnd(float a0, float a1, float a2, float a3) {
    Opnd o;
    o.f[0] = a0;
    o.f[1] = a1;
    o.f[2] = a2;
    o.f[3] = a3;
    if constexpr (PREC == 1) {
        // Packed conversions: two values per instruction for hi and for lo.
        float h0, h1, h2, h3;
        const unsigned hi01 = pack_hi(a0, a1, h0, h1);
        const unsigned hi23 = pack_hi(a2, a3, h2, h3);
        const unsigned lo01 = pack_bf16(a0 - h0, a1 - h1);
        const unsigned lo23 = pack_bf16(a2 - h2, a3 - h3);
        o.hi = __builtin_bit_cast(bf16x4, (static_cast<unsigned long long>(hi23) << 32) | hi01);
        o.lo = __builtin_bit_cast(bf16x4, (static_cast<unsigned long long>(lo23) << 32) | lo01);
    }
    return o;
}

// acc += A * B over one 16-long contraction (4 K slots per lane x 4 lane groups).
// PREC 0: four FP32 MFMA 16x16x4.  PREC 1: split BF16 with FP32 accumulation,
// A*B ~ Ahi*Bhi + Alo*Bhi + Ahi*Blo, three BF16 MFMA 16x16x16.  The K slot pairing
// (lane group, s) is the same in every form, so operands need no reshuffle.
// The K-packed alternative (one 16x16x32 for [Ahi | Alo] x [Bhi | Bhi] plus one 16x16x16)
// returned corrupted products on gfx950 when the compiler scheduled a VALU write to a
// source register of the 16x16x32 right after it; with wait states inserted it was
// correct, and three 16x16x16 ran at the same speed, so this form is used.
template <int PREC>
__device__ __forceinline__ floatx4 mm16(const Opnd& A, const Opnd& B, floatx4 acc) {
    if constexpr (PREC == 0) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            acc = mfma4(A.f[s], B.f[s], acc);
        }
    } else {
        const shortx4 ahi = __builtin_bit_cast(shortx4, A.hi);
        const shortx4 alo = __builtin_bit_cast(shortx4, A.lo);
        const shortx4 bhi = __builtin_bit_cast(shortx4, B.hi);
        const shortx4 blo = __builtin_bit_cast(shortx4, B.lo);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, acc, 0, 0, 0);
    }
    return acc;
}

// Orders this wave's LDS accesses before and after this point, so that values one lane
// writes can be read by another lane of the same wave afterwards.  The LDS executes a
// wave's instructions in issue order, so only the compiler has to be kept from moving
// accesses across the point; nothing waits here (the compiler still waits for a load's
// data before its first use).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int mirror(int v, int n) {
    v = v < 0 ? -v : v;
    return v >= n ? 2 * (n - 1) - v : v;
}

// Values that are the same in every lane of the wave, moved to scalar registers (the
// compiler cannot prove that for values derived from threadIdx.x / 64 or read from the
// job table, and would otherwise keep them, and everything computed from them, per lane).
__device__ __forceinline__ int sgpr(int v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ long long sgpr64(long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(static_cast<unsigned long long>(v) >> 32));
    return static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo);
}

template <typename P>
__device__ __forceinline__ P* sgpr_ptr(P* p) {
    return reinterpret_cast<P*>(sgpr64(reinterpret_cast<long long>(p)));
}

// Source samples are read with raw buffer loads: a descriptor per source frame (base,
// size in bytes; loads past the size return 0 instead of faulting) and a 32-bit byte
// offset per lane, plus an immediate offset per column.  The raw bits stay in a register
// until the sample is used, so a prefetch keeps its loads in flight.
constexpr int kBufferDword3 = 0x00020000; // gfx9-family descriptor word 3 (raw access)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), kBufferDword3);
}

// Raw sample as loaded: the integer sample itself, or the bits of a float sample.
template <typename Tin>
using raw_t = typename std::conditional<sizeof(Tin) == 4, unsigned, Tin>::type;

template <typename Tin>
__device__ __forceinline__ raw_t<Tin> load_raw(__amdgpu_buffer_rsrc_t r, unsigned off_bytes);
template <>
__device__ __forceinline__ uint8_t load_raw<uint8_t>(__amdgpu_buffer_rsrc_t r, unsigned off_bytes) {
    return __builtin_amdgcn_raw_buffer_load_b8(r, off_bytes, 0, 0);
}
template <>
__device__ __forceinline__ uint16_t load_raw<uint16_t>(__amdgpu_buffer_rsrc_t r, unsigned off_bytes) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, off_bytes, 0, 0);
}
template <>
__device__ __forceinline__ unsigned load_raw<float>(__amdgpu_buffer_rsrc_t r, unsigned off_bytes) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, off_bytes, 0, 0);
}

template <typename Tin>
__device__ __forceinline__ float raw_value(raw_t<Tin> v) {
    if constexpr (sizeof(Tin) == 4) {
        return __builtin_bit_cast(float, v);
    } else {
        return static_cast<float>(v);
    }
}

// Source planes live in device global memory; loading through an address-space-1
// pointer gives global (not flat) loads.
template <typename Tin>
using gptr_t = const __attribute__((address_space(1))) Tin*;

template <typename Tin>
__device__ __forceinline__ gptr_t<Tin> to_global(const Tin* p) {
    return (gptr_t<Tin>)(p);
}

// Gain of one coefficient (cpu_source/kernel.hpp:1826-1890).  FTM 0 and 1 are the
// Wiener and hard-threshold filters; FTM 2 dispatches on the filter type at run time.
template <int FTM>
__device__ __forceinline__ float gain_of(float psd, float s, const LaunchArgs& a) {
    if constexpr (FTM == 0) {
        // s arrives as sigma + 1e-15 (gain table of a Wiener plan):
        // (psd - sigma) / (psd + 1e-15) = 1 - (sigma + 1e-15) / (psd + 1e-15).
        return fmaxf(fmaf(-s, __builtin_amdgcn_rcpf(psd + 1e-15f), 1.0f), 0.0f);
    } else if constexpr (FTM == 1) {
        return psd < s ? 0.0f : 1.0f;
    } else {
        switch (a.filter_type) {
        case 0:
            return fmaxf((psd - s) / (psd + 1e-15f), 0.0f);
        case 1:
            return psd < s ? 0.0f : 1.0f;
        case 2:
            return s;
        case 3:
            return (psd >= a.pmin && psd <= a.pmax) ? s : a.sigma2;
        case 4:
            return s * sqrtf(psd * (a.pmax / ((psd + a.pmin) * (psd + a.pmax) + 1e-15f)));
        case 5:
            return powf(fmaxf((psd - s) / (psd + 1e-15f), 0.0f), a.pmin);
        default:
            return sqrtf(fmaxf((psd - s) / (psd + 1e-15f), 0.0f));
        }
    }
}

// Zero-mean, gain and mean re-add on one complex coefficient.  e = (sigma, sigma of the
// conjugate partner, Re WF, Im WF) at this coefficient.  With an asymmetric sigma array
// the coefficients at kx = 0 and kx = 8 take the mean of their own gain and their
// conjugate partner's gain, which is what the reference's C2R (it drops Im at those
// columns) makes of two different gains.
// One gain-table entry from LDS as a single 128-bit read.  Without the empty asm the
// symmetric form, which does not use the partner sigma, gets split 32/64-bit reads at a
// 16-byte lane stride: twice the LDS instructions, with bank conflicts.
template <int ASYM>
__device__ __forceinline__ float4 tab_entry(const float4* p) {
    float4 e = *p;
    if constexpr (ASYM == 0) {
        asm volatile("" : "+v"(e.y));
    }
    return e;
}

// ASYM: 0 = symmetric sigma (no partner gains), 1 = asymmetric, 2 = decided at run time
// from a.asym.  The forms are separate kernel instantiations because the compiler turns
// a run-time test into selects and computes the partner gains for every plan.
template <int FTM, int ASYM>
__device__ __forceinline__ void filter_coef(float& re, float& im, float4 e, float gf, const LaunchArgs& a) {
    float xr = re - gf * e.z;
    float xi = im - gf * e.w;
    float psd = xr * xr + xi * xi;
    float g = gain_of<FTM>(psd, e.x, a);
    if constexpr (ASYM != 0) {
        if ((ASYM == 1 || a.asym) && e.y != e.x) { // only the kx = 0, 8 lanes
            g = 0.5f * (g + gain_of<FTM>(psd, e.y, a));
        }
    }
    re = xr * g + gf * e.z;
    im = xi * g + gf * e.w;
}

// Coefficient (kt, ky, kx) owned by middle-stage lane `lane` at temporal pair g and slot
// (A of evaluation 0, B of 0, A of 1, B of 1); see the middle stage of fused_kernel.
__device__ __forceinline__ int3 slot_position(int g, int slot, int lane, int t_len) {
    const int pp = lane >> 3, q = lane & 7;
    const int e = slot >> 1, k = slot & 1;
    const int ky = pp != 0 ? (e == 0 ? pp : (16 - pp) & 15) : (e == 0 ? 0 : 8);
    if (q != 0) {
        return k == 0 ? make_int3(g, ky, q) : make_int3((t_len - g) % t_len, (16 - ky) & 15, q);
    }
    return make_int3(g, ky, k == 0 ? 0 : 8);
}

// Conjugate partner whose sigma is averaged with `p` at kx = 0, 8 (itself elsewhere).
__device__ __forceinline__ int3 partner_position(int3 p, int t_len) {
    if (p.z == 0 || p.z == 8) {
        return make_int3((t_len - p.x) % t_len, (16 - p.y) & 15, p.z);
    }
    return p;
}

__device__ __forceinline__ int sigma_index(int3 p) {
    return (p.x * 16 + p.y) * 9 + p.z;
}

// Index of the last job whose first work item (or first job-frame) is <= v: the job
// table is sorted by both.  Jobs without work in a chunk share their start with the
// next job and are skipped by taking the last match.
template <bool kByFrame>
__device__ __forceinline__ int find_job(const JobDev* jobs, int n_jobs, long long v) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const long long start = kByFrame ? jobs[mid].frame_begin : jobs[mid].item_begin;
        if (start <= v) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

// Gain table into LDS: the plan's table, with sigma taken from a.sigma_dev when given.
// kBias: sigma + 1e-15 (the form gain_of<0> takes; the plan's table of a Wiener plan
// holds it already) instead of sigma.
template <int R, bool kBias>
__device__ __forceinline__ void fill_tab(float4* tab, const LaunchArgs& a, int tid, int n_threads) {
    constexpr int T = 2 * R + 1;
    if (a.tab_dev) { // built for this launch by gain_table_kernel
        for (int i = tid; i < (R + 1) * 4 * 64; i += n_threads) {
            tab[i] = a.tab_dev[i];
        }
        return;
    }
    const float4* gtab = reinterpret_cast<const float4*>(a.consts + kOffTAB);
    for (int i = tid; i < (R + 1) * 4 * 64; i += n_threads) {
        float4 e = gtab[i];
        if (a.sigma_dev) {
            const int lane = i & 63, slot = (i >> 6) & 3, g = i >> 8;
            const int3 pos = slot_position(g, slot, lane, T);
            e.x = a.sigma_dev[sigma_index(pos)] * a.sigma_scale;
            e.y = a.sigma_dev[sigma_index(partner_position(pos, T))] * a.sigma_scale;
            if constexpr (kBias) {
                e.x += 1e-15f;
                e.y += 1e-15f;
            }
        }
        tab[i] = e;
    }
}

// The gain table of one launch with sigma taken from a.sigma_dev, built once into global
// memory (every workgroup then copies it instead of gathering sigma itself).
template <int R, bool kBias>
__global__ __launch_bounds__(256) void gain_table_kernel(LaunchArgs a, float4* out) {
    a.tab_dev = nullptr;
    fill_tab<R, kBias>(out, a, blockIdx.x * 256 + threadIdx.x, gridDim.x * 256);
}

template <int R, typename Tin, int FTM, int PREC, bool PERSIST, int ASYM, bool SEP>
__global__ __launch_bounds__(kThreads) void fused_kernel(LaunchArgs a) {
    constexpr int T = 2 * R + 1;
    constexpr int kWbuf = (T * 256 > 320) ? T * 256 : 320;
    extern __shared__ float smem[];

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
    float* acc = smem;
    float4* tab = reinterpret_cast<float4*>(smem + a.region_h * rw);
    float* wbuf = reinterpret_cast<float*>(tab + (R + 1) * 4 * 64) + wave * kWbuf;
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
    float as[T]; // SEP: sample scale * temporal window factor of slice t (uniform)
#pragma unroll
    for (int t = 0; t < T; t++) {
        as[t] = SEP ? a.consts[kOffAS + t] : 1.0f;
    }
    const Opnd WXo = make_opnd<PREC>(wx[0], wx[1], wx[2], wx[3]);
    const Opnd WYo = make_opnd<PREC>(wy[0], wy[1], wy[2], wy[3]);
    const Opnd GXo = make_opnd<PREC>(gx[0], gx[1], gx[2], gx[3]);
    const Opnd GYo = make_opnd<PREC>(gy[0], gy[1], gy[2], gy[3]);
    float wt[T][T];
    float ct[T];
#pragma unroll
    for (int j = 0; j < T; j++) {
        ct[j] = a.consts[(SEP ? kOffCTS : kOffCT) + j];
#pragma unroll
        for (int t = 0; t < T; t++) {
            wt[j][t] = a.consts[kOffWT + j * T + t];
        }
    }

    // Middle-stage ownership: lane = 8*p + q owns y pair p (rows m = p, p+8 of the real
    // basis) and x pair q (columns n = q, q+8).  Pair 0 is (cos 0, cos 8) on both axes,
    // two purely real-axis coefficients; pairs 1..7 are (cos k, sin k).
    const int mp = lane >> 3;
    const int mq = lane & 7;
    const bool ey = mp != 0;
    const bool ex = mq != 0;
    const int wq = l16 & 7;
    const int wc = l16 >> 3;

    while (item < a.n_items) {
    // ---- work item -> (job, frame, tile) ----
    const JobDev& job = a.jobs[sgpr(find_job<false>(a.jobs, a.n_jobs, item))];
    const int tiles_x = sgpr(job.tiles_x);
    const int tiles = tiles_x * sgpr(job.tiles_y);
    const int local = item - sgpr(static_cast<int>(job.item_begin));
    const int f_local = local / tiles;
    const int tile = local - f_local * tiles;
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int n = sgpr(job.f_begin) + f_local;
    const int by0 = ty * a.tile_bh;
    const int bx0 = tx * a.tile_bw;
    const int nby = min(a.tile_bh, sgpr(job.nby) - by0);
    const int nbx = min(a.tile_bw, sgpr(job.nbx) - bx0);
    const int nblk = nby * nbx;
    const int rows_used = S * (nby - 1) + 16;
    const int width = sgpr(job.width);
    const int height = sgpr(job.height);
    const int ox = sgpr(job.ox);
    const int oy = sgpr(job.oy);
    // Row stride in samples; a frame holds fewer than 2^31 samples (checked on the host),
    // so sample offsets within a frame are 32-bit.
    const int rs = sgpr(static_cast<int>(job.src_rs));
    const unsigned lane_off = static_cast<unsigned>(l16 * rs + 4 * lg);

    for (int i = tid; i < rows_used * rw; i += kThreads) {
        acc[i] = 0.0f;
    }

    // Source planes of the T slices (temporal clamp, cpu_source/getframe_impl.cpp:294-298).
    __amdgpu_buffer_rsrc_t frame[T];
    {
        const int n_frames = sgpr(job.n_frames);
        const bool use_fptr = sgpr(job.use_fptr) != 0;
        const Tin* src = sgpr_ptr(reinterpret_cast<const Tin*>(job.src));
        const long long fp = sgpr64(job.src_fp);
        const unsigned frame_bytes = static_cast<unsigned>(rs * height) * static_cast<unsigned>(sizeof(Tin));
#pragma unroll
        for (int t = 0; t < T; t++) {
            const int fi = min(max(n - R + t, 0), n_frames - 1);
            const Tin* base = use_fptr ? sgpr_ptr(reinterpret_cast<const Tin*>(job.fptr[fi]))
                                       : src + static_cast<long long>(fi) * fp;
            frame[t] = frame_rsrc(base, frame_bytes);
        }
    }
    // Samples of block (fy, fx) of the tile for this lane: row y = lane%16, columns
    // 4*(lane/16)..+3 of every slice, as raw bits.  Blocks that lie inside the plane
    // (most of them) take one offset per lane; the others apply the mirror padding per
    // sample.
    constexpr unsigned kSz = sizeof(Tin);
    auto fetch = [&](int fy, int fx, raw_t<Tin> (&dst)[T][4]) {
        const int y0 = (by0 + fy) * S - oy;
        const int x0 = (bx0 + fx) * S - ox;
        if (y0 >= 0 && y0 + 15 < height && x0 >= 0 && x0 + 15 < width) {
            const unsigned off = (static_cast<unsigned>(y0 * rs + x0) + lane_off) * kSz;
#pragma unroll
            for (int t = 0; t < T; t++) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    dst[t][s] = load_raw<Tin>(frame[t], off + s * kSz);
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
            for (int t = 0; t < T; t++) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    dst[t][s] = load_raw<Tin>(frame[t], offs[s]);
                }
            }
        }
    };

    // The next block's samples are fetched one iteration ahead, so their load latency
    // overlaps the current block's transforms.
    // Block b of the tile is (b / nbx, b % nbx); the wave's blocks advance by kWaves.
    auto step_block = [&](int& by, int& bx) {
        bx += kWaves;
        while (bx >= nbx) {
            bx -= nbx;
            ++by;
        }
    };
    int nby_ = wave / nbx;
    int nbx_ = wave - nby_ * nbx;
    raw_t<Tin> nxt[T][4];
    if (wave < nblk) {
        fetch(nby_, nbx_, nxt);
    }

    __syncthreads();

    for (int b = wave; b < nblk; b += kWaves) {
        const int bly = nby_;
        const int blx = nbx_;
        float cur[T][4];
#pragma unroll
        for (int t = 0; t < T; t++) {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                cur[t][s] = raw_value<Tin>(nxt[t][s]);
            }
        }
        step_block(nby_, nbx_);
        if (b + kWaves < nblk) {
            fetch(nby_, nbx_, nxt);
        }

        // ---- forward: D2_t = WY * (slice_t * WX) ----
        floatx4 d2[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            float wv[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if constexpr (SEP) {
                    wv[q] = as[t] * cur[t][q];
                } else {
                    wv[q] = win[t][q] * cur[t][q];
                }
            }
            const Opnd in = make_opnd<PREC>(wv[0], wv[1], wv[2], wv[3]);
            const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
            const floatx4 d1 = mm16<PREC>(in, WXo, zero);
            d2[t] = mm16<PREC>(WYo, make_opnd<PREC>(d1[0], d1[1], d1[2], d1[3]), zero);
        }

        // ---- temporal real basis, staged as [j][p][q][b*2+c] ----
        // cos(2 pi j t/T) is even and sin odd under t -> T - t, so the rows need only
        // the sums and differences of the slice pairs (t, T - t).
        floatx4 tsum[R + 1], tdif[R + 1]; // [t] for t = 1..R
#pragma unroll
        for (int t = 1; t <= R; t++) {
            tsum[t] = d2[t] + d2[T - t];
            tdif[t] = d2[t] - d2[T - t];
        }
#pragma unroll
        for (int j = 0; j < T; j++) {
            floatx4 u;
            if (j == 0) {
                u = d2[0];
#pragma unroll
                for (int t = 1; t <= R; t++) {
                    u += tsum[t];
                }
            } else if (j <= R) {
                u = d2[0];
#pragma unroll
                for (int t = 1; t <= R; t++) {
                    u += tsum[t] * wt[j][t];
                }
            } else if constexpr (R > 0) {
                u = tdif[1] * wt[j][1];
#pragma unroll
                for (int t = 2; t <= R; t++) {
                    u += tdif[t] * wt[j][t];
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
                    z1r[c] = cc0[c] - (ey ? ss1[c] : 0.0f);
                    z1i[c] = -(ss0[c] + (ey ? cc1[c] : 0.0f));
                    z2r[c] = ey ? (cc0[c] + ss1[c]) : cc1[c];
                    z2i[c] = ey ? (cc1[c] - ss0[c]) : -ss1[c];
                }
            }

            float zr[2][2], zi[2][2]; // [e][c] after the gains
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const float z0r = e ? z2r[0] : z1r[0];
                const float z0i = e ? z2i[0] : z1i[0];
                const float zxr = e ? z2r[1] : z1r[1];
                const float zxi = e ? z2i[1] : z1i[1];
                // proper x pair: A = Z0 - i Z1 (+kx), B = conj(Z0) - i conj(Z1) (conjugate
                // partner side); pair 0: A = column kx=0, B = column kx=8.
                float ar = ex ? z0r + zxi : z0r;
                float ai = ex ? z0i - zxr : z0i;
                float br = ex ? z0r - zxi : zxr;
                float bi = ex ? -z0i - zxr : zxi;
                filter_coef<FTM, ASYM>(ar, ai, tab_entry<ASYM>(&tab[(g * 4 + 2 * e) * 64 + lane]), gf, a);
                filter_coef<FTM, ASYM>(br, bi, tab_entry<ASYM>(&tab[(g * 4 + 2 * e + 1) * 64 + lane]), gf, a);
                // x slots back from (A, B), doubled for a proper pair (see gx above).
                zr[e][0] = ex ? ar + br : ar;
                zi[e][0] = ex ? ai - bi : ai;
                zr[e][1] = ex ? -(ai + bi) : br;
                zi[e][1] = ex ? ar - br : bi;
            }

            // Back to the real (t, y) basis, then the centre-slice inverse weights.
            const float wc_ = ct[g];
            const float ws_ = g ? ct[R + g] : 0.0f;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float pr = zr[0][c], pi = zi[0][c], qr = zr[1][c], qi = zi[1][c];
                // doubled for a proper y pair (see gy above)
                const float u_c0 = ey ? pr + qr : pr;     // cos t, y slot 0
                const float u_s1 = ey ? qr - pr : -qi;    // sin t, y slot 1
                const float u_s0 = ey ? -(pi + qi) : -pi; // sin t, y slot 0
                const float u_c1 = ey ? qi - pi : qr;     // cos t, y slot 1
                const float y0 = wc_ * u_c0 + ws_ * u_s0;
                const float y1 = wc_ * u_c1 + ws_ * u_s1;
                if (c == 0) {
                    v00 += y0;
                    v10 += y1;
                } else {
                    v01 += y0;
                    v11 += y1;
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

        // ---- synthesis window + overlap-add into the tile accumulator ----
        float* arow = acc + (bly * S + 4 * lg) * rw + blx * S + l16;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            atomicAdd(arow + i * rw, SEP ? o[i] : o[i] * syn[i]);
        }
        wave_lds_sync();
    }
    __syncthreads();

    float* out = a.scratch + static_cast<long long>(item) * (a.region_h * rw);
    for (int i = tid; i < rows_used * rw; i += kThreads) {
        out[i] = acc[i];
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

// ---------------------------------------------------------------------------------------
// Backward pass (Wiener gain, float32 planes): gradients of a loss with respect to the
// input planes and to the sigma array, given the gradient of the output planes.
// It recomputes each block's spectrum (forward transform + middle stage) instead of
// storing it, then applies the adjoint of every stage in reverse:
//   output gradient -> x synthesis window / scale -> adjoint inverse DFT (two MFMA)
//   -> adjoint middle stage (mean re-add, gain, zero-mean, complex<->real basis maps)
//   -> adjoint temporal basis -> adjoint forward DFT (two MFMA) -> x analysis window
//   -> scatter-add into the input gradient through the same mirror / temporal clamp.
// With sigma taken from a device array the gain is always the average of a coefficient's
// own gain and its conjugate partner's (kx = 0, 8), so each sigma entry gets its share.

// Wiener gain g = max((psd - s)/(psd + eps), 0) and its derivatives where g > 0:
// dg/dpsd = (s + eps)/(psd + eps)^2, dg/ds = -1/(psd + eps); both 0 elsewhere.
__device__ __forceinline__ float3 wiener_d(float psd, float s) {
    const float r = __builtin_amdgcn_rcpf(psd + 1e-15f);
    const float g = (psd - s) * r;
    return g > 0.0f ? make_float3(g, (s + 1e-15f) * r * r, -r) : make_float3(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
    }
    return v;
}

// Adjoint of one gain application: x = A - gf*W, y = g(|x|^2) x + gf*W.  (dyr, dyi) in,
// gradient w.r.t. A out in (dyr, dyi); accumulates d/dsigma (own, partner) and d/dgf.
__device__ __forceinline__ void gain_adjoint(float ar, float ai, float& dyr, float& dyi, float4 e, float gf,
                                             float2& dsig, float& dgf) {
    const float xr = ar - gf * e.z;
    const float xi = ai - gf * e.w;
    const float psd = xr * xr + xi * xi;
    const float3 d1 = wiener_d(psd, e.x);
    const float3 d2 = wiener_d(psd, e.y);
    const float g = 0.5f * (d1.x + d2.x);
    const float h = 0.5f * (d1.y + d2.y);
    const float sdot = dyr * xr + dyi * xi;
    const float dxr = g * dyr + 2.0f * h * sdot * xr;
    const float dxi = g * dyi + 2.0f * h * sdot * xi;
    dsig.x += 0.5f * sdot * d1.z;
    dsig.y += 0.5f * sdot * d2.z;
    dgf += (dyr - dxr) * e.z + (dyi - dxi) * e.w;
    dyr = dxr;
    dyi = dxi;
}

template <int R>
__global__ __launch_bounds__(kThreads) void backward_kernel(LaunchArgs a) {
    constexpr int T = 2 * R + 1;
    // Contractions in split BF16 (hi + lo operands, three products, FP32 accumulation):
    // about 16 significant bits per operand, well inside the gradients' 1e-3 bound, at a
    // fraction of the FP32 MFMA cost.
    constexpr int kBP = 1;
    constexpr int kWbuf = T * 256 + 256;
    extern __shared__ float smem[];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = sgpr(tid >> 6);
    const int l16 = lane & 15;
    const int lg = lane >> 4;

    const int item = blockIdx.x;
    const JobDev& job = a.jobs[sgpr(find_job<false>(a.jobs, a.n_jobs, item))];
    const int tiles_x = sgpr(job.tiles_x);
    const int tiles = tiles_x * sgpr(job.tiles_y);
    const int local = item - sgpr(static_cast<int>(job.item_begin));
    const int f_local = local / tiles;
    const int tile = local - f_local * tiles;
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int n = sgpr(job.f_begin) + f_local;
    const int S = a.block_step;
    const int by0 = ty * a.tile_bh;
    const int bx0 = tx * a.tile_bw;
    const int nby = min(a.tile_bh, sgpr(job.nby) - by0);
    const int nbx = min(a.tile_bw, sgpr(job.nbx) - bx0);
    const int nblk = nby * nbx;
    const int width = sgpr(job.width);
    const int height = sgpr(job.height);
    const int ox = sgpr(job.ox);
    const int oy = sgpr(job.oy);
    const int rs = sgpr(static_cast<int>(job.src_rs));
    const int go_rs = sgpr(static_cast<int>(job.go_rs));

    // LDS: gain table | input-gradient regions [T][region_h][region_w] | per-wave staging.
    // After the block loop the staging area holds the sigma gradient of the workgroup.
    const int rw = a.region_w;
    const int rows_used = S * (nby - 1) + 16;
    const int cols_used = S * (nbx - 1) + 16;
    const int region = a.region_h * rw;
    float4* tab = reinterpret_cast<float4*>(smem);
    float* gacc = reinterpret_cast<float*>(tab + (R + 1) * 4 * 64);
    float* stage = gacc + T * region;
    float* wbuf = stage + wave * kWbuf;
    float* vbuf = wbuf + T * 256;
    float* dsig_lds = stage;
    static_assert(kWaves * kWbuf >= T * 144, "sigma gradient must fit the staging area");
    fill_tab<R, false>(tab, a, tid, kThreads);
    for (int i = tid; i < T * region; i += kThreads) {
        gacc[i] = 0.0f;
    }

    // The output-gradient chain uses the unweighted inverse basis (the transposes of
    // which are WX and WY themselves), as the forward does: the adjoint middle stage then
    // carries no 1/2 factors either.
    float wx[4], wy[4], wyT[4], wxT[4], synA[4];
    float win[T][4];
    const float inv_scale = 1.0f / a.out_scale;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        wx[s] = a.consts[kOffWX + (4 * lg + s) * 16 + l16];
        wy[s] = a.consts[kOffWY + l16 * 16 + 4 * lg + s];
        wyT[s] = a.consts[kOffWY + (4 * lg + s) * 16 + l16];
        wxT[s] = a.consts[kOffWX + l16 * 16 + 4 * lg + s];
        synA[s] = a.consts[kOffSYN + l16 * 16 + 4 * lg + s] * inv_scale;
#pragma unroll
        for (int t = 0; t < T; t++) {
            win[t][s] = a.consts[kOffWIN + (t * 16 + l16) * 16 + 4 * lg + s];
        }
    }
    const Opnd WXo = make_opnd<kBP>(wx[0], wx[1], wx[2], wx[3]);
    const Opnd WYo = make_opnd<kBP>(wy[0], wy[1], wy[2], wy[3]);
    const Opnd WYTo = make_opnd<kBP>(wyT[0], wyT[1], wyT[2], wyT[3]);
    const Opnd WXTo = make_opnd<kBP>(wxT[0], wxT[1], wxT[2], wxT[3]);
    float wt[T][T];
    float ct[T];
#pragma unroll
    for (int j = 0; j < T; j++) {
        ct[j] = a.consts[kOffCT + j];
#pragma unroll
        for (int t = 0; t < T; t++) {
            wt[j][t] = a.consts[kOffWT + j * T + t];
        }
    }

    __amdgpu_buffer_rsrc_t frame[T];
    float* gin[T];
    {
        const int n_frames = sgpr(job.n_frames);
        const float* src = sgpr_ptr(reinterpret_cast<const float*>(job.src));
        float* gi = sgpr_ptr(reinterpret_cast<float*>(job.grad_in));
        const long long fp = sgpr64(job.src_fp);
        const long long gfp = sgpr64(job.gi_fp);
        const unsigned frame_bytes = static_cast<unsigned>(rs * height) * 4u;
#pragma unroll
        for (int t = 0; t < T; t++) {
            const int fi = min(max(n - R + t, 0), n_frames - 1);
            frame[t] = frame_rsrc(src + static_cast<long long>(fi) * fp, frame_bytes);
            gin[t] = gi + static_cast<long long>(fi) * gfp;
        }
    }
    const __amdgpu_buffer_rsrc_t gout =
        frame_rsrc(sgpr_ptr(reinterpret_cast<const float*>(job.grad_out)) + static_cast<long long>(n) * sgpr64(job.go_fp),
                   static_cast<unsigned>(go_rs * height) * 4u);
    const unsigned lane_off = static_cast<unsigned>(l16 * rs + 4 * lg) * 4u;
    const unsigned lane_off_go = static_cast<unsigned>(l16 * go_rs + 4 * lg) * 4u;

    const int mp = lane >> 3;
    const int mq = lane & 7;
    const bool ey = mp != 0;
    const bool ex = mq != 0;
    const int wq = l16 & 7;
    const int wc = l16 >> 3;
    float2 dsg[R + 1][4];
#pragma unroll
    for (int g = 0; g <= R; g++) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            dsg[g][k] = make_float2(0.0f, 0.0f);
        }
    }

    __syncthreads();

    int bly = wave / nbx;
    int blx = wave - bly * nbx;
    for (int b = wave; b < nblk; b += kWaves, blx += kWaves) {
        while (blx >= nbx) {
            blx -= nbx;
            ++bly;
        }
        const int y0 = (by0 + bly) * S - oy;
        const int x0 = (bx0 + blx) * S - ox;
        const bool inside = y0 >= 0 && y0 + 15 < height && x0 >= 0 && x0 + 15 < width;
        const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};

        // ---- recompute the forward spectrum, staged as [j][p][q][b*2+c] ----
        float smp[T][4];
        if (inside) {
            const unsigned off = static_cast<unsigned>(y0 * rs + x0) * 4u + lane_off;
#pragma unroll
            for (int t = 0; t < T; t++) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    smp[t][s] = raw_value<float>(load_raw<float>(frame[t], off + 4u * s));
                }
            }
        } else {
            const unsigned roff = static_cast<unsigned>(mirror(y0 + l16, height) * rs);
            unsigned offs[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                offs[s] = (roff + static_cast<unsigned>(mirror(x0 + 4 * lg + s, width))) * 4u;
            }
#pragma unroll
            for (int t = 0; t < T; t++) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    smp[t][s] = raw_value<float>(load_raw<float>(frame[t], offs[s]));
                }
            }
        }
        floatx4 d2[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            const Opnd in = make_opnd<kBP>(win[t][0] * smp[t][0], win[t][1] * smp[t][1], win[t][2] * smp[t][2],
                                         win[t][3] * smp[t][3]);
            const floatx4 d1 = mm16<kBP>(in, WXo, zero);
            d2[t] = mm16<kBP>(WYo, make_opnd<kBP>(d1[0], d1[1], d1[2], d1[3]), zero);
        }
        {
            // temporal basis from the sums / differences of slices t and T - t
            floatx4 tsum[R + 1], tdif[R + 1];
#pragma unroll
            for (int t = 1; t <= R; t++) {
                tsum[t] = d2[t] + d2[T - t];
                tdif[t] = d2[t] - d2[T - t];
            }
#pragma unroll
            for (int j = 0; j < T; j++) {
                floatx4 u = d2[0];
                if (j == 0) {
#pragma unroll
                    for (int t = 1; t <= R; t++) {
                        u += tsum[t];
                    }
                } else if (j <= R) {
#pragma unroll
                    for (int t = 1; t <= R; t++) {
                        u += tsum[t] * wt[j][t];
                    }
                } else if constexpr (R > 0) {
                    u = tdif[1] * wt[j][1];
#pragma unroll
                    for (int t = 2; t <= R; t++) {
                        u += tdif[t] * wt[j][t];
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int m = 4 * lg + i;
                    wbuf[((j * 8 + (m & 7)) * 8 + wq) * 4 + (m >> 3) * 2 + wc] = u[i];
                }
            }
        }

        // ---- output gradient -> dV = IY^T (dO IX^T), dO = dy / scale * synthesis window ----
        {
            float dov[4];
            if (inside) {
                const unsigned off = static_cast<unsigned>(y0 * go_rs + x0) * 4u + lane_off_go;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    dov[s] = raw_value<float>(load_raw<float>(gout, off + 4u * s)) * synA[s];
                }
            } else {
                const int fy = y0 + l16;
                const bool row_ok = fy >= 0 && fy < height;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int fx = x0 + 4 * lg + s;
                    const bool ok = row_ok && fx >= 0 && fx < width;
                    const float v = raw_value<float>(
                        load_raw<float>(gout, ok ? static_cast<unsigned>(fy * go_rs + fx) * 4u : 0u));
                    dov[s] = ok ? v * synA[s] : 0.0f;
                }
            }
            const floatx4 pm = mm16<kBP>(make_opnd<kBP>(dov[0], dov[1], dov[2], dov[3]), WXo, zero);
            const floatx4 dv = mm16<kBP>(WYo, make_opnd<kBP>(pm[0], pm[1], pm[2], pm[3]), zero);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = 4 * lg + i;
                vbuf[((m & 7) * 8 + wq) * 4 + (m >> 3) * 2 + wc] = dv[i];
            }
        }
        wave_lds_sync();

        // ---- adjoint middle stage, one lane per (y pair, x pair) group ----
        float4 U4[T];
#pragma unroll
        for (int j = 0; j < T; j++) {
            U4[j] = *reinterpret_cast<const float4*>(&wbuf[((j * 8 + mp) * 8 + mq) * 4]);
        }
        const float4 dV4 = *reinterpret_cast<const float4*>(&vbuf[(mp * 8 + mq) * 4]);
        const float gf = a.zero_mean ? wbuf[0] * a.inv_wf0 : 0.0f;
        float dgf = 0.0f;
        float4 dU4[T];
#pragma unroll
        for (int g = 0; g <= R; g++) {
            const float4 uc = U4[g];
            const float4 us = g ? U4[R + g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float cc0[2] = {uc.x, uc.y}, cc1[2] = {uc.z, uc.w};
            const float ss0[2] = {us.x, us.y}, ss1[2] = {us.z, us.w};
            float z1r[2], z1i[2], z2r[2], z2i[2];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                z1r[c] = cc0[c] - (ey ? ss1[c] : 0.0f);
                z1i[c] = -(ss0[c] + (ey ? cc1[c] : 0.0f));
                z2r[c] = ey ? (cc0[c] + ss1[c]) : cc1[c];
                z2i[c] = ey ? (cc1[c] - ss0[c]) : -ss1[c];
            }
            // forward coefficients A, B of both evaluations (pre zero-mean)
            float ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const float z0r = e ? z2r[0] : z1r[0];
                const float z0i = e ? z2i[0] : z1i[0];
                const float zxr = e ? z2r[1] : z1r[1];
                const float zxi = e ? z2i[1] : z1i[1];
                ar[e] = ex ? z0r + zxi : z0r;
                ai[e] = ex ? z0i - zxr : z0i;
                br[e] = ex ? z0r - zxi : zxr;
                bi[e] = ex ? -z0i - zxr : zxi;
            }
            // dV -> d(real t,y basis) -> dP, dQ per x slot
            const float wcg = ct[g];
            const float wsg = g ? ct[R + g] : 0.0f;
            const float dv0[2] = {dV4.x, dV4.y}, dv1[2] = {dV4.z, dV4.w};
            float dzr[2][2], dzi[2][2]; // [e][c] gradients of the gained Z'
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float duc0 = wcg * dv0[c], dus0 = wsg * dv0[c];
                const float duc1 = wcg * dv1[c], dus1 = wsg * dv1[c];
                dzr[0][c] = ey ? duc0 - dus1 : duc0;
                dzr[1][c] = ey ? duc0 + dus1 : duc1;
                dzi[0][c] = ey ? -(dus0 + duc1) : -dus0;
                dzi[1][c] = ey ? duc1 - dus0 : -dus1;
            }
            float gz1r[2], gz1i[2], gz2r[2], gz2i[2]; // gradients of Z (pre gain), [c]
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const float d0r = dzr[e][0], d0i = dzi[e][0], d1r = dzr[e][1], d1i = dzi[e][1];
                float dar = ex ? d0r + d1i : d0r;
                float dai = ex ? d0i - d1r : d0i;
                float dbr = ex ? d0r - d1i : d1r;
                float dbi = ex ? -d0i - d1r : d1i;
                gain_adjoint(ar[e], ai[e], dar, dai, tab[(g * 4 + 2 * e) * 64 + lane], gf, dsg[g][2 * e], dgf);
                gain_adjoint(br[e], bi[e], dbr, dbi, tab[(g * 4 + 2 * e + 1) * 64 + lane], gf, dsg[g][2 * e + 1],
                             dgf);
                const float g0r = ex ? dar + dbr : dar;
                const float g0i = ex ? dai - dbi : dai;
                const float g1r = ex ? -dai - dbi : dbr;
                const float g1i = ex ? dar - dbr : dbi;
                if (e == 0) {
                    gz1r[0] = g0r, gz1i[0] = g0i, gz1r[1] = g1r, gz1i[1] = g1i;
                } else {
                    gz2r[0] = g0r, gz2i[0] = g0i, gz2r[1] = g1r, gz2i[1] = g1i;
                }
            }
            float dc0[2], dc1[2], ds0[2], ds1[2];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                dc0[c] = ey ? gz1r[c] + gz2r[c] : gz1r[c];
                ds1[c] = ey ? gz2r[c] - gz1r[c] : -gz2i[c];
                ds0[c] = ey ? -gz1i[c] - gz2i[c] : -gz1i[c];
                dc1[c] = ey ? gz2i[c] - gz1i[c] : gz2r[c];
            }
            dU4[g] = make_float4(dc0[0], dc0[1], dc1[0], dc1[1]);
            if (g) {
                dU4[R + g] = make_float4(ds0[0], ds0[1], ds1[0], ds1[1]);
            }
        }
        dgf = wave_sum(dgf);
        if (a.zero_mean && lane == 0) {
            dU4[0].x += dgf * a.inv_wf0;
        }
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < T; j++) {
            *reinterpret_cast<float4*>(&wbuf[((j * 8 + mp) * 8 + mq) * 4]) = dU4[j];
        }
        wave_lds_sync();

        // ---- adjoint temporal basis and forward DFT, scatter into the input gradient ----
        floatx4 du[T];
#pragma unroll
        for (int j = 0; j < T; j++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = 4 * lg + i;
                du[j][i] = wbuf[((j * 8 + (m & 7)) * 8 + wq) * 4 + (m >> 3) * 2 + wc];
            }
        }
        // adjoint temporal basis: slice t gets C_t + S_t and slice T - t gets C_t - S_t
        // (C: cos rows, even in t; S: sin rows, odd in t)
        floatx4 ddt[T];
        ddt[0] = du[0];
#pragma unroll
        for (int j = 1; j <= R; j++) {
            ddt[0] += du[j];
        }
        if constexpr (R > 0) {
#pragma unroll
            for (int t = 1; t <= R; t++) {
                floatx4 c = du[0];
#pragma unroll
                for (int j = 1; j <= R; j++) {
                    c += du[j] * wt[j][t];
                }
                floatx4 sn = du[R + 1] * wt[R + 1][t];
#pragma unroll
                for (int k = 2; k <= R; k++) {
                    sn += du[R + k] * wt[R + k][t];
                }
                ddt[t] = c + sn;
                ddt[T - t] = c - sn;
            }
        }
#pragma unroll
        for (int t = 0; t < T; t++) {
            const floatx4 dd = ddt[t];
            // dd (m in registers) used as A operand is dD2^T; R1 = dD2^T WY, db^T = WX R1.
            const floatx4 r1 = mm16<kBP>(make_opnd<kBP>(dd[0], dd[1], dd[2], dd[3]), WYTo, zero);
            const floatx4 dbt = mm16<kBP>(WXTo, make_opnd<kBP>(r1[0], r1[1], r1[2], r1[3]), zero);
            float* arow = gacc + t * region + (bly * S + l16) * rw + blx * S + 4 * lg;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                atomicAdd(arow + i, win[t][i] * dbt[i]);
            }
        }
        wave_lds_sync();
    }

    __syncthreads();
    for (int i = tid; i < T * 144; i += kThreads) {
        dsig_lds[i] = 0.0f;
    }
    // input gradient: padded tile regions -> source samples (mirror, temporal clamp).
    // Region element i = (ry, rx) is stepped without a division; tiles whose region lies
    // inside the plane skip the mirror.
    {
        const int gy0 = by0 * S - oy;
        const int gx0 = bx0 * S - ox;
        const bool inside = gy0 >= 0 && gy0 + rows_used <= height && gx0 >= 0 && gx0 + cols_used <= width;
        const long long gi_rs = sgpr64(job.gi_rs);
        const int dy = kThreads / rw;
        const int dx = kThreads - dy * rw;
        int ry = tid / rw;
        int rx = tid - ry * rw;
        for (int i = tid; i < rows_used * rw; i += kThreads) {
            if (rx < cols_used) {
                int fy = gy0 + ry;
                int fx = gx0 + rx;
                if (!inside) {
                    fy = mirror(fy, height);
                    fx = mirror(fx, width);
                }
                const long long off = static_cast<long long>(fy) * gi_rs + fx;
#pragma unroll
                for (int t = 0; t < T; t++) {
                    const float v = gacc[t * region + i];
                    if (v != 0.0f) {
                        atomicAdd(gin[t] + off, v);
                    }
                }
            }
            rx += dx;
            ry += dy;
            if (rx >= rw) {
                rx -= rw;
                ++ry;
            }
        }
    }

    // sigma gradient: lanes -> LDS (by coefficient) -> global
    __syncthreads(); // dsig_lds zeroed by every thread
#pragma unroll
    for (int g = 0; g <= R; g++) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int3 pos = slot_position(g, k, lane, T);
            atomicAdd(&dsig_lds[sigma_index(pos)], dsg[g][k].x);
            atomicAdd(&dsig_lds[sigma_index(partner_position(pos, T))], dsg[g][k].y);
        }
    }
    __syncthreads();
    for (int i = tid; i < T * 144; i += kThreads) {
        const float v = dsig_lds[i];
        if (v != 0.0f) {
            atomicAdd(&a.sigma_grad[i], v * a.sigma_scale);
        }
    }
}

template <typename Tout>
__device__ __forceinline__ void store_sample(Tout* p, float v, const LaunchArgs& a) {
    if constexpr (sizeof(Tout) == 4) {
        *p = v / a.out_scale;
    } else {
        float x = fmaxf(0.0f, fminf(v / a.out_scale + 0.5f, static_cast<float>(a.peak)));
        *p = static_cast<Tout>(static_cast<int>(x));
    }
}

// One workgroup per work item (frame, tile): the output pixels the tile owns (its S*tb x
// S*tb corner of the padded plane; the last tile in a row or column owns its whole
// region) take the tile's own region plus the 16 - S overlapping columns / rows of the
// left, top and top-left neighbours' regions, summed in a fixed order (top-left, top,
// left, own), then converted to the sample type.  Scratch region index = work item.
template <typename Tout>
__global__ __launch_bounds__(256) void finalize_kernel(LaunchArgs a) {
    const int item = blockIdx.x;
    const JobDev& job = a.jobs[sgpr(find_job<false>(a.jobs, a.n_jobs, item))];
    const int tiles_x = sgpr(job.tiles_x);
    const int tiles_y = sgpr(job.tiles_y);
    const int tiles = tiles_x * tiles_y;
    const int local = item - sgpr(static_cast<int>(job.item_begin));
    const int f_local = local / tiles;
    const int tile = local - f_local * tiles;
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int S = a.block_step;
    const int th = S * a.tile_bh;
    const int tw = S * a.tile_bw;
    const int y_lo = ty * th;
    const int x_lo = tx * tw;
    const int y_hi = ty == tiles_y - 1 ? y_lo + S * (sgpr(job.nby) - ty * a.tile_bh - 1) + 16 : y_lo + th;
    const int x_hi = tx == tiles_x - 1 ? x_lo + S * (sgpr(job.nbx) - tx * a.tile_bw - 1) + 16 : x_lo + tw;
    const int oy = sgpr(job.oy);
    const int ox = sgpr(job.ox);
    const int y0 = max(y_lo, oy);
    const int y1 = min(y_hi, oy + sgpr(job.height));
    const int x0 = max(x_lo, ox);
    const int x1 = min(x_hi, ox + sgpr(job.width));
    if (y0 >= y1 || x0 >= x1) {
        return;
    }
    const int rw = a.region_w;
    const long long region = static_cast<long long>(a.region_h) * rw;
    const float* own = a.scratch + static_cast<long long>(item) * region;
    const float* left = own - region;
    const float* top = own - tiles_x * region;
    const float* corner = top - region;
    const int ov = 16 - S;
    const bool has_left = tx > 0;
    const bool has_top = ty > 0;
    Tout* dst = reinterpret_cast<Tout*>(sgpr_ptr(job.dst)) +
                static_cast<long long>(sgpr(job.f_begin) + f_local) * sgpr64(job.dst_fp);
    const long long drs = sgpr64(job.dst_rs);
    const int w = x1 - x0;
    const int h = y1 - y0;
    // pixel i of the w x h area is (r, c) = (i / w, i % w), stepped without a division
    const int dr = 256 / w;
    const int dc = 256 - dr * w;
    int r = threadIdx.x / w;
    int c = threadIdx.x - r * w;
    for (; r < h;) {
        const int ly = y0 + r - y_lo;
        const int lx = x0 + c - x_lo;
        float sum = 0.0f;
        if (has_top && ly < ov) {
            if (has_left && lx < ov) {
                sum += corner[(ly + th) * rw + lx + tw];
            }
            sum += top[(ly + th) * rw + lx];
        }
        if (has_left && lx < ov) {
            sum += left[ly * rw + lx + tw];
        }
        sum += own[ly * rw + lx];
        store_sample(dst + (y0 + r - oy) * drs + (x0 + c - ox), sum, a);
        c += dc;
        r += dr;
        if (c >= w) {
            c -= w;
            ++r;
        }
    }
}

template <int R, typename Tin, int PREC, bool PERSIST, bool SEP>
const void* fused_ptr_ft(int filter_type, bool asym) {
    // Gains: 0 Wiener, 1 hard threshold (each with a symmetric and an asymmetric form),
    // 2 the generic form for types 2..6.
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
        return k.precision == 1 ? fused_ptr_ft<R, Tin, 1, true, true>(k.filter_type, k.asym != 0)
                                : fused_ptr_ft<R, Tin, 0, true, true>(k.filter_type, k.asym != 0);
    } else {
        if (k.sep) {
            return k.precision == 1 ? fused_ptr_ft<R, Tin, 1, false, true>(k.filter_type, k.asym != 0)
                                    : fused_ptr_ft<R, Tin, 0, false, true>(k.filter_type, k.asym != 0);
        }
        return k.precision == 1 ? fused_ptr_ft<R, Tin, 1, false, false>(k.filter_type, k.asym != 0)
                                : fused_ptr_ft<R, Tin, 0, false, false>(k.filter_type, k.asym != 0);
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

template <bool PERSIST>
const void* fused_ptr_radius(const FusedKey& k) {
    switch (k.radius) {
    case 0:
        return fused_ptr_type<0, PERSIST>(k);
    case 1:
        return fused_ptr_type<1, PERSIST>(k);
    case 2:
        return fused_ptr_type<2, PERSIST>(k);
    default:
        return fused_ptr_type<3, PERSIST>(k);
    }
}

const void* fused_ptr(const FusedKey& k) {
    return k.persistent ? fused_ptr_radius<true>(k) : fused_ptr_radius<false>(k);
}

hipError_t allow_lds(const void* kern, size_t lds) {
    if (lds > 65536) {
        return hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    }
    return hipSuccess;
}

} // namespace

size_t backward_lds_bytes(int radius, int region_w, int region_h) {
    const int t_len = 2 * radius + 1;
    return sizeof(float) * (static_cast<size_t>(radius + 1) * 4 * 64 * 4 +
                            static_cast<size_t>(t_len) * region_w * region_h +
                            static_cast<size_t>(kWaves) * (t_len * 256 + 256));
}

hipError_t launch_backward(const LaunchArgs& a, int radius, long long n_items, size_t lds, hipStream_t stream) {
    auto run = [&](auto kern) -> hipError_t {
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
            if (e != hipSuccess) {
                return e;
            }
        }
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(n_items)), dim3(kThreads), lds, stream, a);
        return hipGetLastError();
    };
    switch (radius) {
    case 0:
        return run(backward_kernel<0>);
    case 1:
        return run(backward_kernel<1>);
    case 2:
        return run(backward_kernel<2>);
    default:
        return run(backward_kernel<3>);
    }
}

size_t fused_lds_bytes(int radius, int region_w, int region_h) {
    const int t_len = 2 * radius + 1;
    const int wbuf = (t_len * 256 > 320) ? t_len * 256 : 320;
    return sizeof(float) * (static_cast<size_t>(region_w) * region_h + (radius + 1) * 4 * 64 * 4 + kWaves * wbuf);
}

hipError_t launch_fused(const LaunchArgs& a, const FusedKey& k, unsigned grid, size_t lds_bytes,
                        hipStream_t stream) {
    if (k.persistent != (a.queue != nullptr) || k.asym != a.asym) {
        return hipErrorInvalidValue;
    }
    const void* kern = fused_ptr(k);
    if (!kern) {
        return hipErrorInvalidValue;
    }
    hipError_t e = allow_lds(kern, lds_bytes);
    if (e != hipSuccess) {
        return e;
    }
    LaunchArgs args = a;
    if (!k.persistent) {
        args.n_items = static_cast<int>(grid);
    }
    void* params[] = {&args};
    return hipLaunchKernel(kern, dim3(grid), dim3(kThreads), params, lds_bytes, stream);
}

int fused_workgroups_per_cu(const FusedKey& k, size_t lds_bytes) {
    const void* kern = fused_ptr(k);
    if (!kern || allow_lds(kern, lds_bytes) != hipSuccess) {
        return 0;
    }
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, kThreads, lds_bytes) != hipSuccess) {
        return 0;
    }
    return n;
}

__global__ void queue_reset_kernel(unsigned* q, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        q[i] = 0;
    }
}

hipError_t launch_gain_table(const LaunchArgs& a, int radius, bool bias, float4* out, hipStream_t stream) {
    auto run = [&](auto kern) -> hipError_t {
        hipLaunchKernelGGL(kern, dim3(4), dim3(256), 0, stream, a, out);
        return hipGetLastError();
    };
    switch (radius) {
    case 0:
        return bias ? run(gain_table_kernel<0, true>) : run(gain_table_kernel<0, false>);
    case 1:
        return bias ? run(gain_table_kernel<1, true>) : run(gain_table_kernel<1, false>);
    case 2:
        return bias ? run(gain_table_kernel<2, true>) : run(gain_table_kernel<2, false>);
    default:
        return bias ? run(gain_table_kernel<3, true>) : run(gain_table_kernel<3, false>);
    }
}

hipError_t launch_queue_reset(unsigned* queue, int n, hipStream_t stream) {
    hipLaunchKernelGGL(queue_reset_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, queue, n);
    return hipGetLastError();
}

hipError_t launch_finalize(const LaunchArgs& a, int sample_code, long long n_items, hipStream_t stream) {
    const dim3 grid(static_cast<unsigned>(n_items));
    switch (sample_code) {
    case kU8:
        hipLaunchKernelGGL(finalize_kernel<uint8_t>, grid, dim3(256), 0, stream, a);
        break;
    case kU16:
        hipLaunchKernelGGL(finalize_kernel<uint16_t>, grid, dim3(256), 0, stream, a);
        break;
    default:
        hipLaunchKernelGGL(finalize_kernel<float>, grid, dim3(256), 0, stream, a);
        break;
    }
    return hipGetLastError();
}

} // namespace dft2

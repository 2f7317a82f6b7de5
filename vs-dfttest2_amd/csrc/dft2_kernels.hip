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

template <int PREC>
__device__ __forceinline__ Opnd make_opnd(float a0, float a1, float a2, float a3) {
    Opnd o;
    o.f[0] = a0;
    o.f[1] = a1;
    o.f[2] = a2;
    o.f[3] = a3;
    if constexpr (PREC == 1) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const __bf16 h = static_cast<__bf16>(o.f[s]);
            o.hi[s] = h;
            o.lo[s] = static_cast<__bf16>(o.f[s] - static_cast<float>(h));
        }
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

// Completes this wave's outstanding LDS operations and stops the compiler from moving
// memory accesses across this point, so that LDS values written by one lane can be
// read by another lane of the same wave afterwards.
__device__ __forceinline__ void wave_lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

__device__ __forceinline__ int mirror(int v, int n) {
    v = v < 0 ? -v : v;
    return v >= n ? 2 * (n - 1) - v : v;
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
        return fmaxf((psd - s) * __builtin_amdgcn_rcpf(psd + 1e-15f), 0.0f);
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
template <int FTM>
__device__ __forceinline__ void filter_coef(float& re, float& im, float4 e, float gf, const LaunchArgs& a) {
    float xr = re - gf * e.z;
    float xi = im - gf * e.w;
    float psd = xr * xr + xi * xi;
    float g = gain_of<FTM>(psd, e.x, a);
    if (a.asym) {
        g = 0.5f * (g + gain_of<FTM>(psd, e.y, a));
    }
    re = xr * g + gf * e.z;
    im = xi * g + gf * e.w;
}

template <int R, typename Tin, int FTM, int PREC>
__global__ __launch_bounds__(kThreads) void fused_kernel(LaunchArgs a) {
    constexpr int T = 2 * R + 1;
    constexpr int kWbuf = (T * 256 > 320) ? T * 256 : 320;
    extern __shared__ float smem[];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int l16 = lane & 15;
    const int lg = lane >> 4;

    // ---- work item -> (job, frame, tile) ----
    const long long item = blockIdx.x;
    int jb = 0;
    while (jb + 1 < a.n_jobs && a.jobs[jb + 1].item_begin <= item) {
        ++jb;
    }
    const JobDev& job = a.jobs[jb];
    const int tiles = job.tiles_x * job.tiles_y;
    const long long local = item - job.item_begin;
    const int f_local = static_cast<int>(local / tiles);
    const int tile = static_cast<int>(local - static_cast<long long>(f_local) * tiles);
    const int ty = tile / job.tiles_x;
    const int tx = tile - ty * job.tiles_x;
    const int n = job.f_begin + f_local;
    const int S = a.block_step;
    const int by0 = ty * a.tile_bh;
    const int bx0 = tx * a.tile_bw;
    const int nby = min(a.tile_bh, job.nby - by0);
    const int nbx = min(a.tile_bw, job.nbx - bx0);
    const int nblk = nby * nbx;
    const int rw = a.region_w;
    const int rows_used = S * (nby - 1) + 16;

    float* acc = smem;
    float4* tab = reinterpret_cast<float4*>(smem + a.region_h * rw);
    float* wbuf = reinterpret_cast<float*>(tab + (R + 1) * 4 * 64) + wave * kWbuf;

    for (int i = tid; i < rows_used * rw; i += kThreads) {
        acc[i] = 0.0f;
    }
    const float4* gtab = reinterpret_cast<const float4*>(a.consts + kOffTAB);
    for (int i = tid; i < (R + 1) * 4 * 64; i += kThreads) {
        tab[i] = gtab[i];
    }

    // ---- per-lane constant operands ----
    float wx[4], wy[4], gx[4], gy[4], syn[4];
    float win[T][4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        wx[s] = a.consts[kOffWX + (4 * lg + s) * 16 + l16];
        wy[s] = a.consts[kOffWY + l16 * 16 + 4 * lg + s];
        gx[s] = a.consts[kOffGX + (4 * lg + s) * 16 + l16];
        gy[s] = a.consts[kOffGY + l16 * 16 + 4 * lg + s];
        syn[s] = a.consts[kOffSYN + (4 * lg + s) * 16 + l16];
#pragma unroll
        for (int t = 0; t < T; t++) {
            win[t][s] = a.consts[kOffWIN + (t * 16 + l16) * 16 + 4 * lg + s];
        }
    }
    const Opnd WXo = make_opnd<PREC>(wx[0], wx[1], wx[2], wx[3]);
    const Opnd WYo = make_opnd<PREC>(wy[0], wy[1], wy[2], wy[3]);
    const Opnd GXo = make_opnd<PREC>(gx[0], gx[1], gx[2], gx[3]);
    const Opnd GYo = make_opnd<PREC>(gy[0], gy[1], gy[2], gy[3]);
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

    // Source planes of the T slices (temporal clamp, cpu_source/getframe_impl.cpp:294-298).
    gptr_t<Tin> frame[T];
#pragma unroll
    for (int t = 0; t < T; t++) {
        int fi = min(max(n - R + t, 0), job.n_frames - 1);
        frame[t] = to_global(job.use_fptr ? reinterpret_cast<const Tin*>(job.fptr[fi])
                                          : reinterpret_cast<const Tin*>(job.src) + static_cast<long long>(fi) * job.src_fp);
    }
    // Samples of one block for this lane: row y = lane%16, columns 4*(lane/16)..+3 of
    // every slice, with the mirror padding applied in the address.
    auto fetch = [&](int blk, float (&dst)[T][4]) {
        const int fy = blk / nbx;
        const int fx = blk - fy * nbx;
        const long long roff =
            static_cast<long long>(mirror((by0 + fy) * S + l16 - job.oy, job.height)) * job.src_rs;
        int cols[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            cols[s] = mirror((bx0 + fx) * S + 4 * lg + s - job.ox, job.width);
        }
#pragma unroll
        for (int t = 0; t < T; t++) {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                dst[t][s] = static_cast<float>(frame[t][roff + cols[s]]);
            }
        }
    };

    // Middle-stage ownership: lane = 8*p + q owns y pair p (rows m = p, p+8 of the real
    // basis) and x pair q (columns n = q, q+8).  Pair 0 is (cos 0, cos 8) on both axes,
    // two purely real-axis coefficients; pairs 1..7 are (cos k, sin k).
    const int mp = lane >> 3;
    const int mq = lane & 7;
    const bool ey = mp != 0;
    const bool ex = mq != 0;
    const int wq = l16 & 7;
    const int wc = l16 >> 3;

    // The next block's samples are fetched one iteration ahead, so their load latency
    // overlaps the current block's transforms.
    float nxt[T][4];
    if (wave < nblk) {
        fetch(wave, nxt);
    }

    __syncthreads();

    for (int b = wave; b < nblk; b += kWaves) {
        const int bly = b / nbx;
        const int blx = b - bly * nbx;
        float cur[T][4];
#pragma unroll
        for (int t = 0; t < T; t++) {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                cur[t][s] = nxt[t][s];
            }
        }
        if (b + kWaves < nblk) {
            fetch(b + kWaves, nxt);
        }

        // ---- forward: D2_t = WY * (slice_t * WX) ----
        floatx4 d2[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            const Opnd in = make_opnd<PREC>(win[t][0] * cur[t][0], win[t][1] * cur[t][1], win[t][2] * cur[t][2],
                                            win[t][3] * cur[t][3]);
            const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
            const floatx4 d1 = mm16<PREC>(in, WXo, zero);
            d2[t] = mm16<PREC>(WYo, make_opnd<PREC>(d1[0], d1[1], d1[2], d1[3]), zero);
        }

        // ---- temporal real basis, staged as [j][p][q][b*2+c] ----
#pragma unroll
        for (int j = 0; j < T; j++) {
            floatx4 u = d2[0] * wt[j][0];
#pragma unroll
            for (int t = 1; t < T; t++) {
                u += d2[t] * wt[j][t];
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
                filter_coef<FTM>(ar, ai, tab[(g * 4 + 2 * e) * 64 + lane], gf, a);
                filter_coef<FTM>(br, bi, tab[(g * 4 + 2 * e + 1) * 64 + lane], gf, a);
                zr[e][0] = ex ? 0.5f * (ar + br) : ar;
                zi[e][0] = ex ? 0.5f * (ai - bi) : ai;
                zr[e][1] = ex ? -0.5f * (ai + bi) : br;
                zi[e][1] = ex ? 0.5f * (ar - br) : bi;
            }

            // Back to the real (t, y) basis, then the centre-slice inverse weights.
            const float wc_ = ct[g];
            const float ws_ = g ? ct[R + g] : 0.0f;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float pr = zr[0][c], pi = zi[0][c], qr = zr[1][c], qi = zi[1][c];
                const float u_c0 = ey ? 0.5f * (pr + qr) : pr;   // cos t, y slot 0
                const float u_s1 = ey ? 0.5f * (qr - pr) : -qi;  // sin t, y slot 1
                const float u_s0 = ey ? -0.5f * (pi + qi) : -pi; // sin t, y slot 0
                const float u_c1 = ey ? 0.5f * (qi - pi) : qr;   // cos t, y slot 1
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
            atomicAdd(arow + i * rw, o[i] * syn[i]);
        }
        wave_lds_sync();
    }
    __syncthreads();

    float* out = a.scratch + item * static_cast<long long>(a.region_h * rw);
    for (int i = tid; i < rows_used * rw; i += kThreads) {
        out[i] = acc[i];
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

// Sums the tile regions covering each output pixel (at most 2 per axis) in a fixed
// order and converts to the sample type.
template <typename Tout>
__global__ __launch_bounds__(256) void finalize_kernel(LaunchArgs a) {
    const long long jf = blockIdx.y;
    int jb = 0;
    while (jb + 1 < a.n_jobs && a.jobs[jb + 1].frame_begin <= jf) {
        ++jb;
    }
    const JobDev& job = a.jobs[jb];
    const int f_local = static_cast<int>(jf - job.frame_begin);
    const int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= job.width * job.height) {
        return;
    }
    const int y = pix / job.width;
    const int x = pix - y * job.width;
    const int py = y + job.oy;
    const int px = x + job.ox;
    const int S = a.block_step;
    const int th = S * a.tile_bh;
    const int tw = S * a.tile_bw;
    const int tiles = job.tiles_x * job.tiles_y;
    const long long region = static_cast<long long>(a.region_h) * a.region_w;
    const long long base_item = job.item_begin + static_cast<long long>(f_local) * tiles;

    int ty_list[2], tx_list[2];
    int nty = 0, ntx = 0;
    const int ty1 = min(py / th, job.tiles_y - 1);
    for (int ty = max(ty1 - 1, 0); ty <= ty1; ty++) {
        const int nb = min(a.tile_bh, job.nby - ty * a.tile_bh);
        if (py >= ty * th && py < ty * th + S * (nb - 1) + 16) {
            ty_list[nty++] = ty;
        }
    }
    const int tx1 = min(px / tw, job.tiles_x - 1);
    for (int tx = max(tx1 - 1, 0); tx <= tx1; tx++) {
        const int nb = min(a.tile_bw, job.nbx - tx * a.tile_bw);
        if (px >= tx * tw && px < tx * tw + S * (nb - 1) + 16) {
            tx_list[ntx++] = tx;
        }
    }
    float sum = 0.0f;
    for (int i = 0; i < nty; i++) {
        for (int k = 0; k < ntx; k++) {
            const int ty = ty_list[i], tx = tx_list[k];
            const long long it = base_item + ty * job.tiles_x + tx;
            sum += a.scratch[it * region + static_cast<long long>(py - ty * th) * a.region_w + (px - tx * tw)];
        }
    }
    Tout* dst = reinterpret_cast<Tout*>(job.dst) + static_cast<long long>(job.f_begin + f_local) * job.dst_fp +
                static_cast<long long>(y) * job.dst_rs + x;
    store_sample(dst, sum, a);
}

template <int R, typename Tin, int FTM, int PREC>
hipError_t launch_one(const LaunchArgs& a, long long n_items, size_t lds, hipStream_t stream) {
    auto kern = fused_kernel<R, Tin, FTM, PREC>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) {
            return e;
        }
    }
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(n_items)), dim3(kThreads), lds, stream, a);
    return hipGetLastError();
}

template <int R, typename Tin, int PREC>
hipError_t launch_ft(const LaunchArgs& a, long long n_items, size_t lds, hipStream_t stream) {
    if (a.filter_type == 0) {
        return launch_one<R, Tin, 0, PREC>(a, n_items, lds, stream);
    }
    if (a.filter_type == 1) {
        return launch_one<R, Tin, 1, PREC>(a, n_items, lds, stream);
    }
    return launch_one<R, Tin, 2, PREC>(a, n_items, lds, stream);
}

template <int R, typename Tin>
hipError_t launch_prec(const LaunchArgs& a, int precision, long long n_items, size_t lds, hipStream_t stream) {
    if (precision == 1) {
        return launch_ft<R, Tin, 1>(a, n_items, lds, stream);
    }
    return launch_ft<R, Tin, 0>(a, n_items, lds, stream);
}

template <int R>
hipError_t launch_type(const LaunchArgs& a, int sample_code, int precision, long long n_items, size_t lds,
                       hipStream_t stream) {
    switch (sample_code) {
    case kU8:
        return launch_prec<R, uint8_t>(a, precision, n_items, lds, stream);
    case kU16:
        return launch_prec<R, uint16_t>(a, precision, n_items, lds, stream);
    default:
        return launch_prec<R, float>(a, precision, n_items, lds, stream);
    }
}

} // namespace

size_t fused_lds_bytes(int radius, int region_w, int region_h) {
    const int t_len = 2 * radius + 1;
    const int wbuf = (t_len * 256 > 320) ? t_len * 256 : 320;
    return sizeof(float) * (static_cast<size_t>(region_w) * region_h + (radius + 1) * 4 * 64 * 4 + kWaves * wbuf);
}

hipError_t launch_fused(const LaunchArgs& a, int radius, int sample_code, int precision, long long n_items,
                        size_t lds_bytes, hipStream_t stream) {
    switch (radius) {
    case 0:
        return launch_type<0>(a, sample_code, precision, n_items, lds_bytes, stream);
    case 1:
        return launch_type<1>(a, sample_code, precision, n_items, lds_bytes, stream);
    case 2:
        return launch_type<2>(a, sample_code, precision, n_items, lds_bytes, stream);
    default:
        return launch_type<3>(a, sample_code, precision, n_items, lds_bytes, stream);
    }
}

hipError_t launch_finalize(const LaunchArgs& a, int sample_code, long long n_job_frames, int max_pixels,
                           hipStream_t stream) {
    dim3 grid(static_cast<unsigned>((max_pixels + 255) / 256), static_cast<unsigned>(n_job_frames));
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

// Backward pass of the differentiable (Wiener) filter and the per-launch gain table.

#include "dft2_device.h"

namespace dft2 {
namespace {

// The gain table of one launch with sigma taken from a.sigma_dev, built once into global
// memory (every workgroup then copies it instead of gathering sigma itself).
template <int R, bool kBias>
__global__ __launch_bounds__(256) void gain_table_kernel(LaunchArgs a, float4* out) {
    a.tab_dev = nullptr;
    fill_tab<R, kBias>(out, a, blockIdx.x * 256 + threadIdx.x, gridDim.x * 256);
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

// Sum of v over the wave, valid in lane 0: DPP adds within each row of 16 lanes (quad
// swaps, then rotations by 4 and 8), then the four row sums read into lane 0.
template <int kCtrl>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), kCtrl, 0xf, 0xf, false));
}

__device__ __forceinline__ float wave_sum_lane0(float v) {
    v = dpp_add<0xb1>(v);  // quad_perm (1, 0, 3, 2)
    v = dpp_add<0x4e>(v);  // quad_perm (2, 3, 0, 1)
    v = dpp_add<0x124>(v); // row_ror 4
    v = dpp_add<0x128>(v); // row_ror 8
    const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
    const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
    const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
    return (v + r1) + (r2 + r3);
}

// A constant close to a block's samples (the wave's mean of the four samples each lane
// holds of one slice), subtracted from every sample of the block before the forward
// transform when zero-mean is on.  Zero-mean removes any such constant exactly: with
// samples s - c the DC is DC - c W0 and gf is gf - c, so x = A - gf W is unchanged.  Left
// in, the block's mean dominates the low-frequency coefficients (about the mean times the
// window spectrum), x is a small remainder of them, and the split-BF16 contractions'
// error relative to the coefficients becomes a large one relative to x.
__device__ __forceinline__ float block_offset(const raw4<float>& v) {
    const float sum = (raw4_value<float>(v, 0) + raw4_value<float>(v, 1)) + (raw4_value<float>(v, 2) + raw4_value<float>(v, 3));
    const float total = wave_sum_lane0(sum);
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, total))) * (1.0f / 256.0f);
}

// Adjoint of one gain application: x = A - gf*W, y = g x + gf*W, where the gain is the
// average of the Wiener gains of the coefficient's own sigma and its partner's,
// g = (g1 + g2)/2, g_i = max(1 - c_i, 0), c_i = se_i/(psd + eps), se_i = sigma_i + eps
// (table entries e.x, e.y), psd = |x|^2.  Where g_i > 0, dg_i/dpsd = c_i/(psd + eps) and
// dg_i/dsigma_i = -1/(psd + eps); both are 0 elsewhere.
// (dyr, dyi) in: HALF the gradient w.r.t. y (the factor 1/2 of the average is folded into
// the caller's inverse weights); out: the full gradient w.r.t. A.  Accumulates d/dsigma
// (own, partner) and d/dgf.
__device__ __forceinline__ void gain_adjoint(float ar, float ai, float& dyr, float& dyi, float4 e, float gf,
                                             float2& dsig, float& dgf) {
    const float xr = ar - gf * e.z;
    const float xi = ai - gf * e.w;
    const float r = __builtin_amdgcn_rcpf(fmaf(xr, xr, fmaf(xi, xi, 1e-15f))); // 1 / (psd + eps)
    const float c1 = e.x * r;
    const float c2 = e.y * r;
    const float m1 = c1 < 1.0f ? 1.0f : 0.0f; // g1 > 0
    const float m2 = c2 < 1.0f ? 1.0f : 0.0f;
    const float n = m1 + m2;
    const float q = fmaf(m1, c1, m2 * c2); // 2 (psd + eps) dg/dpsd; 2 g = n - q
    const float sdot = dyr * xr + dyi * xi;
    const float rs = r * sdot;
    // dx = g dy + 2 (dg/dpsd) (x . dy) x with dy = 2 (dyr, dyi), i.e.
    // (n - q) dy + 2 q rs x = n dy + q (2 rs x - dy)
    const float rs2 = rs + rs;
    const float dxr = fmaf(q, fmaf(rs2, xr, -dyr), n * dyr);
    const float dxi = fmaf(q, fmaf(rs2, xi, -dyi), n * dyi);
    dsig.x = fmaf(-m1, rs, dsig.x);
    dsig.y = fmaf(-m2, rs, dsig.y);
    dgf += (2.0f * dyr - dxr) * e.z + (2.0f * dyi - dxi) * e.w;
    dyr = dxr;
    dyi = dxi;
}

// Adjoint middle stage of the wave's block, one lane per (y pair, x pair) group (the
// ownership of the forward's middle stage).  In: the block's temporal spectra U staged in
// wbuf as [j][p][q][b*2+c] and the gradient of V in vbuf as [p][q][b*2+c].  Out: the
// gradient of U in wbuf, same layout.  ct[g]: half the centre-slice inverse weight of
// temporal pair g, times any factor the caller left in V's gradient.  Accumulates the
// sigma gradient of the lane's coefficients into dsg.
template <int R>
__device__ __forceinline__ void adjoint_middle(float* wbuf, const float* vbuf, const float4* tab, int lane,
                                               const float (&ct)[2 * R + 1], float2 (&dsg)[R + 1][4],
                                               const LaunchArgs& a) {
    constexpr int T = 2 * R + 1;
    const int mp = lane >> 3;
    const int mq = lane & 7;
    const bool ey = mp != 0;
    const bool ex = mq != 0;
    // 1/0 factors: "c ? a + b : a" is fma(cf, b, a) (one instruction, same value)
    const float eyf = ey ? 1.0f : 0.0f;
    const float exf = ex ? 1.0f : 0.0f;
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
            z1r[c] = fmaf(-eyf, ss1[c], cc0[c]);
            z1i[c] = -fmaf(eyf, cc1[c], ss0[c]);
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
            ar[e] = fmaf(exf, zxi, z0r);
            ai[e] = fmaf(-exf, zxr, z0i);
            br[e] = ex ? z0r - zxi : zxr;
            bi[e] = ex ? -z0i - zxr : zxi;
        }
        // dV -> d(real t,y basis) -> dP, dQ per x slot; at the centre slice only the
        // cos-t components carry weight (the sin-t rows vanish there)
        const float wcg = ct[g];
        const float dv0[2] = {dV4.x, dV4.y}, dv1[2] = {dV4.z, dV4.w};
        float dzr[2][2], dzi[2][2]; // [e][c] gradients of the gained Z'
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const float duc0 = wcg * dv0[c];
            const float duc1 = wcg * dv1[c];
            dzr[0][c] = duc0;
            dzr[1][c] = ey ? duc0 : duc1;
            dzi[0][c] = ey ? -duc1 : 0.0f;
            dzi[1][c] = ey ? duc1 : 0.0f;
        }
        float gz1r[2], gz1i[2], gz2r[2], gz2i[2]; // gradients of Z (pre gain), [c]
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const float d0r = dzr[e][0], d0i = dzi[e][0], d1r = dzr[e][1], d1i = dzi[e][1];
            float dar = fmaf(exf, d1i, d0r);
            float dai = fmaf(-exf, d1r, d0i);
            float dbr = ex ? d0r - d1i : d1r;
            float dbi = ex ? -d0i - d1r : d1i;
            gain_adjoint(ar[e], ai[e], dar, dai, tab[(g * 4 + 2 * e) * 64 + lane], gf, dsg[g][2 * e], dgf);
            gain_adjoint(br[e], bi[e], dbr, dbi, tab[(g * 4 + 2 * e + 1) * 64 + lane], gf, dsg[g][2 * e + 1], dgf);
            const float g0r = fmaf(exf, dbr, dar);
            const float g0i = fmaf(-exf, dbi, dai);
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
            dc0[c] = fmaf(eyf, gz2r[c], gz1r[c]);
            ds1[c] = ey ? gz2r[c] - gz1r[c] : -gz2i[c];
            ds0[c] = -fmaf(eyf, gz2i[c], gz1i[c]);
            dc1[c] = ey ? gz2i[c] - gz1i[c] : gz2r[c];
        }
        dU4[g] = make_float4(dc0[0], dc0[1], dc1[0], dc1[1]);
        if (g) {
            dU4[R + g] = make_float4(ds0[0], ds0[1], ds1[0], ds1[1]);
        }
    }
    dgf = wave_sum_lane0(dgf);
    if (a.zero_mean && lane == 0) {
        dU4[0].x += dgf * a.inv_wf0;
    }
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < T; j++) {
        *reinterpret_cast<float4*>(&wbuf[((j * 8 + mp) * 8 + mq) * 4]) = dU4[j];
    }
    wave_lds_sync();
}

// Adjoint of the centre-indexed temporal basis: du[j] (gradients of the T temporal rows,
// in the D2 register layout) -> ddt[t] (gradients of the T slice spectra).  Slice R + k
// gets C_k + S_k and slice R - k gets C_k - S_k (C: cos rows, even in k; S: sin rows, odd
// in k); the centre slice the cos rows.  wt: the forward's temporal basis; kRow0 = false
// when its DC row is 1 at every slice (the per-lane window form), true when it carries
// the slice factors (the separable form).
template <int R, bool kRow0>
__device__ __forceinline__ void adjoint_temporal(const floatx4 (&du)[2 * R + 1], const float (&wt)[2 * R + 1][2 * R + 1],
                                                 floatx4 (&ddt)[2 * R + 1]) {
    ddt[R] = du[0];
#pragma unroll
    for (int j = 1; j <= R; j++) {
        ddt[R] += du[j];
    }
    if constexpr (R > 0) {
#pragma unroll
        for (int k = 1; k <= R; k++) {
            floatx4 c = kRow0 ? du[0] * wt[0][k] : du[0];
#pragma unroll
            for (int j = 1; j <= R; j++) {
                c += du[j] * wt[j][k];
            }
            floatx4 sn = du[R + 1] * wt[R + 1][k];
#pragma unroll
            for (int i = 2; i <= R; i++) {
                sn += du[R + i] * wt[R + i][k];
            }
            ddt[R + k] = c + sn;
            ddt[R - k] = c - sn;
        }
    }
}

// Sigma gradient of the workgroup: lanes -> LDS (by coefficient; dsig_lds, T*144 floats,
// zeroed by the caller before its last barrier) -> one global atomic per entry.
template <int R>
__device__ __forceinline__ void sigma_grad_out(float* dsig_lds, const float2 (&dsg)[R + 1][4], int lane, int tid,
                                               const LaunchArgs& a) {
    constexpr int T = 2 * R + 1;
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

template <int R>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(R <= 1 ? 4 : 1))) void backward_kernel(LaunchArgs a) {
    constexpr int T = 2 * R + 1;
    // The recomputed slice spectra on the FP32 MFMA, as the FP32 forward computes them
    // (so every gain, and the side of the Wiener gain's kink a coefficient lies on, is the
    // forward's); the linear adjoint contractions in split BF16 (hi + lo operands, three
    // products, FP32 accumulation: about 16 significant bits per operand).
    constexpr int kBP = 1;
    constexpr int kWbuf = T * 256 + 256;
    extern __shared__ __attribute__((aligned(16))) float smem[]; // float4 accesses

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = sgpr(tid >> 6);
    const int l16 = lane & 15;
    const int lg = lane >> 4;

    const JobDev& job = a.jobs[sgpr(find_job<0>(a.jobs, a.n_jobs, blockIdx.x))];
    const int tiles_x = sgpr(job.tiles_x);
    const int tiles = tiles_x * sgpr(job.tiles_y);
    const int job_begin = sgpr(static_cast<int>(job.item_begin));
    // workgroup -> work item of the same job, each XCD a contiguous part of the job
    const int item = job_xcd_item(blockIdx.x, job_begin, sgpr(job.f_count) * tiles);
    const int local = item - job_begin;
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
    fill_tab<R, true>(tab, a, tid, kThreads);
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
        // halved: gain_adjoint takes half the gradient of the gained coefficients
        ct[j] = 0.5f * a.consts[kOffCT + j];
#pragma unroll
        for (int t = 0; t < T; t++) {
            wt[j][t] = a.consts[kOffWT + j * T + t];
        }
    }

    __amdgpu_buffer_rsrc_t frame[T];
    float* gin[T];
    // wide_ok: plane bases and row pitches allow one dwordx4 per lane and slice for blocks
    // whose first sample is 16-byte aligned too.
    bool wide_ok = (rs % 4) == 0 && (go_rs % 4) == 0;
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
            const float* base = src + static_cast<long long>(fi) * fp;
            frame[t] = frame_rsrc(base, frame_bytes);
            gin[t] = gi + static_cast<long long>(fi) * gfp;
            wide_ok = wide_ok && (reinterpret_cast<unsigned long long>(base) % 16) == 0;
        }
    }
    const float* go_base =
        sgpr_ptr(reinterpret_cast<const float*>(job.grad_out)) + static_cast<long long>(n) * sgpr64(job.go_fp);
    const __amdgpu_buffer_rsrc_t gout = frame_rsrc(go_base, static_cast<unsigned>(go_rs * height) * 4u);
    wide_ok = wide_ok && (reinterpret_cast<unsigned long long>(go_base) % 16) == 0;
    const unsigned lane_off = static_cast<unsigned>(l16 * rs + 4 * lg) * 4u;
    const unsigned lane_off_go = static_cast<unsigned>(l16 * go_rs + 4 * lg) * 4u;

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

    // Samples of the T slices and of the output gradient for block (fy, fx) of the tile,
    // as raw bits, fetched one block ahead (once the current block's samples are consumed)
    // so that their latency overlaps the current block's work; aligned interior blocks take
    // one dwordx4 per slice.  Output-gradient samples outside the plane count as 0: their
    // validity is kept as a 4-bit mask (bit s for column 4*(lane/16) + s) and applied at use.
    auto fetch = [&](int fy, int fx, raw4<float> (&dst)[T], raw4<float>& dgo, unsigned& okm, bool& full) {
        fx = opaque_uniform(fx); // a wave's column is fixed: keep its offsets out of the loop's registers
        const int y0 = (by0 + fy) * S - oy;
        const int x0 = (bx0 + fx) * S - ox;
        if (y0 >= 0 && y0 + 15 < height && x0 >= 0 && x0 + 15 < width) {
            const unsigned off = static_cast<unsigned>(y0 * rs + x0) * 4u + lane_off;
            const unsigned og = static_cast<unsigned>(y0 * go_rs + x0) * 4u + lane_off_go;
            if (wide_ok && ((y0 * rs + x0) % 4) == 0 && ((y0 * go_rs + x0) % 4) == 0) {
#pragma unroll
                for (int t = 0; t < T; t++) {
                    dst[t] = load4_aligned<float>(frame[t], off);
                }
                dgo = load4_aligned<float>(gout, og);
            } else {
                const unsigned offs[4] = {off, off + 4u, off + 8u, off + 12u};
                const unsigned ogs[4] = {og, og + 4u, og + 8u, og + 12u};
#pragma unroll
                for (int t = 0; t < T; t++) {
                    dst[t] = load4_each<float>(frame[t], offs);
                }
                dgo = load4_each<float>(gout, ogs);
            }
            okm = 0xfu;
            full = true;
        } else {
            full = false;
            const unsigned roff = static_cast<unsigned>(mirror(y0 + l16, height) * rs);
            unsigned offs[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                offs[s] = (roff + static_cast<unsigned>(mirror(x0 + 4 * lg + s, width))) * 4u;
            }
#pragma unroll
            for (int t = 0; t < T; t++) {
                dst[t] = load4_each<float>(frame[t], offs);
            }
            const int gy = y0 + l16;
            const bool row_ok = gy >= 0 && gy < height;
            okm = 0u;
            unsigned ogs[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int gx = x0 + 4 * lg + s;
                const bool ok = row_ok && gx >= 0 && gx < width;
                ogs[s] = ok ? static_cast<unsigned>(gy * go_rs + gx) * 4u : 0u;
                okm |= ok ? (1u << s) : 0u;
            }
            dgo = load4_each<float>(gout, ogs);
        }
    };
    // Each block row r of the tile is done in passes of kWaves blocks: in pass p wave w
    // takes block (r, p*kWaves + w).  The wave's next block is fetched one block ahead.
    const int n_pass = (nbx + kWaves - 1) / kWaves;
    const int my_pass = nbx > wave ? (nbx - wave + kWaves - 1) / kWaves : 0; // passes with a block of this wave
    raw4<float> nxt[T], nxt_go;
    unsigned nxt_ok = 0u;
    bool nxt_full = false;
    if (my_pass > 0) {
        fetch(0, wave, nxt, nxt_go, nxt_ok, nxt_full);
    }

    __syncthreads();

    for (int r = 0; r < nby; r++) {
    for (int pass = 0; pass < n_pass; pass++) {
    if (pass < my_pass) {
        const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
        // ---- recompute the forward spectrum, staged as [j][p][q][b*2+c] ----
        const float boff = a.zero_mean ? block_offset(nxt[R]) : 0.0f;
        floatx4 d2[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            float wv[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                wv[s] = win[t][s] * (raw4_value<float>(nxt[t], s) - boff);
            }
            // WXo and WYo carry their FP32 values beside the split halves
            const floatx4 d1 = mm16<0>(make_opnd<0>(wv[0], wv[1], wv[2], wv[3]), WXo, zero);
            d2[t] = mm16<0>(WYo, make_opnd<0>(d1[0], d1[1], d1[2], d1[3]), zero);
        }
        {
            // centre-indexed temporal basis from the sums / differences of slices R + k
            // and R - k (as in the forward kernel)
            floatx4 tsum[R + 1], tdif[R + 1];
#pragma unroll
            for (int k = 1; k <= R; k++) {
                tsum[k] = d2[R + k] + d2[R - k];
                tdif[k] = d2[R + k] - d2[R - k];
            }
#pragma unroll
            for (int j = 0; j < T; j++) {
                floatx4 u = d2[R];
                if (j == 0) {
#pragma unroll
                    for (int k = 1; k <= R; k++) {
                        u += tsum[k];
                    }
                } else if (j <= R) {
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
        }

        // ---- output gradient -> dV = IY^T (dO IX^T), dO = dy / scale * synthesis window ----
        {
            // samples outside the plane count as 0 (border blocks only: a uniform branch)
            float dov[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                dov[s] = raw4_value<float>(nxt_go, s) * synA[s];
            }
            if (!nxt_full) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    dov[s] = ((nxt_ok >> s) & 1u) ? dov[s] : 0.0f;
                }
            }
            // The samples of this block are consumed: fetch the wave's next block into the
            // same registers; the loads have the rest of this block's work to complete.
            const bool row_end = pass + 1 == my_pass;
            if (!row_end || r + 1 < nby) {
                fetch(row_end ? r + 1 : r, (row_end ? 0 : pass + 1) * kWaves + wave, nxt, nxt_go, nxt_ok, nxt_full);
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
        adjoint_middle<R>(wbuf, vbuf, tab, lane, ct, dsg, a);

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
        floatx4 ddt[T];
        adjoint_temporal<R, false>(du, wt, ddt);
        wave_lds_sync(); // du read before the staged outputs overwrite the buffer
#pragma unroll
        for (int t = 0; t < T; t++) {
            const floatx4 dd = ddt[t];
            // dd (m in registers) used as A operand is dD2^T; R1 = dD2^T WY, db^T = WX R1.
            const floatx4 r1 = mm16<kBP>(make_opnd<kBP>(dd[0], dd[1], dd[2], dd[3]), WYTo, zero);
            const floatx4 dbt = mm16<kBP>(WXTo, make_opnd<kBP>(r1[0], r1[1], r1[2], r1[3]), zero);
            // slice t's block gradient (row l16, columns 4*(lane/16)..+3) staged as [y][x]
            // at offset t*256 + wave*S of the wave's buffer
            float* srow = wbuf + t * 256 + wave * S + l16 * 16 + 4 * lg;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                srow[i] = win[t][i] * dbt[i];
            }
        }
    }
        // ---- overlap-add of pass `pass` of block row r into the T gradient regions ----
        // As in the forward kernel: each pixel of the pass's 16-row band sums the staged
        // values of the pass's blocks covering it (block k's value for local column x at
        // k*kWbuf + t*256 + y*16 + x) and its owning thread adds the sum, no atomics.
        __syncthreads();
        {
            const int Sg = opaque_uniform(S); // per-lane bounds stay in the loop
            const int Sm = opaque_uniform(a.step_magic); // for div_small
            const int nb = min(kWaves, nbx - pass * kWaves);
            const int cols = Sg * (nb - 1) + 16;
            for (int x0 = 0; x0 < cols; x0 += 64) {
                const int x = x0 + lane;
                if (x >= cols) {
                    continue;
                }
                // column x is covered by blocks ceil((x - 15) / S) .. floor(x / S); the
                // bounds and masks serve the wave's rows y = wave, wave + 8 of every slice
                const int lo = div_small(max(x + Sg - 16, 0), Sm);
                const int hi = min(div_small(x, Sm), nb - 1);
                // All reads first (one LDS round trip), then the sums and the writes.
                constexpr int kRows = 16 / kWaves;
                const int k0 = max(hi - 3, 0);
                float old[kRows][T], sum[kRows][T];
#pragma unroll
                for (int i = 0; i < kRows; i++) {
                    const int y = i * kWaves + wave;
                    const int g = (r * Sg + y) * rw + pass * kWaves * Sg + x;
#pragma unroll
                    for (int t = 0; t < T; t++) {
                        old[i][t] = gacc[t * region + g];
                        sum[i][t] = gather_head<kWbuf>(stage + t * 256 + y * 16 + x, k0, lo, hi);
                    }
                }
#pragma unroll
                for (int i = 0; i < kRows; i++) {
                    const int y = i * kWaves + wave;
                    const int g = (r * Sg + y) * rw + pass * kWaves * Sg + x;
#pragma unroll
                    for (int t = 0; t < T; t++) {
                        sum[i][t] = gather_tail<kWbuf>(stage + t * 256 + y * 16 + x, k0, lo, sum[i][t]);
                        gacc[t * region + g] = old[i][t] + sum[i][t];
                    }
                }
            }
        }
        __syncthreads(); // staged values consumed before the next pass reuses the buffers
    }
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

    sigma_grad_out<R>(dsig_lds, dsg, lane, tid, a);
}

// ---------------------------------------------------------------------------------------
// Cross-frame backward (separable window, radius >= 1): a work item is a tile over K =
// a.reuse_k consecutive output frames n0 .. n0+K-1 (fewer at the end of a job's range).
// Output frame n0+k reads the source slices n0+k-R .. n0+k+R, so the item touches the
// K + 2R "virtual" input frames v = 0 .. K+2R-1 (source frame n0-R+v, clamped to the clip
// where it is read or written).  Per block position each wave:
//   - slides the T slice spectra D2 from one output frame to the next, transforming one
//     new slice per frame (the separable window makes a slice's D2 independent of its
//     position in the window: the slice factors are in the temporal basis);
//   - per output frame, takes the output gradient through the adjoint inverse DFT, the
//     adjoint middle stage and the adjoint temporal basis to the gradients of the T slice
//     spectra, and adds them to a register ring of the pending spectrum gradients of
//     virtual frames k .. k+2R;
//   - virtual frame k is then complete (later frames of the item do not read it): one
//     adjoint spatial transform takes its summed spectrum gradient to the block's sample
//     gradient (the adjoint of D2 = AS[R] SWY s SWX is linear and the same for every
//     slice), which is staged for the overlap-add into v's accumulator; the item's last
//     frame completes the 2R + 1 remaining virtual frames.
// So a block position costs K + 2R slice transforms each way for K output frames, where
// the one-frame kernel costs 3K (T = 3) of each, and K + 2R overlap-add gathers instead of
// TK.  The input-gradient accumulator of each virtual frame is a ring of ring_bands bands
// of S rows; band r is final once block row r is done and is then added to the input
// gradient with global atomics (mirror padding, temporal clamp) and cleared.
// Basis operands of the cross-frame backward, kept in LDS: [b][lane] as a lane's four
// FP32 values (PREC 0) or the packed hi and lo halves of its split-BF16 Opnd (PREC 1).
// WX, WY: the spatial forward in the recompute's precision; GX, GY: the same factors in
// the output-gradient contraction's; WYT, WXT: the adjoint's.
constexpr int kBaseWX = 0, kBaseWY = 1, kBaseWYT = 2, kBaseWXT = 3, kBaseGX = 4, kBaseGY = 5, kBases = 6;

template <int PREC>
__device__ __forceinline__ u32x4 opnd_bits(const Opnd& o) {
    if constexpr (PREC == 0) {
        return u32x4{__builtin_bit_cast(unsigned, o.f[0]), __builtin_bit_cast(unsigned, o.f[1]),
                     __builtin_bit_cast(unsigned, o.f[2]), __builtin_bit_cast(unsigned, o.f[3])};
    } else {
        const u32x2 h = __builtin_bit_cast(u32x2, o.hi);
        const u32x2 l = __builtin_bit_cast(u32x2, o.lo);
        return u32x4{h.x, h.y, l.x, l.y};
    }
}

template <int PREC>
__device__ __forceinline__ Opnd opnd_from_bits(u32x4 w) {
    Opnd o;
    if constexpr (PREC == 0) {
        // the whole vector cast at once: this compiler's __builtin_bit_cast of one element
        // of an ext_vector (w.y, w[1]) reads element 0
        const floatx4 f = __builtin_bit_cast(floatx4, w);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            o.f[s] = f[s];
        }
    } else {
        o.hi = __builtin_bit_cast(bf16x4, u32x2{w.x, w.y});
        o.lo = __builtin_bit_cast(bf16x4, u32x2{w.z, w.w});
    }
    return o;
}

template <int R>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(R <= 1 ? 4 : 1)))
void backward_reuse_kernel(LaunchArgs a) {
    static_assert(R >= 1, "the cross-frame backward needs radius >= 1");
    constexpr int T = 2 * R + 1;
    // Contraction precision: the recomputed slice spectra on the FP32 MFMA, as the FP32
    // forward computes them, so that every gain (and which side of the Wiener gain's kink
    // at psd = sigma a coefficient lies on) is the forward's; the two linear adjoint
    // contractions (output gradient -> dV, spectrum gradient -> input gradient) in split
    // BF16.  Measured on 8 clips of 256x256x16: 1.07 ms per dispatch, input gradient
    // within 5e-6 of FP64 (norm-wise); split BF16 throughout: 1.03 ms, 3.7e-4 (kink
    // events); FP32 throughout: 1.14 ms, 1.1e-6.
    constexpr int kFP = 0, kDP = 1, kAP = 1;
    constexpr int kWbuf = T * 256 + 256;
    extern __shared__ __attribute__((aligned(16))) float smem[]; // float4 accesses

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = sgpr(tid >> 6);
    const int l16 = lane & 15;
    const int lg = lane >> 4;

    const JobDev& job = a.jobs[sgpr(find_job<2>(a.jobs, a.n_jobs, blockIdx.x))];
    const int K = sgpr(a.reuse_k);
    const int tiles_x = sgpr(job.tiles_x);
    const int tiles = tiles_x * sgpr(job.tiles_y);
    const int job_begin = sgpr(static_cast<int>(job.group_begin));
    // workgroup -> work item of the same job, each XCD a contiguous part of the job
    const int item = job_xcd_item(blockIdx.x, job_begin, (sgpr(job.f_count) + K - 1) / K * tiles);
    const int local = item - job_begin;
    const int group = local / tiles;
    const int tile = local - group * tiles;
    const int f_local = group * K;
    const int kcount = min(K, sgpr(job.f_count) - f_local); // output frames of this item
    const int nv = kcount + 2 * R;                          // its virtual input frames
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int n0 = sgpr(job.f_begin) + f_local;
    const int S = a.block_step;
    const int by0 = ty * a.tile_bh;
    const int bx0 = tx * a.tile_bw;
    const int nby = min(a.tile_bh, sgpr(job.nby) - by0);
    const int nbx = min(a.tile_bw, sgpr(job.nbx) - bx0);
    const int width = sgpr(job.width);
    const int height = sgpr(job.height);
    const int ox = sgpr(job.ox);
    const int oy = sgpr(job.oy);
    const int rs = sgpr(static_cast<int>(job.src_rs));
    const int go_rs = sgpr(static_cast<int>(job.go_rs));

    // LDS: K + 2R rings of nbands bands of S rows | gain table | per-wave staging.  After
    // the block loop the staging area holds the sigma gradient of the workgroup.
    const int rw = a.region_w;
    const int rows_used = S * (nby - 1) + 16;
    const int cols_used = S * (nbx - 1) + 16;
    const int nbands = sgpr(a.ring_bands);
    const int ring_floats = nbands * S * rw;
    float* rings = smem;
    float4* tab = reinterpret_cast<float4*>(smem + (K + 2 * R) * ring_floats);
    u32x4* bases = reinterpret_cast<u32x4*>(tab + (R + 1) * 4 * 64);
    float* stage = reinterpret_cast<float*>(bases + kBases * 64);
    float* wbuf = stage + wave * kWbuf;
    float* vbuf = wbuf + T * 256;
    float* dsig_lds = stage;
    static_assert(kWaves * kWbuf >= T * 144, "sigma gradient must fit the staging area");
    fill_tab<R, true>(tab, a, tid, kThreads);
    {
        float4* rings4 = reinterpret_cast<float4*>(rings);
        for (int i = tid; i < nv * ring_floats / 4; i += kThreads) {
            rings4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }

    // Operands of the separable form: the window's spatial factors are in SWX and SWY and
    // AS[R] (sample scale times the centre slice's temporal factor) in the x operands, as
    // in the forward; the other slices' temporal factors relative to it are in WTS.  Their
    // split-BF16 halves sit in LDS (bases[b][lane]), read where used: held in registers
    // they would take 16 of a lane's 128, and the loop would spill.
    const float as_c = a.consts[kOffAS + R];
    if (wave == 0) {
        float wx[4], wy[4], wyT[4], wxT[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            wx[s] = a.consts[kOffSWX + (4 * lg + s) * 16 + l16] * as_c;
            wy[s] = a.consts[kOffSWY + l16 * 16 + 4 * lg + s];
            wyT[s] = a.consts[kOffSWY + (4 * lg + s) * 16 + l16];
            wxT[s] = a.consts[kOffSWX + l16 * 16 + 4 * lg + s] * as_c;
        }
        bases[kBaseWX * 64 + lane] = opnd_bits<kFP>(make_opnd<kFP>(wx[0], wx[1], wx[2], wx[3]));
        bases[kBaseWY * 64 + lane] = opnd_bits<kFP>(make_opnd<kFP>(wy[0], wy[1], wy[2], wy[3]));
        bases[kBaseGX * 64 + lane] = opnd_bits<kDP>(make_opnd<kDP>(wx[0], wx[1], wx[2], wx[3]));
        bases[kBaseGY * 64 + lane] = opnd_bits<kDP>(make_opnd<kDP>(wy[0], wy[1], wy[2], wy[3]));
        bases[kBaseWYT * 64 + lane] = opnd_bits<kAP>(make_opnd<kAP>(wyT[0], wyT[1], wyT[2], wyT[3]));
        bases[kBaseWXT * 64 + lane] = opnd_bits<kAP>(make_opnd<kAP>(wxT[0], wxT[1], wxT[2], wxT[3]));
    }
    // The operand b of this lane, read from LDS at each use (the index is opaque to the
    // optimizer, which would otherwise hoist the reads out of the loops).
    auto basis = [&](auto prec, int b) {
        int i = b * 64 + lane;
        asm volatile("" : "+v"(i));
        return opnd_from_bits<decltype(prec)::value>(bases[i]);
    };
    using FwdP = std::integral_constant<int, kFP>;
    using DvP = std::integral_constant<int, kDP>;
    using AdjP = std::integral_constant<int, kAP>;
    float wt[T][T];
#pragma unroll
    for (int j = 0; j < T; j++) {
#pragma unroll
        for (int t = 0; t < T; t++) {
            wt[j][t] = a.consts[kOffWTS + j * T + t];
        }
    }
    // The forward's block output is c0 SWY^T V SWX^T (c0 = CTS[0]), divided by the sample
    // scale when rounded.  The output gradient goes through the forward operands, which
    // give P = AS[R] SWY dy SWX, so dV = c0 / (out_scale AS[R]) P: that factor, and the
    // half that gain_adjoint takes, are in ct.
    const float dv_scale = a.consts[kOffCTS] / (a.out_scale * as_c);
    float ct[T];
#pragma unroll
    for (int j = 0; j < T; j++) {
        ct[j] = 0.5f * a.consts[kOffCT + j] * dv_scale;
    }

    const int n_frames = sgpr(job.n_frames);
    const float* src = sgpr_ptr(reinterpret_cast<const float*>(job.src));
    const float* gob = sgpr_ptr(reinterpret_cast<const float*>(job.grad_out));
    float* gi = sgpr_ptr(reinterpret_cast<float*>(job.grad_in));
    const long long fp = sgpr64(job.src_fp);
    const long long go_fp = sgpr64(job.go_fp);
    const long long gi_fp = sgpr64(job.gi_fp);
    const long long gi_rs = sgpr64(job.gi_rs);
    const unsigned frame_bytes = static_cast<unsigned>(rs * height) * 4u;
    const unsigned go_bytes = static_cast<unsigned>(go_rs * height) * 4u;
    // Virtual frame v of the item: source frame n0 - R + v, clamped to the clip
    // (cpu_source/getframe_impl.cpp:294-298).
    auto vframe = [&](int v) { return min(max(n0 - R + v, 0), n_frames - 1); };
    auto src_rsrc = [&](int v) { return frame_rsrc(src + static_cast<long long>(vframe(v)) * fp, frame_bytes); };
    auto go_rsrc = [&](int k) { return frame_rsrc(gob + static_cast<long long>(n0 + k) * go_fp, go_bytes); };
    // wide_ok: plane bases, frame pitches and row pitches allow one dwordx4 per lane and
    // slice for blocks whose first sample is 16-byte aligned too.
    const bool wide_ok = (rs % 4) == 0 && (go_rs % 4) == 0 && (fp % 4) == 0 && (go_fp % 4) == 0 &&
                         (reinterpret_cast<unsigned long long>(src) % 16) == 0 &&
                         (reinterpret_cast<unsigned long long>(gob) % 16) == 0;
    const unsigned lane_off = static_cast<unsigned>(l16 * rs + 4 * lg) * 4u;
    const unsigned lane_off_go = static_cast<unsigned>(l16 * go_rs + 4 * lg) * 4u;

    // Samples of block (fy, fx) of the tile from the N source descriptors fr[] into dst[],
    // as raw bits (lane: row l16, columns 4*(lane/16)..+3).
    auto fetch_src = [&](int fy, int fx, const auto& fr, auto& dst) {
        constexpr int N = sizeof(fr) / sizeof(fr[0]);
        fx = opaque_uniform(fx); // a wave's column is fixed: keep its offsets out of the loop's registers
        const int y0 = (by0 + fy) * S - oy;
        const int x0 = (bx0 + fx) * S - ox;
        if (y0 >= 0 && y0 + 15 < height && x0 >= 0 && x0 + 15 < width) {
            const unsigned off = static_cast<unsigned>(y0 * rs + x0) * 4u + lane_off;
            if (wide_ok && ((y0 * rs + x0) % 4) == 0) {
#pragma unroll
                for (int t = 0; t < N; t++) {
                    dst[t] = load4_aligned<float>(fr[t], off);
                }
            } else {
                const unsigned offs[4] = {off, off + 4u, off + 8u, off + 12u};
#pragma unroll
                for (int t = 0; t < N; t++) {
                    dst[t] = load4_each<float>(fr[t], offs);
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
            for (int t = 0; t < N; t++) {
                dst[t] = load4_each<float>(fr[t], offs);
            }
        }
    };
    // Output gradient of block (fy, fx) of one output frame; samples outside the plane
    // count as 0, their validity kept as a 4-bit mask (bit s: column 4*(lane/16) + s)
    // where the block is not inside the plane (full = false, the same in every lane).
    auto fetch_go = [&](int fy, int fx, __amdgpu_buffer_rsrc_t gr, raw4<float>& dgo, unsigned& okm, bool& full) {
        fx = opaque_uniform(fx);
        const int y0 = (by0 + fy) * S - oy;
        const int x0 = (bx0 + fx) * S - ox;
        if (y0 >= 0 && y0 + 15 < height && x0 >= 0 && x0 + 15 < width) {
            const unsigned og = static_cast<unsigned>(y0 * go_rs + x0) * 4u + lane_off_go;
            if (wide_ok && ((y0 * go_rs + x0) % 4) == 0) {
                dgo = load4_aligned<float>(gr, og);
            } else {
                const unsigned ogs[4] = {og, og + 4u, og + 8u, og + 12u};
                dgo = load4_each<float>(gr, ogs);
            }
            okm = 0xfu;
            full = true;
        } else {
            full = false;
            const int gy = y0 + l16;
            const bool row_ok = gy >= 0 && gy < height;
            okm = 0u;
            unsigned ogs[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int gx = x0 + 4 * lg + s;
                const bool ok = row_ok && gx >= 0 && gx < width;
                ogs[s] = ok ? static_cast<unsigned>(gy * go_rs + gx) * 4u : 0u;
                okm |= ok ? (1u << s) : 0u;
            }
            dgo = load4_each<float>(gr, ogs);
        }
    };

    // Band b of the region (rows b*S .. b*S + S - 1, fewer past rows_used) of every
    // virtual frame: ring slot `slot` -> input gradient, then cleared for band b + nbands.
    // Element i = (ry, rx) of a band is this thread's for i = tid + j * kThreads, stepped
    // without a division; regions inside the plane skip the mirror.
    const int gy0 = by0 * S - oy;
    const int gx0 = bx0 * S - ox;
    const bool inside = gy0 >= 0 && gy0 + rows_used <= height && gx0 >= 0 && gx0 + cols_used <= width;
    const int fl_dy = kThreads / rw;
    const int fl_dx = kThreads - fl_dy * rw;
    const int fl_ry = tid / rw;
    const int fl_rx = tid - fl_ry * rw;
    auto flush_band = [&](int b, int slot) {
        const int nel = min(S, rows_used - b * S) * rw;
        for (int v = 0; v < nv; v++) {
            float* ring = rings + v * ring_floats + slot * S * rw;
            float* gv = gi + static_cast<long long>(vframe(v)) * gi_fp;
            int ry = fl_ry;
            int rx = fl_rx;
            for (int i = tid; i < nel; i += kThreads) {
                const float val = ring[i];
                ring[i] = 0.0f;
                if (rx < cols_used && val != 0.0f) {
                    int fy = gy0 + b * S + ry;
                    int fx = gx0 + rx;
                    if (!inside) {
                        fy = mirror(fy, height);
                        fx = mirror(fx, width);
                    }
                    atomicAdd(gv + static_cast<long long>(fy) * gi_rs + fx, val);
                }
                rx += fl_dx;
                ry += fl_dy;
                if (rx >= rw) {
                    rx -= rw;
                    ++ry;
                }
            }
        }
    };

    // Each block row r of the tile is done in passes of kWaves blocks (wave w: block
    // (r, p*kWaves + w)), and each pass over the item's K output frames.
    const int n_pass = (nbx + kWaves - 1) / kWaves;
    const int my_pass = nbx > wave ? (nbx - wave + kWaves - 1) / kWaves : 0; // passes with a block of this wave
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
    const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    raw4<float> nxt[T], nxt_go;
    unsigned nxt_ok = 0u;
    bool nxt_full = false;
    floatx4 d2[T];     // slice spectra of output frame k (virtual frames k .. k+2R)
    float boff = 0.0f; // sample offset of the block position (block_offset)
    floatx4 gp[2 * R]; // pending spectrum gradients of virtual frames k .. k+2R-1

    // The overlap-add of the pass's staged block gradients of N completed virtual frames
    // v0 .. v0+N-1 (slot s of a wave's buffer: frame v0 + s) into their rings, in a fixed
    // order and without atomics, as backward_kernel: every pixel of the pass's 16-row
    // band sums the staged values of the blocks covering it (block k's value for local
    // column x at k*kWbuf + s*256 + y*16 + x) and its owning thread adds the sum.
    auto gather = [&](auto n_tag, int v0, int pass, int r_band) {
        constexpr int N = decltype(n_tag)::value;
        const int Sg = opaque_uniform(S); // per-lane bounds stay in the loop
        const int Sm = opaque_uniform(a.step_magic); // for div_small
        const int nb = min(kWaves, nbx - pass * kWaves);
        const int cols = Sg * (nb - 1) + 16;
        constexpr int kRows = 16 / kWaves;
        // ring offset of the wave's rows dy = i * kWaves + wave of the band: band
        // r + dy / S, in slot (r + dy / S) mod nbands
        int roff[kRows];
#pragma unroll
        for (int i = 0; i < kRows; i++) {
            const int dy = i * kWaves + wave;
            const int q = div_small(dy, Sm);
            int band = r_band + q;
            band = band >= nbands ? band - nbands : band;
            roff[i] = v0 * ring_floats + (band * Sg + dy - q * Sg) * rw + pass * kWaves * Sg;
        }
        for (int x0 = 0; x0 < cols; x0 += 64) {
            const int x = x0 + lane;
            if (x >= cols) {
                continue;
            }
            // column x is covered by blocks ceil((x - 15) / S) .. floor(x / S)
            const int lo = div_small(max(x + Sg - 16, 0), Sm);
            const int hi = min(div_small(x, Sm), nb - 1);
            const int k0 = max(hi - 3, 0);
            float old[kRows][N], sum[kRows][N];
#pragma unroll
            for (int i = 0; i < kRows; i++) {
                const int y = i * kWaves + wave;
#pragma unroll
                for (int s = 0; s < N; s++) {
                    old[i][s] = rings[roff[i] + s * ring_floats + x];
                    sum[i][s] = gather_head<kWbuf>(stage + s * 256 + y * 16 + x, k0, lo, hi);
                }
            }
#pragma unroll
            for (int i = 0; i < kRows; i++) {
                const int y = i * kWaves + wave;
#pragma unroll
                for (int s = 0; s < N; s++) {
                    sum[i][s] = gather_tail<kWbuf>(stage + s * 256 + y * 16 + x, k0, lo, sum[i][s]);
                    rings[roff[i] + s * ring_floats + x] = old[i][s] + sum[i][s];
                }
            }
        }
    };

    __syncthreads();

    int r_band = 0; // r mod nbands, the ring slot of band r
    for (int r = 0; r < nby; r++) {
    for (int pass = 0; pass < n_pass; pass++) {
    for (int k = 0; k < kcount; k++) {
    const bool last = k + 1 == kcount; // frame k completes virtual frames k .. k+2R
    if (pass < my_pass) {
        const int bx = pass * kWaves + wave;
        if (k == 0) {
            // the block's T slices and the first output gradient, fetched when the block
            // starts (a prefetch would keep them live across all K frames)
            __amdgpu_buffer_rsrc_t fr[T];
#pragma unroll
            for (int t = 0; t < T; t++) {
                fr[t] = src_rsrc(t);
            }
            fetch_src(r, bx, fr, nxt);
            fetch_go(r, bx, go_rsrc(0), nxt_go, nxt_ok, nxt_full);
#pragma unroll
            for (int i = 0; i < 2 * R; i++) {
                gp[i] = zero;
            }
            // one offset for every slice of this block position (see block_offset)
            boff = a.zero_mean ? block_offset(nxt[R]) : 0.0f;
        }
        raw4<float> cur[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            cur[t] = nxt[t];
        }
        const raw4<float> cur_go = nxt_go;
        const unsigned cur_ok = nxt_ok;
        const bool cur_full = nxt_full;
        if (!last) {
            // the same block at the next output frame: its newest slice (virtual frame
            // k + 1 + 2R) and its output gradient, in flight during this frame's work
            const __amdgpu_buffer_rsrc_t fr1[1] = {src_rsrc(k + 1 + 2 * R)};
            raw4<float>(&newest)[1] = *reinterpret_cast<raw4<float>(*)[1]>(&nxt[T - 1]);
            fetch_src(r, bx, fr1, newest);
            fetch_go(r, bx, go_rsrc(k + 1), nxt_go, nxt_ok, nxt_full);
        }

        // ---- slice spectra D2 = WY (s WX): all T at the item's first frame, else the newest ----
        if (k > 0) {
#pragma unroll
            for (int t = 0; t + 1 < T; t++) {
                d2[t] = d2[t + 1];
            }
        }
#pragma unroll
        for (int t = 0; t < T; t++) {
            if (k > 0 && t + 1 < T) {
                continue;
            }
            float sv[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                sv[s] = raw4_value<float>(cur[t], s) - boff;
            }
            const floatx4 d1 = mm16<kFP>(make_opnd<kFP>(sv[0], sv[1], sv[2], sv[3]), basis(FwdP(), kBaseWX), zero);
            d2[t] = mm16<kFP>(basis(FwdP(), kBaseWY), make_opnd<kFP>(d1[0], d1[1], d1[2], d1[3]), zero);
        }
        // ---- temporal real basis (centre-indexed, separable form), staged as [j][p][q][b*2+c] ----
        {
            floatx4 tsum[R + 1], tdif[R + 1];
#pragma unroll
            for (int k2 = 1; k2 <= R; k2++) {
                tsum[k2] = d2[R + k2] + d2[R - k2];
                tdif[k2] = d2[R + k2] - d2[R - k2];
            }
#pragma unroll
            for (int j = 0; j < T; j++) {
                floatx4 u;
                if (j <= R) {
                    u = d2[R];
#pragma unroll
                    for (int k2 = 1; k2 <= R; k2++) {
                        u += tsum[k2] * wt[j][k2];
                    }
                } else {
                    u = tdif[1] * wt[j][1];
#pragma unroll
                    for (int k2 = 2; k2 <= R; k2++) {
                        u += tdif[k2] * wt[j][k2];
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int m = 4 * lg + i;
                    wbuf[((j * 8 + (m & 7)) * 8 + wq) * 4 + (m >> 3) * 2 + wc] = u[i];
                }
            }
        }
        // ---- output gradient -> P = WY (dy WX), dV = P * dv_scale (in ct) ----
        {
            float dov[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                dov[s] = raw4_value<float>(cur_go, s);
            }
            if (!cur_full) { // border blocks only (a uniform branch)
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    dov[s] = ((cur_ok >> s) & 1u) ? dov[s] : 0.0f;
                }
            }
            const floatx4 pm = mm16<kDP>(make_opnd<kDP>(dov[0], dov[1], dov[2], dov[3]), basis(DvP(), kBaseGX), zero);
            const floatx4 dv = mm16<kDP>(basis(DvP(), kBaseGY), make_opnd<kDP>(pm[0], pm[1], pm[2], pm[3]), zero);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = 4 * lg + i;
                vbuf[((m & 7) * 8 + wq) * 4 + (m >> 3) * 2 + wc] = dv[i];
            }
        }
        wave_lds_sync();

        adjoint_middle<R>(wbuf, vbuf, tab, lane, ct, dsg, a);

        // ---- adjoint temporal basis; add to the pending ring of virtual frames k .. k+2R ----
        floatx4 du[T];
#pragma unroll
        for (int j = 0; j < T; j++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = 4 * lg + i;
                du[j][i] = wbuf[((j * 8 + (m & 7)) * 8 + wq) * 4 + (m >> 3) * 2 + wc];
            }
        }
        floatx4 gs[T];
        adjoint_temporal<R, true>(du, wt, gs);
#pragma unroll
        for (int t = 0; t < 2 * R; t++) {
            gs[t] += gp[t];
        }
#pragma unroll
        for (int t = 1; t < T; t++) {
            gp[t - 1] = gs[t];
        }
        wave_lds_sync(); // du read before the staged gradients overwrite the buffer

        // ---- adjoint spatial DFT of the completed virtual frames, staged as [y][x] ----
        // gs (m in registers) as A operand is dD2^T; R1 = dD2^T SWY, ds^T = (AS[R] SWX) R1.
#pragma unroll
        for (int s = 0; s < T; s++) {
            if (s > 0 && !last) {
                continue;
            }
            const floatx4 r1 = mm16<kAP>(make_opnd<kAP>(gs[s][0], gs[s][1], gs[s][2], gs[s][3]), basis(AdjP(), kBaseWYT), zero);
            const floatx4 dbt = mm16<kAP>(basis(AdjP(), kBaseWXT), make_opnd<kAP>(r1[0], r1[1], r1[2], r1[3]), zero);
            // row l16, columns 4*(lane/16)..+3, at offset s*256 + wave*S of the wave's buffer
            float* srow = wbuf + s * 256 + wave * S + l16 * 16 + 4 * lg;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                srow[i] = dbt[i];
            }
        }
    }
        __syncthreads();
        if (last) {
            gather(std::integral_constant<int, T>(), k, pass, r_band);
        } else {
            gather(std::integral_constant<int, 1>(), k, pass, r_band);
        }
        __syncthreads(); // staged gradients consumed before the next block reuses the buffers
    } // k
    } // pass
    // Band r is final in every virtual frame (later block rows start below it).
    flush_band(r, r_band);
    r_band = r_band + 1 == nbands ? 0 : r_band + 1;
    } // r
    for (int b = nby; b * S < rows_used; b++) {
        flush_band(b, r_band);
        r_band = r_band + 1 == nbands ? 0 : r_band + 1;
    }

    for (int i = tid; i < T * 144; i += kThreads) {
        dsig_lds[i] = 0.0f;
    }
    sigma_grad_out<R>(dsig_lds, dsg, lane, tid, a);
}
} // namespace

size_t backward_lds_bytes(int radius, int region_w, int region_h) {
    const int t_len = 2 * radius + 1;
    return sizeof(float) * (static_cast<size_t>(radius + 1) * 4 * 64 * 4 +
                            static_cast<size_t>(t_len) * region_w * region_h +
                            static_cast<size_t>(kWaves) * (t_len * 256 + 256));
}

size_t backward_reuse_lds_bytes(int radius, int region_w, int block_step, int reuse_k) {
    const int t_len = 2 * radius + 1;
    const int ring_rows = (16 + block_step - 1) / block_step * block_step;
    return sizeof(float) * (static_cast<size_t>(reuse_k + 2 * radius) * ring_rows * region_w +
                            static_cast<size_t>(radius + 1) * 4 * 64 * 4 + kBases * 64 * 4 +
                            static_cast<size_t>(kWaves) * (t_len * 256 + 256));
}

namespace {
const void* backward_ptr(int radius, bool reuse) {
    if (reuse) {
        switch (radius) {
        case 0:
            return nullptr;
        case 1:
            return reinterpret_cast<const void*>(backward_reuse_kernel<1>);
        case 2:
            return reinterpret_cast<const void*>(backward_reuse_kernel<2>);
        default:
            return reinterpret_cast<const void*>(backward_reuse_kernel<3>);
        }
    }
    switch (radius) {
    case 0:
        return reinterpret_cast<const void*>(backward_kernel<0>);
    case 1:
        return reinterpret_cast<const void*>(backward_kernel<1>);
    case 2:
        return reinterpret_cast<const void*>(backward_kernel<2>);
    default:
        return reinterpret_cast<const void*>(backward_kernel<3>);
    }
}
} // namespace

int backward_workgroups_per_cu(int radius, bool reuse, size_t lds) {
    const void* kern = backward_ptr(radius, reuse);
    if (!kern) {
        return 0;
    }
    if (lds > 65536 && hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)) !=
                           hipSuccess) {
        return 0;
    }
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, kThreads, lds) != hipSuccess) {
        return 0;
    }
    return n;
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
    if (a.reuse_k > 1) {
        switch (radius) {
        case 0:
            return hipErrorInvalidValue;
        case 1:
            return run(backward_reuse_kernel<1>);
        case 2:
            return run(backward_reuse_kernel<2>);
        default:
            return run(backward_reuse_kernel<3>);
        }
    }
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

} // namespace dft2

// Generic block path for gfx950: block sizes other than 16 and tbsize above 7, which the
// reference serves with its FFT-library backends (hip_source/source.cpp:520-700: window,
// R2C, gain, C2R, then col2im of the centre slice, hip_source/kernel.hpp:88-250).  Same
// pipeline and conventions as the fused kernel, in two kernels:
//   generic_block_kernel: one wave per block.  The wave stages the T x B x B windowed
//     samples in LDS (mirror padding and temporal clamp in the address, as the fused
//     kernel), runs the separable 3-D real DFT as direct sums over twiddle tables (x real
//     -> B/2+1 bins, then y, then t), removes the mean (gf = Re X0 / Re WF0), applies the
//     gain (cpu_source/kernel.hpp:1826-1890), re-adds the mean, and evaluates the inverse
//     at the centre slice only, then over y, then C2R over x (the imaginary parts of the
//     kx = 0 and kx = B/2 columns are ignored); times the synthesis window, the B x B
//     result goes to the block-output buffer.
//   generic_gather_kernel: one thread per output pixel sums, in a fixed order, the
//     outputs of the blocks covering it and rounds as the reference
//     (cpu_source/getframe_impl.cpp:177-221).
// FP32 throughout, sums of at most B or T terms per stage.

#include "dft2_device.h"

namespace dft2 {
namespace {

__device__ __forceinline__ float generic_gain(float psd, float s, const GenericArgs& g) {
    switch (g.filter_type) {
    case 0:
        return fmaxf((psd - s) / (psd + 1e-15f), 0.0f);
    case 1:
        return psd < s ? 0.0f : 1.0f;
    case 2:
        return s;
    case 3:
        return (psd >= g.pmin && psd <= g.pmax) ? s : g.sigma2;
    case 4:
        return s * sqrtf(psd * (g.pmax / ((psd + g.pmin) * (psd + g.pmax) + 1e-15f)));
    case 5:
        return powf(fmaxf((psd - s) / (psd + 1e-15f), 0.0f), g.pmin);
    default:
        return sqrtf(fmaxf((psd - s) / (psd + 1e-15f), 0.0f));
    }
}

template <typename Tin>
__device__ __forceinline__ float load_sample(const Tin* p) {
    return static_cast<float>(*p);
}

// One wave per block; waves of a workgroup work on consecutive blocks, each in its own LDS
// buffers (two complex T x B x (B/2+1) arrays).  The twiddle tables (cos and sin of
// 2 pi m / B and of 2 pi m / T) are copied to LDS behind those buffers once per workgroup:
// read from global memory inside the DFT loops, with a lane-dependent index, every
// multiply-add waited for a load.  After that copy's barrier there is no workgroup
// barrier: a wave's LDS accesses are ordered by wave barriers only.
template <typename Tin>
__global__ __launch_bounds__(256) void generic_block_kernel(GenericArgs g) {
    extern __shared__ __attribute__((aligned(16))) float smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int B = g.block_size, H = g.half, T = g.t_len, R = g.radius, S = g.block_step;
    const int tbh = T * B * H;
    float* tw = smem + static_cast<size_t>(blockDim.x >> 6) * 4 * tbh; // [cos_b B | sin_b B | cos_t T | sin_t T]
    for (int i = threadIdx.x; i < 2 * (B + T); i += blockDim.x) {
        tw[i] = i < B ? g.cos_b[i] : i < 2 * B ? g.sin_b[i - B] : i < 2 * B + T ? g.cos_t[i - 2 * B] : g.sin_t[i - 2 * B - T];
    }
    __syncthreads();
    const float* cos_b = tw;
    const float* sin_b = tw + B;
    const float* cos_t = tw + 2 * B;
    const float* sin_t = cos_t + T;
    const int item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= g.n_items) {
        return;
    }
    float* x0 = smem + static_cast<size_t>(wave) * 4 * tbh; // complex [tbh] (first: real [T][B][B])
    float* x1 = x0 + 2 * tbh;                               // complex [tbh]

    const JobDev& job = g.jobs[sgpr(find_job<0>(g.jobs, g.n_jobs, item))];
    const int nbx = job.nbx, nby = job.nby;
    const int local = item - static_cast<int>(job.item_begin);
    const int f_local = local / (nbx * nby);
    const int rem = local - f_local * nbx * nby;
    const int by = rem / nbx;
    const int bx = rem - by * nbx;
    const int n = job.f_begin + f_local;
    const int width = job.width, height = job.height;

    // ---- windowed samples, [t][y][x] ----
    const Tin* src = reinterpret_cast<const Tin*>(job.src);
    for (int e = lane; e < T * B * B; e += 64) {
        const int t = e / (B * B);
        const int yx = e - t * B * B;
        const int y = yx / B;
        const int x = yx - y * B;
        const int fi = min(max(n - R + t, 0), job.n_frames - 1);
        const int py = mirror(by * S + y - job.oy, height);
        const int px = mirror(bx * S + x - job.ox, width);
        const Tin* base = job.use_fptr ? reinterpret_cast<const Tin*>(job.fptr[fi]) : src + fi * job.src_fp;
        x0[e] = g.win[e] * load_sample<Tin>(base + static_cast<long long>(py) * job.src_rs + px);
    }
    wave_lds_sync();
    // ---- x: real [t][y][x] -> complex [t][y][k], k < B/2 + 1 ----
    for (int e = lane; e < tbh; e += 64) {
        const int ty = e / H;
        const int k = e - ty * H;
        const float* row = x0 + ty * B;
        float re = 0.0f, im = 0.0f;
        for (int x = 0, m = 0; x < B; x++, m = (m + k >= B) ? m + k - B : m + k) {
            re = fmaf(row[x], cos_b[m], re);
            im = fmaf(-row[x], sin_b[m], im);
        }
        x1[2 * e] = re;
        x1[2 * e + 1] = im;
    }
    wave_lds_sync();
    // ---- y: [t][y][k] -> [t][ky][k] ----
    for (int e = lane; e < tbh; e += 64) {
        const int t = e / (B * H);
        const int r2 = e - t * B * H;
        const int ky = r2 / H;
        const int k = r2 - ky * H;
        const float* col = x1 + 2 * (t * B * H + k);
        float re = 0.0f, im = 0.0f;
        for (int y = 0, m = 0; y < B; y++, m = (m + ky >= B) ? m + ky - B : m + ky) {
            const float a = col[2 * y * H], b = col[2 * y * H + 1], c = cos_b[m], s = sin_b[m];
            re = fmaf(a, c, fmaf(b, s, re));   // (a + ib)(c - is)
            im = fmaf(b, c, fmaf(-a, s, im));
        }
        x0[2 * e] = re;
        x0[2 * e + 1] = im;
    }
    wave_lds_sync();
    // ---- t: [t][ky][k] -> [kt][ky][k] ----
    for (int e = lane; e < tbh; e += 64) {
        const int kt = e / (B * H);
        const int r2 = e - kt * B * H;
        float re = 0.0f, im = 0.0f;
        for (int t = 0, m = 0; t < T; t++, m = (m + kt >= T) ? m + kt - T : m + kt) {
            const float a = x0[2 * (t * B * H + r2)], b = x0[2 * (t * B * H + r2) + 1];
            const float c = cos_t[m], s = sin_t[m];
            re = fmaf(a, c, fmaf(b, s, re));
            im = fmaf(b, c, fmaf(-a, s, im));
        }
        x1[2 * e] = re;
        x1[2 * e + 1] = im;
    }
    wave_lds_sync();
    // ---- zero-mean, gain, mean re-add ----
    const float gf = g.zero_mean ? x1[0] * g.inv_wf0 : 0.0f;
    wave_lds_sync();
    unsigned long long* rec =
        g.decisions ? g.decisions + (job.dec_base + local) * static_cast<long long>(g.dec_words) : nullptr;
    for (int e = lane, word = 0; e < tbh; e += 64, word++) {
        const float mr = gf * g.wf[2 * e], mi = gf * g.wf[2 * e + 1];
        const float xr = x1[2 * e] - mr, xi = x1[2 * e + 1] - mi;
        const float psd = xr * xr + xi * xi;
        const float s = g.sig[e];
        const float gn = generic_gain(psd, s, g);
        x1[2 * e] = fmaf(xr, gn, mr);
        x1[2 * e + 1] = fmaf(xi, gn, mi);
        if (rec) {
            // lanes 0..63 of this iteration hold coefficients 64 word .. 64 word + 63
            const bool bit = g.filter_type == 1 ? psd < s : (psd >= g.pmin && psd <= g.pmax);
            const unsigned long long mask = __ballot(bit);
            if (lane == 0) {
                rec[word] = mask;
            }
        }
    }
    wave_lds_sync();
    // ---- inverse over t at the centre slice: [kt][ky][k] -> [ky][k] ----
    for (int e = lane; e < B * H; e += 64) {
        float re = 0.0f, im = 0.0f;
        for (int kt = 0, m = 0; kt < T; kt++, m = (m + R >= T) ? m + R - T : m + R) {
            const float a = x1[2 * (kt * B * H + e)], b = x1[2 * (kt * B * H + e) + 1];
            const float c = cos_t[m], s = sin_t[m];
            re = fmaf(a, c, fmaf(-b, s, re));  // (a + ib)(c + is)
            im = fmaf(a, s, fmaf(b, c, im));
        }
        x0[2 * e] = re;
        x0[2 * e + 1] = im;
    }
    wave_lds_sync();
    // ---- inverse over y: [ky][k] -> [y][k] ----
    for (int e = lane; e < B * H; e += 64) {
        const int y = e / H;
        const int k = e - y * H;
        float re = 0.0f, im = 0.0f;
        for (int ky = 0, m = 0; ky < B; ky++, m = (m + y >= B) ? m + y - B : m + y) {
            const float a = x0[2 * (ky * H + k)], b = x0[2 * (ky * H + k) + 1];
            const float c = cos_b[m], s = sin_b[m];
            re = fmaf(a, c, fmaf(-b, s, re));
            im = fmaf(a, s, fmaf(b, c, im));
        }
        x1[2 * e] = re;
        x1[2 * e + 1] = im;
    }
    wave_lds_sync();
    // ---- C2R over x, synthesis window ----
    float* out = g.blockout + static_cast<long long>(item) * B * B;
    for (int e = lane; e < B * B; e += 64) {
        const int y = e / B;
        const int x = e - y * B;
        const float* row = x1 + 2 * y * H;
        float sum = row[0];
        for (int k = 1, m = x; k < H; k++, m = (m + x >= B) ? m + x - B : m + x) {
            if (2 * k == B) {
                sum = fmaf(row[2 * k], cos_b[m], sum); // Nyquist column: real part only
            } else {
                sum = fmaf(2.0f * row[2 * k], cos_b[m], fmaf(-2.0f * row[2 * k + 1], sin_b[m], sum));
            }
        }
        out[e] = sum * g.syn[e];
    }
}

template <typename Tout>
__device__ __forceinline__ void store_rounded(Tout* p, float v, const GenericArgs& g) {
    if constexpr (sizeof(Tout) == 4) {
        *p = v / g.out_scale;
    } else {
        const float x = fmaxf(0.0f, fminf(v / g.out_scale + 0.5f, static_cast<float>(g.peak)));
        *p = static_cast<Tout>(static_cast<int>(x));
    }
}

// blockIdx.z: job-frame of the launch; (blockIdx.y, x): output pixel.
template <typename Tout>
__global__ __launch_bounds__(256) void generic_gather_kernel(GenericArgs g) {
    const int jf = blockIdx.z;
    const JobDev& job = g.jobs[sgpr(find_job<1>(g.jobs, g.n_jobs, jf))];
    const int y = blockIdx.y;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= job.height || x >= job.width) {
        return;
    }
    const int B = g.block_size, S = g.block_step;
    const int f_local = jf - static_cast<int>(job.frame_begin);
    const int py = y + job.oy, px = x + job.ox;
    // blocks i with i*S <= py < i*S + B (py >= B - S: the padding offset is at least that)
    const int i1 = (py - B + S) / S, i2 = min(py / S, job.nby - 1);
    const int j1 = (px - B + S) / S, j2 = min(px / S, job.nbx - 1);
    const long long base = job.item_begin + static_cast<long long>(f_local) * job.nbx * job.nby;
    float sum = 0.0f;
    for (int i = i1; i <= i2; i++) {
        const float* rowp = g.blockout + (base + static_cast<long long>(i) * job.nbx) * B * B + (py - i * S) * B;
        for (int j = j1; j <= j2; j++) {
            sum += rowp[static_cast<long long>(j) * B * B + px - j * S];
        }
    }
    Tout* dst = reinterpret_cast<Tout*>(job.dst) + (job.f_begin + f_local) * job.dst_fp +
                static_cast<long long>(y) * job.dst_rs + x;
    store_rounded<Tout>(dst, sum, g);
}

} // namespace

hipError_t launch_generic_blocks(const GenericArgs& g, int sample_code, int waves_per_group, hipStream_t stream) {
    const size_t lds = generic_lds_per_wave(g.block_size, g.t_len) * waves_per_group +
                       sizeof(float) * 2 * static_cast<size_t>(g.block_size + g.t_len);
    const dim3 grid(static_cast<unsigned>((g.n_items + waves_per_group - 1) / waves_per_group));
    const dim3 block(64 * waves_per_group);
    const void* kern = sample_code == kU8    ? reinterpret_cast<const void*>(generic_block_kernel<uint8_t>)
                       : sample_code == kU16 ? reinterpret_cast<const void*>(generic_block_kernel<uint16_t>)
                                             : reinterpret_cast<const void*>(generic_block_kernel<float>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) {
            return e;
        }
    }
    GenericArgs args = g;
    void* params[] = {&args};
    return hipLaunchKernel(kern, grid, block, params, lds, stream);
}

hipError_t launch_generic_gather(const GenericArgs& g, int sample_code, int max_width, int max_height, long long frames,
                                 hipStream_t stream) {
    const dim3 grid(static_cast<unsigned>((max_width + 255) / 256), static_cast<unsigned>(max_height),
                    static_cast<unsigned>(frames));
    switch (sample_code) {
    case kU8:
        hipLaunchKernelGGL(generic_gather_kernel<uint8_t>, grid, dim3(256), 0, stream, g);
        break;
    case kU16:
        hipLaunchKernelGGL(generic_gather_kernel<uint16_t>, grid, dim3(256), 0, stream, g);
        break;
    default:
        hipLaunchKernelGGL(generic_gather_kernel<float>, grid, dim3(256), 0, stream, g);
        break;
    }
    return hipGetLastError();
}

} // namespace dft2

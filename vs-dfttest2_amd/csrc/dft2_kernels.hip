// Launch code of the fused kernel (dispatch over the per-radius objects), the finalize
// kernel and the work-queue reset.

#include "dft2_device.h"

namespace dft2 {
namespace {

template <typename Tout>
__device__ __forceinline__ void store_sample(Tout* p, float v, const LaunchArgs& a) {
    if constexpr (sizeof(Tout) == 4) {
        *p = v / a.out_scale;
    } else {
        float x = fmaxf(0.0f, fminf(v / a.out_scale + 0.5f, static_cast<float>(a.peak)));
        *p = static_cast<Tout>(static_cast<int>(x));
    }
}

// Four consecutive converted samples to p, aligned to 4 * sizeof(Tout) bytes: one store.
template <typename Tout>
__device__ __forceinline__ void store_samples4(Tout* p, const float (&v)[4], const LaunchArgs& a) {
    if constexpr (sizeof(Tout) == 4) {
        *reinterpret_cast<float4*>(p) =
            make_float4(v[0] / a.out_scale, v[1] / a.out_scale, v[2] / a.out_scale, v[3] / a.out_scale);
    } else {
        unsigned q[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const float x = fmaxf(0.0f, fminf(v[e] / a.out_scale + 0.5f, static_cast<float>(a.peak)));
            q[e] = static_cast<unsigned>(static_cast<int>(x));
        }
        if constexpr (sizeof(Tout) == 1) {
            *reinterpret_cast<unsigned*>(p) = q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
        } else {
            *reinterpret_cast<uint2*>(p) = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
        }
    }
}

// One workgroup per work item (frame, tile): the output pixels the tile owns (its S*tb x
// S*tb corner of the padded plane; the last tile in a row or column owns its whole
// region) take the tile's own region plus the 16 - S overlapping columns / rows of the
// left, top and top-left neighbours' regions, summed in a fixed order (top-left, top,
// left, own), then converted to the sample type.  Scratch region index = work item.
template <typename Tout>
__global__ __launch_bounds__(256) void finalize_kernel(LaunchArgs a) {
    const JobDev& job = a.jobs[sgpr(find_job<0>(a.jobs, a.n_jobs, blockIdx.x))];
    const int tiles_x = sgpr(job.tiles_x);
    const int tiles_y = sgpr(job.tiles_y);
    const int tiles = tiles_x * tiles_y;
    const int job_begin = sgpr(static_cast<int>(job.item_begin));
    // workgroup -> work item of the same job, each XCD a contiguous part of the job: the
    // neighbours whose overlap strips a tile reads were finalized just before it on the
    // same XCD, so the partly used lines they share are still in its L2
    const int item = job_xcd_item(blockIdx.x, job_begin, sgpr(job.f_count) * tiles);
    const int local = item - job_begin;
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
    const int h = y1 - y0;
    // Columns in groups of four starting at multiples of 4 (padded coordinates): one float4
    // read per region and one store per group, when the tile origin, the padding offset, the
    // tile width and the output rows are aligned for it; sums in the same order as below.
    if (((x_lo | ox | tw) & 3) == 0 && (drs & 3) == 0 &&
        (reinterpret_cast<unsigned long long>(dst) % (4 * sizeof(Tout))) == 0) {
        const int g0 = x0 >> 2;
        const int ng = ((x1 + 3) >> 2) - g0;
        const int dr4 = 256 / ng;
        const int dc4 = 256 - dr4 * ng;
        int r = threadIdx.x / ng;
        int g = threadIdx.x - r * ng;
        for (; r < h;) {
            const int x = (g0 + g) * 4;
            const int ly = y0 + r - y_lo;
            const int lx = x - x_lo; // a multiple of 4, as are rw and tw: float4 reads
            const float4 o4 = *reinterpret_cast<const float4*>(own + ly * rw + lx);
            float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            const float ow[4] = {o4.x, o4.y, o4.z, o4.w};
            const bool lft = has_left && lx < ov;
            if (has_top && ly < ov) {
                const float4 t4 = *reinterpret_cast<const float4*>(top + (ly + th) * rw + lx);
                const float tp[4] = {t4.x, t4.y, t4.z, t4.w};
                if (lft) {
                    const float4 c4 = *reinterpret_cast<const float4*>(corner + (ly + th) * rw + lx + tw);
                    const float cn[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        v[e] += lx + e < ov ? cn[e] : 0.0f;
                    }
                }
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    v[e] += tp[e];
                }
            }
            if (lft) {
                const float4 l4 = *reinterpret_cast<const float4*>(left + ly * rw + lx + tw);
                const float lf[4] = {l4.x, l4.y, l4.z, l4.w};
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    v[e] += lx + e < ov ? lf[e] : 0.0f;
                }
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
                v[e] += ow[e];
            }
            Tout* drow = dst + (y0 + r - oy) * drs;
            if (x >= x0 && x + 4 <= x1) {
                store_samples4(drow + (x - ox), v, a);
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if (x + e >= x0 && x + e < x1) {
                        store_sample(drow + (x + e - ox), v[e], a);
                    }
                }
            }
            g += dc4;
            r += dr4;
            if (g >= ng) {
                g -= ng;
                ++r;
            }
        }
        return;
    }
    const int w = x1 - x0;
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

const void* fused_ptr(const FusedKey& k) {
    switch (k.radius) {
    case 0:
        return fused_kernel_ptr(k, RadiusTag<0>{});
    case 1:
        return fused_kernel_ptr(k, RadiusTag<1>{});
    case 2:
        return fused_kernel_ptr(k, RadiusTag<2>{});
    default:
        return fused_kernel_ptr(k, RadiusTag<3>{});
    }
}

hipError_t allow_lds(const void* kern, size_t lds) {
    if (lds > 65536) {
        return hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    }
    return hipSuccess;
}

} // namespace

size_t fused_lds_bytes(int radius, int region_w, int region_h) {
    const int t_len = 2 * radius + 1;
    const int wbuf = (t_len * 256 > 448) ? t_len * 256 : 448; // as kWbuf in the fused kernel
    return sizeof(float) * (static_cast<size_t>(region_w) * region_h + (radius + 1) * 4 * 64 * 4 + kWaves * wbuf);
}

size_t fused_reuse_lds_bytes(int radius, int region_w, int block_step, int reuse_k) {
    const int t_len = 2 * radius + 1;
    const int wbuf = (t_len * 256 > 448) ? t_len * 256 : 448; // as kWbuf in the fused kernel
    const int ring_rows = (16 + block_step - 1) / block_step * block_step;
    return sizeof(float) * (static_cast<size_t>(reuse_k) * ring_rows * region_w + (radius + 1) * 4 * 64 * 4 +
                            kWaves * wbuf);
}

hipError_t launch_fused(const LaunchArgs& a, const FusedKey& k, unsigned grid, size_t lds_bytes,
                        hipStream_t stream) {
    if (k.persistent != (a.queue != nullptr) || k.asym != a.asym || (k.reuse != 0) != (a.reuse_k > 1)) {
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

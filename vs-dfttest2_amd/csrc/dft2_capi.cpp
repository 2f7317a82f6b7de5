// Host side of the dfttest2_mi355x C ABI (include/dfttest2_mi355x.h).
//
// Plan creation follows the reference DFTTestCreate (hiprtc_source/source.cpp:710-934):
// argument checks, float conversion of window / sigma / window_freq, device upload.
// Instead of compiling a kernel per radius at run time it selects precompiled gfx950
// instantiations, and instead of the reference's host-side padding + per-frame launches
// it describes whole planes (jobs) to the fused kernel.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dft2_host.h"

using namespace dft2;
using namespace dft2::host;

namespace dft2 {
namespace host {

namespace {
thread_local std::string g_last_error;
} // namespace

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() {
    g_last_error.clear();
}

const char* last_error() {
    return g_last_error.c_str();
}

int pad_size(int size, int bs, int step) {
    return size + ((size % bs) ? bs - size % bs : 0) + std::max(bs - step, step) * 2;
}

int pad_num(int size, int bs, int step) {
    return (pad_size(size, bs, step) - bs) / step + 1;
}

bool tile_valid(const dft2_plan* p, int bw, int bh) {
    // A pixel is covered by ceil(16 / S) consecutive blocks per axis; they fall into at
    // most 2 tiles when a tile holds at least ceil(16 / S) - 1 blocks.
    const int S = p->block_step;
    const int need = std::max(1, (16 + S - 1) / S - 1);
    return bw >= need && bh >= need && bw <= 64 && bh <= 64;
}

TileShape choose_tile(const dft2_plan* p) {
    // A block row of a tile is computed in passes of kWaves blocks (one per wave), so
    // widths that are multiples of kWaves keep every wave busy.
    static const int kCand[][2] = {{8, 24}, {16, 12}, {8, 16}, {16, 16}, {16, 8}, {8, 12}, {16, 24}, {8, 8}};
    const int tb_max = 84 / p->block_step + 1; // accumulator region at most ~100 samples per side
    TileShape best = make_tile(p, std::min(16, tb_max), std::min(16, tb_max));
    long long best_score = -1;
    for (const auto& c : kCand) {
        const int bw = std::min(c[0], tb_max), bh = std::min(c[1], tb_max);
        if (!tile_valid(p, bw, bh)) {
            continue;
        }
        const TileShape ts = make_tile(p, bw, bh);
        const int wg = fused_workgroups_per_cu(fused_key(p), ts.lds);
        // Occupancy first (measured on 1080p: 3 workgroups per CU beat 2 at any tile size),
        // then the blocks per wave slot of a tile (fewer tiles, less halo per block).
        const int wave_slots = kWaves * ((bw + kWaves - 1) / kWaves); // per block row
        const long long score = static_cast<long long>(wg) * 100000 + 1000LL * bh * bw / wave_slots;
        if (wg > 0 && score > best_score) {
            best_score = score;
            best = ts;
        }
    }
    return best;
}

TileShape make_tile(const dft2_plan* p, int bw, int bh) {
    TileShape ts;
    ts.bw = bw;
    ts.bh = bh;
    ts.region_w = ((p->block_step * (bw - 1) + 16) + 3) / 4 * 4;
    ts.region_h = p->block_step * (bh - 1) + 16;
    ts.lds = fused_lds_bytes(p->radius, ts.region_w, ts.region_h);
    return ts;
}

// Cross-frame form of plan p: frames per work item and tile.  It exists for separable
// windows at radius >= 1.  Default: 8 frames per item at radius 2 (tbsize 5, where the
// spatial transforms of T slices take most MFMA cycles; whole clips in one launch, 1080p
// and 4K YUV420 of 8-bit, 10-bit and float samples, both precisions: 8 frames 4-14 %
// faster than the one-frame form, 2-10 % faster than 4 frames, the most rings for which
// two workgroups still fit a CU) and at radius 3 for float samples (1080p 1,830 -> 2,259
// fps FP32); radius 3 with integer samples keeps the one-frame form (its tbsize-7 spectra
// take 153-166 VGPRs in the cross-frame form, one workgroup per CU: 1080p 8-bit 2,775 fps
// one-frame against 2,452 at 8 frames, FP32; split BF16 3,229 against 2,631), and so does
// radius 1 (its extra live registers cost a third of the resident waves: 1080p 8-bit
// tbsize 3, 33.2 -> 41.3 ms at 4 frames).  Calls whose jobs have one output frame each
// (per-frame calls) take the one-frame form whatever the plan's (run_jobs).
// DFT2_REUSE=K and DFT2_REUSE_TILE=WxH override.  The tile: 8 blocks wide (one per wave
// and pass), and the tallest candidate among those with the most resident workgroups:
// its accumulators are rings of 16-row strips, so height costs no LDS, and taller tiles
// have fewer overlap strips to write and gather.
void choose_reuse(dft2_plan* p) {
    p->reuse_k = 1;
    if (!p->separable || p->radius < 1) {
        return;
    }
    int k = p->radius == 2 || (p->radius >= 3 && p->sample_code == kF32) ? 8 : 1;
    if (const char* env = std::getenv("DFT2_REUSE")) {
        k = std::max(1, std::min(8, std::atoi(env)));
    }
    if (k <= 1) {
        return;
    }
    const int tb_max = 84 / p->block_step + 1;
    int bw = std::min(8, tb_max);
    int best_wg = -1;
    TileShape best;
    int forced_w = 0, forced_h = 0;
    if (const char* env = std::getenv("DFT2_REUSE_TILE")) {
        std::sscanf(env, "%dx%d", &forced_w, &forced_h);
    }
    static const int kHeights[] = {32, 24, 16, 12, 8};
    for (int bh0 : kHeights) {
        const int w = forced_w > 0 ? forced_w : bw;
        const int h = forced_h > 0 ? forced_h : std::min(bh0, 64);
        if (!tile_valid(p, w, h)) {
            continue;
        }
        TileShape ts = make_tile(p, w, h);
        ts.lds = fused_reuse_lds_bytes(p->radius, ts.region_w, p->block_step, k);
        FusedKey key = fused_key(p);
        key.reuse = 1;
        const int wg = fused_workgroups_per_cu(key, ts.lds);
        if (wg > best_wg) {
            best_wg = wg;
            best = ts;
        }
        if (forced_h > 0) {
            break;
        }
    }
    if (best_wg <= 0) {
        return;
    }
    p->reuse_k = k;
    p->reuse_tile = best;
}

} // namespace host
} // namespace dft2

namespace dft2 {
namespace host {

// Builds the constant table (layout in dft2_internal.h) in FP64 and rounds to float.
void build_consts(dft2_plan* p, const dft2_params* prm, std::vector<float>& c, float scale) {
    const double pi = 3.14159265358979323846;
    const int T = p->t_len, r = p->radius;
    c.assign(kConstFloats, 0.0f);
    for (int x = 0; x < 16; x++) {
        for (int n = 0; n < 16; n++) {
            double v = n <= 8 ? std::cos(2 * pi * n * x / 16.0) : std::sin(2 * pi * (n - 8) * x / 16.0);
            double w = (n == 0 || n == 8) ? 1.0 : 2.0;
            c[kOffWX + x * 16 + n] = static_cast<float>(v);  // WX[x][n]
            c[kOffWY + n * 16 + x] = static_cast<float>(v);  // WY[m][y]
            c[kOffGX + n * 16 + x] = static_cast<float>(w * v); // GX[n][x]
            c[kOffGY + x * 16 + n] = static_cast<float>(w * v); // GY[y][m]
        }
    }
    for (int i = 0; i < T * 256; i++) {
        // load_block computes (scale * window) * sample in float (cpu_source/getframe_impl.cpp:130-152)
        c[kOffWIN + i] = scale * static_cast<float>(prm->window[i]);
    }
    for (int i = 0; i < 256; i++) {
        c[kOffSYN + i] = static_cast<float>(prm->window[r * 256 + i]);
    }
    for (int j = 0; j < T; j++) {
        for (int t = 0; t < T; t++) {
            double v = j <= r ? std::cos(2 * pi * j * t / T) : std::sin(2 * pi * (j - r) * t / T);
            c[kOffWT + j * T + t] = static_cast<float>(v);
        }
        // centre slice t' = 0: cos rows 1, sin rows 0
        c[kOffCT + j] = static_cast<float>(j == 0 ? 1.0 : (j <= r ? 2.0 : 0.0));
    }
    auto sigma_at = [&](int kt, int ky, int kx) -> float {
        if (prm->sigma_len == 1) {
            return static_cast<float>(prm->sigma[0]);
        }
        return static_cast<float>(prm->sigma[(kt * 16 + ky) * 9 + kx]);
    };
    // Window spectrum in the centre-indexed temporal convention: times exp(2 pi i kt r / T).
    auto wf_at = [&](int kt, int ky, int kx, int part) -> float {
        if (!p->zero_mean) {
            return 0.0f;
        }
        const double re = prm->window_freq[((kt * 16 + ky) * 9 + kx) * 2];
        const double im = prm->window_freq[((kt * 16 + ky) * 9 + kx) * 2 + 1];
        const double ph = 2 * pi * kt * r / T;
        return static_cast<float>(part == 0 ? re * std::cos(ph) - im * std::sin(ph) : re * std::sin(ph) + im * std::cos(ph));
    };
    // Middle-stage ownership (see dft2_kernels.hip): lane = 8p + q, evaluations e = 0, 1,
    // slots A (2e) and B (2e+1).
    for (int g = 0; g <= r; g++) {
        for (int lane = 0; lane < 64; lane++) {
            const int pp = lane >> 3, q = lane & 7;
            const bool ey = pp != 0, ex = q != 0;
            for (int e = 0; e < 2; e++) {
                const int ky = ey ? (e == 0 ? pp : (16 - pp) % 16) : (e == 0 ? 0 : 8);
                int pos[2][3];
                if (ex) {
                    pos[0][0] = g, pos[0][1] = ky, pos[0][2] = q;
                    pos[1][0] = (T - g) % T, pos[1][1] = (16 - ky) % 16, pos[1][2] = q;
                } else {
                    pos[0][0] = g, pos[0][1] = ky, pos[0][2] = 0;
                    pos[1][0] = g, pos[1][1] = ky, pos[1][2] = 8;
                }
                for (int k = 0; k < 2; k++) {
                    const int kt = pos[k][0], ky2 = pos[k][1], kx = pos[k][2];
                    const bool edge = kx == 0 || kx == 8;
                    float* e4 = &c[kOffTAB + (((g * 4 + 2 * e + k) * 64) + lane) * 4];
                    // Wiener plans keep sigma + 1e-15, the form the kernel's gain takes.
                    const float bias = p->filter_type == 0 ? 1e-15f : 0.0f;
                    e4[0] = sigma_at(kt, ky2, kx) + bias;
                    e4[1] = (edge ? sigma_at((T - kt) % T, (16 - ky2) % 16, kx) : sigma_at(kt, ky2, kx)) + bias;
                    e4[2] = wf_at(kt, ky2, kx, 0);
                    e4[3] = wf_at(kt, ky2, kx, 1);
                }
            }
        }
    }
    // Separable window: w[t][y][x] = a[t] b[y] c[x] to within 1e-9 of its peak, with a[t]
    // symmetric about the centre slice (every DFTTest window type is).
    {
        const double* w = prm->window;
        int t0 = 0, y0 = 0, x0 = 0;
        double peak = 0.0;
        for (int i = 0; i < T * 256; i++) {
            if (std::fabs(w[i]) > peak) {
                peak = std::fabs(w[i]);
                t0 = i / 256, y0 = (i / 16) % 16, x0 = i % 16;
            }
        }
        p->separable = 0;
        if (peak > 0.0) {
            const double w0 = w[(t0 * 16 + y0) * 16 + x0];
            double at[7], by[16], cx[16];
            for (int t = 0; t < T; t++) {
                at[t] = w[(t * 16 + y0) * 16 + x0];
            }
            for (int i = 0; i < 16; i++) {
                by[i] = w[(t0 * 16 + i) * 16 + x0] / w0;
                cx[i] = w[(t0 * 16 + y0) * 16 + i] / w0;
            }
            bool sep = true;
            for (int k = 1; k <= r; k++) {
                if (std::fabs(at[r + k] - at[r - k]) > 1e-9 * peak) {
                    sep = false;
                }
            }
            for (int t = 0; t < T && sep; t++) {
                for (int y = 0; y < 16 && sep; y++) {
                    for (int x = 0; x < 16; x++) {
                        if (std::fabs(w[(t * 16 + y) * 16 + x] - at[t] * by[y] * cx[x]) > 1e-9 * peak) {
                            sep = false;
                            break;
                        }
                    }
                }
            }
            // The separable form scales the slices by the centre slice's factor and the
            // temporal basis by the others' ratios to it.
            if (sep && at[r] != 0.0) {
                p->separable = 1;
                for (int x = 0; x < 16; x++) {
                    for (int n = 0; n < 16; n++) {
                        const double v = n <= 8 ? std::cos(2 * pi * n * x / 16.0) : std::sin(2 * pi * (n - 8) * x / 16.0);
                        c[kOffSWX + x * 16 + n] = static_cast<float>(v * cx[x]);
                        c[kOffSWY + n * 16 + x] = static_cast<float>(v * by[x]);
                    }
                }
                for (int t = 0; t < T; t++) {
                    c[kOffAS + t] = static_cast<float>(static_cast<double>(scale) * at[t]);
                }
                for (int j = 0; j < T; j++) {
                    c[kOffCTS + j] = static_cast<float>((j == 0 ? 1.0 : (j <= r ? 2.0 : 0.0)) * at[r]);
                    for (int k = 1; k <= r; k++) {
                        const double v = j <= r ? std::cos(2 * pi * j * k / T) : std::sin(2 * pi * (j - r) * k / T);
                        c[kOffWTS + j * T + k] = static_cast<float>(v * at[r + k] / at[r]);
                    }
                }
            }
        }
    }
    // Asymmetric sigma: conjugate partners at kx = 0, 8 carry different sigmas.
    p->asym = 0;
    if (prm->sigma_len != 1) {
        for (int kt = 0; kt < T && !p->asym; kt++) {
            for (int ky = 0; ky < 16 && !p->asym; ky++) {
                for (int kx : {0, 8}) {
                    if (sigma_at(kt, ky, kx) != sigma_at((T - kt) % T, (16 - ky) % 16, kx)) {
                        p->asym = 1;
                        break;
                    }
                }
            }
        }
    }
    // The symmetric Wiener form clamps its damping factor sigma / psd to [0, 1], which holds
    // for sigma >= 0 only; a negative sigma (gains above 1) takes the other form.
    for (int i = 0; i < prm->sigma_len && p->filter_type == 0; i++) {
        if (prm->sigma[i] < 0.0) {
            p->asym = 1;
        }
    }
}

Slot* acquire_slot(dft2_plan* p) {
    std::unique_lock<std::mutex> lk(p->mu);
    for (;;) {
        for (auto& s : p->slots) {
            if (!s.busy) {
                s.busy = true;
                return &s;
            }
        }
        p->cv.wait(lk);
    }
}

void release_slot(dft2_plan* p, Slot* s) {
    {
        std::lock_guard<std::mutex> lk(p->mu);
        s->busy = false;
    }
    p->cv.notify_one();
}

// Holds a stream slot for one call.  Once the call may queue work on stream st it sets
// st; a call that returns early after that (an error) leaves `recorded` false, and the
// slot's done event is recorded here, so that the slot's next call, on whatever stream,
// still waits for the work this one queued before reusing the slot's buffers.
struct SlotGuard {
    dft2_plan* p;
    Slot* s;
    hipStream_t st = nullptr;
    bool recorded = false;
    ~SlotGuard() {
        if (st && !recorded) {
            (void)hipEventRecord(s->done, st);
        }
        release_slot(p, s);
    }
};

int ensure_device(void** ptr, size_t* cap, size_t bytes) {
    if (*cap >= bytes) {
        return DFT2_OK;
    }
    if (*ptr) {
        HIP_TRY(hipFree(*ptr));
        *ptr = nullptr;
        *cap = 0;
    }
    HIP_TRY(hipMalloc(ptr, bytes));
    *cap = bytes;
    return DFT2_OK;
}

int ensure_host(void** ptr, size_t* cap, size_t bytes) {
    if (*cap >= bytes) {
        return DFT2_OK;
    }
    if (*ptr) {
        HIP_TRY(hipHostFree(*ptr));
        *ptr = nullptr;
        *cap = 0;
    }
    HIP_TRY(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    *cap = bytes;
    return DFT2_OK;
}

// True for memory the GPU reaches directly (pinned host, device or managed); false for
// ordinary pageable host memory, whose attributes query fails or reports it unregistered.
bool device_visible(const void* ptr) {
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, ptr) != hipSuccess) {
        (void)hipGetLastError(); // the query's error is not the caller's
        return false;
    }
    return at.type != hipMemoryTypeUnregistered;
}

int check_mirror(int size, int block_size, int step, const char* what) {
    int ps = pad_size(size, block_size, step);
    int o = (ps - size) / 2;
    if (o > size - 1 || ps - o - size > size - 1) {
        return fail(DFT2_ERR_ARG, "%s %d is too small for reflection padding of %d samples", what, size, o);
    }
    return DFT2_OK;
}

// Fills the launch arguments common to every kernel of a plan.
LaunchArgs make_args(const dft2_plan* p, const TileShape& ts) {
    LaunchArgs a{};
    a.consts = p->d_consts;
    a.block_step = p->block_step;
    a.step_magic = (65536 + p->block_step - 1) / p->block_step;
    a.tile_bw = ts.bw;
    a.tile_bh = ts.bh;
    a.region_w = ts.region_w;
    a.region_h = ts.region_h;
    a.inv_wf0 = 1.0f / p->wf0;
    a.sigma2 = p->sigma2;
    a.pmin = p->pmin;
    a.pmax = p->pmax;
    a.filter_type = p->filter_type;
    a.zero_mean = p->zero_mean;
    a.asym = p->asym;
    a.out_scale = p->out_scale;
    a.peak = p->peak;
    a.sigma_dev = nullptr;
    a.sigma_scale = 1.0f;
    a.sigma_grad = nullptr;
    a.tab_dev = nullptr;
    a.queue = nullptr;
    a.n_items = 0;
    a.decisions = nullptr;
    a.reuse_k = 1;
    a.ring_bands = (16 + p->block_step - 1) / p->block_step;
    return a;
}

// Makes `table_bytes` of job descriptors resident in device memory: a cached table with
// the same bytes is reused; otherwise the least recently used cache entry is refilled
// through the slot's pinned ring.  *out_dev receives the device table.
int upload_jobs(Slot* slot, const JobDev* src_table, size_t table_bytes, hipStream_t stream, JobDev** out_dev) {
    const unsigned char* src_bytes = reinterpret_cast<const unsigned char*>(src_table);
    DeviceTable* victim = &slot->tables[0];
    for (auto& t : slot->tables) {
        if (t.dev && t.bytes.size() == table_bytes && std::memcmp(t.bytes.data(), src_bytes, table_bytes) == 0) {
            t.last_use = ++slot->table_clock;
            *out_dev = t.dev;
            return DFT2_OK;
        }
        if (t.last_use < victim->last_use) {
            victim = &t;
        }
    }
    if (victim->cap < table_bytes) {
        // Kernels queued earlier may still read the old buffer.
        HIP_TRY(hipStreamSynchronize(stream));
        int rc = ensure_device(reinterpret_cast<void**>(&victim->dev), &victim->cap, table_bytes);
        if (rc) {
            return rc;
        }
    }
    victim->bytes.assign(src_bytes, src_bytes + table_bytes);
    victim->last_use = ++slot->table_clock;
    *out_dev = victim->dev;
    const int ring = slot->ring_pos;
    slot->ring_pos = (slot->ring_pos + 1) % kRing;
    if (slot->h_jobs_done[ring]) {
        HIP_TRY(hipEventSynchronize(slot->h_jobs_done[ring]));
    } else {
        HIP_TRY(hipEventCreateWithFlags(&slot->h_jobs_done[ring], hipEventDisableTiming));
    }
    int rc = ensure_host(reinterpret_cast<void**>(&slot->h_jobs[ring]), &slot->h_jobs_cap[ring], table_bytes);
    if (rc) {
        return rc;
    }
    std::memcpy(slot->h_jobs[ring], src_table, table_bytes);
    HIP_TRY(hipMemcpyAsync(victim->dev, slot->h_jobs[ring], table_bytes, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipEventRecord(slot->h_jobs_done[ring], stream));
    return DFT2_OK;
}

int plan_chunks(const dft2_plan* p, const TileShape& ts, const std::vector<JobDev>& jobs_in, size_t scratch_cap,
                ChunkPlan& out, int reuse_k) {
    const int n_jobs = static_cast<int>(jobs_in.size());
    const long long region = static_cast<long long>(ts.region_w) * ts.region_h;
    long long per_frame_items = 0;
    int max_frames = 0;
    for (const auto& j : jobs_in) {
        per_frame_items += static_cast<long long>(j.tiles_x) * j.tiles_y;
        max_frames = std::max(max_frames, j.f_count);
    }
    out = ChunkPlan();
    out.n_jobs = n_jobs;
    if (per_frame_items == 0 || max_frames == 0) {
        return DFT2_OK;
    }
    long long frames_per_chunk = static_cast<long long>(scratch_cap / (per_frame_items * region * sizeof(float)));
    // A launch has fewer than 2^31 work items (32-bit item arithmetic in the kernels).
    frames_per_chunk = std::min<long long>(frames_per_chunk, (1LL << 31) / per_frame_items - 1);
    frames_per_chunk = std::max(1LL, std::min<long long>(frames_per_chunk, max_frames));
    // Whole groups of reuse_k frames per chunk: a group cut by a chunk boundary repeats
    // its first slices' transforms in the next chunk.
    if (frames_per_chunk > reuse_k && frames_per_chunk < max_frames) {
        frames_per_chunk -= frames_per_chunk % reuse_k;
    }
    if (per_frame_items >= (1LL << 30)) {
        return fail(DFT2_ERR_ARG, "too many blocks in one frame of the batch");
    }
    const int n_chunks = static_cast<int>((max_frames + frames_per_chunk - 1) / frames_per_chunk);
    out.scratch_bytes = static_cast<size_t>(frames_per_chunk * per_frame_items * region) * sizeof(float);
    out.table.resize(static_cast<size_t>(n_jobs) * n_chunks);
    out.chunks.resize(n_chunks);
    for (int c = 0; c < n_chunks; c++) {
        Chunk& ch = out.chunks[c];
        for (int k = 0; k < n_jobs; k++) {
            JobDev j = jobs_in[k];
            const int f0 = static_cast<int>(c * frames_per_chunk);
            const int cnt = std::max(0, std::min<int>(static_cast<int>(frames_per_chunk), j.f_count - f0));
            j.f_begin = j.f_begin + f0;
            j.f_count = cnt;
            j.dec_base += static_cast<long long>(f0) * j.nby * j.nbx;
            j.item_begin = ch.items;
            j.frame_begin = ch.frames;
            j.group_begin = ch.groups;
            ch.items += static_cast<long long>(cnt) * j.tiles_x * j.tiles_y;
            ch.frames += cnt;
            ch.groups += static_cast<long long>((cnt + reuse_k - 1) / reuse_k) * j.tiles_x * j.tiles_y;
            if (cnt > 0) {
                ch.max_pixels = std::max(ch.max_pixels, j.width * j.height);
            }
            out.table[static_cast<size_t>(c) * n_jobs + k] = j;
        }
    }
    return DFT2_OK;
}

// With timing enabled, hands out three events from the plan's free pool (creating them
// when it is empty); otherwise leaves ev[] null.  The caller records them around its
// launch pair and then passes them to timing_commit.
int timing_events(dft2_plan* p, hipEvent_t ev[3]) {
    std::lock_guard<std::mutex> lk(p->mu);
    if (!p->timing) {
        return DFT2_OK;
    }
    for (int i = 0; i < 3; i++) {
        if (p->timing_free.empty()) {
            hipEvent_t e = nullptr;
            HIP_TRY(hipEventCreate(&e));
            p->timing_free.push_back(e);
        }
        ev[i] = p->timing_free.back();
        p->timing_free.pop_back();
    }
    return DFT2_OK;
}

// Waits for the oldest pending triple, adds its durations to the totals and returns its
// events to the free pool.  Called with p->mu held.
int timing_resolve_oldest(dft2_plan* p) {
    const dft2_plan::TimingRecord rec = p->timing_pending.front();
    const std::array<hipEvent_t, 3>& ev = rec.ev;
    p->timing_pending.pop_front();
    for (hipEvent_t e : ev) {
        p->timing_free.push_back(e);
    }
    HIP_TRY(hipEventSynchronize(ev[2]));
    float f = 0.0f, g = 0.0f;
    HIP_TRY(hipEventElapsedTime(&f, ev[0], ev[1]));
    HIP_TRY(hipEventElapsedTime(&g, ev[1], ev[2]));
    p->timing_totals.fused_launches++;
    p->timing_totals.fused_ms += f;
    p->timing_totals.finalize_launches++;
    p->timing_totals.finalize_ms += g;
    p->timing_totals.blocks += rec.blocks;
    p->timing_totals.slice_spectra += rec.slices;
    return DFT2_OK;
}

// Output blocks of one chunk's jobs and the slice spectra computed for them: T per
// block in the one-frame form; in the cross-frame form each group of up to reuse_k
// consecutive frames computes frames + T - 1 slices per block position.
void chunk_work(const dft2_plan* p, const JobDev* jobs, int n_jobs, int reuse_k, long long& blocks,
                long long& slices) {
    blocks = slices = 0;
    for (int i = 0; i < n_jobs; i++) {
        const JobDev& j = jobs[i];
        const long long nb = static_cast<long long>(j.nbx) * j.nby;
        blocks += nb * j.f_count;
        if (reuse_k > 1) {
            const long long full = j.f_count / reuse_k, rest = j.f_count % reuse_k;
            slices += nb * (full * (reuse_k + 2 * p->radius) + (rest ? rest + 2 * p->radius : 0));
        } else {
            slices += nb * j.f_count * p->t_len;
        }
    }
}

// Queues a recorded triple with its launch pair's work; beyond kTimingPending pending
// triples the oldest is resolved.
int timing_commit(dft2_plan* p, const hipEvent_t ev[3], long long blocks, long long slices) {
    if (!ev[0]) {
        return DFT2_OK;
    }
    std::lock_guard<std::mutex> lk(p->mu);
    p->timing_pending.push_back({{ev[0], ev[1], ev[2]}, blocks, slices});
    while (p->timing_pending.size() > kTimingPending) {
        int rc = timing_resolve_oldest(p);
        if (rc) {
            return rc;
        }
    }
    return DFT2_OK;
}

// Builds the launch's gain table from a.sigma_dev into the slot's buffer and points
// a.tab_dev at it.
int build_gain_table(const dft2_plan* p, Slot* slot, LaunchArgs& a, bool bias, hipStream_t stream) {
    const size_t bytes = sizeof(float) * kTabFloats(p->radius);
    if (slot->tab_dev_cap < bytes) {
        int rc = ensure_device(reinterpret_cast<void**>(&slot->tab_dev), &slot->tab_dev_cap, bytes);
        if (rc) {
            return rc;
        }
    }
    a.tab_dev = nullptr;
    HIP_TRY(launch_gain_table(a, p->radius, bias, slot->tab_dev, stream));
    a.tab_dev = slot->tab_dev;
    return DFT2_OK;
}

// Generic path's table (layout of GenericArgs, dft2_internal.h) in FP64, rounded to float.
void build_generic_consts(dft2_plan* p, const dft2_params* prm, std::vector<float>& c) {
    const double pi = 3.14159265358979323846;
    const int B = prm->block_size, H = B / 2 + 1, T = 2 * prm->radius + 1, r = prm->radius;
    GenericLayout L = generic_layout(B, T);
    c.assign(L.total, 0.0f);
    for (int m = 0; m < B; m++) {
        c[L.cos_b + m] = static_cast<float>(std::cos(2 * pi * m / B));
        c[L.sin_b + m] = static_cast<float>(std::sin(2 * pi * m / B));
    }
    for (int m = 0; m < T; m++) {
        c[L.cos_t + m] = static_cast<float>(std::cos(2 * pi * m / T));
        c[L.sin_t + m] = static_cast<float>(std::sin(2 * pi * m / T));
    }
    for (int i = 0; i < T * B * B; i++) {
        // load_block's (scale * window) in float (cpu_source/getframe_impl.cpp:130-152)
        c[L.win + i] = p->out_scale * static_cast<float>(prm->window[i]);
    }
    for (int i = 0; i < B * B; i++) {
        c[L.syn + i] = static_cast<float>(prm->window[r * B * B + i]);
    }
    for (int i = 0; i < T * B * H; i++) {
        c[L.sig + i] = static_cast<float>(prm->sigma_len == 1 ? prm->sigma[0] : prm->sigma[i]);
        if (p->zero_mean) {
            c[L.wf + 2 * i] = static_cast<float>(prm->window_freq[2 * i]);
            c[L.wf + 2 * i + 1] = static_cast<float>(prm->window_freq[2 * i + 1]);
        }
    }
}

// Generic path: the block kernel writes every block's windowed output to slot scratch,
// the gather kernel sums and rounds, one launch pair per frame chunk.
int run_generic(dft2_plan* p, Slot* slot, const std::vector<JobDev>& jobs_in, hipStream_t stream,
                unsigned long long* decisions) {
    ChunkPlan cp;
    int rc = plan_chunks(p, p->tile, jobs_in, p->scratch_cap, cp);
    if (rc || cp.chunks.empty()) {
        return rc;
    }
    rc = ensure_device(reinterpret_cast<void**>(&slot->scratch), &slot->scratch_bytes, cp.scratch_bytes);
    if (rc) {
        return rc;
    }
    JobDev* d_table = nullptr;
    rc = upload_jobs(slot, cp.table.data(), sizeof(JobDev) * cp.table.size(), stream, &d_table);
    if (rc) {
        return rc;
    }
    const GenericLayout L = generic_layout(p->block_size, p->t_len);
    GenericArgs g{};
    g.cos_b = p->d_consts + L.cos_b;
    g.sin_b = p->d_consts + L.sin_b;
    g.cos_t = p->d_consts + L.cos_t;
    g.sin_t = p->d_consts + L.sin_t;
    g.win = p->d_consts + L.win;
    g.syn = p->d_consts + L.syn;
    g.sig = p->d_consts + L.sig;
    g.wf = p->d_consts + L.wf;
    g.n_jobs = cp.n_jobs;
    g.blockout = slot->scratch;
    g.block_size = p->block_size;
    g.half = p->block_size / 2 + 1;
    g.t_len = p->t_len;
    g.radius = p->radius;
    g.block_step = p->block_step;
    g.filter_type = p->filter_type;
    g.zero_mean = p->zero_mean;
    g.inv_wf0 = 1.0f / p->wf0;
    g.sigma2 = p->sigma2;
    g.pmin = p->pmin;
    g.pmax = p->pmax;
    g.out_scale = p->out_scale;
    g.peak = p->peak;
    g.decisions = decisions;
    g.dec_words = generic_dec_words(p);
    for (size_t c = 0; c < cp.chunks.size(); c++) {
        const Chunk& ch = cp.chunks[c];
        if (ch.items == 0) {
            continue;
        }
        g.jobs = d_table + c * cp.n_jobs;
        g.n_items = static_cast<int>(ch.items);
        int max_w = 0, max_h = 0;
        for (int k = 0; k < cp.n_jobs; k++) {
            const JobDev& j = cp.table[c * cp.n_jobs + k];
            if (j.f_count > 0) {
                max_w = std::max(max_w, j.width);
                max_h = std::max(max_h, j.height);
            }
        }
        hipEvent_t ev[3] = {};
        rc = timing_events(p, ev);
        if (rc) {
            return rc;
        }
        if (ev[0]) {
            HIP_TRY(hipEventRecord(ev[0], stream));
        }
        HIP_TRY(launch_generic_blocks(g, p->sample_code, p->gen_waves, stream));
        if (ev[1]) {
            HIP_TRY(hipEventRecord(ev[1], stream));
        }
        HIP_TRY(launch_generic_gather(g, p->sample_code, max_w, max_h, ch.frames, stream));
        if (ev[2]) {
            HIP_TRY(hipEventRecord(ev[2], stream));
        }
        long long blocks = 0, slices = 0;
        chunk_work(p, &cp.table[c * cp.n_jobs], cp.n_jobs, 1, blocks, slices);
        rc = timing_commit(p, ev, blocks, slices);
        if (rc) {
            return rc;
        }
    }
    return DFT2_OK;
}

// Runs `jobs` (already validated, device memory) on `stream` using slot scratch.
int run_jobs(dft2_plan* p, Slot* slot, const std::vector<JobDev>& jobs_in, hipStream_t stream,
             const float* sigma_dev = nullptr, float sigma_scale = 1.0f,
             unsigned long long* decisions = nullptr) {
    if (p->generic) {
        return run_generic(p, slot, jobs_in, stream, decisions);
    }
    // The cross-frame form when the plan takes it (it records no decisions), for calls
    // with more than one output frame in some job: with one frame per job its items would
    // transform every slice anyway and pay the rings and the form's registers for nothing.
    int max_frames = 0;
    for (const auto& j : jobs_in) {
        max_frames = std::max(max_frames, j.f_count);
    }
    const int reuse_k = decisions || max_frames <= 1 ? 1 : p->reuse_k;
    const TileShape& ts = reuse_k > 1 ? p->reuse_tile : p->tile;
    // Job geometry counts tiles of p->tile; recount them for the form's tile.
    std::vector<JobDev> jobs_t;
    const std::vector<JobDev>* jobs_use = &jobs_in;
    if (reuse_k > 1) {
        jobs_t = jobs_in;
        for (auto& j : jobs_t) {
            j.tiles_x = (j.nbx + ts.bw - 1) / ts.bw;
            j.tiles_y = (j.nby + ts.bh - 1) / ts.bh;
        }
        jobs_use = &jobs_t;
    }
    ChunkPlan cp;
    int rc = plan_chunks(p, ts, *jobs_use, p->scratch_cap, cp, reuse_k);
    if (rc || cp.chunks.empty()) {
        return rc;
    }
    rc = ensure_device(reinterpret_cast<void**>(&slot->scratch), &slot->scratch_bytes, cp.scratch_bytes);
    if (rc) {
        return rc;
    }
    // Device job table: reused across calls on this slot; stream order (or the wait on
    // the slot's previous completion) keeps earlier kernels from seeing new data.
    JobDev* d_table = nullptr;
    rc = upload_jobs(slot, cp.table.data(), sizeof(JobDev) * cp.table.size(), stream, &d_table);
    if (rc) {
        return rc;
    }
    LaunchArgs a = make_args(p, ts);
    a.n_jobs = cp.n_jobs;
    a.scratch = slot->scratch;
    if (sigma_dev) {
        a.sigma_dev = sigma_dev;
        a.sigma_scale = sigma_scale;
        a.asym = 1;
        rc = build_gain_table(p, slot, a, p->filter_type == 0, stream);
        if (rc) {
            return rc;
        }
    }
    FusedKey key = fused_key(p, 0, a.asym);
    if (decisions) {
        a.decisions = decisions;
        key.rec = 1;
    }
    if (reuse_k > 1) {
        a.reuse_k = reuse_k;
        key.reuse = 1;
    }
    for (size_t c = 0; c < cp.chunks.size(); c++) {
        const Chunk& ch = cp.chunks[c];
        if (ch.items == 0) {
            continue;
        }
        a.jobs = d_table + c * cp.n_jobs;
        hipEvent_t ev[3] = {};
        rc = timing_events(p, ev);
        if (rc) {
            return rc;
        }
        if (ev[0]) {
            HIP_TRY(hipEventRecord(ev[0], stream));
        }
        HIP_TRY(launch_fused(a, key, static_cast<unsigned>(reuse_k > 1 ? ch.groups : ch.items), ts.lds, stream));
        if (ev[1]) {
            HIP_TRY(hipEventRecord(ev[1], stream));
        }
        HIP_TRY(launch_finalize(a, p->sample_code, ch.items, stream));
        if (ev[2]) {
            HIP_TRY(hipEventRecord(ev[2], stream));
        }
        long long blocks = 0, slices = 0;
        chunk_work(p, &cp.table[c * cp.n_jobs], cp.n_jobs, reuse_k, blocks, slices);
        rc = timing_commit(p, ev, blocks, slices);
        if (rc) {
            return rc;
        }
    }
    return DFT2_OK;
}

// Fills geometry of a JobDev from the user's job description.
int make_job(const dft2_plan* p, const dft2_job& uj, const TileShape& ts, JobDev& j) {
    if (!uj.src || !uj.dst) {
        return fail(DFT2_ERR_ARG, "job src/dst must not be NULL");
    }
    int rc = check_mirror(uj.width, p->block_size, p->block_step, "width");
    if (rc) {
        return rc;
    }
    rc = check_mirror(uj.height, p->block_size, p->block_step, "height");
    if (rc) {
        return rc;
    }
    if (uj.n_frames <= 0 || uj.frame_begin < 0 || uj.frame_end > uj.n_frames || uj.frame_begin > uj.frame_end) {
        return fail(DFT2_ERR_ARG, "invalid frame range [%d, %d) of %d frames", uj.frame_begin, uj.frame_end,
                    uj.n_frames);
    }
    const int bps = p->bytes_per_sample;
    if (uj.src_stride % bps || uj.dst_stride % bps || uj.src_frame_pitch % bps || uj.dst_frame_pitch % bps) {
        return fail(DFT2_ERR_ARG, "strides and pitches must be multiples of the sample size (%d)", bps);
    }
    if (uj.src_stride < static_cast<long long>(uj.width) * bps || uj.dst_stride < static_cast<long long>(uj.width) * bps) {
        return fail(DFT2_ERR_ARG, "row stride smaller than the row");
    }
    // The kernels address samples within one source frame with 32-bit byte offsets.
    if (uj.src_stride * uj.height >= (1LL << 31)) {
        return fail(DFT2_ERR_ARG, "a source frame must be smaller than 2 GiB");
    }
    std::memset(&j, 0, sizeof(j));
    j.src = uj.src;
    j.dst = uj.dst;
    j.src_fp = uj.src_frame_pitch / bps;
    j.src_rs = uj.src_stride / bps;
    j.dst_fp = uj.dst_frame_pitch / bps;
    j.dst_rs = uj.dst_stride / bps;
    j.width = uj.width;
    j.height = uj.height;
    j.n_frames = uj.n_frames;
    j.f_begin = uj.frame_begin;
    j.f_count = uj.frame_end - uj.frame_begin;
    j.pad_w = pad_size(uj.width, p->block_size, p->block_step);
    j.pad_h = pad_size(uj.height, p->block_size, p->block_step);
    j.nbx = pad_num(uj.width, p->block_size, p->block_step);
    j.nby = pad_num(uj.height, p->block_size, p->block_step);
    j.ox = (j.pad_w - uj.width) / 2;
    j.oy = (j.pad_h - uj.height) / 2;
    j.tiles_x = (j.nbx + ts.bw - 1) / ts.bw;
    j.tiles_y = (j.nby + ts.bh - 1) / ts.bh;
    j.use_fptr = 0;
    return DFT2_OK;
}

int select_device(const dft2_plan* p) {
    HIP_TRY(hipSetDevice(p->device));
    return DFT2_OK;
}

} // namespace host
} // namespace dft2

extern "C" {

const char* dft2_last_error(void) {
    return last_error();
}

const char* dft2_version(void) {
    return "dfttest2_mi355x 0.1.0 (gfx950, fused FP32 / split-BF16 MFMA)";
}

int dft2_plan_create(const dft2_params* prm, int sample_type, int bits_per_sample, int device_id, int num_streams,
                     int precision, dft2_plan** out_plan) {
    clear_error();
    if (!prm || !out_plan) {
        return fail(DFT2_ERR_ARG, "params and out_plan must not be NULL");
    }
    *out_plan = nullptr;
    // Checks of DFTTestCreate (hiprtc_source/source.cpp:719-847).  The fused kernel takes
    // block_size 16 and radius 0..3, the reference's fused-backend limits; other block sizes
    // and radii up to kMaxGenericRadius take the generic block path, as the reference's
    // FFT-library backends do (python/dfttest2.py:283-287 restricts only the fused ones).
    if (prm->radius < 0 || prm->radius > kMaxGenericRadius) {
        return fail(DFT2_ERR_ARG, "radius must be in [0, %d] (0..3 with the fused kernel)", kMaxGenericRadius);
    }
    if (prm->block_size < 1 || prm->block_size > kMaxGenericBlock) {
        return fail(DFT2_ERR_ARG, "block_size must be in [1, %d] (16 with the fused kernel)", kMaxGenericBlock);
    }
    if (prm->block_step < 1 || prm->block_step > prm->block_size) {
        return fail(DFT2_ERR_ARG, "block_step must be in [1, %d]", prm->block_size);
    }
    const bool generic = prm->block_size != 16 || prm->radius > 3;
    if (prm->filter_type < 0 || prm->filter_type > 6) {
        return fail(DFT2_ERR_ARG, "filter_type must be in [0, 6]");
    }
    if (num_streams < 1) {
        return fail(DFT2_ERR_ARG, "num_streams must be at least 1");
    }
    if (precision != DFT2_PRECISION_FP32 && precision != DFT2_PRECISION_SPLIT_BF16) {
        return fail(DFT2_ERR_ARG, "precision must be 0 (FP32) or 1 (split BF16)");
    }
    // Filter types 3 and 4 take gains above 1 (sigma, sigma2): the split form's coefficient
    // error (about 2^-15 of the block's largest) is amplified with them, and type 3 jumps
    // between its two gains at pmin and pmax, so that form would miss the output bars.
    if (precision == DFT2_PRECISION_SPLIT_BF16 && (prm->filter_type == 3 || prm->filter_type == 4)) {
        return fail(DFT2_ERR_ARG, "split BF16 precision supports filter types 0, 1, 2, 5 and 6 (use FP32 for types 3 and 4)");
    }
    const int T = 2 * prm->radius + 1;
    const int B = prm->block_size, Hx = B / 2 + 1;
    if (!prm->window || prm->window_len != T * B * B) {
        return fail(DFT2_ERR_ARG, "window must have %d elements (got %d)", T * B * B, prm->window_len);
    }
    if (!prm->sigma || (prm->sigma_len != 1 && prm->sigma_len != T * B * Hx)) {
        return fail(DFT2_ERR_ARG, "sigma must have 1 or %d elements (got %d)", T * B * Hx, prm->sigma_len);
    }
    if (prm->zero_mean && (!prm->window_freq || prm->window_freq_len != T * B * Hx * 2)) {
        return fail(DFT2_ERR_ARG, "window_freq with %d elements is required when zero_mean is enabled",
                    T * B * Hx * 2);
    }
    if (generic && generic_lds_per_wave(B, T) > kGenericLdsBytes) {
        return fail(DFT2_ERR_ARG, "block_size %d with tbsize %d needs %zu KiB of LDS per block (at most %zu)", B, T,
                    generic_lds_per_wave(B, T) >> 10, kGenericLdsBytes >> 10);
    }
    if (generic && precision != DFT2_PRECISION_FP32) {
        return fail(DFT2_ERR_ARG, "split BF16 precision needs block_size 16 and radius <= 3 (the fused kernel)");
    }
    if (sample_type == DFT2_SAMPLE_INT) {
        if (bits_per_sample < 8 || bits_per_sample > 16) {
            return fail(DFT2_ERR_ARG, "only 8-16 bit integer format input is supported");
        }
    } else if (sample_type == DFT2_SAMPLE_FLOAT) {
        if (bits_per_sample != 32) {
            return fail(DFT2_ERR_ARG, "only 32-bit float format input is supported");
        }
    } else {
        return fail(DFT2_ERR_ARG, "invalid sample_type %d", sample_type);
    }

    auto* p = new dft2_plan();
    p->device = device_id;
    p->generic = generic ? 1 : 0;
    p->block_size = B;
    p->radius = prm->radius;
    p->t_len = T;
    p->block_step = prm->block_step;
    p->filter_type = prm->filter_type;
    p->zero_mean = prm->zero_mean ? 1 : 0;
    p->bits = bits_per_sample;
    p->precision = precision;
    if (sample_type == DFT2_SAMPLE_FLOAT) {
        p->sample_code = kF32;
        p->bytes_per_sample = 4;
        p->out_scale = 255.0f;
        p->peak = 0;
    } else {
        p->sample_code = bits_per_sample <= 8 ? kU8 : kU16;
        p->bytes_per_sample = bits_per_sample <= 8 ? 1 : 2;
        p->out_scale = 1.0f / static_cast<float>(1 << (bits_per_sample - 8));
        p->peak = (1 << bits_per_sample) - 1;
    }
    p->sigma2 = static_cast<float>(prm->sigma2);
    p->pmin = static_cast<float>(prm->pmin);
    p->pmax = static_cast<float>(prm->pmax);
    p->wf0 = p->zero_mean ? static_cast<float>(prm->window_freq[0]) : 1.0f;
    if (p->zero_mean && p->wf0 == 0.0f) {
        delete p;
        return fail(DFT2_ERR_ARG, "window_freq[0] must be non-zero when zero_mean is enabled");
    }
    // Tile of blocks per workgroup (refined by choose_tile once the device is selected):
    // accumulator region at most ~100 samples per side.
    const int tb = std::min(16, 84 / p->block_step + 1);
    p->tile = make_tile(p, tb, tb);
    if (const char* env = std::getenv("DFT2_SCRATCH_MB")) {
        long mb = std::atol(env);
        if (mb > 0) {
            p->scratch_cap = static_cast<size_t>(mb) << 20;
        }
    }

    std::vector<float> consts;
    if (generic) {
        build_generic_consts(p, prm, consts);
    } else {
        build_consts(p, prm, consts, p->out_scale);
    }

    int rc = select_device(p);
    if (rc) {
        delete p;
        return rc;
    }
    if (generic) {
        p->gen_waves = static_cast<int>(std::min<size_t>(4, kGenericLdsBytes / generic_lds_per_wave(B, T)));
        p->tile = TileShape();
        p->tile.bw = p->tile.bh = 1;
        p->tile.region_w = p->tile.region_h = B;
        p->tile.lds = 0;
    } else {
        p->tile = choose_tile(p);
        choose_reuse(p);
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->d_consts), sizeof(float) * consts.size());
    if (e == hipSuccess) {
        e = hipMemcpy(p->d_consts, consts.data(), sizeof(float) * consts.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        rc = fail(DFT2_ERR_HIP, "device setup failed: %s", hipGetErrorString(e));
        dft2_plan_destroy(p);
        return rc;
    }
    p->slots.resize(num_streams);
    for (auto& s : p->slots) {
        e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) {
            e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        }
        if (e == hipSuccess) {
            e = hipEventCreateWithFlags(&s.up_ready, hipEventDisableTiming);
        }
        if (e != hipSuccess) {
            rc = fail(DFT2_ERR_HIP, "stream setup failed: %s", hipGetErrorString(e));
            dft2_plan_destroy(p);
            return rc;
        }
    }
    *out_plan = p;
    return DFT2_OK;
}

int dft2_plan_set_timing(dft2_plan* p, int enable) {
    clear_error();
    if (!p) {
        return fail(DFT2_ERR_ARG, "plan must not be NULL");
    }
    std::lock_guard<std::mutex> lk(p->mu);
    if (enable) {
        // Pending events are waited for and dropped; the totals start from zero.
        while (!p->timing_pending.empty()) {
            int rc = timing_resolve_oldest(p);
            if (rc) {
                return rc;
            }
        }
        p->timing_totals = dft2_kernel_timing{};
    }
    p->timing = enable != 0;
    return DFT2_OK;
}

int dft2_plan_get_timing(dft2_plan* p, dft2_kernel_timing* out) {
    clear_error();
    if (!p || !out) {
        return fail(DFT2_ERR_ARG, "plan and out must not be NULL");
    }
    std::lock_guard<std::mutex> lk(p->mu);
    while (!p->timing_pending.empty()) {
        int rc = timing_resolve_oldest(p);
        if (rc) {
            return rc;
        }
    }
    *out = p->timing_totals;
    return DFT2_OK;
}

int dft2_plan_destroy(dft2_plan* p) {
    if (!p) {
        return DFT2_OK;
    }
    (void)hipSetDevice(p->device);
    for (auto& s : p->slots) {
        if (s.stream) {
            (void)hipStreamSynchronize(s.stream);
        }
        for (int i = 0; i < kRing; i++) {
            if (s.h_jobs_done[i]) {
                (void)hipEventSynchronize(s.h_jobs_done[i]);
                (void)hipEventDestroy(s.h_jobs_done[i]);
            }
            if (s.h_jobs[i]) {
                (void)hipHostFree(s.h_jobs[i]);
            }
        }
        if (s.done) {
            (void)hipEventSynchronize(s.done);
            (void)hipEventDestroy(s.done);
        }
        if (s.up_ready) {
            (void)hipEventDestroy(s.up_ready);
        }
        if (s.scratch) {
            (void)hipFree(s.scratch);
        }
        for (auto& t : s.tables) {
            if (t.dev) {
                (void)hipFree(t.dev);
            }
        }
        if (s.tab_dev) {
            (void)hipFree(s.tab_dev);
        }
        if (s.stage) {
            (void)hipFree(s.stage);
        }
        if (s.h_stage) {
            (void)hipHostFree(s.h_stage);
        }
        if (s.stream) {
            (void)hipStreamDestroy(s.stream);
        }
    }
    for (const auto& tr : p->timing_pending) {
        for (hipEvent_t e : tr.ev) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
    }
    for (hipEvent_t e : p->timing_free) {
        (void)hipEventDestroy(e);
    }
    for (auto& fp : p->frame_planes) {
        if (fp.dev) {
            (void)hipFree(fp.dev);
        }
    }
    if (p->d_consts) {
        (void)hipFree(p->d_consts);
    }
    delete p;
    return DFT2_OK;
}

int dft2_process_batch(dft2_plan* p, const dft2_job* jobs, int n_jobs, void* stream) {
    return dft2_process_batch_sigma(p, jobs, n_jobs, nullptr, 1.0f, stream);
}

int dft2_process_batch_sigma(dft2_plan* p, const dft2_job* jobs, int n_jobs, const float* sigma_dev,
                             float sigma_scale, void* stream) {
    clear_error();
    if (!p) {
        return fail(DFT2_ERR_ARG, "plan must not be NULL");
    }
    if (sigma_dev && p->generic) {
        return fail(DFT2_ERR_ARG, "sigma as a device array needs block_size 16 and radius <= 3 (the fused kernel)");
    }
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) {
        return fail(DFT2_ERR_ARG, "invalid job list");
    }
    std::vector<JobDev> jd(n_jobs);
    for (int i = 0; i < n_jobs; i++) {
        int rc = make_job(p, jobs[i], p->tile, jd[i]);
        if (rc) {
            return rc;
        }
    }
    int rc = select_device(p);
    if (rc) {
        return rc;
    }
    Slot* slot = acquire_slot(p);
    SlotGuard guard{p, slot};
    hipStream_t st = call_stream(stream, slot->stream);
    // The slot's buffers (scratch, job tables, gain table) may still be in use by its
    // previous call, queued on another stream: order this call after it.
    HIP_TRY(hipStreamWaitEvent(st, slot->done, 0));
    guard.st = st;
    rc = run_jobs(p, slot, jd, st, sigma_dev, sigma_scale);
    if (rc) {
        return rc;
    }
    HIP_TRY(hipEventRecord(slot->done, st));
    guard.recorded = true;
    if (!stream) {
        HIP_TRY(hipStreamSynchronize(st));
    }
    return DFT2_OK;
}

int dft2_process_batch_decisions(dft2_plan* p, const dft2_job* jobs, int n_jobs, unsigned long long* decisions_dev,
                                 long long decisions_len, void* stream) {
    clear_error();
    if (!p || !decisions_dev) {
        return fail(DFT2_ERR_ARG, "plan and decisions_dev must not be NULL");
    }
    if (p->filter_type != 1 && p->filter_type != 3) {
        return fail(DFT2_ERR_ARG, "decision records exist for filter types 1 and 3 only");
    }
    if (p->asym && !p->generic) {
        return fail(DFT2_ERR_ARG, "decision records need a symmetric sigma");
    }
    if (p->filter_type == 3 && p->precision != DFT2_PRECISION_FP32) {
        return fail(DFT2_ERR_ARG, "decision records of filter type 3 need FP32 precision");
    }
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) {
        return fail(DFT2_ERR_ARG, "invalid job list");
    }
    std::vector<JobDev> jd(n_jobs);
    long long blocks = 0;
    for (int i = 0; i < n_jobs; i++) {
        int rc = make_job(p, jobs[i], p->tile, jd[i]);
        if (rc) {
            return rc;
        }
        jd[i].dec_base = blocks;
        blocks += static_cast<long long>(jd[i].f_count) * jd[i].nby * jd[i].nbx;
    }
    const long long need = blocks * (p->generic ? generic_dec_words(p) : (p->radius + 1) * 4);
    if (decisions_len < need) {
        return fail(DFT2_ERR_ARG, "decisions_dev holds %lld masks, %lld needed", decisions_len, need);
    }
    int rc = select_device(p);
    if (rc) {
        return rc;
    }
    Slot* slot = acquire_slot(p);
    SlotGuard guard{p, slot};
    hipStream_t st = call_stream(stream, slot->stream);
    HIP_TRY(hipStreamWaitEvent(st, slot->done, 0)); // the slot's previous call, on any stream
    guard.st = st;
    rc = run_jobs(p, slot, jd, st, nullptr, 1.0f, decisions_dev);
    if (rc) {
        return rc;
    }
    HIP_TRY(hipEventRecord(slot->done, st));
    guard.recorded = true;
    if (!stream) {
        HIP_TRY(hipStreamSynchronize(st));
    }
    return DFT2_OK;
}

int dft2_backward_batch(dft2_plan* p, const dft2_grad_job* jobs, int n_jobs, const float* sigma_dev,
                        float sigma_scale, float* sigma_grad_dev, void* stream) {
    clear_error();
    if (!p || !sigma_dev || !sigma_grad_dev) {
        return fail(DFT2_ERR_ARG, "plan, sigma_dev and sigma_grad_dev must not be NULL");
    }
    if (p->sample_code != kF32 || p->filter_type != 0 || p->precision != DFT2_PRECISION_FP32) {
        return fail(DFT2_ERR_ARG, "the backward pass needs a float32 plan with filter_type 0 and FP32 precision");
    }
    if (p->generic) {
        return fail(DFT2_ERR_ARG, "the backward pass needs block_size 16 and radius <= 3 (the fused kernel)");
    }
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) {
        return fail(DFT2_ERR_ARG, "invalid job list");
    }
    // Tiles are 8 blocks wide (one per wave and pass).  The one-frame form keeps T
    // input-gradient regions per tile in LDS, so its tiles are as tall as still lets two
    // workgroups share a CU (measured on 256x256 float clips, T = 3: taller tiles amortise
    // the per-tile work; tiles that leave one workgroup per CU are slower).  The
    // cross-frame form (separable window, radius >= 1) takes 8 output frames per work
    // item and keeps ring accumulators, whose LDS does not grow with the tile height; its
    // tiles are the tallest of 16, 12, 8, 6, 5 and 4 blocks that still give every
    // resident workgroup slot 3.5 work items (8 clips of 256x256x16, tbsize 3: 1.30 ms in
    // the one-frame form; 8 frames per item with 16-, 8-, 6-, 5-, 4-, 3- and 2-block tiles
    // 1.30, 1.11, 1.095, 1.060, 1.079, 1.122 and 1.227 ms, 720, 1,296, 1,728, 2,016,
    // 2,448, 3,264 and 4,896 items for 512 slots; 4 frames per item 1.20, 1.15 and 1.24 ms
    // at 16, 8 and 4 blocks).  DFT2_BWD_REUSE=K and DFT2_BWD_TILE_H=H override (K = 1:
    // one-frame form).
    const int S = p->block_step;
    const int tbw = std::min(8, 84 / S + 1);
    auto region_w_of = [S](int tb) { return ((S * (tb - 1) + 16) + 3) / 4 * 4; };
    auto region_h_of = [S](int tb) { return S * (tb - 1) + 16; };
    int reuse_k = p->separable && p->radius >= 1 ? 8 : 1;
    if (const char* env = std::getenv("DFT2_BWD_REUSE")) {
        reuse_k = p->separable && p->radius >= 1 ? std::max(1, std::min(16, std::atoi(env))) : 1;
    }
    int tbh = tbw;
    if (reuse_k > 1) {
        int forced_h = 0;
        if (const char* env = std::getenv("DFT2_BWD_TILE_H")) {
            forced_h = std::max(1, std::min(64, std::atoi(env)));
        }
        int rc = select_device(p);
        if (rc) {
            return rc;
        }
        int n_cu = 0;
        HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, p->device));
        const int wg = backward_workgroups_per_cu(p->radius, true,
                                                  backward_reuse_lds_bytes(p->radius, region_w_of(tbw), S, reuse_k));
        if (wg <= 0) {
            return fail(DFT2_ERR_HIP, "the cross-frame backward kernel does not fit a CU");
        }
        const long long slots = static_cast<long long>(n_cu) * wg;
        for (int h : {16, 12, 8, 6, 5, 4}) {
            tbh = h;
            if (forced_h > 0) {
                tbh = forced_h;
                break;
            }
            long long n_items = 0;
            for (int i = 0; i < n_jobs; i++) {
                const int nbx = pad_num(jobs[i].width, 16, S), nby = pad_num(jobs[i].height, 16, S);
                const long long frames = std::max(0, jobs[i].frame_end - jobs[i].frame_begin);
                n_items += (frames + reuse_k - 1) / reuse_k * ((nbx + tbw - 1) / tbw) * ((nby + h - 1) / h);
            }
            if (2 * n_items >= 7 * slots) {
                break;
            }
        }
    } else {
        for (int cand : {16, 12}) {
            const int h = std::min(cand, 84 / S + 1);
            if (h > tbh && backward_lds_bytes(p->radius, region_w_of(tbw), region_h_of(h)) <= 80 * 1024) {
                tbh = h;
                break;
            }
        }
    }
    const int region_w = region_w_of(tbw);
    const int region_h = region_h_of(tbh);
    const size_t lds = reuse_k > 1 ? backward_reuse_lds_bytes(p->radius, region_w, S, reuse_k)
                                   : backward_lds_bytes(p->radius, region_w, region_h);
    std::vector<JobDev> jd(n_jobs);
    long long items = 0;
    for (int i = 0; i < n_jobs; i++) {
        const dft2_grad_job& g = jobs[i];
        if (!g.grad_out || !g.grad_in) {
            return fail(DFT2_ERR_ARG, "grad_out and grad_in must not be NULL");
        }
        dft2_job uj{};
        uj.src = g.src;
        uj.dst = g.grad_in;
        uj.src_frame_pitch = g.src_frame_pitch;
        uj.src_stride = g.src_stride;
        uj.dst_frame_pitch = g.gi_frame_pitch;
        uj.dst_stride = g.gi_stride;
        uj.width = g.width;
        uj.height = g.height;
        uj.n_frames = g.n_frames;
        uj.frame_begin = g.frame_begin;
        uj.frame_end = g.frame_end;
        int rc = make_job(p, uj, p->tile, jd[i]);
        if (rc) {
            return rc;
        }
        if (g.go_stride < static_cast<long long>(g.width) * 4 || g.go_stride % 4 || g.go_frame_pitch % 4 ||
            g.gi_frame_pitch % 4 || g.go_stride * g.height >= (1LL << 31)) {
            return fail(DFT2_ERR_ARG, "invalid gradient strides");
        }
        jd[i].grad_out = g.grad_out;
        jd[i].grad_in = g.grad_in;
        jd[i].go_fp = g.go_frame_pitch / 4;
        jd[i].go_rs = g.go_stride / 4;
        jd[i].gi_fp = g.gi_frame_pitch / 4;
        jd[i].gi_rs = g.gi_stride / 4;
        jd[i].tiles_x = (jd[i].nbx + tbw - 1) / tbw;
        jd[i].tiles_y = (jd[i].nby + tbh - 1) / tbh;
        jd[i].item_begin = items;
        jd[i].group_begin = items;
        const long long per_tile = reuse_k > 1 ? (jd[i].f_count + reuse_k - 1) / reuse_k : jd[i].f_count;
        items += per_tile * jd[i].tiles_x * jd[i].tiles_y;
    }
    if (items >= (1LL << 31)) {
        return fail(DFT2_ERR_ARG, "too many backward work items in one call (split the jobs)");
    }
    if (items == 0) {
        return DFT2_OK;
    }
    int rc = select_device(p);
    if (rc) {
        return rc;
    }
    Slot* slot = acquire_slot(p);
    SlotGuard guard{p, slot};
    hipStream_t st = call_stream(stream, slot->stream);
    // The slot's buffers (scratch, job tables, gain table) may still be in use by its
    // previous call, queued on another stream: order this call after it.
    HIP_TRY(hipStreamWaitEvent(st, slot->done, 0));
    guard.st = st;
    JobDev* d_table = nullptr;
    rc = upload_jobs(slot, jd.data(), sizeof(JobDev) * jd.size(), st, &d_table);
    if (rc) {
        return rc;
    }
    LaunchArgs a = make_args(p, p->tile);
    a.jobs = d_table;
    a.n_jobs = n_jobs;
    a.sigma_dev = sigma_dev;
    a.sigma_scale = sigma_scale;
    a.sigma_grad = sigma_grad_dev;
    a.asym = 1;
    a.tile_bw = tbw;
    a.tile_bh = tbh;
    a.region_w = region_w;
    a.region_h = region_h;
    a.reuse_k = reuse_k;
    rc = build_gain_table(p, slot, a, true, st); // sigma + 1e-15, as gain_adjoint takes it
    if (rc) {
        return rc;
    }
    HIP_TRY(launch_backward(a, p->radius, items, lds, st));
    HIP_TRY(hipEventRecord(slot->done, st));
    guard.recorded = true;
    if (!stream) {
        HIP_TRY(hipStreamSynchronize(st));
    }
    return DFT2_OK;
}

namespace {
// Checks of a per-frame call (dft2_process_frame, dft2_process_frame_keyed): the source
// planes and the caller's geometry, before anything is staged or copied (the host paths
// copy height rows of width samples at the caller's strides).
int check_frame_call(const dft2_plan* p, const void* const* src_frames, long long src_stride, long long dst_stride,
                     int width, int height) {
    for (int t = 0; t < p->t_len; t++) {
        if (!src_frames[t]) {
            return fail(DFT2_ERR_ARG, "src_frames[%d] is NULL", t);
        }
    }
    const int bps = p->bytes_per_sample;
    if (width < 1 || height < 1) {
        return fail(DFT2_ERR_ARG, "width and height must be positive (got %d x %d)", width, height);
    }
    if (src_stride % bps || dst_stride % bps) {
        return fail(DFT2_ERR_ARG, "strides and pitches must be multiples of the sample size (%d)", bps);
    }
    if (src_stride < static_cast<long long>(width) * bps || dst_stride < static_cast<long long>(width) * bps) {
        return fail(DFT2_ERR_ARG, "row stride smaller than the row");
    }
    int rc = check_mirror(width, p->block_size, p->block_step, "width");
    if (rc) {
        return rc;
    }
    return check_mirror(height, p->block_size, p->block_step, "height");
}

// The source planes a dft2_process_frame_keyed call uses, released when the call ends.
// Its work is complete then: a call that returns normally has waited for its stream, and
// one that fails waits here, so that no upload or kernel of it still touches a plane
// another call may claim next.  A plane this call was to upload and did not is
// invalidated.
struct FramePlaneClaim {
    dft2_plan* p;
    hipStream_t st = nullptr;
    bool synced = false;
    FramePlane* use[kMaxFramePtrs] = {};
    bool mine[kMaxFramePtrs] = {};
    int n = 0;
    ~FramePlaneClaim() {
        if (!synced && st) {
            (void)hipStreamSynchronize(st); // an error path: the call's own error stands
        }
        {
            std::lock_guard<std::mutex> lk(p->mu);
            for (int t = 0; t < n; t++) {
                bool first = true;
                for (int u = 0; u < t; u++) {
                    first = first && use[u] != use[t];
                }
                if (!use[t] || !first) {
                    continue;
                }
                if (mine[t] && !use[t]->recorded) {
                    use[t]->valid = false;
                }
                use[t]->users--;
            }
        }
    }
};
} // namespace

int dft2_process_frame(dft2_plan* p, const void* const* src_frames, long long src_stride, void* dst,
                       long long dst_stride, int width, int height, int memory, void* stream) {
    clear_error();
    if (!p || !src_frames || !dst) {
        return fail(DFT2_ERR_ARG, "plan, src_frames and dst must not be NULL");
    }
    const int T = p->t_len;
    const int bps = p->bytes_per_sample;
    if (memory != 0 && memory != 1) {
        return fail(DFT2_ERR_ARG, "memory must be 0 (device) or 1 (host)");
    }
    int rc = check_frame_call(p, src_frames, src_stride, dst_stride, width, height);
    if (rc) {
        return rc;
    }
    dft2_job uj{};
    uj.src = src_frames[0];
    uj.dst = dst;
    uj.width = width;
    uj.height = height;
    uj.n_frames = T;
    uj.frame_begin = p->radius;
    uj.frame_end = p->radius + 1;
    uj.src_stride = src_stride;
    uj.dst_stride = dst_stride;
    JobDev j;
    rc = select_device(p);
    if (rc) {
        return rc;
    }
    if (memory == 0) {
        rc = make_job(p, uj, p->tile, j);
        if (rc) {
            return rc;
        }
        j.use_fptr = 1;
        for (int t = 0; t < T; t++) {
            j.fptr[t] = src_frames[t];
        }
        // dst is frame index r of a T-frame clip; shift the base so that r lands on dst.
        j.dst_fp = 0;
        Slot* slot = acquire_slot(p);
        SlotGuard guard{p, slot};
        hipStream_t st = call_stream(stream, slot->stream);
        HIP_TRY(hipStreamWaitEvent(st, slot->done, 0)); // the slot's previous call, on any stream
        guard.st = st;
        rc = run_jobs(p, slot, {j}, st);
        if (rc) {
            return rc;
        }
        HIP_TRY(hipEventRecord(slot->done, st));
        guard.recorded = true;
        if (!stream) {
            HIP_TRY(hipStreamSynchronize(st));
        }
        return DFT2_OK;
    }
    // Host memory: stage T planes and the output through device memory (the reference
    // copies T padded planes host->device per frame, hiprtc_source/source.cpp:570-573).
    // Pinned planes go over with 2-D copies.  Pageable planes (VapourSynth's frames) are
    // first copied row by row into the slot's pinned buffer and moved with one copy each
    // way, as the reference stages through its per-thread pinned h_padded
    // (hiprtc_source/source.cpp:480-510): 2-D copies from pageable memory run at a small
    // fraction of the link rate (74 against 0.7 ms per 1080p YUV420 frame).
    const long long row = static_cast<long long>(width) * bps;
    const long long plane = row * height;
    bool pinned = device_visible(dst);
    for (int t = 0; t < T && pinned; t++) {
        pinned = device_visible(src_frames[t]);
    }
    Slot* slot = acquire_slot(p);
    SlotGuard guard{p, slot};
    hipStream_t st = slot->stream;
    // The slot's previous call may have run on a caller's stream: wait for it before the
    // staging buffers and scratch are reused.
    HIP_TRY(hipStreamWaitEvent(st, slot->done, 0));
    guard.st = st;
    HIP_TRY(hipStreamSynchronize(st));
    rc = ensure_device(&slot->stage, &slot->stage_bytes, static_cast<size_t>(plane) * (T + 1));
    if (rc) {
        return rc;
    }
    char* stage = static_cast<char*>(slot->stage);
    char* h_stage = nullptr;
    if (pinned) {
        for (int t = 0; t < T; t++) {
            HIP_TRY(hipMemcpy2DAsync(stage + t * plane, row, src_frames[t], src_stride, row, height,
                                     hipMemcpyHostToDevice, st));
        }
    } else {
        rc = ensure_host(&slot->h_stage, &slot->h_stage_bytes, static_cast<size_t>(plane) * (T + 1));
        if (rc) {
            return rc;
        }
        h_stage = static_cast<char*>(slot->h_stage);
        for (int t = 0; t < T; t++) {
            const char* src = static_cast<const char*>(src_frames[t]);
            for (int y = 0; y < height; y++) {
                std::memcpy(h_stage + t * plane + y * row, src + y * src_stride, static_cast<size_t>(row));
            }
        }
        HIP_TRY(hipMemcpyAsync(stage, h_stage, static_cast<size_t>(plane) * T, hipMemcpyHostToDevice, st));
    }
    uj.src = stage;
    uj.src_stride = row;
    uj.src_frame_pitch = plane;
    uj.dst = stage + T * plane;
    uj.dst_stride = row;
    uj.dst_frame_pitch = 0;
    rc = make_job(p, uj, p->tile, j);
    if (rc) {
        return rc;
    }
    rc = run_jobs(p, slot, {j}, st);
    if (rc) {
        return rc;
    }
    if (pinned) {
        HIP_TRY(hipMemcpy2DAsync(dst, dst_stride, stage + T * plane, row, row, height, hipMemcpyDeviceToHost, st));
    } else {
        HIP_TRY(hipMemcpyAsync(h_stage + T * plane, stage + T * plane, static_cast<size_t>(plane),
                               hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipEventRecord(slot->done, st));
    guard.recorded = true;
    HIP_TRY(hipStreamSynchronize(st));
    if (!pinned) {
        char* out = static_cast<char*>(dst);
        for (int y = 0; y < height; y++) {
            std::memcpy(out + y * dst_stride, h_stage + T * plane + y * row, static_cast<size_t>(row));
        }
    }
    return DFT2_OK;
}

int dft2_process_frame_keyed(dft2_plan* p, const void* const* src_frames, const unsigned long long* src_keys,
                             long long src_stride, void* dst, long long dst_stride, int width, int height) {
    clear_error();
    if (!p || !src_frames || !src_keys || !dst) {
        return fail(DFT2_ERR_ARG, "plan, src_frames, src_keys and dst must not be NULL");
    }
    int rc = check_frame_call(p, src_frames, src_stride, dst_stride, width, height);
    if (rc) {
        return rc;
    }
    const int T = p->t_len;
    const long long row = static_cast<long long>(width) * p->bytes_per_sample;
    const long long plane = row * height;
    rc = select_device(p);
    if (rc) {
        return rc;
    }
    const bool pinned_dst = device_visible(dst);
    Slot* slot = acquire_slot(p);
    SlotGuard guard{p, slot};
    hipStream_t st = slot->stream;
    // The slot's previous call may have run on a caller's stream: wait for it before the
    // staging buffers and scratch are reused.
    HIP_TRY(hipStreamWaitEvent(st, slot->done, 0));
    guard.st = st;
    HIP_TRY(hipStreamSynchronize(st));
    rc = ensure_device(&slot->stage, &slot->stage_bytes, static_cast<size_t>(plane));
    if (rc) {
        return rc;
    }
    rc = ensure_host(&slot->h_stage, &slot->h_stage_bytes, static_cast<size_t>(plane) * (T + 1));
    if (rc) {
        return rc;
    }
    char* h_stage = static_cast<char*>(slot->h_stage);

    // Claim a device plane per source: one holding its key whose upload is ordered before
    // this call's work (queued on this call's stream, or already complete), or (a miss)
    // the least recently used plane no call uses, which this call then uploads.  A plane
    // another call is still uploading on another stream is not waited for: this call
    // uploads its own copy (waiting, on the host for the upload to be queued and on the
    // GPU across streams, made concurrent calls slower than uploading again; waiting on
    // the GPU only, for uploads already queued, was no faster).  The same
    // key twice (the temporal clamp at a clip's ends) shares one plane.  The pool holds
    // 3 (T + 1) planes per stream slot, so that a caller going through consecutive frames
    // of three planes (YUV) finds the T - 1 planes it uploaded for each plane's previous
    // frame; a call holds at most T, so a free plane always exists.  The pool's device
    // memory and events are made by the first call, at its plane size (a larger plane
    // later grows the plane it lands in), not by the calls that first use each plane.
    FramePlaneClaim claim{p};
    claim.n = T;
    claim.st = st;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        if (p->frame_planes.empty()) {
            p->frame_planes.resize(static_cast<size_t>(3 * (T + 1)) * p->slots.size());
            for (auto& fp : p->frame_planes) {
                rc = ensure_device(&fp.dev, &fp.cap, static_cast<size_t>(plane));
                if (rc) {
                    return rc;
                }
            }
        }
        for (int t = 0; t < T; t++) {
            for (int u = 0; u < t && !claim.use[t]; u++) {
                if (src_keys[u] == src_keys[t]) {
                    claim.use[t] = claim.use[u];
                    claim.mine[t] = claim.mine[u];
                }
            }
            if (claim.use[t]) {
                continue;
            }
            FramePlane* hit = nullptr;
            FramePlane* victim = nullptr;
            for (auto& fp : p->frame_planes) {
                if (fp.valid && !fp.stale && fp.key == src_keys[t] && fp.width == width && fp.height == height &&
                    fp.bps == p->bytes_per_sample && fp.recorded &&
                    (fp.up_stream == st || hipEventQuery(fp.ready) == hipSuccess)) {
                    hit = &fp;
                    break;
                }
                if (fp.users == 0 &&
                    (!victim || (victim->valid && !fp.valid) || (fp.valid == victim->valid && fp.last_use < victim->last_use))) {
                    victim = &fp;
                }
            }
            FramePlane* fp = hit ? hit : victim;
            if (!fp) {
                return fail(DFT2_ERR_STATE, "no free source plane in the frame cache");
            }
            if (!hit) {
                fp->key = src_keys[t];
                fp->valid = true;
                fp->stale = false;
                fp->recorded = false;
                fp->width = width;
                fp->height = height;
                fp->bps = p->bytes_per_sample;
                claim.mine[t] = true;
            }
            fp->users++;
            fp->last_use = ++p->frame_clock;
            claim.use[t] = fp;
        }
    }

    // Upload this call's misses: pinned planes directly, pageable ones through the slot's
    // pinned buffer (one part per plane, as dft2_process_frame stages them).
    int n_staged = 0;
    int uploads = 0;
    for (int t = 0; t < T; t++) {
        bool first = true;
        for (int u = 0; u < t; u++) {
            first = first && claim.use[u] != claim.use[t];
        }
        if (!claim.mine[t] || !first) {
            continue;
        }
        FramePlane& fp = *claim.use[t];
        rc = ensure_device(&fp.dev, &fp.cap, static_cast<size_t>(plane));
        if (rc) {
            return rc;
        }
        if (device_visible(src_frames[t])) {
            HIP_TRY(hipMemcpy2DAsync(fp.dev, row, src_frames[t], src_stride, row, height, hipMemcpyHostToDevice, st));
        } else {
            char* part = h_stage + n_staged * plane;
            const char* src = static_cast<const char*>(src_frames[t]);
            for (int y = 0; y < height; y++) {
                std::memcpy(part + y * row, src + y * src_stride, static_cast<size_t>(row));
            }
            HIP_TRY(hipMemcpyAsync(fp.dev, part, static_cast<size_t>(plane), hipMemcpyHostToDevice, st));
            n_staged++;
        }
        fp.ready = slot->up_ready;
        fp.up_stream = st;
        uploads++;
    }
    // One event after all of this call's uploads (one record per plane: 1080p YUV420
    // pinned planes from four threads 3,087-3,109 fps; one per call 3,256-3,471).
    if (uploads) {
        HIP_TRY(hipEventRecord(slot->up_ready, st));
    }
    // Publish them: later calls may share them from now on.
    {
        std::lock_guard<std::mutex> lk(p->mu);
        for (int t = 0; t < T; t++) {
            if (claim.mine[t]) {
                claim.use[t]->recorded = true;
            }
        }
    }

    // The filter on the device planes (as dft2_process_frame with device pointers), the
    // output staged through the slot's device buffer.
    dft2_job uj{};
    uj.src = claim.use[0]->dev;
    uj.dst = slot->stage;
    uj.width = width;
    uj.height = height;
    uj.n_frames = T;
    uj.frame_begin = p->radius;
    uj.frame_end = p->radius + 1;
    uj.src_stride = row;
    uj.dst_stride = row;
    JobDev j;
    rc = make_job(p, uj, p->tile, j);
    if (rc) {
        return rc;
    }
    j.use_fptr = 1;
    for (int t = 0; t < T; t++) {
        j.fptr[t] = claim.use[t]->dev;
    }
    j.dst_fp = 0;
    rc = run_jobs(p, slot, {j}, st);
    if (rc) {
        return rc;
    }
    if (pinned_dst) {
        HIP_TRY(hipMemcpy2DAsync(dst, dst_stride, slot->stage, row, row, height, hipMemcpyDeviceToHost, st));
    } else {
        HIP_TRY(hipMemcpyAsync(h_stage + T * plane, slot->stage, static_cast<size_t>(plane), hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipEventRecord(slot->done, st));
    guard.recorded = true;
    HIP_TRY(hipStreamSynchronize(st));
    claim.synced = true;
    if (!pinned_dst) {
        char* out = static_cast<char*>(dst);
        for (int y = 0; y < height; y++) {
            std::memcpy(out + y * dst_stride, h_stage + T * plane + y * row, static_cast<size_t>(row));
        }
    }
    return DFT2_OK;
}

int dft2_plan_forget_frames(dft2_plan* p) {
    clear_error();
    if (!p) {
        return fail(DFT2_ERR_ARG, "plan must not be NULL");
    }
    std::lock_guard<std::mutex> lk(p->mu);
    for (auto& fp : p->frame_planes) {
        fp.stale = true;
    }
    return DFT2_OK;
}

int dft2_rdft(const double* data, int data_len, const int* shape, int ndim, double* out, int out_len) {
    clear_error();
    // Same transform as RDFT (cpu_source/source.cpp:269-340): naive separable DFT in FP64,
    // real along the last axis, then complex along the others.
    if (!data || !shape || !out || ndim < 1 || ndim > 3) {
        return fail(DFT2_ERR_ARG, "\"shape\" must be an array of ints with 1, 2 or 3 values");
    }
    int dims[3] = {1, 1, 1};
    long long size = 1;
    for (int i = 0; i < ndim; i++) {
        if (shape[i] <= 0) {
            return fail(DFT2_ERR_ARG, "shape values must be positive");
        }
        dims[3 - ndim + i] = shape[i];
        size *= shape[i];
    }
    if (size != data_len) {
        return fail(DFT2_ERR_ARG, "cannot reshape array");
    }
    const int d0 = dims[0], d1 = dims[1], d2 = dims[2], h = d2 / 2 + 1;
    const long long n_out = static_cast<long long>(d0) * d1 * h;
    if (out_len < 2 * n_out) {
        return fail(DFT2_ERR_ARG, "output needs %lld doubles", 2 * n_out);
    }
    const double pi = 3.14159265358979323846;
    std::vector<double> re(n_out), im(n_out), tre(n_out), tim(n_out);
    for (int a = 0; a < d0; a++) {
        for (int b = 0; b < d1; b++) {
            const double* src = data + (static_cast<long long>(a) * d1 + b) * d2;
            for (int k = 0; k < h; k++) {
                double sr = 0, si = 0;
                for (int x = 0; x < d2; x++) {
                    double ang = -2 * pi * static_cast<double>(k) * x / d2;
                    sr += src[x] * std::cos(ang);
                    si += src[x] * std::sin(ang);
                }
                re[(a * d1 + b) * h + k] = sr;
                im[(a * d1 + b) * h + k] = si;
            }
        }
    }
    // axis 1 (length d1)
    for (int a = 0; a < d0; a++) {
        for (int k = 0; k < h; k++) {
            for (int u = 0; u < d1; u++) {
                double sr = 0, si = 0;
                for (int b = 0; b < d1; b++) {
                    double ang = -2 * pi * static_cast<double>(u) * b / d1;
                    double c = std::cos(ang), s = std::sin(ang);
                    double xr = re[(a * d1 + b) * h + k], xi = im[(a * d1 + b) * h + k];
                    sr += xr * c - xi * s;
                    si += xr * s + xi * c;
                }
                tre[(a * d1 + u) * h + k] = sr;
                tim[(a * d1 + u) * h + k] = si;
            }
        }
    }
    // axis 0 (length d0)
    for (int b = 0; b < d1; b++) {
        for (int k = 0; k < h; k++) {
            for (int u = 0; u < d0; u++) {
                double sr = 0, si = 0;
                for (int a = 0; a < d0; a++) {
                    double ang = -2 * pi * static_cast<double>(u) * a / d0;
                    double c = std::cos(ang), s = std::sin(ang);
                    double xr = tre[(a * d1 + b) * h + k], xi = tim[(a * d1 + b) * h + k];
                    sr += xr * c - xi * s;
                    si += xr * s + xi * c;
                }
                re[(u * d1 + b) * h + k] = sr;
                im[(u * d1 + b) * h + k] = si;
            }
        }
    }
    for (long long i = 0; i < n_out; i++) {
        out[2 * i] = re[i];
        out[2 * i + 1] = im[i];
    }
    return DFT2_OK;
}

} // extern "C"

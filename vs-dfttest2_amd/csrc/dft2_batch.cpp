// Batched launch of many clips (dft2_batch_* in include/dfttest2_mi355x.h).
//
// The reference filters one frame of one plane per GetFrame call, with a launch pair
// per call (hiprtc_source/source.cpp:575-640).  A batch instead describes a whole set of
// clips of any resolution and length once: the job table, the scratch regions and the
// work-queue counters are set up when the batch is created, and the launch sequence
// (counter reset, then fused + finalize per frame chunk) is captured into a hipGraph, so
// that running the batch is one hipGraphLaunch.
//
// The tile shape (blocks per workgroup, which sets the LDS footprint and the halo
// overhead) and the launch form (one work item per workgroup, or a persistent grid of
// resident workgroups that take work items from an atomic counter) are chosen by an
// autotuner: every valid candidate is timed on a sample of the batch (the first output
// frames of every job) and the fastest is kept.

#include <algorithm>
#include <cstring>
#include <vector>

#include "dft2_host.h"

using namespace dft2;
using namespace dft2::host;

struct dft2_batch {
    dft2_plan* plan = nullptr;
    std::vector<dft2_job> jobs;
    TileShape tile;
    int persistent = 0;
    int wg_per_cu = 0;
    int n_cu = 0;
    ChunkPlan cp;
    JobDev* d_table = nullptr;
    size_t d_table_cap = 0;
    float* scratch = nullptr;
    size_t scratch_cap = 0;
    unsigned* d_queue = nullptr;
    size_t d_queue_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float tuned_ms = 0.0f;
    std::vector<dft2_batch_candidate> candidates;
};

namespace dft2 {
namespace host {
namespace {

// Dynamic LDS a fused workgroup may use: 160 KiB per CU less room for the kernel's
// static LDS (the work-queue slot).
constexpr size_t kMaxLds = 160 * 1024 - 256;

// Sets up job table, scratch and counters of `b` for tile `ts`; with sample_frames > 0
// only the first sample_frames output frames of every job are included.
int build_work(dft2_batch* b, const TileShape& ts, int sample_frames) {
    const dft2_plan* p = b->plan;
    std::vector<JobDev> jd(b->jobs.size());
    for (size_t i = 0; i < b->jobs.size(); i++) {
        int rc = make_job(p, b->jobs[i], ts, jd[i]);
        if (rc) {
            return rc;
        }
        if (sample_frames > 0) {
            jd[i].f_count = std::min(jd[i].f_count, sample_frames);
        }
    }
    int rc = plan_chunks(p, ts, jd, p->scratch_cap, b->cp);
    if (rc) {
        return rc;
    }
    if (b->cp.chunks.empty()) {
        return DFT2_OK;
    }
    rc = ensure_device(reinterpret_cast<void**>(&b->scratch), &b->scratch_cap, b->cp.scratch_bytes);
    if (rc) {
        return rc;
    }
    const size_t table_bytes = sizeof(JobDev) * b->cp.table.size();
    rc = ensure_device(reinterpret_cast<void**>(&b->d_table), &b->d_table_cap, table_bytes);
    if (rc) {
        return rc;
    }
    HIP_TRY(hipMemcpyAsync(b->d_table, b->cp.table.data(), table_bytes, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    rc = ensure_device(reinterpret_cast<void**>(&b->d_queue), &b->d_queue_cap,
                       sizeof(unsigned) * b->cp.chunks.size());
    if (rc) {
        return rc;
    }
    b->tile = ts;
    return DFT2_OK;
}

// Queues the whole batch on `stream` (the sequence the graph captures).
int enqueue(dft2_batch* b, int persistent, hipStream_t stream) {
    const dft2_plan* p = b->plan;
    const ChunkPlan& cp = b->cp;
    if (cp.chunks.empty()) {
        return DFT2_OK;
    }
    if (persistent) {
        HIP_TRY(launch_queue_reset(b->d_queue, static_cast<int>(cp.chunks.size()), stream));
    }
    LaunchArgs a = make_args(p, b->tile);
    a.n_jobs = cp.n_jobs;
    a.scratch = b->scratch;
    const FusedKey key = fused_key(p, persistent);
    const long long resident = static_cast<long long>(b->wg_per_cu) * b->n_cu;
    for (size_t c = 0; c < cp.chunks.size(); c++) {
        const Chunk& ch = cp.chunks[c];
        if (ch.items == 0) {
            continue;
        }
        a.jobs = b->d_table + c * cp.n_jobs;
        unsigned grid = static_cast<unsigned>(ch.items);
        if (persistent) {
            a.queue = b->d_queue + c;
            a.n_items = static_cast<int>(ch.items);
            grid = static_cast<unsigned>(std::min(ch.items, resident));
        } else {
            a.queue = nullptr;
        }
        HIP_TRY(launch_fused(a, key, grid, b->tile.lds, stream));
        HIP_TRY(launch_finalize(a, p->sample_code, ch.items, stream));
    }
    return DFT2_OK;
}

// Average time of `reps` runs of the current work setup (after one untimed run).
int time_runs(dft2_batch* b, int persistent, int reps, float* ms) {
    int rc = enqueue(b, persistent, b->stream);
    if (rc) {
        return rc;
    }
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    for (int i = 0; i < reps; i++) {
        rc = enqueue(b, persistent, b->stream);
        if (rc) {
            return rc;
        }
    }
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipEventSynchronize(b->ev1));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = t / reps;
    return DFT2_OK;
}

int autotune(dft2_batch* b, int sample_frames) {
    dft2_plan* p = b->plan;
    // Widths in whole passes of kWaves blocks (every wave busy), heights from tall to
    // short; the plan's own choice is among them.
    static const int kWidths[] = {8, 16};
    static const int kHeights[] = {24, 16, 12, 8, 4};
    float best = -1.0f;
    TileShape best_ts = p->tile;
    int best_persistent = 0;
    for (int bh : kHeights) {
        for (int bw : kWidths) {
            if (!tile_valid(p, bw, bh)) {
                continue;
            }
            const TileShape ts = make_tile(p, bw, bh);
            if (ts.lds > kMaxLds) {
                continue;
            }
            if (fused_workgroups_per_cu(fused_key(p), ts.lds) <= 0) {
                continue;
            }
            int rc = build_work(b, ts, sample_frames);
            if (rc) {
                return rc;
            }
            for (int persistent = 0; persistent < 2; persistent++) {
                const int wg = fused_workgroups_per_cu(fused_key(p, persistent), ts.lds);
                if (wg <= 0) {
                    continue;
                }
                b->wg_per_cu = wg;
                float ms = 0.0f;
                rc = time_runs(b, persistent, 3, &ms);
                if (rc) {
                    return rc;
                }
                dft2_batch_candidate c{};
                c.tile_w = bw;
                c.tile_h = bh;
                c.persistent = persistent;
                c.workgroups_per_cu = wg;
                c.lds_bytes = static_cast<int>(ts.lds);
                c.ms = ms;
                b->candidates.push_back(c);
                if (best < 0.0f || ms < best) {
                    best = ms;
                    best_ts = ts;
                    best_persistent = persistent;
                }
            }
        }
    }
    if (best < 0.0f) {
        return fail(DFT2_ERR_STATE, "autotuner found no tile shape that fits this plan");
    }
    b->tile = best_ts;
    b->persistent = best_persistent;
    b->wg_per_cu = fused_workgroups_per_cu(fused_key(p, best_persistent), best_ts.lds);
    b->tuned_ms = best;
    return DFT2_OK;
}

void release(dft2_batch* b) {
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
    }
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
    }
    if (b->scratch) {
        (void)hipFree(b->scratch);
    }
    if (b->d_table) {
        (void)hipFree(b->d_table);
    }
    if (b->d_queue) {
        (void)hipFree(b->d_queue);
    }
    if (b->ev0) {
        (void)hipEventDestroy(b->ev0);
    }
    if (b->ev1) {
        (void)hipEventDestroy(b->ev1);
    }
    if (b->stream) {
        (void)hipStreamDestroy(b->stream);
    }
    delete b;
}

int create(dft2_plan* p, const dft2_job* jobs, int n_jobs, const dft2_batch_options* opt, dft2_batch* b) {
    b->plan = p;
    b->jobs.assign(jobs, jobs + n_jobs);
    int rc = select_device(p);
    if (rc) {
        return rc;
    }
    HIP_TRY(hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, p->device));
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&b->ev0));
    HIP_TRY(hipEventCreate(&b->ev1));
    // Validate every job up front (also for an empty batch).
    for (int i = 0; i < n_jobs; i++) {
        JobDev j;
        rc = make_job(p, jobs[i], p->tile, j);
        if (rc) {
            return rc;
        }
    }
    bool has_work = false;
    for (int i = 0; i < n_jobs; i++) {
        has_work = has_work || jobs[i].frame_end > jobs[i].frame_begin;
    }
    if (opt->autotune && has_work) {
        rc = autotune(b, opt->sample_frames > 0 ? opt->sample_frames : 2);
        if (rc) {
            return rc;
        }
    } else {
        const int bw = opt->tile_w > 0 ? opt->tile_w : p->tile.bw;
        const int bh = opt->tile_h > 0 ? opt->tile_h : p->tile.bh;
        if (!tile_valid(p, bw, bh)) {
            return fail(DFT2_ERR_ARG, "tile %dx%d does not cover every pixel with at most 2 tiles per axis "
                        "at block_step %d", bw, bh, p->block_step);
        }
        b->tile = make_tile(p, bw, bh);
        if (b->tile.lds > kMaxLds) {
            return fail(DFT2_ERR_ARG, "tile %dx%d needs %zu bytes of LDS (at most %zu)", bw, bh, b->tile.lds, kMaxLds);
        }
        b->persistent = opt->persistent ? 1 : 0;
        if (b->persistent && !p->separable) {
            return fail(DFT2_ERR_ARG, "the persistent launch form needs a separable window (a[t] b[y] c[x])");
        }
        b->wg_per_cu = fused_workgroups_per_cu(fused_key(p, b->persistent), b->tile.lds);
        if (b->wg_per_cu <= 0) {
            return fail(DFT2_ERR_ARG, "tile %dx%d does not fit on a CU", bw, bh);
        }
    }
    rc = build_work(b, b->tile, 0);
    if (rc) {
        return rc;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->cp.chunks.empty()) {
        return DFT2_OK; // nothing to run: no graph
    }
    // Capture the launch sequence.  The counters are zeroed by a kernel node, not a
    // memset node.
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    rc = enqueue(b, b->persistent, b->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) {
            (void)hipGraphDestroy(g);
        }
        return rc;
    }
    HIP_TRY(e);
    b->graph = g;
    HIP_TRY(hipGraphInstantiateWithFlags(&b->exec, b->graph, 0));
    return DFT2_OK;
}

} // namespace
} // namespace host
} // namespace dft2

extern "C" {

int dft2_batch_create(dft2_plan* plan, const dft2_job* jobs, int n_jobs, const dft2_batch_options* options,
                      dft2_batch** out_batch) {
    clear_error();
    if (!plan || !out_batch) {
        return fail(DFT2_ERR_ARG, "plan and out_batch must not be NULL");
    }
    *out_batch = nullptr;
    if (plan->generic) {
        return fail(DFT2_ERR_ARG, "batches need block_size 16 and radius <= 3 (the fused kernel)");
    }
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) {
        return fail(DFT2_ERR_ARG, "invalid job list");
    }
    dft2_batch_options opt{};
    opt.autotune = 1;
    if (options) {
        opt = *options;
    }
    if (opt.tile_w < 0 || opt.tile_h < 0 || opt.sample_frames < 0) {
        return fail(DFT2_ERR_ARG, "tile sizes and sample_frames must not be negative");
    }
    auto* b = new dft2_batch();
    int rc = create(plan, jobs, n_jobs, &opt, b);
    if (rc) {
        release(b);
        return rc;
    }
    *out_batch = b;
    return DFT2_OK;
}

int dft2_batch_launch(dft2_batch* b, void* stream) {
    clear_error();
    if (!b) {
        return fail(DFT2_ERR_ARG, "batch must not be NULL");
    }
    int rc = select_device(b->plan);
    if (rc) {
        return rc;
    }
    if (!b->exec) {
        return DFT2_OK; // nothing to do
    }
    hipStream_t st = call_stream(stream, b->stream);
    HIP_TRY(hipGraphLaunch(b->exec, st));
    if (!stream) {
        HIP_TRY(hipStreamSynchronize(st));
    }
    return DFT2_OK;
}

int dft2_batch_get_info(const dft2_batch* b, dft2_batch_info* info) {
    clear_error();
    if (!b || !info) {
        return fail(DFT2_ERR_ARG, "batch and info must not be NULL");
    }
    std::memset(info, 0, sizeof(*info));
    info->tile_w = b->tile.bw;
    info->tile_h = b->tile.bh;
    info->persistent = b->persistent;
    info->workgroups_per_cu = b->wg_per_cu;
    info->compute_units = b->n_cu;
    info->lds_bytes = static_cast<int>(b->tile.lds);
    info->n_chunks = static_cast<int>(b->cp.chunks.size());
    long long items = 0;
    for (const auto& c : b->cp.chunks) {
        items += c.items;
    }
    info->work_items = items;
    info->scratch_bytes = static_cast<long long>(b->cp.scratch_bytes);
    info->n_candidates = static_cast<int>(b->candidates.size());
    info->tuned_ms = b->tuned_ms;
    return DFT2_OK;
}

int dft2_batch_get_candidate(const dft2_batch* b, int index, dft2_batch_candidate* out) {
    clear_error();
    if (!b || !out) {
        return fail(DFT2_ERR_ARG, "batch and out must not be NULL");
    }
    if (index < 0 || index >= static_cast<int>(b->candidates.size())) {
        return fail(DFT2_ERR_ARG, "candidate index %d out of range [0, %zu)", index, b->candidates.size());
    }
    *out = b->candidates[index];
    return DFT2_OK;
}

int dft2_batch_destroy(dft2_batch* b) {
    if (!b) {
        return DFT2_OK;
    }
    (void)hipSetDevice(b->plan->device);
    release(b);
    return DFT2_OK;
}

} // extern "C"

// dft2_process_batch_noalloc: dft2_process_batch without any allocation, for callers
// that cannot allocate device memory (code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// The working memory of a launch (the job table and the tile regions the fused kernel
// writes and the finalize kernel reads) comes from `scratch`, a file-scope device array
// of kNoAllocScratchBytes that every GPU running this library holds for the life of the
// process.  Its device address is looked up once per device (hipGetSymbolAddress) and
// passed to the kernels; the array's name is never used as an address on the host.
// The job table is written into it by a one-workgroup kernel that receives the rows as
// kernel arguments, so no host staging buffer is needed either.  Frames are split into
// chunks whose tile regions fit the array; the chunks run one after another on the
// caller's stream and reuse it.
//
// Every call on a device shares that device's array: calls that overlap on one GPU (on
// different streams or threads) race and must be serialised by the caller.  Nothing in
// the array needs to start at zero: the fused kernel writes every region it covers
// before finalize reads it, and each chunk's table rows are written before its launch.

#include <algorithm>
#include <cstring>
#include <mutex>

#include "dft2_host.h"

using namespace dft2;
using namespace dft2::host;

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxNoAllocJobs = 64;                         // jobs (planes) per call
constexpr size_t kTableBytes = (kMaxNoAllocJobs * sizeof(JobDev) + 255) / 256 * 256; // job table at the front
constexpr int kRowsPerWrite = 8;                            // rows per table-write launch
constexpr int kMaxDevices = 64;

static_assert(kRowsPerWrite * sizeof(JobDev) + 64 <= 4096, "kernel argument block too large");

} // namespace

// The working memory of the allocation-free entry on each device (see above).
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Per device: the address of that device's `scratch` (null until looked up).
void* g_scratch_addr[kMaxDevices] = {};
std::mutex g_scratch_mu;

struct TableRows {
    JobDev rows[kRowsPerWrite];
    JobDev* dst;
    int n;
};

__global__ void table_write_kernel(TableRows t) {
    // One row per wave, 32-bit words across its lanes: vector stores only.
    const int row = threadIdx.x / 64;
    const int word = threadIdx.x % 64;
    constexpr int kWords = static_cast<int>(sizeof(JobDev) / 4);
    if (row < t.n) {
        const unsigned* src = reinterpret_cast<const unsigned*>(&t.rows[row]);
        unsigned* dst = reinterpret_cast<unsigned*>(t.dst + row);
        for (int w = word; w < kWords; w += 64) {
            dst[w] = src[w];
        }
    }
}
static_assert(sizeof(JobDev) % 4 == 0, "JobDev is copied as 32-bit words");

// Address of the current device's `scratch`, looked up on the device's first call.
int scratch_address(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) {
        return fail(DFT2_ERR_ARG, "device %d is beyond the %d devices the allocation-free entry supports", device,
                    kMaxDevices);
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) {
            return fail(DFT2_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        }
        g_scratch_addr[device] = ptr;
    }
    *out = static_cast<unsigned char*>(g_scratch_addr[device]);
    return DFT2_OK;
}

int write_rows(JobDev* table, const JobDev* rows, int n, hipStream_t stream) {
    for (int i = 0; i < n; i += kRowsPerWrite) {
        TableRows t{};
        t.n = std::min(kRowsPerWrite, n - i);
        std::memcpy(t.rows, rows + i, sizeof(JobDev) * t.n);
        t.dst = table + i;
        hipLaunchKernelGGL(table_write_kernel, dim3(1), dim3(64 * kRowsPerWrite), 0, stream, t);
        HIP_TRY(hipGetLastError());
    }
    return DFT2_OK;
}

} // namespace

extern "C" {

int dft2_process_batch_noalloc(dft2_plan* p, const dft2_job* jobs, int n_jobs, void* stream) {
    clear_error();
    if (!p) {
        return fail(DFT2_ERR_ARG, "plan must not be NULL");
    }
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) {
        return fail(DFT2_ERR_ARG, "invalid job list");
    }
    if (p->generic) {
        return fail(DFT2_ERR_ARG, "the allocation-free entry needs block_size 16 and radius <= 3 (the fused kernel)");
    }
    if (n_jobs > kMaxNoAllocJobs) {
        return fail(DFT2_ERR_ARG, "the allocation-free entry takes at most %d jobs (got %d)", kMaxNoAllocJobs,
                    n_jobs);
    }
    JobDev jd[kMaxNoAllocJobs];
    // The plan's form, as dft2_process_batch takes it: one frame per work item, or the
    // cross-frame form (reuse_k frames per item, its own tile) for calls with more than
    // one output frame in some job.
    int call_frames = 0;
    for (int i = 0; i < n_jobs; i++) {
        call_frames = std::max(call_frames, jobs[i].frame_end - jobs[i].frame_begin);
    }
    const int reuse_k = call_frames <= 1 ? 1 : p->reuse_k;
    const TileShape& ts = reuse_k > 1 ? p->reuse_tile : p->tile;
    long long per_frame_items = 0;
    int max_frames = 0;
    for (int i = 0; i < n_jobs; i++) {
        int rc = make_job(p, jobs[i], ts, jd[i]);
        if (rc) {
            return rc;
        }
        per_frame_items += static_cast<long long>(jd[i].tiles_x) * jd[i].tiles_y;
        max_frames = std::max(max_frames, jd[i].f_count);
    }
    if (per_frame_items == 0 || max_frames == 0) {
        return DFT2_OK;
    }
    const size_t region_bytes = sizeof(float) * static_cast<size_t>(ts.region_w) * ts.region_h;
    const size_t region_cap = kNoAllocScratchBytes - kTableBytes;
    const size_t frame_bytes = static_cast<size_t>(per_frame_items) * region_bytes;
    if (frame_bytes > region_cap) {
        return fail(DFT2_ERR_ARG,
                    "one frame of these jobs needs %.1f MiB of tile scratch; the allocation-free entry holds %.1f MiB",
                    frame_bytes / 1048576.0, region_cap / 1048576.0);
    }
    int frames_per_chunk = static_cast<int>(std::min<size_t>(region_cap / frame_bytes, max_frames));
    // whole groups of reuse_k frames per chunk where a chunk holds more than one group
    if (frames_per_chunk > reuse_k && frames_per_chunk < max_frames) {
        frames_per_chunk -= frames_per_chunk % reuse_k;
    }
    int rc = select_device(p);
    if (rc) {
        return rc;
    }
    unsigned char* base = nullptr;
    rc = scratch_address(p->device, &base);
    if (rc) {
        return rc;
    }
    JobDev* table = reinterpret_cast<JobDev*>(base);
    // NULL: the plan's first stream slot, and the call waits (as dft2_process_batch).
    hipStream_t st = call_stream(stream, p->slots[0].stream);

    LaunchArgs a = make_args(p, ts);
    a.n_jobs = n_jobs;
    a.jobs = table;
    a.scratch = reinterpret_cast<float*>(base + kTableBytes);
    FusedKey key = fused_key(p, 0, a.asym);
    if (reuse_k > 1) {
        a.reuse_k = reuse_k;
        key.reuse = 1;
    }
    JobDev chunk[kMaxNoAllocJobs];
    for (int f0 = 0; f0 < max_frames; f0 += frames_per_chunk) {
        long long items = 0, frames = 0, groups = 0;
        for (int k = 0; k < n_jobs; k++) {
            JobDev j = jd[k];
            const int cnt = std::max(0, std::min(frames_per_chunk, j.f_count - f0));
            j.f_begin += f0;
            j.f_count = cnt;
            j.item_begin = items;
            j.frame_begin = frames;
            j.group_begin = groups;
            items += static_cast<long long>(cnt) * j.tiles_x * j.tiles_y;
            frames += cnt;
            groups += static_cast<long long>((cnt + reuse_k - 1) / reuse_k) * j.tiles_x * j.tiles_y;
            chunk[k] = j;
        }
        if (items == 0) {
            continue;
        }
        // Stream order: this chunk's rows overwrite the table only after the previous
        // chunk's kernels (queued earlier on the same stream) have finished with it.
        rc = write_rows(table, chunk, n_jobs, st);
        if (rc) {
            return rc;
        }
        HIP_TRY(launch_fused(a, key, static_cast<unsigned>(reuse_k > 1 ? groups : items), ts.lds, st));
        HIP_TRY(launch_finalize(a, p->sample_code, items, st));
    }
    if (!stream) {
        HIP_TRY(hipStreamSynchronize(st));
    }
    return DFT2_OK;
}

} // extern "C"

// Host-side internals shared by the C ABI translation units (dft2_capi.cpp,
// dft2_batch.cpp): the plan object, tile shapes, job validation and chunk planning.
#pragma once

#include <array>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dfttest2_mi355x.h"
#include "dft2_internal.h"

namespace dft2 {
namespace host {

// Records the message for dft2_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);
void clear_error();
const char* last_error();

#define HIP_TRY(expr)                                                                                          \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            return ::dft2::host::fail(DFT2_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                                      __FILE__, __LINE__);                                                     \
        }                                                                                                      \
    } while (0)

// cpu_source/getframe_impl.cpp:11-20
int pad_size(int size, int bs, int step);
int pad_num(int size, int bs, int step);

constexpr int kRing = 8;
// Generic block path limits: block sizes 1..64 and radius up to 7, with at most this much
// LDS per block (one wave's two complex T x B x (B/2+1) buffers).
constexpr int kMaxGenericBlock = 64;
constexpr int kMaxGenericRadius = 7;
static_assert(2 * kMaxGenericRadius + 1 <= kMaxFramePtrs, "a job's frame pointers must cover T source frames");
constexpr size_t kGenericLdsBytes = 160 * 1024 - 1024;

// Offsets (floats) of the generic path's table: twiddles, window, synthesis window, sigma,
// window spectrum (GenericArgs in dft2_internal.h).
struct GenericLayout {
    int cos_b, sin_b, cos_t, sin_t, win, syn, sig, wf, total;
};
inline GenericLayout generic_layout(int B, int T) {
    const int H = B / 2 + 1;
    auto up4 = [](int v) { return (v + 3) / 4 * 4; };
    GenericLayout L{};
    L.cos_b = 0;
    L.sin_b = up4(L.cos_b + B);
    L.cos_t = up4(L.sin_b + B);
    L.sin_t = up4(L.cos_t + T);
    L.win = up4(L.sin_t + T);
    L.syn = up4(L.win + T * B * B);
    L.sig = up4(L.syn + B * B);
    L.wf = up4(L.sig + T * B * H);
    L.total = up4(L.wf + 2 * T * B * H);
    return L;
}
constexpr int kTableCache = 4;
constexpr size_t kTimingPending = 64; // launch pairs whose timing events may be pending

// A job table resident in device memory, with the host bytes it was made from.
struct DeviceTable {
    std::vector<unsigned char> bytes;
    JobDev* dev = nullptr;
    size_t cap = 0;
    unsigned long long last_use = 0;
};

// One concurrent call's stream and buffers (num_streams of them per plan).
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // dft2_process_frame_keyed: recorded after a call's source-plane uploads; the planes
    // that call uploaded point at it (a later call's record only delays when they count
    // as complete, since the slot's stream runs its calls in order)
    hipEvent_t up_ready = nullptr;
    float* scratch = nullptr;
    size_t scratch_bytes = 0;
    // Recently uploaded job tables: a call whose table matches one of them byte for
    // byte (the same buffers and geometry as an earlier call) reuses it without a copy.
    DeviceTable tables[kTableCache];
    unsigned long long table_clock = 0;
    JobDev* h_jobs[kRing] = {};
    size_t h_jobs_cap[kRing] = {};
    hipEvent_t h_jobs_done[kRing] = {};
    int ring_pos = 0;
    float4* tab_dev = nullptr; // gain table built per launch when sigma comes from a device array
    size_t tab_dev_cap = 0;
    void* stage = nullptr; // host-memory path: T input planes + output plane
    size_t stage_bytes = 0;
    void* h_stage = nullptr; // host-memory path, pageable planes: pinned copy of the same
    size_t h_stage_bytes = 0;
    bool busy = false;
};

// One device copy of a host source plane for dft2_process_frame_keyed, shared by the
// plan's stream slots (guarded by the plan's mutex).
struct FramePlane {
    unsigned long long key = 0;
    bool valid = false;      // holds (or is being given) the plane named by key, of the geometry below
    bool stale = false;      // forgotten (dft2_plan_forget_frames): no later call matches it
    bool recorded = false;   // its upload is queued and `ready` recorded: other calls may use it
    int width = 0, height = 0, bps = 0;
    int users = 0;           // calls that use the plane now; 0: it may be replaced
    unsigned long long last_use = 0;
    void* dev = nullptr;     // width * height * bps bytes, rows packed
    size_t cap = 0;
    hipEvent_t ready = nullptr; // the uploading slot's up_ready (not owned), recorded after the upload
    hipStream_t up_stream = nullptr; // that stream: calls on it are ordered after the upload
};

// Tile of blocks owned by one workgroup of the fused kernel and its LDS footprint.
struct TileShape {
    int bw = 16, bh = 16;           // blocks per tile (x, y)
    int region_w = 0, region_h = 0; // accumulator region of a full tile (samples)
    size_t lds = 0;                 // dynamic LDS bytes of the fused kernel
};

// Work of one fused + finalize launch pair.
struct Chunk {
    long long items = 0;  // work items (tiles x frames)
    long long frames = 0; // job-frames (finalize grid y)
    long long groups = 0; // cross-frame form: work items (tiles x groups of reuse_k frames)
    int max_pixels = 0;   // largest plane among jobs with work (finalize grid x)
};

// Job table rows [c * n_jobs, (c+1) * n_jobs) describe chunk c.
struct ChunkPlan {
    int n_jobs = 0;
    std::vector<JobDev> table;
    std::vector<Chunk> chunks;
    size_t scratch_bytes = 0;
};

} // namespace host
} // namespace dft2

struct dft2_plan {
    int device = 0;
    int generic = 0;      // block_size != 16 or radius > 3: the generic block path
    int block_size = 16;
    int gen_waves = 1;    // generic path: waves (blocks) per workgroup, limited by LDS
    int radius = 1;
    int t_len = 3;
    int block_step = 4;
    int filter_type = 0;
    int zero_mean = 0;
    int asym = 0;
    int separable = 0; // window = a[t] b[y] c[x] (see kOffSWX)
    int sample_code = dft2::kU8;
    int precision = DFT2_PRECISION_FP32;
    int bits = 8;
    int bytes_per_sample = 1;
    float out_scale = 1.0f;
    int peak = 255;
    float sigma2 = 0, pmin = 0, pmax = 0, wf0 = 1.0f;
    dft2::host::TileShape tile; // default forward tile
    // Cross-frame form (process_batch and frame calls without decision records): output
    // frames per work item (1 = the one-frame form) and its tile.
    int reuse_k = 1;
    dft2::host::TileShape reuse_tile;
    size_t scratch_cap = size_t(4) << 30;
    float* d_consts = nullptr;
    std::vector<dft2::host::Slot> slots;
    std::mutex mu;
    std::condition_variable cv;
    // Kernel timing (dft2_plan_set_timing): events before the fused launch, between the
    // fused and finalize launches and after the finalize launch, one triple per launch
    // pair.  A triple is taken from the free pool before the launch and committed to the
    // pending queue once recorded; at most kTimingPending triples stay pending (older ones
    // are waited for and added to the totals), so a long run holds a bounded number of
    // events.  Guarded by mu.
    bool timing = false;
    std::vector<hipEvent_t> timing_free;
    struct TimingRecord {
        std::array<hipEvent_t, 3> ev;
        long long blocks, slices; // work of the launch pair (dft2_kernel_timing)
    };
    std::deque<TimingRecord> timing_pending;
    dft2_kernel_timing timing_totals{};
    // Source planes of dft2_process_frame_keyed (2 * T * num_streams, made on first use),
    // guarded by mu; frame_clock orders their uses for replacement.
    std::vector<dft2::host::FramePlane> frame_planes;
    unsigned long long frame_clock = 0;
};

namespace dft2 {
namespace host {

// Decision-record masks per block of a generic plan: T x B x (B/2+1) coefficients.
inline int generic_dec_words(const dft2_plan* p) {
    return (p->t_len * p->block_size * (p->block_size / 2 + 1) + 63) / 64;
}

// Kernel selection for plan p; asym = -1 takes the plan's own sigma symmetry.
inline FusedKey fused_key(const dft2_plan* p, int persistent = 0, int asym = -1) {
    return FusedKey{p->radius, p->sample_code, p->precision, p->filter_type, persistent,
                    asym < 0 ? p->asym : asym, p->separable ? 1 : 0};
}

// Forward tile of bw x bh blocks for plan p.  Valid when every pixel is covered by the
// regions of at most 2 tiles per axis (the finalize kernel gathers at most 4).
TileShape make_tile(const dft2_plan* p, int bw, int bh);
bool tile_valid(const dft2_plan* p, int bw, int bh);
// Default forward tile: the candidate with the most blocks in flight per CU (resident
// workgroups from the occupancy API x blocks per tile); needs the plan's device current.
TileShape choose_tile(const dft2_plan* p);
// Sets the plan's cross-frame form (reuse_k, reuse_tile); needs the device current.
void choose_reuse(dft2_plan* p);

// Launch arguments common to every kernel of a plan, for tile `ts`.
LaunchArgs make_args(const dft2_plan* p, const TileShape& ts);

// Validates one user job and fills its geometry (tiles counted for `ts`).
int make_job(const dft2_plan* p, const dft2_job& uj, const TileShape& ts, JobDev& j);

// Splits jobs into frame chunks whose scratch fits `scratch_cap` bytes.
// reuse_k > 1: also numbers the cross-frame form's work items (groups of reuse_k frames).
int plan_chunks(const dft2_plan* p, const TileShape& ts, const std::vector<JobDev>& jobs, size_t scratch_cap,
                ChunkPlan& out, int reuse_k = 1);

int ensure_device(void** ptr, size_t* cap, size_t bytes);

// The stream a call runs on: the caller's `stream`, HIP's null stream for
// DFT2_DEFAULT_STREAM (a value HIP calls do not all accept), or `own` for NULL (then the
// call also waits for completion).
inline hipStream_t call_stream(void* stream, hipStream_t own) {
    if (!stream) {
        return own;
    }
    return stream == DFT2_DEFAULT_STREAM ? nullptr : static_cast<hipStream_t>(stream);
}
int select_device(const dft2_plan* p);

} // namespace host
} // namespace dft2

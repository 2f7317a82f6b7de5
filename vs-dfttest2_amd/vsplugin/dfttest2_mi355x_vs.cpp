// VapourSynth plugin "dfttest2_mi355x": the reference plugin's DFTTest, RDFT and Version
// functions over the dfttest2_mi355x C ABI (include/dfttest2_mi355x.h).
//
// Registration and argument strings follow the reference HIPRTC plugin
// (hiprtc_source/source.cpp:1013-1059), so that the reference Python front end calls it
// with the same keyword arguments (python/dfttest2.py:381-399); DFTTest also takes an
// optional "precision" (0 = FP32, 1 = split BF16, DFT2_PRECISION_*).
//   DFTTestCreate (hiprtc_source/source.cpp:710-934)  -> Create: arguments, defaults and
//       checks, then dft2_plan_create (which repeats the checks with the same messages);
//   DFTTestGetFrame (:445-692)                        -> GetFrame: requests the frames
//       n - r .. n + r, clamps them at the clip ends (cpu_source/getframe_impl.cpp:294-298),
//       copies unprocessed planes from the centre frame, and filters each processed plane
//       with dft2_process_frame_keyed on the frames' host planes, keyed by (frame number,
//       plane): frames of a node are immutable, so each source plane is uploaded once for
//       the T output frames that read it (the reference uploads T planes per call);
//   DFTTestFree (:694-707)                            -> Free;
//   RDFT (:936-1007), Version                         -> dft2_rdft, dft2_version.
// fmParallel: VapourSynth calls GetFrame from several threads; the plan admits
// num_streams calls at once (one stream slot each) and queues the others.
//
// Built by vs-dfttest2_amd/build_native.py when VapourSynth4.h is found (VS_INCLUDE or the
// usual include directories), linked against libdfttest2_mi355x.so.

#include <VSHelper4.h>
#include <VapourSynth4.h>

#include <algorithm>
#include <vector>

#include "../../include/dfttest2_mi355x.h"

namespace {

// Source frames of one output frame at the largest radius dft2_plan_create accepts (7: the
// generic block path), T = 2 * 7 + 1.
constexpr int kMaxSlices = 15;

struct DFTTestData {
    VSNode* node = nullptr;
    int radius = 0;
    bool process[3] = {};
    dft2_plan* plan = nullptr;
};

const VSFrame* VS_CC GetFrame(int n, int activation_reason, void* instance_data, void**, VSFrameContext* frame_ctx,
                              VSCore* core, const VSAPI* vsapi) noexcept {
    auto* d = static_cast<DFTTestData*>(instance_data);
    const VSVideoInfo* vi = vsapi->getVideoInfo(d->node);
    if (activation_reason == arInitial) {
        for (int i = std::max(n - d->radius, 0); i <= std::min(n + d->radius, vi->numFrames - 1); i++) {
            vsapi->requestFrameFilter(i, d->node, frame_ctx);
        }
        return nullptr;
    }
    if (activation_reason != arAllFramesReady) {
        return nullptr;
    }
    const int t_len = 2 * d->radius + 1; // <= kMaxSlices (Create keeps radius <= 7)
    const VSFrame* src[kMaxSlices] = {};
    int src_n[kMaxSlices] = {};
    for (int t = 0; t < t_len; t++) {
        src_n[t] = std::clamp(n - d->radius + t, 0, vi->numFrames - 1);
        src[t] = vsapi->getFrameFilter(src_n[t], d->node, frame_ctx);
    }
    const VSFrame* centre = src[d->radius];
    const VSFrame* copy_from[3] = {d->process[0] ? nullptr : centre, d->process[1] ? nullptr : centre,
                                   d->process[2] ? nullptr : centre};
    const int copy_planes[3] = {0, 1, 2};
    VSFrame* dst = vsapi->newVideoFrame2(&vi->format, vi->width, vi->height, copy_from, copy_planes, centre, core);
    const char* error = nullptr;
    for (int plane = 0; plane < vi->format.numPlanes && !error; plane++) {
        if (!d->process[plane]) {
            continue;
        }
        const void* planes[kMaxSlices] = {};
        unsigned long long keys[kMaxSlices] = {};
        for (int t = 0; t < t_len; t++) {
            planes[t] = vsapi->getReadPtr(src[t], plane);
            keys[t] = static_cast<unsigned long long>(src_n[t]) * 4 + static_cast<unsigned long long>(plane);
        }
        if (dft2_process_frame_keyed(d->plan, planes, keys, vsapi->getStride(src[0], plane),
                                     vsapi->getWritePtr(dst, plane), vsapi->getStride(dst, plane),
                                     vsapi->getFrameWidth(dst, plane), vsapi->getFrameHeight(dst, plane)) != DFT2_OK) {
            error = dft2_last_error();
        }
    }
    for (int t = 0; t < t_len; t++) {
        vsapi->freeFrame(src[t]);
    }
    if (error) {
        vsapi->setFilterError(error, frame_ctx);
        vsapi->freeFrame(dst);
        return nullptr;
    }
    return dst;
}

void VS_CC Free(void* instance_data, VSCore*, const VSAPI* vsapi) noexcept {
    auto* d = static_cast<DFTTestData*>(instance_data);
    dft2_plan_destroy(d->plan);
    vsapi->freeNode(d->node);
    delete d;
}

int opt_int(const VSAPI* vsapi, const VSMap* in, const char* key, int fallback) {
    int error = 0;
    const int v = vsh::int64ToIntS(vsapi->mapGetInt(in, key, 0, &error));
    return error ? fallback : v;
}

void VS_CC Create(const VSMap* in, VSMap* out, void*, VSCore* core, const VSAPI* vsapi) noexcept {
    auto* d = new DFTTestData();
    d->node = vsapi->mapGetNode(in, "clip", 0, nullptr);
    auto fail = [&](const char* message) {
        vsapi->mapSetError(out, message);
        vsapi->freeNode(d->node);
        delete d;
    };
    const VSVideoInfo* vi = vsapi->getVideoInfo(d->node);
    if (!vsh::isConstantVideoFormat(vi)) {
        return fail("only constant format input is supported");
    }
    dft2_params p{};
    p.radius = opt_int(vsapi, in, "radius", 0);
    p.block_size = opt_int(vsapi, in, "block_size", 16);
    p.block_step = opt_int(vsapi, in, "block_step", p.block_size);
    p.zero_mean = opt_int(vsapi, in, "zero_mean", 0);
    p.filter_type = opt_int(vsapi, in, "filter_type", 0);
    int error = 0;
    p.sigma2 = vsapi->mapGetFloat(in, "sigma2", 0, &error);
    if (error) {
        p.sigma2 = 8.0;
    }
    p.pmin = vsapi->mapGetFloat(in, "pmin", 0, &error);
    if (error) {
        p.pmin = 0.0;
    }
    p.pmax = vsapi->mapGetFloat(in, "pmax", 0, &error);
    if (error) {
        p.pmax = 500.0;
    }
    const int n_planes = vsapi->mapNumElements(in, "planes");
    std::fill(std::begin(d->process), std::end(d->process), n_planes <= 0);
    for (int i = 0; i < n_planes; i++) {
        const int plane = vsh::int64ToIntS(vsapi->mapGetInt(in, "planes", i, nullptr));
        if (plane < 0 || plane >= vi->format.numPlanes) {
            return fail("plane index out of range");
        }
        if (d->process[plane]) {
            return fail("plane specified twice");
        }
        d->process[plane] = true;
    }
    const int device_id = opt_int(vsapi, in, "device_id", 0);
    const int num_streams = opt_int(vsapi, in, "num_streams", 1);
    const int precision = opt_int(vsapi, in, "precision", DFT2_PRECISION_FP32);
    if (vsapi->mapNumElements(in, "window") <= 0) {
        return fail("window cannot be empty");
    }
    p.window = vsapi->mapGetFloatArray(in, "window", nullptr);
    p.window_len = vsapi->mapNumElements(in, "window");
    if (vsapi->mapNumElements(in, "sigma") <= 0) {
        return fail("sigma cannot be empty");
    }
    p.sigma = vsapi->mapGetFloatArray(in, "sigma", nullptr);
    p.sigma_len = vsapi->mapNumElements(in, "sigma");
    if (p.zero_mean) {
        if (vsapi->mapNumElements(in, "window_freq") <= 0) {
            return fail("window_freq required when zero_mean is enabled");
        }
        p.window_freq = vsapi->mapGetFloatArray(in, "window_freq", nullptr);
        p.window_freq_len = vsapi->mapNumElements(in, "window_freq");
    }
    const int sample_type = vi->format.sampleType == stFloat ? DFT2_SAMPLE_FLOAT : DFT2_SAMPLE_INT;
    if (dft2_plan_create(&p, sample_type, vi->format.bitsPerSample, device_id, num_streams, precision, &d->plan) !=
        DFT2_OK) {
        return fail(dft2_last_error());
    }
    if (2 * p.radius + 1 > kMaxSlices) { // dft2_plan_create rejects radius > 7 already
        dft2_plan_destroy(d->plan);
        return fail("radius must be at most 7");
    }
    d->radius = p.radius;
    VSFilterDependency deps[] = {{d->node, rpGeneral}};
    vsapi->createVideoFilter(out, "DFTTest", vi, GetFrame, Free, fmParallel, deps, 1, d, core);
}

void VS_CC Rdft(const VSMap* in, VSMap* out, void*, VSCore*, const VSAPI* vsapi) noexcept {
    const int ndim = vsapi->mapNumElements(in, "shape");
    if (ndim < 1 || ndim > 3) {
        return vsapi->mapSetError(out, "\"shape\" must be an array of ints with 1, 2 or 3 values");
    }
    int shape[3] = {0, 0, 0};
    long long outer = 1;
    for (int i = 0; i < ndim; i++) {
        shape[i] = vsh::int64ToIntS(vsapi->mapGetInt(in, "shape", i, nullptr));
        if (i + 1 < ndim) {
            outer *= shape[i];
        }
    }
    const long long n_out = 2 * outer * (shape[ndim - 1] / 2 + 1);
    std::vector<double> ret(static_cast<size_t>(std::max(n_out, 1LL)));
    if (dft2_rdft(vsapi->mapGetFloatArray(in, "data", nullptr), vsapi->mapNumElements(in, "data"), shape, ndim,
                  ret.data(), static_cast<int>(ret.size())) != DFT2_OK) {
        return vsapi->mapSetError(out, dft2_last_error());
    }
    // The reference sets the result under "ret" (cpu_source/source.cpp:305,317,338).
    vsapi->mapSetFloatArray(out, "ret", ret.data(), static_cast<int>(n_out));
}

void VS_CC Version(const VSMap*, VSMap* out, void*, VSCore*, const VSAPI* vsapi) noexcept {
    vsapi->mapSetData(out, "version", dft2_version(), -1, dtUtf8, maReplace);
}

} // namespace

VS_EXTERNAL_API(void) VapourSynthPluginInit2(VSPlugin* plugin, const VSPLUGINAPI* vspapi) {
    vspapi->configPlugin("io.github.dfttest2.mi355x", "dfttest2_mi355x", "DFTTest2 (MI355X)", VS_MAKE_VERSION(1, 0),
                         VAPOURSYNTH_API_VERSION, 0, plugin);
    vspapi->registerFunction("DFTTest",
                             "clip:vnode;"
                             "window:float[];"
                             "sigma:float[];"
                             "sigma2:float;"
                             "pmin:float;"
                             "pmax:float;"
                             "filter_type:int;"
                             "radius:int:opt;"
                             "block_size:int:opt;"
                             "block_step:int:opt;"
                             "zero_mean:int:opt;"
                             "window_freq:float[]:opt;"
                             "planes:int[]:opt;"
                             "device_id:int:opt;"
                             "num_streams:int:opt;"
                             "precision:int:opt;",
                             "clip:vnode;", Create, nullptr, plugin);
    vspapi->registerFunction("RDFT",
                             "data:float[];"
                             "shape:int[];",
                             "window_freq:float[];", Rdft, nullptr, plugin);
    vspapi->registerFunction("Version", "", "any", Version, nullptr, plugin);
}

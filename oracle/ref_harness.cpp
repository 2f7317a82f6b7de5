// Harness around the reference GCC backend's block core -- TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile into oracle/_ref/libdft2ref.so (never committed) from
// /root/reference/gcc_source/kernel.hpp, which is included where it lies.  That header
// is self-contained (<cmath> only) and holds the reference's fused() block pipeline:
// rdft/dft/idft/irdft codelets, zero-mean and the gains (gcc_source/kernel.hpp:1872-2036).
// The reference's per-frame glue (gcc_source/getframe_impl.cpp) needs VapourSynth
// headers that this image lacks, so the few lines of glue it performs are restated
// here, following cpu_source/getframe_impl.cpp:
//   calc_pad_size/calc_pad_num :11-20, reflection_padding :22-65, load_block :105-156,
//   store_block :158-175, store_frame :177-221, block loop :330-402,
//   parameter layouts cpu_source/source.cpp:145-203.
// Only tests, the fixture script and bench.py's cpu_baseline leg load this library.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "kernel.hpp"

namespace {

int pad_size(int size, int bs, int step) {
    return size + ((size % bs) ? bs - size % bs : 0) + std::max(bs - step, step) * 2;
}

int pad_num(int size, int bs, int step) {
    return (pad_size(size, bs, step) - bs) / step + 1;
}

struct Params {
    int radius;
    int step;
    int filter_type;
    bool zero_mean;
    float sigma2, pmin, pmax;
    int syn_slice;
    // 64-byte aligned lane rows, as DFTTestCreate lays them out.
    Vec16f* window;      // (T*16) rows of 16
    Vec16f* sigma;       // (T*16) rows, lanes kx = 0..8 used
    Vec16f* window_freq; // (T*16*2) rows: re lanes, im lanes
};

Vec16f* alloc_rows(int n) {
    void* p = nullptr;
    if (posix_memalign(&p, 64, sizeof(Vec16f) * n) != 0) {
        return nullptr;
    }
    std::memset(p, 0, sizeof(Vec16f) * n);
    return static_cast<Vec16f*>(p);
}

bool make_params(Params& p, const double* window, const double* sigma, double sigma2, double pmin, double pmax,
                 int filter_type, int zero_mean, const double* window_freq, int radius, int step, int syn_slice) {
    int t_len = 2 * radius + 1;
    p.radius = radius;
    p.step = step;
    p.filter_type = filter_type;
    p.zero_mean = zero_mean != 0;
    p.sigma2 = static_cast<float>(sigma2);
    p.pmin = static_cast<float>(pmin);
    p.pmax = static_cast<float>(pmax);
    p.syn_slice = syn_slice;
    p.window = alloc_rows(t_len * 16);
    p.sigma = alloc_rows(t_len * 16);
    p.window_freq = alloc_rows(t_len * 32);
    if (!p.window || !p.sigma || !p.window_freq) {
        return false;
    }
    for (int i = 0; i < t_len * 16; i++) {
        for (int j = 0; j < 16; j++) {
            p.window[i][j] = static_cast<float>(window[i * 16 + j]);
        }
        for (int j = 0; j < 9; j++) {
            p.sigma[i][j] = static_cast<float>(sigma[i * 9 + j]);
            if (window_freq) {
                p.window_freq[2 * i][j] = static_cast<float>(window_freq[(i * 9 + j) * 2]);
                p.window_freq[2 * i + 1][j] = static_cast<float>(window_freq[(i * 9 + j) * 2 + 1]);
            }
        }
    }
    return true;
}

void free_params(Params& p) {
    std::free(p.window);
    std::free(p.sigma);
    std::free(p.window_freq);
}

template <typename T>
void mirror_pad(T* dst, const T* src, int w, int h, int stride, int bs, int step) {
    int pw = pad_size(w, bs, step), ph = pad_size(h, bs, step);
    int oy = (ph - h) / 2, ox = (pw - w) / 2;
    for (int y = 0; y < ph; y++) {
        int sy = y - oy;
        if (sy < 0) sy = -sy;
        if (sy >= h) sy = 2 * (h - 1) - sy;
        for (int x = 0; x < pw; x++) {
            int sx = x - ox;
            if (sx < 0) sx = -sx;
            if (sx >= w) sx = 2 * (w - 1) - sx;
            dst[y * pw + x] = src[sy * stride + sx];
        }
    }
}

template <typename T>
void filter_frame(const T* frames, int n_frames, int n, int w, int h, int stride, int bits, bool is_float,
                  const Params& p, T* dst, std::vector<T>& padded, std::vector<float>& acc) {
    const int bs = 16;
    int t_len = 2 * p.radius + 1;
    int pw = pad_size(w, bs, p.step), ph = pad_size(h, bs, p.step);
    int ny = pad_num(h, bs, p.step), nx = pad_num(w, bs, p.step);
    size_t plane = static_cast<size_t>(pw) * ph;
    padded.resize(plane * t_len);
    acc.assign(plane, 0.0f);
    for (int t = 0; t < t_len; t++) {
        int fi = std::clamp(n - p.radius + t, 0, n_frames - 1);
        mirror_pad(&padded[plane * t], frames + static_cast<size_t>(fi) * h * stride, w, h, stride, bs, p.step);
    }
    float scale = is_float ? 255.0f : 1.0f / static_cast<float>(1 << (bits - 8));
    alignas(64) Vec16f block[7 * 16 * 2];
    const Vec16f* syn = &p.window[(p.syn_slice ? 2 * p.radius : p.radius) * 16];
    for (int by = 0; by < ny; by++) {
        for (int bx = 0; bx < nx; bx++) {
            for (int t = 0; t < t_len; t++) {
                for (int j = 0; j < 16; j++) {
                    const T* row = &padded[plane * t + static_cast<size_t>(by * p.step + j) * pw + bx * p.step];
                    Vec16f v;
                    for (int k = 0; k < 16; k++) {
                        v[k] = static_cast<float>(row[k]);
                    }
                    block[t * 32 + j] = scale * p.window[t * 16 + j] * v;
                }
            }
            fused(block, p.sigma, p.sigma2, p.pmin, p.pmax, p.filter_type, p.zero_mean, p.window_freq, p.radius);
            for (int j = 0; j < 16; j++) {
                float* arow = &acc[static_cast<size_t>(by * p.step + j) * pw + bx * p.step];
                Vec16f out = block[p.radius * 32 + j];
                for (int k = 0; k < 16; k++) {
                    arow[k] = std::fma(out[k], syn[j][k], arow[k]);
                }
            }
        }
    }
    int oy = (ph - h) / 2, ox = (pw - w) / 2;
    int peak = (1 << bits) - 1;
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            float v = acc[static_cast<size_t>(y + oy) * pw + x + ox];
            if (is_float) {
                reinterpret_cast<float*>(dst)[static_cast<size_t>(y) * stride + x] = v / scale;
            } else {
                int q = std::clamp(static_cast<int>(v / scale + 0.5f), 0, peak);
                dst[static_cast<size_t>(y) * stride + x] = static_cast<T>(q);
            }
        }
    }
}

template <typename T>
void run_plane(const T* src, int n_frames, int w, int h, int stride, int bits, bool is_float, const Params& p,
               T* dst, int n_threads) {
    auto worker = [&](int tid) {
        std::vector<T> padded;
        std::vector<float> acc;
        for (int n = tid; n < n_frames; n += n_threads) {
            filter_frame(src, n_frames, n, w, h, stride, bits, is_float, p,
                         dst + static_cast<size_t>(n) * h * stride, padded, acc);
        }
    };
    if (n_threads <= 1) {
        worker(0);
        return;
    }
    std::vector<std::thread> pool;
    for (int i = 0; i < n_threads; i++) {
        pool.emplace_back(worker, i);
    }
    for (auto& th : pool) {
        th.join();
    }
}

} // namespace

extern "C" {

// Filters one plane of a clip (frames [n_frames][height][stride] samples) into dst.
// sigma is the full [T][16][9] array.  Returns 0 on success.
int ref_filter_plane(const void* src, int n_frames, int width, int height, int stride, int bytes_per_sample,
                     int bits, int is_float, const double* window, const double* sigma, double sigma2, double pmin,
                     double pmax, int filter_type, int zero_mean, const double* window_freq, int radius,
                     int block_step, int syn_slice, void* dst, int n_threads) {
    Params p{};
    if (radius < 0 || radius > 3 || block_step < 1 || block_step > 16) {
        return 1;
    }
    if (!make_params(p, window, sigma, sigma2, pmin, pmax, filter_type, zero_mean, window_freq, radius, block_step,
                     syn_slice)) {
        free_params(p);
        return 2;
    }
    if (bytes_per_sample == 1) {
        run_plane(static_cast<const uint8_t*>(src), n_frames, width, height, stride, bits, false, p,
                  static_cast<uint8_t*>(dst), n_threads);
    } else if (bytes_per_sample == 2) {
        run_plane(static_cast<const uint16_t*>(src), n_frames, width, height, stride, bits, false, p,
                  static_cast<uint16_t*>(dst), n_threads);
    } else if (bytes_per_sample == 4 && is_float) {
        run_plane(static_cast<const float*>(src), n_frames, width, height, stride, 32, true, p,
                  static_cast<float*>(dst), n_threads);
    } else {
        free_params(p);
        return 3;
    }
    free_params(p);
    return 0;
}

// Runs the reference fused() on one windowed block [T][16][16] (in place); the first
// 256 floats of `block` then hold the centre output slice [16][16].
int ref_fused_block(float* block, const double* window, const double* sigma, double sigma2, double pmin, double pmax,
                    int filter_type, int zero_mean, const double* window_freq, int radius) {
    Params p{};
    if (!make_params(p, window, sigma, sigma2, pmin, pmax, filter_type, zero_mean, window_freq, radius, 4, 0)) {
        free_params(p);
        return 2;
    }
    int t_len = 2 * radius + 1;
    alignas(64) Vec16f b[7 * 16 * 2];
    std::memset(b, 0, sizeof(b));
    for (int t = 0; t < t_len; t++) {
        for (int j = 0; j < 16; j++) {
            for (int k = 0; k < 16; k++) {
                b[t * 32 + j][k] = block[(t * 16 + j) * 16 + k];
            }
        }
    }
    fused(b, p.sigma, p.sigma2, p.pmin, p.pmax, p.filter_type, p.zero_mean, p.window_freq, p.radius);
    for (int j = 0; j < 16; j++) {
        for (int k = 0; k < 16; k++) {
            block[j * 16 + k] = b[radius * 32 + j][k];
        }
    }
    free_params(p);
    return 0;
}

} // extern "C"

/*
 * A C99 client of the dfttest2_mi355x C ABI (include/dfttest2_mi355x.h), as a caller in
 * another language binds it: it includes only the public header and links only the
 * shared library.  Test infrastructure (tests/test_abi_client.py runs it).
 *
 *   abi_client cpu
 *       Entry points that need no GPU: the version string, dft2_rdft against a direct
 *       FP64 DFT, and plan-creation argument errors with their messages.  Exit 0 when all
 *       checks hold.
 *   abi_client frames PARAMS FRAMES OUT_PLAIN OUT_KEYED
 *       Creates a plan from PARAMS and filters every output frame of the clip in FRAMES
 *       with dft2_process_frame on host planes (OUT_PLAIN) and with
 *       dft2_process_frame_keyed (OUT_KEYED), gathering the T source planes of each output
 *       frame with the temporal clamp, as a VapourSynth GetFrame does.
 *
 * PARAMS (native byte order): 13 int32 (radius, block_size, block_step, filter_type,
 * zero_mean, window_len, sigma_len, window_freq_len, sample_type, bits_per_sample, width,
 * height, n_frames), 3 float64 (sigma2, pmin, pmax), then window_len, sigma_len and
 * window_freq_len float64.  FRAMES: n_frames planes of width x height samples (1, 2 or 4
 * bytes each), rows contiguous.  The outputs have the same layout.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dfttest2_mi355x.h"

static int fail_msg(const char* what) {
    fprintf(stderr, "abi_client: %s (%s)\n", what, dft2_last_error());
    return 1;
}

/* Direct FP64 DFT of a real [n0][n1][n2] array: real along the last axis (n2/2 + 1
 * bins), complex along the others; out interleaved re, im. */
static void direct_rdft(const double* x, const int* n, double* out) {
    const int h = n[2] / 2 + 1;
    const double pi = 3.14159265358979323846;
    for (int k0 = 0; k0 < n[0]; k0++) {
        for (int k1 = 0; k1 < n[1]; k1++) {
            for (int k2 = 0; k2 < h; k2++) {
                double re = 0.0, im = 0.0;
                for (int i0 = 0; i0 < n[0]; i0++) {
                    for (int i1 = 0; i1 < n[1]; i1++) {
                        for (int i2 = 0; i2 < n[2]; i2++) {
                            const double a = -2.0 * pi *
                                             ((double)k0 * i0 / n[0] + (double)k1 * i1 / n[1] + (double)k2 * i2 / n[2]);
                            const double v = x[(i0 * n[1] + i1) * n[2] + i2];
                            re += v * cos(a);
                            im += v * sin(a);
                        }
                    }
                }
                const int o = ((k0 * n[1] + k1) * h + k2) * 2;
                out[o] = re;
                out[o + 1] = im;
            }
        }
    }
}

static int run_cpu(void) {
    const char* v = dft2_version();
    if (!v || !*v) {
        return fail_msg("empty version string");
    }
    printf("version: %s\n", v);

    const int shape[3] = {2, 3, 4};
    double x[24], got[2 * 3 * 3 * 2], want[2 * 3 * 3 * 2];
    for (int i = 0; i < 24; i++) {
        x[i] = sin(0.7 * i) + 0.25 * i;
    }
    if (dft2_rdft(x, 24, shape, 3, got, 36) != DFT2_OK) {
        return fail_msg("dft2_rdft failed");
    }
    direct_rdft(x, shape, want);
    double err = 0.0;
    for (int i = 0; i < 36; i++) {
        err = fmax(err, fabs(got[i] - want[i]));
    }
    printf("rdft max error %.3g\n", err);
    if (err > 1e-9) {
        return fail_msg("dft2_rdft differs from the direct DFT");
    }
    if (dft2_rdft(x, 24, shape, 3, got, 35) == DFT2_OK) {
        return fail_msg("dft2_rdft accepted a short output");
    }

    /* plan creation: argument errors come back as DFT2_ERR_ARG with a message */
    double window[3 * 256], sigma = 8.0;
    for (int i = 0; i < 3 * 256; i++) {
        window[i] = 1.0;
    }
    dft2_params p;
    memset(&p, 0, sizeof p);
    p.radius = 1;
    p.block_size = 16;
    p.block_step = 4;
    p.filter_type = 0;
    p.zero_mean = 0;
    p.window = window;
    p.window_len = 3 * 256;
    p.sigma = &sigma;
    p.sigma_len = 1;
    dft2_plan* plan = NULL;
    p.radius = 9;
    if (dft2_plan_create(&p, DFT2_SAMPLE_INT, 8, 0, 1, DFT2_PRECISION_FP32, &plan) != DFT2_ERR_ARG || plan) {
        return fail_msg("radius 9 was not rejected as an argument error");
    }
    if (!strlen(dft2_last_error())) {
        return fail_msg("no message for the rejected radius");
    }
    printf("radius 9: %s\n", dft2_last_error());
    p.radius = 1;
    p.window_len = 5;
    if (dft2_plan_create(&p, DFT2_SAMPLE_INT, 8, 0, 1, DFT2_PRECISION_FP32, &plan) != DFT2_ERR_ARG || plan) {
        return fail_msg("a short window was not rejected");
    }
    printf("short window: %s\n", dft2_last_error());
    if (dft2_plan_create(NULL, DFT2_SAMPLE_INT, 8, 0, 1, DFT2_PRECISION_FP32, &plan) == DFT2_OK) {
        return fail_msg("NULL parameters were accepted");
    }
    if (dft2_plan_destroy(NULL) != DFT2_OK) { /* documented: NULL is accepted */
        return fail_msg("dft2_plan_destroy(NULL) failed");
    }
    printf("cpu checks passed\n");
    return 0;
}

static void* read_file(const char* path, size_t* size) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        return NULL;
    }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    void* buf = malloc(n > 0 ? (size_t)n : 1);
    if (buf && n > 0 && fread(buf, 1, (size_t)n, f) != (size_t)n) {
        free(buf);
        buf = NULL;
    }
    fclose(f);
    *size = (size_t)n;
    return buf;
}

static int write_file(const char* path, const void* data, size_t size) {
    FILE* f = fopen(path, "wb");
    if (!f) {
        return 1;
    }
    const size_t w = fwrite(data, 1, size, f);
    fclose(f);
    return w != size;
}

static int run_frames(const char* params_path, const char* frames_path, const char* out_plain, const char* out_keyed) {
    size_t psize = 0, fsize = 0;
    unsigned char* pb = read_file(params_path, &psize);
    unsigned char* frames = read_file(frames_path, &fsize);
    if (!pb || !frames || psize < 13 * 4 + 3 * 8) {
        return fail_msg("cannot read the parameter or frame file");
    }
    int hdr[13];
    double scal[3];
    memcpy(hdr, pb, sizeof hdr);
    memcpy(scal, pb + sizeof hdr, sizeof scal);
    const int window_len = hdr[5], sigma_len = hdr[6], wf_len = hdr[7];
    const int sample_type = hdr[8], bits = hdr[9], width = hdr[10], height = hdr[11], n_frames = hdr[12];
    if (psize != sizeof hdr + sizeof scal + 8 * (size_t)(window_len + sigma_len + wf_len)) {
        return fail_msg("parameter file size does not match its header");
    }
    const double* arrays = (const double*)(pb + sizeof hdr + sizeof scal);
    dft2_params p;
    memset(&p, 0, sizeof p);
    p.radius = hdr[0];
    p.block_size = hdr[1];
    p.block_step = hdr[2];
    p.filter_type = hdr[3];
    p.zero_mean = hdr[4];
    p.window = arrays;
    p.window_len = window_len;
    p.sigma = arrays + window_len;
    p.sigma_len = sigma_len;
    p.window_freq = wf_len ? arrays + window_len + sigma_len : NULL;
    p.window_freq_len = wf_len;
    p.sigma2 = scal[0];
    p.pmin = scal[1];
    p.pmax = scal[2];
    const int bps = sample_type == DFT2_SAMPLE_FLOAT ? 4 : (bits <= 8 ? 1 : 2);
    const size_t plane = (size_t)width * height * bps;
    if (fsize != plane * n_frames) {
        return fail_msg("frame file size does not match the geometry");
    }
    dft2_plan* plan = NULL;
    if (dft2_plan_create(&p, sample_type, bits, 0, 1, DFT2_PRECISION_FP32, &plan) != DFT2_OK) {
        return fail_msg("dft2_plan_create failed");
    }
    const int r = p.radius, t_len = 2 * r + 1;
    unsigned char* plain = malloc(plane * n_frames);
    unsigned char* keyed = malloc(plane * n_frames);
    const void* src[15];
    unsigned long long keys[15];
    int rc = 0;
    for (int n = 0; n < n_frames && !rc; n++) {
        for (int t = 0; t < t_len; t++) {
            int i = n - r + t;
            i = i < 0 ? 0 : (i > n_frames - 1 ? n_frames - 1 : i);
            src[t] = frames + plane * i;
            keys[t] = (unsigned long long)i;
        }
        if (dft2_process_frame(plan, src, (long long)width * bps, plain + plane * n, (long long)width * bps, width,
                               height, 1, NULL) != DFT2_OK) {
            rc = fail_msg("dft2_process_frame failed");
        } else if (dft2_process_frame_keyed(plan, src, keys, (long long)width * bps, keyed + plane * n,
                                            (long long)width * bps, width, height) != DFT2_OK) {
            rc = fail_msg("dft2_process_frame_keyed failed");
        }
    }
    if (dft2_plan_destroy(plan) != DFT2_OK) {
        rc = fail_msg("dft2_plan_destroy failed");
    }
    if (!rc && (write_file(out_plain, plain, plane * n_frames) || write_file(out_keyed, keyed, plane * n_frames))) {
        rc = fail_msg("cannot write the outputs");
    }
    free(plain);
    free(keyed);
    free(pb);
    free(frames);
    if (!rc) {
        printf("filtered %d frames of %dx%d\n", n_frames, width, height);
    }
    return rc;
}

int main(int argc, char** argv) {
    if (argc == 2 && !strcmp(argv[1], "cpu")) {
        return run_cpu();
    }
    if (argc == 6 && !strcmp(argv[1], "frames")) {
        return run_frames(argv[2], argv[3], argv[4], argv[5]);
    }
    fprintf(stderr, "usage: abi_client cpu | abi_client frames PARAMS FRAMES OUT_PLAIN OUT_KEYED\n");
    return 2;
}

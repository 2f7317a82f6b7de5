// Device helpers shared by the gfx950 kernels: MFMA operand forms, buffer loads of
// source samples, wave-uniform values, the gain and its table.  Everything here has
// internal linkage; each kernel translation unit includes it.
#pragma once

#include <type_traits>

#include "dft2_internal.h"

namespace dft2 {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short shortx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// One lane's share of a 16-long contraction operand: the value at K slot (lane/16, s)
// for s = 0..3, plus its split into BF16 hi + lo parts for the split-precision form.
struct Opnd {
    float f[4];
    bf16x4 hi, lo;
};

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// BF16 hi part of (a, b) as one packed register, and the FP32 values of its halves (a
// BF16 is the upper half of an FP32 with the same value).
__device__ __forceinline__ unsigned pack_hi(float a, float b, float& ha, float& hb) {
    const unsigned h = __builtin_bit_cast(unsigned, __builtin_convertvector(floatx2{a, b}, bf16x2));
    ha = __builtin_bit_cast(float, h << 16);
    hb = __builtin_bit_cast(float, h & 0xffff0000u);
    return h;
}

__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(floatx2{a, b}, bf16x2));
}

template <int PREC>
__device__ __forceinline__ Opnd make_opnd(float a0, float a1, float a2, float a3) {
    Opnd o;
    o.f[0] = a0;
    o.f[1] = a1;
    o.f[2] = a2;
    o.f[3] = a3;
    if constexpr (PREC == 1) {
        // Packed conversions: two values per instruction for hi and for lo.
        float h0, h1, h2, h3;
        const unsigned hi01 = pack_hi(a0, a1, h0, h1);
        const unsigned hi23 = pack_hi(a2, a3, h2, h3);
        const unsigned lo01 = pack_bf16(a0 - h0, a1 - h1);
        const unsigned lo23 = pack_bf16(a2 - h2, a3 - h3);
        o.hi = __builtin_bit_cast(bf16x4, (static_cast<unsigned long long>(hi23) << 32) | hi01);
        o.lo = __builtin_bit_cast(bf16x4, (static_cast<unsigned long long>(lo23) << 32) | lo01);
    }
    return o;
}

// acc += A * B over one 16-long contraction (4 K slots per lane x 4 lane groups).
// PREC 0: four FP32 MFMA 16x16x4.  PREC 1: split BF16 with FP32 accumulation,
// A*B ~ Ahi*Bhi + Alo*Bhi + Ahi*Blo, three BF16 MFMA 16x16x16.  The K slot pairing
// (lane group, s) is the same in every form, so operands need no reshuffle.
// The K-packed alternative, one 16x16x32 on [Ahi | Alo] x [Blo | Bhi] plus one 16x16x16
// on the hi halves, was measured on gfx950 with this toolchain: chained through one
// accumulator (either order) it returned wrong products; with separate accumulators and
// a VALU add it was correct but no faster than three 16x16x16 (1080p u8, 4K 10-bit and
// the backward pass), so this form is used.
template <int PREC>
__device__ __forceinline__ floatx4 mm16(const Opnd& A, const Opnd& B, floatx4 acc) {
    if constexpr (PREC == 0) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            acc = mfma4(A.f[s], B.f[s], acc);
        }
    } else {
        const shortx4 ahi = __builtin_bit_cast(shortx4, A.hi);
        const shortx4 alo = __builtin_bit_cast(shortx4, A.lo);
        const shortx4 bhi = __builtin_bit_cast(shortx4, B.hi);
        const shortx4 blo = __builtin_bit_cast(shortx4, B.lo);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, acc, 0, 0, 0);
    }
    return acc;
}

// An A operand that is exact in BF16 (integers of at most 8 significant bits): only
// A.hi is set.
__device__ __forceinline__ Opnd make_opnd_exact(float a0, float a1, float a2, float a3) {
    Opnd o;
    o.hi = __builtin_bit_cast(bf16x4, (static_cast<unsigned long long>(pack_bf16(a2, a3)) << 32) | pack_bf16(a0, a1));
    return o;
}

// FP32 contraction whose A operand is exact in BF16, on the BF16 MFMA.  An FP32 value is
// the exact sum of three BF16 values (hi = its rounding to 8 significant bits, then the
// same for the remainder, and the remainder of that, which has at most 8 bits left), so
// A*B = A*Bhi + A*Bmid + A*Blo with every product exact (8 x 8 significant bits) and the
// sums accumulated in FP32: the accuracy of an FP32 contraction, in three 16x16x16 BF16
// MFMA instead of four 16x16x4 FP32 MFMA.  Smallest part first.
struct Opnd3 {
    bf16x4 hi, mid, lo;
};

__device__ __forceinline__ Opnd3 make_opnd3(float a0, float a1, float a2, float a3) {
    float h0, h1, h2, h3, m0, m1, m2, m3;
    const unsigned hi01 = pack_hi(a0, a1, h0, h1);
    const unsigned hi23 = pack_hi(a2, a3, h2, h3);
    const float r0 = a0 - h0, r1 = a1 - h1, r2 = a2 - h2, r3 = a3 - h3;
    const unsigned mid01 = pack_hi(r0, r1, m0, m1);
    const unsigned mid23 = pack_hi(r2, r3, m2, m3);
    const unsigned lo01 = pack_bf16(r0 - m0, r1 - m1);
    const unsigned lo23 = pack_bf16(r2 - m2, r3 - m3);
    Opnd3 o;
    o.hi = __builtin_bit_cast(bf16x4, (static_cast<unsigned long long>(hi23) << 32) | hi01);
    o.mid = __builtin_bit_cast(bf16x4, (static_cast<unsigned long long>(mid23) << 32) | mid01);
    o.lo = __builtin_bit_cast(bf16x4, (static_cast<unsigned long long>(lo23) << 32) | lo01);
    return o;
}

__device__ __forceinline__ floatx4 mm16_exact3(const Opnd& A, const Opnd3& B, floatx4 acc) {
    const shortx4 ahi = __builtin_bit_cast(shortx4, A.hi);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, __builtin_bit_cast(shortx4, B.lo), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, __builtin_bit_cast(shortx4, B.mid), acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, __builtin_bit_cast(shortx4, B.hi), acc, 0, 0, 0);
    return acc;
}

// Contraction whose A operand is made of integers of at most 16 bits, on the 16x16x32
// BF16 MFMA.  A sample v = 256 h + l (h, l bytes, exact in BF16) takes two of a lane's
// eight K slots: slot s (0..3) holds l of its sample s, slot 4 + s holds h; B holds a
// part of the basis in slots 0..3 and 256 times it (exact) in slots 4..7, so one MFMA
// gives A * Bpart with exact products.  FP32: the three parts of make_opnd3 (smallest
// first); split BF16: the hi and lo parts of make_opnd<1>.  Three (two) MFMA of 16
// cycles against four FP32 MFMA of 32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct OpndW {
    bf16x8 hi, mid, lo;
};

__device__ __forceinline__ bf16x8 part_w(float b0, float b1, float b2, float b3) {
    const u32x4 w = {pack_bf16(b0, b1), pack_bf16(b2, b3), pack_bf16(256.0f * b0, 256.0f * b1),
                     pack_bf16(256.0f * b2, 256.0f * b3)};
    return __builtin_bit_cast(bf16x8, w);
}

template <int PREC>
__device__ __forceinline__ OpndW make_opndw(float a0, float a1, float a2, float a3) {
    float h0, h1, h2, h3;
    pack_hi(a0, a1, h0, h1);
    pack_hi(a2, a3, h2, h3);
    const float r0 = a0 - h0, r1 = a1 - h1, r2 = a2 - h2, r3 = a3 - h3;
    OpndW o;
    o.hi = part_w(h0, h1, h2, h3);
    if constexpr (PREC == 1) {
        // lo = bf16(a - hi), as make_opnd<1>
        o.lo = part_w(r0, r1, r2, r3);
    } else {
        float m0, m1, m2, m3;
        pack_hi(r0, r1, m0, m1);
        pack_hi(r2, r3, m2, m3);
        o.mid = part_w(m0, m1, m2, m3);
        o.lo = part_w(r0 - m0, r1 - m1, r2 - m2, r3 - m3);
    }
    return o;
}

// The four 16-bit samples of a raw4 (two per dword) as the A operand: low bytes in slots
// 0..3, high bytes in slots 4..7.
__device__ __forceinline__ bf16x8 bytes_w(unsigned d0, unsigned d1) {
    const u32x4 w = {pack_bf16(static_cast<float>(d0 & 0xffu), static_cast<float>((d0 >> 16) & 0xffu)),
                     pack_bf16(static_cast<float>(d1 & 0xffu), static_cast<float>((d1 >> 16) & 0xffu)),
                     pack_bf16(static_cast<float>((d0 >> 8) & 0xffu), static_cast<float>(d0 >> 24)),
                     pack_bf16(static_cast<float>((d1 >> 8) & 0xffu), static_cast<float>(d1 >> 24))};
    return __builtin_bit_cast(bf16x8, w);
}

template <int PREC>
__device__ __forceinline__ floatx4 mm16_bytes(bf16x8 A, const OpndW& B, floatx4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B.lo, acc, 0, 0, 0);
    if constexpr (PREC == 0) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B.mid, acc, 0, 0, 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B.hi, acc, 0, 0, 0);
    return acc;
}

// Contraction whose A operand is made of integers of at most 8 bits, on the FP16 MFMA
// (the split form's first contraction for 8-bit samples).
// A byte b is exact in FP16, and so are the parts of an FP32 basis value split into FP16
// (11 significant bits each; three parts hold the 24 bits of an FP32 value, two parts 22),
// so every product is exact and the sums are FP32.  The basis is scaled by kBasisH16
// (a power of two) before the split, which keeps its parts clear of FP16's subnormal
// range for all but the smallest values (whose lost bits are far below the FP32 sums'
// rounding); the caller takes the scale out of the next contraction's operand.
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));
constexpr float kBasisH16 = 4096.0f;

struct OpndH {
    halfx4 hi, mid, lo;
};

__device__ __forceinline__ halfx4 pack_h4(float a0, float a1, float a2, float a3) {
    return halfx4{static_cast<_Float16>(a0), static_cast<_Float16>(a1), static_cast<_Float16>(a2),
                  static_cast<_Float16>(a3)};
}

// PREC 0: three FP16 parts of kBasisH16 * a (smallest first in mm16_h); PREC 1: two.
template <int PREC>
__device__ __forceinline__ OpndH make_opndh(float a0, float a1, float a2, float a3) {
    const float v[4] = {a0 * kBasisH16, a1 * kBasisH16, a2 * kBasisH16, a3 * kBasisH16};
    float r[4], q[4];
    OpndH o;
    o.hi = pack_h4(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        r[i] = v[i] - static_cast<float>(o.hi[i]); // exact: the part left below hi's bits
    }
    if constexpr (PREC == 1) {
        o.lo = pack_h4(r[0], r[1], r[2], r[3]);
    } else {
        o.mid = pack_h4(r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            q[i] = r[i] - static_cast<float>(o.mid[i]);
        }
        o.lo = pack_h4(q[0], q[1], q[2], q[3]);
    }
    return o;
}

// The four 8-bit samples of a dword as FP16: each byte placed under the exponent of 1024
// (FP16 bits 0x6400 | b = 1024 + b, exact: 10 mantissa bits) by one byte permute per
// pair, then 1024 taken off two at a time (exact).  Four instructions for four samples.
__device__ __forceinline__ halfx4 bytes_h(unsigned d) {
    const unsigned k = 0x64646464u;
    // v_perm_b32 selectors: 0..3 pick bytes of the second operand (d), 4..7 of the first (k)
    const unsigned p01 = __builtin_amdgcn_perm(k, d, 0x04010400u); // [b0, 0x64, b1, 0x64]
    const unsigned p23 = __builtin_amdgcn_perm(k, d, 0x04030402u); // [b2, 0x64, b3, 0x64]
    const halfx2 off = {static_cast<_Float16>(1024.0f), static_cast<_Float16>(1024.0f)};
    const halfx2 h01 = __builtin_bit_cast(halfx2, p01) - off;
    const halfx2 h23 = __builtin_bit_cast(halfx2, p23) - off;
    return halfx4{h01[0], h01[1], h23[0], h23[1]};
}

template <int PREC>
__device__ __forceinline__ floatx4 mm16_h(halfx4 A, const OpndH& B, floatx4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(A, B.lo, acc, 0, 0, 0);
    if constexpr (PREC == 0) {
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(A, B.mid, acc, 0, 0, 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(A, B.hi, acc, 0, 0, 0);
    return acc;
}

// Orders this wave's LDS accesses before and after this point, so that values one lane
// writes can be read by another lane of the same wave afterwards.  The LDS executes a
// wave's instructions in issue order, so only the compiler has to be kept from moving
// accesses across the point; nothing waits here (the compiler still waits for a load's
// data before its first use).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_wave_barrier();
}

// Returns a wave-uniform value the optimizer cannot see through, so that per-lane values
// derived from it inside a loop are recomputed there rather than hoisted out of the loop
// into registers that stay live across every iteration.
__device__ __forceinline__ int opaque_uniform(int v) {
    asm volatile("" : "+s"(v));
    return v;
}

// x / S for 0 <= x < 256 and 1 <= S <= 16, given m = ceil(65536 / S): exact, since
// x * m / 65536 exceeds x / S by less than 256 / 65536 < 1 / S.  x * m < 2^24, so one
// 24-bit multiply.
__device__ __forceinline__ int div_small(int x, int m) {
    return static_cast<int>(__umul24(static_cast<unsigned>(x), static_cast<unsigned>(m)) >> 16);
}

// Overlap-add gather: the sum of the staged values of blocks lo..hi covering one pixel,
// block k's value at p[k * kStride].  The four blocks k0..k0+3, k0 = max(hi - 3, 0), are
// read together at immediate offsets (for block steps >= 4 they include every covering
// block) and masked to lo..hi; block steps 1 and 2 add the rest one by one.  Every read
// stays below block max(hi, 3) <= 7, inside the staging area.
// The gather is split in two, gather_head (blocks k0..k0+3, k0 = max(hi - 3, 0)) and
// gather_tail (blocks k0-1 down to lo), so that a caller with several pixels issues every
// first-part read, and its accumulator reads, before any sum is formed.
template <int kStride>
__device__ __forceinline__ float gather_head(const float* p, int k0, int lo, int hi) {
    const float* q = p + __umul24(static_cast<unsigned>(k0), static_cast<unsigned>(kStride));
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float v = q[j * kStride];
        // lo <= k0 + j <= hi as one compare, so that the read is not made conditional
        sum += static_cast<unsigned>(k0 + j - lo) <= static_cast<unsigned>(hi - lo) ? v : 0.0f;
    }
    return sum;
}

template <int kStride>
__device__ __forceinline__ float gather_tail(const float* p, int k0, int lo, float sum) {
    for (int k = k0 - 1; k >= lo; k--) {
        sum += p[k * kStride];
    }
    return sum;
}


// The work item of workgroup `bid` of a grid of n: each XCD takes a contiguous range of
// the items, in order.  Workgroups are dispatched to the 8 XCDs round-robin by index and
// each XCD has its own L2, so with the identity map a tile's horizontal neighbours, and
// the same tile of the next frame, which read the same source rows, run on other XCDs
// and fetch those rows again.  A bijection on [0, n): XCD x = bid % 8 takes items
// x*q + min(x, n % 8) + bid / 8, q = n / 8.
__device__ __forceinline__ int xcd_item(int bid, int n) {
    constexpr int kXcds = 8;
    const int q = n / kXcds;
    const int rem = n - q * kXcds;
    const int x = bid % kXcds;
    return x * q + min(x, rem) + bid / kXcds;
}

// The same within one job's items [begin, begin + count), which the workgroups
// begin .. begin + count - 1 take: each XCD takes a contiguous part of every job (plane),
// so that planes whose tiles cost differently (luma and chroma) spread evenly.
__device__ __forceinline__ int job_xcd_item(int bid, int begin, int count) {
    return begin + xcd_item(bid - begin, count);
}

__device__ __forceinline__ int mirror(int v, int n) {
    v = v < 0 ? -v : v;
    return v >= n ? 2 * (n - 1) - v : v;
}

// Values that are the same in every lane of the wave, moved to scalar registers (the
// compiler cannot prove that for values derived from threadIdx.x / 64 or read from the
// job table, and would otherwise keep them, and everything computed from them, per lane).
__device__ __forceinline__ int sgpr(int v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ long long sgpr64(long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(static_cast<unsigned long long>(v) >> 32));
    return static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo);
}

template <typename P>
__device__ __forceinline__ P* sgpr_ptr(P* p) {
    return reinterpret_cast<P*>(sgpr64(reinterpret_cast<long long>(p)));
}

// Source samples are read with raw buffer loads: a descriptor per source frame (base,
// size in bytes; loads past the size return 0 instead of faulting) and a 32-bit byte
// offset per lane, plus an immediate offset per column.  The raw bits stay in a register
// until the sample is used, so a prefetch keeps its loads in flight.
constexpr int kBufferDword3 = 0x00020000; // gfx9-family descriptor word 3 (raw access)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), kBufferDword3);
}

// Four consecutive samples (columns 4*(lane/16)..+3) of one lane and slice as raw bits:
// one dword of 4 bytes, two dwords of 2 halves, or four float dwords.  They stay packed
// until the samples are used, so a prefetch keeps its loads in flight and holds T
// registers per lane for 8-bit samples instead of 4*T.

template <typename Tin>
struct raw4 {
    unsigned v[sizeof(Tin)];
};

// Byte alignment of a raw4 load as one access (dword, dwordx2, dwordx4): its own size.
template <typename Tin>
constexpr unsigned kRaw4Align = 4u * sizeof(Tin);

// The four samples at byte offset `off`, which is a multiple of kRaw4Align<Tin>.
template <typename Tin>
__device__ __forceinline__ raw4<Tin> load4_aligned(__amdgpu_buffer_rsrc_t r, unsigned off) {
    raw4<Tin> x;
    if constexpr (sizeof(Tin) == 1) {
        x.v[0] = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    } else if constexpr (sizeof(Tin) == 2) {
        const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        x.v[0] = w.x;
        x.v[1] = w.y;
    } else {
        const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        x.v[0] = w.x;
        x.v[1] = w.y;
        x.v[2] = w.z;
        x.v[3] = w.w;
    }
    return x;
}

// The four samples at byte offsets off[0..3], one load each.  Integer samples are packed
// as they arrive (their loads complete here); float samples stay in flight.
template <typename Tin>
__device__ __forceinline__ raw4<Tin> load4_each(__amdgpu_buffer_rsrc_t r, const unsigned (&off)[4]) {
    raw4<Tin> x;
    if constexpr (sizeof(Tin) == 1) {
        unsigned b[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            b[s] = __builtin_amdgcn_raw_buffer_load_b8(r, off[s], 0, 0);
        }
        x.v[0] = b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24);
    } else if constexpr (sizeof(Tin) == 2) {
        unsigned h[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            h[s] = __builtin_amdgcn_raw_buffer_load_b16(r, off[s], 0, 0);
        }
        x.v[0] = h[0] | (h[1] << 16);
        x.v[1] = h[2] | (h[3] << 16);
    } else {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            x.v[s] = __builtin_amdgcn_raw_buffer_load_b32(r, off[s], 0, 0);
        }
    }
    return x;
}

// Sample s (0..3) of a raw4 as float: one conversion instruction for integer samples.
template <typename Tin>
__device__ __forceinline__ float raw4_value(const raw4<Tin>& x, int s) {
    if constexpr (sizeof(Tin) == 1) {
        return static_cast<float>((x.v[0] >> (8 * s)) & 0xffu);
    } else if constexpr (sizeof(Tin) == 2) {
        return static_cast<float>((x.v[s >> 1] >> (16 * (s & 1))) & 0xffffu);
    } else {
        return __builtin_bit_cast(float, x.v[s]);
    }
}

// Gain of one coefficient (cpu_source/kernel.hpp:1826-1890).  FTM 0 and 1 are the
// Wiener and hard-threshold filters; FTM 2 dispatches on the filter type at run time.
template <int FTM>
__device__ __forceinline__ float gain_of(float psd, float s, const LaunchArgs& a) {
    if constexpr (FTM == 0) {
        // s arrives as sigma + 1e-15 (gain table of a Wiener plan) and psd as
        // |x|^2 + 1e-15 (filter_coef):
        // (psd - sigma) / (psd + 1e-15) = 1 - (sigma + 1e-15) / (psd + 1e-15).
        return fmaxf(fmaf(-s, __builtin_amdgcn_rcpf(psd), 1.0f), 0.0f);
    } else if constexpr (FTM == 1) {
        return psd < s ? 0.0f : 1.0f;
    } else {
        switch (a.filter_type) {
        case 0:
            return fmaxf((psd - s) / (psd + 1e-15f), 0.0f);
        case 1:
            return psd < s ? 0.0f : 1.0f;
        case 2:
            return s;
        case 3:
            return (psd >= a.pmin && psd <= a.pmax) ? s : a.sigma2;
        case 4:
            return s * sqrtf(psd * (a.pmax / ((psd + a.pmin) * (psd + a.pmax) + 1e-15f)));
        case 5:
            return powf(fmaxf((psd - s) / (psd + 1e-15f), 0.0f), a.pmin);
        default:
            return sqrtf(fmaxf((psd - s) / (psd + 1e-15f), 0.0f));
        }
    }
}

// One gain-table entry from LDS as a single 128-bit read.  Without the empty asm the
// symmetric form, which does not use the partner sigma, gets split 32/64-bit reads at a
// 16-byte lane stride: twice the LDS instructions, with bank conflicts.
template <int ASYM>
__device__ __forceinline__ float4 tab_entry(const float4* p) {
    float4 e = *p;
    if constexpr (ASYM == 0) {
        asm volatile("" : "+v"(e.y));
    }
    return e;
}

// 1 - gain of the Wiener (FTM 0) and hard-threshold (FTM 1) filters, with psd and s as
// gain_of takes them: Wiener 1 - max(1 - s/psd, 0) = min(s/psd, 1).  kNonNeg: s >= 0 (the
// host sends plans with a negative sigma to the asymmetric form), so that the min is a
// clamp to [0, 1], which folds into the multiply as its output modifier.
template <int FTM, bool kNonNeg>
__device__ __forceinline__ float damping_of(float psd, float s) {
    if constexpr (FTM == 0) {
        const float q = s * __builtin_amdgcn_rcpf(psd);
        return kNonNeg ? fminf(fmaxf(q, 0.0f), 1.0f) : fminf(q, 1.0f);
    } else {
        return psd < s ? 1.0f : 0.0f;
    }
}

// Zero-mean, gain and mean re-add on one complex coefficient.  e = (sigma, sigma of the
// conjugate partner, Re WF, Im WF) at this coefficient.  With an asymmetric sigma array
// the coefficients at kx = 0 and kx = 8 take the mean of their own gain and their
// conjugate partner's gain, which is what the reference's C2R (it drops Im at those
// columns) makes of two different gains.
// ASYM: 0 = symmetric sigma (no partner gains), 1 = asymmetric, 2 = decided at run time
// from a.asym.  The forms are separate kernel instantiations because the compiler turns
// a run-time test into selects and computes the partner gains for every plan.
// Returns the coefficient's hard decision: filter type 1, psd < sigma (dropped); type 3
// (FTM 2), pmin <= psd <= pmax; false otherwise.  Callers that do not record decisions
// ignore it, and the compiler drops the comparison.
template <int FTM, int ASYM>
__device__ __forceinline__ bool filter_coef(float& re, float& im, float4 e, float gf, const LaunchArgs& a) {
    const float xr = fmaf(-gf, e.z, re);
    const float xi = fmaf(-gf, e.w, im);
    // The Wiener gain takes psd + 1e-15: the constant rides in the psd's own fma chain.
    const float psd = fmaf(xr, xr, fmaf(xi, xi, FTM == 0 ? 1e-15f : 0.0f));
    if constexpr (FTM == 0 || FTM == 1) {
        // g x + gf WF = (x + gf WF) - (1 - g) x: the mean term gf WF is never formed.
        float h = damping_of<FTM, ASYM == 0>(psd, e.x);
        if constexpr (ASYM != 0) {
            if ((ASYM == 1 || a.asym) && e.y != e.x) { // only the kx = 0, 8 lanes
                h = 0.5f * (h + damping_of<FTM, false>(psd, e.y));
            }
        }
        re = fmaf(-xr, h, re);
        im = fmaf(-xi, h, im);
        return FTM == 1 && psd < e.x;
    } else {
        float g = gain_of<FTM>(psd, e.x, a);
        if constexpr (ASYM != 0) {
            if ((ASYM == 1 || a.asym) && e.y != e.x) {
                g = 0.5f * (g + gain_of<FTM>(psd, e.y, a));
            }
        }
        re = fmaf(xr, g, gf * e.z);
        im = fmaf(xi, g, gf * e.w);
        return a.filter_type == 3 && psd >= a.pmin && psd <= a.pmax;
    }
}

// Coefficient (kt, ky, kx) owned by middle-stage lane `lane` at temporal pair g and slot
// (A of evaluation 0, B of 0, A of 1, B of 1); see the middle stage of fused_kernel.
__device__ __forceinline__ int3 slot_position(int g, int slot, int lane, int t_len) {
    const int pp = lane >> 3, q = lane & 7;
    const int e = slot >> 1, k = slot & 1;
    const int ky = pp != 0 ? (e == 0 ? pp : (16 - pp) & 15) : (e == 0 ? 0 : 8);
    if (q != 0) {
        return k == 0 ? make_int3(g, ky, q) : make_int3((t_len - g) % t_len, (16 - ky) & 15, q);
    }
    return make_int3(g, ky, k == 0 ? 0 : 8);
}

// Conjugate partner whose sigma is averaged with `p` at kx = 0, 8 (itself elsewhere).
__device__ __forceinline__ int3 partner_position(int3 p, int t_len) {
    if (p.z == 0 || p.z == 8) {
        return make_int3((t_len - p.x) % t_len, (16 - p.y) & 15, p.z);
    }
    return p;
}

__device__ __forceinline__ int sigma_index(int3 p) {
    return (p.x * 16 + p.y) * 9 + p.z;
}

// Index of the last job whose first work item (kKey 0), first job-frame (1) or first
// frame-group work item (2, the cross-frame form of the fused kernel) is <= v: the job
// table is sorted by each.  Jobs without work in a chunk share their start with the
// next job and are skipped by taking the last match.
template <int kKey>
__device__ __forceinline__ int find_job(const JobDev* jobs, int n_jobs, long long v) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const long long start = kKey == 2 ? jobs[mid].group_begin : (kKey == 1 ? jobs[mid].frame_begin : jobs[mid].item_begin);
        if (start <= v) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

// Gain table into LDS: the plan's table, with sigma taken from a.sigma_dev when given.
// kBias: sigma + 1e-15 (the form gain_of<0> takes; the plan's table of a Wiener plan
// holds it already) instead of sigma.
template <int R, bool kBias>
__device__ __forceinline__ void fill_tab(float4* tab, const LaunchArgs& a, int tid, int n_threads) {
    constexpr int T = 2 * R + 1;
    if (a.tab_dev) { // built for this launch by gain_table_kernel
        for (int i = tid; i < (R + 1) * 4 * 64; i += n_threads) {
            tab[i] = a.tab_dev[i];
        }
        return;
    }
    const float4* gtab = reinterpret_cast<const float4*>(a.consts + kOffTAB);
    for (int i = tid; i < (R + 1) * 4 * 64; i += n_threads) {
        float4 e = gtab[i];
        if (a.sigma_dev) {
            const int lane = i & 63, slot = (i >> 6) & 3, g = i >> 8;
            const int3 pos = slot_position(g, slot, lane, T);
            e.x = a.sigma_dev[sigma_index(pos)] * a.sigma_scale;
            e.y = a.sigma_dev[sigma_index(partner_position(pos, T))] * a.sigma_scale;
            if constexpr (kBias) {
                e.x += 1e-15f;
                e.y += 1e-15f;
            }
        }
        tab[i] = e;
    }
}

} // namespace
} // namespace dft2

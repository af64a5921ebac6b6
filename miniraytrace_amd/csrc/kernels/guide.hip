// Path-guiding network on gfx950 matrix cores: BF16 operands, FP32 accumulate.
//
// Network: 78 (padded 80) -> 64 -> 64 -> 64 -> 64 -> 40, ReLU on the hidden
// layers (guide_device.h holds the encoding and the decoding of the outputs).
//
// Every matrix product uses v_mfma_f32_32x32x16_bf16.  Fragments of one
// 32x32 output tile C = A (32 x 16) * B (16 x 32) for lane l (0..63),
// h = l / 32:
//   A: row m = l % 32, k = 8 h + {0..7}        (8 bf16, one 16-byte load)
//   B: col n = l % 32, k = 8 h + {0..7}
//   C: col n = l % 32, rows m = 8 j + 4 h + i in acc[4 j + i]
// A and B fill their k slots from the same k, so the product does not depend
// on how the instruction orders k internally; only the C layout matters.
//
// Inference keeps one wave's 32 rows in registers between layers (lane-half
// swaps turn a layer's outputs into the next layer's operands); the training
// forward passes them through LDS (row-major), from where it also stores
// them.  Training stores inputs, activations and deltas row-block-tiled
// (storeTiled) so every operand load of the weight-gradient products is one
// contiguous kilobyte in HBM.
//
// Kernels:
//   kGuidePack        FP32 params -> packed bf16 forward image and transposed image
//   kGuideInfer       fused forward for rendering (features -> mixture)
//   kGuideTrainRows   per 32-row tile of a training batch: encoding, forward,
//                     d(weighted NLL)/d(raw outputs) and the backward pass
//                     delta_{l-1} = (W_l^T delta_l) * relu'(a_{l-1}), l = 4 .. 1
//   kGuideWeightGrad  per row chunk, all layers in one launch after the
//                     backward pass: grad W_l = delta_l a_{l-1}^T, grad b_l
//   kGuideReduceStep  sum of the chunk partials (fixed order, no atomics),
//                     then the Adam update and the bf16 copies of the weights
//   kGuideCollect     path records -> weighted training records
// All counts are read from device memory, so a captured graph replays the
// same launches whatever the batch sizes turn out to be.
#include "guide_device.h"
#include "guide_launch.h"

namespace mrt {
namespace dev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Round to nearest even: one v_cvt_pk_bf16_f32 per one or two values.
__device__ __forceinline__ uint16_t toBf16(float f) { return __builtin_bit_cast(uint16_t, __bf16(f)); }
__device__ __forceinline__ float fromBf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) { return uint32_t(lo) | (uint32_t(hi) << 16); }
// bf16(lo) | bf16(hi) << 16
__device__ __forceinline__ uint32_t pack2f(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2{ lo, hi }), bf16x2));
}

// Eight consecutive bf16 at a 16-byte aligned address (LDS or global).
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
    for (int i = 0; i < 16; ++i) z[i] = 0.0f;
    return z;
}

// Lanes of one wave hand rows to each other through LDS (written by some
// lanes, read by others, often with another access width).  A wavefront-
// scope fence keeps the compiler from moving those LDS accesses across the
// hand-off; the hardware executes one wave's LDS operations in order.
__device__ __forceinline__ void waveHandoff() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

__device__ __forceinline__ float* wordF(uint32_t* w, uint32_t i) { return reinterpret_cast<float*>(w + i); }
__device__ __forceinline__ uint32_t pad32(uint32_t n) { return (n + 31u) & ~31u; }

__device__ __forceinline__ const float4* recRow(const RecView& v, uint32_t i) {
    return v.base + (size_t(v.phase) + size_t(i) * v.period) * v.stride;
}
__device__ __forceinline__ uint32_t batchCount(const GuideBatch& b) {
    uint32_t n = *b.total;
    if (n > b.poolCap) n = b.poolCap;
    const uint32_t c = n > b.phase ? (n - b.phase + b.period - 1) / b.period : 0u;
    return c < b.cap ? c : b.cap;
}
__device__ __forceinline__ RecView batchView(const GuideBatch& b) {
    RecView v;
    v.base = b.pool;
    v.stride = 4;
    v.phase = b.phase;
    v.period = b.period;
    return v;
}

// ------------------------------------------------------------ forward image --
// W_l output-major, 64 rows per layer (the output layer's 40 padded with
// zeros), then 5 x 64 FP32 biases.  The image is copied to LDS as it is, and
// its row strides (88 bf16 for layer 0's 80 inputs, 72 for the 64 of the
// others; 176 and 144 bytes) spread the 32 rows one MFMA operand load reads
// over 16 bank offsets: two lanes share a bank, where 64-element rows (128
// bytes) would put 16 on one.  Pad columns are zero.
constexpr int kWRow0 = 88;
constexpr int kWRow = 72;
constexpr int kWOff1 = 64 * kWRow0;
constexpr int kWTotal = kWOff1 + 4 * 64 * kWRow;         // 24,064 bf16
constexpr int kImageU4 = (kWTotal * 2 + 320 * 4) / 16;  // 3,088
constexpr int kImageWU4 = kWTotal * 2 / 16;              // 3,008
static_assert(kWTotal * 2 + 320 * 4 == int(kGuideImageBytes), "image size");
static_assert(kWRow0 >= kGuideInPad && kWRow >= 64, "image rows");
__host__ __device__ constexpr int wOff(int l) { return l == 0 ? 0 : kWOff1 + (l - 1) * 64 * kWRow; }
__host__ __device__ constexpr int wStride(int l) { return l == 0 ? kWRow0 : kWRow; }
// Per-wave row buffers in LDS: bf16 activations at a row stride of kActStride
// (176 bytes; 80 encoded inputs or 64 hidden outputs), reused after the last
// layer as FP32 raw outputs at a row stride of kRawStride (also 176 bytes).
// Both strides put the 32 rows of a tile on 16 bank offsets.
constexpr int kActStride = 88;
constexpr int kRawStride = 44;
static_assert(kActStride >= kGuideInPad && kRawStride >= kGuideOut && kRawStride * 4 <= kActStride * 2, "row buffers");

// Transposed image: W_l^T of layers 1..4 (inputs x outputs, the output
// layer's 40 outputs padded with zeros) at the forward image's hidden row
// stride, the A operand of the backward products.
constexpr int kWTTotal = 4 * 64 * kWRow;  // 18,432 bf16
constexpr int kImageTU4 = kWTTotal * 2 / 16;
static_assert(kWTTotal * 2 == int(kGuideImageTBytes), "transposed image size");
__host__ __device__ constexpr int wtOff(int l) { return (l - 1) * 64 * kWRow; }

__global__ __launch_bounds__(256) void kGuidePack(const float* __restrict__ params, uint16_t* __restrict__ image,
                                                  uint16_t* __restrict__ imageT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kWTotal + 320) {
        const int t = i - (kWTotal + 320);
        if (t >= kWTTotal) return;
        const int l = 1 + t / (64 * kWRow), in = (t % (64 * kWRow)) / kWRow, o = t % kWRow;
        float v = 0.0f;
        if (o < guideLayerOut(l) && in < guideLayerIn(l)) v = params[guideWeightOffset(l) + o * guideLayerIn(l) + in];
        imageT[t] = toBf16(v);
    } else if (i < kWTotal) {
        const int l = i < kWOff1 ? 0 : 1 + (i - kWOff1) / (64 * kWRow);
        const int r = (i - wOff(l)) / wStride(l), c = (i - wOff(l)) % wStride(l);
        float v = 0.0f;
        if (r < guideLayerOut(l) && c < guideLayerIn(l)) v = params[guideWeightOffset(l) + r * guideLayerIn(l) + c];
        image[i] = toBf16(v);
    } else if (i < kWTotal + 320) {
        const int j = i - kWTotal, l = j / 64, o = j % 64;
        reinterpret_cast<float*>(image + kWTotal)[j] = o < guideLayerOut(l) ? params[guideBiasOffset(l) + o] : 0.0f;
    }
}

__device__ __forceinline__ void loadImage(const u32x4* __restrict__ image, uint16_t* W, float* bias) {
    u32x4* w4 = reinterpret_cast<u32x4*>(W);
    u32x4* b4 = reinterpret_cast<u32x4*>(bias);
    for (int i = threadIdx.x; i < kImageU4; i += blockDim.x) {
        const u32x4 v = image[i];
        if (i < kImageWU4) w4[i] = v;
        else b4[i - kImageWU4] = v;
    }
}

// One layer over the wave's 32 rows in act (row-major, stride kActStride).
// The weights are the A operand and the rows the B operand, so lane (r, h)
// ends with outputs 8 j + 4 h + i (c0) and 32 + 8 j + 4 h + i (c1) of row r:
// four consecutive features of one row, stored back with one 8-byte write.
// Hidden layers write ReLU(W a + b) back into act as bf16; the output layer
// leaves W a + b in the two accumulator tiles.
__device__ __forceinline__ void layerForward(int l, uint16_t* act, const uint16_t* W, const float* bias,
                                             f32x16& c0, f32x16& c1) {
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    const int K = l == 0 ? kGuideInPad : 64;
    const uint16_t* Wl = W + wOff(l);
    const int ws = wStride(l);
    c0 = zero16();
    c1 = zero16();
    waveHandoff();
    for (int kb = 0; kb < K; kb += 16) {
        const bf16x8 x = ld8(act + r * kActStride + kb + 8 * h);
        c0 = mfma(ld8(Wl + r * ws + kb + 8 * h), x, c0);
        c1 = mfma(ld8(Wl + (32 + r) * ws + kb + 8 * h), x, c1);
    }
    for (int j = 0; j < 4; ++j) {
        const float4 b0 = *reinterpret_cast<const float4*>(bias + l * 64 + 8 * j + 4 * h);
        const float4 b1 = *reinterpret_cast<const float4*>(bias + l * 64 + 32 + 8 * j + 4 * h);
        c0[4 * j] += b0.x;
        c0[4 * j + 1] += b0.y;
        c0[4 * j + 2] += b0.z;
        c0[4 * j + 3] += b0.w;
        c1[4 * j] += b1.x;
        c1[4 * j + 1] += b1.y;
        c1[4 * j + 2] += b1.z;
        c1[4 * j + 3] += b1.w;
    }
    if (l < 4) {
        auto relu4 = [](const f32x16& c, int j) {
            return make_uint2(pack2f(fmaxf(c[4 * j], 0.0f), fmaxf(c[4 * j + 1], 0.0f)),
                              pack2f(fmaxf(c[4 * j + 2], 0.0f), fmaxf(c[4 * j + 3], 0.0f)));
        };
        for (int j = 0; j < 4; ++j) {
            *reinterpret_cast<uint2*>(act + r * kActStride + 8 * j + 4 * h) = relu4(c0, j);
            *reinterpret_cast<uint2*>(act + r * kActStride + 32 + 8 * j + 4 * h) = relu4(c1, j);
        }
    }
    waveHandoff();
}

// sin(pi t) and cos(pi t) for |t| <= 1: reflection to a = |t| <= 1/2 (or
// 1 - |t|), then the Taylor series of sin(pi a) to degree 11 and of
// cos(pi a) to degree 12 (truncation error below 1e-7), evaluated without
// contraction so that the oracle repeats it bit for bit.
__device__ __forceinline__ void sinCosPiSeries(float t, float& s, float& c) {
#pragma clang fp contract(off)
    const float a0 = fabsf(t);
    const bool reflect = a0 > 0.5f;
    const float a = reflect ? 1.0f - a0 : a0;
    const float z = a * a;
    float ps = -0x1.e3075p-8f;  // -pi^11 / 11!
    ps = ps * z + 0x1.507834p-4f;
    ps = ps * z + -0x1.32d2ccp-1f;
    ps = ps * z + 0x1.466bc6p+1f;
    ps = ps * z + -0x1.4abbcep+2f;
    ps = ps * z + 0x1.921fb6p+1f;  // pi
    float pc = 0x1.f9d38ap-10f;  // pi^12 / 12!
    pc = pc * z + -0x1.a6d1f2p-6f;
    pc = pc * z + 0x1.e1f506p-3f;
    pc = pc * z + -0x1.55d3c8p+0f;
    pc = pc * z + 0x1.03c1fp+2f;
    pc = pc * z + -0x1.3bd3ccp+2f;  // -pi^2 / 2
    s = ps * a;
    c = pc * z + 1.0f;
    if (reflect) c = -c;
    if (t < 0.0f) s = -s;
}

// The three query float4 of one row (pos01, normal, wo); zero past count.
struct RowIn {
    float4 a, b, c;
};
__device__ __forceinline__ RowIn loadRow(const RecView& rv, uint32_t row, uint32_t count) {
    RowIn q{ make_float4(0, 0, 0, 0), make_float4(0, 0, 0, 0), make_float4(0, 0, 0, 0) };
    if (row < count) {
        const float4* p = recRow(rv, row);
        q.a = p[0];
        q.b = p[1];
        q.c = p[2];
    }
    return q;
}

// Encodes the wave's 32 rows into act; lanes r and r + 32 both hold row r's
// query (in).  Feature pair j = 12 a + b is (sin, cos)(pi 2^b p_a): band 0
// from the series after exact reduction of p_a to [-1, 1], each further band
// by the double-angle step (s, c) -> (2 s c, (c - s)(c + s)).  Lanes 0..31
// write pairs 0..17 of row lane, lanes 32..63 pairs 18..35, the normal, wo
// and two zero pads.  Invalid rows (past count) are zero.
__device__ __forceinline__ void encodeRows(const RowIn& in, bool valid, uint16_t* act) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    const float4 a = in.a, b = in.b, c = in.c;
    const float p[3] = { a.x, a.y, a.z };
    uint32_t* dst = reinterpret_cast<uint32_t*>(act + r * kActStride);
    waveHandoff();  // the previous tile's reads of act come first
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float sn, cs;
        sinCosPiSeries(p[ax] - 2.0f * rintf(0.5f * p[ax]), sn, cs);
#pragma unroll
        for (int band = 0; band < kGuideBands; ++band) {
            if (band > 0) {
                const float s2 = (2.0f * sn) * cs;
                cs = (cs - sn) * (cs + sn);
                sn = s2;
            }
            const int j = 12 * ax + band;  // feature pair; this lane half writes 18 h .. 18 h + 17
            if (j / 18 == h) dst[j] = valid ? pack2f(sn, cs) : 0u;
        }
    }
    if (h == 1) {
        dst[36] = valid ? pack2f(b.x, b.y) : 0u;
        dst[37] = valid ? pack2f(b.z, c.x) : 0u;
        dst[38] = valid ? pack2f(c.y, c.z) : 0u;
        dst[39] = 0u;
    }
    waveHandoff();
}

// Output tiles of rows [row0, row0 + 32) as FP32 into act (reused as
// float[32][40]).
__device__ __forceinline__ void stashOutputs(const f32x16& c0, const f32x16& c1, uint16_t* act) {
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    float* fo = reinterpret_cast<float*>(act) + r * kRawStride;
    for (int j = 0; j < 4; ++j)
        *reinterpret_cast<float4*>(fo + 8 * j + 4 * h) = make_float4(c0[4 * j], c0[4 * j + 1], c0[4 * j + 2],
                                                                     c0[4 * j + 3]);
    // outputs 32..39: j = 0 of the second tile
    *reinterpret_cast<float4*>(fo + 32 + 4 * h) = make_float4(c1[0], c1[1], c1[2], c1[3]);
    waveHandoff();
}

// Copy of the wave's 32 rows into a row-block-tiled array of F features:
// element (f, row) at ((row / 16) F + f) 16 + row % 16, so the 16 rows of
// one feature in a block are 32 contiguous bytes and one MFMA operand load
// of the weight gradient (32 features x 16 rows) is one contiguous
// kilobyte.  Features [0, nfeat) come from act (row-major, stride
// kActStride); features [nfeat, F) are written as zeros.
__device__ __forceinline__ void storeTiled(const uint16_t* act, int nfeat, int F, uint16_t* __restrict__ dst,
                                           uint32_t row0) {
    const int lane = threadIdx.x & 63;
    for (int f = lane; f < F; f += 64) {
        uint32_t w[16];
        for (int m = 0; m < 16; ++m)
            w[m] = f < nfeat ? pack2(act[(2 * m) * kActStride + f], act[(2 * m + 1) * kActStride + f]) : 0u;
        for (int blk = 0; blk < 2; ++blk) {
            u32x4* d = reinterpret_cast<u32x4*>(dst + (size_t(row0 / 16 + blk) * F + f) * 16);
            d[0] = u32x4{ w[8 * blk], w[8 * blk + 1], w[8 * blk + 2], w[8 * blk + 3] };
            d[1] = u32x4{ w[8 * blk + 4], w[8 * blk + 5], w[8 * blk + 6], w[8 * blk + 7] };
        }
    }
}

// ------------------------------------------- register-resident forward pass --
// The inference kernel keeps a tile's activations in registers.  With the
// weights as the A operand, lane (r, h) ends a layer with features
// 8 j + 4 h .. 8 j + 4 h + 3 of row r (j = 0..3 in c0, the same plus 32 in
// c1), and the next layer's B operand of k step k needs features
// 16 k + 8 h .. 16 k + 8 h + 7 of row r: group j = 2 k' + h (k' = k within
// the 32-feature tile) of both lane halves.  A lane keeps its own group and
// trades the other with its partner lane r + 32 (or r - 32).

// v_permlane32_swap: lanes 32..63 of a trade places with lanes 0..31 of b.
// With a = group 2 k' and b = group 2 k' + 1 of every lane, a ends holding
// features +0..3 of the operand (lane half 0: its own group 2 k'; half 1:
// group 2 k' + 1 of its partner, i.e. features +8..11) and b the rest, so
// the operand is (a, b) in every lane.
__device__ __forceinline__ void swapHalves(uint32_t& a, uint32_t& b) {
    const auto s = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = s[0];
    b = s[1];
}

// ReLU of two packed bf16: as int16 a negative bf16 is negative, so max
// with 0 zeroes exactly the negative values (rounding then ReLU equals ReLU
// then rounding).
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t reluBf16x2(uint32_t v) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, v), s16x2{ 0, 0 }));
}

// Layer-0 B operands of the wave's 32 rows, encoded in registers exactly as
// encodeRows writes them: lane (r, h) gets feature pairs 8 k + 4 h ..
// 8 k + 4 h + 3 of row r for k steps k = 0..4 (pairs 0..35 the sin/cos bands,
// 36..38 normal and wo, 39 zero).  Invalid rows are zero.
__device__ __forceinline__ void encodeOperands(const RowIn& in, bool valid, u32x4 (&b)[5]) {
#pragma clang fp contract(off)
    const bool hi = (threadIdx.x & 32) != 0;
    uint32_t pr[40];
    const float p[3] = { in.a.x, in.a.y, in.a.z };
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float sn, cs;
        sinCosPiSeries(p[ax] - 2.0f * rintf(0.5f * p[ax]), sn, cs);
#pragma unroll
        for (int band = 0; band < kGuideBands; ++band) {
            if (band > 0) {
                const float s2 = (2.0f * sn) * cs;
                cs = (cs - sn) * (cs + sn);
                sn = s2;
            }
            pr[12 * ax + band] = pack2f(sn, cs);
        }
    }
    pr[36] = pack2f(in.b.x, in.b.y);
    pr[37] = pack2f(in.b.z, in.c.x);
    pr[38] = pack2f(in.c.y, in.c.z);
    pr[39] = 0u;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        uint32_t q[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i] = valid ? (hi ? pr[8 * k + 4 + i] : pr[8 * k + i]) : 0u;
        b[k] = u32x4{ q[0], q[1], q[2], q[3] };
    }
}

// One layer from register B operands (KS k steps); outputs as in
// layerForward, with the accumulators starting from the bias.
template <int KS>
__device__ __forceinline__ void layerOperands(int l, const u32x4 (&x)[KS], const uint16_t* W, const float* bias,
                                              f32x16& c0, f32x16& c1) {
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    const uint16_t* Wl = W + wOff(l);
    const int ws = wStride(l);
    for (int j = 0; j < 4; ++j) {
        const float4 b0 = *reinterpret_cast<const float4*>(bias + l * 64 + 8 * j + 4 * h);
        const float4 b1 = *reinterpret_cast<const float4*>(bias + l * 64 + 32 + 8 * j + 4 * h);
        c0[4 * j] = b0.x;
        c0[4 * j + 1] = b0.y;
        c0[4 * j + 2] = b0.z;
        c0[4 * j + 3] = b0.w;
        c1[4 * j] = b1.x;
        c1[4 * j + 1] = b1.y;
        c1[4 * j + 2] = b1.z;
        c1[4 * j + 3] = b1.w;
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const bf16x8 xb = __builtin_bit_cast(bf16x8, x[k]);
        c0 = mfma(ld8(Wl + r * ws + 16 * k + 8 * h), xb, c0);
        c1 = mfma(ld8(Wl + (32 + r) * ws + 16 * k + 8 * h), xb, c1);
    }
}

// Hidden outputs -> bf16 -> ReLU -> the next layer's four B operands.
__device__ __forceinline__ void nextOperands(const f32x16& c0, const f32x16& c1, u32x4 (&x)[4]) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const f32x16& c = t ? c1 : c0;
        uint32_t g[4][2];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            g[j][0] = reluBf16x2(pack2f(c[4 * j], c[4 * j + 1]));
            g[j][1] = reluBf16x2(pack2f(c[4 * j + 2], c[4 * j + 3]));
        }
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            uint32_t a0 = g[2 * kk][0], a1 = g[2 * kk][1], b0 = g[2 * kk + 1][0], b1 = g[2 * kk + 1][1];
            swapHalves(a0, b0);
            swapHalves(a1, b1);
            x[2 * t + kk] = u32x4{ a0, a1, b0, b1 };
        }
    }
}

// ------------------------------------------------------------------ inference --
constexpr int kFwdWaves = 16;  // one 1,024-thread block per CU: 48 KB weights + 16 x 5.5 KB rows
__global__ __launch_bounds__(kFwdWaves * 64) void kGuideInfer(RecView rv, const uint32_t* __restrict__ countPtr,
                                                              const u32x4* __restrict__ image, void* __restrict__ out,
                                                              int decoded) {
    __shared__ __attribute__((aligned(16))) uint16_t W[kWTotal];
    __shared__ __attribute__((aligned(16))) float bias[5 * 64];
    __shared__ __attribute__((aligned(16))) uint16_t actAll[kFwdWaves][32 * kActStride];
    const uint32_t count = *countPtr;
    if (blockIdx.x * kFwdWaves * 32 >= count) return;  // uniform per block
    loadImage(image, W, bias);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* act = actAll[wave];
    const float* fo = reinterpret_cast<const float*>(act);
    const uint32_t totalWaves = gridDim.x * kFwdWaves;
    // The next tile's queries are loaded while the current tile computes.
    uint32_t blk = blockIdx.x * kFwdWaves + wave;
    RowIn next = loadRow(rv, blk * 32 + (lane & 31), count);
    for (; blk * 32 < count; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        const RowIn cur = next;
        next = loadRow(rv, (blk + totalWaves) * 32 + (lane & 31), count);
        u32x4 x0[5];
        encodeOperands(cur, row0 + (lane & 31) < count, x0);
        f32x16 c0, c1;
        layerOperands<5>(0, x0, W, bias, c0, c1);
        for (int l = 1; l < 5; ++l) {
            u32x4 x[4];
            nextOperands(c0, c1, x);
            layerOperands<4>(l, x, W, bias, c0, c1);
        }
        waveHandoff();  // the previous tile's reads of the stash come first
        stashOutputs(c0, c1, act);
        if (decoded) {
            // lanes r and r + 32 decode lobes 0..3 and 4..7 of row r
            const int r = lane & 31;
            const uint32_t row = row0 + r;
            storeMixHalf(fo + r * kRawStride, lane >> 5, row < count,
                         static_cast<float4*>(out) + size_t(row) * kGuideMixF4);
        } else {
            float* o = static_cast<float*>(out);
            for (int i = lane; i < 32 * kGuideOut; i += 64) {
                const uint32_t row = row0 + i / kGuideOut;
                if (row < count) o[size_t(row) * kGuideOut + i % kGuideOut] = fo[(i / kGuideOut) * kRawStride + i % kGuideOut];
            }
        }
    }
}

// ------------------------------------------------------------------- training --
// One Adam step is three launches:
//   kGuideTrainRows   per 32-row tile, one wave: encoding, forward, output
//                     gradient and the backward pass through the hidden
//                     layers; stores the inputs, activations and deltas
//                     feature-major for the weight gradient, and the tile's
//                     weight and loss sums
//   kGuideWeightGrad  weight gradients of all five layers, per row chunk
//   kGuideReduceStep  chunk partials summed in chunk order and divided by the
//                     batch weight; with Adam, the update of the FP32 master
//                     weights and of their bf16 copies in both images
// The deltas are derivatives of the unnormalised sum_j w_j (-log q(wi_j)),
// rounded to bf16; the gradient is divided by the batch's weight sum once,
// after the reduction.  Every sum runs in a fixed order, so a step is
// bit-reproducible.

__device__ __forceinline__ void loadImageT(const u32x4* __restrict__ imageT, uint16_t* WT) {
    u32x4* w4 = reinterpret_cast<u32x4*>(WT);
    for (int i = threadIdx.x; i < kImageTU4; i += blockDim.x) w4[i] = imageT[i];
}

// A training row: the query (pos01 + weight, normal, wo) and the direction wi.
struct TrainRow {
    RowIn q;
    float4 wi;
};
__device__ __forceinline__ TrainRow loadTrainRow(const RecView& rv, uint32_t row, uint32_t count) {
    TrainRow t;
    t.q = loadRow(rv, row, count);
    t.wi = row < count ? recRow(rv, row)[3] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return t;
}

// ReLU bits of the hidden outputs a lane holds after layerForward (which
// stored bf16(max(c, 0))): bit 4 j + i for output 8 j + 4 h + i (c0), bit
// 16 + 4 j + i for output 32 + 8 j + 4 h + i (c1), set where the stored
// activation is > 0.
__device__ __forceinline__ uint32_t reluBits(const f32x16& c0, const f32x16& c1) {
    uint32_t m = 0;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        m |= uint32_t(fromBf16(toBf16(fmaxf(c0[s], 0.0f))) > 0.0f) << s;
        m |= uint32_t(fromBf16(toBf16(fmaxf(c1[s], 0.0f))) > 0.0f) << (16 + s);
    }
    return m;
}

// d/d raw of wgt (-log q(wi)) for lobes 4 h .. 4 h + 3 of the lane's row
// (raw outputs 20 h .. 20 h + 19 into d), q the row's decoded mixture.
// Lanes r and r + 32 hold the same row and share the softmax and the
// log-sum-exp through shuffles, so every lane of the wave must call it.
// Returns -wgt log q.
__device__ __forceinline__ float outputGradHalf(const float* raw, int h, float wgt, V3 wi, float (&d)[20]) {
    float mxl = -1e30f;
    for (int k = 0; k < kGuideLobes; ++k) mxl = fmaxf(mxl, raw[5 * k + 4]);
    float e[4], s = 0.0f;
    for (int i = 0; i < 4; ++i) {
        e[i] = expf(raw[5 * (4 * h + i) + 4] - mxl);
        s += e[i];
    }
    const float other = __shfl_xor(s, 32, 64);
    const float inv = 1.0f / (h == 0 ? s + other : other + s);  // lobes 0..3, then 4..7 (decodeMix's order)
    V3 mu[4];
    float kap[4], c[4], lf[4], w[4];
    float mx = -1e30f;
    for (int i = 0; i < 4; ++i) {
        decodeLobe(raw, 4 * h + i, mu[i], kap[i]);
        w[i] = e[i] * inv;
        c[i] = dot(mu[i], wi);
        lf[i] = logf(w[i]) + logf(kap[i]) - logf(kTwoPi) - logf(-expm1f(-2.0f * kap[i])) + kap[i] * (c[i] - 1.0f);
        mx = fmaxf(mx, lf[i]);
    }
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    float se = 0.0f;
    for (int i = 0; i < 4; ++i) se += expf(lf[i] - mx);
    const float seo = __shfl_xor(se, 32, 64);
    const float logq = mx + logf(h == 0 ? se + seo : seo + se);
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * h + i;
        const float resp = expf(lf[i] - logq);
        const float lk = raw[5 * k + 3];
        const bool inside = lk > kLogKappaLo && lk < kLogKappaHi;
        const float dlogf = 1.0f / kap[i] - 2.0f / expm1f(2.0f * kap[i]) + (c[i] - 1.0f);
        const V3 v = v3(raw[5 * k], raw[5 * k + 1], raw[5 * k + 2]);
        const float vn = fmaxf(sqrtf(dot(v, v)), 1e-12f);
        const float gmu = -wgt * resp * kap[i] / vn;
        d[5 * i + 0] = gmu * (wi.x - c[i] * mu[i].x);
        d[5 * i + 1] = gmu * (wi.y - c[i] * mu[i].y);
        d[5 * i + 2] = gmu * (wi.z - c[i] * mu[i].z);
        d[5 * i + 3] = inside ? -wgt * resp * dlogf * kap[i] : 0.0f;
        d[5 * i + 4] = -wgt * (resp - w[i]);
    }
    return -wgt * logq;
}

// delta_{l-1} of the wave's 32 rows (l = 4 .. 1): layer l's inputs are the A
// operand's rows (W_l^T from WT), the rows of delta_l in act (row-major)
// the B operand, so lane (r, h) ends with inputs 8 j + 4 h + i (c0) and
// 32 + 8 j + 4 h + i (c1) of row r: the layout reluBits recorded layer
// l - 1's outputs in.  Masked, rounded to bf16 and written back into act.
__device__ __forceinline__ void layerBackward(int l, uint16_t* act, const uint16_t* WT, uint32_t mask) {
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    const uint16_t* Wl = WT + wtOff(l);
    f32x16 c0 = zero16(), c1 = zero16();
    waveHandoff();
    for (int kb = 0; kb < 64; kb += 16) {
        const bf16x8 x = ld8(act + r * kActStride + kb + 8 * h);
        c0 = mfma(ld8(Wl + r * kWRow + kb + 8 * h), x, c0);
        c1 = mfma(ld8(Wl + (32 + r) * kWRow + kb + 8 * h), x, c1);
    }
    auto masked4 = [mask](const f32x16& c, int j, int bit0) {
        float v[4];
        for (int i = 0; i < 4; ++i) v[i] = (mask >> (bit0 + i)) & 1u ? c[4 * j + i] : 0.0f;
        return make_uint2(pack2f(v[0], v[1]), pack2f(v[2], v[3]));
    };
    for (int j = 0; j < 4; ++j) {
        *reinterpret_cast<uint2*>(act + r * kActStride + 8 * j + 4 * h) = masked4(c0, j, 4 * j);
        *reinterpret_cast<uint2*>(act + r * kActStride + 32 + 8 * j + 4 * h) = masked4(c1, j, 16 + 4 * j);
    }
    waveHandoff();
}

// Everything of a training step that is per row, one wave per 32-row tile
// of the batch's padded rows [0, pad32(count)).  Block 0 publishes the batch
// size and advances the Adam step counter (countStep, non-empty batch).
// Padding rows encode to zero inputs and get zero weight and deltas.
constexpr int kTrainWaves = 8;  // 47 KB forward image + 36 KB transposed image + 8 x 5.5 KB rows
__global__ __launch_bounds__(kTrainWaves * 64) void kGuideTrainRows(GuideBatch b, GuideNet net, int countStep) {
    __shared__ __attribute__((aligned(16))) uint16_t W[kWTotal];
    __shared__ __attribute__((aligned(16))) float bias[5 * 64];
    __shared__ __attribute__((aligned(16))) uint16_t WT[kWTTotal];
    __shared__ __attribute__((aligned(16))) uint16_t actAll[kTrainWaves][32 * kActStride];
    const uint32_t count = batchCount(b);
    const uint32_t rows = pad32(count);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        net.words[kGwBatch] = count;
        if (countStep && count > 0) net.words[kGwStep] += 1u;
    }
    if (blockIdx.x * kTrainWaves * 32 >= rows) return;  // uniform per block
    loadImage(static_cast<const u32x4*>(net.image), W, bias);
    loadImageT(static_cast<const u32x4*>(net.imageT), WT);
    __syncthreads();
    const RecView rv = batchView(b);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    uint16_t* act = actAll[wave];
    const float* fo = reinterpret_cast<const float*>(act);
    const uint32_t stride = net.stride;
    const uint32_t totalWaves = gridDim.x * kTrainWaves;
    uint32_t blk = blockIdx.x * kTrainWaves + wave;
    TrainRow next = loadTrainRow(rv, blk * 32 + r, count);
    for (; blk * 32 < rows; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        const TrainRow cur = next;
        next = loadTrainRow(rv, (blk + totalWaves) * 32 + r, count);
        const bool valid = row0 + r < count;
        encodeRows(cur.q, valid, act);
        storeTiled(act, kGuideInPad, 96, net.actIn, row0);
        uint32_t mask[4];
        f32x16 c0, c1;
        for (int l = 0; l < 5; ++l) {
            layerForward(l, act, W, bias, c0, c1);
            if (l < 4) {
                mask[l] = reluBits(c0, c1);
                storeTiled(act, 64, 64, net.act + size_t(l) * 64 * stride, row0);
            }
        }
        waveHandoff();  // layer 3's stores have read act
        stashOutputs(c0, c1, act);
        const float wgt = valid ? cur.q.a.w : 0.0f;
        float d[20];
        const float loss = outputGradHalf(fo + r * kRawStride, h, wgt, v3(cur.wi), d);
        waveHandoff();  // every lane has read its raw outputs
        // output-layer deltas, row-major: lane half h writes outputs 20 h ..
        // 20 h + 19 of row r, half 1 also the zero features 40..63
        uint32_t* drow = reinterpret_cast<uint32_t*>(act + r * kActStride);
        for (int m = 0; m < 10; ++m) drow[10 * h + m] = valid ? pack2f(d[2 * m], d[2 * m + 1]) : 0u;
        if (h == 1)
            for (int m = 20; m < 32; ++m) drow[m] = 0u;
        waveHandoff();
        storeTiled(act, 64, 64, net.delta[4], row0);
        for (int l = 4; l >= 1; --l) {
            layerBackward(l, act, WT, mask[l - 1]);
            storeTiled(act, 64, 64, net.delta[l - 1], row0);
        }
        // the tile's weight and loss sums (each row counted by its h = 0 lane)
        float sw = h == 0 ? wgt : 0.0f, sl = h == 0 && valid ? loss : 0.0f;
        for (int off = 32; off > 0; off >>= 1) {
            sw += __shfl_xor(sw, off, 64);
            sl += __shfl_xor(sl, off, 64);
        }
        if (lane == 0) *reinterpret_cast<float2*>(net.tilePart + 2 * size_t(blk)) = make_float2(sw, sl);
    }
}

// One launch forms the weight gradients of all five layers.  Block job =
// (row chunk, layer): the block's 8 waves take consecutive eighths of the
// chunk's rows, and each wave forms every 32 x 32 tile of the layer's
// gradient (outputs 32 mt.., inputs 32 nt..; layer 0: 3 input tiles over the
// 96 zero-padded input features) and the bias sums (an all-ones B operand),
// so each delta and activation fragment is read once.  Wave 0 then adds the
// other waves' accumulators in wave order through LDS, two tiles at a time
// (fixed summation order, no atomics).  The layer-0 job of each chunk also
// adds the chunk's 64 tile weight and loss sums, in tile order, into
// weightPart.
constexpr uint32_t kGradWaves = 8;
constexpr uint32_t kGradSlice = kGuideGradChunk / kGradWaves;  // rows per wave
static_assert(kGradSlice % 16 == 0, "whole MFMA k steps per wave");

template <int NT>
__device__ __forceinline__ void weightGradJob(const GuideNet& net, int l, uint32_t chunk, uint32_t rows,
                                              float (*red)[64][32]) {
    constexpr int kTiles = 2 * NT + 2;  // (mt, nt) weight tiles, then the two bias tiles
    const uint16_t* delta = net.delta[l];
    const uint16_t* actPrev = l == 0 ? net.actIn : net.act + size_t(l - 1) * 64 * net.stride;
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t s0 = chunk * kGuideGradChunk + w * kGradSlice, s1 = min(s0 + kGradSlice, rows);
    const int lane = threadIdx.x & 63, lr = lane & 31, h = lane >> 5;
    // row-block-tiled operands (storeTiled): the 16-row block of row r starts
    // at (r / 16) F 16, feature f of it at + 16 f
    constexpr int F = NT == 3 ? 96 : 64;
    const uint16_t* aRow0 = delta + lr * 16 + 8 * h;
    const uint16_t* aRow1 = aRow0 + 32 * 16;
    const uint16_t* bRow = actPrev + lr * 16 + 8 * h;
    constexpr int bStep = 32 * 16;  // next input tile
    const uint32_t one2 = 0x3F803F80u;         // two bf16 1.0
    const bf16x8 ones = __builtin_bit_cast(bf16x8, u32x4{ one2, one2, one2, one2 });
    f32x16 T[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) T[t] = zero16();
    // Four row steps' operands are loaded before their products (one wait
    // per four steps); the products accumulate in row order either way.
    constexpr int kAhead = 4;
    uint32_t r = s0;
    for (; r + 16 * kAhead <= s1; r += 16 * kAhead) {
        bf16x8 a0[kAhead], a1[kAhead], b[kAhead][NT];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const size_t blk = r / 16 + u;
            a0[u] = ld8(aRow0 + blk * 64 * 16);
            a1[u] = ld8(aRow1 + blk * 64 * 16);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[u][nt] = ld8(bRow + blk * F * 16 + nt * bStep);
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                T[nt] = mfma(a0[u], b[u][nt], T[nt]);
                T[NT + nt] = mfma(a1[u], b[u][nt], T[NT + nt]);
            }
            T[2 * NT] = mfma(a0[u], ones, T[2 * NT]);
            T[2 * NT + 1] = mfma(a1[u], ones, T[2 * NT + 1]);
        }
    }
    for (; r < s1; r += 16) {
        const size_t blk = r / 16;
        const bf16x8 a0 = ld8(aRow0 + blk * 64 * 16), a1 = ld8(aRow1 + blk * 64 * 16);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const bf16x8 b = ld8(bRow + blk * F * 16 + nt * bStep);
            T[nt] = mfma(a0, b, T[nt]);
            T[NT + nt] = mfma(a1, b, T[NT + nt]);
        }
        T[2 * NT] = mfma(a0, ones, T[2 * NT]);
        T[2 * NT + 1] = mfma(a1, ones, T[2 * NT + 1]);
    }
#pragma unroll
    for (int p = 0; p < kTiles / 2; ++p) {
        if (w != 0) {
            float4* dst = reinterpret_cast<float4*>(red[w - 1][lane]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                dst[q] = make_float4(T[2 * p][4 * q], T[2 * p][4 * q + 1], T[2 * p][4 * q + 2], T[2 * p][4 * q + 3]);
                dst[4 + q] = make_float4(T[2 * p + 1][4 * q], T[2 * p + 1][4 * q + 1], T[2 * p + 1][4 * q + 2],
                                         T[2 * p + 1][4 * q + 3]);
            }
        }
        __syncthreads();
        if (w == 0)
            for (uint32_t v = 1; v < kGradWaves; ++v) {
                const float4* src = reinterpret_cast<const float4*>(red[v - 1][lane]);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 x = src[q], y = src[4 + q];
                    T[2 * p][4 * q] += x.x;
                    T[2 * p][4 * q + 1] += x.y;
                    T[2 * p][4 * q + 2] += x.z;
                    T[2 * p][4 * q + 3] += x.w;
                    T[2 * p + 1][4 * q] += y.x;
                    T[2 * p + 1][4 * q + 1] += y.y;
                    T[2 * p + 1][4 * q + 2] += y.z;
                    T[2 * p + 1][4 * q + 3] += y.w;
                }
            }
        __syncthreads();
    }
    if (w != 0) return;
    const int nin = guideLayerIn(l), nout = guideLayerOut(l);
    float* dst = net.partial + size_t(chunk) * kGuideParams;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int o = 32 * mt + 8 * j + 4 * h + i;
                if (o >= nout) continue;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int in = 32 * nt + lr;
                    if (in < nin) dst[guideWeightOffset(l) + o * nin + in] = T[mt * NT + nt][4 * j + i];
                }
                if (lr == 0) dst[guideBiasOffset(l) + o] = T[2 * NT + mt][4 * j + i];
            }
}

__global__ __launch_bounds__(kGradWaves * 64) void kGuideWeightGrad(GuideNet net) {
    __shared__ __attribute__((aligned(16))) float red[kGradWaves - 1][64][32];
    const uint32_t rows = pad32(net.words[kGwBatch]);
    const uint32_t chunk = blockIdx.x / 5;
    const int l = int(blockIdx.x % 5);
    const uint32_t r0 = chunk * kGuideGradChunk;
    if (r0 >= rows) return;  // whole block
    if (l == 0 && threadIdx.x < 64) {
        constexpr uint32_t kTiles = kGuideGradChunk / 32;
        static_assert(kTiles <= 64, "one lane per tile");
        const uint32_t t0 = r0 / 32, tn = min(kTiles, (rows - r0) / 32);
        float2 v = threadIdx.x < tn ? *reinterpret_cast<const float2*>(net.tilePart + 2 * size_t(t0 + threadIdx.x))
                                    : make_float2(0.0f, 0.0f);
        for (int off = 32; off > 0; off >>= 1) {
            v.x += __shfl_down(v.x, off, 64);
            v.y += __shfl_down(v.y, off, 64);
        }
        if (threadIdx.x == 0) *reinterpret_cast<float2*>(net.weightPart + 2 * size_t(chunk)) = v;
    }
    if (l == 0) weightGradJob<3>(net, 0, chunk, rows, red);
    else weightGradJob<2>(net, l, chunk, rows, red);
}

// Sum of the chunks' weight and loss sums in chunk order, the same in every
// block (threads add strided slices in order, then an LDS tree).
__device__ float2 batchSums(const float* __restrict__ weightPart, uint32_t chunks) {
    __shared__ float2 part[256];
    float2 s = make_float2(0.0f, 0.0f);
    for (uint32_t i = threadIdx.x; i < chunks; i += 256) {
        s.x += weightPart[2 * i];
        s.y += weightPart[2 * i + 1];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1) {
        if (threadIdx.x < half) {
            part[threadIdx.x].x += part[threadIdx.x + half].x;
            part[threadIdx.x].y += part[threadIdx.x + half].y;
        }
        __syncthreads();
    }
    const float2 total = part[0];
    __syncthreads();
    return total;
}

// 64 parameters per block.  Thread (q, j) sums the partials of chunks q,
// q + 4, ... of parameter j in chunk order (eight loads in flight), the four
// slices are added in q order, and grad = that sum / batch weight sum.  With
// adam, threads q = 0 then apply the Adam update to the FP32 master weight
// and write its bf16 copies into the forward image and, for layers 1..4,
// the transposed image (the images' padding stays zero from
// launchGuidePack).  Block 0 publishes the normalised loss.
constexpr int kStepParams = 64;
constexpr int kStepBlock = 4 * kStepParams;
__global__ __launch_bounds__(kStepBlock) void kGuideReduceStep(GuideNet net, GuideAdam hp, int adam) {
    __shared__ float slice[4][kStepParams];
    const uint32_t n = net.words[kGwBatch];
    const uint32_t chunks = (pad32(n) + kGuideGradChunk - 1) / kGuideGradChunk;
    const float2 sums = batchSums(net.weightPart, chunks);
    const float inv = sums.x > 0.0f ? 1.0f / sums.x : 0.0f;
    const int j = threadIdx.x % kStepParams, q = threadIdx.x / kStepParams;
    const int i = blockIdx.x * kStepParams + j;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *wordF(net.words, kGwWeightSum) = sums.x;
        *wordF(net.words, kGwLoss) = sums.y * inv;
        if (adam && n > 0) {
            *wordF(net.words, kGwLastLoss) = sums.y * inv;
            net.words[kGwLastBatch] = n;
        }
    }
    float s = 0.0f;
    if (i < kGuideParams) {
        const float* src = net.partial + i;
        uint32_t c = q;
        for (; c + 28 < chunks; c += 32) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = src[size_t(c + 4 * u) * kGuideParams];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; c < chunks; c += 4) s += src[size_t(c) * kGuideParams];
    }
    slice[q][j] = s;
    __syncthreads();
    if (q != 0 || i >= kGuideParams) return;
    const float g = (((slice[0][j] + slice[1][j]) + slice[2][j]) + slice[3][j]) * inv;
    net.grad[i] = g;
    if (!adam || n == 0) return;
    const float t = float(net.words[kGwStep] > 0 ? net.words[kGwStep] : 1u);
    const float c1 = 1.0f - powf(hp.beta1, t), c2 = 1.0f - powf(hp.beta2, t);
    const float mi = hp.beta1 * net.adamM[i] + (1.0f - hp.beta1) * g;
    const float vi = hp.beta2 * net.adamV[i] + (1.0f - hp.beta2) * g * g;
    net.adamM[i] = mi;
    net.adamV[i] = vi;
    const float p = net.params[i] - hp.lr * (mi / c1) / (sqrtf(vi / c2) + hp.eps);
    net.params[i] = p;
    int l = 0;
    while (l < 4 && i >= guideWeightOffset(l + 1)) ++l;
    const int k = i - guideWeightOffset(l), nin = guideLayerIn(l);
    uint16_t* image = static_cast<uint16_t*>(net.image);
    if (k < nin * guideLayerOut(l)) {
        const int o = k / nin, in = k % nin;
        const uint16_t v = toBf16(p);
        image[wOff(l) + o * wStride(l) + in] = v;
        if (l > 0) static_cast<uint16_t*>(net.imageT)[wtOff(l) + in * kWRow + o] = v;
    } else {
        reinterpret_cast<float*>(image + kWTotal)[l * 64 + (k - nin * guideLayerOut(l))] = p;
    }
}

// ------------------------------------------------------- sampling queries --
__global__ void kGuideSample(const float* __restrict__ raw, const float* __restrict__ u, uint32_t count,
                             float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    GuideMix m;
    decodeMix(raw + size_t(i) * kGuideOut, m);
    float4 st[kGuideMixF4];
    storeMix(m, st);
    const V3 d = storedMixSample(st, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    out[i] = f4(d, storedMixPdf(st, d));
}

__global__ void kGuidePdf(const float* __restrict__ raw, const float4* __restrict__ dirs, uint32_t count,
                          float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    GuideMix m;
    decodeMix(raw + size_t(i) * kGuideOut, m);
    float4 st[kGuideMixF4];
    storeMix(m, st);
    out[i] = storedMixPdf(st, v3(dirs[i]));
}

// ------------------------------------------------------- training records --
__global__ void kGuideFrameBegin(uint32_t* __restrict__ words, const FrameConsts* __restrict__ g) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        words[kGwCand0 + (g->guideStamp & 1u)] = 0u;
        words[kGwPool] = 0u;
    }
}

__device__ __forceinline__ float hashUnit(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (c + 0x165667B1u) * 0xC2B2AE3Du;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return float(h >> 8) * (1.0f / 16777216.0f);
}

// One thread per training path: a lane gathers the levels of its path that
// hold a record of this frame (its bit mask, set by the shade kernel) with
// positive weight (the candidates) and those kept for the pool, then the
// workgroup reserves its pool slots with one atomic (prefix sums over lanes
// and waves give each record its slot).  The loops over levels are unrolled
// to the largest level count and predicated by the mask, so a lane's record
// reads are all in flight at once.  Every lane of the workgroup runs the
// same number of iterations (the barriers need all of them); the per-wave
// totals are double buffered by iteration parity, so one iteration's writes
// never meet the previous one's reads.
constexpr uint32_t kCollectLevels = 17;  // bounce levels 0..16
constexpr uint32_t kCollectBlock = 1024;  // one pair of pool atomics per 1,024 paths
__global__ __launch_bounds__(kCollectBlock) void kGuideCollect(GuideCollectArgs a) {
    const uint32_t stamp = a.g->guideStamp;
    const uint32_t kStride = a.g->guideTrainStride, phase = a.g->guideTrainPhase;
    uint32_t* cand = a.words + kGwCand0 + (stamp & 1u);
    const uint32_t prev = a.words[kGwCand0 + ((stamp + 1u) & 1u)];
    const float keep = prev > a.poolCap ? float(a.poolCap) / float(prev) : 1.0f;
    const uint32_t stride = gridDim.x * blockDim.x;
    constexpr uint32_t kWaves = kCollectBlock / 64;
    __shared__ uint32_t sKeep[2][kWaves], sTake[2][kWaves], sBase[2];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t levelMask = a.levels >= 32 ? ~0u : (1u << a.levels) - 1u;
    uint32_t buf = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.trainPaths; base += stride) {
        const uint32_t q = base + threadIdx.x;
        const uint32_t p = q * kStride + phase;
        const bool valid = q < a.trainPaths && p < a.paths;
        const uint32_t m = valid ? a.recMask[q] & levelMask : 0u;
        const float lf = m ? luminance(v3(a.radiance[p])) : 0.0f;
        // Levels of this path with a record of this frame and positive
        // weight (candidates), and those kept for the pool.
        float r0w[kCollectLevels], r1w[kCollectLevels];
        uint32_t st[kCollectLevels];
#pragma unroll
        for (uint32_t lv = 0; lv < kCollectLevels; ++lv) {
            r0w[lv] = r1w[lv] = 0.0f;
            st[lv] = 0u;
            if (m >> lv & 1u) {
                const float4* rec = a.pathRec + (size_t(lv) * a.trainPaths + q) * 4;
                r0w[lv] = rec[0].w;
                r1w[lv] = rec[1].w;
                st[lv] = __float_as_uint(rec[2].w);
            }
        }
        uint32_t takeMask = 0, keepMask = 0;
        float wv[kCollectLevels];
#pragma unroll
        for (uint32_t lv = 0; lv < kCollectLevels; ++lv) {
            wv[lv] = (lf - r1w[lv]) / fmaxf(r0w[lv], 1e-20f);
            if (!(m >> lv & 1u) || st[lv] != stamp || !(wv[lv] > 0.0f && isfinite(wv[lv]))) continue;
            takeMask |= 1u << lv;
            if (keep >= 1.0f || hashUnit(p, lv, stamp) < keep) keepMask |= 1u << lv;
        }
        // Inclusive scan of the lanes' kept counts within the wave.
        const uint32_t nKeep = __builtin_popcount(keepMask);
        uint32_t incl = nKeep;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off, 64);
            if (lane >= uint32_t(off)) incl += t;
        }
        uint32_t nTake = __builtin_popcount(takeMask);
        for (int off = 32; off >= 1; off >>= 1) nTake += __shfl_xor(nTake, off, 64);
        if (lane == 63) {
            sKeep[buf][wave] = incl;
            sTake[buf][wave] = nTake;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t take = 0, kept = 0;
            for (uint32_t w = 0; w < kWaves; ++w) {
                take += sTake[buf][w];
                kept += sKeep[buf][w];
            }
            if (take) atomicAdd(cand, take);
            sBase[buf] = kept ? atomicAdd(a.words + kGwPool, kept) : 0u;
        }
        __syncthreads();
        uint32_t slot0 = sBase[buf] + incl - nKeep;
        for (uint32_t w = 0; w < wave; ++w) slot0 += sKeep[buf][w];
        buf ^= 1u;
#pragma unroll
        for (uint32_t lv = 0; lv < kCollectLevels; ++lv) {
            if (!(keepMask >> lv & 1u)) continue;
            const uint32_t slot = slot0 + uint32_t(__builtin_popcount(keepMask & ((1u << lv) - 1u)));
            if (slot >= a.poolCap) continue;
            const float4* rec = a.pathRec + (size_t(lv) * a.trainPaths + q) * 4;
            const float4 r0 = rec[0], r2 = rec[2], r3 = rec[3];
            const float4 r1 = rec[1];
            float4* dst = a.pool + size_t(slot) * 4;
            dst[0] = make_float4(r0.x, r0.y, r0.z, wv[lv]);
            dst[1] = make_float4(r1.x, r1.y, r1.z, 0.0f);
            dst[2] = make_float4(r2.x, r2.y, r2.z, 0.0f);
            dst[3] = make_float4(r3.x, r3.y, r3.z, 0.0f);
        }
    }
}

// ---------------------------------------------------------------- launchers --
static uint32_t cappedGrid(uint32_t rows, uint32_t rowsPerBlock, uint32_t cap) {
    uint32_t g = (rows + rowsPerBlock - 1) / rowsPerBlock;
    if (g > cap) g = cap;
    return g ? g : 1;
}

hipError_t launchGuidePack(const float* params, void* image, void* imageT, hipStream_t s) {
    hipLaunchKernelGGL(kGuidePack, dim3((kWTotal + 320 + kWTTotal + 255) / 256), dim3(256), 0, s, params,
                       static_cast<uint16_t*>(image), static_cast<uint16_t*>(imageT));
    return hipGetLastError();
}

hipError_t launchGuideInfer(const RecView& rec, const uint32_t* countPtr, uint32_t maxCount, const void* image,
                            void* out, bool decoded, hipStream_t s) {
    hipLaunchKernelGGL(kGuideInfer, dim3(cappedGrid(maxCount, kFwdWaves * 32, 512)), dim3(kFwdWaves * 64), 0, s, rec,
                       countPtr, static_cast<const u32x4*>(image), out, decoded ? 1 : 0);
    return hipGetLastError();
}

// The per-row launch and the weight gradient; b.cap bounds the batch, the
// workspace is sized for net.cap rows.
static hipError_t launchRowsAndWeightGrad(const GuideBatch& b, const GuideNet& net, bool countStep, hipStream_t s) {
    if (b.cap > net.cap || b.cap == 0 || net.stride != guideRowStride(net.cap)) return hipErrorInvalidValue;
    const uint32_t rows = guideRowStride(b.cap);
    hipLaunchKernelGGL(kGuideTrainRows, dim3(cappedGrid(rows, kTrainWaves * 32, 256)), dim3(kTrainWaves * 64), 0, s, b,
                       net, countStep ? 1 : 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideWeightGrad, dim3(guideGradChunks(b.cap) * 5), dim3(kGradWaves * 64), 0, s, net);
    return hipGetLastError();
}

hipError_t launchGuideGradient(const GuideBatch& b, const GuideNet& net, bool countStep, hipStream_t s) {
    hipError_t e = launchRowsAndWeightGrad(b, net, countStep, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideReduceStep, dim3((kGuideParams + kStepParams - 1) / kStepParams), dim3(kStepBlock), 0, s,
                       net, GuideAdam{}, 0);
    return hipGetLastError();
}

hipError_t launchGuideStep(const GuideBatch& b, const GuideNet& net, const GuideAdam& h, hipStream_t s) {
    hipError_t e = launchRowsAndWeightGrad(b, net, true, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideReduceStep, dim3((kGuideParams + kStepParams - 1) / kStepParams), dim3(kStepBlock), 0, s,
                       net, h, 1);
    return hipGetLastError();
}

hipError_t launchGuideSample(const float* raw, const float* u, uint32_t count, float4* out, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(kGuideSample, dim3((count + 255) / 256), dim3(256), 0, s, raw, u, count, out);
    return hipGetLastError();
}

hipError_t launchGuidePdf(const float* raw, const float4* dirs, uint32_t count, float* out, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(kGuidePdf, dim3((count + 255) / 256), dim3(256), 0, s, raw, dirs, count, out);
    return hipGetLastError();
}

hipError_t launchGuideFrameBegin(uint32_t* words, const FrameConsts* g, hipStream_t s) {
    hipLaunchKernelGGL(kGuideFrameBegin, dim3(1), dim3(64), 0, s, words, g);
    return hipGetLastError();
}

hipError_t launchGuideCollect(const GuideCollectArgs& a, hipStream_t s) {
    if (a.trainPaths == 0) return hipSuccess;
    hipLaunchKernelGGL(kGuideCollect, dim3(cappedGrid(a.trainPaths, kCollectBlock, 1024)), dim3(kCollectBlock), 0, s, a);
    return hipGetLastError();
}

} // namespace dev
} // namespace mrt

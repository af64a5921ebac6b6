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
// them.  Training stores activations and deltas feature-major ([feature][row],
// rows padded to 32) so every operand of the backward and weight-gradient
// products is a contiguous 16-byte fragment in HBM.
//
// Kernels:
//   kGuidePack        FP32 params -> packed bf16 forward image
//   kGuideInfer       fused forward for rendering (features -> mixture)
//   kGuideStepBegin   batch size, counters and per-chunk weight sums of one step
//   kGuideFwdTrain    forward over the batch, keeping the activations
//   kGuideOutGrad     d(weighted NLL)/d(raw outputs) per row
//   kGuideBackward    delta_{l-1} = (W_l^T delta_l) * relu'(a_{l-1}), l = 4 .. 1
//   kGuideWeightGrad  per row chunk, all layers in one launch after the
//                     backward pass: grad W_l = delta_l a_{l-1}^T, grad b_l
//   kGuideGradReduce  sum of the chunk partials (fixed order, no atomics)
//   kGuideAdam        Adam update of the FP32 master weights
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

__global__ __launch_bounds__(256) void kGuidePack(const float* __restrict__ params, uint16_t* __restrict__ image) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kWTotal) {
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

// Feature-major copy of the wave's 32 rows: features [0, nfeat) of act to
// dst[f * stride + row0 + 0..31] (64 contiguous bytes per feature).
__device__ __forceinline__ void storeFeatureMajor(const uint16_t* act, int nfeat, uint16_t* __restrict__ dst,
                                                  uint32_t stride, uint32_t row0) {
    const int lane = threadIdx.x & 63;
    for (int f = lane; f < nfeat; f += 64) {
        uint32_t w[16];
        for (int m = 0; m < 16; ++m) w[m] = pack2(act[(2 * m) * kActStride + f], act[(2 * m + 1) * kActStride + f]);
        u32x4* d = reinterpret_cast<u32x4*>(dst + size_t(f) * stride + row0);
        for (int q = 0; q < 4; ++q) d[q] = u32x4{ w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3] };
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
// Start of a training step.  Block 0 publishes the batch size, clears the
// loss and advances the Adam step counter (countStep, non-empty batch); the
// later kernels of the step read them.  Every block sums the weights of its
// row chunk in a fixed order: block k sums chunk k (kGuideGradChunk rows, four
// per thread, then an LDS tree) into weightPart[k], and kGuideOutGrad adds
// the chunk sums in index order.  The sum scales every delta before its bf16
// rounding, so an order-dependent sum would make the gradient differ from run
// to run.
constexpr int kWeightSumBlock = 256;
__global__ __launch_bounds__(kWeightSumBlock) void kGuideStepBegin(GuideBatch b, uint32_t* __restrict__ words,
                                                                   float* __restrict__ weightPart, int countStep) {
    __shared__ float part[kWeightSumBlock];
    const uint32_t n = batchCount(b);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        words[kGwBatch] = n;
        *wordF(words, kGwLoss) = 0.0f;
        if (countStep && n > 0) words[kGwStep] += 1u;
    }
    const RecView rv = batchView(b);
    const uint32_t row0 = blockIdx.x * kGuideGradChunk + threadIdx.x;
    float s = 0.0f;
    for (uint32_t i = 0; i < kGuideGradChunk / kWeightSumBlock; ++i) {
        const uint32_t row = row0 + i * kWeightSumBlock;
        if (row < n) s += recRow(rv, row)[0].w;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int half = kWeightSumBlock / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half) part[threadIdx.x] += part[threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) weightPart[blockIdx.x] = part[0];
}

// Sum of the chunk weight sums in index order, the same in every block
// (threads add strided slices in order, then an LDS tree).
__device__ float batchWeightSum(const float* __restrict__ weightPart, uint32_t chunks) {
    __shared__ float part[256];
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < chunks; i += 256) s += weightPart[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1) {
        if (threadIdx.x < half) part[threadIdx.x] += part[threadIdx.x + half];
        __syncthreads();
    }
    const float total = part[0];
    __syncthreads();
    return total;
}

// Forward over the batch's padded rows [0, pad32(count)): keeps the encoded
// inputs (features 0..79 of actIn), the hidden outputs and the raw outputs,
// feature-major.  Padding rows encode to zero inputs (finite activations).
__global__ __launch_bounds__(kFwdWaves * 64) void kGuideFwdTrain(GuideBatch b, const uint32_t* __restrict__ words,
                                                                 const u32x4* __restrict__ image, GuideNet net) {
    __shared__ __attribute__((aligned(16))) uint16_t W[kWTotal];
    __shared__ __attribute__((aligned(16))) float bias[5 * 64];
    __shared__ __attribute__((aligned(16))) uint16_t actAll[kFwdWaves][32 * kActStride];
    const uint32_t count = words[kGwBatch];
    const uint32_t rows = pad32(count);
    if (blockIdx.x * kFwdWaves * 32 >= rows) return;  // uniform per block
    loadImage(image, W, bias);
    __syncthreads();
    const RecView rv = batchView(b);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* act = actAll[wave];
    const float* fo = reinterpret_cast<const float*>(act);
    const uint32_t stride = net.stride;
    const uint32_t totalWaves = gridDim.x * kFwdWaves;
    uint32_t blk = blockIdx.x * kFwdWaves + wave;
    RowIn next = loadRow(rv, blk * 32 + (lane & 31), count);
    for (; blk * 32 < rows; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        const RowIn cur = next;
        next = loadRow(rv, (blk + totalWaves) * 32 + (lane & 31), count);
        encodeRows(cur, row0 + (lane & 31) < count, act);
        storeFeatureMajor(act, kGuideInPad, net.actIn, stride, row0);
        f32x16 c0, c1;
        for (int l = 0; l < 5; ++l) {
            layerForward(l, act, W, bias, c0, c1);
            if (l < 4) storeFeatureMajor(act, 64, net.act + size_t(l) * 64 * stride, stride, row0);
        }
        stashOutputs(c0, c1, act);
        if (lane < kGuideOut) {
            float* d = net.raw + size_t(lane) * stride + row0;
            for (int m = 0; m < 32; m += 4)
                *reinterpret_cast<float4*>(d + m) = make_float4(fo[m * kRawStride + lane], fo[(m + 1) * kRawStride + lane],
                                                                fo[(m + 2) * kRawStride + lane],
                                                                fo[(m + 3) * kRawStride + lane]);
        }
    }
}

// delta (feature-major bf16, 64 features, 40..63 zero) = d L / d raw with
// L = sum_j w_j (-log q(wi_j)) / sum_j w_j; padding rows get zero deltas.
__global__ __launch_bounds__(256) void kGuideOutGrad(GuideBatch b, uint32_t* __restrict__ words, GuideNet net,
                                                     uint32_t chunks) {
    const uint32_t count = words[kGwBatch];
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = net.stride;
    const float wsum = batchWeightSum(net.weightPart, chunks);
    if (blockIdx.x == 0 && threadIdx.x == 0) *wordF(words, kGwWeightSum) = wsum;
    float loss = 0.0f;
    if (row < pad32(count)) {
        uint16_t* dcol = net.delta[4] + row;
        if (row < count) {
            const float scale = 1.0f / fmaxf(wsum, 1e-30f);
            const float4* q = recRow(batchView(b), row);
            const float wgt = q[0].w * scale;
            const V3 wi = v3(q[3]);
            float o[kGuideOut];
            for (int k = 0; k < kGuideOut; ++k) o[k] = net.raw[size_t(k) * stride + row];
            GuideMix m;
            decodeMix(o, m);
            float lf[kGuideLobes], mx = -1e30f;
            for (int k = 0; k < kGuideLobes; ++k) {
                const float c = dot(m.mu[k], wi);
                lf[k] = logf(m.w[k]) + logf(m.kappa[k]) - logf(kTwoPi) - logf(-expm1f(-2.0f * m.kappa[k])) +
                        m.kappa[k] * (c - 1.0f);
                mx = fmaxf(mx, lf[k]);
            }
            float s = 0.0f;
            for (int k = 0; k < kGuideLobes; ++k) s += expf(lf[k] - mx);
            const float logq = mx + logf(s);
            loss = -wgt * logq;
            for (int k = 0; k < kGuideLobes; ++k) {
                const float resp = expf(lf[k] - logq);
                const float kap = m.kappa[k];
                const float c = dot(m.mu[k], wi);
                const float lk = o[5 * k + 3];
                const bool inside = lk > kLogKappaLo && lk < kLogKappaHi;
                const float dlogf = 1.0f / kap - 2.0f / expm1f(2.0f * kap) + (c - 1.0f);
                const V3 v = v3(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
                const float vn = fmaxf(sqrtf(dot(v, v)), 1e-12f);
                const float gmu = -wgt * resp * kap / vn;
                dcol[size_t(5 * k + 0) * stride] = toBf16(gmu * (wi.x - c * m.mu[k].x));
                dcol[size_t(5 * k + 1) * stride] = toBf16(gmu * (wi.y - c * m.mu[k].y));
                dcol[size_t(5 * k + 2) * stride] = toBf16(gmu * (wi.z - c * m.mu[k].z));
                dcol[size_t(5 * k + 3) * stride] = toBf16(inside ? -wgt * resp * dlogf * kap : 0.0f);
                dcol[size_t(5 * k + 4) * stride] = toBf16(-wgt * (resp - m.w[k]));
            }
        } else {
            for (int k = 0; k < kGuideOut; ++k) dcol[size_t(k) * stride] = 0;
        }
        for (int k = kGuideOut; k < 64; ++k) dcol[size_t(k) * stride] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if ((threadIdx.x & 63) == 0 && loss != 0.0f) atomicAdd(wordF(words, kGwLoss), loss);
}

// Backward pass through the hidden layers in one launch: for l = 4 .. 1,
// delta_{l-1}[in][row] = (sum_o W_l[o][in] delta_l[o][row]) * (a_{l-1}[in][row] > 0),
// one wave per 32-row tile.  The four W_l^T are staged in LDS once per block
// (WT[l-1][in][o]); the wave's tile of delta_4 is read from HBM and staged
// transposed (sD[row][o]), and each following layer's tile is written
// straight into sD from the accumulators, so both MFMA operands are 16-byte
// LDS loads (144-byte row stride) and only delta_4 is read back.  Every
// delta_{l-1} is also written out for the weight gradient.
__global__ __launch_bounds__(256) void kGuideBackward(GuideNet net) {
    // W^T rows at the forward image's hidden row stride (144 bytes): the 32
    // rows of one operand load fall on 16 LDS bank offsets.
    __shared__ __attribute__((aligned(16))) uint16_t WT[4][64 * kWRow];
    __shared__ __attribute__((aligned(16))) uint16_t sDAll[4][32 * kWRow];
    const uint32_t rows = pad32(net.words[kGwBatch]);
    if (blockIdx.x * 128 >= rows) return;  // uniform per block
    const uint32_t stride = net.stride;
    for (int l = 1; l <= 4; ++l) {
        const int nin = guideLayerIn(l), nout = guideLayerOut(l);
        for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
            const int in = i / 64, o = i % 64;
            WT[l - 1][in * kWRow + o] =
                (o < nout && in < nin) ? toBf16(net.params[guideWeightOffset(l) + o * nin + in]) : uint16_t(0);
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    uint16_t* sD = sDAll[wave];
    const uint32_t totalWaves = gridDim.x * 4;
    for (uint32_t blk = blockIdx.x * 4 + wave; blk * 32 < rows; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        // stage delta_4[0..63][row0 .. row0 + 31] transposed: 256 16-byte
        // pieces (8 rows of one feature), 4 per lane
        for (int q = lane; q < 256; q += 64) {
            const int o = q >> 2, part = q & 3;
            const u32x4 v = *reinterpret_cast<const u32x4*>(net.delta[4] + size_t(o) * stride + row0 + 8 * part);
            const uint32_t w[4] = { v.x, v.y, v.z, v.w };
            for (int t = 0; t < 8; ++t) sD[(8 * part + t) * kWRow + o] = uint16_t(w[t >> 1] >> (16 * (t & 1)));
        }
        waveHandoff();
        for (int l = 4; l >= 1; --l) {
            bf16x8 b[4];
            for (int k = 0; k < 4; ++k) b[k] = ld8(sD + r * kWRow + 16 * k + 8 * h);
            waveHandoff();  // every lane has its operand before sD is overwritten
            const uint16_t* actPrev = net.act + size_t(l - 1) * 64 * stride;
            uint16_t* deltaPrev = net.delta[l - 1];
            const uint16_t* W = WT[l - 1];
            // The delta rows are the A operand and W^T the B operand, so lane
            // (r, h) ends with rows 8 j + 4 h + i of input feature 32 mt + r:
            // four consecutive rows of deltaPrev and actPrev per 8-byte access.
            for (int mt = 0; mt < 2; ++mt) {
                f32x16 acc = zero16();
                for (int k = 0; k < 4; ++k) acc = mfma(b[k], ld8(W + (32 * mt + r) * kWRow + 16 * k + 8 * h), acc);
                for (int j = 0; j < 4; ++j) {
                    const size_t at = size_t(32 * mt + r) * stride + row0 + 8 * j + 4 * h;
                    const uint2 ap = *reinterpret_cast<const uint2*>(actPrev + at);
                    const uint32_t a4[4] = { ap.x & 0xFFFFu, ap.x >> 16, ap.y & 0xFFFFu, ap.y >> 16 };
                    uint16_t d4[4];
                    for (int i = 0; i < 4; ++i)
                        d4[i] = fromBf16(uint16_t(a4[i])) > 0.0f ? toBf16(acc[4 * j + i]) : uint16_t(0);
                    *reinterpret_cast<uint2*>(deltaPrev + at) = make_uint2(pack2(d4[0], d4[1]), pack2(d4[2], d4[3]));
                    if (l > 1)
                        for (int i = 0; i < 4; ++i) sD[(8 * j + 4 * h + i) * kWRow + 32 * mt + r] = d4[i];
                }
            }
            waveHandoff();  // the next layer (or tile) reads sD
        }
    }
}

// One launch forms the weight gradients of all five layers.  Block job =
// (chunk, layer, tile); tile (mt, nt) covers outputs 32 mt.. and inputs
// 32 nt.. (layer 0: 3 input tiles over the 96 zero-padded input features).
// The block's 4 waves take consecutive quarters of the chunk's rows, and
// wave 0 adds their accumulators in wave order (fixed summation order, no
// atomics).  Tiles with nt == 0 also form the bias sums with an all-ones B
// operand.
constexpr uint32_t kGradWaves = 4;
constexpr uint32_t kGradSlice = kGuideGradChunk / kGradWaves;  // rows per wave
constexpr uint32_t kGradJobsPerChunk = 2 * 3 + 4 * (2 * 2);    // layer 0: 6 tiles, layers 1..4: 4 each
__global__ __launch_bounds__(kGradWaves * 64) void kGuideWeightGrad(GuideNet net) {
    // per wave and lane, the 16 accumulator values contiguous (16-byte LDS accesses)
    __shared__ __attribute__((aligned(16))) float sAcc[kGradWaves][64][16];
    __shared__ __attribute__((aligned(16))) float sBias[kGradWaves][64][16];
    const uint32_t rows = pad32(net.words[kGwBatch]);
    const uint32_t chunk = blockIdx.x / kGradJobsPerChunk, job = blockIdx.x % kGradJobsPerChunk;
    const uint32_t r0 = chunk * kGuideGradChunk;
    if (r0 >= rows) return;  // whole block
    const int l = job < 6 ? 0 : 1 + int(job - 6) / 4;
    const uint32_t tile = job < 6 ? job : (job - 6) % 4;
    const uint32_t stride = net.stride;
    const uint16_t* delta = net.delta[l];
    const uint16_t* actPrev = l == 0 ? net.actIn : net.act + size_t(l - 1) * 64 * stride;
    const int nin = guideLayerIn(l), nout = guideLayerOut(l);
    const int ntIn = l == 0 ? 3 : 2;
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t s0 = r0 + w * kGradSlice, s1 = min(s0 + kGradSlice, rows);
    const int mt = int(tile) / ntIn, nt = int(tile) % ntIn;
    const int lane = threadIdx.x & 63, lr = lane & 31, h = lane >> 5;
    const uint16_t* aRow = delta + size_t(32 * mt + lr) * stride + 8 * h;
    const uint16_t* bRow = actPrev + size_t(32 * nt + lr) * stride + 8 * h;
    const uint32_t one2 = 0x3F803F80u;  // two bf16 1.0
    const bf16x8 ones = __builtin_bit_cast(bf16x8, u32x4{ one2, one2, one2, one2 });
    f32x16 acc = zero16(), accB = zero16();
    // Eight row steps' operands are loaded before their products, so a wave
    // waits once per eight steps instead of once per step; the products are
    // accumulated in row order either way.
    constexpr int kAhead = 8;
    uint32_t r = s0;
    for (; r + 16 * kAhead <= s1; r += 16 * kAhead) {
        bf16x8 a[kAhead], b[kAhead];
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            a[u] = ld8(aRow + r + 16 * u);
            b[u] = ld8(bRow + r + 16 * u);
        }
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            acc = mfma(a[u], b[u], acc);
            if (nt == 0) accB = mfma(a[u], ones, accB);
        }
    }
    for (; r < s1; r += 16) {
        const bf16x8 a = ld8(aRow + r);
        acc = mfma(a, ld8(bRow + r), acc);
        if (nt == 0) accB = mfma(a, ones, accB);
    }
    for (int q = 0; q < 4; ++q) {
        reinterpret_cast<float4*>(sAcc[w][lane])[q] = make_float4(acc[4 * q], acc[4 * q + 1], acc[4 * q + 2],
                                                                  acc[4 * q + 3]);
        if (nt == 0)
            reinterpret_cast<float4*>(sBias[w][lane])[q] = make_float4(accB[4 * q], accB[4 * q + 1],
                                                                       accB[4 * q + 2], accB[4 * q + 3]);
    }
    __syncthreads();
    if (w != 0) return;
    for (uint32_t v = 1; v < kGradWaves; ++v)
        for (int q = 0; q < 4; ++q) {
            const float4 x = reinterpret_cast<const float4*>(sAcc[v][lane])[q];
            acc[4 * q] += x.x;
            acc[4 * q + 1] += x.y;
            acc[4 * q + 2] += x.z;
            acc[4 * q + 3] += x.w;
            if (nt == 0) {
                const float4 y = reinterpret_cast<const float4*>(sBias[v][lane])[q];
                accB[4 * q] += y.x;
                accB[4 * q + 1] += y.y;
                accB[4 * q + 2] += y.z;
                accB[4 * q + 3] += y.w;
            }
        }
    float* dst = net.partial + size_t(chunk) * kGuideParams;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            const int o = 32 * mt + 8 * j + 4 * h + i, in = 32 * nt + lr;
            if (o < nout && in < nin) dst[guideWeightOffset(l) + o * nin + in] = acc[4 * j + i];
            if (nt == 0 && lr == 0 && o < nout) dst[guideBiasOffset(l) + o] = accB[4 * j + i];
        }
}

// 64 parameters per block; thread (q, j) sums chunks q, q + 4, ... of
// parameter j in order, then the four slices are added in q order.
constexpr int kReduceParams = 64;
__global__ __launch_bounds__(256) void kGuideGradReduce(const uint32_t* __restrict__ words,
                                                        const float* __restrict__ partial, float* __restrict__ grad) {
    __shared__ float slice[4][kReduceParams];
    const int j = threadIdx.x & (kReduceParams - 1), q = threadIdx.x / kReduceParams;
    const int i = blockIdx.x * kReduceParams + j;
    const uint32_t chunks = (pad32(words[kGwBatch]) + kGuideGradChunk - 1) / kGuideGradChunk;
    float s = 0.0f;
    if (i < kGuideParams)
        for (uint32_t c = q; c < chunks; c += 4) s += partial[size_t(c) * kGuideParams + i];
    slice[q][j] = s;
    __syncthreads();
    if (q == 0 && i < kGuideParams) grad[i] = ((slice[0][j] + slice[1][j]) + slice[2][j]) + slice[3][j];
}

__global__ __launch_bounds__(256) void kGuideAdam(float* __restrict__ params, float* __restrict__ m,
                                                  float* __restrict__ v, const float* __restrict__ grad,
                                                  uint32_t* __restrict__ words, GuideAdam h) {
    const uint32_t n = words[kGwBatch];
    if (n == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        words[kGwLastLoss] = words[kGwLoss];
        words[kGwLastBatch] = n;
    }
    if (i >= kGuideParams) return;
    const float t = float(words[kGwStep] > 0 ? words[kGwStep] : 1u);
    const float c1 = 1.0f - powf(h.beta1, t), c2 = 1.0f - powf(h.beta2, t);
    const float g = grad[i];
    const float mi = h.beta1 * m[i] + (1.0f - h.beta1) * g;
    const float vi = h.beta2 * v[i] + (1.0f - h.beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= h.lr * (mi / c1) / (sqrtf(vi / c2) + h.eps);
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
// hold a record of this frame with positive weight (the candidates) and those
// kept for the pool, then the workgroup reserves its pool slots with one
// atomic (prefix sums over lanes and waves give each record its slot).  Every
// lane of the workgroup runs the same number of iterations (the barriers need
// all of them); the per-wave totals are double buffered by iteration parity,
// so one iteration's writes never meet the previous one's reads.
__global__ __launch_bounds__(256) void kGuideCollect(GuideCollectArgs a) {
    const uint32_t stamp = a.g->guideStamp;
    const uint32_t kStride = a.g->guideTrainStride, phase = a.g->guideTrainPhase;
    uint32_t* cand = a.words + kGwCand0 + (stamp & 1u);
    const uint32_t prev = a.words[kGwCand0 + ((stamp + 1u) & 1u)];
    const float keep = prev > a.poolCap ? float(a.poolCap) / float(prev) : 1.0f;
    const uint32_t stride = gridDim.x * blockDim.x;
    constexpr uint32_t kWaves = 256 / 64;
    __shared__ uint32_t sKeep[2][kWaves], sTake[2][kWaves], sBase[2];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t buf = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.trainPaths; base += stride) {
        const uint32_t q = base + threadIdx.x;
        const uint32_t p = q * kStride + phase;
        const bool valid = q < a.trainPaths && p < a.paths;
        const float lf = valid ? luminance(v3(a.radiance[p])) : 0.0f;
        // Levels of this path with a record of this frame and positive
        // weight (candidates), and those kept for the pool.
        uint32_t takeMask = 0, keepMask = 0;
        if (valid) {
            for (uint32_t lv = 0; lv < a.levels; ++lv) {
                const float4* rec = a.pathRec + (size_t(lv) * a.trainPaths + q) * 4;
                if (__float_as_uint(rec[2].w) != stamp) continue;
                const float w = (lf - rec[1].w) / fmaxf(rec[0].w, 1e-20f);
                if (!(w > 0.0f && isfinite(w))) continue;
                takeMask |= 1u << lv;
                if (keep >= 1.0f || hashUnit(p, lv, stamp) < keep) keepMask |= 1u << lv;
            }
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
        uint32_t slot = sBase[buf] + incl - nKeep;
        for (uint32_t w = 0; w < wave; ++w) slot += sKeep[buf][w];
        buf ^= 1u;
        while (keepMask) {
            const uint32_t lv = uint32_t(__builtin_ctz(keepMask));
            keepMask &= keepMask - 1u;
            if (slot < a.poolCap) {
                const float4* rec = a.pathRec + (size_t(lv) * a.trainPaths + q) * 4;
                const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
                const float w = (lf - r1.w) / fmaxf(r0.w, 1e-20f);
                float4* dst = a.pool + size_t(slot) * 4;
                dst[0] = make_float4(r0.x, r0.y, r0.z, w);
                dst[1] = make_float4(r1.x, r1.y, r1.z, 0.0f);
                dst[2] = make_float4(r2.x, r2.y, r2.z, 0.0f);
                dst[3] = make_float4(r3.x, r3.y, r3.z, 0.0f);
            }
            ++slot;
        }
    }
}

// ---------------------------------------------------------------- launchers --
static uint32_t cappedGrid(uint32_t rows, uint32_t rowsPerBlock, uint32_t cap) {
    uint32_t g = (rows + rowsPerBlock - 1) / rowsPerBlock;
    if (g > cap) g = cap;
    return g ? g : 1;
}

hipError_t launchGuidePack(const float* params, void* image, hipStream_t s) {
    hipLaunchKernelGGL(kGuidePack, dim3((kWTotal + 320 + 255) / 256), dim3(256), 0, s, params,
                       static_cast<uint16_t*>(image));
    return hipGetLastError();
}

hipError_t launchGuideInfer(const RecView& rec, const uint32_t* countPtr, uint32_t maxCount, const void* image,
                            void* out, bool decoded, hipStream_t s) {
    hipLaunchKernelGGL(kGuideInfer, dim3(cappedGrid(maxCount, kFwdWaves * 32, 512)), dim3(kFwdWaves * 64), 0, s, rec,
                       countPtr, static_cast<const u32x4*>(image), out, decoded ? 1 : 0);
    return hipGetLastError();
}

// b.cap bounds the batch; the workspace is sized for net.cap rows.
hipError_t launchGuideGradient(const GuideBatch& b, const GuideNet& net, bool countStep, hipStream_t s) {
    if (b.cap > net.cap || b.cap == 0 || net.stride != guideRowStride(net.cap)) return hipErrorInvalidValue;
    const uint32_t rows = guideRowStride(b.cap);
    hipError_t e;
    const uint32_t chunks = guideGradChunks(b.cap);
    hipLaunchKernelGGL(kGuideStepBegin, dim3(chunks), dim3(kWeightSumBlock), 0, s, b, net.words, net.weightPart,
                       countStep ? 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideFwdTrain, dim3(cappedGrid(rows, kFwdWaves * 32, 1024)), dim3(kFwdWaves * 64), 0, s, b,
                       net.words, static_cast<const u32x4*>(net.image), net);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideOutGrad, dim3(cappedGrid(rows, 256, 1u << 20)), dim3(256), 0, s, b, net.words, net,
                       chunks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideBackward, dim3(cappedGrid(rows, 128, 1024)), dim3(256), 0, s, net);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideWeightGrad, dim3(chunks * kGradJobsPerChunk), dim3(kGradWaves * 64), 0, s, net);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideGradReduce, dim3((kGuideParams + kReduceParams - 1) / kReduceParams), dim3(256), 0, s,
                       net.words, net.partial, net.grad);
    return hipGetLastError();
}

hipError_t launchGuideAdam(const GuideNet& net, const GuideAdam& h, hipStream_t s) {
    hipLaunchKernelGGL(kGuideAdam, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, net.params, net.adamM,
                       net.adamV, net.grad, net.words, h);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launchGuidePack(net.params, net.image, s);
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
    hipLaunchKernelGGL(kGuideCollect, dim3(cappedGrid(a.trainPaths, 256, 2048)), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace dev
} // namespace mrt

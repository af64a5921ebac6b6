// Path-guiding network on gfx950 matrix cores: BF16 operands, FP32 accumulate.
//
// Network: 78 (padded 80) -> 64 -> 64 -> 64 -> 64 -> 40, ReLU on the hidden
// layers (guide_device.h holds the encoding and the decoding of the outputs).
//
// Every matrix product uses v_mfma_f32_32x32x8bf16_1k.  Operand fragments of
// one 32x32 output tile C = A (32 x 8) * B (8 x 32) for lane l (0..63):
//   A: row m = l % 32, k = 4 (l / 32) + {0..3}        (4 bf16)
//   B: col n = l % 32, k = 4 (l / 32) + {0..3}        (4 bf16)
//   C: col n = l % 32, rows m = 8 j + 4 (l / 32) + i in acc[4 j + i]
// A wave owns 32 rows (records); its activations live in LDS between layers.
//
// Kernels:
//   kGuideInfer       fused forward for rendering (features -> mixture)
//   kGuideStepBegin   batch size and counters of one training step
//   kGuideWeightSum   sum of the batch's training weights
//   kGuideFwdTrain    forward over the batch, keeping the activations
//   kGuideOutGrad     d(weighted NLL)/d(raw outputs) per row
//   kGuideBackLayer   delta_{l-1} = (delta_l W_l) * relu'(a_{l-1})
//   kGuideWeightGrad  grad W_l += delta_l^T a_{l-1}, grad b_l += sum delta_l
//   kGuideAdam        Adam update of the FP32 master weights
//   kGuideCollect     path records -> weighted training records
// All counts are read from device memory, so a captured graph replays the
// same launches whatever the batch sizes turn out to be.
#include "guide_device.h"
#include "guide_launch.h"

namespace mrt {
namespace dev {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Round to nearest even (inputs are finite).
__device__ __forceinline__ uint16_t toBf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float fromBf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// Four consecutive bf16 at an 8-byte aligned address.
__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) {
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = short(v.x & 0xFFFFu);
    r[1] = short(v.x >> 16);
    r[2] = short(v.y & 0xFFFFu);
    r[3] = short(v.y >> 16);
    return r;
}
__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
    for (int i = 0; i < 16; ++i) z[i] = 0.0f;
    return z;
}

__device__ __forceinline__ float* wordF(uint32_t* w, uint32_t i) { return reinterpret_cast<float*>(w + i); }

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

// LDS image of the forward weights: W_l output-major, layer 0 rows padded to
// 80 inputs, the output layer padded to 64 rows.
constexpr int kWOff1 = 64 * kGuideInPad;
constexpr int kWTotal = kWOff1 + 4 * 64 * 64;  // 21,504 bf16
__host__ __device__ constexpr int wOff(int l) { return l == 0 ? 0 : kWOff1 + (l - 1) * 64 * 64; }
__host__ __device__ constexpr int wStride(int l) { return l == 0 ? kGuideInPad : 64; }

__device__ void loadForwardWeights(const float* __restrict__ params, uint16_t* W, float* bias) {
    for (int i = threadIdx.x; i < kWTotal; i += blockDim.x) {
        const int l = i < kWOff1 ? 0 : 1 + (i - kWOff1) / (64 * 64);
        const int r = (i - wOff(l)) / wStride(l), c = (i - wOff(l)) % wStride(l);
        float v = 0.0f;
        if (r < guideLayerOut(l) && c < guideLayerIn(l)) v = params[guideWeightOffset(l) + r * guideLayerIn(l) + c];
        W[i] = toBf16(v);
    }
    for (int i = threadIdx.x; i < 5 * 64; i += blockDim.x) {
        const int l = i / 64, o = i % 64;
        bias[i] = o < guideLayerOut(l) ? params[guideBiasOffset(l) + o] : 0.0f;
    }
}

// One layer over the wave's 32 rows in act (row stride kGuideInPad).
// Hidden layers write ReLU(W a + b) back into act as bf16; the output layer
// leaves W a + b in the two accumulator tiles (columns 0..31, 32..63).
__device__ __forceinline__ void layerForward(int l, uint16_t* act, const uint16_t* W, const float* bias,
                                             f32x16& c0, f32x16& c1) {
    const int lane = threadIdx.x & 63, r = lane & 31, kq = 4 * (lane >> 5);
    const int K = l == 0 ? kGuideInPad : 64;
    const uint16_t* Wl = W + wOff(l);
    const int ws = wStride(l);
    c0 = zero16();
    c1 = zero16();
    for (int kb = 0; kb < K; kb += 8) {
        const bf16x4 a = ld4(act + r * kGuideInPad + kb + kq);
        const bf16x4 b0 = ld4(Wl + r * ws + kb + kq);
        const bf16x4 b1 = ld4(Wl + (32 + r) * ws + kb + kq);
        c0 = mfma(a, b0, c0);
        c1 = mfma(a, b1, c1);
    }
    const float bb0 = bias[l * 64 + r], bb1 = bias[l * 64 + 32 + r];
    for (int i = 0; i < 16; ++i) {
        c0[i] += bb0;
        c1[i] += bb1;
    }
    if (l < 4) {
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                const int m = 8 * j + kq + i;
                act[m * kGuideInPad + r] = toBf16(fmaxf(c0[4 * j + i], 0.0f));
                act[m * kGuideInPad + 32 + r] = toBf16(fmaxf(c1[4 * j + i], 0.0f));
            }
    }
}

// Encodes rows [row0, row0 + 32) into act: lanes 0..31 write features 0..39
// of row lane, lanes 32..63 features 40..79.  Rows past count are zero.
__device__ __forceinline__ void encodeRows(const RecView& rv, uint32_t row0, uint32_t count, uint16_t* act) {
    const int lane = threadIdx.x & 63, r = lane & 31, half = lane >> 5;
    const uint32_t row = row0 + r;
    float p[3] = { 0, 0, 0 }, n[3] = { 0, 0, 0 }, wo[3] = { 0, 0, 0 };
    const bool valid = row < count;
    if (valid) {
        const float4* q = recRow(rv, row);
        const float4 a = q[0], b = q[1], c = q[2];
        p[0] = a.x; p[1] = a.y; p[2] = a.z;
        n[0] = b.x; n[1] = b.y; n[2] = b.z;
        wo[0] = c.x; wo[1] = c.y; wo[2] = c.z;
    }
    for (int k = half * 40; k < half * 40 + 40; ++k)
        act[r * kGuideInPad + k] = valid ? toBf16(guideFeature(p, n, wo, k)) : uint16_t(0);
}

// Writes the output tiles of rows [row0, row0 + 32) as FP32 into act (reused
// as float[32][40]).
__device__ __forceinline__ void stashOutputs(const f32x16& c0, const f32x16& c1, uint16_t* act) {
    const int lane = threadIdx.x & 63, r = lane & 31, kq = 4 * (lane >> 5);
    float* fo = reinterpret_cast<float*>(act);
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            const int m = 8 * j + kq + i;
            fo[m * kGuideOut + r] = c0[4 * j + i];
            if (r < kGuideOut - 32) fo[m * kGuideOut + 32 + r] = c1[4 * j + i];
        }
}

// ------------------------------------------------------------------ inference --
constexpr int kInferBlock = 256;  // 4 waves
__global__ __launch_bounds__(kInferBlock) void kGuideInfer(RecView rv, const uint32_t* __restrict__ countPtr,
                                                           const float* __restrict__ params, void* __restrict__ out,
                                                           int decoded) {
    __shared__ __attribute__((aligned(16))) uint16_t W[kWTotal];
    __shared__ float bias[5 * 64];
    __shared__ __attribute__((aligned(16))) uint16_t actAll[4][32 * kGuideInPad];
    const uint32_t count = *countPtr;
    if (blockIdx.x * 128 >= count) return;  // uniform per block
    loadForwardWeights(params, W, bias);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* act = actAll[wave];
    const float* fo = reinterpret_cast<const float*>(act);
    const uint32_t totalWaves = gridDim.x * 4;
    for (uint32_t blk = blockIdx.x * 4 + wave; blk * 32 < count; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        encodeRows(rv, row0, count, act);
        f32x16 c0, c1;
        for (int l = 0; l < 5; ++l) layerForward(l, act, W, bias, c0, c1);
        stashOutputs(c0, c1, act);
        if (decoded) {
            const uint32_t row = row0 + lane;
            if (lane < 32 && row < count) {
                GuideMix m;
                decodeMix(fo + lane * kGuideOut, m);
                storeMix(m, static_cast<float4*>(out) + size_t(row) * kGuideMixF4);
            }
        } else {
            float* o = static_cast<float*>(out);
            for (int i = lane; i < 32 * kGuideOut; i += 64) {
                const uint32_t row = row0 + i / kGuideOut;
                if (row < count) o[size_t(row) * kGuideOut + i % kGuideOut] = fo[i];
            }
        }
    }
}

// ------------------------------------------------------------------- training --
__global__ void kGuideStepBegin(GuideBatch b, uint32_t* __restrict__ words, int countStep) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t n = batchCount(b);
    words[kGwBatch] = n;
    *wordF(words, kGwWeightSum) = 0.0f;
    *wordF(words, kGwLoss) = 0.0f;
    if (countStep && n > 0) words[kGwStep] += 1u;
}

// Sum of the batch's weights in a fixed order (one block, LDS tree): the
// sum scales every delta before its bf16 rounding, so an order-dependent sum
// would make the gradient differ from run to run.
constexpr int kWeightSumBlock = 1024;
__global__ __launch_bounds__(kWeightSumBlock) void kGuideWeightSum(GuideBatch b, uint32_t* __restrict__ words) {
    __shared__ float part[kWeightSumBlock];
    const uint32_t n = words[kGwBatch];
    const RecView rv = batchView(b);
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += kWeightSumBlock) s += recRow(rv, i)[0].w;
    part[threadIdx.x] = s;
    __syncthreads();
    for (int half = kWeightSumBlock / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half) part[threadIdx.x] += part[threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) *wordF(words, kGwWeightSum) = part[0];
}

// Forward over the batch; keeps the bf16 inputs (cap x 80), the hidden
// activations (4 x cap x 64) and the raw FP32 outputs (cap x 40), where
// cap = b.cap is the workspace's row capacity.
__global__ __launch_bounds__(256) void kGuideFwdTrain(GuideBatch b, const uint32_t* __restrict__ words,
                                                      const float* __restrict__ params, uint16_t* __restrict__ acts,
                                                      float* __restrict__ raw) {
    __shared__ __attribute__((aligned(16))) uint16_t W[kWTotal];
    __shared__ float bias[5 * 64];
    __shared__ __attribute__((aligned(16))) uint16_t actAll[4][32 * kGuideInPad];
    const uint32_t count = words[kGwBatch];
    if (blockIdx.x * 128 >= count) return;  // uniform per block
    loadForwardWeights(params, W, bias);
    __syncthreads();
    const RecView rv = batchView(b);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* act = actAll[wave];
    const float* fo = reinterpret_cast<const float*>(act);
    const uint32_t totalWaves = gridDim.x * 4;
    for (uint32_t blk = blockIdx.x * 4 + wave; blk * 32 < count; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        encodeRows(rv, row0, count, act);
        for (int k = lane; k < 32 * kGuideInPad; k += 64) {
            const uint32_t row = row0 + k / kGuideInPad;
            if (row < count) acts[size_t(row) * kGuideInPad + k % kGuideInPad] = act[k];
        }
        f32x16 c0, c1;
        for (int l = 0; l < 5; ++l) {
            layerForward(l, act, W, bias, c0, c1);
            if (l < 4) {
                uint16_t* dst = acts + size_t(b.cap) * kGuideInPad + size_t(l) * b.cap * 64;
                for (int k = lane; k < 32 * 64; k += 64) {
                    const uint32_t row = row0 + k / 64;
                    if (row < count) dst[size_t(row) * 64 + k % 64] = act[(k / 64) * kGuideInPad + k % 64];
                }
            }
        }
        stashOutputs(c0, c1, act);
        for (int i = lane; i < 32 * kGuideOut; i += 64) {
            const uint32_t row = row0 + i / kGuideOut;
            if (row < count) raw[size_t(row) * kGuideOut + i % kGuideOut] = fo[i];
        }
    }
}

// delta (bf16, 64 columns, 40..63 zero) = d L / d raw with
// L = sum_j w_j (-log q(wi_j)) / sum_j w_j; the loss is accumulated too.
__global__ __launch_bounds__(256) void kGuideOutGrad(GuideBatch b, uint32_t* __restrict__ words,
                                                     const float* __restrict__ raw, uint16_t* __restrict__ delta) {
    const uint32_t count = words[kGwBatch];
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    float loss = 0.0f;
    if (row < count) {
        const float scale = 1.0f / fmaxf(*wordF(words, kGwWeightSum), 1e-30f);
        const float4* q = recRow(batchView(b), row);
        const float wgt = q[0].w * scale;
        const V3 wi = v3(q[3]);
        const float* o = raw + size_t(row) * kGuideOut;
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
        uint16_t* d = delta + size_t(row) * 64;
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
            d[5 * k + 0] = toBf16(gmu * (wi.x - c * m.mu[k].x));
            d[5 * k + 1] = toBf16(gmu * (wi.y - c * m.mu[k].y));
            d[5 * k + 2] = toBf16(gmu * (wi.z - c * m.mu[k].z));
            d[5 * k + 3] = toBf16(inside ? -wgt * resp * dlogf * kap : 0.0f);
            d[5 * k + 4] = toBf16(-wgt * (resp - m.w[k]));
        }
        for (int k = kGuideOut; k < 64; ++k) d[k] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if ((threadIdx.x & 63) == 0 && loss != 0.0f) atomicAdd(wordF(words, kGwLoss), loss);
}

// deltaPrev (count x 64) = (delta . W_l) * (actPrev > 0), l >= 1.  W_l is
// staged transposed in LDS (WT[in][out]) so B fragments are contiguous.
__global__ __launch_bounds__(256) void kGuideBackLayer(int l, const uint32_t* __restrict__ words,
                                                       const float* __restrict__ params,
                                                       const uint16_t* __restrict__ delta,
                                                       const uint16_t* __restrict__ actPrev,
                                                       uint16_t* __restrict__ deltaPrev) {
    __shared__ __attribute__((aligned(16))) uint16_t WT[64 * 64];
    const uint32_t count = words[kGwBatch];
    if (blockIdx.x * 128 >= count) return;
    const int nin = guideLayerIn(l), nout = guideLayerOut(l);
    for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
        const int in = i / 64, o = i % 64;
        WT[i] = (o < nout && in < nin) ? toBf16(params[guideWeightOffset(l) + o * nin + in]) : uint16_t(0);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, kq = 4 * (lane >> 5);
    const uint32_t totalWaves = gridDim.x * 4;
    for (uint32_t blk = blockIdx.x * 4 + wave; blk * 32 < count; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        const uint32_t ra = min(row0 + r, count - 1);
        f32x16 c0 = zero16(), c1 = zero16();
        for (int kb = 0; kb < 64; kb += 8) {
            const bf16x4 a = ld4(delta + size_t(ra) * 64 + kb + kq);
            const bf16x4 b0 = ld4(WT + r * 64 + kb + kq);
            const bf16x4 b1 = ld4(WT + (32 + r) * 64 + kb + kq);
            c0 = mfma(a, b0, c0);
            c1 = mfma(a, b1, c1);
        }
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = row0 + 8 * j + kq + i;
                if (row >= count) continue;
                const uint16_t p0 = actPrev[size_t(row) * 64 + r], p1 = actPrev[size_t(row) * 64 + 32 + r];
                deltaPrev[size_t(row) * 64 + r] = fromBf16(p0) > 0.0f ? toBf16(c0[4 * j + i]) : uint16_t(0);
                deltaPrev[size_t(row) * 64 + 32 + r] = fromBf16(p1) > 0.0f ? toBf16(c1[4 * j + i]) : uint16_t(0);
            }
    }
}

// grad W_l += delta^T actPrev and grad b_l += column sums of delta over rows
// [blockIdx.x * kGradRowsPerBlock, ...); FP32 atomics into grad.  Output
// tiles (32 outputs x 32 inputs) go to the 4 waves round-robin (4 or 6 tiles).
constexpr uint32_t kGradRowsPerBlock = 512;
__global__ __launch_bounds__(256) void kGuideWeightGrad(int l, const uint32_t* __restrict__ words,
                                                        const uint16_t* __restrict__ delta,
                                                        const uint16_t* __restrict__ actPrev, uint32_t actStride,
                                                        float* __restrict__ grad) {
    __shared__ __attribute__((aligned(16))) uint16_t dT[64 * 32];  // [out][row]
    __shared__ __attribute__((aligned(16))) uint16_t aT[96 * 32];  // [in][row]
    const uint32_t count = words[kGwBatch];
    const uint32_t begin = blockIdx.x * kGradRowsPerBlock;
    if (begin >= count) return;
    const uint32_t end = min(begin + kGradRowsPerBlock, count);
    const int nin = guideLayerIn(l), nout = guideLayerOut(l);
    const int ntIn = (nin + 31) / 32;  // 3 for layer 0, else 2
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, kq = 4 * (lane >> 5);
    f32x16 acc[2] = { zero16(), zero16() };
    float bsum = 0.0f;
    for (uint32_t row0 = begin; row0 < end; row0 += 32) {
        __syncthreads();
        for (int i = threadIdx.x; i < 64 * 32; i += blockDim.x) {
            const int o = i / 32, rr = i % 32;
            const uint32_t row = row0 + rr;
            dT[i] = row < end ? delta[size_t(row) * 64 + o] : uint16_t(0);
        }
        for (int i = threadIdx.x; i < 96 * 32; i += blockDim.x) {
            const int in = i / 32, rr = i % 32;
            const uint32_t row = row0 + rr;
            aT[i] = (row < end && in < nin) ? actPrev[size_t(row) * actStride + in] : uint16_t(0);
        }
        __syncthreads();
        for (int t = 0; t < 2; ++t) {
            const int tile = wave + 4 * t;
            if (tile >= 2 * ntIn) break;
            const int mt = tile / ntIn, nt = tile % ntIn;
            for (int kb = 0; kb < 32; kb += 8) {
                const bf16x4 a = ld4(dT + (32 * mt + r) * 32 + kb + kq);
                const bf16x4 bb = ld4(aT + (32 * nt + r) * 32 + kb + kq);
                acc[t] = mfma(a, bb, acc[t]);
            }
        }
        if (wave == 0)
            for (int rr = 0; rr < 32; ++rr) bsum += fromBf16(dT[lane * 32 + rr]);
    }
    for (int t = 0; t < 2; ++t) {
        const int tile = wave + 4 * t;
        if (tile >= 2 * ntIn) break;
        const int mt = tile / ntIn, nt = tile % ntIn;
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                const int o = 32 * mt + 8 * j + kq + i, in = 32 * nt + r;
                if (o < nout && in < nin) atomicAdd(&grad[guideWeightOffset(l) + o * nin + in], acc[t][4 * j + i]);
            }
    }
    if (wave == 0 && lane < nout) atomicAdd(&grad[guideBiasOffset(l) + lane], bsum);
}

__global__ __launch_bounds__(256) void kGuideAdam(float* __restrict__ params, float* __restrict__ m,
                                                  float* __restrict__ v, float* __restrict__ grad,
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
    grad[i] = 0.0f;
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

// One thread per path, levels visited in order; every lane of a wave runs
// the same number of iterations (the wave-wide appends need all lanes).
__global__ __launch_bounds__(256) void kGuideCollect(GuideCollectArgs a) {
    const uint32_t stamp = a.g->guideStamp;
    const uint32_t frame = stamp;
    uint32_t* cand = a.words + kGwCand0 + (frame & 1u);
    const uint32_t prev = a.words[kGwCand0 + ((frame + 1u) & 1u)];
    const float keep = prev > a.poolCap ? float(a.poolCap) / float(prev) : 1.0f;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.paths; base += stride) {
        const uint32_t p = base + threadIdx.x;
        const bool valid = p < a.paths;
        const float lf = valid ? luminance(v3(a.radiance[p])) : 0.0f;
        for (uint32_t b = 0; b < a.levels; ++b) {
            const float4* rec = a.pathRec + (size_t(b) * a.paths + (valid ? p : 0u)) * 4;
            bool take = false;
            float w = 0.0f;
            if (valid) {
                const float4 r2 = rec[2];
                if (__float_as_uint(r2.w) == stamp) {
                    w = (lf - rec[1].w) / fmaxf(rec[0].w, 1e-20f);
                    take = w > 0.0f && isfinite(w);
                }
            }
            (void)waveAppend(cand, take);
            const bool kept = take && (keep >= 1.0f || hashUnit(p, b, frame) < keep);
            const uint32_t slot = waveAppend(a.words + kGwPool, kept);
            if (kept && slot < a.poolCap) {
                const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
                float4* dst = a.pool + size_t(slot) * 4;
                dst[0] = make_float4(r0.x, r0.y, r0.z, w);
                dst[1] = make_float4(r1.x, r1.y, r1.z, 0.0f);
                dst[2] = make_float4(r2.x, r2.y, r2.z, 0.0f);
                dst[3] = make_float4(r3.x, r3.y, r3.z, 0.0f);
            }
        }
    }
}

// ---------------------------------------------------------------- launchers --
static uint32_t cappedGrid(uint32_t rows, uint32_t rowsPerBlock, uint32_t cap) {
    uint32_t g = (rows + rowsPerBlock - 1) / rowsPerBlock;
    if (g > cap) g = cap;
    return g ? g : 1;
}

hipError_t launchGuideInfer(const RecView& rec, const uint32_t* countPtr, uint32_t maxCount, const float* params,
                            void* out, bool decoded, hipStream_t s) {
    hipLaunchKernelGGL(kGuideInfer, dim3(cappedGrid(maxCount, 128, 512)), dim3(kInferBlock), 0, s, rec, countPtr,
                       params, out, decoded ? 1 : 0);
    return hipGetLastError();
}

// b.cap bounds the batch; the workspace stride is net.cap.
hipError_t launchGuideGradient(const GuideBatch& b, const GuideNet& net, bool countStep, hipStream_t s) {
    if (b.cap > net.cap || b.cap == 0) return hipErrorInvalidValue;
    hipError_t e;
    hipLaunchKernelGGL(kGuideStepBegin, dim3(1), dim3(64), 0, s, b, net.words, countStep ? 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideWeightSum, dim3(1), dim3(kWeightSumBlock), 0, s, b, net.words);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    GuideBatch ws = b;
    ws.cap = net.cap;
    hipLaunchKernelGGL(kGuideFwdTrain, dim3(cappedGrid(b.cap, 128, 1024)), dim3(256), 0, s, ws, net.words,
                       net.params, net.acts, net.raw);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideOutGrad, dim3(cappedGrid(b.cap, 256, 1u << 20)), dim3(256), 0, s, b, net.words, net.raw,
                       net.delta[0]);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const size_t actBase = size_t(net.cap) * kGuideInPad;
    int cur = 0;
    for (int l = 4; l >= 0; --l) {
        const uint16_t* prev = l == 0 ? net.acts : net.acts + actBase + size_t(l - 1) * net.cap * 64;
        const uint32_t stride = l == 0 ? kGuideInPad : 64;
        hipLaunchKernelGGL(kGuideWeightGrad, dim3(cappedGrid(b.cap, kGradRowsPerBlock, 1u << 20)), dim3(256), 0, s, l,
                           net.words, net.delta[cur], prev, stride, net.grad);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (l > 0) {
            hipLaunchKernelGGL(kGuideBackLayer, dim3(cappedGrid(b.cap, 128, 1024)), dim3(256), 0, s, l, net.words,
                               net.params, net.delta[cur], prev, net.delta[cur ^ 1]);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            cur ^= 1;
        }
    }
    return hipSuccess;
}

hipError_t launchGuideAdam(const GuideNet& net, const GuideAdam& h, hipStream_t s) {
    hipLaunchKernelGGL(kGuideAdam, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, net.params, net.adamM,
                       net.adamV, net.grad, net.words, h);
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
    if (a.paths == 0) return hipSuccess;
    hipLaunchKernelGGL(kGuideCollect, dim3(cappedGrid(a.paths, 256, 2048)), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace dev
} // namespace mrt

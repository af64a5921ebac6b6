// Path-guiding MLP on gfx950 matrix cores: BF16 operands, FP32 accumulate.
//
// Network: 78 (padded 80) -> 64 -> 64 -> 64 -> 64 -> 40, ReLU on hidden
// layers (see guide_device.h for the encoding and the output decoding).
//
// All matrix products use v_mfma_f32_32x32x8bf16_1k.  Operand fragments for
// a 32x32 output tile C = A (32 x 8) * B (8 x 32), lane l (0..63):
//   A: rows m = l % 32, k = 4 * (l / 32) + {0..3}      (4 bf16)
//   B: cols n = l % 32, k = 4 * (l / 32) + {0..3}      (4 bf16)
//   C: col n = l % 32, rows m = 8 j + 4 (l / 32) + i at acc[4 j + i]
//
// Kernels:
//   kGuideInfer      fused forward for rendering: features -> 40 outputs
//   kGuideFwdTrain   forward over a training batch, storing activations
//   kGuideOutGrad    d(weighted NLL)/d(raw outputs) per row
//   kGuideBackLayer  delta_{l-1} = (delta_l W_l) * relu'(a_{l-1})
//   kGuideWeightGrad dW_l += delta_l^T a_{l-1}, db_l += sum delta_l (K split over rows)
//   kGuideAdam       Adam update of the FP32 master weights
#include "guide_device.h"
#include "guide_launch.h"

namespace mrt {
namespace dev {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t toBf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float fromBf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) {  // 8-byte aligned LDS / global
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = short(v.x & 0xFFFFu); r[1] = short(v.x >> 16);
    r[2] = short(v.y & 0xFFFFu); r[3] = short(v.y >> 16);
    return r;
}
__device__ __forceinline__ bf16x4 ld4s(const uint16_t* p, int stride) {  // strided
    bf16x4 r;
    r[0] = short(p[0]); r[1] = short(p[stride]); r[2] = short(p[2 * stride]); r[3] = short(p[3 * stride]);
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

// LDS image of the weights for the forward pass: W_l output-major with the
// input dimension padded (layer 0: 80; output layer padded to 64 rows).
constexpr int kWOff0 = 0;
constexpr int kWOff1 = kWOff0 + 64 * kGuideInPad;
constexpr int kWOff2 = kWOff1 + 64 * 64;
constexpr int kWOff3 = kWOff2 + 64 * 64;
constexpr int kWOff4 = kWOff3 + 64 * 64;
constexpr int kWTotal = kWOff4 + 64 * 64;  // 21,504 bf16
__host__ __device__ constexpr int wOff(int l) {
    return l == 0 ? kWOff0 : (l == 1 ? kWOff1 : (l == 2 ? kWOff2 : (l == 3 ? kWOff3 : kWOff4)));
}
__host__ __device__ constexpr int wStride(int l) { return l == 0 ? kGuideInPad : 64; }

// Cooperative load of the forward weight image and biases into LDS.
__device__ void loadForwardWeights(const float* __restrict__ params, uint16_t* W, float* bias) {
    for (int i = threadIdx.x; i < kWTotal; i += blockDim.x) {
        int l = 4;
        while (l > 0 && i < wOff(l)) --l;
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

// One layer for the wave's 32 rows held in `act` (row stride kGuideInPad):
// out = W_l act + b; hidden layers write ReLU(out) back into act as bf16;
// the last layer returns the two accumulator tiles.
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
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            c0[4 * j + i] += bb0;
            c1[4 * j + i] += bb1;
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

// Encodes rows [row0, row0 + 32) of the query/training records into act.
__device__ __forceinline__ void encodeRows(const float4* __restrict__ rec, uint32_t recStride, uint32_t row0,
                                           uint32_t count, uint16_t* act) {
    const int lane = threadIdx.x & 63, r = lane & 31, half = lane >> 5;
    const uint32_t row = row0 + r;
    float p[3] = { 0, 0, 0 }, n[3] = { 0, 0, 0 }, wo[3] = { 0, 0, 0 };
    const bool valid = row < count;
    if (valid) {
        const float4 a = rec[size_t(row) * recStride], b = rec[size_t(row) * recStride + 1],
                     c = rec[size_t(row) * recStride + 2];
        p[0] = a.x; p[1] = a.y; p[2] = a.z;
        n[0] = b.x; n[1] = b.y; n[2] = b.z;
        wo[0] = c.x; wo[1] = c.y; wo[2] = c.z;
    }
    for (int k = half * 40; k < half * 40 + 40; ++k)
        act[r * kGuideInPad + k] = valid ? toBf16(guideFeature(p, n, wo, k)) : uint16_t(0);
}

// ------------------------------------------------------------------ inference --
// rec: recStride float4 per row, first three = (pos01, n, wo).  count read
// from the device.  out: 40 floats per row.
__global__ __launch_bounds__(256) void kGuideInfer(const float4* __restrict__ rec, uint32_t recStride,
                                                   const uint32_t* __restrict__ countPtr,
                                                   const float* __restrict__ params, float* __restrict__ out) {
    __shared__ uint16_t W[kWTotal];
    __shared__ float bias[5 * 64];
    __shared__ __attribute__((aligned(16))) uint16_t actAll[4][32 * kGuideInPad];
    loadForwardWeights(params, W, bias);
    __syncthreads();
    const uint32_t count = *countPtr;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, kq = 4 * (lane >> 5);
    uint16_t* act = actAll[wave];
    const uint32_t totalWaves = gridDim.x * 4;
    for (uint32_t blk = blockIdx.x * 4 + wave; blk * 32 < count; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        encodeRows(rec, recStride, row0, count, act);
        f32x16 c0, c1;
        for (int l = 0; l < 5; ++l) layerForward(l, act, W, bias, c0, c1);
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = row0 + 8 * j + kq + i;
                if (row < count) {
                    out[size_t(row) * kGuideOut + r] = c0[4 * j + i];
                    if (r < kGuideOut - 32) out[size_t(row) * kGuideOut + 32 + r] = c1[4 * j + i];
                }
            }
    }
}

// ------------------------------------------------------------------- training --
// Forward over a batch; stores act_0 (N x 80), act_1..act_4 (N x 64 each, bf16)
// and the raw outputs (N x 40, FP32).
__global__ __launch_bounds__(256) void kGuideFwdTrain(const float4* __restrict__ rec, uint32_t recStride,
                                                      uint32_t count, const float* __restrict__ params,
                                                      uint16_t* __restrict__ acts, float* __restrict__ raw) {
    __shared__ uint16_t W[kWTotal];
    __shared__ float bias[5 * 64];
    __shared__ __attribute__((aligned(16))) uint16_t actAll[4][32 * kGuideInPad];
    loadForwardWeights(params, W, bias);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, kq = 4 * (lane >> 5);
    uint16_t* act = actAll[wave];
    const uint32_t totalWaves = gridDim.x * 4;
    uint16_t* act0 = acts;
    for (uint32_t blk = blockIdx.x * 4 + wave; blk * 32 < count; blk += totalWaves) {
        const uint32_t row0 = blk * 32;
        encodeRows(rec, recStride, row0, count, act);
        for (int k = lane; k < 32 * kGuideInPad; k += 64) {
            const uint32_t row = row0 + k / kGuideInPad;
            if (row < count) act0[size_t(row) * kGuideInPad + k % kGuideInPad] = act[k];
        }
        f32x16 c0, c1;
        for (int l = 0; l < 5; ++l) {
            layerForward(l, act, W, bias, c0, c1);
            if (l < 4) {
                uint16_t* dst = acts + size_t(count) * kGuideInPad + size_t(l) * count * 64;
                for (int k = lane; k < 32 * 64; k += 64) {
                    const uint32_t row = row0 + k / 64;
                    if (row < count) dst[size_t(row) * 64 + k % 64] = act[(k / 64) * kGuideInPad + k % 64];
                }
            }
        }
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = row0 + 8 * j + kq + i;
                if (row < count) {
                    raw[size_t(row) * kGuideOut + r] = c0[4 * j + i];
                    if (r < kGuideOut - 32) raw[size_t(row) * kGuideOut + 32 + r] = c1[4 * j + i];
                }
            }
    }
}

// delta_5 = d(sum_j w_j * -log q(wi_j)) / d raw_j, scaled by 1/batchWeight,
// stored as bf16 rows of 64 (columns 40..63 zero).  lossOut accumulates the
// (scaled) loss.  rec row: (pos01, weight), (n), (wo), (wi).
__global__ __launch_bounds__(256) void kGuideOutGrad(const float4* __restrict__ rec, uint32_t recStride,
                                                     uint32_t count, const float* __restrict__ raw,
                                                     const float* __restrict__ weightSum,
                                                     uint16_t* __restrict__ delta, float* __restrict__ lossOut) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    float loss = 0.0f;
    if (row < count) {
        const float scale = 1.0f / fmaxf(*weightSum, 1e-20f);
        const float wgt = rec[size_t(row) * recStride].w * scale;
        const float4 wi4 = rec[size_t(row) * recStride + 3];
        const V3 wi = v3(wi4);
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
    // wave-reduce the loss, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(lossOut, loss);
}

// delta_prev (N x 64) = (delta (N x 64) . W_l (out x in)) * (act_prev > 0), for l >= 1.
// W_l is staged transposed in LDS (WT[in][out]) so B fragments are contiguous.
__global__ __launch_bounds__(256) void kGuideBackLayer(int l, uint32_t count, const float* __restrict__ params,
                                                       const uint16_t* __restrict__ delta,
                                                       const uint16_t* __restrict__ actPrev,
                                                       uint16_t* __restrict__ deltaPrev) {
    __shared__ __attribute__((aligned(16))) uint16_t WT[64 * 64];
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

// grad[W_l] += delta^T act_prev and grad[b_l] += column sums of delta, over
// rows [chunk0, chunk0 + rowsPerBlock) of each block; atomics into FP32.
// Wave w of the block computes output tiles (mt, nt) with index w, w+4.
__global__ __launch_bounds__(256) void kGuideWeightGrad(int l, uint32_t count, uint32_t rowsPerBlock,
                                                        const uint16_t* __restrict__ delta,
                                                        const uint16_t* __restrict__ actPrev, uint32_t actStride,
                                                        float* __restrict__ grad) {
    __shared__ __attribute__((aligned(16))) uint16_t dT[64 * 32];   // [out][row]
    __shared__ __attribute__((aligned(16))) uint16_t aT[96 * 32];   // [in][row]
    const int nin = guideLayerIn(l), nout = guideLayerOut(l);
    const int ntIn = (nin + 31) / 32;  // 3 for layer 0, else 2
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, kq = 4 * (lane >> 5);
    f32x16 acc[2] = { zero16(), zero16() };
    float bsum = 0.0f;
    const uint32_t begin = blockIdx.x * rowsPerBlock;
    const uint32_t end = min(begin + rowsPerBlock, count);
    for (uint32_t row0 = begin; row0 < end; row0 += 32) {
        __syncthreads();
        for (int i = threadIdx.x; i < 64 * 32; i += blockDim.x) {
            const int o = i / 32, rr = i % 32;
            const uint32_t row = row0 + rr;
            dT[i] = (row < end && o < 64) ? delta[size_t(row) * 64 + o] : uint16_t(0);
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
                const bf16x4 b = ld4(aT + (32 * nt + r) * 32 + kb + kq);
                acc[t] = mfma(a, b, acc[t]);
            }
        }
        if (wave == 0) {
            for (int rr = 0; rr < 32; ++rr) bsum += fromBf16(dT[lane * 32 + rr]);
        }
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

__global__ void kGuideAdam(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                           float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGuideParams) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    grad[i] = 0.0f;  // ready for the next step
}

// Sum of the training weights (for normalisation) and clearing of the loss.
__global__ void kGuideWeightSum(const float4* __restrict__ rec, uint32_t recStride, uint32_t count,
                                float* __restrict__ sum) {
    float s = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        s += rec[size_t(i) * recStride].w;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(sum, s);
}

// ---------------------------------------------------------------- launchers --
hipError_t launchGuideInfer(const float4* rec, uint32_t recStride, const uint32_t* countPtr, uint32_t maxCount,
                            const float* params, float* out, uint32_t grid, hipStream_t s) {
    uint32_t g = (maxCount + 127) / 128;
    if (g > grid) g = grid;
    if (g == 0) g = 1;
    hipLaunchKernelGGL(kGuideInfer, dim3(g), dim3(256), 0, s, rec, recStride, countPtr, params, out);
    return hipGetLastError();
}

hipError_t launchGuideTrainStep(const GuideTrainArgs& a, hipStream_t s) {
    const uint32_t n = a.count;
    if (n == 0) return hipSuccess;
    hipError_t e;
    hipLaunchKernelGGL(kGuideWeightSum, dim3(64), dim3(256), 0, s, a.rec, a.recStride, n, a.scratch);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t fwdGrid = (n + 127) / 128 < a.grid ? (n + 127) / 128 : a.grid;
    hipLaunchKernelGGL(kGuideFwdTrain, dim3(fwdGrid), dim3(256), 0, s, a.rec, a.recStride, n, a.params, a.acts, a.raw);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(kGuideOutGrad, dim3((n + 255) / 256), dim3(256), 0, s, a.rec, a.recStride, n, a.raw,
                       a.scratch, a.delta[0], a.scratch + 1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const size_t actBase = size_t(n) * kGuideInPad;
    int cur = 0;
    for (int l = 4; l >= 0; --l) {
        const uint16_t* prev = l == 0 ? a.acts : a.acts + actBase + size_t(l - 1) * n * 64;
        const uint32_t stride = l == 0 ? kGuideInPad : 64;
        const uint32_t rowsPerBlock = 512;
        hipLaunchKernelGGL(kGuideWeightGrad, dim3((n + rowsPerBlock - 1) / rowsPerBlock), dim3(256), 0, s, l, n,
                           rowsPerBlock, a.delta[cur], prev, stride, a.grad);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (l > 0) {
            hipLaunchKernelGGL(kGuideBackLayer, dim3(fwdGrid), dim3(256), 0, s, l, n, a.params, a.delta[cur], prev,
                               a.delta[cur ^ 1]);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            cur ^= 1;
        }
    }
    hipLaunchKernelGGL(kGuideAdam, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, a.params, a.adamM, a.adamV,
                       a.grad, a.lr, a.beta1, a.beta2, a.eps, a.biasCorr1, a.biasCorr2);
    return hipGetLastError();
}

hipError_t launchGuideForwardBatch(const float4* rec, uint32_t recStride, uint32_t count, const float* params,
                                   uint16_t* acts, float* raw, uint32_t grid, hipStream_t s) {
    const uint32_t g = (count + 127) / 128 < grid ? (count + 127) / 128 : grid;
    hipLaunchKernelGGL(kGuideFwdTrain, dim3(g ? g : 1), dim3(256), 0, s, rec, recStride, count, params, acts, raw);
    return hipGetLastError();
}

} // namespace dev
} // namespace mrt

// Kernels of the warm-start network (layout and parameter vector in mlp.h).
//
// One wavefront evaluates 16 samples at a time on the matrix cores. Lane
// l = 16*g + c (g = l / 16, c = l % 16). With the 16x16x16 BF16 instruction:
//   A operand (weights, 16 rows x 16 k): row c, k = 4g..4g+3
//   B operand (activations, 16 k x 16 samples): k = 4g..4g+3, sample c
//   accumulator: rows 4g..4g+3, sample c
// so a layer's accumulator, after the activation and conversion to BF16, is
// the B operand of the next layer's product over the same 16 rows. Products
// over 64 features use gfx950's 16x16x32 BF16 instruction (two per 16
// output rows instead of four), whose lane l holds k = 8g..8g+7: its B
// operand for features [32t, 32t + 32) is the pair of accumulator tiles 2t
// and 2t + 1 side by side (features 32t + 4g + v, then 32t + 16 + 4g + v),
// and its A operand takes the weight columns in that same order (two runs of
// four from LDS). The sum runs over the same features in another order.
#include "mlp.h"

#include <hip/hip_ext.h>

#include <math.h>

namespace b2mlp {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;           // wavefronts per workgroup
constexpr int kMaxGradBlocks = 256; // workgroups of one gradient launch at most
constexpr int kThreads = 64 * kWaves;
constexpr int kBiasCount = 3 * kHidden + kOutPad;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two 4-element halves side by side as an 8-element operand.
__device__ inline bf16x8 join8(bf16x4 lo, bf16x4 hi) {
    const short s[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    bf16x8 r;
    __builtin_memcpy(&r, s, sizeof(r));
    return r;
}

// 16x16x32: features [32t, 32t + 32) of a 64-feature product, as above.
__device__ inline f32x4 mfma32(bf16x4 aLo, bf16x4 aHi, bf16x4 bLo, bf16x4 bHi, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(join8(aLo, aHi), join8(bLo, bHi), c, 0, 0, 0);
}

// Round to nearest even BF16 by the hardware conversion (the toBf16 rule for
// finite inputs).
__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
    const __bf16 h[4] = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
    bf16x4 r;
    __builtin_memcpy(&r, h, sizeof(r));
    return r;
}

// Four consecutive BF16 values (8-byte aligned) from LDS.
__device__ inline bf16x4 frag(const uint16_t* p) {
    bf16x4 r;
    __builtin_memcpy(&r, p, sizeof(r));
    return r;
}

// GELU with the Gaussian CDF, 0.5 x (1 + erf(x / sqrt 2)), and its
// derivative cdf(x) + x pdf(x). erf is Abramowitz and Stegun's rational
// approximation 7.1.26 (Handbook of Mathematical Functions, 1964; absolute
// error below 1.5e-7, far under the BF16 rounding of the activations), whose
// exponential exp(-z^2) at z = x / sqrt 2 is also the density's exp(-x^2 / 2),
// so both come from one hardware exponential and one reciprocal instead of
// the library's erff and expf.
struct Cdf {
    float cdf, pdf;
};
__device__ inline Cdf gaussCdf(float x) {
    const float z = fabsf(x) * 0.70710678118654752f;
    const float e = __expf(-0.5f * x * x);
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * z);
    const float poly =
        t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    const float erfAbs = 1.0f - poly * e;
    Cdf r;
    r.cdf = 0.5f * (1.0f + (x < 0.0f ? -erfAbs : erfAbs));
    r.pdf = 0.39894228040143268f * e;
    return r;
}
__device__ inline float gelu(float x) { return x * gaussCdf(x).cdf; }
__device__ inline float geluGrad(float x) {
    const Cdf c = gaussCdf(x);
    return c.cdf + x * c.pdf;
}

struct Smem {
    uint16_t w[hTotal];
    float b[kBiasCount];   // b1 | b2 | b3 | b4 (padded)
    float norm[kNorm];
};

__device__ inline void loadSmem(Smem& S, const uint16_t* wh, const float* params, const float* norm) {
    const uint32_t* src = (const uint32_t*)wh;
    uint32_t* dst = (uint32_t*)S.w;
    for (int i = threadIdx.x; i < hTotal / 2; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < kHidden; i += blockDim.x) {
        S.b[i] = params[kB1 + i];
        S.b[kHidden + i] = params[kB2 + i];
        S.b[2 * kHidden + i] = params[kB3 + i];
    }
    for (int i = threadIdx.x; i < kOutPad; i += blockDim.x) S.b[3 * kHidden + i] = params[kB4 + i];
    for (int i = threadIdx.x; i < kNorm; i += blockDim.x) S.norm[i] = norm[i];
}

// Input fragment of sample s (features 4g..4g+3), standardised; zero when
// the sample is out of range.
__device__ inline bf16x4 inputFrag(const Smem& S, const float* x, long long s, bool valid, int g) {
    if (!valid) return pack4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 v = *(const float4*)(x + s * kIn + 4 * g);
    const int f = 4 * g;
    return pack4((v.x - S.norm[kNormXMean + f]) * S.norm[kNormXInv + f],
                 (v.y - S.norm[kNormXMean + f + 1]) * S.norm[kNormXInv + f + 1],
                 (v.z - S.norm[kNormXMean + f + 2]) * S.norm[kNormXInv + f + 2],
                 (v.w - S.norm[kNormXMean + f + 3]) * S.norm[kNormXInv + f + 3]);
}

__device__ inline f32x4 biasTile(const Smem& S, int base, int g) {
    f32x4 r;
    for (int j = 0; j < 4; ++j) r[j] = S.b[base + 4 * g + j];
    return r;
}

// The product of weight rows [16mt, 16mt + 16) (row stride ldw, in LDS at
// wOff) with the four 16-feature tiles in[0..3] of a 64-feature input, added
// to acc: two 16x16x32 products.
__device__ inline f32x4 product64(const Smem& S, int wOff, int ldw, int mt, const bf16x4* in, f32x4 acc, int g,
                                  int c) {
    const uint16_t* row = &S.w[wOff + (16 * mt + c) * ldw];
#pragma unroll
    for (int t = 0; t < 2; ++t)
        acc = mfma32(frag(row + 32 * t + 4 * g), frag(row + 32 * t + 16 + 4 * g), in[2 * t], in[2 * t + 1], acc);
    return acc;
}

// Hidden layer: acc[mt] = W[16mt + c][.] x in + bias; the 16-feature input
// layer with one 16x16x16 product, the 64-feature ones with product64.
template <int KT>
__device__ inline void hiddenLayer(const Smem& S, int wOff, int ldw, int bOff, const bf16x4* in, f32x4* z,
                                   bf16x4* a, int g, int c) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 acc = biasTile(S, bOff + 16 * mt, g);
        if (KT == 1) acc = mfma(frag(&S.w[wOff + (16 * mt + c) * ldw + 4 * g]), in[0], acc);
        else acc = product64(S, wOff, ldw, mt, in, acc, g, c);
        z[mt] = acc;
        a[mt] = pack4(gelu(acc[0]), gelu(acc[1]), gelu(acc[2]), gelu(acc[3]));
    }
}

__device__ inline f32x4 outputLayer(const Smem& S, const bf16x4* a3, int g, int c) {
    return product64(S, hW4, kHidden, 0, a3, biasTile(S, 3 * kHidden, g), g, c);
}

// ------------------------------------------------------------ inference
// MODE 0: y[2s], y[2s+1] = outputs (raw units).
// MODE 1: impulses of fresh contact points (see mlpPredictImpulses).
// MODE 2: impulses of any contact point as the carried impulse (features 9
//         and 15, velocity-change units) plus the outputs.
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_mlp_infer(const uint16_t* wh, const float* params, const float* norm,
                                                        const float* x, int count, const int* countDev, float* y,
                                                        const int* ref, const float4* nf, float4* imp) {
    __shared__ Smem S;
    loadSmem(S, wh, params, norm);
    __syncthreads();
    const int n = countDev ? min(*countDev, count) : count;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    const int tiles = (n + 15) / 16;
    for (int t = blockIdx.x * kWaves + wave; t < tiles; t += gridDim.x * kWaves) {
        const long long s = (long long)t * 16 + c;
        const bool valid = s < n;
        bf16x4 x0[1] = {inputFrag(S, x, s, valid, g)};
        f32x4 z[4];
        bf16x4 a1[4], a2[4], a3[4];
        hiddenLayer<1>(S, hW1, kIn, 0, x0, z, a1, g, c);
        hiddenLayer<4>(S, hW2, kHidden, kHidden, a1, z, a2, g, c);
        hiddenLayer<4>(S, hW3, kHidden, 2 * kHidden, a2, z, a3, g, c);
        const f32x4 o = outputLayer(S, a3, g, c);
        if (g == 0 && valid) {
            const float y0 = o[0] * S.norm[kNormYStd] + S.norm[kNormYMean];
            const float y1 = o[1] * S.norm[kNormYStd + 1] + S.norm[kNormYMean + 1];
            if (MODE == 0) {
                y[2 * s] = y0;
                y[2 * s + 1] = y1;
            } else {
                // outputs are in velocity-change units: scale by the
                // point's normal and tangent mass (features 10, 11)
                const int r = ref[s];
                const int k = r >> 1, j = r & 1;
                const float mu = nf[k].z;
                const float un = MODE == 2 ? y0 + x[s * kIn + 9] : y0;
                const float ut = MODE == 2 ? y1 + x[s * kIn + 15] : y1;
                const float yn = un * x[s * kIn + 10], yt = ut * x[s * kIn + 11];
                const float ln = yn > 0.0f ? yn : 0.0f;
                const float lim = mu * ln;
                const float lt = yt < -lim ? -lim : (yt > lim ? lim : yt);
                float* p = (float*)&imp[k];
                p[2 * j] = ln;
                p[2 * j + 1] = lt;
            }
        }
    }
}

// ------------------------------------------------------------ training
// Per-wave 16x16 transpose through LDS: in holds rows 4g..4g+3 of column c;
// the result holds row c, columns 4g..4g+3, as BF16.
__device__ inline bf16x4 transposeF(float (*T)[17], f32x4 d, int g, int c) {
    for (int j = 0; j < 4; ++j) T[4 * g + j][c] = d[j];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const bf16x4 r = pack4(T[c][4 * g], T[c][4 * g + 1], T[c][4 * g + 2], T[c][4 * g + 3]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return r;
}
__device__ inline bf16x4 transposeH(uint16_t (*T)[18], bf16x4 d, int g, int c) {
    for (int j = 0; j < 4; ++j) T[4 * g + j][c] = (uint16_t)d[j];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    bf16x4 r;
    for (int j = 0; j < 4; ++j) r[j] = (short)T[c][4 * g + j];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return r;
}

__device__ inline bf16x4 packTile(f32x4 d) { return pack4(d[0], d[1], d[2], d[3]); }

// Sum over the 16 lanes that share g (lanes 16g..16g+15).
__device__ inline float groupSum16(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

__global__ __launch_bounds__(kThreads) void k_mlp_grad(const uint16_t* wh, const float* params, const float* norm,
                                                       const float* x, const float* tgt, const int* order, int first,
                                                       int count, float invN, float* partial) {
    __shared__ Smem S;
    __shared__ float red[kParams];
    __shared__ float lossShared;
    __shared__ float tf[kWaves][16][17];
    __shared__ uint16_t th[kWaves][16][18];
    loadSmem(S, wh, params, norm);
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) red[i] = 0.0f;
    if (threadIdx.x == 0) lossShared = 0.0f;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    float (*TF)[17] = tf[wave];
    uint16_t (*TH)[18] = th[wave];

    f32x4 gW1[4], gW2[4][4], gW3[4][4], gW4[4];
    float gb1[4][4], gb2[4][4], gb3[4][4], gb4[4];
    const f32x4 zero4 = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 4; ++i) {
        gW1[i] = zero4;
        gW4[i] = zero4;
        gb4[i] = 0.0f;
        for (int k = 0; k < 4; ++k) {
            gW2[i][k] = zero4;
            gW3[i][k] = zero4;
            gb1[i][k] = gb2[i][k] = gb3[i][k] = 0.0f;
        }
    }
    float loss = 0.0f;
    const int tiles = (count + 15) / 16;
    for (int t = blockIdx.x * kWaves + wave; t < tiles; t += gridDim.x * kWaves) {
        const int local = t * 16 + c;
        const bool valid = local < count;
        const long long s = valid ? (order ? (long long)order[first + local] : (long long)first + local) : 0;
        // forward, keeping pre-activations and activations
        bf16x4 x0[1] = {inputFrag(S, x, s, valid, g)};
        f32x4 z1[4], z2[4], z3[4];
        bf16x4 a1[4], a2[4], a3[4];
        hiddenLayer<1>(S, hW1, kIn, 0, x0, z1, a1, g, c);
        hiddenLayer<4>(S, hW2, kHidden, kHidden, a1, z2, a2, g, c);
        hiddenLayer<4>(S, hW3, kHidden, 2 * kHidden, a2, z3, a3, g, c);
        const f32x4 o = outputLayer(S, a3, g, c);
        // dL/dy on standardised targets (rows 0, 1 of the output tile are in lanes g == 0)
        f32x4 d4 = zero4;
        if (g == 0 && valid) {
            const float t0 = (tgt[2 * s] - S.norm[kNormYMean]) / S.norm[kNormYStd];
            const float t1 = (tgt[2 * s + 1] - S.norm[kNormYMean + 1]) / S.norm[kNormYStd + 1];
            const float e0 = o[0] - t0, e1 = o[1] - t1;
            loss += 0.5f * (e0 * e0 + e1 * e1) * invN;
            d4[0] = e0 * invN;
            d4[1] = e1 * invN;
        }
        // layer 4: weight gradient and backward-data
        {
            const bf16x4 dT = transposeF(TF, d4, g, c);
#pragma unroll
            for (int ni = 0; ni < 4; ++ni) gW4[ni] = mfma(dT, transposeH(TH, a3[ni], g, c), gW4[ni]);
            for (int j = 0; j < 4; ++j) gb4[j] += d4[j];
        }
        f32x4 d3[4];
        {
            const bf16x4 dB = packTile(d4);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                f32x4 acc = mfma(frag(&S.w[hW4t + (16 * mt + c) * kOutPad + 4 * g]), dB, zero4);
                for (int j = 0; j < 4; ++j) acc[j] *= geluGrad(z3[mt][j]);
                d3[mt] = acc;
            }
        }
        // layer 3
        {
            bf16x4 hT[4], dB[4];
            for (int ni = 0; ni < 4; ++ni) hT[ni] = transposeH(TH, a2[ni], g, c);
#pragma unroll
            for (int mo = 0; mo < 4; ++mo) {
                const bf16x4 dT = transposeF(TF, d3[mo], g, c);
#pragma unroll
                for (int ni = 0; ni < 4; ++ni) gW3[mo][ni] = mfma(dT, hT[ni], gW3[mo][ni]);
                for (int j = 0; j < 4; ++j) gb3[mo][j] += d3[mo][j];
                dB[mo] = packTile(d3[mo]);
            }
            f32x4 d2[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                f32x4 acc = product64(S, hW3t, kHidden, mt, dB, zero4, g, c);
                for (int j = 0; j < 4; ++j) acc[j] *= geluGrad(z2[mt][j]);
                d2[mt] = acc;
            }
            // layer 2
            for (int ni = 0; ni < 4; ++ni) hT[ni] = transposeH(TH, a1[ni], g, c);
#pragma unroll
            for (int mo = 0; mo < 4; ++mo) {
                const bf16x4 dT = transposeF(TF, d2[mo], g, c);
#pragma unroll
                for (int ni = 0; ni < 4; ++ni) gW2[mo][ni] = mfma(dT, hT[ni], gW2[mo][ni]);
                for (int j = 0; j < 4; ++j) gb2[mo][j] += d2[mo][j];
                dB[mo] = packTile(d2[mo]);
            }
            // layer 1
            const bf16x4 xT = transposeH(TH, x0[0], g, c);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                f32x4 acc = product64(S, hW2t, kHidden, mt, dB, zero4, g, c);
                for (int j = 0; j < 4; ++j) acc[j] *= geluGrad(z1[mt][j]);
                const bf16x4 dT = transposeF(TF, acc, g, c);
                gW1[mt] = mfma(dT, xT, gW1[mt]);
                for (int j = 0; j < 4; ++j) gb1[mt][j] += acc[j];
            }
        }
    }
    // Reduce the four wavefronts in LDS, one wavefront after another (each
    // lane owns distinct entries, so plain read-add-writes): accumulator
    // element j of tile (mo, ni) is dW[16mo + 4g + j][16ni + c]. Then the
    // workgroup's sums go to its row of `partial`, which k_mlp_reduce adds up
    // in a fixed order: the gradient does not depend on the schedule.
    float l = loss;
    for (int off = 1; off < 64; off <<= 1) l += __shfl_xor(l, off);
    for (int wv = 0; wv < kWaves; ++wv) {
        if (wave == wv) {
            for (int mo = 0; mo < 4; ++mo)
                for (int j = 0; j < 4; ++j) {
                    const int o = 16 * mo + 4 * g + j;
                    red[kW1 + o * kIn + c] += gW1[mo][j];
                    for (int ni = 0; ni < 4; ++ni) {
                        red[kW2 + o * kHidden + 16 * ni + c] += gW2[mo][ni][j];
                        red[kW3 + o * kHidden + 16 * ni + c] += gW3[mo][ni][j];
                    }
                }
            for (int ni = 0; ni < 4; ++ni)
                for (int j = 0; j < 4; ++j) red[kW4 + (4 * g + j) * kHidden + 16 * ni + c] += gW4[ni][j];
            for (int mt = 0; mt < 4; ++mt)
                for (int j = 0; j < 4; ++j) {
                    const float s1 = groupSum16(gb1[mt][j]), s2 = groupSum16(gb2[mt][j]), s3 = groupSum16(gb3[mt][j]);
                    if (c == 0) {
                        red[kB1 + 16 * mt + 4 * g + j] += s1;
                        red[kB2 + 16 * mt + 4 * g + j] += s2;
                        red[kB3 + 16 * mt + 4 * g + j] += s3;
                    }
                }
            for (int j = 0; j < 4; ++j) {
                const float s4 = groupSum16(gb4[j]);
                if (c == 0) red[kB4 + 4 * g + j] += s4;
            }
            if (lane == 0) lossShared += l;
        }
        __syncthreads();
    }
    float* row = partial + (size_t)blockIdx.x * (kParams + 1);
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) row[i] = red[i];
    if (threadIdx.x == 0) row[kParams] = lossShared;
}

// Sum of the workgroups' rows added to the gradient and the loss, in a
// fixed order: a workgroup takes kReduceCols consecutive entries; its
// kReduceSplit lane groups each sum every kReduceSplit-th row in row order
// (coalesced across the entries), then the groups' sums are added in group
// order.
constexpr int kReduceCols = 64, kReduceSplit = 4;
__global__ __launch_bounds__(kReduceCols * kReduceSplit) void k_mlp_reduce(const float* partial, int rows,
                                                                            float* grad, float* loss) {
    __shared__ float part[kReduceSplit][kReduceCols];
    const int col = threadIdx.x % kReduceCols, grp = threadIdx.x / kReduceCols;
    const int i = blockIdx.x * kReduceCols + col;
    float acc = 0.0f;
    if (i <= kParams)
        for (int r = grp; r < rows; r += kReduceSplit) acc += partial[(size_t)r * (kParams + 1) + i];
    part[grp][col] = acc;
    __syncthreads();
    if (grp != 0 || i > kParams) return;
    float sum = part[0][col];
    for (int k = 1; k < kReduceSplit; ++k) sum += part[k][col];
    if (i < kParams) grad[i] += sum;
    else *loss += sum;
}

// BF16 image entries of parameter i (weights only; biases stay FP32).
__device__ inline void writeImage(uint16_t* wh, int i, float v) {
    const uint16_t h = toBf16(v);
    if (i < kB1) {
        wh[hW1 + i] = h;
    } else if (i >= kW2 && i < kB2) {
        const int r = i - kW2, o = r / kHidden, in = r % kHidden;
        wh[hW2 + r] = h;
        wh[hW2t + in * kHidden + o] = h;
    } else if (i >= kW3 && i < kB3) {
        const int r = i - kW3, o = r / kHidden, in = r % kHidden;
        wh[hW3 + r] = h;
        wh[hW3t + in * kHidden + o] = h;
    } else if (i >= kW4 && i < kB4) {
        const int r = i - kW4, o = r / kHidden, in = r % kHidden;
        wh[hW4 + r] = h;
        wh[hW4t + in * kOutPad + o] = h;
    }
}

__global__ void k_mlp_image(const float* params, uint16_t* wh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParams) writeImage(wh, i, params[i]);
}

__global__ void k_mlp_adam(float* params, float* grad, float* m, float* v, uint16_t* wh, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float gi = grad[i];
    grad[i] = 0.0f;
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float p = params[i] - lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
    params[i] = p;
    writeImage(wh, i, p);
}

int gridFor(int count) {
    const int tiles = (count + 15) / 16;
    int g = (tiles + kWaves - 1) / kWaves;
    if (g < 1) g = 1;
    return g < 1024 ? g : 1024;
}

}  // namespace

hipError_t mlpAlloc(Net& n) {
    hipError_t e;
    if ((e = hipMalloc((void**)&n.params, sizeof(float) * kParams)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&n.grad, sizeof(float) * kParams)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&n.m, sizeof(float) * kParams)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&n.v, sizeof(float) * kParams)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&n.wh, sizeof(uint16_t) * hTotal)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&n.norm, sizeof(float) * kNorm)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&n.loss, sizeof(float))) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&n.partial, sizeof(float) * kMaxGradBlocks * (kParams + 1))) != hipSuccess) return e;
    if ((e = hipMemset(n.params, 0, sizeof(float) * kParams)) != hipSuccess) return e;
    if ((e = hipMemset(n.grad, 0, sizeof(float) * kParams)) != hipSuccess) return e;
    if ((e = hipMemset(n.m, 0, sizeof(float) * kParams)) != hipSuccess) return e;
    if ((e = hipMemset(n.v, 0, sizeof(float) * kParams)) != hipSuccess) return e;
    if ((e = hipMemset(n.wh, 0, sizeof(uint16_t) * hTotal)) != hipSuccess) return e;
    if ((e = hipMemset(n.norm, 0, sizeof(float) * kNorm)) != hipSuccess) return e;
    return hipMemset(n.loss, 0, sizeof(float));
}

void mlpFree(Net& n) {
    void* ps[] = {n.params, n.grad, n.m, n.v, n.wh, n.norm, n.loss, n.partial};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    n = Net();
}

hipError_t mlpPrepare(const Net& n, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_image, dim3((kParams + 255) / 256), dim3(256), 0, s, n.params, n.wh);
    return hipGetLastError();
}

hipError_t mlpForward(const Net& n, const float* x, int count, float* y, hipStream_t s, hipEvent_t e0,
                      hipEvent_t e1) {
    if (count <= 0) return hipSuccess;
    hipExtLaunchKernelGGL(k_mlp_infer<0>, dim3(gridFor(count)), dim3(kThreads), 0, s, e0, e1, 0, n.wh, n.params,
                          n.norm, x, count, (const int*)nullptr, y, (const int*)nullptr, (const float4*)nullptr,
                          (float4*)nullptr);
    return hipGetLastError();
}

hipError_t mlpPredictImpulses(const Net& n, const float* feat, const int* ref, const int* countDev, int maxCount,
                              const float4* normalFriction, float4* imp, bool delta, hipStream_t s, hipEvent_t e0,
                              hipEvent_t e1) {
    if (maxCount <= 0) return hipSuccess;
    if (delta)
        hipExtLaunchKernelGGL(k_mlp_infer<2>, dim3(gridFor(maxCount)), dim3(kThreads), 0, s, e0, e1, 0, n.wh,
                              n.params, n.norm, feat, maxCount, countDev, (float*)nullptr, ref, normalFriction, imp);
    else
        hipExtLaunchKernelGGL(k_mlp_infer<1>, dim3(gridFor(maxCount)), dim3(kThreads), 0, s, e0, e1, 0, n.wh,
                              n.params, n.norm, feat, maxCount, countDev, (float*)nullptr, ref, normalFriction, imp);
    return hipGetLastError();
}

hipError_t mlpGradient(const Net& n, const float* x, const float* t, const int* order, int first, int count,
                       hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (count <= 0) return hipSuccess;
    // one workgroup per CU at most (the kernel's LDS allows one), each
    // wavefront taking as many 16-sample tiles as needed
    const int tiles = (count + 15) / 16;
    int grid = (tiles + kWaves - 1) / kWaves;
    if (grid > kMaxGradBlocks) grid = kMaxGradBlocks;
    if (grid < 1) grid = 1;
    hipExtLaunchKernelGGL(k_mlp_grad, dim3(grid), dim3(kThreads), 0, s, e0, e1, 0, n.wh, n.params, n.norm, x, t,
                          order, first, count, 1.0f / (float)count, n.partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 1 + kReduceCols - 1) / kReduceCols),
                       dim3(kReduceCols * kReduceSplit), 0, s, n.partial, grid, n.grad, n.loss);
    return hipGetLastError();
}

hipError_t mlpAdam(Net& n, float lr, hipStream_t s) {
    ++n.adamStep;
    const double c1 = 1.0 / (1.0 - pow(0.9, n.adamStep));
    const double c2 = 1.0 / (1.0 - pow(0.999, n.adamStep));
    hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, n.params, n.grad, n.m, n.v, n.wh, lr,
                       (float)c1, (float)c2);
    return hipGetLastError();
}

}  // namespace b2mlp

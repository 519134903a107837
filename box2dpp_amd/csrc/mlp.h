// Warm-start network of the learned warm start: a 16-64-64-64-2 multilayer
// perceptron with GELU activations, BF16 weights and activations on the
// matrix cores (v_mfma_f32_16x16x32_bf16 for the 64-feature products,
// v_mfma_f32_16x16x16_bf16 for the 16-wide ones), FP32 accumulation, FP32
// master weights trained with Adam by the kernels in mlp.hip.
//
// Activations are kept in a "feature x sample" layout: a wavefront owns 16
// samples, and the accumulator tile of one layer (lane = 16*g + c holds
// features 4g..4g+3 of sample c) is directly the B operand of the next
// layer, so forward and backward-data products chain without data
// movement. Weight gradients need the transposed tile and go through LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b2mlp {

constexpr int kIn = 16, kHidden = 64, kOut = 2, kOutPad = 16;

// FP32 parameter vector: weights row-major [out][in], then the bias.
constexpr int kW1 = 0;                          // [64][16]
constexpr int kB1 = kW1 + kHidden * kIn;        // [64]
constexpr int kW2 = kB1 + kHidden;              // [64][64]
constexpr int kB2 = kW2 + kHidden * kHidden;    // [64]
constexpr int kW3 = kB2 + kHidden;              // [64][64]
constexpr int kB3 = kW3 + kHidden * kHidden;    // [64]
constexpr int kW4 = kB3 + kHidden;              // [16][64], rows >= kOut stay zero
constexpr int kB4 = kW4 + kOutPad * kHidden;    // [16], entries >= kOut stay zero
constexpr int kParams = kB4 + kOutPad;

// BF16 weight image (element offsets): forward matrices [out][in] and the
// transposed ones [in][out] used by the backward-data products.
constexpr int hW1 = 0;
constexpr int hW2 = hW1 + kHidden * kIn;
constexpr int hW3 = hW2 + kHidden * kHidden;
constexpr int hW4 = hW3 + kHidden * kHidden;
constexpr int hW2t = hW4 + kOutPad * kHidden;
constexpr int hW3t = hW2t + kHidden * kHidden;
constexpr int hW4t = hW3t + kHidden * kHidden;
constexpr int hTotal = hW4t + kHidden * kOutPad;

// Normalisation block: feature mean[16], 1/std[16], target mean[2], std[2].
constexpr int kNormXMean = 0, kNormXInv = 16, kNormYMean = 32, kNormYStd = 34, kNorm = 36;

struct Net {
    float* params = nullptr;  // kParams, FP32 master weights
    float* grad = nullptr;    // kParams, accumulated by mlpGradient, cleared by mlpAdam
    float* m = nullptr;       // Adam first moment
    float* v = nullptr;       // Adam second moment
    uint16_t* wh = nullptr;   // hTotal BF16 weights
    float* norm = nullptr;    // kNorm
    float* loss = nullptr;    // one float: loss sum of the last mlpGradient calls
    float* partial = nullptr; // kMaxGradBlocks x (kParams + 1): per-workgroup gradient and loss
    int adamStep = 0;
    bool ready = false;       // parameters set (initialised, trained or loaded)
};

hipError_t mlpAlloc(Net& n);
void mlpFree(Net& n);
// BF16 image from the FP32 master weights.
hipError_t mlpPrepare(const Net& n, hipStream_t s);
// y[2i..2i+1] = network(x[16i..16i+15]) for i < count; raw (unnormalised)
// features in, raw impulses out.
hipError_t mlpForward(const Net& n, const float* x, int count, float* y, hipStream_t s, hipEvent_t e0 = nullptr,
                      hipEvent_t e1 = nullptr);
// Adds to n.grad the gradient of L = 1/(2N) sum_i sum_o (y_io - t_io)^2 over
// the samples order[first..first+count) (order may be null: identity) on
// standardised targets, N = count, and adds L to *n.loss.
hipError_t mlpGradient(const Net& n, const float* x, const float* t, const int* order, int first, int count,
                       hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// One Adam step from n.grad (then cleared); refreshes the BF16 image.
hipError_t mlpAdam(Net& n, float lr, hipStream_t s);
// Impulses for the candidate contact points of the step: rows of feat
// (count x 16) through the network, clamped to the friction cone, written
// into the constraint impulses imp[ref[i] >> 1] component (ref[i] & 1). With
// delta, the outputs are added to the carried impulse (features 9 and 15,
// velocity-change units) before scaling by the point's masses.
hipError_t mlpPredictImpulses(const Net& n, const float* feat, const int* ref, const int* countDev, int maxCount,
                              const float4* normalFriction, float4* imp, bool delta, hipStream_t s,
                              hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Matrix-core operations per sample as executed (padded output layer):
// forward 2 * (16*64 + 64*64 + 64*64 + 64*16); a training sample runs the
// forward, the backward-data products of layers 4..2 and the four
// weight-gradient products.
constexpr double kForwardFlops = 2.0 * (kIn * kHidden + 2 * kHidden * kHidden + kHidden * kOutPad);
constexpr double kTrainFlops = 2.0 * kForwardFlops + 2.0 * (kOutPad * kHidden + 2 * kHidden * kHidden);

// Round to nearest even BF16 (finite inputs).
__host__ __device__ inline uint16_t toBf16(float f) {
    union { float f; uint32_t u; } v;
    v.f = f;
    const uint32_t u = v.u + 0x7FFFu + ((v.u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float fromBf16(uint16_t h) {
    union { float f; uint32_t u; } v;
    v.u = (uint32_t)h << 16;
    return v.f;
}

}  // namespace b2mlp

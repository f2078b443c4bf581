// The warm-start network on the device: 4 layers of 3x3 convolutions over the element
// grid, BF16 activations and weights, FP32 accumulation on MFMA, trained with Adam.
#pragma once
#include "nsdg_internal.hpp"

namespace nsdg {
namespace cnn {

constexpr int C = 32;                 // channels of every activation
constexpr int OUT = 8;                // used outputs of the last layer
constexpr int LAYERS = 4;
constexpr int TAPS = 9;
constexpr int WSZ = C * C * TAPS;     // weights of one layer, [co][ci][ky][kx]
constexpr int PSZ = WSZ + C;          // + bias
constexpr int NPARAM = LAYERS * PSZ;
constexpr int FRAG = TAPS * 2 * 64 * 8; // bf16 weight fragments of one layer

struct Net {
    hipStream_t s = nullptr;
    Profiler prof;
    DBuf<float> params, m, v, grad;   // FP32 master values, Adam moments, gradient
    DBuf<uint16_t> wf, wb;            // BF16 fragments: forward, transposed-flipped (backward data)
    DBuf<uint16_t> act[LAYERS];       // act[0] unused (the input), act[1..3] post-ReLU
    DBuf<uint16_t> dz0, dz1;          // gradients of the pre-activations (ping-pong)
    DBuf<float> out, part;            // network output, weight-gradient partials
    DBuf<double> lossPart;            // loss partial sums per block
    size_t pixels = 0;                // capacity of the activation buffers
    int adamStep = 0;

    explicit Net(uint64_t seed);
    ~Net();
    void pack();
    void reserve(size_t npix);
    void forward(const uint16_t* x, int n, int ny, int nx, float* y, bool keep);
    double lossGrad(const uint16_t* x, const float* t, const float* w, int n, int ny, int nx);
    void adam(float lr);
    void sync();
};

} // namespace cnn
} // namespace nsdg

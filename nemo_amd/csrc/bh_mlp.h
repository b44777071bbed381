// Sizes and parameter layout of the learned far-field correction.
#pragma once

namespace bh {

constexpr int kMlpIn = 12;      // features
constexpr int kMlpHid = 64;     // hidden width (3 layers)
constexpr int kMlpOut = 9;      // force residual (3) and its Jacobian at the point (xx xy xz yy yz zz)
constexpr int kMlpOutR = 3;     // outputs 0..2: residual; 3..8: Jacobian
// FP32 master parameter layout
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHid * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid * kMlpHid;
constexpr int kOffW4 = kOffB3 + kMlpHid;
constexpr int kOffB4 = kOffW4 + kMlpOut * kMlpHid;
constexpr int kMlpParams = kOffB4 + kMlpOut;
// normalisation block: output scales of the residual and Jacobian blocks,
// and the size/distance range the model was trained on (cells outside it
// keep the plain multipole force)
constexpr int kMlpNorm = 4;     // residual scale, rlo, rhi, Jacobian scale

}  // namespace bh

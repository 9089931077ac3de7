// ntx_mlp.h -- layout of the BC1 endpoint-proposal network shared by host and
// device code: 48 inputs (a 4x4 block's R, G, B planes, centred on the block
// mean and scaled by the largest deviation, floored), two hidden layers of 64
// with ReLU, 25 outputs: for each of four candidates, offsets of its endpoint
// pair from one diagonal of the block's colour bounding box (same frame), and
// the predicted RMSE of the cluster fit, scaled alike.
#pragma once

namespace ntx_mlp {

constexpr int kIn = 48, kHidden = 64, kOut = 25, kOutPad = 32;
constexpr int kCandidates = 4;

// Offsets into the flat FP32 parameter vector (row-major, out x in).
constexpr int kW1 = 0;                       // [64][48]
constexpr int kB1 = kW1 + kHidden * kIn;     // [64]
constexpr int kW2 = kB1 + kHidden;           // [64][64]
constexpr int kB2 = kW2 + kHidden * kHidden; // [64]
constexpr int kW3 = kB2 + kHidden;           // [25][64]
constexpr int kB3 = kW3 + kOut * kHidden;    // [25]
constexpr int kParams = kB3 + kOut;          // 8921

// Features are scaled by 1 / max(largest deviation, kMinSpread): about one
// 565 step, so that nearly flat blocks keep labels of moderate size.
constexpr float kMinSpread = 1.0f / 32.0f;

// Per-block training label: endpoints A, B of the reference cluster fit and
// its RMSE, all in the block's normalised frame; [7] unused.
constexpr int kLabel = 8;

// Loss weights: the nearest candidate takes the full endpoint error, every
// candidate a kSpread share of its own; kRmseWeight scales the RMSE head.
constexpr float kSpread = 0.05f;
constexpr float kRmseWeight = 1.0f;

}  // namespace ntx_mlp

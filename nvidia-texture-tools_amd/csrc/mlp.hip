// mlp.hip -- learned endpoint proposal for BC1 on gfx950: a 48-64-64-25
// ReLU network in BF16 with FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16_bf16), its training step, and the compressor that
// refines the network's four candidate endpoint pairs by least squares,
// keeps the best, and sends blocks the network is unsure about to the full
// cluster fit (the per-image BC1 encoder's own device code).
//
// The network predicts offsets from the four diagonals of each block's
// colour bounding box, so with a zero output layer it is exactly the
// bounding-box method it refines.
//
// This part has no counterpart in the reference.  Matrix layout: the weights
// are the A operand (rows = outputs, k = inputs) and 16 blocks are the N
// columns, so a layer's 16x16 FP32 result fragment (lane l: rows 4(l/16)+j of
// column l%16) is exactly the next layer's B fragment after bias, ReLU and
// rounding to BF16: the forward pass never leaves registers.
#include "bc1_block.hpp"
#include "ntx_mlp.h"

#include <cfloat>

namespace {

using namespace ntx_mlp;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32, two values per
// instruction).
__device__ __forceinline__ uint16_t bf16_bits(float x) { return __builtin_bit_cast(uint16_t, __bf16(x)); }
__device__ __forceinline__ float bf16_round(float x) { return float(__bf16(x)); }
__device__ __forceinline__ float bf16_float(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

__device__ __forceinline__ v4s pack4(float a, float b, float c, float d)
{
    const v2f lo = { a, b }, hi = { c, d };
    const v2u r = { __builtin_bit_cast(unsigned, __builtin_convertvector(lo, v2bf)),
                    __builtin_bit_cast(unsigned, __builtin_convertvector(hi, v2bf)) };
    return __builtin_bit_cast(v4s, r);
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Fragment slots in LDS: forward A fragments of W1 (4 x 3), W2 (4 x 4), W3
// (2 x 4), then the transposed W3^T (4 x 2) and W2^T (4 x 4) for backward.
constexpr int kF1 = 0, kF2 = 12, kF3 = 28, kF3T = 36, kF2T = 44, kFrags = 60;

// The workgroup's copy of the network: F fragments (all kFrags for
// training, the forward ones, kF3T, for inference) and the biases (16-byte
// aligned: a lane reads four consecutive biases at once).
template <int F>
struct alignas(16) NetT {
    v4s frag[F][64];
    float b1[kHidden], b2[kHidden], b3[kOutPad];
};
using Net = NetT<kFrags>;
using ForwardNet = NetT<kF3T>;

// A fragment (tm, ts) of a rows x cols row-major matrix M (or of its
// transpose): lane l holds M[16 tm + l % 16][16 ts + 4 (l / 16) + j].
__device__ __forceinline__ v4s load_frag(const float *M, int rows, int cols, bool transpose, int tm, int ts, int lane)
{
    const int r = 16 * tm + (lane & 15);
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c = 16 * ts + 4 * (lane >> 4) + j;
        float x = 0.0f;
        if (!transpose && r < rows && c < cols) x = M[r * cols + c];
        if (transpose && c < rows && r < cols) x = M[c * cols + r];  // M^T[r][c] = M[c][r]
        v[j] = x;
    }
    return pack4(v[0], v[1], v[2], v[3]);
}

// Builds the workgroup's copy of the network from the FP32 parameters.
template <int F>
__device__ void load_net(NetT<F> &net, const float *__restrict__ p, int t, int nthreads)
{
    for (int i = t; i < F * 64; i += nthreads) {
        const int f = i / 64, lane = i % 64;
        v4s v;
        if (f < kF2) v = load_frag(p + kW1, kHidden, kIn, false, f / 3, f % 3, lane);
        else if (f < kF3) v = load_frag(p + kW2, kHidden, kHidden, false, (f - kF2) / 4, (f - kF2) % 4, lane);
        else if (f < kF3T) v = load_frag(p + kW3, kOut, kHidden, false, (f - kF3) / 4, (f - kF3) % 4, lane);
        else if (f < kF2T) v = load_frag(p + kW3, kOut, kHidden, true, (f - kF3T) / 2, (f - kF3T) % 2, lane);
        else v = load_frag(p + kW2, kHidden, kHidden, true, (f - kF2T) / 4, (f - kF2T) % 4, lane);
        net.frag[f][lane] = v;
    }
    for (int i = t; i < kHidden; i += nthreads) {
        net.b1[i] = p[kB1 + i];
        net.b2[i] = p[kB2 + i];
    }
    for (int i = t; i < kOutPad; i += nthreads) net.b3[i] = i < kOut ? p[kB3 + i] : 0.0f;
}

// ReLU of four BF16 values: as 16-bit integers a negative BF16 value (sign
// bit set) is negative and a positive one keeps its order, so the packed
// integer max with 0 is the ReLU (v_pk_max_i16, two values per instruction).
__device__ __forceinline__ v4s relu_bf16(v4s x)
{
    const v4s zero = { 0, 0, 0, 0 };
    return __builtin_elementwise_max(x, zero);
}

// Forward pass for the 16 blocks whose input fragments are bx[0..2] (k-step
// s = colour channel s; lane l holds texels 4(l/16)..+3 of block l % 16).
// Each layer accumulates onto its bias (the matrix cores' C operand), and
// the ReLU acts on the BF16 rounding.  Returns the post-ReLU hidden
// activations as FP32 values of their BF16 roundings (what the next layer
// and the weight gradients consume) and the 32 output rows (25 used).
template <int F>
__device__ __forceinline__ void forward(const NetT<F> &net, const v4s bx[3], int lane, v4f h1[4], v4f h2[4], v4f y[2])
{
    const int g = lane >> 4;
    v4s hb[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        v4f acc = *reinterpret_cast<const v4f *>(&net.b1[16 * m + 4 * g]);
#pragma unroll
        for (int s = 0; s < 3; s++) acc = mfma(net.frag[kF1 + 3 * m + s][lane], bx[s], acc);
        hb[m] = relu_bf16(pack4(acc[0], acc[1], acc[2], acc[3]));
#pragma unroll
        for (int j = 0; j < 4; j++) h1[m][j] = bf16_float(uint16_t(hb[m][j]));
    }
    v4s hb2[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        v4f acc = *reinterpret_cast<const v4f *>(&net.b2[16 * m + 4 * g]);
#pragma unroll
        for (int s = 0; s < 4; s++) acc = mfma(net.frag[kF2 + 4 * m + s][lane], hb[s], acc);
        hb2[m] = relu_bf16(pack4(acc[0], acc[1], acc[2], acc[3]));
#pragma unroll
        for (int j = 0; j < 4; j++) h2[m][j] = bf16_float(uint16_t(hb2[m][j]));
    }
#pragma unroll
    for (int m = 0; m < 2; m++) {
        v4f acc = *reinterpret_cast<const v4f *>(&net.b3[16 * m + 4 * g]);
#pragma unroll
        for (int s = 0; s < 4; s++) acc = mfma(net.frag[kF3 + 4 * m + s][lane], hb2[s], acc);
        y[m] = acc;
    }
}

// Block features: the 16 texels' R, G, B (outside the image: the mean),
// minus the mean of the texels inside, times 1 / max(largest deviation,
// kMinSpread).
// aux = (mean R, G, B, scale).
__device__ __forceinline__ void block_features(const float *src, int w, int h, int bx, int by, float f[kIn], float4 &aux)
{
    const size_t plane = size_t(w) * size_t(h);
    float c[3][16];
    bool in[16];
    float sum[3] = { 0.0f, 0.0f, 0.0f };
    int cnt = 0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const int x = 4 * bx + (t & 3), y = 4 * by + (t >> 2);
        in[t] = x < w && y < h;
        const size_t o = in[t] ? size_t(y) * size_t(w) + size_t(x) : 0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            c[k][t] = in[t] ? src[o + k * plane] : 0.0f;
            sum[k] += c[k][t];
        }
        cnt += in[t] ? 1 : 0;
    }
    float m[3], d = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; k++) m[k] = sum[k] / float(cnt);
    // All texels inside equal (texel 0 always is inside): exactly zero
    // features, whatever the rounding of the mean.
    bool uni = true;
#pragma unroll
    for (int t = 1; t < 16; t++)
        if (in[t] && (c[0][t] != c[0][0] || c[1][t] != c[1][0] || c[2][t] != c[2][0])) uni = false;
    if (uni) {
#pragma unroll
        for (int k = 0; k < 3; k++) m[k] = c[k][0];
    }
#pragma unroll
    for (int t = 0; t < 16; t++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (in[t]) d = fmaxf(d, fabsf(c[k][t] - m[k]));
    const float s = 1.0f / fmaxf(d, kMinSpread);
#pragma unroll
    for (int t = 0; t < 16; t++)
#pragma unroll
        for (int k = 0; k < 3; k++) f[k * 16 + t] = in[t] ? (c[k][t] - m[k]) * s : 0.0f;
    aux = make_float4(m[0], m[1], m[2], s);
}

__global__ __launch_bounds__(256) void mlp_features_kernel(const float *__restrict__ src, int w, int h, int bw, int n,
                                                           uint16_t *__restrict__ feat, float4 *__restrict__ aux)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    float f[kIn];
    float4 a;
    block_features(src, w, h, b % bw, b / bw, f, a);
#pragma unroll
    for (int i = 0; i < kIn; i++) feat[size_t(b) * kIn + i] = bf16_bits(f[i]);
    aux[b] = a;
}

// D3D10 BC1 palette of (c0, c1) as floats in [0, 1].
__device__ __forceinline__ void bc1_palette(uint32_t c0, uint32_t c1, float p[4][3])
{
    int q[4][3];
    ntx_bc1::expand565(c0, q[0][0], q[0][1], q[0][2]);
    ntx_bc1::expand565(c1, q[1][0], q[1][1], q[1][2]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (c0 > c1) {
            q[2][k] = (2 * q[0][k] + q[1][k]) / 3;
            q[3][k] = (2 * q[1][k] + q[0][k]) / 3;
        } else {
            q[2][k] = (q[0][k] + q[1][k]) / 2;
            q[3][k] = 0;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) p[i][k] = float(q[i][k]) / 255.0f;
}

// Training label of a block from its cluster-fit BC1 block: both endpoints
// and the decoded RMSE over the texels inside the image, in the block frame.
__global__ __launch_bounds__(256) void mlp_labels_kernel(const float *__restrict__ src, int w, int h, int bw, int n,
                                                         const uint2 *__restrict__ blocks, const float4 *__restrict__ aux,
                                                         float *__restrict__ labels)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const uint2 blk = blocks[b];
    const uint32_t c0 = blk.x & 0xffff, c1 = blk.x >> 16;
    float p[4][3];
    bc1_palette(c0, c1, p);
    const float4 a = aux[b];
    const int bx = b % bw, by = b / bw;
    const size_t plane = size_t(w) * size_t(h);
    float se = 0.0f;
    int cnt = 0;
    for (int t = 0; t < 16; t++) {
        const int x = 4 * bx + (t & 3), y = 4 * by + (t >> 2);
        if (x >= w || y >= h) continue;
        const size_t o = size_t(y) * size_t(w) + size_t(x);
        const int idx = (blk.y >> (2 * t)) & 3;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float d = p[idx][k] - src[o + k * plane];
            se += d * d;
        }
        cnt++;
    }
    float *L = labels + size_t(b) * kLabel;
    const float m[3] = { a.x, a.y, a.z };
#pragma unroll
    for (int k = 0; k < 3; k++) {
        L[k] = (p[0][k] - m[k]) * a.w;
        L[3 + k] = (p[1][k] - m[k]) * a.w;
    }
    L[6] = sqrtf(se / float(3 * cnt)) * a.w;
    L[7] = 0.0f;
}

// Candidate k's starting pair before the network's offsets: diagonal k of
// the block's colour bounding box (k = 0: every channel low -> high; k =
// 1..3: channel k - 1 runs high -> low against the others), given the box
// corners lo, hi.
__device__ __forceinline__ void box_diagonal(int k, const float lo[3], const float hi[3], float a[3], float b[3])
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const bool flip = k == c + 1;
        a[c] = flip ? hi[c] : lo[c];
        b[c] = flip ? lo[c] : hi[c];
    }
}

// Loss of one block and its gradient with respect to the 25 outputs:
// candidate k's endpoints are diagonal k of the block's bounding box (in
// the feature frame, corners lo, hi) plus outputs 6k..6k+5; the loss is the
// min over candidates of the unordered endpoint-pair error, plus a kSpread
// share of every candidate's error, plus the squared RMSE-head error.
__device__ __forceinline__ float block_loss(const float *y, const float *L, const float lo[3], const float hi[3],
                                            float *dy)
{
    float lk[kCandidates];
    bool swp[kCandidates];
    float ya[kCandidates][6];
#pragma unroll
    for (int k = 0; k < kCandidates; k++) {
        float ba[3], bb[3];
        box_diagonal(k, lo, hi, ba, bb);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            ya[k][i] = ba[i] + y[6 * k + i];
            ya[k][3 + i] = bb[i] + y[6 * k + 3 + i];
        }
        float ds = 0.0f, dw = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float a = ya[k][i], b = ya[k][3 + i];
            ds += (a - L[i]) * (a - L[i]) + (b - L[3 + i]) * (b - L[3 + i]);
            dw += (a - L[3 + i]) * (a - L[3 + i]) + (b - L[i]) * (b - L[i]);
        }
        swp[k] = dw < ds;
        lk[k] = swp[k] ? dw : ds;
    }
    int kb = 0;
#pragma unroll
    for (int k = 1; k < kCandidates; k++)
        if (lk[k] < lk[kb]) kb = k;
    float loss = lk[kb];
#pragma unroll
    for (int k = 0; k < kCandidates; k++) {
        loss += kSpread * lk[k] / kCandidates;
        const float coef = (k == kb ? 1.0f : 0.0f) + kSpread / kCandidates;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float ta = swp[k] ? L[3 + i] : L[i], tb = swp[k] ? L[i] : L[3 + i];
            dy[6 * k + i] = coef * 2.0f * (ya[k][i] - ta);
            dy[6 * k + 3 + i] = coef * 2.0f * (ya[k][3 + i] - tb);
        }
    }
    const float dr = y[24] - L[6];
    loss += kRmseWeight * dr * dr;
    dy[24] = kRmseWeight * 2.0f * dr;
#pragma unroll
    for (int i = kOut; i < kOutPad; i++) dy[i] = 0.0f;
    return loss;
}

// Input fragments of 16 blocks from the feature array: k-step s = channel s,
// lane l: texels 4(l/16)..+3 of block base + l % 16 (zero past n).
__device__ __forceinline__ void input_frags(const uint16_t *__restrict__ feat, int base, int n, int lane, v4s bx[3])
{
    const int b = base + (lane & 15), g = lane >> 4;
#pragma unroll
    for (int s = 0; s < 3; s++) {
        v4s v = { 0, 0, 0, 0 };
        if (b < n) {
            const uint16_t *f = feat + size_t(b) * kIn + 16 * s + 4 * g;
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = short(f[j]);
        }
        bx[s] = v;
    }
}

__global__ __launch_bounds__(64) void mlp_forward_kernel(const float *__restrict__ params, const uint16_t *__restrict__ feat,
                                                         int n, float *__restrict__ out)
{
    __shared__ Net net;
    const int lane = threadIdx.x;
    load_net(net, params, lane, 64);
    __syncthreads();
    const int base = blockIdx.x * 16;
    v4s bx[3];
    input_frags(feat, base, n, lane, bx);
    v4f h1[4], h2[4], y[2];
    forward(net, bx, lane, h1, h2, y);
    const int b = base + (lane & 15), g = lane >> 4;
    if (b < n)
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) out[size_t(b) * kOutPad + 16 * m + 4 * g + j] = y[m][j];
}

// One wave: forward, loss, backward over `tiles` tiles of 16 blocks from
// block tile0 * 16, accumulating this wave's weight gradients (summed over
// its blocks) into partial row `blockIdx.x`.  Deterministic: the partials
// are reduced in a fixed order by mlp_adam_kernel.
__global__ __launch_bounds__(64) void mlp_train_kernel(const float *__restrict__ params, const uint16_t *__restrict__ feat,
                                                       const float *__restrict__ labels, int n, int tiles,
                                                       float *__restrict__ partials, float *__restrict__ loss_out)
{
    __shared__ Net net;
    __shared__ float s_h1[16][kHidden + 1], s_h2[16][kHidden + 1], s_d1[16][kHidden + 1], s_d2[16][kHidden + 1];
    __shared__ float s_y[16][kOutPad + 1], s_dy[16][kOutPad + 1];
    const int lane = threadIdx.x, g = lane >> 4, col = lane & 15;
    load_net(net, params, lane, 64);
    __syncthreads();
    const v4f zero = { 0.0f, 0.0f, 0.0f, 0.0f };
    v4f gW1[4][3], gW2[4][4], gW3[2][4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
#pragma unroll
        for (int t = 0; t < 3; t++) gW1[m][t] = zero;
#pragma unroll
        for (int t = 0; t < 4; t++) gW2[m][t] = zero;
    }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int t = 0; t < 4; t++) gW3[m][t] = zero;
    float gb1[4][4] = {}, gb2[4][4] = {}, gb3[2][4] = {};
    float loss = 0.0f;

    for (int tile = 0; tile < tiles; tile++) {
        const int base = (blockIdx.x * tiles + tile) * 16;
        if (base >= n) break;
        const int b = base + col;
        const bool valid = b < n;
        v4s bx[3];
        input_frags(feat, base, n, lane, bx);
        v4f h1[4], h2[4], y[2];
        forward(net, bx, lane, h1, h2, y);
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                s_h1[col][16 * m + 4 * g + j] = valid ? h1[m][j] : 0.0f;
                s_h2[col][16 * m + 4 * g + j] = valid ? h2[m][j] : 0.0f;
            }
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) s_y[col][16 * m + 4 * g + j] = y[m][j];
        __syncthreads();
        if (lane < 16) {
            float dy[kOutPad];
            if (valid) {
                float yy[kOutPad];
#pragma unroll
                for (int i = 0; i < kOutPad; i++) yy[i] = s_y[lane][i];
                // Box corners in the feature frame (outside texels are 0, the
                // mean, which lies inside the box anyway).
                float lo[3], hi[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    lo[c] = FLT_MAX;
                    hi[c] = -FLT_MAX;
                    for (int t = 0; t < 16; t++) {
                        const float f = bf16_float(feat[size_t(b) * kIn + 16 * c + t]);
                        lo[c] = fminf(lo[c], f);
                        hi[c] = fmaxf(hi[c], f);
                    }
                }
                loss += block_loss(yy, labels + size_t(b) * kLabel, lo, hi, dy);
            } else {
#pragma unroll
                for (int i = 0; i < kOutPad; i++) dy[i] = 0.0f;
            }
            // The backward pass consumes dy as BF16.
#pragma unroll
            for (int i = 0; i < kOutPad; i++) s_dy[lane][i] = bf16_round(dy[i]);
        }
        __syncthreads();
        // delta2 = (W3^T dy) * relu'(h2); delta1 = (W2^T delta2) * relu'(h1).
        v4s bdy[2];
#pragma unroll
        for (int s = 0; s < 2; s++)
            bdy[s] = pack4(s_dy[col][16 * s + 4 * g], s_dy[col][16 * s + 4 * g + 1], s_dy[col][16 * s + 4 * g + 2],
                           s_dy[col][16 * s + 4 * g + 3]);
        v4s bd2[4];
        float d2[4][4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            v4f acc = zero;
#pragma unroll
            for (int s = 0; s < 2; s++) acc = mfma(net.frag[kF3T + 2 * m + s][lane], bdy[s], acc);
#pragma unroll
            for (int j = 0; j < 4; j++) d2[m][j] = bf16_round(h2[m][j] > 0.0f ? acc[j] : 0.0f);
            bd2[m] = pack4(d2[m][0], d2[m][1], d2[m][2], d2[m][3]);
        }
        float d1[4][4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            v4f acc = zero;
#pragma unroll
            for (int s = 0; s < 4; s++) acc = mfma(net.frag[kF2T + 4 * m + s][lane], bd2[s], acc);
#pragma unroll
            for (int j = 0; j < 4; j++) d1[m][j] = bf16_round(h1[m][j] > 0.0f ? acc[j] : 0.0f);
        }
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                s_d2[col][16 * m + 4 * g + j] = valid ? d2[m][j] : 0.0f;
                s_d1[col][16 * m + 4 * g + j] = valid ? d1[m][j] : 0.0f;
                gb2[m][j] += valid ? d2[m][j] : 0.0f;
                gb1[m][j] += valid ? d1[m][j] : 0.0f;
            }
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) gb3[m][j] += s_dy[col][16 * m + 4 * g + j];
        __syncthreads();
        // Weight gradients, k = the tile's 16 blocks: A = delta (row = output
        // index l % 16, k = block 4(l/16)+j), B = activation (k = block,
        // column = input index l % 16).
        v4s bh2[4], bh1[4], bxT[3];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            bh2[t] = pack4(s_h2[4 * g][16 * t + col], s_h2[4 * g + 1][16 * t + col], s_h2[4 * g + 2][16 * t + col],
                           s_h2[4 * g + 3][16 * t + col]);
            bh1[t] = pack4(s_h1[4 * g][16 * t + col], s_h1[4 * g + 1][16 * t + col], s_h1[4 * g + 2][16 * t + col],
                           s_h1[4 * g + 3][16 * t + col]);
        }
#pragma unroll
        for (int t = 0; t < 3; t++) {
            v4s v;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int bb = base + 4 * g + j;
                v[j] = bb < n ? short(feat[size_t(bb) * kIn + 16 * t + col]) : short(0);
            }
            bxT[t] = v;
        }
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const v4s a = pack4(s_dy[4 * g][16 * m + col], s_dy[4 * g + 1][16 * m + col], s_dy[4 * g + 2][16 * m + col],
                                s_dy[4 * g + 3][16 * m + col]);
#pragma unroll
            for (int t = 0; t < 4; t++) gW3[m][t] = mfma(a, bh2[t], gW3[m][t]);
        }
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const v4s a2 = pack4(s_d2[4 * g][16 * m + col], s_d2[4 * g + 1][16 * m + col], s_d2[4 * g + 2][16 * m + col],
                                 s_d2[4 * g + 3][16 * m + col]);
            const v4s a1 = pack4(s_d1[4 * g][16 * m + col], s_d1[4 * g + 1][16 * m + col], s_d1[4 * g + 2][16 * m + col],
                                 s_d1[4 * g + 3][16 * m + col]);
#pragma unroll
            for (int t = 0; t < 4; t++) gW2[m][t] = mfma(a2, bh1[t], gW2[m][t]);
#pragma unroll
            for (int t = 0; t < 3; t++) gW1[m][t] = mfma(a1, bxT[t], gW1[m][t]);
        }
        __syncthreads();
    }
    // Partial row: the parameter layout, D fragments (row 16m + 4(l/16) + j,
    // column 16t + l % 16).
    float *P = partials + size_t(blockIdx.x) * kParams;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = 16 * m + 4 * g + j;
#pragma unroll
            for (int t = 0; t < 3; t++) P[kW1 + r * kIn + 16 * t + col] = gW1[m][t][j];
#pragma unroll
            for (int t = 0; t < 4; t++) P[kW2 + r * kHidden + 16 * t + col] = gW2[m][t][j];
        }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = 16 * m + 4 * g + j;
            if (r < kOut)
#pragma unroll
                for (int t = 0; t < 4; t++) P[kW3 + r * kHidden + 16 * t + col] = gW3[m][t][j];
        }
    // Bias gradients: sums over the 16 columns (lanes l % 16) of each row.
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                gb1[m][j] += __shfl_xor(gb1[m][j], off);
                gb2[m][j] += __shfl_xor(gb2[m][j], off);
            }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) gb3[m][j] += __shfl_xor(gb3[m][j], off);
    if (col == 0) {
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                P[kB1 + 16 * m + 4 * g + j] = gb1[m][j];
                P[kB2 + 16 * m + 4 * g + j] = gb2[m][j];
            }
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (16 * m + 4 * g + j < kOut) P[kB3 + 16 * m + 4 * g + j] = gb3[m][j];
    }
    for (int off = 1; off < 64; off <<= 1) loss += __shfl_xor(loss, off);
    if (lane == 0) loss_out[blockIdx.x] = loss;
}

// Mean gradient over the n blocks (partials summed in row order) and, when
// lr > 0, one Adam step on the FP32 parameters.
__global__ __launch_bounds__(256) void mlp_adam_kernel(float *__restrict__ params, float *__restrict__ m1,
                                                       float *__restrict__ m2, const float *__restrict__ partials,
                                                       int rows, float inv_n, float lr, float beta1, float beta2,
                                                       float eps, float c1, float c2, float *__restrict__ grad_out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float gsum = 0.0f;
    for (int r = 0; r < rows; r++) gsum += partials[size_t(r) * kParams + p];
    const float gr = gsum * inv_n;
    if (grad_out) grad_out[p] = gr;
    if (lr > 0.0f) {
        const float a = beta1 * m1[p] + (1.0f - beta1) * gr;
        const float b = beta2 * m2[p] + (1.0f - beta2) * gr * gr;
        m1[p] = a;
        m2[p] = b;
        params[p] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
    }
}

// ------------------------------------------------------------ compression --

// Least-squares refinement of one endpoint pair for the block's texels held
// as 8-bit RGB in LDS, t8[t] = R | G << 8 | B << 16 (4-colour index by projection
// onto the segment, then the 2x2 normal equations), `iters` times; texel t
// takes part when it lies inside the image, (t & 3) < xe and (t >> 2) < ye.
// Endpoints stay in [0, 1].
__device__ __forceinline__ void refine8(const uint32_t *t8, int xe, int ye, float a[3], float b[3], int iters)
{
    for (int it = 0; it < iters; it++) {
        const float d[3] = { b[0] - a[0], b[1] - a[1], b[2] - a[2] };
        const float len2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const float il = len2 > 0.0f ? 1.0f / len2 : 0.0f;
        float saa = 0.0f, sbb = 0.0f, sab = 0.0f, sac[3] = { 0, 0, 0 }, sbc[3] = { 0, 0, 0 };
#pragma unroll 4
        for (int t = 0; t < 16; t++) {
            const float v[3] = { float(t8[t] & 255u) * (1.0f / 255.0f), float((t8[t] >> 8) & 255u) * (1.0f / 255.0f),
                                 float(t8[t] >> 16) * (1.0f / 255.0f) };
            const bool in = (t & 3) < xe && (t >> 2) < ye;
            const float u = ((v[0] - a[0]) * d[0] + (v[1] - a[1]) * d[1] + (v[2] - a[2]) * d[2]) * il;
            const float q = fminf(fmaxf(rintf(3.0f * u), 0.0f), 3.0f);
            const float be = in ? q * (1.0f / 3.0f) : 0.0f, al = in ? 1.0f - q * (1.0f / 3.0f) : 0.0f;
            saa += al * al; sbb += be * be; sab += al * be;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                sac[k] += al * v[k];
                sbc[k] += be * v[k];
            }
        }
        const float det = saa * sbb - sab * sab;
        if (!(det > 1e-8f)) break;
        const float id = 1.0f / det;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            a[k] = fminf(fmaxf((sbb * sac[k] - sab * sbc[k]) * id, 0.0f), 1.0f);
            b[k] = fminf(fmaxf((saa * sbc[k] - sab * sac[k]) * id, 0.0f), 1.0f);
        }
    }
}

// The block's four palette entries for 565 endpoints c0, c1 (four-colour
// order when c0 > c1, the caller orders them; the D3D10 integer palette) as
// 8-bit RGB pk[i] = R | G << 8 | B << 16, and the keys
// nk[i] = 3 - i - 4 |p_i|^2 of eval_pair8.
__device__ __forceinline__ void palette8(uint32_t c0, uint32_t c1, uint32_t pk[4], int nk[4])
{
    int e[2][3];
    ntx_bc1::expand565(c0, e[0][0], e[0][1], e[0][2]);
    ntx_bc1::expand565(c1, e[1][0], e[1][1], e[1][2]);
    uint32_t q[4][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t a = uint32_t(e[0][k]), b = uint32_t(e[1][k]);
        q[0][k] = a;
        q[1][k] = b;
        // x / 3 = (x * 43691) >> 17 for every x <= 765 (24-bit multiply)
        q[2][k] = c0 > c1 ? __umul24(2 * a + b, 43691u) >> 17 : (a + b) >> 1;
        q[3][k] = c0 > c1 ? __umul24(2 * b + a, 43691u) >> 17 : 0u;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        pk[i] = q[i][0] | q[i][1] << 8 | q[i][2] << 16;
        nk[i] = 3 - i - 4 * int(__builtin_amdgcn_udot4(pk[i], pk[i], 0u, false));
        // Opaque to the optimiser (no instruction): each key is then one
        // shift-add of the texel's dot product with this constant, where
        // entry 3's "0 - 4|p|^2" would otherwise be folded into a shift and
        // a subtract per texel.
        asm("" : "+v"(nk[i]));
    }
}

// Sum of |t|^2 over the block's texels, t8[t] = R | G << 8 | B << 16 (zero
// for a texel outside the image): the part of every candidate's error that
// does not depend on the endpoints.
__device__ __forceinline__ int texel_norms(const uint32_t t8[16])
{
    uint32_t v2 = 0;
#pragma unroll
    for (int t = 0; t < 16; t++) v2 = __builtin_amdgcn_udot4(t8[t], t8[t], v2, false);
    return int(v2);
}

// Nearest palette entry of each texel and the summed squared distance
// |p|^2 - 2 p.t over the texels inside the image; the n_out texels outside
// it hold t = 0.  Per texel and entry, e = 8 p.t + nk = -4 (|p|^2 - 2 p.t) +
// (3 - i): one v_dot4_u32_u8 and one shift-add; the largest e is the
// nearest entry and, among equally near ones, the lowest i (distinct
// distances differ by 4 in e, 3 - i by at most 3).  Its low two bits, 3 - i,
// are gathered into `inv`.  The sum runs over all 16 texels; a texel with
// t = 0 takes e0 = max nk, which is taken back out n_out times.
__device__ __forceinline__ int nearest8(const uint32_t t8[16], int n_out, const uint32_t pk[4], const int nk[4],
                                        uint32_t &inv)
{
    int sum = 0;
    inv = 0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const uint32_t v = t8[t];
        int e = int(__builtin_amdgcn_udot4(pk[0], v, 0u, false) << 3) + nk[0];
#pragma unroll
        for (int i = 1; i < 4; i++) e = max(e, int(__builtin_amdgcn_udot4(pk[i], v, 0u, false) << 3) + nk[i]);
        inv = __builtin_amdgcn_alignbit(uint32_t(e), inv, 2);  // (inv >> 2) | e << 30: texel 0 ends in bits 0-1
        sum += e;
    }
    // sum = -4 D + (sum of the 3 - i fields) over all 16 texels
    const int fields = __builtin_popcount(inv & 0x55555555u) + 2 * __builtin_popcount(inv & 0xaaaaaaaau);
    const int e0 = max(max(nk[0], nk[1]), max(nk[2], nk[3]));
    return (fields - sum + n_out * (e0 - (e0 & 3))) >> 2;
}

// Error (exact squared error in 8-bit units) and index bits of the block
// with 565 endpoints c0, c1, each texel taking its nearest palette entry
// (the lowest entry on ties), on the block's texels held as 8-bit RGB
// (each channel round(255 * saturate(v)); 0 for the n_out texels outside
// the image) with v2 = texel_norms(t8).  For 8-bit sources (every value
// k / 255) this is evaluate_mse's measure without rounding.
__device__ __forceinline__ float eval_pair8(const uint32_t t8[16], int n_out, int v2, uint32_t c0, uint32_t c1,
                                            uint32_t *bits_out)
{
    uint32_t pk[4];
    int nk[4];
    palette8(c0, c1, pk, nk);
    uint32_t inv;
    const int d = nearest8(t8, n_out, pk, nk, inv);
    if (bits_out) *bits_out = ~inv;
    return float(d + v2);
}

__device__ __forceinline__ float encode_pair8(const uint32_t t8[16], int n_out, int v2, const float a[3],
                                              const float b[3], const float *mid5, const float *mid6, uint2 *blk)
{
    uint32_t c0 = ntx_bc1::to565(a[0], a[1], a[2], mid5, mid6);
    uint32_t c1 = ntx_bc1::to565(b[0], b[1], b[2], mid5, mid6);
    if (c0 < c1) { const uint32_t t = c0; c0 = c1; c1 = t; }
    uint32_t bits;
    const float err = eval_pair8(t8, n_out, v2, c0, c1, &bits);
    blk->x = c0 | (c1 << 16);
    blk->y = bits;
    return err;
}

// R | G << 8 | B << 16 with each channel round(255 * saturate(v)) (round to
// nearest even): v_cvt_pk_u8_f32 rounds and saturates 255 v into its byte.
__device__ __forceinline__ uint32_t rgb_u8(float r, float g, float b)
{
    uint32_t p = __builtin_amdgcn_cvt_pk_u8_f32(r * 255.0f, 0, 0u);
    p = __builtin_amdgcn_cvt_pk_u8_f32(g * 255.0f, 1, p);
    return __builtin_amdgcn_cvt_pk_u8_f32(b * 255.0f, 2, p);
}

// Exchanges across the four 16-lane rows of a wavefront without LDS
// (v_permlane16_swap_b32 and v_permlane32_swap_b32 with both operands x):
// rows16(x) gives each lane x of the even row (lo) and of the odd row (hi)
// of its row pair {0, 1} or {2, 3}, same column; halves32(x) gives x of the
// lower (lo) and upper (hi) lane of its pair l, l ^ 32.
struct LanePair {
    uint32_t lo, hi;
};
__device__ __forceinline__ LanePair rows16(uint32_t x)
{
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return { uint32_t(r[0]), uint32_t(r[1]) };
}
__device__ __forceinline__ LanePair halves32(uint32_t x)
{
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return { uint32_t(r[0]), uint32_t(r[1]) };
}

// op over the four lanes l % 16 + 16 r (r = 0..3) of a column, in each of
// them (op commutative: the result is the same in all four).
template <typename Op>
__device__ __forceinline__ float column_reduce(float x, Op op)
{
    LanePair p = rows16(__float_as_uint(x));
    x = op(__uint_as_float(p.lo), __uint_as_float(p.hi));
    p = halves32(__float_as_uint(x));
    return op(__uint_as_float(p.lo), __uint_as_float(p.hi));
}
template <typename Op>
__device__ __forceinline__ int column_reduce_int(int x, Op op)
{
    LanePair p = rows16(uint32_t(x));
    x = op(int(p.lo), int(p.hi));
    p = halves32(uint32_t(x));
    return op(int(p.lo), int(p.hi));
}
// x of row 0 of the lane's column.
__device__ __forceinline__ float column_row0(float x)
{
    return __uint_as_float(halves32(rows16(__float_as_uint(x)).lo).lo);
}

// One wave per 16 blocks at a time, 4 waves per workgroup.  Lane (g, n) of a
// tile gathers row g of block n, the network runs on the matrix cores, and
// then lane (g, n) refines candidate g of block n; the four candidates of a
// block are compared across lanes.  A block whose predicted cluster-fit RMSE
// exceeds tau_pred, or whose best candidate is worse than that prediction by
// more than tau_gap (both in 8-bit units), is queued for the cluster fit.
// Lane g refines diagonal g of the block's colour bounding box both as it is
// and moved by the network's offsets, keeping the better; box_candidates
// drops the moved candidates (the same method without the network; no
// prediction, so only tau_pred < 0 sends blocks to the fallback).
__global__ __launch_bounds__(256, 5) void bc1_mlp_kernel(const float *__restrict__ params, const float *__restrict__ src,
                                                      int w, int h, int bw, int n, const ntx::Tables *__restrict__ tab,
                                                      float tau_pred, float tau_gap, int refine_iters, int greedy_rounds,
                                                      int box_candidates,
                                                      uint2 *__restrict__ out, int *__restrict__ fb_count,
                                                      int *__restrict__ fb_list, float *__restrict__ err_out)
{
    __shared__ ForwardNet net;
    // Each wave's 16 blocks: the texels as 8-bit RGB and the network
    // outputs.  Only the wave that owns them touches them after the network
    // is loaded.
    __shared__ __align__(16) uint32_t s_t8[4][16][16];
    __shared__ float s_y[4][16][kOutPad + 1];
    // The 5:6:5 quantisation midpoints, read per endpoint channel.
    __shared__ float s_mid5[32], s_mid6[64];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, g = lane >> 4, col = lane & 15;
    load_net(net, params, t, 256);
    if (t < 32) s_mid5[t] = tab->mid5[t];
    if (t < 64) s_mid6[t] = tab->mid6[t];
    __syncthreads();
    const size_t plane = size_t(w) * size_t(h);
    // Rows of 4 texels start on 16-byte boundaries in every plane.
    const bool aligned = (w & 3) == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0;
    const int tiles = (n + 15) / 16;
    // A wave past the last tile computes on no valid block.
    for (int group = blockIdx.x; group * 4 < tiles; group += gridDim.x) {
        const int tile = group * 4 + wave;
        const int b = tile * 16 + col;
        const bool valid = b < n;
        const int bx = valid ? b % bw : 0, by = valid ? b / bw : 0;
        // Row g of block b, and the block mean / scale over the texels inside.
        float c[3][4];
        bool inr[4];
        const int y = 4 * by + g;
        if (aligned && valid && 4 * bx + 3 < w && y < h) {
            // The whole row inside the image: one 16-byte load per plane.
            const float4 *row = reinterpret_cast<const float4 *>(src + size_t(y) * size_t(w) + size_t(4 * bx));
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float4 v = row[k * (plane >> 2)];
                c[k][0] = v.x; c[k][1] = v.y; c[k][2] = v.z; c[k][3] = v.w;
            }
#pragma unroll
            for (int j = 0; j < 4; j++) inr[j] = true;
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int x = 4 * bx + j;
                inr[j] = valid && x < w && y < h;
                const size_t o = inr[j] ? size_t(y) * size_t(w) + size_t(x) : 0;
#pragma unroll
                for (int k = 0; k < 3; k++) c[k][j] = inr[j] ? src[o + k * plane] : 0.0f;
            }
        }
        float sum[3] = { 0, 0, 0 };
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int k = 0; k < 3; k++) sum[k] += c[k][j];
            cnt += inr[j] ? 1 : 0;
        }
        const auto add = [](auto a, auto b) { return a + b; };
#pragma unroll
        for (int k = 0; k < 3; k++) sum[k] = column_reduce(sum[k], add);
        cnt = column_reduce_int(cnt, add);
        float m[3];
        if (cnt == 16) {
            // sum / 16, exactly
#pragma unroll
            for (int k = 0; k < 3; k++) m[k] = sum[k] * 0.0625f;
        } else {
#pragma unroll
            for (int k = 0; k < 3; k++) m[k] = cnt > 0 ? sum[k] / float(cnt) : 0.0f;
        }
        float d = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (inr[j]) d = fmaxf(d, fabsf(c[k][j] - m[k]));
        d = column_reduce(d, [](float a, float b) { return fmaxf(a, b); });
        const float s = 1.0f / fmaxf(d, kMinSpread);
        v4s bx3[3];
#pragma unroll
        for (int k = 0; k < 3; k++)
            bx3[k] = pack4(inr[0] ? (c[k][0] - m[k]) * s : 0.0f, inr[1] ? (c[k][1] - m[k]) * s : 0.0f,
                           inr[2] ? (c[k][2] - m[k]) * s : 0.0f, inr[3] ? (c[k][3] - m[k]) * s : 0.0f);
        // Block-level facts from the four lanes that hold its rows: the
        // colour bounding box of the texels inside the image, whether they
        // all equal texel 0 (row 0's lane, lane col), and the texels as 8-bit
        // RGB for the candidate errors.
        float lo[3] = { 1.0f, 1.0f, 1.0f }, hi[3] = { 0.0f, 0.0f, 0.0f };
        const float t0[3] = { column_row0(c[0][0]), column_row0(c[1][0]), column_row0(c[2][0]) };
        int same = 1;
        uint32_t p8[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (inr[j]) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    lo[k] = fminf(lo[k], c[k][j]);
                    hi[k] = fmaxf(hi[k], c[k][j]);
                }
                if ((g | j) != 0 && (c[0][j] != t0[0] || c[1][j] != t0[1] || c[2][j] != t0[2])) same = 0;
            }
            p8[j] = rgb_u8(c[0][j], c[1][j], c[2][j]);  // 0 outside (c is 0 there)
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = column_reduce(lo[k], [](float a, float b) { return fminf(a, b); });
            hi[k] = column_reduce(hi[k], [](float a, float b) { return fmaxf(a, b); });
        }
        same = column_reduce_int(same, [](int a, int b) { return a & b; });
        const bool uniform = valid && same;

        v4f h1[4], h2[4], yv[2];
        forward(net, bx3, lane, h1, h2, yv);
        *reinterpret_cast<uint4 *>(&s_t8[wave][col][4 * g]) = make_uint4(p8[0], p8[1], p8[2], p8[3]);
#pragma unroll
        for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int mm = 0; mm < 2; mm++) s_y[wave][col][16 * mm + 4 * g + j] = yv[mm][j];
        }
        ntx_wave_sync();
        // Lane (g, col): candidate g of block col, on the block's 16 texels
        // in registers.
        uint32_t t8[16];
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++) {
            const uint4 v = *reinterpret_cast<const uint4 *>(&s_t8[wave][col][4 * q4]);
            t8[4 * q4] = v.x; t8[4 * q4 + 1] = v.y; t8[4 * q4 + 2] = v.z; t8[4 * q4 + 3] = v.w;
        }
        const int v2 = texel_norms(t8), n_out = 16 - cnt;
        const int xe = min(4, w - 4 * bx), ye = min(4, h - 4 * by);
        const float is = 1.0f / s;
        // Candidate g: diagonal g of the block's colour bounding box, plus
        // the network's offsets (in the block frame, hence divided by the
        // scale) unless box_candidates.
        float a[3], bb[3];
        box_diagonal(g, lo, hi, a, bb);
        float pred = fmaxf(s_y[wave][col][24], 0.0f) * is * 255.0f;  // predicted RMSE, 8-bit units
        // The diagonal as it is, refined and encoded ...
        if (refine_iters > 0) refine8(s_t8[wave][col], xe, ye, a, bb, refine_iters);
        uint2 blk;
        float err = valid ? encode_pair8(t8, n_out, v2, a, bb, s_mid5, s_mid6, &blk) : FLT_MAX;
        if (box_candidates) {
            pred = tau_pred < 0.0f ? FLT_MAX : 0.0f;
        } else {
            // ... and moved by the network's offsets; the better one is kept
            // (the diagonal on equal error), so the network can only help.
            float na[3], nb[3];
            box_diagonal(g, lo, hi, na, nb);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                na[k] = fminf(fmaxf(na[k] + s_y[wave][col][6 * g + k] * is, 0.0f), 1.0f);
                nb[k] = fminf(fmaxf(nb[k] + s_y[wave][col][6 * g + 3 + k] * is, 0.0f), 1.0f);
            }
            if (refine_iters > 0) refine8(s_t8[wave][col], xe, ye, na, nb, refine_iters);
            uint2 nblk;
            const float nerr = valid ? encode_pair8(t8, n_out, v2, na, nb, s_mid5, s_mid6, &nblk) : FLT_MAX;
            if (nerr < err) { err = nerr; blk = nblk; }
        }
        // Best of the four candidates (lower candidate on equal error): of
        // each exchanged pair, lo holds the lower candidates.
        {
            LanePair e = rows16(__float_as_uint(err)), px = rows16(blk.x), py = rows16(blk.y);
            bool up = __uint_as_float(e.hi) < __uint_as_float(e.lo);
            err = __uint_as_float(up ? e.hi : e.lo);
            blk = make_uint2(up ? px.hi : px.lo, up ? py.hi : py.lo);
            e = halves32(__float_as_uint(err)); px = halves32(blk.x); py = halves32(blk.y);
            up = __uint_as_float(e.hi) < __uint_as_float(e.lo);
            err = __uint_as_float(up ? e.hi : e.lo);
            blk = make_uint2(up ? px.hi : px.lo, up ? py.hi : py.lo);
        }
        // A block whose texels inside the image are all equal gets what the
        // cluster fit gives it: reduce_colors' running weighted mean of the
        // equal colours (icbc.h:1638-1683, weights 1, texel order) and the
        // single-colour tables.  Texel 0 is always inside the image.
        // The block's lane g runs the mean of channel g (lane 3 repeats
        // channel 2): one division per step instead of three; a whole block
        // of an 8-bit value takes the tabulated mean.
        if (uniform) {
            const float tc = g == 0 ? t0[0] : (g == 1 ? t0[1] : t0[2]);
            const int k8 = int(rintf(fminf(fmaxf(tc, 0.0f), 1.0f) * 255.0f));
            float kc = tc;
            if (cnt == 16 && tab->div255[k8] == tc) {
                kc = tab->mean16[k8];
            } else {
                float kw = 1.0f;
                for (int i = 1; i < cnt; i++) {
                    const float s1 = kw + 1.0f;
                    kc = (kc * kw + tc) / s1;
                    kw = s1;
                }
            }
            const LanePair a = rows16(__float_as_uint(kc));  // lo: channel 0 / 2, hi: channel 1 / 2
            const LanePair rb = halves32(a.lo), gg = halves32(a.hi);
            blk = ntx_bc1::single_color_block(tab, __uint_as_float(rb.lo), __uint_as_float(gg.lo),
                                              __uint_as_float(rb.hi));
        }
        // Greedy search on the quantised endpoints (+-1 in one 565 channel of
        // one endpoint; the 12 moves spread 3 per lane over the block's lanes),
        // taking the best improving move each round, lowest move on ties.
        // q[3e + c]: channel c (5, 6, 5 bits) of endpoint e; move mv changes
        // q[mv >> 1] by -1 (mv even) or +1 (mv odd).
        int q[6];
        {
            const uint32_t e0 = blk.x & 0xffffu, e1 = blk.x >> 16;
            q[0] = int(e0 >> 11); q[1] = int((e0 >> 5) & 63u); q[2] = int(e0 & 31u);
            q[3] = int(e1 >> 11); q[4] = int((e1 >> 5) & 63u); q[5] = int(e1 & 31u);
        }
        // Each move's evaluation keeps its index bits, which travel with the
        // winning (error, move), so the final block needs no re-evaluation.
        const int rounds = uniform ? 0 : greedy_rounds;  // uniform across the block's lanes
        uint32_t bits = blk.y;
        for (int round = 0; round < rounds; round++) {
            float be = err;
            int bm = 64;
            uint32_t bb = 0;
#pragma unroll 1
            for (int i = 0; i < 3; i++) {
                const int mv = 3 * g + i, slot = mv >> 1, step = (mv & 1) ? 1 : -1;
                int t2[6];
                bool ok = true;
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    t2[k] = q[k] + (k == slot ? step : 0);
                    ok = ok && t2[k] >= 0 && t2[k] <= ((k % 3) == 1 ? 63 : 31);
                }
                if (!ok) continue;
                uint32_t c0 = uint32_t(t2[0]) << 11 | uint32_t(t2[1]) << 5 | uint32_t(t2[2]);
                uint32_t c1 = uint32_t(t2[3]) << 11 | uint32_t(t2[4]) << 5 | uint32_t(t2[5]);
                if (c0 < c1) { const uint32_t tt = c0; c0 = c1; c1 = tt; }
                uint32_t mbits = 0;
                const float e = valid ? eval_pair8(t8, n_out, v2, c0, c1, &mbits) : FLT_MAX;
                if (e < be) { be = e; bm = mv; bb = mbits; }
            }
            // The lowest (error, move) over the block's four lanes, with its bits.
            {
                LanePair e = rows16(__float_as_uint(be)), m = rows16(uint32_t(bm)), pb = rows16(bb);
                float el = __uint_as_float(e.lo), eh = __uint_as_float(e.hi);
                bool up = eh < el || (eh == el && int(m.hi) < int(m.lo));
                be = up ? eh : el;
                bm = int(up ? m.hi : m.lo);
                bb = up ? pb.hi : pb.lo;
                e = halves32(__float_as_uint(be)); m = halves32(uint32_t(bm)); pb = halves32(bb);
                el = __uint_as_float(e.lo); eh = __uint_as_float(e.hi);
                up = eh < el || (eh == el && int(m.hi) < int(m.lo));
                be = up ? eh : el;
                bm = int(up ? m.hi : m.lo);
                bb = up ? pb.hi : pb.lo;
            }
            if (bm < 64) {
#pragma unroll
                for (int k = 0; k < 6; k++) q[k] += k == (bm >> 1) ? ((bm & 1) ? 1 : -1) : 0;
                err = be;
                bits = bb;
            }
        }
        if (rounds > 0) {
            uint32_t c0 = uint32_t(q[0]) << 11 | uint32_t(q[1]) << 5 | uint32_t(q[2]);
            uint32_t c1 = uint32_t(q[3]) << 11 | uint32_t(q[4]) << 5 | uint32_t(q[5]);
            if (c0 < c1) { const uint32_t tt = c0; c0 = c1; c1 = tt; }
            blk.x = c0 | (c1 << 16);
            blk.y = bits;
        }
        if (g == 0 && valid) {
            // The block's RMSE only where the gap test or err_out needs it
            // (both launch arguments: the branch is taken by whole waves).
            const bool need_rmse = tau_gap >= 0.0f || err_out != nullptr;
            const float rmse = need_rmse ? sqrtf(err / float(3 * cnt)) : 0.0f;
            const bool fallback = !uniform && (pred > tau_pred || (tau_gap >= 0.0f && rmse > pred + tau_gap));
            out[b] = blk;
            if (err_out) err_out[b] = rmse;
            if (fallback) fb_list[atomicAdd(fb_count, 1)] = b;
        }
        ntx_wave_sync();
    }
}

}  // namespace

namespace ntx {

hipError_t launch_mlp_features(const float *d_src, int w, int h, void *d_feat, float *d_aux, hipStream_t stream)
{
    const int bw = (w + 3) / 4, n = bw * ((h + 3) / 4);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_features_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d_src, w, h, bw, n,
                       static_cast<uint16_t *>(d_feat), reinterpret_cast<float4 *>(d_aux));
    return hipGetLastError();
}

hipError_t launch_mlp_labels(const float *d_src, int w, int h, const void *d_bc1, const float *d_aux, float *d_labels,
                             hipStream_t stream)
{
    const int bw = (w + 3) / 4, n = bw * ((h + 3) / 4);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_labels_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d_src, w, h, bw, n,
                       static_cast<const uint2 *>(d_bc1), reinterpret_cast<const float4 *>(d_aux), d_labels);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const float *d_params, const void *d_feat, int n, float *d_out, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 15) / 16), dim3(64), 0, stream, d_params,
                       static_cast<const uint16_t *>(d_feat), n, d_out);
    return hipGetLastError();
}

// Rows of partial gradients a training step over n blocks writes.
int mlp_train_rows(int n, int tiles_per_wave) { return (((n + 15) / 16) + tiles_per_wave - 1) / tiles_per_wave; }

hipError_t launch_mlp_train(const float *d_params, const void *d_feat, const float *d_labels, int n, int tiles_per_wave,
                            float *d_partials, float *d_loss, hipStream_t stream)
{
    const int rows = mlp_train_rows(n, tiles_per_wave);
    hipLaunchKernelGGL(mlp_train_kernel, dim3(rows), dim3(64), 0, stream, d_params, static_cast<const uint16_t *>(d_feat),
                       d_labels, n, tiles_per_wave, d_partials, d_loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *d_params, float *d_m1, float *d_m2, const float *d_partials, int rows, int n, float lr,
                           int step, float *d_grad, hipStream_t stream)
{
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f / (1.0f - powf(beta1, float(step))), c2 = 1.0f / (1.0f - powf(beta2, float(step)));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((ntx_mlp::kParams + 255) / 256), dim3(256), 0, stream, d_params, d_m1, d_m2,
                       d_partials, rows, 1.0f / float(n), lr, beta1, beta2, eps, c1, c2, d_grad);
    return hipGetLastError();
}

hipError_t launch_bc1_mlp(const float *d_params, const float *d_src, int w, int h, const Tables *d_tab, float tau_pred,
                          float tau_gap, int refine_iters, int greedy_rounds, int box_candidates, void *d_out,
                          int *d_fb_count, int *d_fb_list, float *d_err, int cus, hipStream_t stream)
{
    const int bw = (w + 3) / 4, n = bw * ((h + 3) / 4);
    if (n == 0) return hipSuccess;
    const int tiles = (n + 15) / 16;
    // The persistent grid: as many workgroups as stay resident on each
    // compute unit (the kernel's LDS and registers decide), times the CUs.
    static int per_cu = 0;
    if (per_cu <= 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bc1_mlp_kernel, 256, 0) != hipSuccess || per_cu <= 0))
        per_cu = 2;
    const int grid = std::max(1, std::min(per_cu * cus, (tiles + 3) / 4));
    hipLaunchKernelGGL(bc1_mlp_kernel, dim3(grid), dim3(256), 0, stream, d_params, d_src, w, h, bw, n, d_tab, tau_pred,
                       tau_gap, refine_iters, greedy_rounds, box_candidates, static_cast<uint2 *>(d_out), d_fb_count,
                       d_fb_list, d_err);
    return hipGetLastError();
}

}  // namespace ntx

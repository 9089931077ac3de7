// mip.hip -- mipmap downsampling for gfx950, bit-exact with NVTT's
// single-threaded FloatImage filters (same tap order, no fused multiply-add):
//   Box, width 0.5: FloatImage::fastDownSample (FloatImage.cpp:559-737)
//   otherwise: FloatImage::resize (FloatImage.cpp:761-808) = applyKernelX into
//   a w x H temporary, then applyKernelY (FloatImage.cpp:1115-1176), with the
//   host-built PolyphaseKernel weights (Filter.cpp:563-608).
// Planar FP32 RGBA, four planes.  Every kernel is a coalesced elementwise pass
// (one output texel per lane, consecutive lanes on consecutive x).
#include "mip_taps.hpp"

#include <cmath>

namespace {

using ntx_mip::wrap_index;

__global__ __launch_bounds__(256) void box_kernel(const float *__restrict__ src, int W, int H, float *__restrict__ dst,
                                                  int w, int h)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int c = blockIdx.z;
    if (x >= w) return;
    const float *s = src + size_t(c) * size_t(W) * size_t(H);
    float *d = dst + size_t(c) * size_t(w) * size_t(h);
    if (W == 1 || H == 1) {
        // 1D case: the w outputs (h == 1 here) run along the longer axis.
        d[x] = ntx_mip::box_texel(s, W, H, x, 0);
        return;
    }
    d[size_t(y) * size_t(w) + x] = ntx_mip::box_texel(s, W, H, x, y);
}

// applyKernelX: tmp(i, y) = sum_j k(i, j) * src(wrap(left_i + j), y).
__global__ __launch_bounds__(256) void poly_x_kernel(const float *__restrict__ src, int W, int H, float *__restrict__ tmp,
                                                     int w, const float *__restrict__ kw, const int *__restrict__ left,
                                                     int window, int wrap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int c = blockIdx.z;
    if (i >= w) return;
    const float *row = src + (size_t(c) * size_t(H) + size_t(y)) * size_t(W);
    const float *k = kw + size_t(i) * window;
    const int l = left[i];
    float sum = 0.0f;
    if (l >= 0 && l + window <= W) {
        for (int j = 0; j < window; ++j) sum += k[j] * row[l + j];
    } else {
        for (int j = 0; j < window; ++j) sum += k[j] * row[wrap_index(l + j, W, wrap)];
    }
    tmp[(size_t(c) * size_t(H) + size_t(y)) * size_t(w) + i] = sum;
}

// applyKernelY on the temporary: dst(x, i) = sum_j k(i, j) * tmp(x, wrap(left_i + j)).
__global__ __launch_bounds__(256) void poly_y_kernel(const float *__restrict__ tmp, int w, int H, float *__restrict__ dst,
                                                     int h, const float *__restrict__ kw, const int *__restrict__ left,
                                                     int window, int wrap)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int c = blockIdx.z;
    if (x >= w) return;
    const float *col = tmp + size_t(c) * size_t(w) * size_t(H) + x;
    const float *k = kw + size_t(i) * window;
    const int l = left[i];
    float sum = 0.0f;
    if (l >= 0 && l + window <= H) {
        for (int j = 0; j < window; ++j) sum += k[j] * col[size_t(l + j) * size_t(w)];
    } else {
        for (int j = 0; j < window; ++j) sum += k[j] * col[size_t(wrap_index(l + j, H, wrap)) * size_t(w)];
    }
    dst[(size_t(c) * size_t(h) + size_t(i)) * size_t(w) + x] = sum;
}

// Fused separable resize: one workgroup per kTX x kTY output tile.  The
// source rows/columns the tile's taps touch are staged in LDS (with the wrap
// applied while loading), the X taps fill an LDS temporary for every source
// row the Y taps need, then the Y taps produce the outputs.  Each temporary
// value is exactly the reference's applyKernelX result for that (wrapped)
// row, and each output sums its Y taps in the reference order, so the result
// is bit-identical to the two-pass form while the w x H temporary never
// touches HBM.
constexpr int kTX = 64, kTY = 16;
constexpr int kSrcW = 160, kSrcH = 56;  // largest staged source tile
constexpr int kMaxWin = 16;             // largest window held in registers
constexpr int kWaves = 8;               // wavefronts per workgroup

// Output tile of workgroup `tile` among tx x ty tiles.  Workgroups are
// dispatched round-robin over the eight XCDs (each with its own L2), so when
// tx divides into eight strips XCD k takes strip k, walked row by row: the
// source rows shared by vertically adjacent tiles are then read by the same
// L2 at about the same time instead of twice from HBM.
__device__ __forceinline__ void tile_coords(int tile, int tx, int &tX, int &tY)
{
    if ((tx & 7) == 0) {
        const int strip = tx >> 3, local = tile >> 3;
        tX = (tile & 7) * strip + local % strip;
        tY = local / strip;
    } else {
        tX = tile % tx;
        tY = tile / tx;
    }
}

// The taps of exact halving with the same weights for every output, passed
// by value (they arrive in scalar registers): weights kx, ky of output 0.
struct SlideTaps {
    float kx[kMaxWin], ky[kMaxWin];
};

// kPair: every output's first tap is two source columns after its left
// neighbour's (downsampling by exactly 2), so lane xi's taps start at column
// 2*xi of the staged row and are read two at a time with 8-byte LDS reads
// (consecutive lanes on consecutive 8-byte words: no bank conflicts).
// kWin > 0: both passes have exactly kWin taps, known when compiling, so the
// tap loops unroll without a test per tap; 0: wx and wy taps as passed.
// kUni: exact halving with the same taps for every output (kPair, kWin > 0):
// the weights come from `taps` and output i's first tap is l0 + 2 i, so no
// table is read before the staging loads (small levels are latency-bound).
template <bool kPair, int kWin, bool kUni = false>
__global__ __launch_bounds__(64 * kWaves) void poly_fused_kernel(const float *__restrict__ src, int W, int H,
                                                         float *__restrict__ dst, int w, int h, int tx,
                                                         const float *__restrict__ kx, const int *__restrict__ lx, int wx,
                                                         const float *__restrict__ ky, const int *__restrict__ ly, int wy,
                                                         int wrap, SlideTaps taps, int lx0, int ly0)
{
    static_assert(2 * (kTX - 1) + kMaxWin <= kSrcW, "paired X reads stay inside a staged row");
    static_assert(kWin >= 0 && kWin <= kMaxWin, "window held in registers");
    static_assert(!kUni || (kPair && kWin > 0), "uniform taps: exact halving of a fixed window");
    constexpr int kTaps = kWin > 0 ? kWin : kMaxWin;  // unrolled trip count
    if (kWin > 0) {
        wx = kWin;
        wy = kWin;
    }
    __shared__ __align__(16) float s_src[kSrcH * kSrcW];
    __shared__ float s_tmp[kSrcH * kTX];
    __shared__ float s_ky[kTY * kMaxWin];
    __shared__ int s_ly[kTY];
    const int t = threadIdx.x;
    int tX, tY;
    tile_coords(blockIdx.x, tx, tX, tY);
    const int ox0 = tX * kTX, oy0 = tY * kTY;
    const int nx = min(kTX, w - ox0), ny = min(kTY, h - oy0);
    const int cx0 = kUni ? lx0 + 2 * ox0 : lx[ox0], ry0 = kUni ? ly0 + 2 * oy0 : ly[oy0];
    const int sw = kUni ? 2 * (nx - 1) + wx : lx[ox0 + nx - 1] - cx0 + wx;  // staged columns
    const int sh = kUni ? 2 * (ny - 1) + wy : ly[oy0 + ny - 1] - ry0 + wy;  // staged rows

    // The tile's Y weights and first rows, shared by every plane.
    static_assert(kTY * kMaxWin <= 64 * kWaves, "one thread per staged Y weight");
    if (!kUni) {
        if (t < ny * kMaxWin) {
            const int o = t / kMaxWin, j = t % kMaxWin;
            s_ky[t] = j < wy ? ky[size_t(oy0 + o) * wy + j] : 0.0f;
        }
        if (t < ny) s_ly[t] = ly[oy0 + t] - ry0;
    }

    // Per-thread X-pass column and its weights, reused for every row and plane.
    const int xi = t & (kTX - 1);
    const bool xlive = xi < nx;
    float kxr[kMaxWin];
    int xoff = 0;
    if (xlive) {
        xoff = kPair ? 2 * xi : lx[ox0 + xi] - cx0;
#pragma unroll
        for (int j = 0; j < kMaxWin; j++)
            kxr[j] = j < wx ? (kUni ? taps.kx[j] : kx[size_t(ox0 + xi) * wx + j]) : 0.0f;
    }

    // Staging: lane (t & 63) owns source columns (t & 63) + 64k of the tile,
    // whose wrapped image columns are fixed; wave (t >> 6) owns rows
    // (t >> 6) + kWaves i.  A plane's loads are all issued before its LDS stores
    // (out-of-tile slots load a valid in-tile address and are not stored),
    // and the next plane's loads are issued before this plane's taps run, so
    // their latency hides behind the arithmetic.
    constexpr int kColSlots = (kSrcW + kTX - 1) / kTX;
    constexpr int kRowSlots = (kSrcH + kWaves - 1) / kWaves;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    int gx[kColSlots];
    bool cvalid[kColSlots];
#pragma unroll
    for (int k = 0; k < kColSlots; k++) {
        const int col = xi + kTX * k;
        cvalid[k] = col < sw;
        gx[k] = wrap_index(cx0 + (cvalid[k] ? col : 0), W, wrap);
    }
    size_t rowoff[kRowSlots];
#pragma unroll
    for (int i = 0; i < kRowSlots; i++) {
        const int r = min(wave + kWaves * i, sh - 1);
        rowoff[i] = size_t(wrap_index(ry0 + r, H, wrap)) * size_t(W);
    }

    // Planes [blockIdx.z * planes, +planes): all four per workgroup on large
    // images, one per workgroup (four times the workgroups, a quarter of the
    // serial work each) on images too small to fill the device.
    const int planes = 4 / gridDim.z;
    const int c0 = blockIdx.z * planes;
    float v[kRowSlots][kColSlots];
    {
        const float *plane = src + size_t(c0) * size_t(W) * size_t(H);
#pragma unroll
        for (int i = 0; i < kRowSlots; i++)
#pragma unroll
            for (int k = 0; k < kColSlots; k++) v[i][k] = plane[rowoff[i] + gx[k]];
    }
    for (int c = c0; c < c0 + planes; c++) {
#pragma unroll
        for (int i = 0; i < kRowSlots; i++) {
            const int r = wave + kWaves * i;
#pragma unroll
            for (int k = 0; k < kColSlots; k++)
                if (r < sh && cvalid[k]) s_src[r * kSrcW + xi + kTX * k] = v[i][k];
        }
        __syncthreads();
        if (c + 1 < c0 + planes) {
            const float *plane = src + size_t(c + 1) * size_t(W) * size_t(H);
#pragma unroll
            for (int i = 0; i < kRowSlots; i++)
#pragma unroll
                for (int k = 0; k < kColSlots; k++) v[i][k] = plane[rowoff[i] + gx[k]];
        }
        if (xlive) {
            for (int r = wave; r < sh; r += kWaves) {
                const float *row = s_src + r * kSrcW + xoff;
                float sum = 0.0f;
                if (kPair) {
                    // All kMaxWin columns from 2*xi are inside the staged row
                    // (2 * (kTX - 1) + kMaxWin <= kSrcW), so the reads need no
                    // guard; only the taps below wx are summed.
                    // 64-bit words, so each pair is one ds_read_b64.
                    const uint64_t *row2 = reinterpret_cast<const uint64_t *>(row);
                    constexpr int kPairs = (kTaps + 1) / 2;
                    uint64_t q[kPairs];
#pragma unroll
                    for (int p = 0; p < kPairs; ++p) q[p] = row2[p];
#pragma unroll
                    for (int p = 0; p < kPairs; ++p) {
                        if (kWin > 0 || 2 * p < wx) sum += kxr[2 * p] * __uint_as_float(uint32_t(q[p]));
                        if ((kWin > 0 && 2 * p + 1 < kWin) || (kWin == 0 && 2 * p + 1 < wx))
                            sum += kxr[2 * p + 1] * __uint_as_float(uint32_t(q[p] >> 32));
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < kTaps; ++j)
                        if (kWin > 0 || j < wx) sum += kxr[j] * row[j];
                }
                s_tmp[r * kTX + xi] = sum;
            }
        }
        __syncthreads();
        if (xlive) {
            float *out = dst + size_t(c) * size_t(w) * size_t(h);
            for (int o = wave; o < ny; o += kWaves) {
                const float *k = s_ky + o * kMaxWin;
                const float *col = s_tmp + (kUni ? 2 * o : s_ly[o]) * kTX + xi;
                float sum = 0.0f;
#pragma unroll
                for (int j = 0; j < kTaps; ++j)
                    if (kWin > 0 || j < wy) sum += (kUni ? taps.ky[j] : k[j]) * col[j * kTX];
                out[size_t(oy0 + o) * size_t(w) + ox0 + xi] = sum;
            }
        }
        __syncthreads();
    }
}

// Sliding-window form of the fused resize for exact halving with the same
// taps for every output (the polyphase tables of Box, Triangle and Kaiser at
// 2:1 on even extents): each wavefront owns one plane's strip of 64 output
// columns by `rows` output rows and walks down the source rows the strip
// needs, in order.  A source row is staged in the wavefront's own LDS row,
// its X taps give the lane's temporary value (applyKernelX), and that value
// is added into the (kWin + 1) / 2 outputs whose Y windows contain the row,
// whose running sums stay in registers.  Each output therefore sums its Y
// taps in tap order from the same temporaries as applyKernelY: bit-identical
// to the two-pass form, with each staged row read by X taps once (no
// temporary in LDS, no workgroup barriers).  Source rows are fetched eight
// row pairs ahead of their use.
template <int kWin>
__global__ __launch_bounds__(256) void poly_slide_kernel(const float *__restrict__ src, int W, int H,
                                                         float *__restrict__ dst, int w, int h, int lx0, int ly0,
                                                         int rows, int ctiles, int strips, int wrap, SlideTaps taps)
{
    static_assert(kWin & 1, "odd window: (kWin + 1) / 2 outputs per source row");
    constexpr int K = (kWin + 1) / 2;
    constexpr int kSeg = 2 * 63 + kWin;  // staged source columns of 64 outputs
    constexpr int kSlots = (kSeg + 63) / 64;
    __shared__ __align__(16) float s_row[4][2][kSlots * 64];
    // The wavefront index as a scalar: every job coordinate and source row
    // below is then wave-uniform (scalar registers, scalar wrap of rows).
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Workgroups are dispatched round-robin over the eight XCDs, each with
    // its own L2: when the grid divides by eight, XCD k takes the k-th
    // contiguous eighth of the jobs (column tile fastest, then strip), so
    // the halo columns and rows that neighbouring strips share are fetched
    // into the same L2 at about the same time.
    const int nwg = gridDim.x;
    const int wg = (nwg & 7) == 0 ? (blockIdx.x & 7) * (nwg >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const int job = wg * 4 + wv;
    if (job >= 4 * strips * ctiles) return;  // whole wavefront
    const int ct = job % ctiles, rest = job / ctiles, st = rest % strips, c = rest / strips;
    const int ox0 = ct * 64, oy0 = st * rows;
    const int nrows = min(rows, h - oy0);
    const int x = ox0 + lane;
    const bool xlive = x < w;
    // Wrapped source columns of the lane's staging slots (always a valid
    // column, also past the segment, whose slots no lane reads).
    int gx[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; k++) gx[k] = wrap_index(lx0 + 2 * ox0 + lane + 64 * k, W, wrap);
    const float *plane = src + size_t(c) * size_t(W) * size_t(H);
    float *out = dst + size_t(c) * size_t(w) * size_t(h) + size_t(oy0) * size_t(w) + x;
    const int r0 = ly0 + 2 * oy0;              // first source row, before wrapping
    const int last = 2 * (nrows - 1) + kWin - 1;  // index of the strip's last source row (even)
    float *const buf0 = s_row[wv][0], *const buf1 = s_row[wv][1];

    // Register pipeline: rows 2p, 2p + 1 of pair p in r[p % kDepth].
    constexpr int kDepth = 8;
    float r[kDepth][2][kSlots];
    auto fetch = [&](float (&rr)[2][kSlots], int q) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int qq = min(q + i, last);
            const float *row = plane + size_t(wrap_index(r0 + qq, H, wrap)) * size_t(W);
#pragma unroll
            for (int k = 0; k < kSlots; k++) rr[i][k] = row[gx[k]];
        }
    };
    // Fetches are unconditional (rows past the strip's last read the last
    // row again, from cache): the same number of loads on every path keeps
    // each pair's wait at the loads of its own registers instead of all.
#pragma unroll
    for (int d = 0; d < kDepth; d++) fetch(r[d], 2 * d);

    float acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.0f;
    // X taps of the staged row at `b` for the lane's output column.
    auto xtaps = [&](const float *b) {
        const uint64_t *t = reinterpret_cast<const uint64_t *>(b + 2 * lane);
        uint64_t q[(kWin + 1) / 2];
#pragma unroll
        for (int p = 0; p < (kWin + 1) / 2; p++) q[p] = t[p];
        float sum = 0.0f;
#pragma unroll
        for (int p = 0; p < (kWin + 1) / 2; p++) {
            sum += taps.kx[2 * p] * __uint_as_float(uint32_t(q[p]));
            if (2 * p + 1 < kWin) sum += taps.kx[2 * p + 1] * __uint_as_float(uint32_t(q[p] >> 32));
        }
        return sum;
    };
    // One pair of source rows 2p (even: the last Y tap of acc[0], which is
    // then complete) and 2p + 1, from the registers rr; then rr is refilled
    // with pair p + kDepth.  False after the strip's last row.
    auto pair = [&](float (&rr)[2][kSlots], int p) -> bool {
#pragma unroll
        for (int k = 0; k < kSlots; k++) {
            buf0[lane + 64 * k] = rr[0][k];
            buf1[lane + 64 * k] = rr[1][k];
        }
        ntx_wave_sync();
        fetch(rr, 2 * (p + kDepth));
        const float xe = xtaps(buf0);
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += taps.ky[2 * (K - 1 - k)] * xe;
        const int s = p - (K - 1);  // output row completed by row 2p
        if (s >= 0 && xlive) out[size_t(s) * size_t(w)] = acc[0];
        if (2 * p == last) return false;
#pragma unroll
        for (int k = 0; k + 1 < K; k++) acc[k] = acc[k + 1];
        acc[K - 1] = 0.0f;
        const float xo = xtaps(buf1);
#pragma unroll
        for (int k = 0; k + 1 < K; k++) acc[k] += taps.ky[2 * K - 3 - 2 * k] * xo;
        ntx_wave_sync();  // this pair's LDS reads before the next pair's writes
        return true;
    };
    bool more = true;
    for (int p = 0; more; p += kDepth) {
#pragma unroll
        for (int d = 0; d < kDepth; d++)
            if (more) more = pair(r[d], p + d);
    }
}

// The last levels of a chain in one launch: one workgroup per plane holds
// the level in LDS and produces every following level from it, writing each
// to its place in HBM.  These levels are small enough that a launch per
// level costs more than its work.  Each level is the same arithmetic as
// above: box (FloatImage::fastDownSample) through box_texel, or the
// polyphase X taps into a temporary and the Y taps from it, in tap order.
// Source extents: at most kTailSrc texels per plane.
// The tail kernel keeps, for every level whose tables fit (kTailTaps
// entries over all levels), each tap as (weight, source offset) with the
// wrap already applied: the X taps' offsets within a source row, the Y
// taps' offsets of the temporary's rows (packed on the host, MipTail::packed).
// Other levels read the tables from HBM and wrap per tap.
using ntx::kTailTaps;
constexpr int kTailThreads = 512;

// A workgroup barrier that orders LDS only.  The tail kernel's global stores
// are never read back by the kernel, so its barriers need not wait for them
// (__syncthreads would wait for every outstanding store to reach memory).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One polyphase level from packed taps (see kTailTaps): tx holds w rows of
// wx taps, ty h rows of wy taps.
__device__ __forceinline__ void tail_packed_level(const float *cur, float *nxt, float *out, float *s_t, const float2 *tx,
                                                  int wx, const float2 *ty, int wy, int W, int w, int h, int H, int t,
                                                  int nt)
{
    for (int i = t; i < w * H; i += nt) {
        const int y = i / w, x = i - y * w;
        const float2 *kr = tx + x * wx;
        const float *row = cur + y * W;
        float sum = 0.0f;
#pragma unroll 4
        for (int j = 0; j < wx; ++j) {
            const float2 q = kr[j];
            sum += q.x * row[__float_as_int(q.y)];
        }
        s_t[i] = sum;
    }
    lds_barrier();
    for (int i = t; i < w * h; i += nt) {
        const int y = i / w, x = i - y * w;
        const float2 *kr = ty + y * wy;
        float sum = 0.0f;
#pragma unroll 4
        for (int j = 0; j < wy; ++j) {
            const float2 q = kr[j];
            sum += q.x * s_t[__float_as_int(q.y) + x];
        }
        nxt[i] = sum;
        out[i] = sum;
    }
}

// One polyphase level inside the tail kernel: the X taps of every row of
// cur (W x H) into s_t (w x H), then the Y taps of s_t into nxt and out.
// Called with the tap tables either in LDS or in HBM, so that each call
// reads them through one kind of address.
__device__ __forceinline__ void tail_poly_level(const float *cur, float *nxt, float *out, float *s_t,
                                                const float *kx, const int *lx, int wx, const float *ky, const int *ly,
                                                int wy, int W, int H, int w, int h, int wrap, int t, int nt)
{
    for (int i = t; i < w * H; i += nt) {
        const int y = i / w, x = i - y * w;
        const float *kr = kx + x * wx;
        const float *row = cur + y * W;
        const int l = lx[x];
        float sum = 0.0f;
        if (l >= 0 && l + wx <= W) {
#pragma unroll 4
            for (int j = 0; j < wx; ++j) sum += kr[j] * row[l + j];
        } else {
            for (int j = 0; j < wx; ++j) sum += kr[j] * row[wrap_index(l + j, W, wrap)];
        }
        s_t[i] = sum;
    }
    lds_barrier();
    for (int i = t; i < w * h; i += nt) {
        const int y = i / w, x = i - y * w;
        const float *kr = ky + y * wy;
        const int l = ly[y];
        float sum = 0.0f;
        if (l >= 0 && l + wy <= H) {
#pragma unroll 4
            for (int j = 0; j < wy; ++j) sum += kr[j] * s_t[(l + j) * w + x];
        } else {
            for (int j = 0; j < wy; ++j) sum += kr[j] * s_t[wrap_index(l + j, H, wrap) * w + x];
        }
        nxt[i] = sum;
        out[i] = sum;
    }
}

__global__ __launch_bounds__(kTailThreads) void mip_tail_kernel(const float *__restrict__ src, ntx::MipTail tail,
                                                                int wrap)
{
    __shared__ float s_a[ntx::kTailSrc];
    __shared__ float s_b[ntx::kTailSrc / 2];
    __shared__ float s_t[ntx::kTailSrc];
    __shared__ float2 s_k[kTailTaps];
    const int t = threadIdx.x, nt = kTailThreads, c = blockIdx.x;
    // The source plane and the packed taps into LDS: every load issued
    // before the first store, so the kernel waits for one memory latency.
    {
        constexpr int kSrcPer = ntx::kTailSrc / kTailThreads, kTapPer = kTailTaps / kTailThreads;
        const float *plane = src + size_t(c) * size_t(tail.W) * size_t(tail.H);
        const int n = tail.W * tail.H, np = tail.box ? 0 : tail.npacked;
        const float2 *packed = reinterpret_cast<const float2 *>(tail.packed);
        float v[kSrcPer];
        float2 q[kTapPer];
#pragma unroll
        for (int r = 0; r < kSrcPer; r++) v[r] = plane[min(t + r * nt, n - 1)];
#pragma unroll
        for (int r = 0; r < kTapPer; r++)
            if (t + r * nt < np) q[r] = packed[t + r * nt];
#pragma unroll
        for (int r = 0; r < kSrcPer; r++)
            if (t + r * nt < n) s_a[t + r * nt] = v[r];
#pragma unroll
        for (int r = 0; r < kTapPer; r++)
            if (t + r * nt < np) s_k[t + r * nt] = q[r];
    }
    lds_barrier();
    int W = tail.W, H = tail.H, ko = 0;
    for (int k = 0; k < tail.levels; k++) {
        // Level k reads s_a and writes s_b when k is even, the reverse when odd.
        float *cur = (k & 1) ? s_b : s_a, *nxt = (k & 1) ? s_a : s_b;
        const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
        float *out = tail.dst[k] + size_t(c) * size_t(w) * size_t(h);
        if (tail.box) {
            // 1D sources: box_texel takes the output index along the longer axis.
            for (int i = t; i < w * h; i += nt) {
                const int y = (W == 1 || H == 1) ? 0 : i / w, x = (W == 1 || H == 1) ? i : i - y * w;
                const float v = ntx_mip::box_texel(cur, W, H, x, y);
                nxt[i] = v;
                out[i] = v;
            }
        } else {
            const int wx = tail.wx[k], wy = tail.wy[k];
            if (ko + w * wx + h * wy <= kTailTaps) {
                tail_packed_level(cur, nxt, out, s_t, s_k + ko, wx, s_k + ko + w * wx, wy, W, w, h, H, t, nt);
                ko += w * wx + h * wy;
            } else {
                tail_poly_level(cur, nxt, out, s_t, tail.kx[k], tail.lx[k], wx, tail.ky[k], tail.ly[k], wy, W, H, w, h,
                                wrap, t, nt);
            }
        }
        lds_barrier();
        W = w;
        H = h;
    }
}

// Surface.cpp:1562-1569 fromSrgb.  Values that are exactly k/255 (every
// texel of an 8-bit source) take the host-computed table entry, so they match
// the host powf bit for bit; other values use the device powf.
__device__ __forceinline__ float srgb_to_linear(float f, const float *lut)
{
    const float k = rintf(f * 255.0f);
    if (k >= 0.0f && k <= 255.0f && __float_as_uint(k / 255.0f) == __float_as_uint(f)) return lut[int(k)];
    if (f < 0.0f) return 0.0f;
    if (f < 0.04045f) return f / 12.92f;
    if (f <= 1.0f) return powf((f + 0.055f) / 1.055f, 2.4f);
    return 1.0f;
}

// vec: four consecutive values per lane (16-byte loads and stores; d is
// 16-byte aligned), the n % 4 values after the last whole quad one per lane
// on the first lanes.  Otherwise one value per lane.
__global__ __launch_bounds__(256) void srgb_kernel(float *__restrict__ d, size_t n, const float *__restrict__ lut, int vec)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (!vec) {
        if (i < n) d[i] = srgb_to_linear(d[i], lut);
        return;
    }
    const size_t quads = n / 4;
    if (i < quads) {
        float4 v = reinterpret_cast<float4 *>(d)[i];
        v.x = srgb_to_linear(v.x, lut);
        v.y = srgb_to_linear(v.y, lut);
        v.z = srgb_to_linear(v.z, lut);
        v.w = srgb_to_linear(v.w, lut);
        reinterpret_cast<float4 *>(d)[i] = v;
    }
    if (i < n % 4) d[4 * quads + i] = srgb_to_linear(d[4 * quads + i], lut);
}

// Surface.cpp:745-761: BGRA8 texels to planar float / 255.  With lut, also
// Surface::toLinearFromSrgb on R, G, B (Surface.cpp:1562-1592): every value
// is exactly c / 255, whose conversion is lut[c] (srgb_kernel's table path),
// so the result equals the two passes'.  vec: four texels per lane (source
// and planes 16-byte aligned, n % 4 == 0), else one.
__device__ __forceinline__ float unpack_rgb(uint32_t c, const float *lut)
{
    return lut ? lut[c] : float(c) / 255.0f;
}

__global__ __launch_bounds__(256) void unpack_kernel(const uchar4 *__restrict__ src, size_t n, float *__restrict__ dst,
                                                     const float *__restrict__ lut, int vec)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (vec) {
        if (4 * i >= n) return;
        const uint4 q = reinterpret_cast<const uint4 *>(src)[i];  // 4 texels, bytes b, g, r, a
        const uint32_t t[4] = { q.x, q.y, q.z, q.w };
        float4 r, g, b, a;
        float *pr = &r.x, *pg = &g.x, *pb = &b.x, *pa = &a.x;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            pb[k] = unpack_rgb(t[k] & 0xffu, lut);
            pg[k] = unpack_rgb((t[k] >> 8) & 0xffu, lut);
            pr[k] = unpack_rgb((t[k] >> 16) & 0xffu, lut);
            pa[k] = float(t[k] >> 24) / 255.0f;
        }
        reinterpret_cast<float4 *>(dst)[i] = r;
        reinterpret_cast<float4 *>(dst + n)[i] = g;
        reinterpret_cast<float4 *>(dst + 2 * n)[i] = b;
        reinterpret_cast<float4 *>(dst + 3 * n)[i] = a;
        return;
    }
    if (i >= n) return;
    const uchar4 c = src[i];  // x=b, y=g, z=r, w=a
    dst[i] = unpack_rgb(c.z, lut);
    dst[i + n] = unpack_rgb(c.y, lut);
    dst[i + 2 * n] = unpack_rgb(c.x, lut);
    dst[i + 3 * n] = float(c.w) / 255.0f;
}

}  // namespace

namespace ntx {

hipError_t launch_box(const float *d_src, int W, int H, float *d_dst, hipStream_t stream)
{
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    if (W == 1 || H == 1) {
        const int n = w * h;
        hipLaunchKernelGGL(box_kernel, dim3((n + 255) / 256, 1, 4), dim3(256), 0, stream, d_src, W, H, d_dst, n, 1);
    } else {
        hipLaunchKernelGGL(box_kernel, dim3((w + 255) / 256, h, 4), dim3(256), 0, stream, d_src, W, H, d_dst, w, h);
    }
    return hipGetLastError();
}

hipError_t launch_poly_x(const float *d_src, int W, int H, float *d_tmp, const float *d_kx, const int *d_lx, int wx, int wrap,
                         hipStream_t stream)
{
    const int w = W / 2 > 1 ? W / 2 : 1;
    hipLaunchKernelGGL(poly_x_kernel, dim3((w + 255) / 256, H, 4), dim3(256), 0, stream, d_src, W, H, d_tmp, w, d_kx, d_lx,
                       wx, wrap);
    return hipGetLastError();
}

hipError_t launch_poly_y(const float *d_tmp, int W, int H, float *d_dst, const float *d_ky, const int *d_ly, int wy, int wrap,
                         hipStream_t stream)
{
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    hipLaunchKernelGGL(poly_y_kernel, dim3((w + 255) / 256, h, 4), dim3(256), 0, stream, d_tmp, w, H, d_dst, h, d_ky, d_ly,
                       wy, wrap);
    return hipGetLastError();
}

// Whether the fused kernel's staging fits: sw/sh are the largest staged
// extents over all tiles (computed on the host from the tap tables).
bool poly_fused_fits(int max_sw, int max_sh, int wx, int wy)
{
    return max_sw <= kSrcW && max_sh <= kSrcH && wx <= kMaxWin && wy <= kMaxWin;
}

int poly_tile_x() { return kTX; }

// The last tail.levels levels of a chain from the level at d_src (tail.W x
// tail.H, at most kTailSrc texels): one workgroup per plane.
hipError_t launch_mip_tail(const float *d_src, const MipTail &tail, int wrap, hipStream_t stream)
{
    if (tail.levels < 1 || tail.levels > kTailMaxLevels || size_t(tail.W) * size_t(tail.H) > size_t(kTailSrc))
        return hipErrorInvalidValue;
    if (!tail.box && (tail.npacked < 0 || tail.npacked > kTailTaps || (tail.npacked > 0 && !tail.packed)))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(mip_tail_kernel, dim3(4), dim3(kTailThreads), 0, stream, d_src, tail, wrap);
    return hipGetLastError();
}
int poly_tile_y() { return kTY; }

// Strip height of the sliding-window kernel for a W x H source: as tall as
// keeps sixteen wavefronts per compute unit busy (each strip re-reads the
// window's overlap with its neighbour; with fewer wavefronts each one's
// serial walk down its rows sets the time), 0 when that takes strips under
// 16 rows: a wavefront then walks too few rows to hide its row fetches, and
// the tiled kernel, which issues a whole tile's loads at once, is faster.
// (Kaiser(3) at 4096^2 and 2048^2 on 256 compute units: 64 and 16 rows.)
int poly_slide_rows(int W, int H, int cus)
{
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    const size_t ctiles = size_t((w + 63) / 64);
    for (int rows = 64; rows >= 16; rows /= 2)
        if (size_t(4) * ctiles * size_t((h + rows - 1) / rows) >= size_t(16) * size_t(cus)) return rows;
    return 0;
}

hipError_t launch_poly_slide(const float *d_src, int W, int H, float *d_dst, const float *kx, int lx0, const float *ky,
                             int ly0, int win, int wrap, int rows, hipStream_t stream)
{
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    if (win < 3 || win > 15 || (win & 1) == 0 || rows < 1) return hipErrorInvalidValue;
    SlideTaps taps{};
    for (int j = 0; j < win; j++) {
        taps.kx[j] = kx[j];
        taps.ky[j] = ky[j];
    }
    const int ctiles = (w + 63) / 64;
    const int strips = (h + rows - 1) / rows;
    const size_t jobs = size_t(4) * size_t(ctiles) * size_t(strips);
    const dim3 grid(unsigned((jobs + 3) / 4));
    auto kern = poly_slide_kernel<13>;
    switch (win) {
    case 3: kern = poly_slide_kernel<3>; break;
    case 5: kern = poly_slide_kernel<5>; break;
    case 7: kern = poly_slide_kernel<7>; break;
    case 9: kern = poly_slide_kernel<9>; break;
    case 11: kern = poly_slide_kernel<11>; break;
    case 15: kern = poly_slide_kernel<15>; break;
    default: break;
    }
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, stream, d_src, W, H, d_dst, w, h, lx0, ly0, rows, ctiles, strips, wrap,
                       taps);
    return hipGetLastError();
}

hipError_t launch_poly_fused(const float *d_src, int W, int H, float *d_dst, const float *d_kx, const int *d_lx, int wx,
                             bool pair_x, const float *d_ky, const int *d_ly, int wy, int wrap, hipStream_t stream)
{
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    const int tx = (w + kTX - 1) / kTX, ty = (h + kTY - 1) / kTY;
    const int split = tx * ty < 1024 ? 4 : 1;  // workgroups per tile (planes split between them)
    const dim3 grid(unsigned(tx * ty), 1, split);
    // Exact halving with the windows of Kaiser(3), Triangle(2) and
    // Box / Triangle(1): taps unrolled for that count.
    auto kern = pair_x ? poly_fused_kernel<true, 0> : poly_fused_kernel<false, 0>;
    if (pair_x && wx == wy && wx == 13) kern = poly_fused_kernel<true, 13>;
    if (pair_x && wx == wy && wx == 9) kern = poly_fused_kernel<true, 9>;
    if (pair_x && wx == wy && wx == 5) kern = poly_fused_kernel<true, 5>;
    hipLaunchKernelGGL(kern, grid, dim3(64 * kWaves), 0, stream, d_src, W, H, d_dst, w, h, tx, d_kx, d_lx, wx, d_ky, d_ly,
                       wy, wrap, SlideTaps{}, 0, 0);
    return hipGetLastError();
}

bool poly_fused_uniform_window(int win) { return win == 13 || win == 9 || win == 5; }

// The fused kernel with uniform taps (exact halving, the same `win` weights
// kx / ky for every output, first taps lx0 + 2 i and ly0 + 2 i).
hipError_t launch_poly_fused_uniform(const float *d_src, int W, int H, float *d_dst, const float *kx, int lx0,
                                     const float *ky, int ly0, int win, int wrap, hipStream_t stream)
{
    if (!poly_fused_uniform_window(win)) return hipErrorInvalidValue;
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    const int tx = (w + kTX - 1) / kTX, ty = (h + kTY - 1) / kTY;
    const int split = tx * ty < 1024 ? 4 : 1;
    const dim3 grid(unsigned(tx * ty), 1, split);
    SlideTaps taps{};
    for (int j = 0; j < win; j++) {
        taps.kx[j] = kx[j];
        taps.ky[j] = ky[j];
    }
    auto kern = poly_fused_kernel<true, 13, true>;
    if (win == 9) kern = poly_fused_kernel<true, 9, true>;
    if (win == 5) kern = poly_fused_kernel<true, 5, true>;
    hipLaunchKernelGGL(kern, grid, dim3(64 * kWaves), 0, stream, d_src, W, H, d_dst, w, h, tx, nullptr, nullptr, win,
                       nullptr, nullptr, win, wrap, taps, lx0, ly0);
    return hipGetLastError();
}

hipError_t launch_srgb(float *d, size_t n, const float *d_lut, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const int vec = (reinterpret_cast<uintptr_t>(d) & 15) == 0 && n >= 4;
    const size_t lanes = vec ? n / 4 : n;
    hipLaunchKernelGGL(srgb_kernel, dim3(unsigned((lanes + 255) / 256)), dim3(256), 0, stream, d, n, d_lut, vec);
    return hipGetLastError();
}

hipError_t launch_unpack(const uint8_t *d_bgra, size_t n, float *d_dst, const float *d_srgb_lut, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const int vec = (n & 3) == 0 && (reinterpret_cast<uintptr_t>(d_bgra) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(d_dst) & 15) == 0;
    const size_t lanes = vec ? n / 4 : n;
    hipLaunchKernelGGL(unpack_kernel, dim3(unsigned((lanes + 255) / 256)), dim3(256), 0, stream,
                       reinterpret_cast<const uchar4 *>(d_bgra), n, d_dst, d_srgb_lut, vec);
    return hipGetLastError();
}

}  // namespace ntx

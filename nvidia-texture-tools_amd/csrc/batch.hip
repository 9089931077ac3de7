// batch.hip -- the kernels of a batch of textures: every texture's mip chain
// (fastDownSample or polyphase resize) and the BC1 / BC4 / BC5 compression of
// every level, from one work list, in one of two schedules.
//
// Staged: launch s makes level s + 1 of every texture from level s
// (batch_mip_kernel, one workgroup per item), so each launch's sources are
// complete when it starts and hundreds of textures' tiles fill the device
// together; the compression items of every level then run in one persistent
// launch (batch_kernel<..., false>) whose workgroups take items from a queue
// in list order (one atomic per item), largest levels first.
//
// One launch: batch_kernel<..., true> takes mip and compression items from
// the same queue.  Each item reads one level; levels produced inside the
// launch are guarded by a ready counter that every mip item writing the
// level increments once its outputs are globally visible.  The host orders
// the list so an item only ever depends on items before it: the earliest
// unfinished item in the list always has its inputs complete, so the queue
// drains whatever the number of resident workgroups.  A wait that exceeds its
// poll limit sets a status bit and continues, so every wave reaches the end
// of the kernel.
//
// Per-item arithmetic is the per-image kernels' own device code (bc1_block.hpp,
// channel_block.hpp, mip_taps.hpp), so results are bit-identical to them.
#include "bc1_block.hpp"
#include "channel_block.hpp"
#include "mip_taps.hpp"
#include "ntx_batch.h"

namespace {

using ntx::BatchArgs;
using ntx::BatchItem;
using ntx::BatchLevel;

// A descriptor every lane reads at the same address, copied into scalar
// registers.  The loads themselves are vector loads: in the one-launch
// schedule descriptors are read after other workgroups' writes in the same
// launch, which the scalar cache does not observe.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T *p)
{
    static_assert(sizeof(T) % 4 == 0, "descriptor of whole dwords");
    T r;
    const int *src = reinterpret_cast<const int *>(p);
    int *dst = reinterpret_cast<int *>(&r);
#pragma unroll
    for (int i = 0; i < int(sizeof(T) / 4); i++) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return r;
}

// Level buffers are device (global) memory: say so, so that accesses through
// them are global rather than flat instructions.
template <typename T>
__device__ __forceinline__ T *global_ptr(T *p)
{
    return (T *)(__attribute__((address_space(1))) T *)p;
}

__device__ __forceinline__ void wait_ready(int *counter, int target, int spin_limit, int *status)
{
    // Gives up after spin_limit polls, or at once when another wait already
    // gave up (the launch is failing; later waits need not spin as well).
    int polls = 0;
    while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++polls > spin_limit || __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            atomicOr(status, 1);
            return;
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

// LDS for a polyphase tile in the separable form: the tile's source region
// (wrap applied while staging) and the X taps' temporary of its rows.
constexpr int kMipSrc = 2048, kMipTmp = 1024, kMipTaps = 768;

// Tiles of the next level: thread t of a (1 << shift) x (256 >> shift) tile
// computes one output texel on all four planes.  A polyphase tile whose
// source region fits kMipSrc (and its temporary kMipTmp) runs separably: the
// region is staged in LDS, the X taps of every source row the tile's Y taps
// read fill the temporary, and the Y taps read the temporary.  Every X value
// is the reference's applyKernelX value of that (wrapped) row and every
// output sums its Y taps in tap order, exactly as poly_texel does per texel,
// so both forms give the same bits; poly_texel handles the tiles that do not
// fit.
__device__ void mip_item(const BatchArgs &a, const BatchItem &it, const BatchLevel &src, int t, float *s_src,
                         float *s_tmp, float *s_taps, int *s_first)
{
    const BatchLevel dst = uniform_copy(a.levels + src.next);
    const int W = src.w, H = src.h, w = dst.w, h = dst.h;
    const int tw = 1 << a.tile_shift, th = 256 >> a.tile_shift;
    const int tiles_x = (w + tw - 1) / tw;
    const size_t splane = size_t(W) * size_t(H), dplane = size_t(w) * size_t(h);
    const float *s = global_ptr(src.data);
    float *d = global_ptr(dst.data);
    for (int k = 0; k < it.count; k++) {
        const int tile = it.first + k;
        const int ox0 = (tile % tiles_x) * tw, oy0 = (tile / tiles_x) * th;
        const int x = ox0 + (t & (tw - 1));
        const int y = oy0 + (t >> a.tile_shift);
        const bool live = x < w && y < h;
        const size_t o = size_t(y) * size_t(w) + size_t(x);
        if (src.mip_kind == 0) {
            if (!live) continue;
            const bool line = W == 1 || H == 1;
            const int bx = line ? x + y : x, by = line ? 0 : y;  // one of x, y is 0 for a line
#pragma unroll
            for (int c = 0; c < 4; c++) d[c * dplane + o] = ntx_mip::box_texel(s + c * splane, W, H, bx, by);
            continue;
        }
        // The tile's source region: columns lx[ox0] .. lx[ox0 + nx - 1] + wx,
        // rows likewise (the first taps grow with the output index).
        const int nx = min(tw, w - ox0), ny = min(th, h - oy0);
        const int cx0 = src.lx[ox0], ry0 = src.ly[oy0];
        const int sw = src.lx[ox0 + nx - 1] - cx0 + src.wx, sh = src.ly[oy0 + ny - 1] - ry0 + src.wy;
        if (sw * sh > kMipSrc || sh * nx > kMipTmp || nx * src.wx + ny * src.wy > kMipTaps || splane > 0xffffffffu) {
            if (!live) continue;
#pragma unroll 1
            for (int c = 0; c < 4; c++)
                d[c * dplane + o] = ntx_mip::poly_texel(s + c * splane, W, H, x, y, src.kx, src.lx, src.wx, src.ky,
                                                        src.ly, src.wy, src.wrap);
            continue;
        }
        // The tile's taps and first taps, shared by its four planes: the X
        // weights of outputs ox0.., then the Y weights of rows oy0.. (both
        // contiguous in the tables), the first taps relative to the region.
        float *s_kx = s_taps, *s_ky = s_taps + nx * src.wx;
        for (int i = t; i < nx * src.wx; i += 256) s_kx[i] = src.kx[size_t(ox0) * src.wx + i];
        for (int i = t; i < ny * src.wy; i += 256) s_ky[i] = src.ky[size_t(oy0) * src.wy + i];
        if (t < nx) s_first[t] = src.lx[ox0 + t] - cx0;
        if (t < ny) s_first[256 + t] = src.ly[oy0 + t] - ry0;
        // Thread t stages region elements t + 256 k; their (wrapped) source
        // offsets are the same in every plane, and plane c + 1's values are
        // fetched into registers while plane c is filtered.  Element
        // t + 256 k sits at row r + k q, column col + k m of the region
        // (256 = q sw + m), stepped without a division per element; slots
        // past the region take its last element.  A region inside the image
        // needs no wrap (a branch uniform over the workgroup).
        constexpr int kSlots = kMipSrc / 256;
        const int nsrc = sw * sh;
        const int q = 256 / sw, m = 256 - q * sw;
        const bool inside = cx0 >= 0 && cx0 + sw <= W && ry0 >= 0 && ry0 + sh <= H;
        uint32_t off[kSlots];
        float v[kSlots];
        {
            int r = t / sw, col = t - r * sw;
#pragma unroll
            for (int k = 0; k < kSlots; k++) {
                const bool in = t + 256 * k < nsrc;
                const int rr = in ? r : sh - 1, cc = in ? col : sw - 1;
                if (inside)
                    off[k] = uint32_t(ry0 + rr) * uint32_t(W) + uint32_t(cx0 + cc);
                else
                    off[k] = uint32_t(ntx_mip::wrap_index(ry0 + rr, H, src.wrap)) * uint32_t(W) +
                             uint32_t(ntx_mip::wrap_index(cx0 + cc, W, src.wrap));
                r += q;
                col += m;
                if (col >= sw) {
                    col -= sw;
                    r++;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kSlots; k++) v[k] = s[off[k]];
#pragma unroll 1
        for (int c = 0; c < 4; c++) {
#pragma unroll
            for (int k = 0; k < kSlots; k++)
                if (t + 256 * k < nsrc) s_src[t + 256 * k] = v[k];
            __syncthreads();
            if (c + 1 < 4) {
                const float *pl = s + size_t(c + 1) * splane;
#pragma unroll
                for (int k = 0; k < kSlots; k++) v[k] = pl[off[k]];
            }
            if (nx == tw) {
                // Full-width tile: the thread's column is t mod tw in every
                // row it takes (256 is a multiple of tw).
                const int xi = t & (tw - 1);
                const float *kr = s_kx + xi * src.wx;
                const float *row = s_src + (t >> a.tile_shift) * sw + s_first[xi];
                for (int i = t; i < sh * nx; i += 256, row += th * sw) {
                    float sum = 0.0f;
                    for (int j = 0; j < src.wx; ++j) sum += kr[j] * row[j];
                    s_tmp[i] = sum;
                }
            } else {
                for (int i = t; i < sh * nx; i += 256) {
                    const int r = i / nx, xi = i - r * nx;
                    const float *kr = s_kx + xi * src.wx;
                    const float *row = s_src + r * sw + s_first[xi];
                    float sum = 0.0f;
                    for (int j = 0; j < src.wx; ++j) sum += kr[j] * row[j];
                    s_tmp[i] = sum;
                }
            }
            __syncthreads();
            if (live) {
                const int xi = x - ox0, yi = y - oy0;
                const float *kr = s_ky + yi * src.wy;
                const float *col = s_tmp + s_first[256 + yi] * nx + xi;
                float sum = 0.0f;
                for (int j = 0; j < src.wy; ++j) sum += kr[j] * col[j * nx];
                d[c * dplane + o] = sum;
            }
            __syncthreads();  // the next plane's staging overwrites both arrays (and the next tile the taps)
        }
    }
}

// BC1 blocks [first, first + count): block first + 16 i + g on group g.
template <int LEVEL>
__device__ void bc1_item(const BatchArgs &a, const BatchItem &it, const BatchLevel &lv, ntx_bc1::Bc1Shared &sh, int t)
{
    const int bw = (lv.w + 3) / 4;
    const int g = t >> 4, lane = t & 15;
    const float *data = global_ptr(lv.data);
    uint2 *out = global_ptr(reinterpret_cast<uint2 *>(lv.out));
    for (int k = g; k < it.count; k += ntx_bc1::kBlocksPerWG) {
        const int blk = it.first + k;
        int n, m;
        const uint2 r = ntx_bc1::encode_block<LEVEL>(sh, data, lv.w, lv.h, blk % bw, blk / bw, lv.alpha_mode, lv.cw[0],
                                                     lv.cw[1], lv.cw[2], a.tab, g, lane, &n, &m, lv.bc1_level == 9);
        if (lane == 0) out[blk] = r;
    }
}

// BC4 (R) or BC5 (R then G) channel jobs, one per thread.
__device__ void channel_item(const BatchItem &it, const BatchLevel &lv, int t)
{
    const float *data = global_ptr(lv.data);
    uint8_t *out = global_ptr(lv.out);
    const int bw = (lv.w + 3) / 4;
    const int njobs = lv.format == NTX_FORMAT_BC5 ? 2 : 1;
    const int stride = 8 * njobs;
    const size_t plane = size_t(lv.w) * size_t(lv.h);
    for (int k = t; k < it.count; k += 256) {
        const int job = it.first + k;
        const int blk = job / njobs, part = job - blk * njobs;
        const int bx = blk % bw, by = blk / bw;
        const float *p = data + size_t(part) * plane;
        uint32_t vq[4];
        ntx_bc4::gather_channel(p, lv.w, lv.h, bx, by, vq);
        const uint64_t r = ntx_bc4::compress_channel(vq);
        *reinterpret_cast<uint2 *>(out + size_t(blk) * stride + 8 * part) = make_uint2(uint32_t(r), uint32_t(r >> 32));
    }
}

// The mip items of one stage (their sources complete before the launch), one
// per workgroup.  Workgroups are dispatched round-robin over the eight XCDs,
// each with its own L2: XCD x takes a contiguous run of items, so the halo
// rows and columns that neighbouring tiles share are fetched into one L2.
__global__ __launch_bounds__(256) void batch_mip_kernel(BatchArgs a)
{
    __shared__ float s_src[kMipSrc], s_tmp[kMipTmp], s_taps[kMipTaps];
    __shared__ int s_first[512];
    const int t = threadIdx.x;
    const int nwg = int(gridDim.x), b = int(blockIdx.x);
    const int xcd = b & 7, q = nwg >> 3, rem = nwg & 7;
    const int idx = (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + (b >> 3);
    if (idx < a.nitems) {
        const BatchItem it = uniform_copy(a.items + idx);
        const BatchLevel lv = uniform_copy(a.levels + it.level);
        mip_item(a, it, lv, t, s_src, s_tmp, s_taps, s_first);
    }
}

// The persistent kernel.  kMips false: the compression items of the staged
// schedule, every level already complete.  kMips true: the one-launch
// schedule, mip items included and dependencies waited for.
// kWaves: waves per SIMD the register allocation is held to (0: the
// compiler's choice); the autotuner picks among the instances.  kTiers: the
// batch holds BC1 levels at icbc levels other than 8 (Fastest, Production),
// whose code would otherwise cost every batch registers.
template <int kWaves, bool kTiers, bool kMips>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWaves > 0 ? kWaves : 1)))
void batch_kernel(BatchArgs a)
{
    __shared__ ntx_bc1::Bc1Shared sh;
    __shared__ float s_src[kMips ? kMipSrc : 1], s_tmp[kMips ? kMipTmp : 1], s_taps[kMips ? kMipTaps : 1];
    __shared__ int s_first[kMips ? 512 : 1];
    __shared__ int s_item;
    const int t = threadIdx.x;
    ntx_bc1::load_shared_tables(sh, a.tab, t, 256);
    for (;;) {
        __syncthreads();  // the previous item is finished by every thread
        if (t == 0) {
            const int it = atomicAdd(a.head, 1);
            if (kMips && it < a.nitems) {
                const BatchLevel &lv = a.levels[a.items[it].level];
                if (lv.ready >= 0) wait_ready(a.counters + lv.ready, lv.target, a.spin_limit, a.status);
            }
            s_item = it;
        }
        __syncthreads();
        const int idx = __builtin_amdgcn_readfirstlane(s_item);
        if (idx >= a.nitems) break;
        const BatchItem it = uniform_copy(a.items + idx);
        const BatchLevel lv = uniform_copy(a.levels + it.level);
        if (kMips && lv.ready >= 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (kMips && it.kind == ntx::kItemMip) {
            mip_item(a, it, lv, t, s_src, s_tmp, s_taps, s_first);
            // Every thread makes its stores visible device-wide before the
            // workgroup counts the item as done.
            __threadfence();
            __syncthreads();
            if (t == 0) {
                const int ready = __builtin_amdgcn_readfirstlane(a.levels[lv.next].ready);
                __hip_atomic_fetch_add(a.counters + ready, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (it.kind == ntx::kItemBc1) {
            if constexpr (kTiers) {
                // One cluster-fit instance for levels 8 and 9 (two would
                // spill): level 9's refinement is switched per level.
                if (lv.bc1_level == 1)
                    bc1_item<1>(a, it, lv, sh, t);
                else
                    bc1_item<9>(a, it, lv, sh, t);
            } else {
                bc1_item<8>(a, it, lv, sh, t);
            }
        } else {
            channel_item(it, lv, t);
        }
    }
}

// Clears the ready counters, the queue position and the status word (a
// kernel rather than a memset node, so the run can be captured in a graph).
__global__ __launch_bounds__(256) void batch_reset_kernel(int *p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

namespace ntx {

// variant 0-2: the tuned register budget (compiler's choice, 5 or 6 waves per
// SIMD); tiers: the instance with the BC1 Fastest / Production paths; mips:
// the one-launch schedule's instance.
using BatchKernel = void (*)(BatchArgs);
template <bool kMips>
BatchKernel batch_kernel_sched(int variant, bool tiers)
{
    switch (variant) {
    case 1: return tiers ? batch_kernel<5, true, kMips> : batch_kernel<5, false, kMips>;
    case 2: return tiers ? batch_kernel<6, true, kMips> : batch_kernel<6, false, kMips>;
    default: return tiers ? batch_kernel<0, true, kMips> : batch_kernel<0, false, kMips>;
    }
}

BatchKernel batch_kernel_of(int variant, bool tiers, bool mips)
{
    return mips ? batch_kernel_sched<true>(variant, tiers) : batch_kernel_sched<false>(variant, tiers);
}

const void *batch_kernel_variant(int variant, bool tiers, bool mips)
{
    return reinterpret_cast<const void *>(batch_kernel_of(variant, tiers, mips));
}

int batch_kernel_variants() { return 3; }

hipError_t launch_batch(const BatchArgs &args, int variant, bool tiers, bool mips, int workgroups, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_kernel_of(variant, tiers, mips), dim3(workgroups), dim3(256), 0, stream, args);
    return hipGetLastError();
}

// One stage of mips: args.items / args.nitems are the stage's items.
hipError_t launch_batch_mips(const BatchArgs &args, hipStream_t stream)
{
    if (args.nitems <= 0) return hipSuccess;
    // One workgroup per item (batch_mip_kernel's item mapping assumes it).
    hipLaunchKernelGGL(batch_mip_kernel, dim3(unsigned(args.nitems)), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launch_batch_reset(int *p, int n, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

// Workgroups of one batch_kernel instance that fit on the device at once.
hipError_t batch_resident_workgroups(int device, int variant, bool tiers, bool mips, int *count)
{
    int per_cu = 0, cus = 0;
    hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_variant(variant, tiers, mips), 256, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    *count = per_cu * cus;
    return e;
}

}  // namespace ntx

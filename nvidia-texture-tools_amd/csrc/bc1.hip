// bc1.hip -- BC1 encoder for gfx950, bit-exact with NVTT's CPU path
// (icbc v1.05 compress_dxt1 at Quality_Level8; Level9 for Quality_Production,
// Level1 for Quality_Fastest), icbc.h:3485-3559.
//
// Two mappings:
//  * bc1_kernel: one 4x4 block per 16-lane group, four blocks per wave64,
//    sixteen per 256-thread workgroup (bc1_block.hpp); used for small
//    images, the other quality levels and as the reference form.
//  * bc1_classify_kernel + bc1_search_kernel (Level8, large images): one
//    block per lane.  The classify pass reduces each block's colours,
//    finishes blocks of 0 or 1 colours and appends the others to the list of
//    their colour count n; the search pass takes whole wavefronts from one
//    list, so every lane runs the same cluster-table loops (bc1_lane.hpp)
//    with scalar loop control, no cross-lane traffic and, for unit-weight
//    blocks, the candidates' weight terms from tables.
#include "bc1_lane.hpp"
#include "bucket.hpp"

namespace {

using namespace ntx_bc1;

template <int LEVEL>
__global__ __launch_bounds__(256) void bc1_kernel(const float *__restrict__ src, int w, int h, int bw, int nblocks,
                                                  int alpha_mode, float cwx, float cwy, float cwz,
                                                  const ntx::Tables *__restrict__ tab, uint2 *__restrict__ out,
                                                  short2 *__restrict__ nm, int urgent)
{
    if (urgent) __builtin_amdgcn_s_setprio(3);
    __shared__ Bc1Shared sh;
    load_shared_tables(sh, tab, threadIdx.x, blockDim.x);
    __syncthreads();

    const int g = threadIdx.x >> 4;
    const int lane = threadIdx.x & 15;
    const int blk = blockIdx.x * kBlocksPerWG + g;
    if (blk >= nblocks) return;  // whole group leaves together
    int n, m;
    const uint2 result =
        encode_block<LEVEL>(sh, src, w, h, blk % bw, blk / bw, alpha_mode, cwx, cwy, cwz, tab, g, lane, &n, &m);
    if (lane == 0) {
        out[blk] = result;
        if (nm) nm[blk] = make_short2(short(n), short(m));
    }
}

// The blocks listed in list[0, min(*count, cap)) (indices below cap), one
// per 16-lane group as bc1_kernel encodes them, the groups striding over the
// list.  The encoder synchronises only within a wavefront, so groups may
// finish their strides at different times.  For short lists (the learned
// form's fallback blocks) this has a group's latency, where the one-block-
// per-lane search pass has a whole block's search per lane.
template <int LEVEL>
__global__ __launch_bounds__(256) void bc1_list_kernel(const float *__restrict__ src, int w, int h, int bw,
                                                       const int *__restrict__ list, const int *__restrict__ count,
                                                       int cap, int alpha_mode, float cwx, float cwy, float cwz,
                                                       const ntx::Tables *__restrict__ tab, uint2 *__restrict__ out)
{
    __shared__ Bc1Shared sh;
    load_shared_tables(sh, tab, threadIdx.x, blockDim.x);
    __syncthreads();
    const int g = threadIdx.x >> 4;
    const int lane = threadIdx.x & 15;
    const int items = min(*count, cap);
    for (int item = blockIdx.x * kBlocksPerWG + g; item < items; item += gridDim.x * kBlocksPerWG) {
        const int blk = list[item];
        if (unsigned(blk) >= unsigned(cap)) continue;  // the whole group skips it
        int n, m;
        const uint2 result =
            encode_block<LEVEL>(sh, src, w, h, blk % bw, blk / bw, alpha_mode, cwx, cwy, cwz, tab, g, lane, &n, &m);
        if (lane == 0) out[blk] = result;
    }
}

// One item of bc1_classify_kernel per thread; the whole workgroup calls it.
__device__ __forceinline__ void classify_chunk(const float *__restrict__ src, int w, int h, int bw,
                                               const int *__restrict__ sel, int items, int alpha_mode,
                                               const ntx::Tables *__restrict__ tab, uint2 *__restrict__ out,
                                               short2 *__restrict__ nm, int *__restrict__ counts,
                                               int *__restrict__ lists, int cap, int rows16, float4 (*s_rows)[16][64],
                                               int *s_cnt, int *s_base, int item, int tag = 0)
{
    using namespace ntx_bucket;
    const int t = threadIdx.x;
    if (t < kCountSlots) s_cnt[t] = 0;
    __syncthreads();
    bool has = item < items;
    int blk = has ? (sel ? sel[item] : item) : 0;
    if (unsigned(blk) >= unsigned(cap)) has = false, blk = 0;  // listed indices are below cap
    int k = 0, local = -1;
    if (has) {
        float4 tx[16];
        lane_gather(src, w, h, blk % bw, blk / bw, alpha_mode, rows16 != 0, tx);
        bool any_black;
        LaneRows rows = s_rows[t >> 6];
        const int lane = t & 63;
        const int n = lane_reduce(rows, lane, tx, any_black);
        if (n <= 1) {
            uint2 r = make_uint2(0u, 0u);
            if (n == 1) {
                const float4 c = rows[0][lane];
                r = single_color_block(tab, c.x, c.y, c.z);
            }
            out[blk] = r;
            if (nm) nm[blk] = make_short2(short(n), short(-1));
        } else {
            k = n;
            local = atomicAdd(&s_cnt[n], 1);
        }
    }
    append_to_lists(s_cnt, s_base, counts, lists, cap, k, local, tag | blk);
}

// Level8 pass 1, one block per lane: FloatColorCompressor's gather and
// reduce_colors; a block of no colour (every weight 0) is the zero block, a
// block of one colour compress_dxt1_single_color_optimal (icbc.h:3515-3525),
// every other block goes to list n.  The clusters live in LDS only while
// they are counted.  Blocks 0..nblocks-1 of the image, or, when sel is not
// null, the *sel_count blocks listed in sel.
__global__ __launch_bounds__(256) void bc1_classify_kernel(const float *__restrict__ src, int w, int h, int bw,
                                                           int nblocks, const int *__restrict__ sel,
                                                           const int *__restrict__ sel_count, int alpha_mode,
                                                           const ntx::Tables *__restrict__ tab,
                                                           uint2 *__restrict__ out, short2 *__restrict__ nm,
                                                           int *__restrict__ counts, int *__restrict__ lists, int cap,
                                                           int rows16, int urgent)
{
    using namespace ntx_bucket;
    if (urgent) __builtin_amdgcn_s_setprio(3);
    __shared__ float4 s_rows[4][16][64];
    __shared__ int s_cnt[kCountSlots], s_base[kCountSlots];
    const int t = threadIdx.x;
    const int items = sel ? min(*sel_count, cap) : nblocks;
    // Workgroups stride over the items (the grid of a list is sized for its
    // capacity, not its count, which only the device knows).
    for (int base = blockIdx.x * blockDim.x; base < items; base += gridDim.x * blockDim.x) {
        classify_chunk(src, w, h, bw, sel, items, alpha_mode, tab, out, nm, counts, lists, cap, rows16, s_rows, s_cnt,
                       s_base, base + t);
    }
}


// The image a search lane encodes a block of: the single image of
// bc1_search_kernel (the same in every lane) or, in the multi-image form,
// the image its list entry names.
struct LaneImage {
    const float *src;
    uint2 *out;
    int w, h, bw, alpha_mode, rows16;
    float cwx, cwy, cwz;
};

// One block per lane of a wavefront that holds blocks of count k: block blk
// of image im, written when live (lanes past the end of their list encode
// the wave's first block again, with the same count and valid weights, and
// discard it).
__device__ __forceinline__ void search_lane(const LaneImage &im, int blk, bool live,
                                            const ntx::Tables *__restrict__ tab, short2 *__restrict__ nm,
                                            LaneRows s_rows, int lane, int k)
{
    float4 tx[16];
    lane_gather(im.src, im.w, im.h, blk % im.bw, blk / im.bw, im.alpha_mode, im.rows16 != 0, tx);
    bool unit = true;
#pragma unroll
    for (int i = 0; i < 16; i++) unit = unit && tx[i].w == 1.0f;
    // Every live lane's texels weigh 1: cluster weights are integer counts.
    const bool units = __builtin_amdgcn_ballot_w64(live && !unit) == 0;

    bool any_black;
    const int n = lane_reduce(s_rows, lane, tx, any_black);  // == k in live lanes
    const bool rebuild = any_black && lane_has_black_cluster(s_rows, lane, n);
    const float4 sum = lane_sat(s_rows, lane, n);
    const float mx = im.cwx * im.cwx, my = im.cwy * im.cwy, mz = im.cwz * im.cwz;
    const float *mid5 = tab->mid5, *mid6 = tab->mid6, *div255 = tab->div255;

    // Four colours over the block's n clusters (n is k in every live lane).
    const Best b4 = lane_search4(s_rows, lane, k, sum, mx, my, mz, tab, units);
    uint2 result;
    float best = lane_output<true>(lane_winner<true>(s_rows, lane, b4, sum, mx, my, mz, tab), tx, im.cwx, im.cwy,
                                   im.cwz, mid5, mid6, div255, &result);

    // Three colours: the same clusters, or the non-black ones.
    int m = n;
    float4 sum3 = sum;
    if (rebuild) {
        bool unused;
        lane_reduce(s_rows, lane, tx, unused);
        m = lane_skip_blacks(s_rows, lane, n);
        if (m > 0) sum3 = lane_sat(s_rows, lane, m);
    }
    const bool run3 = m > 0;
    const bool units3 = units && __builtin_amdgcn_ballot_w64(live && run3 && sum3.w != 16.0f) == 0;
    const bool uniform3 = __builtin_amdgcn_ballot_w64(live && m != k) == 0;
    if (run3) {
        const Best b3 = uniform3 ? lane_search3(s_rows, lane, k, sum3, mx, my, mz, tab, units3)
                                 : lane_search3(s_rows, lane, m, sum3, mx, my, mz, tab, units3);
        uint2 r3;
        const float e3 = lane_output<false>(lane_winner<false>(s_rows, lane, b3, sum3, mx, my, mz, tab), tx, im.cwx,
                                            im.cwy, im.cwz, mid5, mid6, div255, &r3);
        if (e3 < best) { best = e3; result = r3; }
    }
    if (!(best < FLT_MAX)) result = make_uint2(0u, 0u);
    if (live) {
        im.out[blk] = result;
        if (nm) nm[blk] = make_short2(short(n), short(run3 ? m : -1));
    }
}

// The wave of bc1_search_kernel's search: entries first .. first + 63 of
// list k, k and first found from the wave's index among the lists' waves
// (walked from count 16 down).  False past the last wave.
__device__ __forceinline__ bool list_wave(const int *__restrict__ counts, int cap, int wave, int &k, int &first,
                                          int &cnt)
{
    int wv = wave;
    k = 0;
    for (int b = 16; b >= 2; b--) {
        const int c = min(counts[b], cap);
        const int waves = (c + 63) >> 6;
        if (k == 0) {
            if (wv < waves) { k = b; first = wv * 64; cnt = c; }
            else wv -= waves;
        }
    }
    return k != 0;
}

// Level8 pass 2, one block per lane (icbc.h:3292-3326 + :3543-3550).  The
// lists are cut into waves of 64 consecutive entries, walked from n = 16
// down; workgroup b takes waves b, b + grid, ...; lanes past the end of
// their list repeat its first block and write nothing.  The
// block is gathered and reduced again (the same n), its prefix-sum table
// built in LDS, the four-colour fit searched, then (after skip_blacks when
// some cluster is black) the three-colour fit, and the better block kept.
// The last workgroup clears the tallies.
__global__ __launch_bounds__(64) void bc1_search_kernel(const float *__restrict__ src, int w, int h, int bw,
                                                        int alpha_mode, float cwx, float cwy, float cwz,
                                                        const ntx::Tables *__restrict__ tab, uint2 *__restrict__ out,
                                                        short2 *__restrict__ nm, int *__restrict__ counts,
                                                        const int *__restrict__ lists, int cap, int rows16)
{
    using namespace ntx_bucket;
    __shared__ float4 s_rows[17][64];
    const int lane = threadIdx.x;
    const LaneImage im{ src, out, w, h, bw, alpha_mode, rows16, cwx, cwy, cwz };
    // A list never holds more than cap entries.
    int total = 0;
    for (int b = 2; b <= 16; b++) total += (min(counts[b], cap) + 63) >> 6;
    for (int wave = blockIdx.x; wave < total; wave += gridDim.x) {
        int k = 0, first = 0, cnt = 0;
        if (!list_wave(counts, cap, wave, k, first, cnt)) break;
        const int item = first + lane;
        bool live = item < cnt;
        // Every caller's lists hold block indices below cap; anything else
        // is written nowhere.
        int blk = lists[size_t(k) * cap + (live ? item : first)];
        if (unsigned(blk) >= unsigned(cap)) {
            blk = 0;
            live = false;
        }
        search_lane(im, blk, live, tab, nm, s_rows, lane, k);
    }
    release_counts(counts);
}

// The multi-image form (many images' blocks in one pair of launches, e.g.
// every BC1 level of a batch): an image table, list entries
// image << kImageShift | block, and for pass 1 one workgroup per chunk of
// 256 blocks of one image (map[g] = (image, first block)).
constexpr int kImageShift = 20;

__global__ __launch_bounds__(256) void bc1_classify_multi_kernel(const ntx::Bc1Image *__restrict__ imgs,
                                                                 const int2 *__restrict__ map,
                                                                 const ntx::Tables *__restrict__ tab,
                                                                 int *__restrict__ counts, int *__restrict__ lists,
                                                                 int cap)
{
    __shared__ float4 s_rows[4][16][64];
    __shared__ int s_cnt[ntx_bucket::kCountSlots], s_base[ntx_bucket::kCountSlots];
    const int2 m = map[blockIdx.x];
    const ntx::Bc1Image im = imgs[m.x];
    const int bw = (im.w + 3) / 4;
    const int blocks = bw * ((im.h + 3) / 4);
    classify_chunk(im.src, im.w, im.h, bw, nullptr, blocks, im.alpha_mode, tab, reinterpret_cast<uint2 *>(im.out),
                   nullptr, counts, lists, cap, im.rows16, s_rows, s_cnt, s_base, m.y + int(threadIdx.x),
                   m.x << kImageShift);
}

__global__ __launch_bounds__(64) void bc1_search_multi_kernel(const ntx::Bc1Image *__restrict__ imgs,
                                                              const ntx::Tables *__restrict__ tab,
                                                              int *__restrict__ counts, const int *__restrict__ lists,
                                                              int cap, int nimgs)
{
    using namespace ntx_bucket;
    __shared__ float4 s_rows[17][64];
    const int lane = threadIdx.x;
    int total = 0;
    for (int b = 2; b <= 16; b++) total += (min(counts[b], cap) + 63) >> 6;
    for (int wave = blockIdx.x; wave < total; wave += gridDim.x) {
        int k = 0, first = 0, cnt = 0;
        if (!list_wave(counts, cap, wave, k, first, cnt)) break;
        const int item = first + lane;
        bool live = item < cnt;
        const int e = lists[size_t(k) * cap + (live ? item : first)];
        int img = int(unsigned(e) >> kImageShift), blk = e & ((1 << kImageShift) - 1);
        if (img >= nimgs) {
            img = 0;
            blk = 0;
            live = false;
        }
        const ntx::Bc1Image d = imgs[img];
        const LaneImage im{ d.src, reinterpret_cast<uint2 *>(d.out), d.w, d.h, (d.w + 3) / 4, d.alpha_mode, d.rows16,
                            d.cw[0], d.cw[1], d.cw[2] };
        if (blk >= im.bw * ((im.h + 3) / 4)) {
            blk = 0;
            live = false;
        }
        search_lane(im, blk, live, tab, nullptr, s_rows, lane, k);
    }
    release_counts(counts);
}

}  // namespace

namespace ntx {

// urgent: the waves raise their issue priority (s_setprio), for small
// latency-bound images compressed beside large ones; level: icbc quality
// level 8 (nvtt Normal / Highest), 9 (Production) or 1 (Fastest).
hipError_t launch_bc1(const float *d_src, int w, int h, int alpha_mode, const float *cw, const Tables *d_tab,
                      void *d_out, int16_t *d_nm, hipStream_t stream, bool urgent, int level)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int nblocks = bw * bh;
    if (nblocks == 0) return hipSuccess;
    const int grid = (nblocks + kBlocksPerWG - 1) / kBlocksPerWG;
    auto kernel = level == 1 ? bc1_kernel<1> : level == 9 ? bc1_kernel<9> : bc1_kernel<8>;
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlocksPerWG * kGroup), 0, stream, d_src, w, h, bw, nblocks, alpha_mode,
                       cw[0], cw[1], cw[2], d_tab, reinterpret_cast<uint2 *>(d_out), reinterpret_cast<short2 *>(d_nm),
                       urgent ? 1 : 0);
    return hipGetLastError();
}

// Level8 pass 1 into the block lists in `counts` / `lists` (zero tallies on
// entry; lists: 17 x cap ints, cap >= the number of blocks classified): the
// image's blocks, or the *d_sel_count blocks listed in d_sel (at most cap;
// then at most max_groups workgroups, striding over the list).
hipError_t launch_bc1_classify(const float *d_src, int w, int h, int alpha_mode, const Tables *d_tab, void *d_out,
                               int16_t *d_nm, const int *d_sel, const int *d_sel_count, int *d_counts, int *d_lists,
                               int cap, int max_groups, hipStream_t stream, bool urgent)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4, nblocks = bw * bh;
    const int items = d_sel ? cap : nblocks;
    if (items == 0) return hipSuccess;
    const int rows16 = (reinterpret_cast<uintptr_t>(d_src) & 15) == 0 && (w & 3) == 0;
    int grid = (items + 255) / 256;
    if (d_sel && grid > max_groups) grid = max_groups;
    hipLaunchKernelGGL(bc1_classify_kernel, dim3(grid), dim3(256), 0, stream, d_src, w, h, bw, nblocks,
                       d_sel, d_sel_count, alpha_mode, d_tab, reinterpret_cast<uint2 *>(d_out),
                       reinterpret_cast<short2 *>(d_nm), d_counts, d_lists, cap, rows16, urgent ? 1 : 0);
    return hipGetLastError();
}

// Level8 pass 2 over the lists pass 1 filled (at most `cap` blocks in all);
// leaves the tallies zero.  max_groups bounds the grid (the workgroups loop
// over the wavefronts).
hipError_t launch_bc1_search(const float *d_src, int w, int h, int alpha_mode, const float *cw, const Tables *d_tab,
                             void *d_out, int16_t *d_nm, int *d_counts, const int *d_lists, int cap, int max_groups,
                             hipStream_t stream)
{
    const int bw = (w + 3) / 4;
    if (cap == 0) return hipSuccess;
    const int rows16 = (reinterpret_cast<uintptr_t>(d_src) & 15) == 0 && (w & 3) == 0;
    // Each list fills whole wavefronts: at most one partial wave per list.
    const int waves = (cap + 63) / 64 + (ntx_bucket::kCountSlots - 2);
    const int grid = max_groups > 0 && max_groups < waves ? max_groups : waves;
    hipLaunchKernelGGL(bc1_search_kernel, dim3(grid), dim3(64), 0, stream, d_src, w, h, bw, alpha_mode, cw[0], cw[1],
                       cw[2], d_tab, reinterpret_cast<uint2 *>(d_out), reinterpret_cast<short2 *>(d_nm), d_counts,
                       d_lists, cap, rows16);
    return hipGetLastError();
}

int bc1_scratch_ints(int nblocks) { return ntx_bucket::kListsOffset + ntx_bucket::kCountSlots * nblocks; }

int bc1_multi_max_images() { return 1 << (31 - kImageShift); }
int bc1_multi_max_blocks() { return 1 << kImageShift; }

// Pass 1 of the multi-image form: d_map holds nchunks (image, first block)
// pairs, one per workgroup of 256 blocks; the lists (zero tallies on entry)
// take up to cap entries each, cap >= all the images' blocks.
hipError_t launch_bc1_classify_multi(const Bc1Image *d_imgs, const void *d_map, int nchunks, const Tables *d_tab,
                                     int *d_counts, int *d_lists, int cap, hipStream_t stream)
{
    if (nchunks == 0) return hipSuccess;
    hipLaunchKernelGGL(bc1_classify_multi_kernel, dim3(nchunks), dim3(256), 0, stream, d_imgs,
                       static_cast<const int2 *>(d_map), d_tab, d_counts, d_lists, cap);
    return hipGetLastError();
}

// Pass 2 of the multi-image form over the lists pass 1 filled; leaves the
// tallies zero.
hipError_t launch_bc1_search_multi(const Bc1Image *d_imgs, int nimgs, const Tables *d_tab, int *d_counts,
                                   const int *d_lists, int cap, int max_groups, hipStream_t stream)
{
    if (cap == 0) return hipSuccess;
    const int waves = (cap + 63) / 64 + (ntx_bucket::kCountSlots - 2);
    const int grid = max_groups > 0 && max_groups < waves ? max_groups : waves;
    hipLaunchKernelGGL(bc1_search_multi_kernel, dim3(grid), dim3(64), 0, stream, d_imgs, d_tab, d_counts, d_lists, cap,
                       nimgs);
    return hipGetLastError();
}

// icbc Level8 of the *d_count blocks listed in d_list (at most cap, each
// below cap, cap = the image's block count), at most max_groups workgroups.
hipError_t launch_bc1_list(const float *d_src, int w, int h, int alpha_mode, const float *cw, const Tables *d_tab,
                           void *d_out, const int *d_list, const int *d_count, int cap, int max_groups,
                           hipStream_t stream)
{
    if (cap == 0) return hipSuccess;
    const int bw = (w + 3) / 4;
    int grid = (cap + kBlocksPerWG - 1) / kBlocksPerWG;
    if (max_groups > 0 && grid > max_groups) grid = max_groups;
    hipLaunchKernelGGL(bc1_list_kernel<8>, dim3(grid), dim3(kBlocksPerWG * kGroup), 0, stream, d_src, w, h, bw, d_list,
                       d_count, cap, alpha_mode, cw[0], cw[1], cw[2], d_tab, reinterpret_cast<uint2 *>(d_out));
    return hipGetLastError();
}

}  // namespace ntx

// bc1.hip -- BC1 encoder for gfx950, bit-exact with NVTT's CPU path
// (icbc v1.05 compress_dxt1 at Quality_Level8; Level9 for Quality_Production,
// Level1 for Quality_Fastest), icbc.h:3485-3559.
//
// Mapping: one 4x4 block per 16-lane group, four blocks per wave64, sixteen
// per 256-thread workgroup; the per-block device code is in bc1_block.hpp.
#include "bc1_block.hpp"

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

}  // namespace ntx

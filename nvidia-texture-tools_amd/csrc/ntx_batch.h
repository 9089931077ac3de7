// ntx_batch.h -- device-side layout of a batch: many textures, each with its
// own mip chain and block format, run from a work list built on the host
// (batch_api.cpp) in one of two schedules (batch.hip): a launch per mip stage
// then one persistent launch over the compression items, or everything in
// one persistent launch whose mip items complete per-level ready counters.
#pragma once

#include <stdint.h>

namespace ntx {

struct Tables;

enum BatchItemKind : int {
    kItemMip = 0,      // tiles of the next mip level, from this level
    kItemBc1 = 1,      // BC1 blocks of this level
    kItemChannel = 2,  // BC4 / BC5 channel blocks of this level
};

// One mip level of one texture.
struct BatchLevel {
    float *data;        // planar FP32 RGBA: four planes of w * h
    uint8_t *out;       // compressed blocks of this level, block-raster order
    const float *kx;    // polyphase weights producing the next level (mip_kind 1)
    const float *ky;
    const int *lx;      // first source tap per output column / row
    const int *ly;
    int w, h;
    int format;         // NTX_FORMAT_BC1 / BC4 / BC5
    int alpha_mode;
    float cw[3];        // colour weights (BC1)
    int bc1_level;      // BC1: icbc quality level (1, 8 or 9)
    int ready;          // one-launch schedule: counter that completes `data` (-1: complete before launch)
    int target;         // its value once every item writing `data` has finished
    int next;           // index of the next level, -1 for the last one
    int mip_kind;       // 0: fastDownSample, 1: polyphase resize
    int wrap;
    int wx, wy;         // taps per output column / row
};

// One unit of work taken from the queue by a workgroup.
struct BatchItem {
    int kind;   // BatchItemKind
    int level;  // level read (mip: source of the next level; compression: the level itself)
    int first;  // first tile / block / channel job
    int count;  // tiles / blocks / channel jobs
};

struct BatchArgs {
    const BatchItem *items;  // the launch's items
    const BatchLevel *levels;
    const Tables *tab;
    int *counters;    // ready counters of the levels (one-launch schedule)
    int *head;        // queue position of the persistent launch
    int *status;      // bit 0: a dependency wait gave up
    int nitems;
    int tile_shift;   // mip tile: (1 << tile_shift) x (256 >> tile_shift) outputs
    int spin_limit;   // polls before a dependency wait gives up
};

}  // namespace ntx

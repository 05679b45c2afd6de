// fsg_kernels.hpp -- kernel entry points shared by the kernels and the C-ABI.
#pragma once
#include "fsg.h"
#include "fsg_device.hpp"

namespace fsg {

// One world of a rigid-body stamp batch.  Pixels of all worlds are numbered
// consecutively; world i owns [pix_start, pix_start + npix).
struct StampWorld {
    uint64_t *cells;           // current state buffer
    uint64_t *claims;          // spare state buffer, used as claim slots
    uint32_t *flags;           // chunk flags the next tick reads (null: no skipping)
    uint64_t *pixels;          // body pixels, body-major, row-major per body
    const uint32_t *pix_body;  // body index per pixel (0xFFFFFFFF = none)
    const fsg_body *bodies;
    uint8_t *state;            // per pixel: 0 open, 1 placed/found, 2/3 placed over sand/soup
    uint64_t *disp;            // displaced cell per pixel
    uint64_t *disp_xy;         // its position (x | y << 32)
    uint64_t *tgt;             // per pixel: its own cell under the pose, x | y << 32 (x = INT32_MIN: not live)
    uint64_t *save;            // per pixel x 5: the claim slots' contents before the claims (null: not kept)
    int32_t W, H, tiles_x;
    uint32_t npix, pix_start, pad;
};

// Stamp kernels run on a 2D grid: blockIdx.y = world, x over its pixels.
struct StampArgs {
    const StampWorld *worlds;  // device array
    int nworlds;
    uint32_t total;            // pixels over all worlds
    MatTab mat;                // shared material table
    uint64_t *open;            // pixels open after round 0: world << 32 | pixel
    uint32_t *nopen;           // their count
};

template <int SK, int RE, int ZN>
__global__ void k_tick_world(DevWorld w, uint32_t k, int nfull);
template <int RE, int ZN>
__global__ void k_tick_small(DevWorld w, uint32_t k);
template <int RE, int ZN>
__global__ void k_tick_quarter(DevWorld w, uint32_t k);
__global__ void k_tick_batch(const DevWorld *__restrict__ worlds, const uint32_t *__restrict__ tile_start,
                             int nworlds, uint32_t k);
__global__ void k_tick_batch4(const DevWorld *__restrict__ worlds, const uint32_t *__restrict__ tile_start,
                              int nworlds, uint32_t k);
__global__ void k_advance(DevWorld *worlds, int nworlds, uint32_t ticks);
__global__ void k_temperature(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int W, int H,
                              MatTab M, int4 zone);
__global__ void k_wake(DevWorld w, uint32_t k);
template <int RE, int ZN>
__global__ void k_tick_list(DevWorld w, uint32_t k);
__global__ void k_stats(const uint64_t *__restrict__ cells, size_t n, unsigned long long *out);
__global__ void k_fill(uint64_t *cells, size_t n, uint64_t v);
__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16);
__global__ void k_set_cells(uint64_t *cells, int W, int H, const int32_t *xy, const uint64_t *vals, size_t n);
__global__ void k_claims_init(StampArgs a, int save);
__global__ void k_claims_clear(StampArgs a);
__global__ void k_claims_restore(StampArgs a);
// Per-pixel steps of a stamp claim round (k_stamp_step, k_stamp_step_list).
enum StampOp { OP_PROPOSE = 1, OP_RESOLVE, OP_UNPROPOSE, OP_UNRESOLVE, OP_UNFINISH };
__global__ void k_stamp_step(StampArgs a, int op, int d);
__global__ void k_stamp_step_list(StampArgs a, int op, int d);
constexpr int STAMP_CPT = 8;  // pixels per thread of k_stamp_compact (grid x: npix / (256 * STAMP_CPT))
__global__ void k_stamp_compact(StampArgs a);
__global__ void k_zero_u32(uint32_t *p);

} // namespace fsg

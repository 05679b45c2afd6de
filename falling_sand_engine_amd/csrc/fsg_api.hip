// fsg_api.hip -- the C-ABI (include/fsg.h) over the HIP kernels.
//
// Ownership follows ME::world (world.cpp:65, 110-133, 3532-3636): the world
// owns its device buffers and frees them on destroy.  Host buffers are
// caller-owned and copied.  Every entry point returns a status and never
// throws.  All work of a context goes to one HIP stream of default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <new>
#include <unordered_map>
#include <vector>

#include "fsg.h"
#include "fsg_device.hpp"
#include "fsg_kernels.hpp"

#include "fsg_internal.hpp"

using namespace fsg_impl;

namespace fsg_impl {


// Device memory with a per-context cache of freed blocks.  hipFree waits for
// the whole device, which would stall a serving loop that creates and evicts
// worlds while other worlds tick; a freed block instead goes back to its
// context and is handed out again for the same size class.  A context's
// work runs on its one stream, so a reused block is only touched after the
// work that used it before.  The cache is trimmed past cache_cap and emptied
// by fsg_ctx_destroy.
namespace {
std::mutex g_alloc_mu;
std::unordered_map<void *, std::pair<fsg_ctx *, size_t>> g_live;  // block -> (ctx, rounded bytes)

void trim_cache_locked(fsg_ctx *c, size_t keep) {
    while (c->cached > keep && !c->cache.empty()) {
        auto it = std::prev(c->cache.end());  // largest first
        (void)hipFree(it->second);
        c->cached -= it->first;
        c->cache.erase(it);
    }
}
} // namespace

void trim_cache(fsg_ctx *c, size_t keep) {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    trim_cache_locked(c, keep);
}

// Rounded block size: 256-byte steps up to 1 MiB, above that 8 size classes
// per power of two (at most 12.5% over).  Worlds of many sizes come and go in
// a serving loop; with exact sizes almost no freed block was ever reused, and
// the cache overflowed into hipFree calls (~0.4 ms each, device-wide waits).
size_t size_class(size_t bytes) {
    if (bytes <= ((size_t)1 << 20)) return (bytes + 255) & ~(size_t)255;
    const int k = 63 - __builtin_clzll(bytes - 1);  // 2^k < bytes <= 2^(k+1)
    const size_t step = (size_t)1 << (k - 3);
    return (bytes + step - 1) & ~(step - 1);
}

int dev_alloc(fsg_ctx *c, void **p, size_t bytes) {
    *p = nullptr;
    for (int k = c->inject.load(); k > 0;)
        if (c->inject.compare_exchange_weak(k, k - 1)) return FSG_ERR_NOMEM;
    if (bytes == 0) bytes = 8;
    bytes = size_class(bytes);
    {
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        auto it = c->cache.find(bytes);
        if (it != c->cache.end()) {
            *p = it->second;
            c->cache.erase(it);
            c->cached -= bytes;
            g_live[*p] = {c, bytes};
            return FSG_OK;
        }
    }
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        trim_cache(c, 0);  // give the cached blocks back and try once more
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            (void)hipGetLastError();
            return FSG_ERR_NOMEM;
        }
    }
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    g_live[*p] = {c, bytes};
    return FSG_OK;
}

void dev_free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {
        (void)hipFree(p);
        return;
    }
    fsg_ctx *c = it->second.first;
    const size_t bytes = it->second.second;
    g_live.erase(it);
    c->cache.emplace(bytes, p);
    c->cached += bytes;
    if (c->cached > c->cache_cap) trim_cache_locked(c, c->cache_cap);
}

int validate_config(const fsg_config &c) {
    if (c.cell_iter < 1 || c.cell_iter > (uint32_t)fsg::MAX_IT) return FSG_ERR_INVALID;
    if (c.n_materials < 1 || c.n_materials > FSG_MAX_MATERIALS) return FSG_ERR_INVALID;
    if (c.temp_every > 0 && c.temp_phase >= c.temp_every) return FSG_ERR_INVALID;
    const uint8_t roles[6] = {c.mat_air, c.mat_water, c.mat_lava, c.mat_steam, c.mat_stone, c.mat_wood};
    for (uint8_t r : roles)
        if (r >= c.n_materials) return FSG_ERR_INVALID;
    for (uint32_t i = 0; i < c.n_materials; i++) {
        if (c.materials[i].type > FSG_GAS) return FSG_ERR_INVALID;
        if (c.materials[i].rank > 14) return FSG_ERR_INVALID;
    }
    if (c.materials[c.mat_air].type != FSG_AIR) return FSG_ERR_INVALID;
    return FSG_OK;
}

// Material nibble tables and masks (Material fields, game_datastruct.hpp:130-168).
MatTab build_mattab(const fsg_config &c) {
    MatTab M;
    std::memset(&M, 0, sizeof(M));
    for (int id = 0; id < 16; id++) {
        int type = FSG_SOLID, rank = 15, iters = 0, add = 0;
        uint32_t rgba = 0;
        if ((uint32_t)id < c.n_materials) {
            type = c.materials[id].type;
            rank = c.materials[id].rank;
            iters = c.materials[id].iterations;
            add = c.materials[id].add_temp;
            rgba = c.materials[id].rgba;
        }
        const uint64_t sh = (uint64_t)id * 4;
        M.pass |= (uint64_t)(type == FSG_SOLID ? 15 : rank) << sh;
        for (int it = 0; it < fsg::MAX_IT; it++) {
            const bool act = it < iters;
            if ((type == FSG_SAND || type == FSG_SOUP) && act) M.sink[it] |= (uint64_t)rank << sh;
            if (type == FSG_SOUP && act) M.flow[it] |= (uint64_t)rank << sh;
            if (type == FSG_GAS && act) M.gas[it] |= 1u << id;
        }
        if (type == FSG_AIR) M.air |= 1u << id;
        if (type == FSG_SAND) M.sand |= 1u << id;
        if (type == FSG_SOUP) M.soup |= 1u << id;
        M.rgba[id] = rgba;
        M.add_temp[id] = add;
        if (add) M.any_add_temp = 1;
    }
    M.burn = c.wood_burn_temp;
    M.water = c.mat_water;
    M.lava = c.mat_lava;
    M.steam = c.mat_steam;
    M.stone = c.mat_stone;
    M.wood = c.mat_wood;
    M.air_id = c.mat_air;
    M.reactions = c.reactions ? 1 : 0;
    M.cell_iter = c.cell_iter;
    M.temp_every = c.temp_every;
    M.temp_phase = c.temp_phase;
    return M;
}

void fill_pair_table(DevWorld &d) {
    for (int it = 0; it < fsg::MAX_IT; it++)
        for (uint32_t i = 0; i < 256; i++) d.pair[it][i] = (uint16_t)fsg::pair_code(d.mat, it, i >> 4, i & 15u);
}

// The block-class tables of the classification pass (DevWorld::ct), for a
// world of at most 8 materials: the class of every block of 3-bit material
// ids at every iteration, from the pair table.  bit 0: some pair relation
// (K_SINK, K_GRISE, K_HSWAP, K_REACT); bit 1: plain (no gas or reaction in
// the columns, no gas diagonal, no horizontal swap); bits 2-3: K_TSANDACT of
// the top left / right cell.
std::vector<uint32_t> class_tables(const DevWorld &d) {
    std::vector<uint32_t> t((size_t)fsg::MAX_IT * 512, 0u);
    for (int it = 0; it < fsg::MAX_IT; it++)
        for (uint32_t i = 0; i < 4096; i++) {
            const uint32_t i0 = i >> 9, i1 = (i >> 6) & 7u, i2 = (i >> 3) & 7u, i3 = i & 7u;
            auto P = [&](uint32_t a, uint32_t b) { return (uint32_t)d.pair[it][a * 16 + b]; };
            const uint32_t VL = P(i0, i2), VR = P(i1, i3), DL = P(i0, i3), DR = P(i1, i2), HT = P(i0, i1),
                           HB = P(i2, i3);
            const uint32_t rel = (VL | VR | DL | DR | HT | HB) & (fsg::K_SINK | fsg::K_GRISE | fsg::K_HSWAP | fsg::K_REACT);
            const uint32_t other = ((VL | VR) & (fsg::K_GRISE | fsg::K_REACT)) | ((DL | DR) & fsg::K_GRISE) |
                                   ((HT | HB) & fsg::K_HSWAP);
            const uint32_t nib = (rel ? 1u : 0u) | (other ? 0u : 2u) | (((VL >> 8) & 1u) << 2) | (((VR >> 8) & 1u) << 3);
            t[(size_t)it * 512 + (i >> 3)] |= nib << ((i & 7u) * 4);
        }
    return t;
}

bool cell_ok(const fsg_config &c, uint16_t mat) {
    return (mat & 0xFE00u) == 0 && (uint32_t)(mat & 0xFFu) < c.n_materials;
}

inline uint64_t as_u64(const fsg_cell &c) {
    uint64_t v;
    std::memcpy(&v, &c, 8);
    return v;
}

unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

int launch_tick(fsg_world *w) {
    const unsigned tiles = (unsigned)(w->d.tiles_x * w->d.tiles_y);
    // chunk flags left stale by a tick without skipping: process every chunk once
    if (w->d.skip && !w->flags_valid) w->d.force_tick = w->d.tick;
    const bool sk = w->d.skip != 0, re = w->d.mat.reactions != 0;
    // a zone smaller than the world takes the kernels that test it
    const bool zoned = w->d.zx0 > 0 || w->d.zy0 > 0 || w->d.zx1 < w->d.W || w->d.zy1 < w->d.H;
    hipStream_t st = w->ctx->stream;
    if (sk) {
        // the awake chunks listed, then one workgroup per listed chunk; the
        // count of this tick's parity must start at zero (a tick run another
        // way, or a jump of the tick counter, leaves it stale)
        if (!w->list_any && w->list_next != w->d.tick) {
            fsg::k_zero_u32<<<1, 1, 0, st>>>(w->d.awake_n);
            fsg::k_zero_u32<<<1, 1, 0, st>>>(w->d.awake_n + 1);
        }
        fsg::k_wake<<<grid_for(tiles, 256), 256, 0, st>>>(w->d, 0u);
        if (zoned) {
            if (re) fsg::k_tick_list<1, 1><<<tiles, fsg::NT, 0, st>>>(w->d, 0u);
            else fsg::k_tick_list<0, 1><<<tiles, fsg::NT, 0, st>>>(w->d, 0u);
        } else {
            if (re) fsg::k_tick_list<1, 0><<<tiles, fsg::NT, 0, st>>>(w->d, 0u);
            else fsg::k_tick_list<0, 0><<<tiles, fsg::NT, 0, st>>>(w->d, 0u);
        }
        w->list_any = false;
        w->list_next = w->d.tick + 1;
    } else if (4 * tiles <= (unsigned)w->ctx->n_cu) {
        // at most a quarter chunk per compute unit: four workgroups per chunk
        // (512^2 world: 9.7 us a tick instead of 11.2)
        if (zoned) {
            if (re) fsg::k_tick_quarter<1, 1><<<4 * tiles, fsg::QUARTER_NT, 0, st>>>(w->d, 0u);
            else fsg::k_tick_quarter<0, 1><<<4 * tiles, fsg::QUARTER_NT, 0, st>>>(w->d, 0u);
        } else {
            if (re) fsg::k_tick_quarter<1, 0><<<4 * tiles, fsg::QUARTER_NT, 0, st>>>(w->d, 0u);
            else fsg::k_tick_quarter<0, 0><<<4 * tiles, fsg::QUARTER_NT, 0, st>>>(w->d, 0u);
        }
    } else if (tiles <= (unsigned)w->ctx->n_cu) {
        // at most one chunk per compute unit: 1024 threads per chunk (a 512^2
        // world ticks in 11.4 us instead of 16.0; with more chunks than
        // compute units the big workgroups take two rounds and lose)
        if (zoned) {
            if (re) fsg::k_tick_small<1, 1><<<tiles, fsg::SMALL_NT, 0, st>>>(w->d, 0u);
            else fsg::k_tick_small<0, 1><<<tiles, fsg::SMALL_NT, 0, st>>>(w->d, 0u);
        } else {
            if (re) fsg::k_tick_small<1, 0><<<tiles, fsg::SMALL_NT, 0, st>>>(w->d, 0u);
            else fsg::k_tick_small<0, 0><<<tiles, fsg::SMALL_NT, 0, st>>>(w->d, 0u);
        }
    } else {
        // whole chunks in full rounds of resident workgroups (five per CU),
        // the remainder as quarters
        const unsigned slots = 5u * (unsigned)w->ctx->n_cu;
        const unsigned nfull = tiles <= slots ? tiles : tiles / slots * slots;
        const unsigned grid = nfull + 4u * (tiles - nfull);
        if (zoned) {
            if (re) fsg::k_tick_world<0, 1, 1><<<grid, fsg::NT, 0, st>>>(w->d, 0u, (int)nfull);
            else fsg::k_tick_world<0, 0, 1><<<grid, fsg::NT, 0, st>>>(w->d, 0u, (int)nfull);
        } else {
            if (re) fsg::k_tick_world<0, 1, 0><<<grid, fsg::NT, 0, st>>>(w->d, 0u, (int)nfull);
            else fsg::k_tick_world<0, 0, 0><<<grid, fsg::NT, 0, st>>>(w->d, 0u, (int)nfull);
        }
    }
    HIPCHK(hipGetLastError());
    w->d.tick++;
    w->offered += tiles;
    w->flags_valid = w->d.skip != 0;
    w->prev_valid = true;
    return FSG_OK;
}

} // namespace

extern "C" int fsg_rule_eval(const fsg_config *cfg, int it, uint32_t r, uint32_t s, uint64_t q[4], int form) {
    if (!cfg || !q || it < 0 || it >= (int)cfg->cell_iter || validate_config(*cfg)) return FSG_ERR_INVALID;
    const MatTab M = build_mattab(*cfg);
    if (form == 0) return fsg::block_rule(M, it, r, s, q) ? 1 : 0;
    if (form != 1) return FSG_ERR_INVALID;
    // The tick kernel's path: compact cells whose origins are the block's
    // own cells, the pair-table decision or the staged rule for reacting
    // blocks, then the colour of each cell taken from its origin (or the
    // base colour of a reaction product).
    uint32_t c[4];
    for (int k = 0; k < 4; k++) c[k] = (uint32_t)q[k] | ((uint32_t)(k & 1) << 9) | ((uint32_t)(k >> 1) << 12);
    auto code = [&](uint32_t t, uint32_t u) { return fsg::pair_code(M, it, fsg::cid(t), fsg::cid(u)); };
    const uint32_t VL = code(c[0], c[2]), VR = code(c[1], c[3]), DL = code(c[0], c[3]), DR = code(c[1], c[2]);
    const uint32_t HT = code(c[0], c[1]), HTR = code(c[1], c[0]), HB = code(c[2], c[3]);
    bool react = ((VL | VR) & fsg::K_REACT) != 0;
    for (int k = 0; k < 4; k++)
        react = react || (M.reactions && fsg::cbyte(c[k]) == M.wood && fsg::ctemp(c[k]) > M.burn);
    const uint32_t movA = (c[0] >> 8) & 1u, movB = (c[1] >> 8) & 1u, movC = (c[2] >> 8) & 1u, movD = (c[3] >> 8) & 1u;
    if (react) {
        if (!fsg::block_react(M, c)) return FSG_ERR_STATE;  // the kernel queues only blocks that react
    } else {
        if (fsg::block_inert(VL, VR, DL, DR, HT, HB, movA, movB)) return 0;
        fsg::apply_decision(fsg::block_decide(VL, VR, DL, DR, HT, HTR, HB, movA, movB, movC, movD, r, s), c);
    }
    uint64_t o[4];
    bool ch = false;
    for (int k = 0; k < 4; k++) {
        const uint32_t org = ((c[k] >> 9) & 1u) | (((c[k] >> 12) & 1u) << 1);
        const uint32_t colour = (c[k] & fsg::C_XF) ? M.rgba[c[k] & 15u] : (uint32_t)(q[org] >> 32);
        o[k] = (uint64_t)(c[k] & fsg::C_KEEP) | ((uint64_t)colour << 32);
    }
    for (int k = 0; k < 4; k++) {
        ch = ch || o[k] != q[k];
        q[k] = o[k];
    }
    return ch ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
extern "C" int fsg_ctx_create(const fsg_ctx_config *cfg, fsg_ctx **out) {
    if (!out) return FSG_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return FSG_ERR_DEVICE;
    }
    const int dev = cfg ? cfg->device : 0;
    if (dev < 0 || dev >= ndev) return FSG_ERR_INVALID;
    fsg_ctx *c = new (std::nothrow) fsg_ctx();
    if (!c) return FSG_ERR_NOMEM;
    c->device = dev;
    c->budget = cfg ? cfg->hbm_budget : 0;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess ||
        hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess ||
        hipEventCreateWithFlags(&c->staged, hipEventDisableTiming) != hipSuccess) {
        delete c;
        return FSG_ERR_DEVICE;
    }
    // the fixed device array of fsg_world_set_cells_noalloc: its address on this device
    if (fsg_impl::noalloc_setup(dev) != FSG_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return FSG_ERR_DEVICE;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
        c->n_cu = ncu;
    // freed blocks kept for reuse: up to a quarter of the device's memory
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b / 4 > c->cache_cap) c->cache_cap = total_b / 4;
    *out = c;
    return FSG_OK;
}

extern "C" int fsg_ctx_destroy(fsg_ctx *c) {
    if (!c) return FSG_ERR_INVALID;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    dev_free(c->d_worlds);
    dev_free(c->d_tile_start);
    dev_free(c->d_open);
    dev_free(c->d_nopen);
    trim_cache(c, 0);
    (void)hipEventDestroy(c->e0);
    (void)hipEventDestroy(c->e1);
    (void)hipEventDestroy(c->t0);
    (void)hipEventDestroy(c->t1);
    (void)hipEventDestroy(c->staged);
    if (c->h_worlds) (void)hipHostFree(c->h_worlds);
    if (c->h_tile_start) (void)hipHostFree(c->h_tile_start);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return FSG_OK;
}

extern "C" int fsg_ctx_usage(fsg_ctx *c, uint64_t *bytes) {
    if (!c || !bytes) return FSG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    *bytes = c->used;
    return FSG_OK;
}

extern "C" int fsg_ctx_inject_alloc_failures(fsg_ctx *c, int n) {
    if (!c || n < 0) return FSG_ERR_INVALID;
    c->inject = n;
    return FSG_OK;
}

extern "C" int fsg_ctx_timer_start(fsg_ctx *c) {
    if (!c) return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventRecord(c->t0, c->stream));
    return FSG_OK;
}

extern "C" int fsg_ctx_timer_stop(fsg_ctx *c, double *ms) {
    if (!c || !ms) return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventRecord(c->t1, c->stream));
    HIPCHK(hipEventSynchronize(c->t1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->t0, c->t1));
    *ms = f;
    return FSG_OK;
}

extern "C" int fsg_ctx_copy_timed(fsg_ctx *c, size_t bytes, int reps, double *ms) {
    if (!c || bytes < 16 || reps < 1) return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    const size_t n16 = bytes / 16;
    if (n16 > (size_t)0x7FFFFFFF * 256) return FSG_ERR_INVALID;
    void *a = nullptr, *b = nullptr;
    if (dev_alloc(c, &a, n16 * 16)) return FSG_ERR_NOMEM;
    if (dev_alloc(c, &b, n16 * 16)) {
        dev_free(a);
        return FSG_ERR_NOMEM;
    }
    int rc = FSG_OK;
    const unsigned grid = grid_for(n16, 256);
    (void)hipMemsetAsync(a, 1, n16 * 16, c->stream);
    fsg::k_copy<<<grid, 256, 0, c->stream>>>((const uint4 *)a, (uint4 *)b, n16);  // warm-up
    std::unique_lock<std::mutex> tl(c->time_mu);
    if (hipEventRecord(c->e0, c->stream) != hipSuccess) rc = FSG_ERR_DEVICE;
    for (int r = 0; r < reps && !rc; r++)
        fsg::k_copy<<<grid, 256, 0, c->stream>>>((const uint4 *)a, (uint4 *)b, n16);
    float f = 0.f;
    if (!rc && (hipEventRecord(c->e1, c->stream) != hipSuccess || hipEventSynchronize(c->e1) != hipSuccess ||
                hipEventElapsedTime(&f, c->e0, c->e1) != hipSuccess))
        rc = FSG_ERR_DEVICE;
    (void)hipStreamSynchronize(c->stream);
    dev_free(a);
    dev_free(b);
    if (!rc && ms) *ms = f / reps;
    return rc;
}

// ---------------------------------------------------------------------------
// World
// ---------------------------------------------------------------------------
extern "C" int fsg_world_create(fsg_ctx *c, int W, int H, uint64_t seed, const fsg_config *cfg,
                                fsg_world **out) {
    if (!c || !out || W <= 0 || H <= 0 || W > FSG_MAX_SIDE || H > FSG_MAX_SIDE) return FSG_ERR_INVALID;
    *out = nullptr;
    fsg_config conf;
    if (cfg) conf = *cfg;
    else fsg_default_config(&conf);
    if (int rc = validate_config(conf)) return rc;

    const int tx = (W + fsg::TW - 1) / fsg::TW, ty = (H + fsg::TH - 1) / fsg::TH;
    const size_t cells = (size_t)W * H;
    const size_t bytes = world_state_bytes(W, H, conf.n_materials);
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->budget && c->used + bytes > c->budget) return FSG_ERR_BUDGET;
        c->used += bytes;  // reserve
    }
    fsg_world *w = new (std::nothrow) fsg_world();
    auto fail = [&](int rc) {
        if (w) {
            dev_free(w->d.buf[0]);
            dev_free(w->d.buf[1]);
            dev_free(w->d.flags[0]);
            dev_free(w->d.flags[1]);
            dev_free(w->d.awake);
            dev_free(w->d.awake_n);
            dev_free(const_cast<uint32_t *>(w->d.ct));
            dev_free(w->d_stats);
            delete w;
        }
        std::lock_guard<std::mutex> lk(c->mu);
        c->used -= bytes;
        return rc;
    };
    if (!w) return fail(FSG_ERR_NOMEM);
    w->ctx = c;
    w->cfg = conf;
    w->bytes = bytes;
    if (hipSetDevice(c->device) != hipSuccess) return fail(FSG_ERR_DEVICE);
    void *p;
    if (dev_alloc(c, &p, cells * 8)) return fail(FSG_ERR_NOMEM);
    w->d.buf[0] = (uint64_t *)p;
    if (dev_alloc(c, &p, cells * 8)) return fail(FSG_ERR_NOMEM);
    w->d.buf[1] = (uint64_t *)p;
    if (dev_alloc(c, &p, (size_t)tx * ty * 4)) return fail(FSG_ERR_NOMEM);
    w->d.flags[0] = (uint32_t *)p;
    if (dev_alloc(c, &p, (size_t)tx * ty * 4)) return fail(FSG_ERR_NOMEM);
    w->d.flags[1] = (uint32_t *)p;
    if (dev_alloc(c, &p, (size_t)tx * ty * 4)) return fail(FSG_ERR_NOMEM);
    w->d.awake = (uint32_t *)p;
    if (dev_alloc(c, &p, 16)) return fail(FSG_ERR_NOMEM);
    w->d.awake_n = (uint32_t *)p;
    if (dev_alloc(c, &p, 18 * 8)) return fail(FSG_ERR_NOMEM);
    w->d_stats = (unsigned long long *)p;
    w->d.W = W;
    w->d.H = H;
    w->d.tiles_x = tx;
    w->d.tiles_y = ty;
    w->d.seed = seed;
    w->d.tick = 0;
    w->d.force_tick = 0;
    w->d.skip = conf.chunk_skip ? 1 : 0;
    w->d.processed = w->d_stats + 17;
    w->d.zx0 = 0;
    w->d.zy0 = 0;
    w->d.zx1 = W;
    w->d.zy1 = H;
    w->d.mat = build_mattab(conf);
    fill_pair_table(w->d);
    w->d.ct = nullptr;
    if (conf.n_materials <= 8) {
        w->ct_host = class_tables(w->d);  // kept until the world goes (the copy is asynchronous)
        if (dev_alloc(c, &p, w->ct_host.size() * 4)) return fail(FSG_ERR_NOMEM);
        w->d.ct = (const uint32_t *)p;
        if (hipMemcpyAsync(p, w->ct_host.data(), w->ct_host.size() * 4, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess)
            return fail(FSG_ERR_DEVICE);
    }
    const uint64_t air = conf.mat_air;
    fsg::k_fill<<<grid_for(cells, 256), 256, 0, c->stream>>>(w->d.buf[0], cells, air);
    fsg::k_fill<<<grid_for(cells, 256), 256, 0, c->stream>>>(w->d.buf[1], cells, air);
    // no wait: everything later on this world is ordered behind these on the
    // context's stream (a serving loop creates worlds while a step runs)
    if (hipMemsetAsync(w->d.flags[0], 0, (size_t)tx * ty * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(w->d.flags[1], 0, (size_t)tx * ty * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(w->d.awake_n, 0, 16, c->stream) != hipSuccess ||
        hipMemsetAsync(w->d_stats, 0, 18 * 8, c->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return fail(FSG_ERR_DEVICE);
    *out = w;
    return FSG_OK;
}

extern "C" int fsg_world_destroy(fsg_world *w) {
    if (!w) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    (void)hipSetDevice(c->device);
    // The world's device blocks go back to the context's cache, whose reuse is
    // ordered behind any pending work on the one stream, so destroying needs no
    // wait -- except for a captured graph over this world, which may be running.
    {
        // the graph state belongs to the batch launches (batch_mu)
        std::lock_guard<std::mutex> bl(c->batch_mu);
        if (std::find(c->graph_key.begin(), c->graph_key.end(), w) != c->graph_key.end()) {
            (void)hipStreamSynchronize(c->stream);  // the graph may be running
            // a graph captured over this world must not be replayed
            if (c->graph) (void)hipGraphExecDestroy(c->graph);
            c->graph = nullptr;
            c->graph_key.clear();
        }
    }
    {
        std::lock_guard<std::mutex> lk(c->mu);
        c->used -= w->bytes;
    }
    dev_free(w->d.buf[0]);
    dev_free(w->d.buf[1]);
    dev_free(w->d.flags[0]);
    dev_free(w->d.flags[1]);
    dev_free(w->d.awake);
    dev_free(w->d.awake_n);
    dev_free(const_cast<uint32_t *>(w->d.ct));
    dev_free(w->d_stats);
    dev_free(w->d_pixels);
    dev_free(w->d_pixbody);
    dev_free(w->d_bodies);
    dev_free(w->d_state);
    dev_free(w->d_disp);
    dev_free(w->d_dispxy);
    dev_free(w->d_tgt);
    dev_free(w->d_save);
    dev_free(w->d_mesh_regs);
    dev_free(w->d_mesh_scratch);
    dev_free(w->d_mesh_pool);
    dev_free(w->d_part);
    dev_free(w->d_part2);
    dev_free(w->d_part_alive);
    dev_free(w->d_part_cnt);
    dev_free(w->d_pround);
    for (uint32_t *l : w->d_layer) dev_free(l);
    dev_free(w->d_tally);
    dev_free(w->d_isl);
    dev_free(w->d_isl_pix);
    dev_free(w->d_gen);
    dev_free(w->d_sw);
    for (void *p : {w->d_pack_slots, w->d_pack_img, w->d_pack_small, w->d_pack_aux, w->d_pack_in, w->d_pack_raw,
                    w->d_pack_dec})
        dev_free(p);
    delete w;
    return FSG_OK;
}

extern "C" int fsg_world_info(fsg_world *w, int *width, int *height, uint64_t *tick) {
    if (!w) return FSG_ERR_INVALID;
    if (width) *width = w->d.W;
    if (height) *height = w->d.H;
    if (tick) *tick = w->d.tick;
    return FSG_OK;
}

extern "C" int fsg_world_set_tick(fsg_world *w, uint64_t tick) {
    if (!w) return FSG_ERR_INVALID;
    HIPCHK(hipStreamSynchronize(w->ctx->stream));
    if ((tick & 1) != (w->d.tick & 1)) std::swap(w->d.buf[0], w->d.buf[1]);
    w->d.tick = tick;
    w->d.force_tick = tick;
    w->prev_valid = false;
    return FSG_OK;
}

extern "C" int fsg_world_upload(fsg_world *w, const fsg_cell *cells, size_t n) {
    if (!w || !cells || n != (size_t)w->d.W * w->d.H) return FSG_ERR_INVALID;
    for (size_t i = 0; i < n; i++)
        if (!cell_ok(w->cfg, cells[i].mat)) return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(w->ctx->device));
    HIPCHK(hipMemcpyAsync(w->d.buf[w->d.tick & 1], cells, n * 8, hipMemcpyHostToDevice, w->ctx->stream));
    HIPCHK(hipStreamSynchronize(w->ctx->stream));
    w->d.force_tick = w->d.tick;
    w->stamped = false;
    w->prev_valid = false;
    return FSG_OK;
}

extern "C" int fsg_world_download(fsg_world *w, fsg_cell *out, int x0, int y0, int rw, int rh) {
    if (!w || !out || rw <= 0 || rh <= 0 || x0 < 0 || y0 < 0 || x0 + rw > w->d.W || y0 + rh > w->d.H)
        return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(w->ctx->device));
    const uint64_t *src = w->d.buf[w->d.tick & 1] + (size_t)y0 * w->d.W + x0;
    HIPCHK(hipMemcpy2DAsync(out, (size_t)rw * 8, src, (size_t)w->d.W * 8, (size_t)rw * 8, (size_t)rh,
                            hipMemcpyDeviceToHost, w->ctx->stream));
    HIPCHK(hipStreamSynchronize(w->ctx->stream));
    return FSG_OK;
}

extern "C" int fsg_world_set_cells(fsg_world *w, const fsg_cell_edit *edits, size_t n) {
    if (!w || (!edits && n)) return FSG_ERR_INVALID;
    if (n == 0) return FSG_OK;
    for (size_t i = 0; i < n; i++)
        if (!cell_ok(w->cfg, edits[i].cell.mat)) return FSG_ERR_INVALID;
    // last edit of a position wins (array order), like consecutive setTile calls
    std::unordered_map<uint64_t, size_t> last;
    last.reserve(n * 2);
    for (size_t i = 0; i < n; i++) {
        const fsg_cell_edit &e = edits[i];
        if (e.x < 0 || e.y < 0 || e.x >= w->d.W || e.y >= w->d.H) continue;
        last[(uint64_t)(uint32_t)e.y << 32 | (uint32_t)e.x] = i;
    }
    std::vector<int32_t> xy;
    std::vector<uint64_t> vals;
    xy.reserve(2 * last.size());
    vals.reserve(last.size());
    for (auto &kv : last) {
        xy.push_back(edits[kv.second].x);
        xy.push_back(edits[kv.second].y);
        vals.push_back(as_u64(edits[kv.second].cell));
    }
    if (vals.empty()) return FSG_OK;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    void *dxy = nullptr, *dv = nullptr;
    if (dev_alloc(c, &dxy, xy.size() * 4)) return FSG_ERR_NOMEM;
    if (dev_alloc(c, &dv, vals.size() * 8)) {
        dev_free(dxy);
        return FSG_ERR_NOMEM;
    }
    int rc = FSG_OK;
    if (hipMemcpyAsync(dxy, xy.data(), xy.size() * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(dv, vals.data(), vals.size() * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = FSG_ERR_DEVICE;
    if (!rc) {
        fsg::k_set_cells<<<grid_for(vals.size(), 256), 256, 0, c->stream>>>(
            w->d.buf[w->d.tick & 1], w->d.W, w->d.H, (const int32_t *)dxy, (const uint64_t *)dv, vals.size());
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) rc = FSG_ERR_DEVICE;
    }
    dev_free(dxy);
    dev_free(dv);
    w->d.force_tick = w->d.tick;
    return rc;
}

// ticks ticks of one world, one launch each.  (A 512^2 world's tick is the
// kernel's own latency -- a dozen barriered phases over 64 workgroups -- not
// launch overhead: replaying the batch path's 16-tick graphs for it was
// slower, 20.5 vs 17.0 ms for 1000 ticks.)
int step_world(fsg_world *w, int ticks) {
    for (int t = 0; t < ticks; t++)
        if (int rc = launch_tick(w)) return rc;
    return FSG_OK;
}

extern "C" int fsg_world_step(fsg_world *w, int ticks) {
    if (!w || ticks < 0) return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(w->ctx->device));
    return step_world(w, ticks);
}

extern "C" int fsg_world_step_timed(fsg_world *w, int ticks, double *ms) {
    if (!w || ticks < 0) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> tl(c->time_mu);
    HIPCHK(hipEventRecord(c->e0, c->stream));
    if (int rc = step_world(w, ticks)) return rc;
    HIPCHK(hipEventRecord(c->e1, c->stream));
    HIPCHK(hipEventSynchronize(c->e1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->e0, c->e1));
    if (ms) *ms = f;
    return FSG_OK;
}

extern "C" int fsg_world_set_tick_zone(fsg_world *w, int32_t x, int32_t y, int32_t width, int32_t height) {
    if (!w || x < 0 || y < 0 || width < 1 || height < 1 || width > w->d.W - x || height > w->d.H - y)
        return FSG_ERR_INVALID;
    w->d.zx0 = x;
    w->d.zy0 = y;
    w->d.zx1 = x + width;
    w->d.zy1 = y + height;
    // a grown zone wakes chunks that slept outside it
    w->d.force_tick = w->d.tick;
    return FSG_OK;
}

extern "C" int fsg_world_get_tick_zone(fsg_world *w, int32_t *x, int32_t *y, int32_t *width, int32_t *height) {
    if (!w) return FSG_ERR_INVALID;
    if (x) *x = w->d.zx0;
    if (y) *y = w->d.zy0;
    if (width) *width = w->d.zx1 - w->d.zx0;
    if (height) *height = w->d.zy1 - w->d.zy0;
    return FSG_OK;
}

extern "C" int fsg_world_tick_temperature(fsg_world *w) {
    if (!w) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)w->d.W * w->d.H;
    uint64_t *cur = w->d.buf[w->d.tick & 1], *spare = w->d.buf[(w->d.tick + 1) & 1];
    fsg::k_temperature<<<grid_for(n, fsg::NT), fsg::NT, 0, c->stream>>>(
        cur, spare, w->d.W, w->d.H, w->d.mat, make_int4(w->d.zx0, w->d.zy0, w->d.zx1, w->d.zy1));
    HIPCHK(hipGetLastError());
    // the result becomes the current state without advancing the tick
    std::swap(w->d.buf[0], w->d.buf[1]);
    w->d.force_tick = w->d.tick;
    return FSG_OK;
}

extern "C" int fsg_world_sync(fsg_world *w) {
    if (!w) return FSG_ERR_INVALID;
    HIPCHK(hipStreamSynchronize(w->ctx->stream));
    return FSG_OK;
}

namespace fsg_impl {
// Enqueues the census of w's grid (counts [0, 16), checksum [16], processed
// chunks [17]) into host[18]; valid after the stream is synchronised.
int stats_enqueue(fsg_world *w, unsigned long long *host) {
    fsg_ctx *c = w->ctx;
    const size_t n = (size_t)w->d.W * w->d.H;
    HIPCHK(hipMemsetAsync(w->d_stats, 0, 17 * 8, c->stream));
    const unsigned blocks = (unsigned)std::min<size_t>(grid_for(n, fsg::NT), 4096);
    fsg::k_stats<<<blocks, fsg::NT, 0, c->stream>>>(w->d.buf[w->d.tick & 1], n, w->d_stats);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, w->d_stats, 18 * 8, hipMemcpyDeviceToHost, c->stream));
    return FSG_OK;
}
}  // namespace fsg_impl

extern "C" int fsg_world_stats(fsg_world *w, fsg_stats *out) {
    if (!w || !out) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    unsigned long long h[18];
    if (int rc = fsg_impl::stats_enqueue(w, h)) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    std::memset(out, 0, sizeof(*out));
    out->tick = w->d.tick;
    for (int i = 0; i < 16; i++) out->count[i] = h[i];
    out->checksum = h[16];
    out->tiles_total = (uint64_t)w->d.tiles_x * w->d.tiles_y;
    out->tiles_processed = w->d.skip ? h[17] : w->offered;
    out->tiles_offered = w->offered;
    return FSG_OK;
}

extern "C" int fsg_world_reset_activity(fsg_world *w) {
    if (!w) return FSG_ERR_INVALID;
    HIPCHK(hipMemsetAsync(w->d_stats + 17, 0, 8, w->ctx->stream));
    HIPCHK(hipStreamSynchronize(w->ctx->stream));
    w->offered = 0;
    return FSG_OK;
}

// ---------------------------------------------------------------------------
// Rigid-body pixels
// ---------------------------------------------------------------------------
namespace fsg_impl {

// The world's entry in a stamp batch (pixels numbered from pix_start).
fsg::StampWorld stamp_world_of(fsg_world *w, uint32_t pix_start) {
    fsg::StampWorld sw;
    const uint64_t t = w->d.tick;
    sw.cells = w->d.buf[t & 1];
    sw.claims = w->d.buf[(t + 1) & 1];
    sw.flags = w->flags_valid ? w->d.flags[(t + 1) & 1] : nullptr;  // flags the next tick reads
    sw.pixels = w->d_pixels;
    sw.pix_body = w->d_pixbody;
    sw.bodies = w->d_bodies;
    sw.state = w->d_state;
    sw.disp = w->d_disp;
    sw.disp_xy = w->d_dispxy;
    sw.tgt = w->d_tgt;
    sw.save = nullptr;
    sw.W = w->d.W;
    sw.H = w->d.H;
    sw.tiles_x = w->d.tiles_x;
    sw.npix = (uint32_t)w->npix;
    sw.pix_start = pix_start;
    sw.pad = 0;
    return sw;
}

// Claim rounds over a batch of worlds whose table is already on the device.
// Round 0 runs over every pixel (y: world, x: its pixels); the pixels it
// leaves open are then compacted into one list that rounds 1-4 (and the
// unstamp's last step) walk, so those rounds cost what is left, not npix.
int stamp_rounds(fsg_ctx *c, const fsg::StampWorld *d_table, int n, uint32_t total, uint32_t max_npix,
                 const MatTab &M, bool unstamp, bool save) {
    if (total == 0 || max_npix == 0) return FSG_OK;
    std::lock_guard<std::mutex> lk(c->stamp_mu);
    if (!c->d_nopen) {
        void *p;
        if (dev_alloc(c, &p, sizeof(uint32_t))) return FSG_ERR_NOMEM;
        c->d_nopen = (uint32_t *)p;
    }
    if (c->open_cap < total) {
        dev_free(c->d_open);
        c->d_open = nullptr;
        c->open_cap = 0;
        void *p;
        if (dev_alloc(c, &p, (size_t)total * sizeof(uint64_t))) return FSG_ERR_NOMEM;
        c->d_open = (uint64_t *)p;
        c->open_cap = total;
    }
    fsg::StampArgs a;
    a.worlds = d_table;
    a.nworlds = n;
    a.total = total;
    a.mat = M;
    a.open = c->d_open;
    a.nopen = c->d_nopen;
    const int propose = unstamp ? fsg::OP_UNPROPOSE : fsg::OP_PROPOSE;
    const int resolve = unstamp ? fsg::OP_UNRESOLVE : fsg::OP_RESOLVE;
    const dim3 g(grid_for(max_npix, 256), (unsigned)n);  // y: world, x: its pixels
    // every target slot is cleared up front; with save its old contents are
    // copied out first and put back after the last round
    fsg::k_claims_init<<<g, 256, 0, c->stream>>>(a, save ? 1 : 0);
    if (save) fsg::k_claims_clear<<<g, 256, 0, c->stream>>>(a);
    fsg::k_stamp_step<<<g, 256, 0, c->stream>>>(a, propose, 0);
    fsg::k_stamp_step<<<g, 256, 0, c->stream>>>(a, resolve, 0);
    fsg::k_zero_u32<<<1, 1, 0, c->stream>>>(a.nopen);
    const dim3 gc(grid_for(max_npix, 256 * fsg::STAMP_CPT), (unsigned)n);
    fsg::k_stamp_compact<<<gc, 256, 0, c->stream>>>(a);
    // the list holds at most `total` entries; the loop kernels stride over it
    const unsigned gl = std::min(grid_for(total, 256), 1024u);
    for (int d = 1; d < 5; d++) {
        fsg::k_stamp_step_list<<<gl, 256, 0, c->stream>>>(a, propose, d);
        fsg::k_stamp_step_list<<<gl, 256, 0, c->stream>>>(a, resolve, d);
    }
    if (unstamp) fsg::k_stamp_step_list<<<gl, 256, 0, c->stream>>>(a, fsg::OP_UNFINISH, 0);
    if (save) fsg::k_claims_restore<<<g, 256, 0, c->stream>>>(a);
    HIPCHK(hipGetLastError());
    return FSG_OK;
}

// One-world batch for the single-world entry points.
int stamp_single(fsg_world *w, bool unstamp) {
    fsg_ctx *c = w->ctx;
    if (!w->d_sw) {
        void *p;
        if (dev_alloc(c, &p, sizeof(fsg::StampWorld))) return FSG_ERR_NOMEM;
        w->d_sw = (fsg::StampWorld *)p;
    }
    if (w->save_cap < w->npix) {  // five saved claim slots per pixel
        dev_free(w->d_save);
        w->d_save = nullptr;
        w->save_cap = 0;
        void *p;
        if (dev_alloc(c, &p, w->npix * 5 * 8)) return FSG_ERR_NOMEM;
        w->d_save = (uint64_t *)p;
        w->save_cap = w->npix;
    }
    w->sw_host = stamp_world_of(w, 0);
    // the spare buffer's claim slots are put back afterwards, so it still
    // holds the state before the last tick (the texture pass's census)
    w->sw_host.save = w->d_save;
    HIPCHK(hipMemcpyAsync(w->d_sw, &w->sw_host, sizeof(fsg::StampWorld), hipMemcpyHostToDevice, c->stream));
    return stamp_rounds(c, w->d_sw, 1, (uint32_t)w->npix, (uint32_t)w->npix, w->d.mat, unstamp, true);
}

int ensure_stamp_buffers(fsg_world *w, size_t npix, int nb) {
    fsg_ctx *c = w->ctx;
    void *p;
    if (npix > w->pix_cap) {
        dev_free(w->d_pixels);
        dev_free(w->d_pixbody);
        dev_free(w->d_state);
        dev_free(w->d_disp);
        dev_free(w->d_dispxy);
        dev_free(w->d_tgt);
        w->d_pixels = nullptr;
        w->d_pixbody = nullptr;
        w->d_state = nullptr;
        w->d_disp = w->d_dispxy = w->d_tgt = nullptr;
        w->pix_cap = 0;
        if (dev_alloc(c, &p, npix * 8)) return FSG_ERR_NOMEM;
        w->d_pixels = (uint64_t *)p;
        if (dev_alloc(c, &p, npix * 4)) return FSG_ERR_NOMEM;
        w->d_pixbody = (uint32_t *)p;
        if (dev_alloc(c, &p, npix)) return FSG_ERR_NOMEM;
        w->d_state = (uint8_t *)p;
        if (dev_alloc(c, &p, npix * 8)) return FSG_ERR_NOMEM;
        w->d_disp = (uint64_t *)p;
        if (dev_alloc(c, &p, npix * 8)) return FSG_ERR_NOMEM;
        w->d_dispxy = (uint64_t *)p;
        if (dev_alloc(c, &p, npix * 8)) return FSG_ERR_NOMEM;
        w->d_tgt = (uint64_t *)p;
        w->pix_cap = npix;
    }
    if (nb > w->body_cap) {
        dev_free(w->d_bodies);
        w->d_bodies = nullptr;
        w->body_cap = 0;
        if (dev_alloc(c, &p, (size_t)nb * sizeof(fsg_body))) return FSG_ERR_NOMEM;
        w->d_bodies = (fsg_body *)p;
        w->body_cap = nb;
    }
    return FSG_OK;
}

// Validates bodies/pixels and uploads them (with the pixel -> body map).
// The host copies must stay valid until the stream is synchronised.
int set_bodies(fsg_world *w, const fsg_body *bodies, int nb, const fsg_cell *pixels, size_t npix) {
    if (npix >= 0xFFFFFFFFull) return FSG_ERR_INVALID;
    w->pix_body_host.assign(npix, 0xFFFFFFFFu);
    std::vector<uint32_t> &pix_body = w->pix_body_host;
    for (int b = 0; b < nb; b++) {
        const fsg_body &B = bodies[b];
        if (B.mat_w <= 0 || B.mat_h <= 0 || B.mat_w > 4096 || B.mat_h > 4096) return FSG_ERR_INVALID;
        const size_t cnt = (size_t)B.mat_w * B.mat_h;
        if ((size_t)B.pixel_offset + cnt > npix) return FSG_ERR_INVALID;
        for (size_t k = 0; k < cnt; k++) {
            if (pix_body[B.pixel_offset + k] != 0xFFFFFFFFu) return FSG_ERR_INVALID;  // overlapping bodies
            pix_body[B.pixel_offset + k] = (uint32_t)b;
        }
    }
    for (size_t p = 0; p < npix; p++)
        if (!cell_ok(w->cfg, pixels[p].mat)) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (int rc = ensure_stamp_buffers(w, npix ? npix : 1, nb ? nb : 1)) return rc;
    w->npix = npix;
    w->nbodies = nb;
    w->bodies.assign(bodies, bodies + nb);
    if (npix) {
        HIPCHK(hipMemcpyAsync(w->d_pixels, pixels, npix * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(w->d_pixbody, pix_body.data(), npix * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (nb) HIPCHK(hipMemcpyAsync(w->d_bodies, w->bodies.data(), (size_t)nb * sizeof(fsg_body), hipMemcpyHostToDevice, c->stream));
    return FSG_OK;
}

} // namespace

extern "C" int fsg_world_stamp_bodies(fsg_world *w, const fsg_body *bodies, int nb, const fsg_cell *pixels,
                                      size_t npix, fsg_stamp_result *res) {
    if (!w || nb < 0 || (nb && !bodies) || (npix && !pixels)) return FSG_ERR_INVALID;
    if (w->stamped) return FSG_ERR_STATE;
    if (int rc = set_bodies(w, bodies, nb, pixels, npix)) return rc;
    const std::vector<uint32_t> &pix_body = w->pix_body_host;
    fsg_ctx *c = w->ctx;
    if (npix)
        if (int rc = stamp_single(w, false)) return rc;
    std::vector<uint8_t> state(npix);
    std::vector<uint64_t> disp(npix), dispxy(npix);
    if (npix) {
        HIPCHK(hipMemcpyAsync(state.data(), w->d_state, npix, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(disp.data(), w->d_disp, npix * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(dispxy.data(), w->d_dispxy, npix * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    w->stamped = true;
    if (res) {
        uint32_t placed = 0, nd = 0, unplaced = 0;
        if (res->hits_sand) std::memset(res->hits_sand, 0, sizeof(uint32_t) * (size_t)nb);
        if (res->hits_soup) std::memset(res->hits_soup, 0, sizeof(uint32_t) * (size_t)nb);
        for (size_t p = 0; p < npix; p++) {
            const uint32_t b = pix_body[p];
            const bool live = b != 0xFFFFFFFFu && bodies[b].enabled &&
                              !((w->d.mat.air >> (pixels[p].mat & 15)) & 1u);
            if (!live) continue;
            if (!state[p]) {
                unplaced++;
                continue;
            }
            placed++;
            if (state[p] >= 2) {
                if (res->displaced_out && nd < res->displaced_cap) {
                    res->displaced_out[nd].x = (int32_t)(uint32_t)dispxy[p];
                    res->displaced_out[nd].y = (int32_t)(uint32_t)(dispxy[p] >> 32);
                    std::memcpy(&res->displaced_out[nd].cell, &disp[p], 8);
                }
                nd++;
                if (state[p] == 2 && res->hits_sand) res->hits_sand[b]++;
                if (state[p] == 3 && res->hits_soup) res->hits_soup[b]++;
            }
        }
        res->placed = placed;
        res->unplaced = unplaced;
        res->displaced = nd;
    }
    return FSG_OK;
}

extern "C" int fsg_world_unstamp_bodies(fsg_world *w, fsg_cell *pixels_out, size_t npix) {
    if (!w || (npix && !pixels_out)) return FSG_ERR_INVALID;
    if (!w->stamped) return FSG_ERR_STATE;
    if (npix != w->npix) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (npix) {
        if (int rc = stamp_single(w, true)) return rc;
        HIPCHK(hipMemcpyAsync(pixels_out, w->d_pixels, npix * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    w->stamped = false;
    return FSG_OK;
}

// ---------------------------------------------------------------------------
// Batching
// ---------------------------------------------------------------------------
namespace fsg_impl {

int batch_prepare(fsg_ctx *c, fsg_world *const *ws, int n, bool *graph_ok) {
    for (int i = 0; i < n; i++)
        if (!ws[i] || ws[i]->ctx != c) return FSG_ERR_INVALID;
    // the pinned staging may still feed the previous batch's upload
    if (c->staged_pending) {
        HIPCHK(hipEventSynchronize(c->staged));
        c->staged_pending = false;
    }
    if (n > c->batch_cap) {
        dev_free(c->d_worlds);
        dev_free(c->d_tile_start);
        if (c->h_worlds) (void)hipHostFree(c->h_worlds);
        if (c->h_tile_start) (void)hipHostFree(c->h_tile_start);
        c->d_worlds = nullptr;
        c->d_tile_start = nullptr;
        c->h_worlds = nullptr;
        c->h_tile_start = nullptr;
        c->batch_cap = 0;
        void *p;
        if (dev_alloc(c, &p, (size_t)n * sizeof(DevWorld))) return FSG_ERR_NOMEM;
        c->d_worlds = (DevWorld *)p;
        if (dev_alloc(c, &p, (size_t)(n + 1) * 4)) return FSG_ERR_NOMEM;
        c->d_tile_start = (uint32_t *)p;
        if (hipHostMalloc(&p, (size_t)n * sizeof(DevWorld)) != hipSuccess) return FSG_ERR_NOMEM;
        c->h_worlds = (DevWorld *)p;
        if (hipHostMalloc(&p, (size_t)(n + 1) * 4) != hipSuccess) return FSG_ERR_NOMEM;
        c->h_tile_start = (uint32_t *)p;
        c->batch_cap = n;
        if (c->graph) (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
        c->graph_key.clear();
    }
    DevWorld *dw = c->h_worlds;
    std::vector<uint32_t> start(n + 1);
    uint64_t total = 0;
    for (int i = 0; i < n; i++) {
        dw[i] = ws[i]->d;
        // the tuning may override the world's skipping; stale flags force a full tick
        if (c->tune.chunk_skip >= 0) dw[i].skip = (uint32_t)c->tune.chunk_skip;
        if (dw[i].skip && !ws[i]->flags_valid) dw[i].force_tick = ws[i]->d.tick;
        start[i] = (uint32_t)total;
        total += (uint64_t)ws[i]->d.tiles_x * ws[i]->d.tiles_y;
    }
    if (total >= 0x7FFFFFFFull) return FSG_ERR_INVALID;
    start[n] = (uint32_t)total;
    std::copy(start.begin(), start.end(), c->h_tile_start);
    // no wait for the upload: the staging is pinned and reused only after `staged`
    HIPCHK(hipMemcpyAsync(c->d_worlds, dw, (size_t)n * sizeof(DevWorld), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_tile_start, c->h_tile_start, (size_t)(n + 1) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->staged, c->stream));
    c->staged_pending = true;
    const std::vector<fsg_world *> key(ws, ws + n);
    *graph_ok = c->graph && key == c->graph_key && start == c->graph_tiles && c->graph_ticks == c->tune.graph_ticks &&
                c->graph_occ == c->tune.occupancy;
    if (!*graph_ok) {
        c->graph_key = key;
        c->graph_tiles = start;
    }
    c->total_tiles = (uint32_t)total;
    return FSG_OK;
}

int batch_run(fsg_ctx *c, fsg_world *const *ws, int n, int ticks, bool graph_ok) {
    const unsigned tiles = c->total_tiles;
    const int T = c->tune.graph_ticks;
    auto tick_launch = [&](uint32_t k) {
        if (c->tune.occupancy == 4)
            fsg::k_tick_batch4<<<tiles, fsg::NT, 0, c->stream>>>(c->d_worlds, c->d_tile_start, n, k);
        else
            fsg::k_tick_batch<<<tiles, fsg::NT, 0, c->stream>>>(c->d_worlds, c->d_tile_start, n, k);
    };
    int done = 0;
    if (T > 1 && ticks >= T) {
        if (!graph_ok) {
            if (c->graph) (void)hipGraphExecDestroy(c->graph);
            c->graph = nullptr;
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < T; k++) tick_launch((uint32_t)k);
            fsg::k_advance<<<grid_for(n, 256), 256, 0, c->stream>>>(c->d_worlds, n, (uint32_t)T);
            HIPCHK(hipStreamEndCapture(c->stream, &g));
            const hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                c->graph = nullptr;
                return FSG_ERR_DEVICE;
            }
            c->graph_ticks = T;
            c->graph_occ = c->tune.occupancy;
        }
        for (; done + T <= ticks; done += T) HIPCHK(hipGraphLaunch(c->graph, c->stream));
    }
    const int rem = ticks - done;
    if (rem > 0) {
        for (int k = 0; k < rem; k++) tick_launch((uint32_t)k);
        fsg::k_advance<<<grid_for(n, 256), 256, 0, c->stream>>>(c->d_worlds, n, (uint32_t)rem);
        HIPCHK(hipGetLastError());
    }
    const bool skip_all = c->tune.chunk_skip > 0, skip_none = c->tune.chunk_skip == 0;
    for (int i = 0; i < n; i++) {
        ws[i]->d.tick += (uint64_t)ticks;
        if (ticks > 0) ws[i]->prev_valid = true;
        ws[i]->offered += (uint64_t)ticks * ws[i]->d.tiles_x * ws[i]->d.tiles_y;
        ws[i]->flags_valid = skip_all || (!skip_none && ws[i]->d.skip);
    }
    return FSG_OK;
}

} // namespace

extern "C" int fsg_batch_step(fsg_ctx *c, fsg_world *const *ws, int n, int ticks) {
    if (!c || (n && !ws) || n < 0 || ticks < 0) return FSG_ERR_INVALID;
    if (n == 0 || ticks == 0) return FSG_OK;
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> lk(c->batch_mu);
    bool graph_ok = false;
    if (int rc = batch_prepare(c, ws, n, &graph_ok)) return rc;
    return batch_run(c, ws, n, ticks, graph_ok);
}

extern "C" int fsg_batch_step_timed(fsg_ctx *c, fsg_world *const *ws, int n, int ticks, double *ms) {
    if (!c || (n && !ws) || n < 0 || ticks < 0) return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> lk(c->batch_mu);
    bool graph_ok = false;
    if (n > 0)
        if (int rc = batch_prepare(c, ws, n, &graph_ok)) return rc;
    std::lock_guard<std::mutex> tl(c->time_mu);
    HIPCHK(hipEventRecord(c->e0, c->stream));
    if (n > 0 && ticks > 0)
        if (int rc = batch_run(c, ws, n, ticks, graph_ok)) return rc;
    HIPCHK(hipEventRecord(c->e1, c->stream));
    HIPCHK(hipEventSynchronize(c->e1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->e0, c->e1));
    if (ms) *ms = f;
    return FSG_OK;
}

extern "C" int fsg_perworld_step_timed(fsg_ctx *c, fsg_world *const *ws, int n, int ticks, double *ms) {
    if (!c || (n && !ws) || n < 0 || ticks < 0) return FSG_ERR_INVALID;
    for (int i = 0; i < n; i++)
        if (!ws[i] || ws[i]->ctx != c) return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> tl(c->time_mu);
    HIPCHK(hipEventRecord(c->e0, c->stream));
    for (int t = 0; t < ticks; t++)
        for (int i = 0; i < n; i++)
            if (int rc = launch_tick(ws[i])) return rc;
    HIPCHK(hipEventRecord(c->e1, c->stream));
    HIPCHK(hipEventSynchronize(c->e1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->e0, c->e1));
    if (ms) *ms = f;
    return FSG_OK;
}

// ---------------------------------------------------------------------------
// Batch autotuner
// ---------------------------------------------------------------------------
namespace {

static_assert(sizeof(fsg_tuning) == 16 && sizeof(fsg_tune_report) == 800, "ABI layout of the tuning structs");

bool tuning_ok(const fsg_tuning &t) {
    const bool ticks_ok = t.graph_ticks == 1 || t.graph_ticks == 4 || t.graph_ticks == 8 || t.graph_ticks == 16 ||
                          t.graph_ticks == 32;
    return ticks_ok && t.chunk_skip >= -1 && t.chunk_skip <= 1 && (t.occupancy == 4 || t.occupancy == 5) &&
           t.reserved == 0;
}

// World-size class of a batch: for each log2 area bucket (2^10 .. 2^25
// cells), the bit length of the number of worlds in it, 4 bits per bucket.
uint64_t size_class(fsg_world *const *ws, int n) {
    int count[16] = {0};
    for (int i = 0; i < n; i++) {
        const uint64_t a = (uint64_t)ws[i]->d.W * ws[i]->d.H;
        int lg = 0;
        while ((2ull << lg) <= a) lg++;
        count[std::min(std::max(lg - 10, 0), 15)]++;
    }
    uint64_t key = 0;
    for (int b = 0; b < 16; b++) {
        int bits = 0;
        while ((1 << bits) <= count[b] && bits < 15) bits++;
        key |= (uint64_t)bits << (4 * b);
    }
    return key;
}

std::mutex g_tune_mu;
std::map<std::pair<std::string, uint64_t>, fsg_tuning> g_tune_cache;

// Copies the state of src (both buffers, flags, tick) into dst, a world of
// the same size and configuration.
int copy_state(fsg_world *dst, const fsg_world *src) {
    fsg_ctx *c = dst->ctx;
    const size_t cells = (size_t)src->d.W * src->d.H, nt = (size_t)src->d.tiles_x * src->d.tiles_y;
    for (int b = 0; b < 2; b++) {
        HIPCHK(hipMemcpyAsync(dst->d.buf[b], src->d.buf[b], cells * 8, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(dst->d.flags[b], src->d.flags[b], nt * 4, hipMemcpyDeviceToDevice, c->stream));
    }
    dst->d.tick = src->d.tick;
    dst->d.force_tick = src->d.force_tick;
    dst->flags_valid = src->flags_valid;
    return FSG_OK;
}

}  // namespace

extern "C" int fsg_ctx_get_tuning(fsg_ctx *c, fsg_tuning *out) {
    if (!c || !out) return FSG_ERR_INVALID;
    *out = c->tune;
    return FSG_OK;
}

extern "C" int fsg_ctx_set_tuning(fsg_ctx *c, const fsg_tuning *t) {
    if (!c || !t || !tuning_ok(*t)) return FSG_ERR_INVALID;
    c->tune = *t;
    return FSG_OK;
}

extern "C" int fsg_autotune_batch(fsg_ctx *c, fsg_world *const *ws, int n, int sample_ticks,
                                  fsg_tune_report *rep) {
    if (!c || !ws || n <= 0 || sample_ticks <= 0 || sample_ticks > 4096) return FSG_ERR_INVALID;
    for (int i = 0; i < n; i++)
        if (!ws[i] || ws[i]->ctx != c) return FSG_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->device));
    const auto key = std::make_pair(std::string(prop.gcnArchName), size_class(ws, n));
    fsg_tune_report r{};
    r.size_class = key.second;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto hit = g_tune_cache.find(key);
        if (hit != g_tune_cache.end()) {
            c->tune = hit->second;
            r.best = hit->second;
            r.from_cache = 1;
            if (rep) *rep = r;
            return FSG_OK;
        }
    }
    // copies of the worlds, restored before every timed run
    std::vector<fsg_world *> cl(n, nullptr);
    int rc = FSG_OK;
    for (int i = 0; i < n && !rc; i++) {
        rc = fsg_world_create(c, ws[i]->d.W, ws[i]->d.H, ws[i]->d.seed, &ws[i]->cfg, &cl[i]);
        if (!rc) rc = copy_state(cl[i], ws[i]);
    }
    std::unique_lock<std::mutex> blk(c->batch_mu);
    std::unique_lock<std::mutex> tl(c->time_mu);
    const fsg_tuning before = c->tune;
    std::vector<fsg_tuning> cand;
    for (int occ : {5, 4})
        for (int skip : {1, 0})
            for (int T : {16, 32, 8, 4, 1}) cand.push_back(fsg_tuning{T, skip, occ, 0});
    double best_ms = 1e30;
    for (size_t k = 0; k < cand.size() && !rc; k++) {
        c->tune = cand[k];
        float f = 0.f;
        for (int pass = 0; pass < 2 && !rc; pass++) {  // pass 0 builds the graph, pass 1 is timed
            for (int i = 0; i < n && !rc; i++) rc = copy_state(cl[i], ws[i]);
            bool graph_ok = false;
            if (!rc) rc = batch_prepare(c, cl.data(), n, &graph_ok);
            if (!rc && hipEventRecord(c->e0, c->stream) != hipSuccess) rc = FSG_ERR_DEVICE;
            if (!rc) rc = batch_run(c, cl.data(), n, sample_ticks, graph_ok);
            if (!rc && (hipEventRecord(c->e1, c->stream) != hipSuccess || hipEventSynchronize(c->e1) != hipSuccess ||
                        hipEventElapsedTime(&f, c->e0, c->e1) != hipSuccess))
                rc = FSG_ERR_DEVICE;
        }
        if (rc) break;
        r.candidate[k] = cand[k];
        r.ms[k] = f;
        r.n_candidates = (int)k + 1;
        if (f < best_ms) {
            best_ms = f;
            r.best = cand[k];
        }
    }
    tl.unlock();
    blk.unlock();  // fsg_world_destroy takes batch_mu
    for (fsg_world *w : cl)
        if (w) fsg_world_destroy(w);
    if (rc) {
        c->tune = before;
        return rc;
    }
    c->tune = r.best;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = r.best;
    }
    if (rep) *rep = r;
    return FSG_OK;
}

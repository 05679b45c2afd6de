// fsg_internal.hpp -- library-private state shared by the C-ABI and the
// serving engine (not part of the public boundary).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <vector>

#include "fsg.h"
#include "fsg_device.hpp"
#include "fsg_generate.hpp"
#include "fsg_kernels.hpp"

using fsg::DevWorld;
using fsg::MatTab;

#define HIPCHK(call)                                     \
    do {                                                 \
        if ((call) != hipSuccess) return FSG_ERR_DEVICE; \
    } while (0)

struct fsg_ctx {
    int device = 0;
    int n_cu = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    // events of the *_timed calls, held under time_mu from the first record
    // to the elapsed-time read
    std::mutex time_mu;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;  // fsg_ctx_timer_start/stop
    uint64_t budget = 0, used = 0;
    std::atomic<int> inject{0};  // allocations still to fail (fsg_ctx_inject_alloc_failures)
    std::mutex mu;
    // batch launch state: device world table + chunk prefix sums + graph,
    // shared by every batched launch of the context (held under batch_mu from
    // batch_prepare to the end of the batch_run that follows it)
    std::mutex batch_mu;
    DevWorld *d_worlds = nullptr;
    uint32_t *d_tile_start = nullptr;
    int batch_cap = 0;
    // their pinned host staging, rewritten only after `staged` (recorded after
    // the last upload from it) has completed
    DevWorld *h_worlds = nullptr;
    uint32_t *h_tile_start = nullptr;
    hipEvent_t staged = nullptr;
    bool staged_pending = false;
    std::vector<fsg_world *> graph_key;
    std::vector<uint32_t> graph_tiles;
    hipGraphExec_t graph = nullptr;
    uint32_t total_tiles = 0;
    // batch tuning (fsg_ctx_set_tuning / fsg_autotune_batch) and the tuning
    // the cached graph was captured with
    fsg_tuning tune{16, -1, 5, 0};
    int graph_ticks = 0, graph_occ = 0;
    // freed device blocks by size class (dev_alloc / dev_free); the cap is at
    // least 16 GiB and a quarter of the device's memory
    std::multimap<size_t, void *> cache;
    size_t cached = 0, cache_cap = (size_t)16 << 30;
    // stamp claim rounds: pixels still open after round 0 and their count,
    // held under stamp_mu from the list's sizing to the last launch reading it
    // (one claim sequence at a time on the stream, so no two interleave)
    std::mutex stamp_mu;
    uint64_t *d_open = nullptr;
    uint32_t *d_nopen = nullptr;
    size_t open_cap = 0;
};

struct fsg_world {
    fsg_ctx *ctx = nullptr;
    fsg_config cfg{};
    DevWorld d{};
    bool flags_valid = false;  // the chunk flags describe the last tick (it ran with skipping on)
    bool prev_valid = false;   // buf[(tick + 1) & 1] holds the state before the last tick
    // the awake-list counts are zero for the tick list_next (after a listed
    // tick) or for any tick (list_any: both zero)
    bool list_any = true;
    uint64_t list_next = 0;
    std::vector<uint32_t> ct_host;  // host copy of the class tables (DevWorld::ct)
    uint32_t *d_layer[3] = {nullptr, nullptr, nullptr};  // texture layers (fsg_world_render)
    unsigned long long *d_tally = nullptr;                 // render census replicas
    void *d_isl = nullptr;                                 // island check scratch
    size_t isl_cap = 0;
    uint64_t *d_isl_pix = nullptr;                         // body images of the last island check
    size_t isl_pix_cap = 0, isl_pix_used = 0;
    void *d_gen = nullptr;                                 // device generator scratch
    size_t gen_cap = 0;
    std::vector<fsg_gen_disc> gen_discs;                   // host copy of the last disc upload
    // chunk .pack codec (fsg_pack.hip): encoder slots, assembled images, sizes/offsets,
    // layer2/background upload, decoder input, record staging, descriptors/statuses
    void *d_pack_slots = nullptr, *d_pack_img = nullptr, *d_pack_small = nullptr, *d_pack_aux = nullptr;
    void *d_pack_in = nullptr, *d_pack_raw = nullptr, *d_pack_dec = nullptr;
    size_t pack_slots_cap = 0, pack_img_cap = 0, pack_small_cap = 0, pack_aux_cap = 0;
    size_t pack_in_cap = 0, pack_raw_cap = 0, pack_dec_cap = 0;
    size_t pack_n = 0, pack_total = 0;     // chunks and bytes of the last fsg_world_pack
    std::vector<int32_t> pack_sizes;       // its block sizes (A, B per chunk)
    std::vector<int8_t> pack_phase;        // its generation phases
    // outlines (fsg_mesh.hip): region table, per-workgroup loop scratch, output
    // pools, their sizes from the last call, and the ordered result
    void *d_mesh_regs = nullptr, *d_mesh_scratch = nullptr, *d_mesh_pool = nullptr;
    size_t mesh_regs_cap = 0, mesh_scratch_cap = 0, mesh_pool_cap = 0;
    uint32_t mesh_loop_hint = 0, mesh_vert_hint = 0;
    // the last mesh call's ordered result, on the device (in d_mesh_pool)
    fsg_mesh_loop *d_mesh_loops = nullptr;
    float *d_mesh_verts = nullptr;
    uint32_t mesh_nloops = 0, mesh_nverts = 0;
    // free-fall particles (fsg_particles.hip): the list, the compaction
    // target, per-particle kept flags, counters
    fsg_particle *d_part = nullptr, *d_part2 = nullptr;
    uint8_t *d_part_alive = nullptr;
    unsigned long long *d_part_cnt = nullptr;
    size_t part_cap = 0, npart = 0;
    void *d_pround = nullptr;  // the particle pass's round buffers
    size_t pround_cap = 0;
    bool pack_map_set = false;             // pack_map holds a caller's map (else the default)
    fsg_pack_map pack_map{};
    size_t bytes = 0;
    uint64_t offered = 0;
    unsigned long long *d_stats = nullptr;  // [17] census scratch + [1] processed counter
    // rigid-body stamp state
    uint64_t *d_pixels = nullptr;
    uint32_t *d_pixbody = nullptr;
    fsg_body *d_bodies = nullptr;
    uint8_t *d_state = nullptr;
    uint64_t *d_disp = nullptr, *d_dispxy = nullptr;
    uint64_t *d_tgt = nullptr;  // per pixel: own cell under the pose (stamp kernels)
    uint64_t *d_save = nullptr; // per pixel x 5: claim slot contents kept across a stamp/unstamp
    size_t save_cap = 0;        // pixels d_save holds (single-world stamps only)
    size_t pix_cap = 0;
    int body_cap = 0;
    size_t npix = 0;
    int nbodies = 0;
    bool stamped = false;
    std::vector<fsg_body> bodies;
    fsg::StampWorld *d_sw = nullptr;  // one-entry stamp table on the device
    fsg::StampWorld sw_host{};
    std::vector<uint32_t> pix_body_host;
};


namespace fsg_impl {
// Device bytes a world holds from fsg_world_create on (both state buffers,
// chunk flags and awake list, counts, stats, and the class tables of a world
// of at most 8 materials): what the context's HBM budget counts.
inline size_t world_state_bytes(int W, int H, uint32_t n_materials) {
    const size_t tiles = (size_t)((W + fsg::TW - 1) / fsg::TW) * (size_t)((H + fsg::TH - 1) / fsg::TH);
    return 2 * (size_t)W * H * 8 + 3 * tiles * 4 + 16 + 18 * 8 +
           (n_materials <= 8 ? (size_t)fsg::MAX_IT * 512 * 4 : 0);
}
int dev_alloc(fsg_ctx *c, void **p, size_t bytes);
void dev_free(void *p);
void trim_cache(fsg_ctx *c, size_t keep);  // hipFree cached blocks down to `keep` bytes
unsigned grid_for(size_t n, unsigned block);
// Uploads the world table of a batch; *graph_ok tells whether the cached
// graph matches this world set.
int batch_prepare(fsg_ctx *c, fsg_world *const *ws, int n, bool *graph_ok);
// Enqueues `ticks` ticks of the prepared batch (16-tick graphs + remainder).
int batch_run(fsg_ctx *c, fsg_world *const *ws, int n, int ticks, bool graph_ok);
// The world's entry in a stamp batch (pixels numbered from pix_start).
fsg::StampWorld stamp_world_of(fsg_world *w, uint32_t pix_start);
// Stamp (or unstamp) claim rounds over a device table of n worlds.
// max_npix: the largest pixel count of one world (the launch's x extent).
// save: the claim slots of the spare buffer are restored afterwards.
int stamp_rounds(fsg_ctx *c, const fsg::StampWorld *d_table, int n, uint32_t total, uint32_t max_npix,
                 const MatTab &M, bool unstamp, bool save);
// Enqueues w's census (counts [0,16), checksum [16], processed chunks [17])
// into host[18], valid after the context's stream is synchronised.
int stats_enqueue(fsg_world *w, unsigned long long *host);
// Uploads pixels/bodies of a world for stamping (validates them).
int set_bodies(fsg_world *w, const fsg_body *bodies, int nb, const fsg_cell *pixels, size_t npix);
// Looks up the device address of fsg_world_set_cells_noalloc's array on the
// current device (once per device).
int noalloc_setup(int device);
}  // namespace fsg_impl

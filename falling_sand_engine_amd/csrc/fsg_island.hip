// fsg_island.hip -- world::physicsCheck (world.cpp:3330-3429) for a list of
// seed cells.
//
// The reference flood-fills the 4-connected SOLID component of one cell with
// a recursive DFS that gives up past 1000 cells, after clearing a W*H visited
// array and a W*H colour array (5 bytes per cell of the world per call).  A
// component of at most 1000 cells is cleared to Tiles_NOTHING and, above 10
// cells, becomes a rigid body textured with its colours.  The outcome does not
// depend on the fill order: either the whole component (<= 1000 cells) or
// "more than 1000".  So here each seed gets one workgroup running a BFS with
// its visited set in LDS (a hash set, never more than half full) and its
// frontier as a range of the visited list, reading the grid only; the host
// then keeps, in seed order, the first seed of each island (later seeds of a
// removed island find air, as consecutive calls would), and one more launch
// writes the body images and clears the islands.
#include "fsg_internal.hpp"

#include <algorithm>
#include <unordered_set>
#include <vector>

static_assert(sizeof(fsg_island) == 40, "ABI size");

namespace {

constexpr int IT = 256;                          // threads per seed
constexpr uint32_t CAP = FSG_ISLAND_MAX_CELLS;   // islands larger than this stay
constexpr uint32_t LIST = 1024;                  // visited cells kept in order (>= CAP + 1)
constexpr uint32_t HT = 4096;                    // LDS hash set slots
// An insert starts only while the count is at most CAP and a thread has one
// insert in flight, so the set never holds more than CAP + 1 + IT cells.
static_assert(LIST >= CAP + 1 && HT >= 2 * (CAP + 1 + IT), "hash set stays at most half full");

struct SeedOut {
    int32_t status;
    uint32_t count;
    int32_t minx, miny, maxx, maxy;
    uint32_t minidx;  // smallest cell index of the island: its identity
    uint32_t pad;
};

__device__ __forceinline__ uint32_t hslot(uint32_t key) { return (key * 2654435761u) >> 20; }  // 12 bits

__device__ __forceinline__ int wave_min(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// One workgroup per seed; read-only on the grid.  lists: CAP cells per seed.
__global__ __launch_bounds__(IT) void k_island_search(const uint64_t *__restrict__ cells, int W, int H,
                                                      uint32_t solid, const int32_t *__restrict__ xy,
                                                      uint32_t *__restrict__ lists, SeedOut *__restrict__ out) {
    __shared__ uint32_t ht[HT];
    __shared__ uint32_t list[LIST];
    __shared__ uint32_t cnt;
    __shared__ int red[5][IT / 64];
    const int tid = threadIdx.x, s = blockIdx.x;
    for (uint32_t i = tid; i < HT; i += IT) ht[i] = 0;
    if (tid == 0) cnt = 0;
    __syncthreads();
    auto is_solid = [&](uint32_t c) { return (solid >> ((uint32_t)cells[c] & 15u)) & 1u; };
    auto insert = [&](uint32_t c) {  // key c + 1; 0 marks an empty slot
        if (*static_cast<volatile uint32_t *>(&cnt) > CAP) return;  // the island is too large already
        uint32_t h = hslot(c + 1);
        while (true) {
            const uint32_t old = atomicCAS(&ht[h], 0u, c + 1);
            if (old == 0u) {
                const uint32_t k = atomicAdd(&cnt, 1u);
                if (k < LIST) list[k] = c;
                return;
            }
            if (old == c + 1) return;
            h = (h + 1) & (HT - 1);
        }
    };
    const int sx = xy[2 * s], sy = xy[2 * s + 1];
    const bool ok = sx >= 0 && sy >= 0 && sx < W && sy < H && is_solid((uint32_t)sy * (uint32_t)W + (uint32_t)sx);
    if (ok && tid == 0) insert((uint32_t)sy * (uint32_t)W + (uint32_t)sx);
    __syncthreads();
    uint32_t lo = 0, hi = cnt;
    while (ok && lo < hi && hi <= CAP) {
        for (uint32_t i = lo + tid; i < hi; i += IT) {
            const uint32_t c = list[i];
            const uint32_t y = c / (uint32_t)W, x = c - y * (uint32_t)W;
            if (x + 1 < (uint32_t)W && is_solid(c + 1)) insert(c + 1);
            if (y + 1 < (uint32_t)H && is_solid(c + W)) insert(c + W);
            if (x > 0 && is_solid(c - 1)) insert(c - 1);
            if (y > 0 && is_solid(c - W)) insert(c - W);
        }
        __syncthreads();
        lo = hi;
        hi = cnt;
        __syncthreads();  // every thread has read cnt before the next level inserts
    }
    const uint32_t total = ok ? cnt : 0u;
    SeedOut r{};
    r.status = !ok ? FSG_ISLAND_NONE : total > CAP ? FSG_ISLAND_ATTACHED
                                     : total < (uint32_t)FSG_ISLAND_MIN_BODY ? FSG_ISLAND_REMOVED : FSG_ISLAND_BODY;
    r.count = total;
    if (r.status == FSG_ISLAND_REMOVED || r.status == FSG_ISLAND_BODY) {
        int mnx = INT32_MAX, mny = INT32_MAX, mxx = -1, mxy = -1, mni = INT32_MAX;
        for (uint32_t i = tid; i < total; i += IT) {
            const uint32_t c = list[i];
            const int y = (int)(c / (uint32_t)W), x = (int)(c - (uint32_t)y * (uint32_t)W);
            mnx = min(mnx, x);
            mxx = max(mxx, x);
            mny = min(mny, y);
            mxy = max(mxy, y);
            mni = min(mni, (int)c);
            lists[(size_t)s * CAP + i] = c;
        }
        mnx = wave_min(mnx);
        mny = wave_min(mny);
        mni = wave_min(mni);
        mxx = wave_max(mxx);
        mxy = wave_max(mxy);
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = mnx;
            red[1][tid >> 6] = mny;
            red[2][tid >> 6] = mxx;
            red[3][tid >> 6] = mxy;
            red[4][tid >> 6] = mni;
        }
        __syncthreads();
        if (tid == 0) {
            for (int k = 1; k < IT / 64; k++) {
                red[0][0] = min(red[0][0], red[0][k]);
                red[1][0] = min(red[1][0], red[1][k]);
                red[2][0] = max(red[2][0], red[2][k]);
                red[3][0] = max(red[3][0], red[3][k]);
                red[4][0] = min(red[4][0], red[4][k]);
            }
            r.minx = red[0][0];
            r.miny = red[1][0];
            r.maxx = red[2][0];
            r.maxy = red[3][0];
            r.minidx = (uint32_t)red[4][0];
        }
    }
    if (tid == 0) out[s] = r;
}

struct Winner {
    uint32_t seed;   // index into lists
    uint32_t count;
    int32_t minx, miny, w;
    uint32_t body;   // 1: write the image
    uint64_t off;    // first image cell
};

// One workgroup per kept island: its cells go to the body image, then the
// grid cell becomes Tiles_NOTHING.  Images were filled with `air` before.
__global__ __launch_bounds__(IT) void k_island_apply(uint64_t *__restrict__ cells, int W,
                                                     const uint32_t *__restrict__ lists,
                                                     const Winner *__restrict__ win, uint64_t air,
                                                     uint64_t *__restrict__ pixels) {
    const Winner v = win[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < v.count; i += IT) {
        const uint32_t c = lists[(size_t)v.seed * CAP + i];
        const int y = (int)(c / (uint32_t)W), x = (int)(c - (uint32_t)y * (uint32_t)W);
        if (v.body) pixels[v.off + (size_t)(y - v.miny) * v.w + (x - v.minx)] = cells[c];
        cells[c] = air;
    }
}

__global__ void k_fill_u64(uint64_t *p, size_t n, uint64_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

} // namespace

// Grows *buf to at least `bytes` (contents are not kept).
static int grow(fsg_ctx *c, void **buf, size_t *cap, size_t bytes) {
    if (*cap >= bytes) return FSG_OK;
    fsg_impl::dev_free(*buf);
    *buf = nullptr;
    *cap = 0;
    if (fsg_impl::dev_alloc(c, buf, bytes)) return FSG_ERR_NOMEM;
    *cap = bytes;
    return FSG_OK;
}

extern "C" int fsg_world_physics_check(fsg_world *w, fsg_island *seeds, int n, size_t *pixels_needed) {
    if (!w || n < 0 || (n > 0 && !seeds) || n > (1 << 20)) return FSG_ERR_INVALID;
    if (pixels_needed) *pixels_needed = 0;
    w->isl_pix_used = 0;
    if (n == 0) return FSG_OK;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int W = w->d.W, H = w->d.H;
    uint32_t solid = 0;
    for (uint32_t m = 0; m < w->cfg.n_materials; m++)
        if (w->cfg.materials[m].type == FSG_SOLID) solid |= 1u << m;
    std::vector<int32_t> xy(2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        xy[2 * i] = seeds[i].x;
        xy[2 * i + 1] = seeds[i].y;
    }
    // scratch (kept by the world): seed cells, island cell lists, per-seed
    // results, kept islands
    const size_t b_xy = (xy.size() * 4 + 255) & ~(size_t)255, b_lists = (size_t)n * CAP * 4;
    const size_t b_out = (size_t)n * sizeof(SeedOut), b_win = (size_t)n * sizeof(Winner);
    if (int rc = grow(c, &w->d_isl, &w->isl_cap, b_xy + b_lists + b_out + b_win)) return rc;
    char *scratch = static_cast<char *>(w->d_isl);
    int32_t *d_xy = reinterpret_cast<int32_t *>(scratch);
    uint32_t *d_lists = reinterpret_cast<uint32_t *>(scratch + b_xy);
    SeedOut *d_out = reinterpret_cast<SeedOut *>(scratch + b_xy + b_lists);
    Winner *d_win = reinterpret_cast<Winner *>(scratch + b_xy + b_lists + b_out);
    std::vector<SeedOut> res((size_t)n);
    uint64_t *cur = w->d.buf[w->d.tick & 1];
    HIPCHK(hipMemcpyAsync(d_xy, xy.data(), xy.size() * 4, hipMemcpyHostToDevice, c->stream));
    k_island_search<<<n, IT, 0, c->stream>>>(cur, W, H, solid, d_xy, d_lists, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(res.data(), d_out, b_out, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // keep the first seed of each island, in seed order
    std::vector<Winner> win;
    std::unordered_set<uint32_t> gone;
    size_t needed = 0;
    for (int i = 0; i < n; i++) {
        SeedOut &r = res[(size_t)i];
        if (r.status != FSG_ISLAND_REMOVED && r.status != FSG_ISLAND_BODY) continue;
        if (!gone.insert(r.minidx).second) {
            r.status = FSG_ISLAND_NONE;  // removed by an earlier seed: air now
            continue;
        }
        Winner v{};
        v.seed = (uint32_t)i;
        v.count = r.count;
        v.minx = r.minx;
        v.miny = r.miny;
        v.w = r.maxx - r.minx + 1;
        v.body = r.status == FSG_ISLAND_BODY;
        v.off = needed;
        if (v.body) needed += (size_t)v.w * (size_t)(r.maxy - r.miny + 1);
        win.push_back(v);
    }
    if (!win.empty()) {
        if (needed) {
            void *pb = w->d_isl_pix;
            if (int rc = grow(c, &pb, &w->isl_pix_cap, needed * 8)) {
                w->d_isl_pix = nullptr;
                return rc;
            }
            w->d_isl_pix = static_cast<uint64_t *>(pb);
        }
        const uint64_t air = (uint64_t)w->cfg.mat_air;  // Tiles_NOTHING: colour 0, temperature 0
        HIPCHK(hipMemcpyAsync(d_win, win.data(), win.size() * sizeof(Winner), hipMemcpyHostToDevice, c->stream));
        if (needed) {
            k_fill_u64<<<fsg_impl::grid_for(needed, 256), 256, 0, c->stream>>>(w->d_isl_pix, needed, air);
            HIPCHK(hipGetLastError());
        }
        k_island_apply<<<(unsigned)win.size(), IT, 0, c->stream>>>(cur, W, d_lists, d_win, air, w->d_isl_pix);
        HIPCHK(hipGetLastError());
        // the grid changed outside a tick: every chunk runs at the next tick
        w->d.force_tick = w->d.tick;
    }
    w->isl_pix_used = needed;
    if (pixels_needed) *pixels_needed = needed;
    for (int i = 0; i < n; i++) {
        const SeedOut &r = res[(size_t)i];
        fsg_island &s = seeds[i];
        const bool kept = r.status == FSG_ISLAND_REMOVED || r.status == FSG_ISLAND_BODY;
        s.status = r.status;
        s.count = kept ? r.count : 0u;
        s.min_x = kept ? r.minx : 0;
        s.min_y = kept ? r.miny : 0;
        s.w = kept ? r.maxx - r.minx + 1 : 0;
        s.h = kept ? r.maxy - r.miny + 1 : 0;
        s.pixel_offset = 0;
    }
    for (const Winner &v : win)
        if (v.body) seeds[v.seed].pixel_offset = v.off;
    return FSG_OK;
}

extern "C" int fsg_world_island_pixels(fsg_world *w, fsg_cell *out, size_t cap) {
    if (!w || (w->isl_pix_used && (!out || cap < w->isl_pix_used))) return FSG_ERR_INVALID;
    if (!w->isl_pix_used) return FSG_OK;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(out, w->d_isl_pix, w->isl_pix_used * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FSG_OK;
}

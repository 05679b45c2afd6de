// fsg_generate.hip -- the benchmark world generator on the device.
//
// fsg_generate (fsg_host.cpp) builds the synthetic worlds in host memory at
// about 7 Mcell/s per core; a serving loop that creates worlds on request
// would spend its time there and in the upload.  Here the same worlds are
// written straight into a world's grid: the per-column terrain heights in one
// launch, then one workgroup per 64x64 tile that keeps, in drawing order, the
// discs overlapping its tile in LDS and gives each cell the material of the
// last disc covering it (discs never replace stone), as the host's disc-by-
// disc painting does.  The disc list comes from the host plan shared with
// fsg_generate, so both paths produce the same cells.
#include "fsg_generate.hpp"
#include "fsg_internal.hpp"

#include <vector>

namespace {

constexpr int GT = 256;      // threads per tile workgroup
constexpr int MAXD = 1024;   // overlapping discs kept in LDS (more: read from the global list)
enum { M_AIR = 0, M_STONE = 1, M_WATER = 3, M_SAND = 2, M_STEAM = 5, M_LAVA = 6 };

struct GenConst {
    uint32_t rgba[8];
    int32_t ambient, lava_temp;
    uint64_t seed;
};

__device__ __forceinline__ uint64_t make_cell_d(const GenConst &g, int m, int x, int y) {
    const uint64_t h = fsg::splitmix64(g.seed ^ ((uint64_t)(uint32_t)x << 20) ^ ((uint64_t)(uint32_t)y << 42) ^ 0xC0105ull);
    uint32_t rgba = 0;
    if (m != M_AIR) {  // +-8 per colour channel, alpha kept
        const uint32_t base = g.rgba[m];
        rgba = base & 0xFF000000u;
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            int v = (int)((base >> (8 * ch)) & 0xFF) + (int)((h >> (8 * ch)) & 15) - 8;
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            rgba |= (uint32_t)v << (8 * ch);
        }
    }
    const int t = m == M_LAVA ? g.lava_temp : g.ambient;
    return (uint64_t)(uint32_t)m | ((uint64_t)(uint16_t)(int16_t)t << 16) | ((uint64_t)rgba << 32);
}

// Terrain height of column x: 4-octave integer value noise (weights 8:4:2:1)
// scaled to 0.4 H, as fsg_host.cpp's value_noise.
__global__ void k_gen_heights(int W, int H, uint64_t seed, int32_t *hgt) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= W) return;
    int acc = 0;
    for (int o = 0; o < 4; o++) {
        int period = (W / 4) >> o;
        if (period < 2) period = 2;
        const int i = x / period, f = x % period;
        const int v0 = (int)(fsg::splitmix64(seed ^ ((uint64_t)(o + 1) << 56) ^ (uint64_t)(uint32_t)i) >> 48);
        const int v1 = (int)(fsg::splitmix64(seed ^ ((uint64_t)(o + 1) << 56) ^ (uint64_t)(uint32_t)(i + 1)) >> 48);
        acc += (8 >> o) * (v0 + (int)(((int64_t)(v1 - v0) * f) / period));
    }
    const int n = acc / 15;
    hgt[x] = (int)(((int64_t)n * 2 * H) / (5 * 65536));
}

__global__ __launch_bounds__(GT) void k_gen_basin(uint64_t *__restrict__ cells, int W, int H, GenConst g,
                                                  const int32_t *__restrict__ hgt,
                                                  const fsg_gen_disc *__restrict__ discs, int nd, int tiles_x) {
    __shared__ fsg_gen_disc loc[MAXD];
    __shared__ int cnt[GT / 64], nloc;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = (blockIdx.x % tiles_x) * 64, y0 = (blockIdx.x / tiles_x) * 64;
    const int x1 = min(x0 + 64, W), y1 = min(y0 + 64, H);
    if (tid == 0) nloc = 0;
    __syncthreads();
    // ordered compaction of the discs whose box meets the tile
    for (int b0 = 0; b0 < nd; b0 += GT) {
        const int i = b0 + tid;
        bool hit = false;
        fsg_gen_disc d{};
        if (i < nd) {
            d = discs[i];
            hit = d.cx + d.r >= x0 && d.cx - d.r < x1 && d.cy + d.r >= y0 && d.cy - d.r < y1;
        }
        const uint64_t m = __ballot(hit);
        if (lane == 0) cnt[wv] = __popcll(m);
        __syncthreads();
        int base = nloc;
        for (int k = 0; k < wv; k++) base += cnt[k];
        const int rank = base + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (hit && rank < MAXD) loc[rank] = d;
        __syncthreads();
        if (tid == 0) nloc += cnt[0] + cnt[1] + cnt[2] + cnt[3];
        __syncthreads();
    }
    const int nl = nloc;
    const bool all_local = nl <= MAXD;
    const int level = H - (H * 55) / 100;  // the water basin's top row
    for (int k = tid; k < 64 * 64; k += GT) {
        const int x = x0 + (k & 63), y = y0 + (k >> 6);
        if (x >= x1 || y >= y1) continue;
        int m = y >= H - hgt[x] ? M_STONE : M_AIR;
        if (m == M_AIR && y >= level) m = M_WATER;
        if (m != M_STONE) {
            // the last disc in drawing order that covers the cell wins
            for (int i = (all_local ? nl : nd) - 1; i >= 0; i--) {
                const fsg_gen_disc d = all_local ? loc[i] : discs[i];
                const int dx = x - d.cx, dy = y - d.cy;
                if (dx < -d.r || dx > d.r || dy < -d.r || dy > d.r) continue;
                if (dx * dx + dy * dy > d.r * d.r) continue;
                m = d.m;
                break;
            }
        }
        cells[(size_t)y * W + x] = make_cell_d(g, m, x, y);
    }
}

// The striped all-active world of fsg_host.cpp's gen_striped.
__global__ void k_gen_striped(uint64_t *__restrict__ cells, int W, int H, GenConst g) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)W * H) return;
    const int y = (int)(i / (size_t)W), x = (int)(i - (size_t)y * W);
    int band = -1;
    if (y * 10 >= H && y * 10 < 2 * H) band = M_WATER;
    if (y * 10 >= 8 * H && y * 10 < 9 * H) band = M_STEAM;
    int m;
    if (y == H - 1) m = M_STONE;
    else if (band >= 0) m = band;
    else if ((y & 1) == 0)
        m = (fsg::splitmix64(g.seed ^ ((uint64_t)(uint32_t)y << 32) ^ (uint64_t)(uint32_t)x) >> 63) ? M_SAND : M_AIR;
    else m = M_AIR;
    const uint32_t rgba = m == M_AIR ? 0u : g.rgba[m];
    cells[i] = (uint64_t)(uint32_t)m | ((uint64_t)(uint16_t)(int16_t)g.ambient << 16) | ((uint64_t)rgba << 32);
}

} // namespace

extern "C" int fsg_world_generate(fsg_world *w, int kind, uint64_t seed) {
    if (!w) return FSG_ERR_INVALID;
    std::vector<fsg_gen_disc> discs;
    GenConst g{};
    int16_t amb, lava;
    if (fsg_gen_plan(kind, w->d.W, w->d.H, seed, discs, g.rgba, amb, lava) != FSG_OK) return FSG_ERR_INVALID;
    g.ambient = amb;
    g.lava_temp = lava;
    g.seed = seed;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    const int W = w->d.W, H = w->d.H;
    uint64_t *cur = w->d.buf[w->d.tick & 1];
    if (kind == 3) {
        k_gen_striped<<<fsg_impl::grid_for((size_t)W * H, 256), 256, 0, c->stream>>>(cur, W, H, g);
        HIPCHK(hipGetLastError());
    } else {
        // per-world scratch: column heights, then the disc list
        const size_t b_h = ((size_t)W * 4 + 255) & ~(size_t)255, b_d = discs.size() * sizeof(fsg_gen_disc);
        if (w->gen_cap < b_h + b_d) {
            fsg_impl::dev_free(w->d_gen);
            w->d_gen = nullptr;
            w->gen_cap = 0;
            if (fsg_impl::dev_alloc(c, &w->d_gen, b_h + b_d)) return FSG_ERR_NOMEM;
            w->gen_cap = b_h + b_d;
        }
        int32_t *d_h = static_cast<int32_t *>(w->d_gen);
        fsg_gen_disc *d_d = reinterpret_cast<fsg_gen_disc *>(static_cast<char *>(w->d_gen) + b_h);
        w->gen_discs = std::move(discs);  // the host copy outlives the asynchronous upload
        if (b_d)
            HIPCHK(hipMemcpyAsync(d_d, w->gen_discs.data(), b_d, hipMemcpyHostToDevice, c->stream));
        k_gen_heights<<<fsg_impl::grid_for((size_t)W, 256), 256, 0, c->stream>>>(W, H, seed, d_h);
        k_gen_basin<<<(unsigned)(w->d.tiles_x * w->d.tiles_y), GT, 0, c->stream>>>(
            cur, W, H, g, d_h, d_d, (int)w->gen_discs.size(), w->d.tiles_x);
        HIPCHK(hipGetLastError());
    }
    // a new grid, as after fsg_world_upload
    w->d.force_tick = w->d.tick;
    w->stamped = false;
    w->prev_valid = false;
    return FSG_OK;
}

// fsg_render.hip -- the texture pass of game.cpp:1994-2067 on the device.
//
// The reference walks the whole grid on three pool tasks after every tick
// and rewrites the texels of dirty cells: the cell colour with the
// material's alpha (AIR: transparent), the material's emitColor, and the
// fire layer (fire cells; AIR clears; any other cell keeps its texel).  Each
// of those texels is a function of the cell alone (the fire texel of the cell
// and its old value), so one streaming pass over a rectangle yields the same
// texture as the dirty-driven updates without per-cell dirty flags; the
// movingTiles census (cells that changed in the last tick, by material)
// compares with the state before the last tick, which the tick's double
// buffer already holds.
//
// k_render: a workgroup covers 1024 cells of one row; a thread 4 cells (two
// 16-byte cell loads, one 16-byte store per layer) where the quad lies inside
// the rectangle and the row pitch keeps 16-byte alignment, else per cell.
// HBM bytes per cell: 8 read (16 with the census) + 4 per colour layer
// written + 4 per fire-layer texel that changes (AIR or fire cells).
#include "fsg_internal.hpp"

#include <algorithm>

static_assert(sizeof(fsg_render_params) == 92 && sizeof(fsg_render_stats) == 144, "ABI sizes");

namespace {

constexpr int RT = 256;  // threads per workgroup; one quad of 4 cells each
// The census lands in NREP replicas of the 17 counters, one 256-byte line
// each, picked by workgroup: a whole-world pass has ~10^5 workgroups, and
// their atomics on one line would serialise in the L2.
constexpr int NREP = 64, REP_STRIDE = 32;  // u64 per replica slot

struct RenderTab {
    uint32_t emit[16];  // emission texel (R,G,B,A bytes) per material id
    uint32_t air;       // bit per material id: PhysicsType AIR
    uint32_t fire;      // fire-layer material id; 16 = none
    uint32_t layers;    // FSG_LAYER_* mask
    uint32_t vec;       // width % 4 == 0: 16-byte accesses are aligned
};

// tally (replica r at r * REP_STRIDE): [0..15] changed cells by material,
// [16] fire cells drawn.
__global__ __launch_bounds__(RT) void k_render(const uint64_t *__restrict__ cur, const uint64_t *__restrict__ prev,
                                               int W, int x0, int x1, int y0, RenderTab t,
                                               uint32_t *__restrict__ lc, uint32_t *__restrict__ le,
                                               uint32_t *__restrict__ lf, unsigned long long *__restrict__ tally) {
    __shared__ uint32_t emit[16];
    __shared__ uint32_t hist[17];
    const int tid = threadIdx.x;
    if (tid < 16) emit[tid] = t.emit[tid];
    if (tid < 17) hist[tid] = 0;
    __syncthreads();
    const size_t row = (size_t)(y0 + (int)blockIdx.y) * (size_t)W;
    const int xs = 4 * ((x0 >> 2) + (int)blockIdx.x * RT + tid);  // first cell of the quad
    const bool want_c = t.layers & FSG_LAYER_CELLS, want_e = t.layers & FSG_LAYER_EMISSION;
    const bool want_f = t.layers & FSG_LAYER_FIRE;
    uint32_t key[4] = {16u, 16u, 16u, 16u};  // census: material of a changed cell, 16 = none
    uint32_t fires = 0;                      // fire cells drawn by this thread
    if (xs < x1) {
        const bool full = t.vec && xs >= x0 && xs + 4 <= x1;
        uint64_t c[4], p[4];
        bool in[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            in[k] = full || (xs + k >= x0 && xs + k < x1);
            c[k] = p[k] = 0;
        }
        if (full) {
            const uint4 a = *reinterpret_cast<const uint4 *>(cur + row + xs);
            const uint4 b = *reinterpret_cast<const uint4 *>(cur + row + xs + 2);
            c[0] = a.x | (uint64_t)a.y << 32;
            c[1] = a.z | (uint64_t)a.w << 32;
            c[2] = b.x | (uint64_t)b.y << 32;
            c[3] = b.z | (uint64_t)b.w << 32;
            if (prev) {
                const uint4 pa = *reinterpret_cast<const uint4 *>(prev + row + xs);
                const uint4 pb = *reinterpret_cast<const uint4 *>(prev + row + xs + 2);
                p[0] = pa.x | (uint64_t)pa.y << 32;
                p[1] = pa.z | (uint64_t)pa.w << 32;
                p[2] = pb.x | (uint64_t)pb.y << 32;
                p[3] = pb.z | (uint64_t)pb.w << 32;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (in[k]) {
                    c[k] = cur[row + xs + k];
                    if (prev) p[k] = prev[row + xs + k];
                }
        }
        uint32_t tc[4], te[4], fmask = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t m = (uint32_t)c[k] & 15u;
            const bool air = (t.air >> m) & 1u, fire = m == t.fire;
            tc[k] = air ? 0u : (uint32_t)(c[k] >> 32);
            te[k] = air ? 0u : emit[m];
            if (in[k]) {
                fmask |= (uint32_t)(air || fire) << k;  // fire-layer texels that change
                fires += fire && !air;
                if (prev && ((c[k] ^ p[k]) & 0xFFFFFFFF000000FFull)) key[k] = m;
            }
        }
        if (full) {
            if (want_c) *reinterpret_cast<uint4 *>(lc + row + xs) = make_uint4(tc[0], tc[1], tc[2], tc[3]);
            if (want_e) *reinterpret_cast<uint4 *>(le + row + xs) = make_uint4(te[0], te[1], te[2], te[3]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (in[k]) {
                    if (want_c) lc[row + xs + k] = tc[k];
                    if (want_e) le[row + xs + k] = te[k];
                }
        }
        if (want_f) {
            if (full && fmask == 15u) {
                *reinterpret_cast<uint4 *>(lf + row + xs) = make_uint4(tc[0], tc[1], tc[2], tc[3]);
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if ((fmask >> k) & 1u) lf[row + xs + k] = tc[k];
            }
        }
    }
    if (tally) {
        // Census per wave: one ballot per distinct changed material (a wave
        // sees a handful), one LDS add each, instead of an atomic per cell.
        const int lane = tid & 63;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint64_t todo = __ballot(key[k] != 16u);
            while (todo) {
                const uint32_t mat = (uint32_t)__builtin_amdgcn_readlane((int)key[k], __builtin_ctzll(todo));
                const uint64_t same = __ballot(key[k] == mat);
                if (lane == 0) atomicAdd(&hist[mat], (uint32_t)__popcll(same));
                todo &= ~same;
            }
        }
        uint32_t wf = fires;
        for (int off = 32; off > 0; off >>= 1) wf += __shfl_xor(wf, off, 64);
        if (lane == 0 && wf) atomicAdd(&hist[16], wf);
        __syncthreads();
        unsigned long long *rep = tally + (size_t)((blockIdx.x + blockIdx.y * gridDim.x) % NREP) * REP_STRIDE;
        if (tid < 17 && hist[tid]) atomicAdd(&rep[tid], (unsigned long long)hist[tid]);
    }
}

int layer_index(uint32_t layer) {
    return layer == FSG_LAYER_CELLS ? 0 : layer == FSG_LAYER_EMISSION ? 1 : layer == FSG_LAYER_FIRE ? 2 : -1;
}

// Allocates layer li of w (zeroed) on first use, under the ctx budget.
int ensure_layer(fsg_world *w, int li) {
    if (w->d_layer[li]) return FSG_OK;
    fsg_ctx *c = w->ctx;
    const size_t bytes = (size_t)w->d.W * (size_t)w->d.H * 4;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->budget && c->used + bytes > c->budget) return FSG_ERR_BUDGET;
        c->used += bytes;
    }
    void *p = nullptr;
    if (fsg_impl::dev_alloc(c, &p, bytes) || hipMemsetAsync(p, 0, bytes, c->stream) != hipSuccess) {
        fsg_impl::dev_free(p);
        std::lock_guard<std::mutex> lk(c->mu);
        c->used -= bytes;
        return FSG_ERR_NOMEM;
    }
    w->d_layer[li] = static_cast<uint32_t *>(p);
    w->bytes += bytes;  // released with the world
    return FSG_OK;
}

// Reference emitColor 0xAARRGGBB (game.cpp:2027-2030) -> texel bytes R,G,B,A.
uint32_t emit_texel(uint32_t e) {
    return ((e >> 16) & 0xFFu) | (((e >> 8) & 0xFFu) << 8) | ((e & 0xFFu) << 16) | (e & 0xFF000000u);
}

} // namespace

extern "C" int fsg_world_render(fsg_world *w, const fsg_render_params *p, fsg_render_stats *stats) {
    if (!w || !p || (p->layers & ~(FSG_LAYER_CELLS | FSG_LAYER_EMISSION | FSG_LAYER_FIRE)) || p->fire_mat > 255)
        return FSG_ERR_INVALID;
    if (stats) *stats = fsg_render_stats{};
    const int W = w->d.W, H = w->d.H;
    int64_t x0 = 0, y0 = 0, x1 = W, y1 = H;
    if (p->w > 0 && p->h > 0) {
        x0 = std::max<int64_t>(p->x0, 0);
        y0 = std::max<int64_t>(p->y0, 0);
        x1 = std::min<int64_t>((int64_t)p->x0 + p->w, W);
        y1 = std::min<int64_t>((int64_t)p->y0 + p->h, H);
    }
    if (x0 >= x1 || y0 >= y1) return FSG_OK;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    for (int li = 0; li < 3; li++)
        if (p->layers & (1u << li))
            if (int rc = ensure_layer(w, li)) return rc;
    RenderTab t{};
    for (int m = 0; m < 16; m++) t.emit[m] = emit_texel(p->emit[m]);
    for (uint32_t m = 0; m < w->cfg.n_materials; m++)
        if (w->cfg.materials[m].type == FSG_AIR) t.air |= 1u << m;
    t.fire = p->fire_mat < 16 ? p->fire_mat : 16u;
    t.layers = p->layers;
    t.vec = (W % 4) == 0;
    const uint64_t *cur = w->d.buf[w->d.tick & 1];
    const uint64_t *prev = (stats && p->count_moved && w->prev_valid) ? w->d.buf[(w->d.tick + 1) & 1] : nullptr;
    constexpr size_t tally_bytes = (size_t)NREP * REP_STRIDE * sizeof(unsigned long long);
    if (stats && !w->d_tally) {
        void *q = nullptr;
        if (fsg_impl::dev_alloc(c, &q, tally_bytes)) return FSG_ERR_NOMEM;
        w->d_tally = static_cast<unsigned long long *>(q);
    }
    unsigned long long *tally = stats ? w->d_tally : nullptr;
    if (tally) HIPCHK(hipMemsetAsync(tally, 0, tally_bytes, c->stream));
    const int quads = (int)(((x1 + 3) >> 2) - (x0 >> 2));
    const dim3 grid((unsigned)((quads + RT - 1) / RT), (unsigned)(y1 - y0));
    k_render<<<grid, RT, 0, c->stream>>>(cur, prev, W, (int)x0, (int)x1, (int)y0, t, w->d_layer[0], w->d_layer[1],
                                         w->d_layer[2], tally);
    HIPCHK(hipGetLastError());
    if (stats) {
        unsigned long long rep[NREP * REP_STRIDE], h[17] = {};
        HIPCHK(hipMemcpyAsync(rep, tally, tally_bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int r = 0; r < NREP; r++)
            for (int k = 0; k < 17; k++) h[k] += rep[r * REP_STRIDE + k];
        for (int m = 0; m < 16; m++) {
            stats->moved[m] = h[m];
            stats->had_dirty |= h[m] != 0;
        }
        stats->fire_cells = h[16];
    }
    return FSG_OK;
}

extern "C" int fsg_world_layer(fsg_world *w, uint32_t layer, uint32_t **device_ptr) {
    const int li = layer_index(layer);
    if (!w || !device_ptr || li < 0) return FSG_ERR_INVALID;
    *device_ptr = w->d_layer[li];
    return FSG_OK;
}

extern "C" int fsg_world_layer_download(fsg_world *w, uint32_t layer, int x0, int y0, int rw, int rh,
                                        uint32_t *out) {
    const int li = layer_index(layer);
    if (!w || !out || li < 0 || rw <= 0 || rh <= 0 || x0 < 0 || y0 < 0 || x0 + rw > w->d.W || y0 + rh > w->d.H)
        return FSG_ERR_INVALID;
    if (!w->d_layer[li]) return FSG_ERR_STATE;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    const uint32_t *src = w->d_layer[li] + (size_t)y0 * w->d.W + x0;
    HIPCHK(hipMemcpy2DAsync(out, (size_t)rw * 4, src, (size_t)w->d.W * 4, (size_t)rw * 4, (size_t)rh,
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FSG_OK;
}

// fsg_particles.hip -- free-fall particles (CellData, game_utils/cells.h:15-36)
// and world::tickCells (world.cpp:2030-2189) on the device.
//
// The reference moves its particles one after another, and a particle that
// lands changes the cells the particles after it move through and search.
// The device keeps that order exactly.  One workgroup walks the list in
// batches of 256: every lane moves its particle against the grid as the
// earlier batches left it, and records the bounding box of every cell it
// read (its path, the cell it started in, the spiral it searched).  A lane
// whose box holds the landing cell of an earlier lane of the batch might
// have seen something else had it run after that lane; the first such lane
// ends the batch, the lanes before it commit (their landings touch no cell
// another committed lane read), and the next batch starts at it.  The first
// lane of a batch always commits, so the walk ends, and a batch of
// particles that land apart commits whole.
#include "fsg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace {

constexpr int PT = 256;  // particles per batch (threads of the one workgroup)
constexpr int NSPIRAL = 1024;

// tickCells' landing spiral (world.cpp:2113-2139): offsets in search order,
// filled by the host.
__constant__ int8_t kSpiral[2 * NSPIRAL];

struct PState {
    fsg_particle p;
    int kind;         // 0 keep, 1 drop, 2 land
    int land;         // cell index of a landing
    int bx0, by0, bx1, by1;  // bounding box of the cells read
};

__device__ __forceinline__ void grow_box(PState &s, int x, int y) {
    s.bx0 = min(s.bx0, x);
    s.by0 = min(s.by0, y);
    s.bx1 = max(s.bx1, x);
    s.by1 = max(s.by1, y);
}

__device__ __forceinline__ bool is_air(const uint64_t *cells, uint32_t airmask, int i) {
    return (airmask >> (uint32_t)(cells[i] & 15u)) & 1u;
}

// One particle's tick against the grid (the reference's loop body).
__device__ void move_particle(PState &s, const uint64_t *cells, int W, int H, uint32_t airmask, int4 zone) {
#pragma clang fp contract(off)
    fsg_particle c = s.p;
    s.kind = 0;
    s.land = -1;
    s.bx0 = s.by0 = 1 << 30;
    s.bx1 = s.by1 = -(1 << 30);
    if ((c.flags & FSG_PARTICLE_TEMPORARY) && c.lifetime <= 0) {
        s.kind = 1;
        return;
    }
    if (c.target_force != 0.0f) {
        const float tdx = c.target_x - c.x, tdy = c.target_y - c.y;
        const float nf = sqrtf(tdx * tdx + tdy * tdy);
        c.vx += tdx / nf * c.target_force;
        c.vy += tdy / nf * c.target_force;
        if (nf < 100.0f) {
            c.vx *= 0.95f;
            c.vy *= 0.95f;
        }
    }
    const int lx = (int)c.x, ly = (int)c.y;
    if (c.x < 0.0f || (int)c.x >= W || c.y < 0.0f || (int)c.y >= H) {
        s.kind = 1;
        return;
    }
    if (lx < zone.x || ly < zone.y || lx >= zone.z || ly >= zone.w) {  // outside the tick zone: kept as it is
        s.p = c;
        return;
    }
    c.vx += c.ax;
    c.vy += c.ay;
    const int div = (int)((fabsf(c.vx) + fabsf(c.vy)) + 1.0f);
    const float dvx = c.vx / (float)div, dvy = c.vy / (float)div;
    bool done = false;
    for (int k = 0; k < div && !done; k++) {
        c.x += dvx;
        c.y += dvy;
        if (c.x < 0.0f || (int)c.x >= W || c.y < 0.0f || (int)c.y >= H) {
            s.kind = 1;
            return;
        }
        const int px = (int)c.x, py = (int)c.y;
        if (c.flags & FSG_PARTICLE_PHASE) continue;
        grow_box(s, px, py);
        if (is_air(cells, airmask, px + py * W)) continue;
        if (c.in_object_state <= 1) c.in_object_state = 2;  // no OBJECT cells on this grid
        done = true;
        if (c.flags & FSG_PARTICLE_TEMPORARY) {
            s.kind = 1;
            return;
        }
        grow_box(s, lx, ly);
        if (!is_air(cells, airmask, lx + ly * W)) {
            for (int k2 = 0; k2 < NSPIRAL; k2++) {
                const int cx = (int)(c.x + (float)kSpiral[2 * k2]), cy = (int)(c.y + (float)kSpiral[2 * k2 + 1]);
                if (cx < 0 || cy < 0 || cx >= W || cy >= H) continue;
                grow_box(s, cx, cy);
                if (is_air(cells, airmask, cx + cy * W)) {
                    s.kind = 2;
                    s.land = cx + cy * W;
                    return;
                }
            }
            c.vy = -4.0f;
            c.y -= 16.0f;
        } else {
            s.kind = 2;
            s.land = lx + ly * W;
            return;
        }
    }
    if (!done && c.lifetime > 0) c.lifetime--;
    s.p = c;
}

struct PArgs {
    fsg_particle *p;       // the list, updated in place
    uint8_t *alive;        // per particle: kept after the pass
    uint32_t n;
    uint64_t *cells;
    int W, H, tiles_x;
    uint32_t airmask;
    int4 zone;             // tick zone [x, z) x [y, w)
    uint32_t *flags;       // chunk flags the next tick reads (null: no skipping)
    unsigned long long *landed;
};

__global__ __launch_bounds__(PT) void k_particles(PArgs a) {
    __shared__ int land_at[PT];
    __shared__ int first_bad, next;
    const int t = threadIdx.x;
    uint32_t base = 0;
    unsigned long long nl = 0;
    while (base < a.n) {  // base is the same in every thread
        const uint32_t i = base + (uint32_t)t;
        PState s;
        const bool mine = i < a.n;
        if (mine) {
            s.p = a.p[i];
            move_particle(s, a.cells, a.W, a.H, a.airmask, a.zone);
        }
        land_at[t] = (mine && s.kind == 2) ? s.land : -1;
        if (t == 0) first_bad = PT;
        __syncthreads();
        // the first lane whose reads cover an earlier lane's landing
        bool bad = false;
        if (mine && s.bx0 <= s.bx1) {
            for (int j = 0; j < t && !bad; j++) {
                const int c = land_at[j];
                if (c < 0) continue;
                const int cy = c / a.W, cx = c - cy * a.W;
                bad = cx >= s.bx0 && cx <= s.bx1 && cy >= s.by0 && cy <= s.by1;
            }
        }
        if (bad) atomicMin(&first_bad, t);
        __syncthreads();
        const int fb = first_bad;
        if (mine && t < fb) {
            if (s.kind == 2) {
                const fsg_cell &tl = s.p.tile;
                a.cells[s.land] = (uint64_t)tl.mat | ((uint64_t)(uint16_t)tl.temp << 16) | ((uint64_t)tl.rgba << 32);
                if (a.flags) {
                    const int cy = s.land / a.W, cx = s.land - cy * a.W;
                    atomicOr(&a.flags[(cy / fsg::TH) * a.tiles_x + cx / fsg::TW], fsg::F_WAKE_ALL);
                }
                nl++;
            }
            a.alive[i] = s.kind == 0 && !(s.p.y > (float)a.H);  // erase_if(y > height), world.cpp:2183
            if (s.kind == 0) a.p[i] = s.p;
        }
        if (t == 0) next = (int)min((uint32_t)fb, a.n - base);
        __syncthreads();
        base += (uint32_t)next;
        __threadfence_block();  // the landings are seen by the next batch's loads
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) nl += __shfl_xor(nl, off, 64);
    if ((t & 63) == 0 && nl) atomicAdd(a.landed, nl);
}

// Survivors in list order: one workgroup, a running offset over blocks.
__global__ __launch_bounds__(1024) void k_particles_compact(const fsg_particle *in, const uint8_t *alive,
                                                            uint32_t n, fsg_particle *out, uint32_t *count) {
    __shared__ uint32_t wsum[16], base_s;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (t == 0) base_s = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < n; b0 += 1024) {
        const uint32_t i = b0 + (uint32_t)t;
        const bool keep = i < n && alive[i];
        const uint64_t m = __ballot(keep);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (lane == 0) wsum[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = base_s;
        for (int k = 0; k < wv; k++) off += wsum[k];
        if (keep) out[off + below] = in[i];
        __syncthreads();
        if (t == 0) {
            uint32_t tot = 0;
            for (int k = 0; k < 16; k++) tot += wsum[k];
            base_s += tot;
        }
        __syncthreads();
    }
    if (t == 0) *count = base_s;
}

using fsg_impl::dev_alloc;
using fsg_impl::dev_free;

bool g_spiral_ready[64];
std::mutex g_spiral_mu;

int ensure_spiral(int device) {
    if (device < 0 || device >= 64) return FSG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_spiral_mu);
    if (g_spiral_ready[device]) return FSG_OK;
    int8_t sp[2 * NSPIRAL];
    int x = 0, y = 0, dx = 0, dy = -1, n = 0;
    for (int j = 0; j < 32 * 32; j++) {  // X = Y = 32 (world.cpp:2115-2139)
        if (-16 <= x && x <= 16 && -16 <= y && y <= 16) {
            sp[2 * n] = (int8_t)x;
            sp[2 * n + 1] = (int8_t)y;
            n++;
        }
        if (x == y || (x < 0 && x == -y) || (x > 0 && x == 1 - y)) {
            const int t = dx;
            dx = -dy;
            dy = t;
        }
        x += dx;
        y += dy;
    }
    if (n != NSPIRAL) return FSG_ERR_INVALID;
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(kSpiral), sp, sizeof(sp)));
    g_spiral_ready[device] = true;
    return FSG_OK;
}

// Grows the list buffers to hold `want` particles, keeping the list.
int grow_lists(fsg_world *w, size_t want) {
    if (w->part_cap >= want) return FSG_OK;
    fsg_ctx *c = w->ctx;
    const size_t nc = std::max(want, w->part_cap * 2);
    void *a = nullptr, *b = nullptr, *f = nullptr;
    if (dev_alloc(c, &a, nc * sizeof(fsg_particle))) return FSG_ERR_NOMEM;
    if (dev_alloc(c, &b, nc * sizeof(fsg_particle))) {
        dev_free(a);
        return FSG_ERR_NOMEM;
    }
    if (dev_alloc(c, &f, nc)) {
        dev_free(a);
        dev_free(b);
        return FSG_ERR_NOMEM;
    }
    if (w->npart) HIPCHK(hipMemcpyAsync(a, w->d_part, w->npart * sizeof(fsg_particle), hipMemcpyDeviceToDevice, c->stream));
    dev_free(w->d_part);
    dev_free(w->d_part2);
    dev_free(w->d_part_alive);
    w->d_part = static_cast<fsg_particle *>(a);
    w->d_part2 = static_cast<fsg_particle *>(b);
    w->d_part_alive = static_cast<uint8_t *>(f);
    w->part_cap = nc;
    return FSG_OK;
}

}  // namespace

extern "C" int fsg_world_particles_add(fsg_world *w, const fsg_particle *p, size_t n) {
    if (!w || (n && !p)) return FSG_ERR_INVALID;
    for (size_t i = 0; i < n; i++) {
        const uint16_t m = p[i].tile.mat;
        if ((m & 0xFE00u) || (uint32_t)(m & 0xFFu) >= w->cfg.n_materials) return FSG_ERR_INVALID;
        if (!std::isfinite(p[i].x) || !std::isfinite(p[i].y) || !std::isfinite(p[i].vx) || !std::isfinite(p[i].vy) ||
            !std::isfinite(p[i].ax) || !std::isfinite(p[i].ay) || !std::isfinite(p[i].target_x) ||
            !std::isfinite(p[i].target_y) || !std::isfinite(p[i].target_force) || p[i].in_object_state > 2 ||
            (p[i].flags & ~(FSG_PARTICLE_PHASE | FSG_PARTICLE_TEMPORARY)))
            return FSG_ERR_INVALID;
    }
    if (n == 0) return FSG_OK;
    if (w->npart + n > 0xFFFFFFFFull) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (int rc = grow_lists(w, w->npart + n)) return rc;
    HIPCHK(hipMemcpyAsync(w->d_part + w->npart, p, n * sizeof(fsg_particle), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    w->npart += n;
    return FSG_OK;
}

extern "C" int fsg_world_particles_tick(fsg_world *w, size_t *alive, size_t *landed) {
    if (!w) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (int rc = ensure_spiral(c->device)) return rc;
    const size_t n = w->npart;
    unsigned long long nl = 0;
    if (n) {
        if (!w->d_part_cnt) {
            void *q = nullptr;
            if (dev_alloc(c, &q, 16)) return FSG_ERR_NOMEM;
            w->d_part_cnt = static_cast<unsigned long long *>(q);
        }
        PArgs a;
        a.p = w->d_part;
        a.alive = w->d_part_alive;
        a.n = (uint32_t)n;
        a.cells = w->d.buf[w->d.tick & 1];
        a.W = w->d.W;
        a.H = w->d.H;
        a.tiles_x = w->d.tiles_x;
        a.zone = make_int4(w->d.zx0, w->d.zy0, w->d.zx1, w->d.zy1);
        a.airmask = 0;
        for (uint32_t m = 0; m < w->cfg.n_materials; m++)
            if (w->cfg.materials[m].type == FSG_AIR) a.airmask |= 1u << m;
        // landings wake their chunks for the next tick (no flags: it processes every chunk)
        a.flags = w->flags_valid ? w->d.flags[(w->d.tick + 1) & 1] : nullptr;
        a.landed = w->d_part_cnt;
        HIPCHK(hipMemsetAsync(w->d_part_cnt, 0, 16, c->stream));
        k_particles<<<1, PT, 0, c->stream>>>(a);
        k_particles_compact<<<1, 1024, 0, c->stream>>>(w->d_part, w->d_part_alive, (uint32_t)n, w->d_part2,
                                                        reinterpret_cast<uint32_t *>(w->d_part_cnt + 1));
        HIPCHK(hipGetLastError());
        unsigned long long h[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(h, w->d_part_cnt, 16, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        nl = h[0];
        std::swap(w->d_part, w->d_part2);  // the survivors, in order, are the list
        w->npart = (size_t)(uint32_t)h[1];
    }
    if (alive) *alive = w->npart;
    if (landed) *landed = (size_t)nl;
    return FSG_OK;
}

extern "C" int fsg_world_particles_get(fsg_world *w, fsg_particle *out, size_t cap, size_t *n) {
    if (!w || (cap && !out)) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    const size_t k = std::min(cap, w->npart);
    if (k) {
        HIPCHK(hipMemcpyAsync(out, w->d_part, k * sizeof(fsg_particle), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (n) *n = w->npart;
    return FSG_OK;
}

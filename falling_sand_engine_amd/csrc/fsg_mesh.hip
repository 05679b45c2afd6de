// fsg_mesh.hip -- outlines of solid cells for the host physics engine:
// the marching-squares step of world::updateChunkMesh (world.cpp:722-895)
// and of updateRigidBodyHitbox (world.cpp:288-510), on the device.
//
// One wave per region (a 128x128 chunk of the grid, or an island's body
// image).  The region's mask, the reference's edge candidates (a set cell
// whose right, lower and lower-right neighbours are not all set, and whose
// marching-squares vertex value is not 15) and the edgeSeen marks are bit
// planes in LDS, built by all 64 lanes.  The reference then walks candidates
// in row order, skipping marked ones; each walk (FindPerimeter,
// physics_math.cpp:1881-1965) is serial, so the wave runs it in lockstep
// with wave-uniform values, and the lanes share the parallel parts: the
// search for the next candidate (64 words at a time), the edgeSeen marks
// of each run, and the farthest-point search of Ramer-Douglas-Peucker
// (simplify, physics_math.cpp:1766-1812; f32 arithmetic as the reference
// does it, with no fused multiply-add, so its decisions are the reference's).
// Kept loops (3 or more vertices) are appended to output pools with atomics;
// the host orders them by region and trace order.
#include "fsg_internal.hpp"

#include <algorithm>
#include <cstring>

namespace {

constexpr int MW = 64;                                  // threads per workgroup: one wave per region
constexpr int MBITS = FSG_MESH_MAX_CELLS;               // mask bits per region
constexpr int MWORDS = MBITS / 32;
constexpr int MAXV = (128 + 1) * (128 + 1) + 2;         // runs of one loop of a 16384-cell region, bound
constexpr int LCAP = 2048;                              // loops up to this many runs stay in LDS

struct MeshSrc {
    const uint64_t *cells;  // region r's cell (x, y) is cells[off[r] + y * pitch[r] + x]
    uint32_t setmask;       // material ids whose cells are set (SOLID for the grid, not AIR for images)
};

struct MeshRegionDev {
    int64_t off;            // first cell
    int32_t pitch, w, h;
    uint32_t region;        // index in the caller's array
};

struct MeshOut {
    fsg_mesh_loop *loops;   // pool of loops (region, seq in first_loop order via `seq`)
    uint32_t *seq;          // per pooled loop: its trace order within its region
    uint32_t *verts;        // pool of vertices, (x | y << 16)
    uint32_t loop_cap, vert_cap;
    unsigned int *counters; // [0] loops, [1] vertices, [2] overflow, [3] walk failures
    uint32_t *scratch;      // per workgroup: MAXV vertices + MAXV stack entries
    uint32_t *regcnt;       // per caller's region: its loops (zero for regions not traced)
};

__device__ __forceinline__ bool bit_at(const uint32_t *b, int i) { return (b[i >> 5] >> (i & 31)) & 1u; }

// isSet of vertex coordinates (x, y): cell (x - 1, y - 1) inside the region.
__device__ __forceinline__ uint32_t vset(const uint32_t *m, int w, int h, int x, int y) {
    return (x <= 0 || x > w || y <= 0 || y > h) ? 0u : (uint32_t)bit_at(m, (y - 1) * w + (x - 1));
}
__device__ __forceinline__ int ms_value(const uint32_t *m, int w, int h, int x, int y) {
    return (int)(vset(m, w, h, x, y) | vset(m, w, h, x + 1, y) << 1 | vset(m, w, h, x, y + 1) << 2 |
                 vset(m, w, h, x + 1, y + 1) << 3);
}

// pDistance (physics_math.cpp:1814-1843) in f32, each operation rounded on
// its own, as the reference's compiled code rounds it.
__device__ float p_distance(float x, float y, float x1, float y1, float x2, float y2) {
#pragma clang fp contract(off)
    const float A = x - x1, B = y - y1, C = x2 - x1, D = y2 - y1;
    const float dot = A * C + B * D;
    const float len_sq = C * C + D * D;
    float param = -1.0f;
    if (len_sq != 0.0f) param = dot / len_sq;
    float xx, yy;
    if (param < 0.0f) {
        xx = x1;
        yy = y1;
    } else if (param > 1.0f) {
        xx = x2;
        yy = y2;
    } else {
        xx = x1 + param * C;
        yy = y1 + param * D;
    }
    const float dx = x - xx, dy = y - yy;
    return sqrtf(dx * dx + dy * dy);
}

__device__ __forceinline__ float vx(uint32_t v) { return (float)(v & 0xFFFFu); }
__device__ __forceinline__ float vy(uint32_t v) { return (float)(v >> 16); }

__global__ __launch_bounds__(MW) void k_mesh(MeshSrc src, const MeshRegionDev *regs, int nreg, MeshOut out) {
    __shared__ uint32_t mask[MWORDS], cand[MWORDS], seen[MWORDS];
    __shared__ uint32_t keep[MAXV / 32 + 1];
    __shared__ uint32_t lmesh[LCAP], lstack[LCAP];
    const int lane = threadIdx.x;
    // the loop being traced: its run ends (vertices) and the RDP sections
    // (i | j << 16), in LDS for loops of up to LCAP runs, else in the
    // workgroup's global scratch (the walk writes both until it knows)
    uint32_t *gmesh = out.scratch + (size_t)blockIdx.x * 2 * MAXV;
    uint32_t *gstack = gmesh + MAXV;
    for (int ri = blockIdx.x; ri < nreg; ri += gridDim.x) {
        const MeshRegionDev R = regs[ri];
        const int w = R.w, h = R.h, size = w * h, nwords = (size + 31) >> 5;
        // the mask, 64 cells per ballot
        for (int i0 = 0; i0 < size; i0 += 64) {
            const int i = i0 + lane;
            bool on = false;
            if (i < size) {
                const int y = i / w, x = i - y * w;
                const uint32_t m = (uint32_t)src.cells[R.off + (int64_t)y * R.pitch + x] & 15u;
                on = (src.setmask >> m) & 1u;
            }
            const uint64_t b = __ballot(on);
            if (lane == 0) mask[i0 >> 5] = (uint32_t)b;
            if (lane == 1 && (i0 >> 5) + 1 < MWORDS) mask[(i0 >> 5) + 1] = (uint32_t)(b >> 32);
        }
        __syncthreads();
        // edge candidates (world.cpp:774-789) whose vertex value is neither
        // 0 nor 15 (world.cpp:826-831); edgeSeen cleared
        for (int i0 = 0; i0 < size; i0 += 64) {
            const int i = i0 + lane;
            bool c = false;
            if (i < size && bit_at(mask, i)) {
                const int y = i / w, x = i - y * w;
                int nb = 0;
                if (x + 1 < w) nb += bit_at(mask, i + 1);
                if (y + 1 < h) nb += bit_at(mask, i + w);
                if (y + 1 < h && x + 1 < w) nb += bit_at(mask, i + w + 1);
                const int v = ms_value(mask, w, h, x, y);
                c = nb != 3 && v != 0 && v != 15;
            }
            const uint64_t b = __ballot(c);
            if (lane == 0) {
                cand[i0 >> 5] = (uint32_t)b;
                seen[i0 >> 5] = 0u;
            }
            if (lane == 1 && (i0 >> 5) + 1 < MWORDS) {
                cand[(i0 >> 5) + 1] = (uint32_t)(b >> 32);
                seen[(i0 >> 5) + 1] = 0u;
            }
        }
        __syncthreads();
        uint32_t nloops_region = 0;
        int look = 0;  // wave-uniform from here on
        for (;;) {
            // next candidate at or after look that no loop has marked
            int found = -1;
            for (int w0 = look >> 5; w0 < nwords && found < 0; w0 += 64) {
                const int wi = w0 + lane;
                uint32_t bits = 0;
                if (wi < nwords) {
                    bits = cand[wi] & ~seen[wi];
                    if (wi == (look >> 5)) bits &= ~0u << (look & 31);
                }
                const uint64_t any = __ballot(bits != 0);
                if (any) {
                    const int first = __builtin_ctzll(any);
                    const uint32_t fb = (uint32_t)__shfl((int)bits, first, 64);
                    found = ((w0 + first) << 5) + __builtin_ctz(fb);
                }
            }
            if (found < 0) break;
            look = found + 1;
            const int x0 = found % w, y0 = found / w;
            // FindPerimeter from vertex (x0, y0); each run's end vertex goes to
            // mesh[], and its cells to edgeSeen from its start vertex
            int x = x0, y = y0, px = 0, py = 0, nr = 0, rdx = 0, rdy = 0;
            int sx = x0, sy = y0;  // start vertex of the open run
            bool ok = true;
            const int limit = 4 * (w + 1) * (h + 1);
            auto close_run = [&]() {
                const int len = rdx != 0 ? abs(rdx) : abs(rdy);
                for (int k = lane; k < len; k += 64) {
                    int ilx = sx + (rdx != 0 ? (rdx < 0 ? -k : k) : 0);
                    int ily = sy - (rdy != 0 ? (rdy < 0 ? -k : k) : 0);
                    ilx = min(max(ilx, 0), w - 1);
                    ily = min(max(ily, 0), h - 1);
                    const int ind = ilx + ily * w;
                    atomicOr(&seen[ind >> 5], 1u << (ind & 31));
                }
                sx += rdx;
                sy -= rdy;
                if (lane == 0) {
                    const uint32_t v = (uint32_t)sx | ((uint32_t)sy << 16);
                    if (nr < LCAP) lmesh[nr] = v;
                    if (nr < MAXV) gmesh[nr] = v;
                }
                nr++;
            };
            for (int step = 0;; step++) {
                if (step > limit) {
                    ok = false;
                    break;
                }
                int dx = 0, dy = 0;
                switch (ms_value(mask, w, h, x, y)) {
                case 1: case 5: case 13: dy = 1; break;
                case 2: case 3: case 7: dx = 1; break;
                case 4: case 12: case 14: dx = -1; break;
                case 8: case 10: case 11: dy = -1; break;
                case 6: dx = (px == 0 && py == 1) ? -1 : 1; break;
                case 9: dy = (px == 1 && py == 0) ? 1 : -1; break;
                default: ok = false; break;
                }
                if (!ok) break;
                if (step > 0 && dx == px && dy == py) {
                    rdx += dx;
                    rdy += dy;
                } else {
                    if (step > 0) close_run();
                    rdx = dx;
                    rdy = dy;
                    px = dx;
                    py = dy;
                }
                x += dx;
                y -= dy;
                if (x == x0 && y == y0) break;
            }
            if (ok) close_run();
            __syncthreads();  // edgeSeen and mesh[] written
            if (!ok || nr > MAXV) {
                if (lane == 0) atomicAdd(&out.counters[3], 1u);
                continue;
            }
            uint32_t *mesh = nr <= LCAP ? lmesh : gmesh;
            uint32_t *stack = nr <= LCAP ? lstack : gstack;
            // Ramer-Douglas-Peucker, tolerance 1, over mesh[0 .. nr)
            for (int k = lane; k < (nr + 31) / 32; k += 64) keep[k] = ~0u;
            __syncthreads();
            if (nr > 2) {
                int sp = 0;
                if (lane == 0) stack[0] = 0u | ((uint32_t)(nr - 1) << 16);
                sp = 1;
                __syncthreads();
                while (sp > 0) {
                    const uint32_t e = stack[sp - 1];
                    __syncthreads();
                    sp--;
                    const int i = (int)(e & 0xFFFFu), j = (int)(e >> 16);
                    if (i + 1 == j) continue;
                    const uint32_t vi = mesh[i], vj = mesh[j];
                    float best = -1.0f;
                    int bk = i;
                    for (int k = i + 1 + lane; k < j; k += 64) {
                        const uint32_t vk = mesh[k];
                        const float d = p_distance(vx(vk), vy(vk), vx(vi), vy(vi), vx(vj), vy(vj));
                        if (d > best) {
                            best = d;
                            bk = k;
                        }
                    }
                    // the first largest distance, as the reference's strict > scan finds it
                    for (int off = 32; off > 0; off >>= 1) {
                        const float ob = __shfl_xor(best, off, 64);
                        const int ok2 = __shfl_xor(bk, off, 64);
                        if (ob > best || (ob == best && ok2 < bk)) {
                            best = ob;
                            bk = ok2;
                        }
                    }
                    if (best <= 1.0f) {
                        for (int k = i + 1 + lane; k < j; k += 64) atomicAnd(&keep[k >> 5], ~(1u << (k & 31)));
                    } else {
                        if (lane == 0) {
                            stack[sp] = (uint32_t)bk | ((uint32_t)j << 16);
                            stack[sp + 1] = (uint32_t)i | ((uint32_t)bk << 16);
                        }
                        sp += 2;
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
            // kept vertices: count, orientation of the reversed list (GetOrientation)
            int ns = 0;
            for (int k = lane; k < nr; k += 64) ns += bit_at(keep, k);
            for (int off = 32; off > 0; off >>= 1) ns += __shfl_xor(ns, off, 64);
            if (ns < 3) continue;
            // compact the kept vertices to the front of mesh[], in order, 64 at
            // a time (every lane reads before any writes, and writes land at or
            // before the positions read)
            {
                int m = 0;
                for (int k0 = 0; k0 < nr; k0 += 64) {
                    const int k = k0 + lane;
                    const bool kp = k < nr && bit_at(keep, k);
                    const uint32_t v = kp ? mesh[k] : 0u;
                    const uint64_t b = __ballot(kp);
                    if (kp) mesh[m + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] = v;
                    m += __popcll(b);
                }
            }
            __syncthreads();
            long long area = 0;
            for (int a = lane; a < ns; a += 64) {
                const int p = ns - 1 - a, q = (a + 1 == ns) ? ns - 1 : ns - 2 - a;
                const uint32_t vp = mesh[p], vq = mesh[q];
                area += (long long)(vp & 0xFFFFu) * (long long)(vq >> 16) - (long long)(vp >> 16) * (long long)(vq & 0xFFFFu);
            }
            for (int off = 32; off > 0; off >>= 1) area += __shfl_xor(area, off, 64);
            uint32_t lslot = 0, vslot = 0;
            if (lane == 0) {
                lslot = atomicAdd(&out.counters[0], 1u);
                vslot = atomicAdd(&out.counters[1], (unsigned)ns);
            }
            lslot = (uint32_t)__shfl((int)lslot, 0, 64);
            vslot = (uint32_t)__shfl((int)vslot, 0, 64);
            if (lslot >= out.loop_cap || (uint64_t)vslot + (uint64_t)ns > out.vert_cap) {
                if (lane == 0) atomicOr(&out.counters[2], 1u);
            } else {
                for (int k = lane; k < ns; k += 64) out.verts[vslot + k] = mesh[k];
                if (lane == 0) {
                    fsg_mesh_loop L;
                    L.first = vslot;
                    L.count = (uint32_t)ns;
                    L.hole = area < 0 ? 1 : 0;
                    L.start_x = x0;
                    L.start_y = y0;
                    L.region = R.region;
                    out.loops[lslot] = L;
                    out.seq[lslot] = nloops_region;
                }
            }
            nloops_region++;
            __syncthreads();
        }
        if (lane == 0) out.regcnt[R.region] = nloops_region;
        __syncthreads();
    }
}

// ---- ordering on the device: regions in the caller's order, loops in trace
// order within each ---------------------------------------------------------
// Exclusive scan of n counts in one workgroup (1024 at a time, carried).
__global__ __launch_bounds__(1024) void k_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < n; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < n ? in[i] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {
            const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < n) out[i] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += part[1023];
        __syncthreads();
    }
}

// Pooled loop l goes to place regbase[region] + seq: its pool index and count.
__global__ void k_mesh_place(const fsg_mesh_loop *loops, const uint32_t *seq, uint32_t nl, const uint32_t *regbase,
                             uint32_t *perm, uint32_t *cnt) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nl) return;
    const uint32_t f = regbase[loops[l].region] + seq[l];
    perm[f] = l;
    cnt[f] = loops[l].count;
}

// The ordered loops and their vertices as float pairs.
__global__ void k_mesh_emit(const fsg_mesh_loop *loops, const uint32_t *verts, uint32_t nl, const uint32_t *perm,
                            const uint32_t *vbase, fsg_mesh_loop *oloops, float2 *overts) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nl) return;
    fsg_mesh_loop L = loops[perm[f]];
    const uint32_t from = L.first;
    L.first = vbase[f];
    oloops[f] = L;
    for (uint32_t v = 0; v < L.count; v++) {
        const uint32_t p = verts[from + v];
        overts[L.first + v] = make_float2((float)(p & 0xFFFFu), (float)(p >> 16));
    }
}

using fsg_impl::dev_alloc;
using fsg_impl::dev_free;

int grow_buf(fsg_ctx *c, void **p, size_t *cap, size_t bytes) {
    if (*cap >= bytes) return FSG_OK;
    dev_free(*p);
    *p = nullptr;
    *cap = 0;
    if (dev_alloc(c, p, bytes)) return FSG_ERR_NOMEM;
    *cap = bytes;
    return FSG_OK;
}

// Runs the mesher over regions (already validated) and leaves the ordered
// result in w->mesh_loops / w->mesh_verts.
int run_mesh(fsg_world *w, const MeshSrc &src, std::vector<MeshRegionDev> &regs, fsg_mesh_region *out_regions,
             int nout, uint32_t *nloops, uint32_t *nverts) {
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    w->mesh_nloops = 0;
    w->mesh_nverts = 0;
    for (int i = 0; i < nout; i++) {
        out_regions[i].first_loop = 0;
        out_regions[i].nloops = 0;
    }
    const int n = (int)regs.size();
    if (n > 0) {
        const int grid = std::min(n, 1024);
        auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
        const size_t b_regs = al((size_t)n * sizeof(MeshRegionDev));
        if (grow_buf(c, &w->d_mesh_regs, &w->mesh_regs_cap, b_regs + 2 * al((size_t)nout * 4)) ||
            grow_buf(c, &w->d_mesh_scratch, &w->mesh_scratch_cap, (size_t)grid * 2 * MAXV * 4))
            return FSG_ERR_NOMEM;
        uint32_t *regcnt = reinterpret_cast<uint32_t *>(static_cast<char *>(w->d_mesh_regs) + b_regs);
        uint32_t *regbase = regcnt + al((size_t)nout * 4) / 4;
        HIPCHK(hipMemcpyAsync(w->d_mesh_regs, regs.data(), (size_t)n * sizeof(MeshRegionDev), hipMemcpyHostToDevice,
                              c->stream));
        // pools sized from the last call, grown and the launch repeated on overflow
        uint32_t loop_cap = std::max<uint32_t>(w->mesh_loop_hint, 4096u);
        uint32_t vert_cap = std::max<uint32_t>(w->mesh_vert_hint, 65536u);
        bool done = false;
        for (int attempt = 0; attempt < 2 && !done; attempt++) {
            // the pool: traced loops, their seq, vertices; then the ordered
            // result: place -> pool index, counts, vertex bases, loops, vertices
            const size_t b_loops = al((size_t)loop_cap * sizeof(fsg_mesh_loop)), b_l4 = al((size_t)loop_cap * 4);
            const size_t b_v4 = al((size_t)vert_cap * 4);
            const size_t total = 2 * b_loops + 4 * b_l4 + b_v4 + al((size_t)vert_cap * 8) + 256;
            if (grow_buf(c, &w->d_mesh_pool, &w->mesh_pool_cap, total)) return FSG_ERR_NOMEM;
            char *pool = static_cast<char *>(w->d_mesh_pool);
            MeshOut o;
            o.loops = reinterpret_cast<fsg_mesh_loop *>(pool);
            pool += b_loops;
            o.seq = reinterpret_cast<uint32_t *>(pool);
            pool += b_l4;
            o.verts = reinterpret_cast<uint32_t *>(pool);
            pool += b_v4;
            o.counters = reinterpret_cast<unsigned int *>(pool);
            pool += 256;
            uint32_t *perm = reinterpret_cast<uint32_t *>(pool);
            pool += b_l4;
            uint32_t *lcnt = reinterpret_cast<uint32_t *>(pool);
            pool += b_l4;
            uint32_t *vbase = reinterpret_cast<uint32_t *>(pool);
            pool += b_l4;
            fsg_mesh_loop *oloops = reinterpret_cast<fsg_mesh_loop *>(pool);
            pool += b_loops;
            float2 *overts = reinterpret_cast<float2 *>(pool);
            o.loop_cap = loop_cap;
            o.vert_cap = vert_cap;
            o.scratch = static_cast<uint32_t *>(w->d_mesh_scratch);
            o.regcnt = regcnt;
            HIPCHK(hipMemsetAsync(o.counters, 0, 16, c->stream));
            HIPCHK(hipMemsetAsync(regcnt, 0, (size_t)nout * 4, c->stream));
            k_mesh<<<grid, MW, 0, c->stream>>>(src, static_cast<const MeshRegionDev *>(w->d_mesh_regs), n, o);
            HIPCHK(hipGetLastError());
            unsigned int cnt[4];
            HIPCHK(hipMemcpyAsync(cnt, o.counters, 16, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (cnt[3]) return FSG_ERR_DEVICE;  // a walk that did not close: not a perimeter of a finite mask
            if (cnt[2]) {
                loop_cap = std::max(loop_cap, cnt[0] + 1024u);
                vert_cap = std::max(vert_cap, cnt[1] + 65536u);
                continue;
            }
            w->mesh_loop_hint = cnt[0];
            w->mesh_vert_hint = cnt[1];
            std::vector<uint32_t> rc((size_t)nout);
            if (cnt[0]) {
                k_scan_u32<<<1, 1024, 0, c->stream>>>(regcnt, regbase, (uint32_t)nout);
                k_mesh_place<<<(cnt[0] + 255) / 256, 256, 0, c->stream>>>(o.loops, o.seq, cnt[0], regbase, perm, lcnt);
                k_scan_u32<<<1, 1024, 0, c->stream>>>(lcnt, vbase, cnt[0]);
                k_mesh_emit<<<(cnt[0] + 255) / 256, 256, 0, c->stream>>>(o.loops, o.verts, cnt[0], perm, vbase,
                                                                         oloops, overts);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(rc.data(), regcnt, (size_t)nout * 4, hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
            }
            uint32_t first = 0;
            for (int r = 0; r < nout; r++) {
                out_regions[r].first_loop = rc[r] ? first : 0;
                out_regions[r].nloops = rc[r];
                first += rc[r];
            }
            w->d_mesh_loops = oloops;
            w->d_mesh_verts = reinterpret_cast<float *>(overts);
            w->mesh_nloops = cnt[0];
            w->mesh_nverts = cnt[1];
            done = true;
        }
        if (!done) return FSG_ERR_NOMEM;  // the second launch had pools of the first one's totals
    }
    if (nloops) *nloops = w->mesh_nloops;
    if (nverts) *nverts = w->mesh_nverts;
    return FSG_OK;
}

uint32_t type_mask(const fsg_config &cfg, bool solid) {
    uint32_t m = 0;
    for (uint32_t i = 0; i < cfg.n_materials; i++) {
        const uint8_t t = cfg.materials[i].type;
        if (solid ? t == FSG_SOLID : t != FSG_AIR) m |= 1u << i;
    }
    if (solid) m |= 1u << fsg::WALL;
    return m;
}

}  // namespace

extern "C" int fsg_world_mesh(fsg_world *w, fsg_mesh_region *regions, int n, uint32_t *nloops, uint32_t *nverts) {
    if (!w || n < 0 || (n && !regions)) return FSG_ERR_INVALID;
    std::vector<MeshRegionDev> regs;
    for (int i = 0; i < n; i++) {
        fsg_mesh_region &R = regions[i];
        if (R.w <= 0 || R.h <= 0 || (int64_t)R.w * R.h > FSG_MESH_MAX_CELLS) return FSG_ERR_INVALID;
    }
    for (int i = 0; i < n; i++) {
        fsg_mesh_region &R = regions[i];
        // the reference meshes a chunk only when it ends before the world's
        // last row and column (world.cpp:734-736)
        if (R.x0 < 0 || R.y0 < 0 || (int64_t)R.x0 + R.w >= w->d.W || (int64_t)R.y0 + R.h >= w->d.H) {
            R.status = FSG_MESH_EDGE;
            continue;
        }
        R.status = FSG_OK;
        regs.push_back(MeshRegionDev{(int64_t)R.y0 * w->d.W + R.x0, w->d.W, R.w, R.h, (uint32_t)i});
    }
    MeshSrc src{w->d.buf[w->d.tick & 1], type_mask(w->cfg, true)};
    return run_mesh(w, src, regs, regions, n, nloops, nverts);
}

extern "C" int fsg_world_mesh_islands(fsg_world *w, const fsg_island *islands, int n, fsg_mesh_region *regions,
                                      uint32_t *nloops, uint32_t *nverts) {
    if (!w || n < 0 || (n && (!islands || !regions))) return FSG_ERR_INVALID;
    std::vector<MeshRegionDev> regs;
    for (int i = 0; i < n; i++) {
        const fsg_island &I = islands[i];
        fsg_mesh_region &R = regions[i];
        R.x0 = 0;
        R.y0 = 0;
        R.w = I.w;
        R.h = I.h;
        // no outline for other statuses, or for an image one pixel wide or
        // high (world.cpp:331, 365)
        if (I.status != FSG_ISLAND_BODY || I.w <= 1 || I.h <= 1) {
            R.status = FSG_MESH_EDGE;
            continue;
        }
        if ((int64_t)I.w * I.h > FSG_MESH_MAX_CELLS ||
            I.pixel_offset + (uint64_t)I.w * I.h > (uint64_t)w->isl_pix_used)
            return FSG_ERR_INVALID;
        R.status = FSG_OK;
        regs.push_back(MeshRegionDev{(int64_t)I.pixel_offset, I.w, I.w, I.h, (uint32_t)i});
    }
    MeshSrc src{w->d_isl_pix, type_mask(w->cfg, false)};
    return run_mesh(w, src, regs, regions, n, nloops, nverts);
}

extern "C" int fsg_world_mesh_download(fsg_world *w, fsg_mesh_loop *loops, uint32_t loop_cap, float *verts,
                                       uint32_t vert_cap) {
    if (!w) return FSG_ERR_INVALID;
    const size_t nl = w->mesh_nloops, nv = w->mesh_nverts;
    if ((nl && (!loops || loop_cap < nl)) || (nv && (!verts || vert_cap < nv))) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (nl) HIPCHK(hipMemcpyAsync(loops, w->d_mesh_loops, nl * sizeof(fsg_mesh_loop), hipMemcpyDeviceToHost, c->stream));
    if (nv) HIPCHK(hipMemcpyAsync(verts, w->d_mesh_verts, nv * 2 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FSG_OK;
}

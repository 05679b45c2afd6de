// fsg_noalloc.hip -- fsg_world_set_cells_noalloc: the batched setTile of
// fsg_world_set_cells for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  fsg_world_set_cells stages its edits in device blocks it
// allocates per call and drops repeated positions on the host; this variant
// stages the caller's edits in a file-scope device array and drops repeated
// positions on the device, with a hash table in the same array, so it calls
// no hipMalloc, hipFree or allocating routine.
//
// The array is one per GPU (a __device__ variable exists once per device the
// module is loaded on).  Host code never passes its name to a kernel or a
// copy: its device address, per device, comes from hipGetSymbolAddress at
// context creation (fsg_noalloc_setup) and is what the entry point uses.
// Calls on one GPU are serialised by a per-device lock held for the whole
// (synchronous) call, so two overlapping calls cannot interleave on it.
#include "fsg_internal.hpp"

#include <mutex>

namespace {

constexpr uint32_t SET_CELLS_SLOTS = 2u * FSG_SET_CELLS_NOALLOC_MAX;  // hash slots: load factor <= 1/2
constexpr uint32_t EMPTY = 0xFFFFFFFFu;
static_assert((SET_CELLS_SLOTS & (SET_CELLS_SLOTS - 1)) == 0, "power-of-two hash table");

// Layout of the array: the staged edits (fsg_cell_edit, 16 bytes each), then
// the hash table's keys (cell index, u32) and values (1 + the last edit's
// index, u32).
constexpr size_t OFF_KEYS = (size_t)FSG_SET_CELLS_NOALLOC_MAX * sizeof(fsg_cell_edit);
constexpr size_t OFF_VALS = OFF_KEYS + (size_t)SET_CELLS_SLOTS * 4;
constexpr size_t SCRATCH_BYTES = OFF_VALS + (size_t)SET_CELLS_SLOTS * 4;
static_assert(sizeof(fsg_cell_edit) == 16, "edit layout");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void *g_scratch[MAX_DEVICES];  // device address of `scratch` per device (set up at context creation)
std::mutex g_mu[MAX_DEVICES];  // one call at a time per device
std::mutex g_setup_mu;

__device__ __forceinline__ uint32_t slot_of(uint32_t key) { return (key * 0x9E3779B1u) & (SET_CELLS_SLOTS - 1u); }

__global__ void k_hash_clear(uint32_t *keys, uint32_t *vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < SET_CELLS_SLOTS) {
        keys[i] = EMPTY;
        vals[i] = 0;
    }
}

// Each in-world edit records its index against its cell: the largest index
// of a cell (its last edit, as consecutive setTile calls leave it) wins.
// Linear probing in a table at most half full: every probe sequence ends
// within SET_CELLS_SLOTS steps.
__global__ void k_hash_insert(const fsg_cell_edit *e, uint32_t n, int W, int H, uint32_t *keys, uint32_t *vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = e[i].x, y = e[i].y;
    if (x < 0 || y < 0 || x >= W || y >= H) return;  // setTile ignores OOB (world.cpp:1005)
    const uint32_t key = (uint32_t)y * (uint32_t)W + (uint32_t)x;
    uint32_t s = slot_of(key);
    for (uint32_t probe = 0; probe < SET_CELLS_SLOTS; probe++) {
        const uint32_t prev = atomicCAS(&keys[s], EMPTY, key);
        if (prev == EMPTY || prev == key) {
            atomicMax(&vals[s], i + 1u);
            return;
        }
        s = (s + 1u) & (SET_CELLS_SLOTS - 1u);
    }
}

// The last edit of each cell writes it.
__global__ void k_apply_last(const fsg_cell_edit *e, uint32_t n, int W, int H, const uint32_t *keys,
                             const uint32_t *vals, uint64_t *cells) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = e[i].x, y = e[i].y;
    if (x < 0 || y < 0 || x >= W || y >= H) return;
    const uint32_t key = (uint32_t)y * (uint32_t)W + (uint32_t)x;
    uint32_t s = slot_of(key);
    for (uint32_t probe = 0; probe < SET_CELLS_SLOTS; probe++) {
        const uint32_t k = keys[s];
        if (k == key) {
            if (vals[s] == i + 1u) {
                const fsg_cell c = e[i].cell;
                cells[key] = (uint64_t)c.mat | ((uint64_t)(uint16_t)c.temp << 16) | ((uint64_t)c.rgba << 32);
            }
            return;
        }
        if (k == EMPTY) return;  // not reached: every in-world edit was inserted
        s = (s + 1u) & (SET_CELLS_SLOTS - 1u);
    }
}

}  // namespace

namespace fsg_impl {
// Looks up the array's address on `device` (made current by the caller) once
// per device; called by fsg_ctx_create.
int noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return FSG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (g_scratch[device]) return FSG_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return FSG_ERR_DEVICE;
    }
    g_scratch[device] = p;
    return FSG_OK;
}
}  // namespace fsg_impl

extern "C" int fsg_world_set_cells_noalloc(fsg_world *w, const fsg_cell_edit *edits, size_t n) {
    if (!w || (!edits && n)) return FSG_ERR_INVALID;
    if (n > FSG_SET_CELLS_NOALLOC_MAX) return FSG_ERR_INVALID;  // more edits than the array stages
    if (n == 0) return FSG_OK;
    for (size_t i = 0; i < n; i++) {
        const uint16_t m = edits[i].cell.mat;
        if ((m & 0xFE00u) || (uint32_t)(m & 0xFFu) >= w->cfg.n_materials) return FSG_ERR_INVALID;
    }
    fsg_ctx *c = w->ctx;
    const int dev = c->device;
    if (dev < 0 || dev >= MAX_DEVICES || !g_scratch[dev]) return FSG_ERR_STATE;  // no context set it up
    std::lock_guard<std::mutex> lk(g_mu[dev]);
    HIPCHK(hipSetDevice(dev));
    unsigned char *base = static_cast<unsigned char *>(g_scratch[dev]);
    fsg_cell_edit *d_e = reinterpret_cast<fsg_cell_edit *>(base);
    uint32_t *keys = reinterpret_cast<uint32_t *>(base + OFF_KEYS);
    uint32_t *vals = reinterpret_cast<uint32_t *>(base + OFF_VALS);
    const uint32_t nn = (uint32_t)n;
    const unsigned g = (nn + 255u) / 256u;
    HIPCHK(hipMemcpyAsync(d_e, edits, n * sizeof(fsg_cell_edit), hipMemcpyHostToDevice, c->stream));
    k_hash_clear<<<SET_CELLS_SLOTS / 256u, 256, 0, c->stream>>>(keys, vals);
    k_hash_insert<<<g, 256, 0, c->stream>>>(d_e, nn, w->d.W, w->d.H, keys, vals);
    k_apply_last<<<g, 256, 0, c->stream>>>(d_e, nn, w->d.W, w->d.H, keys, vals, w->d.buf[w->d.tick & 1]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    w->d.force_tick = w->d.tick;
    return FSG_OK;
}

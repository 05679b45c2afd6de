// fsg_pack.hip -- chunk .pack images made and read on the device.
//
// A .pack image (chunk.cpp:219-330 ChunkWrite, read back by chunk.cpp:74-217
// ChunkRead) holds one 128x128 chunk:
//   i8 generation phase, i32 src_size (393216), i32 compressed_size,
//   i32 src_size2 (65536), i32 compressed_size2,
//   LZ4 block of the 12-byte MaterialInstanceData records {u32 material
//   index, u32 colour, i16 temperature, 2 pad bytes} of tiles[] then layer2[],
//   LZ4 block of the u32 background.
// A record's index is the file's material index of the cell's material id
// (the world's pack map; by default the id itself), its colour and
// temperature are the cell's, and the cell's MOVED flag travels in bit 0 of
// pad byte 10, which ChunkRead ignores; so an image round-trips a chunk
// exactly, and every index it holds is a material the reader knows.
//
// Encoding: one workgroup per chunk and block.  The reference asks LZ4 for a
// hash-chain search (LZ4_compress_fast, acceleration 10); any parse of the
// block format decodes the same, so the encoder uses one that parallelises:
// every byte is compared with the byte one record back and one record row
// (128 records) back, runs of equal bytes of 4 or more become matches with
// that offset, and the rest are literals.  A thread parses one record row;
// matches do not cross rows, literal runs do (a prefix scan over the rows
// gives each row the literals carried into it).  Runs of identical cells,
// identical rows, and the fields that repeat between neighbouring records
// (material, temperature, alpha, padding) all become matches.  The last 12
// bytes of a block are literals, as the format requires.
//
// Decoding: one wave per block (fsg_lz4.hpp), four waves per workgroup, into
// a record staging area; then one workgroup per chunk checks every record and
// writes the cells.
#include "fsg_internal.hpp"
#include "fsg_lz4.hpp"

#include <algorithm>
#include <cstring>

namespace {

constexpr int CW = FSG_CHUNK;                   // chunk side (const.h:19-20)
constexpr int RA = 12, RB = 4;                  // record sizes of the two blocks
constexpr int NA = CW * CW * 2 * RA;            // tiles + layer2 records: 393216
constexpr int NB = CW * CW * RB;                // background: 65536
constexpr int bound(int n) { return n + n / 255 + 16; }  // LZ4_compressBound (lz4.h)
constexpr int CAPA = bound(NA), CAPB = bound(NB);
constexpr size_t SLOT = ((size_t)CAPA + CAPB + 255) & ~(size_t)255;  // encoder output per chunk
constexpr size_t RAW = (size_t)NA + NB;          // decoder output per chunk
constexpr int HDR = 17;
constexpr int KX = 0, KL = 1, KU = 2;            // byte kinds: literal, as one record back, as one row back
constexpr int PT = 256;                          // encoder threads (one per record row of block A)

struct PackSrc {
    const uint64_t *cells;  // world grid
    int W;
    const uint64_t *l2;     // region layer2 cells, row width rw (null: empty)
    const uint32_t *bg;     // region background (null: 0)
    int rw, x0, y0, ncx;
    uint64_t keep;          // cell bits a record carries: ~0x1FF, plus MOVED unless FSG_PACK_NO_MOVED
    uint8_t to_file[16];    // file index of each material id
};

// A cell in record form: the material id replaced by its file index (< 256,
// bits 0-7) and MOVED kept or dropped (bit 8); temperature and colour as
// stored.  Record bytes are then plain bytes of this word (rbyte).
__device__ __forceinline__ uint64_t rec_form(const PackSrc &s, uint64_t v) {
    return (v & s.keep) | s.to_file[v & 15u];
}

// Record i of block A (tiles rows 0..127, then layer2) or B of chunk (cx, cy).
__device__ __forceinline__ uint64_t rec_a(const PackSrc &s, int cx, int cy, int i) {
    const int row = (i >> 7) & (CW - 1), col = i & (CW - 1);
    if (i < CW * CW) return rec_form(s, s.cells[(size_t)(s.y0 + cy * CW + row) * s.W + (s.x0 + cx * CW + col)]);
    return rec_form(s, s.l2 ? s.l2[(size_t)(cy * CW + row) * s.rw + cx * CW + col] : 0ull);
}
__device__ __forceinline__ uint64_t rec_b(const PackSrc &s, int cx, int cy, int i) {
    const int row = i >> 7, col = i & (CW - 1);
    return s.bg ? s.bg[(size_t)(cy * CW + row) * s.rw + cx * CW + col] : 0u;
}

// Byte b of the record of record-form value v: material index (u32, < 256),
// colour (u32), temperature (i16), pad byte 10 (MOVED), pad byte 11 (0); for
// block B the u32 itself.
template <int R>
__device__ __forceinline__ uint32_t rbyte(uint64_t v, int b) {
    if (R == RB) return (uint32_t)(v >> (8 * b)) & 0xFFu;
    switch (b) {
    case 0: return (uint32_t)v & 0xFFu;
    case 4: case 5: case 6: case 7: return (uint32_t)(v >> (32 + 8 * (b - 4))) & 0xFFu;
    case 8: return (uint32_t)(v >> 16) & 0xFFu;
    case 9: return (uint32_t)(v >> 24) & 0xFFu;
    case 10: return (uint32_t)(v >> 8) & 0xFFu;
    default: return 0u;
    }
}

// Bit b set where byte b of two records is equal (x = their xor).
template <int R>
__device__ __forceinline__ uint32_t eqmask(uint64_t x) {
    uint32_t m = 0;
#pragma unroll
    for (int b = 0; b < R; b++) m |= (rbyte<R>(x, b) == 0u ? 1u : 0u) << b;
    return m;
}

__device__ __forceinline__ int extlen(int v) { return v >= 15 ? (v - 15) / 255 + 1 : 0; }

// Records of every row staged at a time, per column block (LDS, coalesced
// loads: consecutive threads read consecutive cells of a row).  One thread
// per row reading its own records touched 64 cache lines per wave load:
// 17.1 ms for 16384 chunks, against 12.1 ms staged 4 columns at a time (8:
// 13.9 ms, 16: 18.8, 2: 15.7; the staging's LDS sets the occupancy).
constexpr int CB = 4;

// Parses the block's record rows, thread t taking row t (bytes
// [t*128*R, (t+1)*128*R) of an n-byte block) into runs and handing each match
// to the sink; returns the literal bytes pending at the row's end.  Every
// thread of the workgroup calls it (the rows' records are staged in LDS by
// all of them, CB columns at a time); threads t >= rows only stage.
template <int R, class Rec, class Sink>
__device__ int walk_row(int rows, int n, const Rec &rec, uint64_t (*stage)[CB], Sink &sink) {
    constexpr uint32_t FULL = (1u << R) - 1;
    const int t = threadIdx.x;
    const bool mine = t < rows;
    const int i0 = t * CW;
    const int forced = n - fsg_lz4::MFLIMIT;  // bytes from here on are literals
    uint64_t left = mine && i0 > 0 ? rec(i0 - 1) : 0ull;
    int kind = -1, len = 0, pend = 0;
    auto close = [&](int end) {
        if (kind != KX && len >= fsg_lz4::MINMATCH) {
            sink.match(pend, end - len, len, kind == KL ? R : CW * R);
            pend = 0;
        } else {
            pend += len;
        }
    };
    for (int jb = 0; jb < CW; jb += CB) {
        __syncthreads();  // the last column block's readers are done
        for (int k = t; k < rows * CB; k += PT) {
            const int r = k / CB, c = k - r * CB;
            stage[r][c] = rec(r * CW + jb + c);
        }
        __syncthreads();
        if (!mine) continue;
        sink.window(i0 + jb, stage[t], left);
        for (int jj = 0; jj < CB; jj++) {
            const int j = jb + jj, i = i0 + j;
            const uint64_t v = stage[t][jj];
            const uint32_t mL = i > 0 ? eqmask<R>(v ^ left) : 0u;
            const uint32_t mU = i >= CW ? eqmask<R>(v ^ stage[t - 1][jj]) : 0u;  // the row above: thread t - 1's
            left = v;
            const int p0 = i * R;
            uint32_t live = FULL;
            if (p0 + R > forced) live = p0 >= forced ? 0u : ((1u << (forced - p0)) - 1u);
            const uint32_t isL = mL & live, isU = mU & ~mL & live;
            const uint32_t same = kind == KL ? isL : kind == KU ? isU : kind == KX ? (~(isL | isU) & FULL) : 0u;
            if (same == FULL) {
                len += R;
                continue;
            }
            // runs of one kind inside the record, found with bit scans
            const uint32_t isX = ~(isL | isU) & FULL;
            for (int b = 0; b < R;) {
                const int k = (isL >> b & 1u) ? KL : (isU >> b & 1u) ? KU : KX;
                const uint32_t m = k == KL ? isL : k == KU ? isU : isX;
                const int run = min(R - b, __builtin_ctz(~(m >> b)));  // bytes of kind k from b on
                if (k == kind) {
                    len += run;
                } else {
                    if (kind >= 0) close(p0 + b);
                    kind = k;
                    len = run;
                }
                b += run;
            }
        }
    }
    if (mine) close((i0 + CW) * R);
    return pend;
}
// Pass 1: the sizes of the sequences whose match lies in the row.
struct Sizer {
    bool hm = false;
    int lead = 0, rest = 0;
    __device__ void window(int, const uint64_t *, uint64_t) {}
    __device__ void match(int pend, int, int ml, int) {
        if (!hm) {
            hm = true;
            lead = pend;
            rest += 3 + extlen(ml - 4);
        } else {
            rest += 1 + extlen(pend) + pend + 2 + extlen(ml - 4);
        }
    }
};

// Output of one block: every store is checked against the block's capacity
// (the size pass keeps the writes inside it; the check keeps a slot's
// neighbours safe whatever happens).
struct Out {
    uint8_t *p;
    int cap;
    __device__ void put(int at, uint32_t v) const {
        if (at >= 0 && at < cap) p[at] = (uint8_t)v;
    }
};

__device__ __forceinline__ int put_len(const Out &o, int oc, int v) {
    if (v >= 15) {
        v -= 15;
        while (v >= 255) {
            o.put(oc++, 255);
            v -= 255;
        }
        o.put(oc++, (uint32_t)v);
    }
    return oc;
}

// Copies raw block bytes [p, p + cnt) to output position at.
template <int R, class Rec>
__device__ void put_raw(const Out &o, int at, int p, int cnt, const Rec &rec) {
    if (cnt <= 0) return;
    int i = p / R, b = p - i * R;
    uint64_t v = rec(i);
    for (int k = 0; k < cnt; k++) {
        o.put(at + k, rbyte<R>(v, b));
        if (++b == R && k + 1 < cnt) {
            b = 0;
            v = rec(++i);
        }
    }
}

// Record i from the thread's staged column block (records base .. base +
// CB - 1, LDS) or the record just before it (a register) when it lies there,
// else from the block's source.
template <class Rec>
struct Near {
    const Rec *rec;
    int base = 0;
    const uint64_t *win = nullptr;
    uint64_t prev = 0;
    __device__ uint64_t operator()(int i) const {
        const int d = i - base;
        return win && d >= 0 && d < CB ? win[d] : win && d == -1 ? prev : (*rec)(i);
    }
};

// Pass 2: writes the row's sequences.
template <int R, class Rec>
struct Writer {
    Out o;
    int oc, carry;
    bool first = true;
    Near<Rec> near;
    __device__ void window(int base, const uint64_t *win, uint64_t prev) {
        near.base = base;
        near.win = win;
        near.prev = prev;
    }
    __device__ void match(int pend, int mstart, int ml, int off) {
        const int ll = first ? carry + pend : pend;
        o.put(oc++, (uint32_t)((min(ll, 15) << 4) | min(ml - 4, 15)));
        oc = put_len(o, oc, ll);
        put_raw<R>(o, oc + (ll - pend), mstart - pend, pend, near);  // carried bytes come from earlier rows
        oc += ll;
        o.put(oc++, (uint32_t)off & 0xFFu);
        o.put(oc++, (uint32_t)off >> 8);
        oc = put_len(o, oc, ml - 4);
        first = false;
    }
};

template <int R, class Rec>
__device__ void encode_block(int n, const Rec &rec, uint8_t *out, int cap, int32_t *size_out) {
    constexpr int S = CW * R;
    const int rows = n / S, L = rows - 1, t = threadIdx.x;
    __shared__ int hm[PT], lead[PT], trail[PT], size[PT], carry[PT], start[PT], nxt[PT];
    __shared__ int total;
    __shared__ uint64_t stage[PT][CB];
    Sizer sz;
    const int pend = walk_row<R>(rows, n, rec, stage, sz);
    if (t < rows) {
        hm[t] = sz.hm;
        lead[t] = sz.hm ? sz.lead : S;
        trail[t] = pend;
    }
    __syncthreads();
    if (t == 0) {
        int c = 0;
        for (int u = 0; u < rows; u++) {
            carry[u] = c;
            c = trail[u] + (hm[u] ? 0 : c);
        }
        int nx = L;
        for (int u = L; u >= 0; u--) {
            nxt[u] = nx;
            if (hm[u]) nx = u;
        }
    }
    __syncthreads();
    if (t < rows) {
        int s = 0;
        if (hm[t]) s = sz.rest + extlen(carry[t] + lead[t]) + carry[t] + lead[t];
        if (t == L) {
            const int llf = hm[t] ? pend : carry[t] + S;
            s += 1 + extlen(llf) + llf;
        }
        size[t] = s;
    }
    __syncthreads();
    if (t == 0) {
        int o = 0;
        for (int u = 0; u < rows; u++) {
            start[u] = o;
            o += size[u];
        }
        total = o;
    }
    __syncthreads();
    if (total > cap) {  // cannot happen for these parses (sizes stay under the LZ4 bound)
        if (t == 0) *size_out = -1;
        return;
    }
    const Out o{out, cap};
    Writer<R, Rec> w{o, t < rows ? start[t] : 0, t < rows ? carry[t] : 0, true, Near<Rec>{&rec}};
    walk_row<R>(rows, n, rec, stage, w);
    if (t < rows) {
        const int row0 = t * S, tstart = row0 + S - pend;  // first trailing literal byte
        if (t == L) {
            // the last sequence: literals only
            const int llf = hm[t] ? pend : carry[t] + S;
            int oc = w.oc;
            o.put(oc++, (uint32_t)(min(llf, 15) << 4));
            oc = put_len(o, oc, llf);
            put_raw<R>(o, oc + (llf - pend), tstart, pend, w.near);  // the last staged block, or the source
        } else {
            // trailing literals belong to the first sequence of the next row with a match
            // (or to the last sequence)
            const int u = nxt[t];
            const int ll0 = carry[u] + lead[u];
            const int lit_out = start[u] + 1 + extlen(ll0);
            const int lit_raw = u * S - carry[u];
            put_raw<R>(o, lit_out + (tstart - lit_raw), tstart, pend, w.near);
        }
    }
    if (t == 0) *size_out = total;
}

struct RecA {
    PackSrc s;
    int cx, cy;
    __device__ uint64_t operator()(int i) const { return rec_a(s, cx, cy, i); }
};
struct RecB {
    PackSrc s;
    int cx, cy;
    __device__ uint64_t operator()(int i) const { return rec_b(s, cx, cy, i); }
};

// grid: x = chunk, y = block (0: tiles + layer2, 1: background)
__global__ __launch_bounds__(PT) void k_pack_encode(PackSrc s, uint8_t *slots, int32_t *sizes) {
    const int c = blockIdx.x, cx = c % s.ncx, cy = c / s.ncx;
    uint8_t *slot = slots + (size_t)c * SLOT;
    if (blockIdx.y == 0) encode_block<RA>(NA, RecA{s, cx, cy}, slot, CAPA, &sizes[2 * c]);
    else encode_block<RB>(NB, RecB{s, cx, cy}, slot + CAPA, CAPB, &sizes[2 * c + 1]);
}

// Image offsets (exclusive scan of 17 + A + B over the chunks); off[n] = total,
// or -1 when a block failed.  One workgroup.
__global__ __launch_bounds__(1024) void k_pack_offsets(const int32_t *sizes, int n, int64_t *off) {
    __shared__ int64_t part[1024];
    __shared__ int bad;
    const int t = threadIdx.x, per = (n + 1023) / 1024, a = min(n, t * per), b = min(n, a + per);
    if (t == 0) bad = 0;
    __syncthreads();
    int64_t s = 0;
    for (int c = a; c < b; c++) {
        if (sizes[2 * c] < 0 || sizes[2 * c + 1] < 0) bad = 1;
        s += HDR + sizes[2 * c] + sizes[2 * c + 1];
    }
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t o = 0;
        for (int k = 0; k < 1024; k++) {
            const int64_t v = part[k];
            part[k] = o;
            o += v;
        }
        off[n] = bad ? -1 : o;
    }
    __syncthreads();
    int64_t o = part[t];
    for (int c = a; c < b; c++) {
        off[c] = o;
        o += HDR + sizes[2 * c] + sizes[2 * c + 1];
    }
}

__device__ __forceinline__ void put_i32(uint8_t *p, int32_t v) {
    for (int k = 0; k < 4; k++) p[k] = (uint8_t)((uint32_t)v >> (8 * k));
}

// Writes each chunk's image: header, then block A, then block B.
__global__ __launch_bounds__(256) void k_pack_assemble(const uint8_t *slots, const int32_t *sizes,
                                                       const int64_t *off, const int8_t *phase, uint8_t *img) {
    const int c = blockIdx.x, t = threadIdx.x;
    if (off[gridDim.x] < 0) return;
    const int a = sizes[2 * c], b = sizes[2 * c + 1];
    uint8_t *o = img + off[c];
    const uint8_t *sa = slots + (size_t)c * SLOT, *sb = sa + CAPA;
    if (t == 0) {
        o[0] = (uint8_t)phase[c];
        put_i32(o + 1, NA);
        put_i32(o + 5, a);
        put_i32(o + 9, NB);
        put_i32(o + 13, b);
    }
    for (int k = t; k < a; k += 256) o[HDR + k] = sa[k];
    for (int k = t; k < b; k += 256) o[HDR + a + k] = sb[k];
}

// Decoder input: per block, its bytes in[off, off + n) (n < 0: no block, the
// result is -1) and its output out[dst, dst + cap).
struct DecDesc {
    int64_t off, dst;
    int32_t n, cap;
};

// The match table of the decode kernel (LDS, filled by each workgroup before
// its waves decode; read only after that).
__shared__ uint8_t dec_modtab[fsg_lz4::MODTAB];

// The lane policy of the decode kernel: this thread's own lane.  A wave's
// window and ring are its own and a wave's LDS accesses execute in order, so
// ordering the lanes' accesses needs only wavefront-scope fences around a
// wave barrier (which keeps the compiler from moving accesses across it); a
// workgroup barrier would tie the workgroup's waves together and also wait
// for every outstanding output store.
struct WaveLane {
    int lane;
    template <class F>
    __device__ void each(F f) const {
        f(lane);
    }
    __device__ void barrier() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __device__ uint32_t uniform(uint32_t x) const { return __builtin_amdgcn_readfirstlane(x); }
    __device__ uint32_t lane_value(uint32_t v, int k, const uint8_t *) const {
        return (uint32_t)__builtin_amdgcn_readlane((int)v, k);
    }
    __device__ uint32_t own(uint32_t v, int, const uint8_t *) const { return v; }
    __device__ uint32_t mod(int i) const { return dec_modtab[i]; }
    // lane vectors are the lane's own register; called with every lane active
    using V = uint32_t;
    template <class F>
    __device__ V make(F f) const {
        return f(lane);
    }
    __device__ uint32_t get(V v, int) const { return v; }
    __device__ uint32_t lane_of(V v, int k) const { return (uint32_t)__builtin_amdgcn_readlane((int)v, k); }
    __device__ uint64_t ballot(V v) const { return __builtin_amdgcn_ballot_w64(v != 0u); }
    __device__ V shfl(V v, V i) const { return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(i << 2), (int)v); }
    __device__ uint32_t below(uint64_t m, int) const {
        return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    }
};

// One wave per block, four blocks per workgroup (a workgroup per wave would
// leave CUs short of waves: their workgroup slots run out first).  Wave
// v = blockIdx.x * 4 + wave takes block desc[v]; the unpack lists every
// chunk's tiles-and-layer2 block (384 KiB decoded) before the background
// blocks (64 KiB), so a workgroup holds blocks of one kind and its waves end
// together (a mixed one stays resident until its long blocks end), and the
// long blocks start first.  Each wave keeps its own block window and output
// ring in LDS and decodes alone (no workgroup barrier); result[v] is the
// decoded size, or -1 for a block lz4.c rejects.  A 2 KB ring (this
// encoder's offsets are at most one record row, 1536 bytes) keeps the LDS at
// 16.6 KB per workgroup, so 8 waves per SIMD fit (58 VGPRs): unpack 14.6 ->
// 13.9 ms on 16384 chunks against a 4 KB ring at 6 waves.
constexpr int DW = 4;  // waves (blocks) per decode workgroup

__global__ __launch_bounds__(64 * DW) __attribute__((amdgpu_waves_per_eu(8))) void k_pack_decode(
    const uint8_t *in, const DecDesc *desc, uint8_t *out, int32_t *result, int nblocks) {
    __shared__ __align__(16) uint8_t win[DW][fsg_lz4::IW];
    __shared__ uint8_t ring[DW][fsg_lz4::RING];
    __shared__ uint32_t gdesc[DW][fsg_lz4::GMAX];
    for (int i = threadIdx.x; i < fsg_lz4::MODTAB; i += 64 * DW) dec_modtab[i] = fsg_lz4::modtab_entry(i);
    __syncthreads();
    // the wave's index as a uniform, so that the block's descriptor and the
    // whole parse (positions, lengths, limits) live in scalar registers
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const int v = blockIdx.x * DW + wv;
    if (v >= nblocks) return;
    const DecDesc d = desc[v];
    int r = -1;
    if (d.n >= 0) {
        fsg_lz4::Staged<WaveLane> M{in + d.off, d.n, out + d.dst, win[wv], ring[wv], gdesc[wv], -(1 << 30),
                                    WaveLane{lane}, 0u, 0};
        r = fsg_lz4::decode_block(d.n, d.cap, M);
    }
    if (lane == 0) result[v] = r;
}

// Little-endian u32 at a 4-byte aligned staging address (records are 12 or
// 4 bytes and each chunk's staging starts at a multiple of 4).
__device__ __forceinline__ uint32_t ld32(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }
static_assert(RAW % 4 == 0 && NA % 4 == 0 && RA % 4 == 0 && RB % 4 == 0, "aligned record words");

struct UnpackDst {
    uint64_t *cells;  // world grid
    int W, x0, y0, ncx, rw;
    uint64_t *l2;     // region layer2 out (null: not kept)
    uint32_t *bg;     // region background out (null: not kept)
    uint32_t moved;   // MOVED taken from pad byte 10 (else clear)
    int n;            // chunks: status[c] and status[n + c] are chunk c's decoded block sizes
    int16_t from_file[256];  // material id of each file index, or -1
};

// One workgroup per chunk: records -> cells, after checking every material
// index against the map (a bad one fails the chunk and leaves its cells as
// they were).
__global__ __launch_bounds__(256) void k_pack_cells(const uint8_t *raw, const int32_t *status, UnpackDst d,
                                                    int32_t *chunk_status) {
    __shared__ int bad;
    const int c = blockIdx.x, t = threadIdx.x, cx = c % d.ncx, cy = c / d.ncx;
    if (status[c] != NA || status[d.n + c] != NB) {
        if (t == 0) chunk_status[c] = FSG_ERR_INVALID;
        return;
    }
    if (t == 0) bad = 0;
    __syncthreads();
    const uint8_t *r = raw + (size_t)c * RAW;
    int mybad = 0;
    for (int i = t; i < 2 * CW * CW; i += 256) {
        const uint32_t idx = ld32(r + (size_t)i * RA);
        mybad |= idx > 0xFFu || d.from_file[idx & 0xFFu] < 0;
    }
    if (mybad) bad = 1;
    __syncthreads();
    if (bad) {
        if (t == 0) chunk_status[c] = FSG_ERR_INVALID;
        return;
    }
    const int nrec = d.l2 ? 2 * CW * CW : CW * CW;  // layer2 records are read only when kept
    for (int i = t; i < nrec; i += 256) {
        const uint8_t *q = r + (size_t)i * RA;
        const uint32_t w2 = ld32(q + 8);  // temperature, pad byte 10, pad byte 11
        const uint64_t v = (uint64_t)(uint16_t)d.from_file[ld32(q)] | (d.moved ? (w2 >> 8) & 0x100u : 0u) |
                           ((uint64_t)(w2 & 0xFFFFu) << 16) | ((uint64_t)ld32(q + 4) << 32);
        const int row = (i >> 7) & (CW - 1), col = i & (CW - 1);
        if (i < CW * CW) d.cells[(size_t)(d.y0 + cy * CW + row) * d.W + (d.x0 + cx * CW + col)] = v;
        else if (d.l2) d.l2[(size_t)(cy * CW + row) * d.rw + cx * CW + col] = v;
    }
    if (d.bg)
        for (int i = t; i < CW * CW; i += 256)
            d.bg[(size_t)(cy * CW + (i >> 7)) * d.rw + cx * CW + (i & (CW - 1))] = ld32(r + NA + (size_t)i * RB);
    if (t == 0) chunk_status[c] = FSG_OK;
}

using fsg_impl::dev_alloc;
using fsg_impl::dev_free;

// Grows a world-owned device buffer to `bytes` (contents not kept).
int grow(fsg_ctx *c, void **p, size_t *cap, size_t bytes) {
    if (*cap >= bytes) return FSG_OK;
    dev_free(*p);
    *p = nullptr;
    *cap = 0;
    if (dev_alloc(c, p, bytes)) return FSG_ERR_NOMEM;
    *cap = bytes;
    return FSG_OK;
}

int region_ok(const fsg_world *w, int x0, int y0, int ncx, int ncy) {
    if (x0 < 0 || y0 < 0 || ncx <= 0 || ncy <= 0 || ncx > w->d.W / CW || ncy > w->d.H / CW) return 0;
    return (int64_t)x0 + (int64_t)ncx * CW <= w->d.W && (int64_t)y0 + (int64_t)ncy * CW <= w->d.H;
}

int32_t rd_i32(const uint8_t *p) {
    int32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

void default_map(const fsg_config &cfg, fsg_pack_map *m) {
    std::memset(m, 0, sizeof(*m));
    for (int i = 0; i < 16; i++) m->to_file[i] = (uint32_t)i;
    for (int i = 0; i < 256; i++) m->from_file[i] = (int16_t)((uint32_t)i < cfg.n_materials ? i : -1);
}

// The map the world's pack calls use.
const fsg_pack_map &world_map(fsg_world *w) {
    if (!w->pack_map_set) default_map(w->cfg, &w->pack_map);
    return w->pack_map;
}

}  // namespace

extern "C" int fsg_pack_map_default(const fsg_config *cfg, fsg_pack_map *map) {
    if (!cfg || !map || cfg->n_materials > FSG_MAX_MATERIALS) return FSG_ERR_INVALID;
    default_map(*cfg, map);
    return FSG_OK;
}

extern "C" int fsg_world_set_pack_map(fsg_world *w, const fsg_pack_map *map) {
    if (!w) return FSG_ERR_INVALID;
    if (!map) {
        w->pack_map_set = false;
        return FSG_OK;
    }
    for (uint32_t m = 0; m < w->cfg.n_materials; m++)
        if (map->to_file[m] > 0xFFu) return FSG_ERR_INVALID;
    for (int i = 0; i < 256; i++)
        if (map->from_file[i] < -1 || map->from_file[i] >= (int)w->cfg.n_materials) return FSG_ERR_INVALID;
    if (map->flags & ~FSG_PACK_NO_MOVED) return FSG_ERR_INVALID;
    w->pack_map = *map;
    w->pack_map_set = true;
    return FSG_OK;
}

extern "C" int fsg_world_pack(fsg_world *w, int32_t x0, int32_t y0, int32_t ncx, int32_t ncy, const int8_t *phase,
                              const fsg_cell *layer2, const uint32_t *background, size_t *image_bytes) {
    if (!w || !image_bytes || !region_ok(w, x0, y0, ncx, ncy)) return FSG_ERR_INVALID;
    const size_t n = (size_t)ncx * ncy, rcells = n * CW * CW;
    if (layer2)  // layer2 records are cells of the world's materials too
        for (size_t i = 0; i < rcells; i++)
            if ((layer2[i].mat & 0xFE00u) || (uint32_t)(layer2[i].mat & 0xFFu) >= w->cfg.n_materials)
                return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    // per-chunk device state: encoder slots, images (worst case), sizes, offsets, phases;
    // the region's layer2 / background when given
    const size_t b_small = ((n * 8 + (n + 1) * 8 + n + 255) & ~(size_t)255);
    if (grow(c, &w->d_pack_slots, &w->pack_slots_cap, n * SLOT) ||
        grow(c, &w->d_pack_img, &w->pack_img_cap, n * (HDR + CAPA + CAPB)) ||
        grow(c, &w->d_pack_small, &w->pack_small_cap, b_small))
        return FSG_ERR_NOMEM;
    const size_t b_l2 = layer2 ? rcells * 8 : 0, b_bg = background ? rcells * 4 : 0;
    if (b_l2 + b_bg && grow(c, &w->d_pack_aux, &w->pack_aux_cap, b_l2 + b_bg)) return FSG_ERR_NOMEM;
    int32_t *d_sizes = static_cast<int32_t *>(w->d_pack_small);
    int64_t *d_off = reinterpret_cast<int64_t *>(d_sizes + 2 * n);
    int8_t *d_phase = reinterpret_cast<int8_t *>(d_off + n + 1);
    w->pack_phase.assign(n, 0);
    if (phase) std::copy(phase, phase + n, w->pack_phase.begin());
    HIPCHK(hipMemcpyAsync(d_phase, w->pack_phase.data(), n, hipMemcpyHostToDevice, c->stream));
    PackSrc s{};
    s.cells = w->d.buf[w->d.tick & 1];
    s.W = w->d.W;
    s.rw = ncx * CW;
    s.x0 = x0;
    s.y0 = y0;
    s.ncx = ncx;
    const fsg_pack_map &map = world_map(w);
    s.keep = (map.flags & FSG_PACK_NO_MOVED) ? ~0x1FFull : ~0xFFull;
    for (int m = 0; m < 16; m++) s.to_file[m] = (uint8_t)map.to_file[m];
    char *aux = static_cast<char *>(w->d_pack_aux);
    if (layer2) {
        HIPCHK(hipMemcpyAsync(aux, layer2, b_l2, hipMemcpyHostToDevice, c->stream));
        s.l2 = reinterpret_cast<const uint64_t *>(aux);
    }
    if (background) {
        HIPCHK(hipMemcpyAsync(aux + b_l2, background, b_bg, hipMemcpyHostToDevice, c->stream));
        s.bg = reinterpret_cast<const uint32_t *>(aux + b_l2);
    }
    uint8_t *slots = static_cast<uint8_t *>(w->d_pack_slots), *img = static_cast<uint8_t *>(w->d_pack_img);
    k_pack_encode<<<dim3((unsigned)n, 2), PT, 0, c->stream>>>(s, slots, d_sizes);
    k_pack_offsets<<<1, 1024, 0, c->stream>>>(d_sizes, (int)n, d_off);
    k_pack_assemble<<<(unsigned)n, 256, 0, c->stream>>>(slots, d_sizes, d_off, d_phase, img);
    HIPCHK(hipGetLastError());
    w->pack_sizes.resize(2 * n);
    int64_t total = 0;
    HIPCHK(hipMemcpyAsync(w->pack_sizes.data(), d_sizes, 2 * n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&total, d_off + n, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (total < 0) {
        w->pack_n = 0;
        return FSG_ERR_DEVICE;
    }
    for (size_t k = 0; k < n; k++) image_bytes[k] = HDR + (size_t)w->pack_sizes[2 * k] + w->pack_sizes[2 * k + 1];
    w->pack_n = n;
    w->pack_total = (size_t)total;
    return FSG_OK;
}

extern "C" int fsg_world_pack_download(fsg_world *w, void *out, size_t cap) {
    if (!w || !out) return FSG_ERR_INVALID;
    if (w->pack_n == 0) return FSG_ERR_STATE;
    if (cap < w->pack_total) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(out, w->d_pack_img, w->pack_total, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FSG_OK;
}

extern "C" int fsg_world_unpack(fsg_world *w, int32_t x0, int32_t y0, int32_t ncx, int32_t ncy, const void *images,
                                const size_t *image_bytes, int8_t *phase, fsg_cell *layer2, uint32_t *background,
                                int32_t *status) {
    if (!w || !status || !region_ok(w, x0, y0, ncx, ncy)) return FSG_ERR_INVALID;
    const size_t n = (size_t)ncx * ncy, rcells = n * CW * CW;
    if (images && !image_bytes) return FSG_ERR_INVALID;
    if (!images && w->pack_n != n) return FSG_ERR_STATE;  // the last pack's images, same chunk count
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    // descriptors of the 2n blocks from the headers (chunk.cpp:83-145)
    // chunk k's blocks: desc[k] (tiles and layer2, NA bytes) and desc[n + k] (background, NB)
    std::vector<DecDesc> desc(2 * n);
    std::vector<int8_t> ph(n, 0);
    const uint8_t *src = nullptr;
    size_t in_total = 0;
    if (images) {
        const uint8_t *p = static_cast<const uint8_t *>(images);
        for (size_t k = 0; k < n; k++) {
            const size_t len = image_bytes[k];
            DecDesc a{(int64_t)in_total + HDR, (int64_t)(k * RAW), -1, NA}, b{0, (int64_t)(k * RAW + NA), -1, NB};
            if (len >= (size_t)HDR) {
                ph[k] = (int8_t)p[in_total];
                const int32_t sa = rd_i32(p + in_total + 1), ca = rd_i32(p + in_total + 5);
                const int32_t sb = rd_i32(p + in_total + 9), cb = rd_i32(p + in_total + 13);
                if (sa == NA && sb == NB && ca >= 0 && cb >= 0 && (size_t)HDR + (size_t)ca + (size_t)cb <= len) {
                    a.n = ca;
                    b.off = a.off + ca;
                    b.n = cb;
                }
            }
            desc[k] = a;
            desc[n + k] = b;
            in_total += len;
        }
        if (grow(c, &w->d_pack_in, &w->pack_in_cap, std::max<size_t>(in_total, 1))) return FSG_ERR_NOMEM;
        HIPCHK(hipMemcpyAsync(w->d_pack_in, images, in_total, hipMemcpyHostToDevice, c->stream));
        src = static_cast<const uint8_t *>(w->d_pack_in);
    } else {
        int64_t o = 0;
        for (size_t k = 0; k < n; k++) {
            const int32_t ca = w->pack_sizes[2 * k], cb = w->pack_sizes[2 * k + 1];
            desc[k] = DecDesc{o + HDR, (int64_t)(k * RAW), ca, NA};
            desc[n + k] = DecDesc{o + HDR + ca, (int64_t)(k * RAW + NA), cb, NB};
            ph[k] = w->pack_phase[k];
            o += HDR + ca + cb;
        }
        src = static_cast<const uint8_t *>(w->d_pack_img);
    }
    const size_t b_desc = (2 * n * sizeof(DecDesc) + 255) & ~(size_t)255, b_st = (3 * n * 4 + 255) & ~(size_t)255;
    const size_t b_l2 = layer2 ? rcells * 8 : 0, b_bg = background ? rcells * 4 : 0;
    if (grow(c, &w->d_pack_raw, &w->pack_raw_cap, n * RAW) ||
        grow(c, &w->d_pack_dec, &w->pack_dec_cap, b_desc + b_st + b_l2 + b_bg))
        return FSG_ERR_NOMEM;
    char *dec = static_cast<char *>(w->d_pack_dec);
    DecDesc *d_desc = reinterpret_cast<DecDesc *>(dec);
    int32_t *d_bst = reinterpret_cast<int32_t *>(dec + b_desc), *d_cst = d_bst + 2 * n;
    UnpackDst d{};
    d.cells = w->d.buf[w->d.tick & 1];
    d.W = w->d.W;
    d.x0 = x0;
    d.y0 = y0;
    d.ncx = ncx;
    d.rw = ncx * CW;
    d.l2 = layer2 ? reinterpret_cast<uint64_t *>(dec + b_desc + b_st) : nullptr;
    d.bg = background ? reinterpret_cast<uint32_t *>(dec + b_desc + b_st + b_l2) : nullptr;
    const fsg_pack_map &map = world_map(w);
    d.moved = (map.flags & FSG_PACK_NO_MOVED) ? 0u : 1u;
    d.n = (int)n;
    std::memcpy(d.from_file, map.from_file, sizeof(d.from_file));
    HIPCHK(hipMemcpyAsync(d_desc, desc.data(), 2 * n * sizeof(DecDesc), hipMemcpyHostToDevice, c->stream));
    if (b_l2 + b_bg) HIPCHK(hipMemsetAsync(dec + b_desc + b_st, 0, b_l2 + b_bg, c->stream));  // failed chunks: 0
    uint8_t *raw = static_cast<uint8_t *>(w->d_pack_raw);
    k_pack_decode<<<(unsigned)((2 * n + DW - 1) / DW), 64 * DW, 0, c->stream>>>(src, d_desc, raw, d_bst, (int)(2 * n));
    k_pack_cells<<<(unsigned)n, 256, 0, c->stream>>>(raw, d_bst, d, d_cst);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(status, d_cst, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (layer2) HIPCHK(hipMemcpyAsync(layer2, d.l2, b_l2, hipMemcpyDeviceToHost, c->stream));
    if (background) HIPCHK(hipMemcpyAsync(background, d.bg, b_bg, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (phase) std::copy(ph.begin(), ph.end(), phase);
    // new cells, as after fsg_world_upload
    w->d.force_tick = w->d.tick;
    w->stamped = false;
    w->prev_valid = false;
    return FSG_OK;
}

namespace {
struct DeviceBuffer {  // freed on every return path
    void *p = nullptr;
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

extern "C" int fsg_lz4_decode_blocks(fsg_ctx *c, const void *src, size_t src_len, const int64_t *src_off,
                                     const int32_t *src_bytes, const int32_t *cap, int32_t n, void *out,
                                     size_t out_len, const int64_t *out_off, int32_t *result) {
    if (!c || n < 0) return FSG_ERR_INVALID;
    if (n == 0) return FSG_OK;
    if (!src || !src_off || !src_bytes || !cap || !out || !out_off || !result) return FSG_ERR_INVALID;
    std::vector<DecDesc> desc((size_t)n);
    for (int32_t k = 0; k < n; k++) {
        if (src_off[k] < 0 || src_bytes[k] < 0 || (uint64_t)src_off[k] + (uint64_t)src_bytes[k] > src_len ||
            cap[k] < 0 || out_off[k] < 0 || (uint64_t)out_off[k] + (uint64_t)cap[k] > out_len)
            return FSG_ERR_INVALID;
        desc[k] = DecDesc{src_off[k], out_off[k], src_bytes[k], cap[k]};
    }
    HIPCHK(hipSetDevice(c->device));
    DeviceBuffer d_src, d_out, d_desc, d_res;
    HIPCHK(hipMalloc(&d_src.p, std::max<size_t>(src_len, 1)));
    HIPCHK(hipMalloc(&d_out.p, std::max<size_t>(out_len, 1)));
    HIPCHK(hipMalloc(&d_desc.p, (size_t)n * sizeof(DecDesc)));
    HIPCHK(hipMalloc(&d_res.p, (size_t)n * 4));
    if (src_len) HIPCHK(hipMemcpyAsync(d_src.p, src, src_len, hipMemcpyHostToDevice, c->stream));
    if (out_len) HIPCHK(hipMemcpyAsync(d_out.p, out, out_len, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_desc.p, desc.data(), (size_t)n * sizeof(DecDesc), hipMemcpyHostToDevice, c->stream));
    k_pack_decode<<<(unsigned)((n + DW - 1) / DW), 64 * DW, 0, c->stream>>>(
        static_cast<const uint8_t *>(d_src.p), static_cast<const DecDesc *>(d_desc.p),
        static_cast<uint8_t *>(d_out.p), static_cast<int32_t *>(d_res.p), n);
    HIPCHK(hipGetLastError());
    if (out_len) HIPCHK(hipMemcpyAsync(out, d_out.p, out_len, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(result, d_res.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FSG_OK;
}

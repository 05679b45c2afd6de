// fsg_lz4.hpp -- LZ4 block decoding by one wave.
//
// The chunk .pack images (chunk.cpp:74-217) hold LZ4 blocks that the reader
// opens with LZ4_decompress_safe (chunk.cpp:151, 191; lz4.c:1949-2360).  A
// block is a list of sequences: a token (literal count in the high nibble,
// match length - 4 in the low one, 15 meaning "more bytes follow, 255 each
// until one is smaller"), the literal bytes, a 2-byte little-endian offset and
// the match, copied byte by byte from `offset` bytes back; the last sequence
// has literals only.  The acceptance rules are lz4.c's for a full block:
//   * a literal run that ends within MFLIMIT (12) bytes of the output end, or
//     whose input ends within 2 + 1 + LASTLITERALS (8) bytes of the input
//     end, must be the last sequence and end the input exactly;
//   * a match may not start before the output start and must end at least
//     LASTLITERALS (5) bytes before the output end; an offset of 0 (invalid
//     in the format) gives zeros, as lz4.c, which zeroes the first word
//     before copying such a match onto itself (lz4.c:2328-2334);
//   * length bytes are read only while the input limits of lz4.c's
//     read_variable_length hold.
// Every lane of a wave runs the same parse on wave-uniform values; the
// literal and match copies are split over the lanes.  The parse reaches
// memory only through a Mem object (a lookahead of the block's next bytes,
// single byte reads, literal and match copies), so the same parse runs on
// the host for tests and in the decode kernel, whose Mem (Staged) keeps a
// window of the block and the recent output in LDS and the lookahead in the
// lanes' registers: one LDS read per token, its bytes read across lanes.
// Runs of short sequences (a few literals and a short match, one per record
// of a noisy chunk) are decoded a group at a time (Staged::group): the lanes
// parse every candidate start of the lookahead at once and the group's
// output bytes are produced together, one per lane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsg_lz4 {

constexpr int MINMATCH = 4, LASTLITERALS = 5, MFLIMIT = 12, RUN_MASK = 15;

__host__ __device__ inline int popc64(uint64_t x) { return __builtin_popcountll(x); }
__host__ __device__ inline int clz64(uint64_t x) { return __builtin_clzll(x); }  // x != 0

// Lane `lane` of `nl` copies bytes lane, lane + nl, ... of a literal run.
__host__ __device__ inline void copy_lane(uint8_t *d, const uint8_t *s, int len, int lane, int nl) {
    for (int j = lane; j < len; j += nl) d[j] = s[j];
}

// Lane `lane` of `nl` writes bytes lane, lane + nl, ... of a match that
// starts at d: d[j] = d[j - off] in order, i.e. d[j] = d[(j mod off) - off]
// (zeros for off = 0).
__host__ __device__ inline void match_lane(uint8_t *d, int off, int len, int lane, int nl) {
    if (off == 0) {
        for (int j = lane; j < len; j += nl) d[j] = 0;
        return;
    }
    if (off >= len) {
        for (int j = lane; j < len; j += nl) d[j] = d[j - off];
        return;
    }
    int r = lane % off;
    const int step = nl % off;
    for (int j = lane; j < len; j += nl) {
        d[j] = d[r - off];
        r += step;
        if (r >= off) r -= off;
    }
}

// Block bytes and recent output staged in two small buffers (LDS in the
// decode kernel).  The window holds IW block bytes starting at a 16-byte
// aligned address; the ring holds output byte q at ring[q % RING], i.e. the
// last RING bytes written.  Lanes is the lane policy: each(f) runs f for the
// lanes this thread stands for (its own lane on the device, all 64 in turn
// on the host), barrier() orders the lanes' buffer accesses, uniform(x)
// marks a value that is the same in every lane (it then lives in a scalar
// register on the device), lane_value(v, k, row) is lane k's v in every lane
// and own(v, lane, row) this lane's v (row: the bytes the lanes hold, which
// the host reads instead of the lanes' registers); mod(i) is match table
// entry i (LDS on the device, filled before the decode).  Lane vectors
// (type V: one 32-bit value per lane, a register on the device, an array on
// the host) serve the group decode: make(f) is f(lane) in every lane,
// get(v, lane) a lane's value inside such an f, lane_of(v, k) lane k's value
// as a uniform, ballot(v) the mask of lanes whose v is nonzero, shfl(v, i)
// lane i's v in each lane, below(m, lane) the number of bits of m below the
// lane's.
constexpr int IW = 1024, RING = 2048, NL = 64;
constexpr int GMAX = NL / MINMATCH;  // sequences in one group at most

// The match table: entry off * NL + j is j mod off for 1 <= off < NL and
// j < NL (MODTAB bytes; entry i is modtab_entry(i)).  A short-offset match
// reads it instead of dividing.
constexpr int MODTAB = NL * NL;
__host__ __device__ inline uint8_t modtab_entry(int i) {
    const int off = i / NL, j = i % NL;
    return off ? (uint8_t)(j % off) : (uint8_t)0;
}

template <class Lanes>
struct Staged {
    const uint8_t *src;  // the block (n bytes)
    int n;
    uint8_t *dst;        // the output
    uint8_t *win;        // IW bytes, 16-byte aligned
    uint8_t *ring;       // RING bytes
    uint32_t *gdesc;     // GMAX words: the sequences of a group
    int wbase;           // the window holds block bytes [wbase, wbase + IW)
    Lanes L;

    uint32_t ahead;      // this lane's byte of the lookahead: block byte at0 + lane
    int at0;

    __host__ __device__ uint32_t byte(int i) {
        if (i < wbase || i >= wbase + IW) refill(i);
        return L.uniform(win[i - wbase]);
    }
    // The NL block bytes from i on, one per lane (0 past the block's end).
    __host__ __device__ void look(int i) {
        if (i < wbase || i + NL > wbase + IW) refill(i);
        at0 = i;
        L.each([&](int lane) { ahead = win[i - wbase + lane]; });
    }
    // Byte at0 + k (k < NL) of the lookahead, in every lane.
    __host__ __device__ uint32_t at(int k) { return L.lane_value(ahead, k, win + (at0 - wbase)); }
    // Reloads the window to start at or just below block byte i: lane k
    // loads 16 bytes (NL * 16 = IW); bytes outside the block read as 0.
    __host__ __device__ void refill(int i) {
        L.barrier();
        wbase = i - (int)((uintptr_t)(src + i) & 15);
        L.each([&](int lane) {
            const int a = wbase + 16 * lane;
            if (a >= 0 && a + 16 <= n) {
                *reinterpret_cast<uint4 *>(win + 16 * lane) = *reinterpret_cast<const uint4 *>(src + a);
            } else {
                for (int k = 0; k < 16; k++) win[16 * lane + k] = (a + k >= 0 && a + k < n) ? src[a + k] : 0;
            }
        });
        L.barrier();
    }
    __host__ __device__ void put(int q, uint8_t b) {
        dst[q] = b;
        ring[q & (RING - 1)] = b;
    }
    // Literals: from the window when the run lies in it, else from the block.
    // Output byte q goes to ring slot q % RING; the slots of one lane are
    // q = lane (mod NL), so a long run's later bytes win in each slot.
    // k: the literals' index in the lookahead; a run inside it is stored
    // straight from the lanes that hold it.
    __host__ __device__ void lit(int op, int ip, int len, int k) {
        if (k + len <= NL) {
            L.each([&](int lane) {
                if (lane >= k && lane < k + len) put(op + lane - k, (uint8_t)L.own(ahead, lane, win + (at0 - wbase)));
            });
            return;
        }
        const bool inwin = ip >= wbase && ip + len <= wbase + IW;
        L.each([&](int lane) {
            for (int j = lane; j < len; j += NL) put(op + j, inwin ? win[ip + j - wbase] : src[ip + j]);
        });
    }
    // Match of len bytes at distance off.  Offsets up to RING - NL read the
    // ring: byte j < P comes from [op - off, op) (j mod off), byte j >= P from
    // op + j - P, P the smallest multiple of off that is >= NL, so every
    // source byte was written before the current group of NL bytes and no
    // group overwrites a slot it still reads.  Larger offsets read the output
    // itself (off > NL: the source bytes are from earlier groups).
    __host__ __device__ void match(int op, int off, int len) {
        L.barrier();
        if (off == 0) {
            L.each([&](int lane) {
                for (int j = lane; j < len; j += NL) put(op + j, 0);
            });
            return;
        }
        if (off >= NL) {  // byte j comes from op + j - off, written before j's group
            const bool near = off <= RING - NL;
            for (int g = 0; g < len; g += NL) {
                L.each([&](int lane) {
                    const int j = g + lane;
                    if (j < len) put(op + j, near ? ring[(op + j - off) & (RING - 1)] : dst[op + j - off]);
                });
                if (g + NL < len) L.barrier();
            }
            return;
        }
        // P, the smallest multiple of off that is >= NL (only matches longer
        // than NL need it): NL - 1 = q * off + r gives P = NL - 1 - r + off.
        // The first group's bytes j = lane < P come from op - off + (j mod off)
        // (the table); a later byte j from op + j - P, which for j < P is the
        // same as op - off + (j mod off), since then j mod off = j - (P - off).
        const int P = len > NL ? NL - 1 - (int)L.uniform(L.mod(off * NL + NL - 1)) + off : 0;
        for (int g = 0; g < len; g += NL) {
            L.each([&](int lane) {
                const int j = g + lane;
                if (j < len) put(op + j, ring[(g == 0 ? op - off + (int)L.mod(off * NL + lane) : op + j - P) & (RING - 1)]);
            });
            if (g + NL < len) L.barrier();
        }
    }
    // A group of short sequences from the lookahead (block byte ip = at0 on):
    // sequences with no length bytes, all of whose bytes lie in the
    // lookahead, with a match offset in [1, min(op, RING - NL)], while their
    // output fits NL bytes.  Lane s parses the sequence that would start at
    // lookahead byte s; the real chain from byte 0 is then followed in
    // uniform registers (H marks each sequence's first output byte, S its
    // token byte).  Output lane j finds its sequence by counting the bits of
    // H up to j, and its byte: a literal from the window, a match byte
    // written before the group from the ring, or one of the group's own
    // bytes, a lane p < j (the match source, op - off + (m mod off) for the
    // m-th match byte, lies before the sequence's match).  Pointer jumping
    // over p resolves the chains of such lanes.  The caller guarantees
    // ip + NL + 6 <= n and op + NL + MFLIMIT <= cap, so no sequence here is
    // the last one (its literals end at most ip + NL - 2 <= n - 8 and its
    // output at most op + NL <= cap - MFLIMIT) and every match ends before
    // cap - LASTLITERALS.  Returns false, consuming nothing, when the first
    // sequence is not a short one.
    __host__ __device__ bool group(int &ip, int &op) {
        const int base = at0 - wbase, op0 = op;
        // lane s: the sequence's literals << 5 | s << 9 | offset << 16 (desc)
        // and its output length (127 when not short) | next token byte << 7 (link)
        const auto desc = L.make([&](int s) -> uint32_t {
            const int ll = (int)(L.own(ahead, s, win + base) >> 4), k = s + 1 + ll < NL - 2 ? s + 1 + ll : NL - 2;
            const int off = (int)win[base + k] | (int)win[base + k + 1] << 8;
            return (uint32_t)ll << 5 | (uint32_t)s << 9 | (uint32_t)off << 16;
        });
        const auto link = L.make([&](int s) -> uint32_t {
            const uint32_t t = L.own(ahead, s, win + base);
            const int ll = (int)(t >> 4), mc = (int)(t & 15u);
            const int off = (int)(L.get(desc, s) >> 16);
            const bool ok = ll != RUN_MASK && mc != RUN_MASK && s + 3 + ll <= NL && off != 0 && off <= op0 &&
                            off <= RING - NL;
            return (ok ? (uint32_t)(ll + mc + MINMATCH) : 127u) | (uint32_t)(s + 3 + ll) << 7;
        });
        uint64_t H = 0, S = 0;
        int at = 0, tot = 0;
        uint32_t lk = L.lane_of(link, 0);
        for (int t2 = (int)(lk & 127u); t2 <= NL; t2 = tot + (int)(lk & 127u)) {
            H |= 1ull << tot;
            S |= 1ull << at;
            tot = t2;
            at = (int)(lk >> 7);
            lk = L.lane_of(link, at & (NL - 1));
            lk = at < NL ? lk : 127u;  // the next token is not in the lookahead
        }
        if (tot == 0) return false;
        L.barrier();
        L.each([&](int lane) {
            if ((S >> lane) & 1u) gdesc[L.below(S, lane)] = L.get(desc, lane);
        });
        L.barrier();
        // output lane j: 256 | byte when known, else (group lane of its source) << 9
        auto w = L.make([&](int j) -> uint32_t {
            const uint64_t upto = H & (~0ull >> (63 - j));  // bit 0 is set
            const uint32_t d = gdesc[popc64(upto) - 1];
            const int o = 63 - clz64(upto), ll = (int)((d >> 5) & 15u), s = (int)((d >> 9) & 63u);
            const int off = (int)(d >> 16), rel = j - o, m = rel - ll;
            // m mod off for the match bytes (m < 15 + MINMATCH < 63): the table row of min(off, 63)
            const int mm = (int)L.mod((off < NL - 1 ? off : NL - 1) * NL + (m & (NL - 1)));
            const int src = op0 + o + ll - off + mm;
            const bool lit = rel < ll;
            // one read for every lane (in bounds for lanes past the group too)
            const uint8_t *p = lit ? win + (base + s + 1 + rel) : ring + (src & (RING - 1));
            const uint32_t v = *p;
            return j >= tot ? 256u : !lit && src >= op0 ? (uint32_t)(src - op0) << 9 : 256u | v;
        });
        while (L.ballot(L.make([&](int j) { return (L.get(w, j) & 256u) ^ 256u; }))) {
            const auto got = L.shfl(w, L.make([&](int j) { return L.get(w, j) >> 9; }));
            w = L.make([&](int j) { return (L.get(w, j) & 256u) ? L.get(w, j) : L.get(got, j); });
        }
        L.barrier();
        L.each([&](int j) {
            if (j < tot) put(op0 + j, (uint8_t)L.get(w, j));
        });
        op = op0 + tot;
        ip = at0 + at;
        return true;
    }
};

// Decodes the LZ4 block of n bytes into an output of cap bytes.  Each token
// starts with M.look(ip), which makes the next NL block bytes readable as
// M.at(k); length bytes and offsets beyond the lookahead come from M.byte(i).
// Literals go through M.lit, matches through M.match.  Away from the block's
// ends (ip + NL + 6 <= n, op + NL + MFLIMIT <= cap), M.group(ip, op) may
// decode a run of short sequences in one step and advance ip and op past
// them; the rules below then hold for each of its sequences by construction.
// Returns the decoded size, or -1 for a block that lz4.c's rules reject.
// Reads only block bytes [0, n) and writes only output bytes [0, cap).
template <class Mem>
__host__ __device__ inline int decode_block(int n, int cap, Mem &M) {
    if (n <= 0 || cap < 0) return -1;
    if (cap == 0) return (n == 1 && M.byte(0) == 0) ? 0 : -1;
    int ip = 0, op = 0;
    for (;;) {
        if (ip >= n) return -1;
        M.look(ip);
        if (ip + NL + 6 <= n && op + NL + MFLIMIT <= cap && M.group(ip, op)) continue;
        auto next = [&](int &k) -> int {  // the byte at ip + k, then k + 1
            const int v = (int)(k < NL ? M.at(k) : M.byte(ip + k));
            k++;
            return v;
        };
        int k = 0;
        const int token = next(k);
        int ll = token >> 4;
        if (ll == RUN_MASK) {  // read_variable_length(ip, iend - RUN_MASK, initial check)
            const int lim = n - RUN_MASK;
            if (ip + k >= lim) return -1;
            int s;
            do {
                s = next(k);
                ll += s;
                if (ip + k > lim || ll > cap) return -1;
            } while (s == 255);
        }
        const int lp = ip + k;  // first literal
        const int cpy = op + ll;
        if (cpy > cap - MFLIMIT || lp + ll > n - (2 + 1 + LASTLITERALS)) {
            // the last sequence: it must end the input exactly and fit
            if (lp + ll != n || cpy > cap) return -1;
            M.lit(op, lp, ll, k);
            return cpy;
        }
        M.lit(op, lp, ll, k);
        op = cpy;
        k += ll;  // lp + ll + 2 <= n - 6 here
        const int lo = next(k);
        const int off = lo | (next(k) << 8);
        int ml = token & 15;
        if (ml == 15) {  // read_variable_length(ip, iend - LASTLITERALS + 1, no initial check)
            const int lim = n - LASTLITERALS + 1;
            int s;
            do {
                s = next(k);
                ml += s;
                if (ip + k > lim || ml > cap) return -1;
            } while (s == 255);
        }
        ml += MINMATCH;
        if (off > op) return -1;
        if (op + ml > cap - LASTLITERALS) return -1;
        M.match(op, off, ml);
        op += ml;
        ip += k;
    }
}

}  // namespace fsg_lz4

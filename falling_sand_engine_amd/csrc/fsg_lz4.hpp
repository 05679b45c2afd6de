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
// Every lane of the wave runs the same parse on wave-uniform values; the
// literal and match copies are split over the lanes.  A match copies from
// bytes before its own start only (its output is periodic with period
// `offset` when offset < length), so no lane reads a byte written by the
// same copy; bytes of earlier copies are visible to every lane of the wave
// because one wave's vector memory operations go through one L1 in order.
// The same code runs on the host for tests, with the lanes taken in turn.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsg_lz4 {

constexpr int MINMATCH = 4, LASTLITERALS = 5, MFLIMIT = 12, RUN_MASK = 15;

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

// One lane of a wave of 64.
struct WaveLane {
    int lane;
    __host__ __device__ void copy(uint8_t *d, const uint8_t *s, int len) const { copy_lane(d, s, len, lane, 64); }
    __host__ __device__ void match(uint8_t *d, int off, int len) const { match_lane(d, off, len, lane, 64); }
};

// Decodes the LZ4 block src[0, n) into dst[0, cap).  Returns the decoded
// size, or -1 for a block that lz4.c's rules reject.  Reads only src[0, n),
// writes only dst[0, cap).
template <class Lanes>
__host__ __device__ inline int decode_block(const uint8_t *src, int n, uint8_t *dst, int cap, const Lanes &L) {
    if (n <= 0 || cap < 0) return -1;
    if (cap == 0) return (n == 1 && src[0] == 0) ? 0 : -1;
    int ip = 0, op = 0;
    for (;;) {
        if (ip >= n) return -1;
        const int token = src[ip++];
        int ll = token >> 4;
        if (ll == RUN_MASK) {  // read_variable_length(ip, iend - RUN_MASK, initial check)
            const int lim = n - RUN_MASK;
            if (ip >= lim) return -1;
            int s;
            do {
                s = src[ip++];
                ll += s;
                if (ip > lim || ll > cap) return -1;
            } while (s == 255);
        }
        const int cpy = op + ll;
        if (cpy > cap - MFLIMIT || ip + ll > n - (2 + 1 + LASTLITERALS)) {
            // the last sequence: it must end the input exactly and fit
            if (ip + ll != n || cpy > cap) return -1;
            L.copy(dst + op, src + ip, ll);
            return cpy;
        }
        L.copy(dst + op, src + ip, ll);
        ip += ll;
        op = cpy;
        const int off = src[ip] | (src[ip + 1] << 8);  // ip + 2 <= n - 6 here
        ip += 2;
        int ml = token & 15;
        if (ml == 15) {  // read_variable_length(ip, iend - LASTLITERALS + 1, no initial check)
            const int lim = n - LASTLITERALS + 1;
            int s;
            do {
                s = src[ip++];
                ml += s;
                if (ip > lim || ml > cap) return -1;
            } while (s == 255);
        }
        ml += MINMATCH;
        if (off > op) return -1;
        if (op + ml > cap - LASTLITERALS) return -1;
        L.match(dst + op, off, ml);
        op += ml;
    }
}

}  // namespace fsg_lz4

// Host build of the library's LZ4 block decoder (falling_sand_engine_amd/
// csrc/fsg_lz4.hpp) for CPU tests: the same parse and per-lane copy code the
// decode kernel runs, with the 64 lanes of a wave taken one after another in
// each copy (a copy reads only bytes written before it starts, so the order
// of lanes within it does not matter).
#include "fsg_lz4.hpp"

namespace {
struct HostWave {
    void copy(uint8_t *d, const uint8_t *s, int len) const {
        for (int lane = 0; lane < 64; lane++) fsg_lz4::copy_lane(d, s, len, lane, 64);
    }
    void match(uint8_t *d, int off, int len) const {
        for (int lane = 0; lane < 64; lane++) fsg_lz4::match_lane(d, off, len, lane, 64);
    }
};
}  // namespace

extern "C" int fsg_test_lz4_decode(const uint8_t *src, int n, uint8_t *dst, int cap) {
    return fsg_lz4::decode_block(src, n, dst, cap, HostWave{});
}

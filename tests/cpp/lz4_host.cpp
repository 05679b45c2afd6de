// Host build of the library's LZ4 block decoder (falling_sand_engine_amd/
// csrc/fsg_lz4.hpp) for CPU tests: the parse the decode kernel runs, with
//   * fsg_test_lz4_decode: plain memory, each copy running the 64 lanes'
//     shares one after another;
//   * fsg_test_lz4_decode_staged: the kernel's staged access (block window
//     and output ring, LDS on the device) and its group decode of short
//     sequences, with the 64 lanes taken in turn inside each step, as the
//     device's lanes take each step together.
#include "fsg_lz4.hpp"

namespace {

// Lane policy of the host build: all 64 lanes in turn.
struct HostLanes {
    template <class F>
    void each(F f) const {
        for (int lane = 0; lane < fsg_lz4::NL; lane++) f(lane);
    }
    void barrier() const {}
    uint32_t uniform(uint32_t x) const { return x; }
    // the lanes' registers are the staged row on the host
    uint32_t lane_value(uint32_t, int k, const uint8_t *row) const { return row[k]; }
    uint32_t own(uint32_t, int lane, const uint8_t *row) const { return row[lane]; }
    uint32_t mod(int i) const { return fsg_lz4::modtab_entry(i); }
    struct V {
        uint32_t v[fsg_lz4::NL];
    };
    template <class F>
    V make(F f) const {
        V r;
        for (int lane = 0; lane < fsg_lz4::NL; lane++) r.v[lane] = f(lane);
        return r;
    }
    uint32_t get(const V &x, int lane) const { return x.v[lane]; }
    uint32_t lane_of(const V &x, int k) const { return x.v[k]; }
    uint64_t ballot(const V &x) const {
        uint64_t m = 0;
        for (int lane = 0; lane < fsg_lz4::NL; lane++) m |= (uint64_t)(x.v[lane] != 0) << lane;
        return m;
    }
    V shfl(const V &x, const V &i) const {
        V r;
        for (int lane = 0; lane < fsg_lz4::NL; lane++) r.v[lane] = x.v[i.v[lane] & (fsg_lz4::NL - 1)];
        return r;
    }
    uint32_t below(uint64_t m, int lane) const { return (uint32_t)fsg_lz4::popc64(m & ((1ull << lane) - 1)); }
};

struct AllLanes {
    const uint8_t *src;
    int n;
    uint8_t *dst;
    int at0 = 0;
    uint32_t byte(int i) { return src[i]; }
    void look(int i) { at0 = i; }
    uint32_t at(int k) { return at0 + k < n ? src[at0 + k] : 0u; }
    void lit(int op, int ip, int len, int) {
        for (int l = 0; l < fsg_lz4::NL; l++) fsg_lz4::copy_lane(dst + op, src + ip, len, l, fsg_lz4::NL);
    }
    void match(int op, int off, int len) {
        for (int l = 0; l < fsg_lz4::NL; l++) fsg_lz4::match_lane(dst + op, off, len, l, fsg_lz4::NL);
    }
    bool group(int &, int &) { return false; }  // one sequence at a time
};

}  // namespace

extern "C" int fsg_test_lz4_decode(const uint8_t *src, int n, uint8_t *dst, int cap) {
    AllLanes m{src, n, dst};
    return fsg_lz4::decode_block(n, cap, m);
}

extern "C" int fsg_test_lz4_decode_staged(const uint8_t *src, int n, uint8_t *dst, int cap) {
    alignas(16) static uint8_t win[fsg_lz4::IW];
    static uint8_t ring[fsg_lz4::RING];
    static uint32_t gdesc[fsg_lz4::GMAX];
    fsg_lz4::Staged<HostLanes> m{src, n, dst, win, ring, gdesc, -(1 << 30), HostLanes{}, 0u, 0};
    return fsg_lz4::decode_block(n, cap, m);
}

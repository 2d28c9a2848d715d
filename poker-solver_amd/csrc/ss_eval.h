// ss_eval.h -- 7-card hand ranks, bit-exact with the reference encoding.
//
// The reference (src/hand_eval.h:40-156, device copy street_solve.cu:467-510)
// ranks a 7-card hand as the max over its 21 five-card subsets of
// cat<<20 | rank nibbles (cat 1 high card .. 9 straight flush). This evaluator
// computes the same value directly from per-suit 13-bit rank masks and rank
// counts: one pass over the 7 cards, a straight lookup table indexed by a rank
// mask (kept in LDS by the device kernels), and bit scans for kickers.
// Category precedence used: a 5+-card suit excludes quads and full houses,
// so flush is tested after them and before straights.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SS_HD __host__ __device__
#else
#define SS_HD
#endif

namespace ss {

// Highest straight in a 13-bit rank mask: top rank + 1 (wheel -> 3 + 1), 0 if none.
SS_HD inline uint8_t straight_top_slow(uint32_t m) {
    for (int hi = 12; hi >= 4; --hi) {
        uint32_t need = 0x1Fu << (hi - 4);
        if ((m & need) == need) return (uint8_t)(hi + 1);
    }
    if ((m & 0x100Fu) == 0x100Fu) return 4;  // A-2-3-4-5: five-high
    return 0;
}

// Build the 8192-entry straight table (host or device).
SS_HD inline void fill_straight_table(uint8_t *tab, int begin, int stride) {
    for (int m = begin; m < 8192; m += stride) tab[m] = straight_top_slow((uint32_t)m);
}

// Top k set bits of a rank mask, high to low, packed as nibbles into the
// positions 16,12,8,4,0 starting at `shift`.
SS_HD inline uint32_t top_ranks(uint32_t m, int k, int shift) {
    uint32_t v = 0;
    for (int i = 0; i < k && m; ++i) {
        int r = 31 - __builtin_clz(m);
        v |= (uint32_t)r << shift;
        shift -= 4;
        m &= ~(1u << r);
    }
    return v;
}

SS_HD inline int top_bit(uint32_t m) { return 31 - __builtin_clz(m); }

// tab: straight table (8192 bytes), in LDS on the device.
SS_HD inline uint32_t eval7_lut(const int *c, const uint8_t *tab) {
    uint32_t suit[4] = {0, 0, 0, 0};
    uint32_t seen1 = 0, seen2 = 0, seen3 = 0, seen4 = 0;  // ranks seen >=1,2,3,4 times
    for (int i = 0; i < 7; ++i) {
        uint32_t b = 1u << (c[i] >> 2);
        suit[c[i] & 3] |= b;
        seen4 |= seen3 & b;
        seen3 |= seen2 & b;
        seen2 |= seen1 & b;
        seen1 |= b;
    }
    // flush suit, if any (at most one suit can hold 5+ of 7 cards)
    uint32_t fm = 0;
    for (int s = 0; s < 4; ++s)
        if (__builtin_popcount(suit[s]) >= 5) fm = suit[s];
    if (fm) {
        uint8_t sf = tab[fm];
        if (sf) return (9u << 20) | ((uint32_t)(sf - 1) << 16);
    }
    uint32_t quads = seen4;
    uint32_t trips = seen3 & ~seen4;
    uint32_t pairs = seen2 & ~seen3;
    if (quads) {
        int q = top_bit(quads);
        uint32_t rest = seen1 & ~(1u << q);
        return (8u << 20) | ((uint32_t)q << 16) | ((uint32_t)top_bit(rest) << 12);
    }
    if (trips) {
        int t = top_bit(trips);
        uint32_t other = (trips & ~(1u << t)) | pairs;  // second trips or any pair
        if (other) return (7u << 20) | ((uint32_t)t << 16) | ((uint32_t)top_bit(other) << 12);
    }
    if (fm) return (6u << 20) | top_ranks(fm, 5, 16);
    uint8_t st = tab[seen1];
    if (st) return (5u << 20) | ((uint32_t)(st - 1) << 16);
    if (trips) {
        int t = top_bit(trips);
        return (4u << 20) | ((uint32_t)t << 16) | top_ranks(seen1 & ~(1u << t), 2, 12);
    }
    if (pairs) {
        int p1 = top_bit(pairs);
        uint32_t rest = pairs & ~(1u << p1);
        if (rest) {
            int p2 = top_bit(rest);
            uint32_t kick = seen1 & ~(1u << p1) & ~(1u << p2);
            return (3u << 20) | ((uint32_t)p1 << 16) | ((uint32_t)p2 << 12) | ((uint32_t)top_bit(kick) << 8);
        }
        return (2u << 20) | ((uint32_t)p1 << 16) | top_ranks(seen1 & ~(1u << p1), 3, 12);
    }
    return (1u << 20) | top_ranks(seen1, 5, 16);
}

}  // namespace ss

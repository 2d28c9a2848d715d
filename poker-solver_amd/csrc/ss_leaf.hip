// ss_leaf.hip -- depth-limit leaf values on the GPU (turn and flop leaves).
//
// Reference: python/leaf_values.py
//   _accumulate_nplayer_equity :573-691   per-runout equity of every hand of a
//                                         hero player against the joint
//                                         opponent hands (exact cross product
//                                         when it has <= 2000 tuples, else 2000
//                                         tuples drawn by random.Random(42))
//   compute_turn_leaf_values   :694-816   runouts = river cards (12 drawn by
//   compute_flop_leaf_equity   :408-570   numpy RandomState when any range has
//                                         > 50 hands); equity averaged over the
//                                         runouts where the hand was counted
//   biased_leaf_value          :91-122    first-order bias model per
//                                         (s_self, s_opp) profile pair, written
//                                         for all 4^NP profiles of every leaf
//
// Work split:
//   host   per (runout, hero) the opponents' unblocked hand lists and the
//          opponent-tuple pattern: the mixed-radix product (decoded on the
//          device from the tuple index) or the 2000 draws of a Python-
//          compatible Mersenne twister (one pattern per distinct list-length
//          tuple: the generator is reseeded with 42 for every hero and runout,
//          so draws depend only on the lengths)
//   device k_leaf_strength  eval7 of every (runout, player, hand)
//          k_leaf_tuples    one thread per (runout, hero, tuple): the record
//                           {card mask, best opponent rank, multiplicity of
//                           that rank}
//          k_leaf_equity    one 256-lane workgroup per (runout, hero, 256 hero
//                           hands): the records copied to LDS, then every
//                           hero hand walks them in order, summing 1 or
//                           1/ties in float64 exactly as the reference's
//                           loop does
//          k_leaf_reduce    per hand, the runout average in runout order
//          k_leaf_bias      one thread per output value (leaf, profile,
//                           player, hand), float64 without contraction, then
//                           rounded to float32
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <map>
#include <cstring>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "ss_eval.h"
#include "ss_internal.h"
#include "../../include/street_solve_leaf.h"

namespace ss {

namespace {

constexpr int LEAF_BLOCK = 256;

// Python's random.Random: MT19937 (Matsumoto & Nishimura, mt19937ar) seeded
// by init_by_array with the seed's 32-bit words, and _randbelow(n) drawing
// bit_length(n) top bits of a 32-bit output until the value is below n.
struct PyRandom {
    uint32_t mt[624];
    int mti = 625;

    void init_genrand(uint32_t s) {
        mt[0] = s;
        for (mti = 1; mti < 624; ++mti) mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + (uint32_t)mti;
    }
    explicit PyRandom(uint32_t seed) {
        init_genrand(19650218u);
        int i = 1;
        for (int k = 624; k; --k) {  // key = {seed}, length 1
            mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + seed;
            if (++i >= 624) { mt[0] = mt[623]; i = 1; }
        }
        for (int k = 623; k; --k) {
            mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
            if (++i >= 624) { mt[0] = mt[623]; i = 1; }
        }
        mt[0] = 0x80000000u;
        mti = 624;
    }
    uint32_t next() {
        if (mti >= 624) {
            for (int k = 0; k < 624; ++k) {
                uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            mti = 0;
        }
        uint32_t y = mt[mti++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    int below(int n) {
        int k = 32 - __builtin_clz((uint32_t)n);
        uint32_t r;
        do r = next() >> (32 - k); while (r >= (uint32_t)n);
        return (int)r;
    }
};

// Per (runout, hero) tuple source.
struct RpDesc {
    int count;    // tuples to walk (0: hero skipped)
    int pattern;  // offset into the draw pool (sampled), -1 for the product
    long long rec;  // first record of this (runout, hero) in the packed record array
    int lens[SS_MAX_PLAYERS - 1];
};

// One opponent tuple of a (runout, hero) walk, 32 bytes.
struct TupleRec {
    uint32_t mlo, mhi;  // opponents' card mask; all ones when opponents share a card
    uint32_t best;      // best opponent rank
    uint32_t mult;      // opponents holding it
    double tie;         // 1 / (mult + 1), rounded once as the reference's 1.0 / ties
    double unused;
};

// Device buffers of the leaf calls, per device: grown to the largest call so
// far and kept until the process ends (no allocation on repeated calls of a
// size seen before). The calls are synchronous; the mutex serialises callers.
enum LeafSlot { L_HC, L_RUN, L_STR, L_VALID, L_DESC, L_POOL, L_EQ, L_CNT, L_RAW, L_RECS,
                V_RAW, V_POT, V_NH, V_OUT, V_ROW, L_NSLOTS };

struct LeafArena {
    std::mutex mu;
    void *p[L_NSLOTS] = {};
    size_t cap[L_NSLOTS] = {};
    void *pinned = nullptr;  // host staging for large read-backs (hipHostMalloc)
    size_t pinned_cap = 0;
};

// Large results are read back into pinned staging at full link speed and
// copied to the caller's (pageable) buffer by several host threads.
constexpr size_t PINNED_MIN_BYTES = 4u << 20;
constexpr int COPY_THREADS = 8;

int read_back(LeafArena &a, void *dst, const void *src, size_t bytes) {
    if (bytes < PINNED_MIN_BYTES) {
        SS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return SS_OK;
    }
    if (a.pinned_cap < bytes) {
        if (a.pinned) (void)hipHostFree(a.pinned);
        a.pinned = nullptr;
        a.pinned_cap = 0;
        SS_HIP(hipHostMalloc(&a.pinned, bytes, hipHostMallocDefault));
        a.pinned_cap = bytes;
    }
    SS_HIP(hipMemcpy(a.pinned, src, bytes, hipMemcpyDeviceToHost));
    // Chunk t goes to thread t; a chunk whose thread could not be started
    // (std::system_error from the constructor) is copied on this thread, so
    // no exception crosses the C-ABI and every byte is copied exactly once.
    const size_t chunk = (bytes + COPY_THREADS - 1) / COPY_THREADS;
    auto copy_chunk = [&a, dst, bytes, chunk](int t) {
        size_t b = (size_t)t * chunk, e = std::min(bytes, b + chunk);
        if (b < e) std::memcpy((char *)dst + b, (const char *)a.pinned + b, e - b);
    };
    std::vector<std::thread> th;
    int started = 0;
    try {
        th.reserve(COPY_THREADS);
        for (; started < COPY_THREADS; ++started) th.emplace_back(copy_chunk, started);
    } catch (...) {
    }
    for (int t = started; t < COPY_THREADS; ++t) copy_chunk(t);
    for (auto &x : th) x.join();
    return SS_OK;
}

constexpr int LEAF_MAX_DEVICES = 16;

int leaf_arena(LeafArena **out) {
    static LeafArena arenas[LEAF_MAX_DEVICES];
    int dev = 0;
    SS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= LEAF_MAX_DEVICES) return set_error(SS_ERR_UNSUPPORTED, "device %d", dev);
    *out = &arenas[dev];
    return SS_OK;
}

template <class T>
int arena_buf(LeafArena &a, int slot, size_t n, T **out) {
    size_t b = std::max<size_t>(n, 1) * sizeof(T);
    if (a.cap[slot] < b) {
        if (a.p[slot]) (void)hipFree(a.p[slot]);
        a.p[slot] = nullptr;
        a.cap[slot] = 0;
        SS_HIP(hipMalloc(&a.p[slot], b));
        a.cap[slot] = b;
    }
    *out = static_cast<T *>(a.p[slot]);
    return SS_OK;
}

template <class T>
int arena_upload(LeafArena &a, int slot, const T *h, size_t n, T **out) {
    int rc = arena_buf(a, slot, n, out);
    if (rc) return rc;
    if (n) SS_HIP(hipMemcpy(*out, h, n * sizeof(T), hipMemcpyHostToDevice));
    return SS_OK;
}

}  // namespace

__global__ __launch_bounds__(LEAF_BLOCK) void k_leaf_strength(const uint16_t *__restrict__ hc,
                                                              const int *__restrict__ runouts,
                                                              uint32_t *__restrict__ str, int np, int mh,
                                                              int total) {
    __shared__ uint8_t tab[8192];
    fill_straight_table(tab, threadIdx.x, blockDim.x);
    __syncthreads();
    int i = blockIdx.x * LEAF_BLOCK + threadIdx.x;
    if (i >= total) return;
    int h = i % mh, p = (i / mh) % np, r = i / (mh * np);
    uint32_t v = hc[p * mh + h];
    const int *b = runouts + 5 * r;
    int c[7] = {b[0], b[1], b[2], b[3], b[4], (int)(v & 0xff), (int)(v >> 8)};
    bool dead = v == 0xffffu;
    for (int k = 0; k < 5; ++k) dead |= (c[k] == c[5]) | (c[k] == c[6]);
    str[i] = dead ? 0u : eval7_lut(c, tab);
}

// biased_leaf_value's tables (python/leaf_values.py:84-88), by bias profile
// 0 unmodified, 1 fold, 2 call, 3 raise.
__constant__ double c_fold_prob[4] = {0.0, 0.35, 0.0, 0.0};
__constant__ double c_pot_factor[4] = {1.0, 0.7, 1.0, 1.4};


// One record per (runout, hero, tuple): every tuple source resolved to
// hand indices, the opponents' card mask, best rank and its multiplicity.
__global__ __launch_bounds__(LEAF_BLOCK) void k_leaf_tuples(const uint16_t *__restrict__ hc,
                                                            const uint32_t *__restrict__ str,
                                                            const int16_t *__restrict__ valid,
                                                            const RpDesc *__restrict__ desc,
                                                            const int16_t *__restrict__ pool,
                                                            TupleRec *__restrict__ recs, int np, int mh) {
    const int rp = blockIdx.y, r = rp / np, p = rp % np;
    const int t = blockIdx.x * LEAF_BLOCK + threadIdx.x;
    const RpDesc d = desc[rp];
    if (t >= d.count) return;
    const int no = np - 1;
    // Opponents taken last to first: the product index is mixed radix with
    // the last opponent fastest; mask, best rank and multiplicity do not
    // depend on the order.
    uint64_t m = 0;
    uint32_t best = 0, mult = 0;
    int x = t;
    for (int k = no - 1; k >= 0; --k) {
        int pos;
        if (d.pattern < 0) {
            int len = desc[rp].lens[k];
            pos = x % len;
            x /= len;
        } else {
            pos = pool[d.pattern + t * no + k];
        }
        int q = k + (k >= p);
        int hidx = valid[((size_t)r * np + q) * mh + pos];
        uint32_t v = hc[q * mh + hidx];
        m |= (1ull << (v & 0xff)) | (1ull << (v >> 8));
        uint32_t s = str[((size_t)r * np + q) * mh + hidx];
        if (s > best) { best = s; mult = 1; }
        else if (s == best) ++mult;
    }
    if (__builtin_popcountll(m) != 2 * no) m = ~0ull;  // every hero hand collides with it
    const double tie = mult == 1 ? 1.0 / 2 : mult == 2 ? 1.0 / 3 : mult == 3 ? 1.0 / 4 : mult == 4 ? 1.0 / 5 : 1.0 / 6;
    recs[d.rec + t] =
        TupleRec{(uint32_t)m, (uint32_t)(m >> 32), best, mult, tie, 0.0};
}

// One tuple of the walk: counted when the hero holds none of its cards (a
// tuple whose opponents share a card has an all-ones mask); then 1 for a win,
// the tuple's 1 / ties for a tie. Adding 0.0 for a loss or a skipped tuple
// leaves acc as the reference's loop leaves it.
__device__ __forceinline__ void tuple_step(const TupleRec &e, uint32_t hs, uint32_t hlo, uint32_t hhi,
                                           double &acc, int &nvalid) {
    const bool use = ((e.mlo & hlo) | (e.mhi & hhi)) == 0;
    const double add = hs > e.best ? 1.0 : (hs == e.best ? e.tie : 0.0);
    nvalid += use ? 1 : 0;
    acc += use ? add : 0.0;
}

// One 256-lane workgroup per (runout, hero, 256 hero hands): the (runout,
// hero) records are copied to LDS (64 KB), then each lane walks them in order
// for its hand, four records per LDS wait.
constexpr int EQ_BLOCK = 256;

__global__ __launch_bounds__(EQ_BLOCK) void k_leaf_equity(const uint16_t *__restrict__ hc,
                                                          const uint32_t *__restrict__ str,
                                                          const RpDesc *__restrict__ desc,
                                                          const TupleRec *__restrict__ recs,
                                                          double *__restrict__ eq,
                                                          uint8_t *__restrict__ cnt, int np, int mh) {
    __shared__ TupleRec srec[SS_LEAF_OPP_COMBO_LIMIT];
    const int rp = blockIdx.y, r = rp / np, p = rp % np;
    const int count = desc[rp].count;
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(recs + desc[rp].rec);
        uint4 *dst = reinterpret_cast<uint4 *>(srec);
        const int n16 = count * (int)(sizeof(TupleRec) / sizeof(uint4));
        for (int i = threadIdx.x; i < n16; i += EQ_BLOCK) dst[i] = src[i];
    }
    __syncthreads();
    const int h = blockIdx.x * EQ_BLOCK + threadIdx.x;
    if (h >= mh) return;
    const size_t o = ((size_t)r * np + p) * mh + h;
    const uint32_t hs = str[o];
    double acc = 0.0;
    int nvalid = 0;
    if (hs && count) {
        const uint32_t v = hc[p * mh + h];
        const uint64_t hm = (1ull << (v & 0xff)) | (1ull << (v >> 8));
        const uint32_t hlo = (uint32_t)hm, hhi = (uint32_t)(hm >> 32);
        int t = 0;
        for (; t + 4 <= count; t += 4) {
            const TupleRec e0 = srec[t], e1 = srec[t + 1], e2 = srec[t + 2], e3 = srec[t + 3];
            tuple_step(e0, hs, hlo, hhi, acc, nvalid);
            tuple_step(e1, hs, hlo, hhi, acc, nvalid);
            tuple_step(e2, hs, hlo, hhi, acc, nvalid);
            tuple_step(e3, hs, hlo, hhi, acc, nvalid);
        }
        for (; t < count; ++t) tuple_step(srec[t], hs, hlo, hhi, acc, nvalid);
    }
    eq[o] = nvalid ? acc / (double)nvalid : 0.0;
    cnt[o] = nvalid ? 1 : 0;
}

// Per hand, the runout average in runout order. Loads are unconditional and
// an uncounted runout adds 0.0 (exact), so eight runouts' loads go out
// together.
__global__ __launch_bounds__(LEAF_BLOCK) void k_leaf_reduce(const double *__restrict__ eq,
                                                            const uint8_t *__restrict__ cnt,
                                                            double *__restrict__ raw, int nrun, int n) {
    int i = blockIdx.x * LEAF_BLOCK + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    int c = 0;
#pragma unroll 8
    for (int r = 0; r < nrun; ++r) {
        const double e = eq[(size_t)r * n + i];
        const int k = cnt[(size_t)r * n + i];
        s += k ? e : 0.0;
        c += k;
    }
    raw[i] = c > 0 ? s / (double)c : 0.0;
}

__device__ __forceinline__ double biased_value(double equity, double pot, int s_self, int s_opp) {
#pragma clang fp contract(off)
    double fp = c_fold_prob[s_opp];
    double showdown_pot = pot * (c_pot_factor[s_self] * c_pot_factor[s_opp]);
    double fold_payoff = 0.5 * pot;
    double showdown_payoff = (equity - 0.5) * showdown_pot;
    return fp * fold_payoff + (1.0 - fp) * showdown_payoff;
}

// Value of profile row `row` (leaf * 4^np + combo) for player p's hand h,
// element j = p * mh + h of the equity: biased_leaf_value with the profile
// digits of the row (player 0 least significant), 0 for rows no leaf info
// names (pot LLONG_MIN) and hands past the player's count.
__device__ __forceinline__ float leaf_row_value(const double *__restrict__ raw, long long pot, long long row,
                                                int j, int p, int h, int np, const int *__restrict__ nh) {
    if (pot == LLONG_MIN || h >= nh[p]) return 0.0f;
    const int combo = (int)(row & ((1ll << (2 * np)) - 1));
    int s_self = (combo >> (2 * p)) & 3, s_opp;
    if (np > 2) {
        int other = 0;
        for (int q = 0; q < np; ++q)
            if (q != p) other += (combo >> (2 * q)) & 3;
        s_opp = (int)rint((double)other / (double)(np - 1));  // round half to even
        s_opp = s_opp < 0 ? 0 : (s_opp > 3 ? 3 : s_opp);
    } else {
        s_opp = (combo >> (2 * (1 - p))) & 3;
    }
    return (float)biased_value(raw[j], (double)pot, s_self, s_opp);
}

__global__ __launch_bounds__(LEAF_BLOCK) void k_leaf_bias(const double *__restrict__ raw,
                                                          const long long *__restrict__ flat_pot,
                                                          const int *__restrict__ nh, float *__restrict__ out,
                                                          int np, int mh) {
    const int j = blockIdx.y * LEAF_BLOCK + threadIdx.x;  // p * mh + h
    if (j >= np * mh) return;
    const long long leaf = blockIdx.x;
    const int p = j / mh, h = j - p * mh;
    out[leaf * np * mh + j] = leaf_row_value(raw, flat_pot[leaf], leaf, j, p, h, np, nh);
}

// The same values gathered into a solve's device leaf buffer [num_leaves][np]
// [mh]: tree leaf l takes profile row row_of[l] (row_of == nullptr: row l),
// rows outside [0, rows) give 0.
__global__ __launch_bounds__(LEAF_BLOCK) void k_leaf_bias_into(const double *__restrict__ raw,
                                                               const long long *__restrict__ flat_pot,
                                                               const int *__restrict__ row_of,
                                                               const int *__restrict__ nh, float *__restrict__ out,
                                                               int np, int mh, long long rows) {
    const int j = blockIdx.y * LEAF_BLOCK + threadIdx.x;
    if (j >= np * mh) return;
    const long long l = blockIdx.x;
    const long long row = row_of ? (long long)row_of[l] : l;
    const int p = j / mh, h = j - p * mh;
    const long long pot = row >= 0 && row < rows ? flat_pot[row] : LLONG_MIN;
    out[l * np * mh + j] = leaf_row_value(raw, pot, row, j, p, h, np, nh);
}

namespace {

int check_hands(int np, const int *hands, const int *nh, int mh) {
    if (np < 2 || np > SS_MAX_PLAYERS) return set_error(SS_ERR_ARG, "num_players %d (2..6)", np);
    if (mh < 1 || mh > SS_LEAF_MAX_HANDS) return set_error(SS_ERR_ARG, "max_hands %d (1..1326)", mh);
    if (!nh) return set_error(SS_ERR_ARG, "num_hands is NULL");
    for (int p = 0; p < np; ++p) {
        if (nh[p] < 0 || nh[p] > mh)
            return set_error(SS_ERR_ARG, "player %d has %d hands (max_hands %d)", p, nh[p], mh);
        for (int h = 0; h < nh[p]; ++h) {
            int a = hands[(p * mh + h) * 2], b = hands[(p * mh + h) * 2 + 1];
            if (a < 0 || a > 51 || b < 0 || b > 51 || a == b)
                return set_error(SS_ERR_ARG, "player %d hand %d has cards (%d, %d)", p, h, a, b);
        }
    }
    return SS_OK;
}

}  // namespace

// Host half of an equity call: argument checks, the packed hole cards, the
// unblocked hand list of every (runout, player), the tuple source of every
// (runout, hero) and the record offsets.
struct EquityPlan {
    int np = 0, mh = 0, nrun = 0;
    std::vector<uint16_t> hc;
    std::vector<int16_t> valid;
    std::vector<RpDesc> desc;
    std::vector<int16_t> pool;
    long long nrec = 0;
    size_t nrp() const { return (size_t)nrun * np; }
    size_t nall() const { return nrp() * mh; }
};

// Device buffers of one equity call.
struct EquityBufs {
    uint16_t *hc;
    int *run;
    int16_t *valid, *pool;
    RpDesc *desc;
    uint32_t *str;
    double *eq, *raw;
    uint8_t *cnt;
    TupleRec *recs;
};

int plan_equity(int np, const int *hands, const int *nh, int mh, const int *runouts, int nrun, const double *raw,
                EquityPlan &pl) {
    int rc = check_hands(np, hands, nh, mh);
    if (rc) return rc;
    if (nrun < 0 || (nrun > 0 && !runouts)) return set_error(SS_ERR_ARG, "bad runouts");
    if (!raw) return set_error(SS_ERR_ARG, "raw_equity is NULL");
    if ((long long)nrun * np > 65535)
        return set_error(SS_ERR_ARG, "%d runouts x %d players exceeds 65535 (runout, player) pairs", nrun, np);
    for (int r = 0; r < nrun; ++r) {
        uint64_t m = 0;
        for (int k = 0; k < 5; ++k) {
            int c = runouts[5 * r + k];
            if (c < 0 || c > 51 || (m >> c & 1)) return set_error(SS_ERR_ARG, "runout %d: bad board", r);
            m |= 1ull << c;
        }
    }
    pl.np = np;
    pl.mh = mh;
    pl.nrun = nrun;
    const int no = np - 1;
    pl.hc.assign((size_t)np * mh, 0xffffu);
    for (int p = 0; p < np; ++p)
        for (int h = 0; h < nh[p]; ++h)
            pl.hc[p * mh + h] = (uint16_t)(hands[(p * mh + h) * 2] | hands[(p * mh + h) * 2 + 1] << 8);

    // Unblocked hand lists per (runout, player) and the tuple source per (runout, hero).
    pl.valid.assign((size_t)nrun * np * mh, 0);
    std::vector<int> vlen((size_t)nrun * np, 0);
    pl.desc.assign((size_t)nrun * np, RpDesc{});
    pl.pool.clear();
    std::map<std::vector<int>, int> patterns;
    for (int r = 0; r < nrun; ++r) {
        uint64_t bm = 0;
        for (int k = 0; k < 5; ++k) bm |= 1ull << runouts[5 * r + k];
        for (int p = 0; p < np; ++p) {
            int n = 0;
            for (int h = 0; h < nh[p]; ++h) {
                uint16_t v = pl.hc[p * mh + h];
                if (!((bm >> (v & 0xff)) & 1) && !((bm >> (v >> 8)) & 1))
                    pl.valid[((size_t)r * np + p) * mh + n++] = (int16_t)h;
            }
            vlen[r * np + p] = n;
        }
        for (int p = 0; p < np; ++p) {
            RpDesc &d = pl.desc[r * np + p];
            d = RpDesc{0, -1, 0, {0, 0, 0, 0, 0}};
            if (!vlen[r * np + p]) continue;
            std::vector<int> lens;
            long long total = 1;
            for (int k = 0; k < no; ++k) {
                int q = k + (k >= p);
                d.lens[k] = vlen[r * np + q];
                lens.push_back(d.lens[k]);
                total = std::min<long long>(total * d.lens[k], 1ll << 40);
            }
            if (total == 0) continue;
            if (total <= SS_LEAF_OPP_COMBO_LIMIT) {
                d.count = (int)total;
                continue;
            }
            auto it = patterns.find(lens);
            if (it == patterns.end()) {
                int off = (int)pl.pool.size();
                PyRandom rng(SS_LEAF_SAMPLE_SEED);
                for (int t = 0; t < SS_LEAF_OPP_COMBO_LIMIT; ++t)
                    for (int k = 0; k < no; ++k) pl.pool.push_back((int16_t)rng.below(lens[k]));
                it = patterns.emplace(lens, off).first;
            }
            d.count = SS_LEAF_OPP_COMBO_LIMIT;
            d.pattern = it->second;
        }
    }
    pl.nrec = 0;
    for (RpDesc &d : pl.desc) {
        d.rec = pl.nrec;
        pl.nrec += d.count;
    }
    return SS_OK;
}

// Device half: uploads, the four kernels, the read-back of the equity. Every
// element a kernel reads was written earlier in the same call, so buffers
// need no clearing (they may hold an earlier call's contents).
int run_equity(const EquityPlan &pl, const int *runouts, const EquityBufs &b, double *raw) {
    const int np = pl.np, mh = pl.mh, nrun = pl.nrun;
    const size_t nrp = pl.nrp(), nall = pl.nall();
    auto up = [](void *dst, const void *src, size_t bytes) {
        return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    SS_HIP(up(b.hc, pl.hc.data(), pl.hc.size() * sizeof(uint16_t)));
    SS_HIP(up(b.run, runouts, (size_t)nrun * 5 * sizeof(int)));
    SS_HIP(up(b.valid, pl.valid.data(), pl.valid.size() * sizeof(int16_t)));
    SS_HIP(up(b.desc, pl.desc.data(), pl.desc.size() * sizeof(RpDesc)));
    SS_HIP(up(b.pool, pl.pool.data(), pl.pool.size() * sizeof(int16_t)));
    k_leaf_strength<<<(unsigned)((nall + LEAF_BLOCK - 1) / LEAF_BLOCK), LEAF_BLOCK>>>(b.hc, b.run, b.str, np, mh,
                                                                                     (int)nall);
    SS_HIP(hipGetLastError());
    k_leaf_tuples<<<dim3((SS_LEAF_OPP_COMBO_LIMIT + LEAF_BLOCK - 1) / LEAF_BLOCK, (unsigned)nrp), LEAF_BLOCK>>>(
        b.hc, b.str, b.valid, b.desc, b.pool, b.recs, np, mh);
    SS_HIP(hipGetLastError());
    k_leaf_equity<<<dim3((mh + EQ_BLOCK - 1) / EQ_BLOCK, (unsigned)nrp), EQ_BLOCK>>>(b.hc, b.str, b.desc, b.recs,
                                                                                  b.eq, b.cnt, np, mh);
    SS_HIP(hipGetLastError());
    k_leaf_reduce<<<(np * mh + LEAF_BLOCK - 1) / LEAF_BLOCK, LEAF_BLOCK>>>(b.eq, b.cnt, b.raw, nrun, np * mh);
    SS_HIP(hipGetLastError());
    SS_HIP(hipMemcpy(raw, b.raw, (size_t)np * mh * sizeof(double), hipMemcpyDeviceToHost));
    return SS_OK;
}

int leaf_equity(int np, const int *hands, const int *nh, int mh, const int *runouts, int nrun, double *raw) {
    EquityPlan pl;
    int rc = plan_equity(np, hands, nh, mh, runouts, nrun, raw, pl);
    if (rc) return rc;
    if (nrun == 0) {
        std::fill(raw, raw + (size_t)np * mh, 0.0);
        return SS_OK;
    }
    LeafArena *A = nullptr;
    if ((rc = leaf_arena(&A))) return rc;
    std::lock_guard<std::mutex> lock(A->mu);
    EquityBufs b;
    if ((rc = arena_buf(*A, L_HC, pl.hc.size(), &b.hc)) || (rc = arena_buf(*A, L_RUN, (size_t)nrun * 5, &b.run)) ||
        (rc = arena_buf(*A, L_VALID, pl.valid.size(), &b.valid)) ||
        (rc = arena_buf(*A, L_DESC, pl.desc.size(), &b.desc)) || (rc = arena_buf(*A, L_POOL, pl.pool.size(), &b.pool)) ||
        (rc = arena_buf(*A, L_STR, pl.nall(), &b.str)) || (rc = arena_buf(*A, L_EQ, pl.nall(), &b.eq)) ||
        (rc = arena_buf(*A, L_CNT, pl.nall(), &b.cnt)) || (rc = arena_buf(*A, L_RAW, (size_t)np * mh, &b.raw)) ||
        (rc = arena_buf(*A, L_RECS, (size_t)pl.nrec, &b.recs)))
        return rc;
    return run_equity(pl, runouts, b, raw);
}

// ---------------------------------------------------------------- allocation-free equity
// The device memory of ss_leaf_equity_static: a fixed array in the code
// object, one copy per GPU, held until the process ends. Host code never
// passes `scratch` itself to a kernel or copy (in host code the name is a
// stand-in, not a device address); it uses the address hipGetSymbolAddress
// returns for the current device.
static __device__ __align__(256) unsigned char scratch[SS_LEAF_STATIC_SCRATCH_BYTES];

struct StaticScratch {
    std::once_flag once;
    std::mutex mu;               // serialises calls on one device
    unsigned char *ptr = nullptr;
    hipError_t err = hipSuccess;
};

// Device address of this device's copy of `scratch`, resolved once per
// device on its first call, with that device current.
int static_scratch(StaticScratch **out) {
    static StaticScratch per_dev[LEAF_MAX_DEVICES];
    int dev = 0;
    SS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= LEAF_MAX_DEVICES) return set_error(SS_ERR_UNSUPPORTED, "device %d", dev);
    StaticScratch &s = per_dev[dev];
    std::call_once(s.once, [&s]() {
        void *p = nullptr;
        s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        s.ptr = s.err == hipSuccess ? static_cast<unsigned char *>(p) : nullptr;
    });
    if (s.err != hipSuccess || !s.ptr)
        return set_error(SS_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(s.err));
    *out = &s;
    return SS_OK;
}

// Carves the buffers of `pl` out of [base, base + cap) at 256-byte
// alignment; returns the bytes needed (also when they exceed cap, with the
// pointers then unset).
size_t carve_static(const EquityPlan &pl, unsigned char *base, size_t cap, EquityBufs &b) {
    size_t off = 0;
    auto take = [&](auto *&ptr, size_t n) {
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(ptr)>>;
        off = (off + 255) & ~(size_t)255;
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        ptr = off + bytes <= cap ? reinterpret_cast<T *>(base + off) : nullptr;
        off += bytes;
    };
    take(b.hc, pl.hc.size());
    take(b.run, (size_t)pl.nrun * 5);
    take(b.valid, pl.valid.size());
    take(b.desc, pl.desc.size());
    take(b.pool, pl.pool.size());
    take(b.str, pl.nall());
    take(b.eq, pl.nall());
    take(b.cnt, pl.nall());
    take(b.raw, (size_t)pl.np * pl.mh);
    take(b.recs, (size_t)pl.nrec);
    return off;
}

int leaf_equity_static(int np, const int *hands, const int *nh, int mh, const int *runouts, int nrun, double *raw) {
    EquityPlan pl;
    int rc = plan_equity(np, hands, nh, mh, runouts, nrun, raw, pl);
    if (rc) return rc;
    EquityBufs b;
    const size_t need = carve_static(pl, nullptr, 0, b);
    if (need > SS_LEAF_STATIC_SCRATCH_BYTES)
        return set_error(SS_ERR_ARG,
                         "equity call needs %zu bytes of device scratch, more than the fixed %zu "
                         "(%d runouts x %d players x %d hands, %lld tuple records); use ss_leaf_equity",
                         need, (size_t)SS_LEAF_STATIC_SCRATCH_BYTES, nrun, np, mh, pl.nrec);
    if (nrun == 0) {
        std::fill(raw, raw + (size_t)np * mh, 0.0);
        return SS_OK;
    }
    StaticScratch *s = nullptr;
    if ((rc = static_scratch(&s))) return rc;
    std::lock_guard<std::mutex> lock(s->mu);
    carve_static(pl, s->ptr, SS_LEAF_STATIC_SCRATCH_BYTES, b);
    return run_equity(pl, runouts, b, raw);
}

int leaf_values(int np, const double *raw, const int *nh, int mh, const int *leaf_idx, const int *leaf_pot,
                int ninfo, float *out) {
    if (np < 2 || np > SS_MAX_PLAYERS) return set_error(SS_ERR_ARG, "num_players %d (2..6)", np);
    if (mh < 1 || mh > SS_LEAF_MAX_HANDS) return set_error(SS_ERR_ARG, "max_hands %d (1..1326)", mh);
    if (!raw || !nh || !out || ninfo < 0 || (ninfo > 0 && (!leaf_idx || !leaf_pot)))
        return set_error(SS_ERR_ARG, "bad arguments");
    for (int p = 0; p < np; ++p)
        if (nh[p] < 0 || nh[p] > mh) return set_error(SS_ERR_ARG, "player %d has %d hands (max_hands %d)", p, nh[p], mh);
    const long long cont = 1ll << (2 * np), nleaves = (long long)ninfo * cont;
    // Pot of each output leaf: the last leaf info naming it wins, leaves no
    // info names stay zero (LLONG_MIN marks them).
    std::vector<long long> flat(std::max<long long>(nleaves, 1), LLONG_MIN);
    for (int i = 0; i < ninfo; ++i) {
        if (leaf_idx[i] < 0) return set_error(SS_ERR_ARG, "leaf info %d has leaf_idx %d", i, leaf_idx[i]);
        for (long long c = 0; c < cont; ++c) {
            long long f = (long long)leaf_idx[i] * cont + c;
            if (f < nleaves) flat[f] = leaf_pot[i];
        }
    }
    const long long total = nleaves * np * mh;
    if (total == 0) return SS_OK;
    LeafArena *A = nullptr;
    int rc = leaf_arena(&A);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(A->mu);
    double *d_raw;
    long long *d_pot;
    int *d_nh;
    float *d_out;
    if ((rc = arena_upload(*A, V_RAW, raw, (size_t)np * mh, &d_raw)) ||
        (rc = arena_upload(*A, V_POT, flat.data(), (size_t)nleaves, &d_pot)) ||
        (rc = arena_upload(*A, V_NH, nh, (size_t)np, &d_nh)) || (rc = arena_buf(*A, V_OUT, (size_t)total, &d_out)))
        return rc;
    k_leaf_bias<<<dim3((unsigned)nleaves, (np * mh + LEAF_BLOCK - 1) / LEAF_BLOCK), LEAF_BLOCK>>>(d_raw, d_pot, d_nh,
                                                                                               d_out, np, mh);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(SS_ERR_HIP, "leaf values: %s", hipGetErrorString(e));
    return read_back(*A, out, d_out, (size_t)total * sizeof(float));
}

int leaf_values_into(float *d_leaf, int num_leaves, int np, int mh, const double *raw, const int *nh,
                     const int *leaf_idx, const int *leaf_pot, int ninfo, const int *row_of, hipStream_t st) {
    const long long cont = 1ll << (2 * np), rows = (long long)ninfo * cont;
    std::vector<long long> flat(std::max<long long>(rows, 1), LLONG_MIN);
    for (int i = 0; i < ninfo; ++i) {
        if (leaf_idx[i] < 0) return set_error(SS_ERR_ARG, "leaf info %d has leaf_idx %d", i, leaf_idx[i]);
        for (long long c = 0; c < cont; ++c) {
            long long f = (long long)leaf_idx[i] * cont + c;
            if (f < rows) flat[f] = leaf_pot[i];
        }
    }
    if (num_leaves <= 0) return SS_OK;
    LeafArena *A = nullptr;
    int rc = leaf_arena(&A);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(A->mu);
    double *d_raw;
    long long *d_pot;
    int *d_nh, *d_row = nullptr;
    if ((rc = arena_upload(*A, V_RAW, raw, (size_t)np * mh, &d_raw)) ||
        (rc = arena_upload(*A, V_POT, flat.data(), flat.size(), &d_pot)) ||
        (rc = arena_upload(*A, V_NH, nh, (size_t)np, &d_nh)) ||
        (row_of && (rc = arena_upload(*A, V_ROW, row_of, (size_t)num_leaves, &d_row))))
        return rc;
    k_leaf_bias_into<<<dim3((unsigned)num_leaves, (np * mh + LEAF_BLOCK - 1) / LEAF_BLOCK), LEAF_BLOCK, 0, st>>>(
        d_raw, d_pot, d_row, d_nh, d_leaf, np, mh, rows);
    SS_HIP(hipGetLastError());
    SS_HIP(hipStreamSynchronize(st));
    return SS_OK;
}

int py_choice_draws(int seed, const int *lens, int nlists, int count, int *out) {
    if (!lens || !out || nlists < 1 || count < 0) return set_error(SS_ERR_ARG, "bad arguments");
    for (int k = 0; k < nlists; ++k)
        if (lens[k] < 1) return set_error(SS_ERR_ARG, "list %d is empty", k);
    PyRandom rng((uint32_t)seed);
    for (int t = 0; t < count; ++t)
        for (int k = 0; k < nlists; ++k) out[t * nlists + k] = rng.below(lens[k]);
    return SS_OK;
}

}  // namespace ss

extern "C" {

int ss_leaf_equity(int num_players, const int *hands, const int *num_hands, int max_hands, const int *runouts,
                   int num_runouts, double *raw_equity) {
    ss::clear_error();
    if (!hands) return ss::set_error(SS_ERR_ARG, "hands is NULL");
    return ss::leaf_equity(num_players, hands, num_hands, max_hands, runouts, num_runouts, raw_equity);
}

int ss_leaf_values(int num_players, const double *raw_equity, const int *num_hands, int max_hands,
                   const int *leaf_idx, const int *leaf_pot, int num_leaf_infos, float *leaf_values) {
    ss::clear_error();
    return ss::leaf_values(num_players, raw_equity, num_hands, max_hands, leaf_idx, leaf_pot, num_leaf_infos,
                           leaf_values);
}

int ss_leaf_equity_static(int num_players, const int *hands, const int *num_hands, int max_hands,
                          const int *runouts, int num_runouts, double *raw_equity) {
    ss::clear_error();
    if (!hands) return ss::set_error(SS_ERR_ARG, "hands is NULL");
    return ss::leaf_equity_static(num_players, hands, num_hands, max_hands, runouts, num_runouts, raw_equity);
}

int ssdbg_py_choice(int seed, const int *lens, int num_lists, int count, int *out) {
    ss::clear_error();
    return ss::py_choice_draws(seed, lens, num_lists, count, out);
}

}  // extern "C"

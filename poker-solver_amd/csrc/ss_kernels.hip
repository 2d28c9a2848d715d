// ss_kernels.hip -- CDNA4 (gfx950) kernels of the single-street Linear-CFR solver.
//
// Semantics follow src/cuda/street_solve.cu (kernels :575-1185, loop
// :1471-1550); the data layout is this library's own (ss_device.h).
//
// Per traverser pass:
//   forward   one launch per BFS level of decision nodes: regret matching of
//             the acting player fused into the reach sweep (the reference's
//             separate regret-match pass :575 and reach pass :650), every
//             non-traverser's reach advanced in the same item.
//   terminals one workgroup per terminal node. An opponent's reach is
//             scattered into combo space (1326 slots, LDS); card removal then
//             has closed forms (no O(M^k) enumeration):
//               1 opponent:  Z(h) = S - C[x] - C[y] + M[x,y]
//               2 opponents: D(h) = Z1 Z2 - sum_c (C1[c]-M1[c,x]-M1[c,y])(C2[c]-..)
//                                   + Z_{M1*M2}
//             and showdowns split the same sums by strength class (<, =)
//             with prefix sums over strength-sorted combo lists. 3+ opponents
//             use the reference odometer (small ranges only; host-guarded).
//   backward  one launch per level, bottom-up: value, regret update with the
//             Linear-CFR discount t/(t+1), strategy sum += t * sigma (:1138-1185).
// The batched kernel runs all of this for one subgame per workgroup with
// barriers between levels; the FP64 sweeps evaluate the best response.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "../../include/street_solve.h"
#include "ss_device.h"
#include "ss_eval.h"

namespace ss {

#define SS_BATCH_BLOCK 512
#define NT_DEC 8

// ---------------------------------------------------------------------------
// hand-rank kernels
// ---------------------------------------------------------------------------

// cstr[k] for all 1326 combos k on a 5-card board (0 when a card is on the
// board). The straight table lives in LDS.
__global__ void __launch_bounds__(SS_BLOCK) k_combo_ranks(const int *board, uint32_t *cstr) {
    __shared__ uint8_t tab[8192];
    fill_straight_table(tab, threadIdx.x, blockDim.x);
    __syncthreads();
    int b0 = board[0], b1 = board[1], b2 = board[2], b3 = board[3], b4 = board[4];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < SS_NCOMBO; k += gridDim.x * blockDim.x) {
        // invert combo_id
        int a = 0, rem = k;
        while (rem >= 51 - a) { rem -= 51 - a; ++a; }
        int b = a + 1 + rem;
        bool dead = a == b0 || a == b1 || a == b2 || a == b3 || a == b4 || b == b0 || b == b1 || b == b2 ||
                    b == b3 || b == b4;
        int c[7] = {b0, b1, b2, b3, b4, a, b};
        cstr[k] = dead ? 0u : eval7_lut(c, tab);
    }
}

// ranks of explicit 7-card hands (test entry)
__global__ void __launch_bounds__(SS_BLOCK) k_eval7(const int *cards, int n, uint32_t *out) {
    __shared__ uint8_t tab[8192];
    fill_straight_table(tab, threadIdx.x, blockDim.x);
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = eval7_lut(cards + 7 * i, tab);
}

// Census over all C(52,7) hands: grid over the three lowest cards.
__global__ void __launch_bounds__(SS_BLOCK) k_eval_census(unsigned long long *counts, unsigned long long *sum) {
    __shared__ uint8_t tab[8192];
    __shared__ unsigned long long lc[10];
    __shared__ unsigned long long ls;
    fill_straight_table(tab, threadIdx.x, blockDim.x);
    if (threadIdx.x < 10) lc[threadIdx.x] = 0;
    if (threadIdx.x == 0) ls = 0;
    __syncthreads();
    // blockIdx.x indexes (c0, c1); threads stride over c2
    int c0 = 0, rem = blockIdx.x;
    while (rem >= 51 - c0) { rem -= 51 - c0; ++c0; }
    int c1 = c0 + 1 + rem;
    unsigned long long my[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    unsigned long long mysum = 0;
    for (int c2 = c1 + 1 + threadIdx.x; c2 < 52; c2 += blockDim.x) {
        int c[7] = {c0, c1, c2, 0, 0, 0, 0};
        for (c[3] = c2 + 1; c[3] < 52; ++c[3])
            for (c[4] = c[3] + 1; c[4] < 52; ++c[4])
                for (c[5] = c[4] + 1; c[5] < 52; ++c[5])
                    for (c[6] = c[5] + 1; c[6] < 52; ++c[6]) {
                        uint32_t v = eval7_lut(c, tab);
                        my[v >> 20]++;
                        mysum += v;
                    }
    }
    for (int i = 0; i < 10; ++i)
        if (my[i]) atomicAdd(&lc[i], my[i]);
    atomicAdd(&ls, mysum);
    __syncthreads();
    if (threadIdx.x < 10) atomicAdd(&counts[threadIdx.x], lc[threadIdx.x]);
    if (threadIdx.x == 0) atomicAdd(sum, ls);
}

// ---------------------------------------------------------------------------
// regret matching (street_solve.cu:575-614, :621-644). The reference's
// pruning predicate maps regrets < -10000 to 0, which matches what any
// negative regret contributes, so it is not evaluated.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void regret_match(const float *__restrict__ R, int row0, int na, int Mp, int h,
                                             const int *frozen, int n, float *sig) {
    float sum = 0.f;
#pragma unroll
    for (int a = 0; a < NT_DEC; ++a) {
        if (a < na) {
            float r = R[(size_t)(row0 + a) * Mp + h];
            sig[a] = r > 0.f ? r : 0.f;
            sum += sig[a];
        }
    }
    if (sum > 0.f) {
        float inv = 1.0f / sum;
#pragma unroll
        for (int a = 0; a < NT_DEC; ++a)
            if (a < na) sig[a] *= inv;
    } else {
        float u = 1.0f / (float)na;
#pragma unroll
        for (int a = 0; a < NT_DEC; ++a)
            if (a < na) sig[a] = u;
    }
    if (frozen) {
        int fa = frozen[n];
        if (fa >= 0) {
#pragma unroll
            for (int a = 0; a < NT_DEC; ++a)
                if (a < na) sig[a] = a == fa ? 1.f : 0.f;
        }
    }
}

// ---------------------------------------------------------------------------
// forward: reach of every non-traverser through decision node n, hand h
// ---------------------------------------------------------------------------
// The reach of a player who does not act at n passes to the children
// unchanged and is found through the reach-row table (rrow), so only the
// actor's child rows are written, and nothing when the traverser acts.
__device__ __forceinline__ void forward_one(const SlotDesc *__restrict__ s, int n, int h, int trav) {
    const int Mp = s->Mp, N = s->N, NP = s->NP;
    const DevNode &nd = s->nodes[n];
    int actor = nd.player, na = nd.nchild, fc = nd.first_child;
    if (actor == trav || actor < 0 || actor >= NP || h >= s->nh[actor]) return;
    float sig[NT_DEC];
    regret_match(s->R, nd.act_base, na, Mp, h, s->frozen, n, sig);
    float *rr = s->reach + (size_t)actor * N * Mp;
    const float pr = rr[(size_t)s->rrow[(size_t)actor * N + n] * Mp + h];
    for (int a = 0; a < na; ++a) {
        int c = s->children[fc + a];
        rr[(size_t)c * Mp + h] = pr * sig[a];
    }
}

__device__ __forceinline__ void forward_items(const SlotDesc *__restrict__ s, int level, int trav, int begin,
                                              int stride) {
    const int Mp = s->Mp;
    const int lo = s->dec_off[level];
    const int total = (s->dec_off[level + 1] - lo) * Mp;
    for (int it = begin; it < total; it += stride) {
        int ni = it / Mp, h = it - ni * Mp;
        forward_one(s, s->dec_lvl[lo + ni], h, trav);
    }
}

// ---------------------------------------------------------------------------
// backward: value + regret / strategy-sum update at decision node n, hand h
// ---------------------------------------------------------------------------
__device__ __forceinline__ void backward_one(const SlotDesc *__restrict__ s, int n, int h, int trav, int iter) {
    const int Mp = s->Mp;
    if (h >= s->nh[trav]) return;
    const DevNode &nd = s->nodes[n];
    int na = nd.nchild, fc = nd.first_child;
    float v = 0.f;
    if (nd.player == trav) {
        const float d = (float)iter / ((float)iter + 1.0f);
        float sig[NT_DEC], cv[NT_DEC];
        regret_match(s->R, nd.act_base, na, Mp, h, s->frozen, n, sig);
        for (int a = 0; a < na; ++a) {
            cv[a] = s->cfv[(size_t)s->children[fc + a] * Mp + h];
            v += sig[a] * cv[a];
        }
        for (int a = 0; a < na; ++a) {
            size_t row = (size_t)(nd.act_base + a) * Mp + h;
            float r = s->R[row];
            r += cv[a] - v;
            r *= d;
            s->R[row] = r;
            s->SS[row] += (float)iter * sig[a];
        }
    } else {
        for (int a = 0; a < na; ++a) v += s->cfv[(size_t)s->children[fc + a] * Mp + h];
    }
    s->cfv[(size_t)n * Mp + h] = v;
}

__device__ __forceinline__ void backward_items(const SlotDesc *__restrict__ s, int level, int trav, int iter,
                                               int begin, int stride) {
    const int Mp = s->Mp;
    const int lo = s->dec_off[level];
    const int total = (s->dec_off[level + 1] - lo) * Mp;
    for (int it = begin; it < total; it += stride) {
        int ni = it / Mp, h = it - ni * Mp;
        backward_one(s, s->dec_lvl[lo + ni], h, trav, iter);
    }
}

// strategy at node 0, kept for the output's root_strategy (street_solve.cu:1563-1573)
__device__ __forceinline__ void root_sigma_items(const SlotDesc *__restrict__ s, int begin, int stride) {
    const DevNode &nd = s->nodes[0];
    if (nd.act_base < 0 || nd.player < 0 || nd.player >= s->NP) return;
    for (int h = begin; h < s->Mp; h += stride) {
        float sig[NT_DEC];
        if (h < s->nh[nd.player]) {
            regret_match(s->R, nd.act_base, nd.nchild, s->Mp, h, s->frozen, 0, sig);
            for (int a = 0; a < nd.nchild; ++a) s->rsig[(size_t)a * s->Mp + h] = sig[a];
        }
    }
}

// ---------------------------------------------------------------------------
// terminal values
//
// An opponent's reach at a node is scattered into a dense symmetric 52x52
// card-pair matrix M (row c, column d = reach of the hand {c,d}; board cards
// give zero rows), with row sums C[c] and total S = sum(C)/2.
// ---------------------------------------------------------------------------
#define DN 52
#define DN2 (52 * 52)

// element of a row addressed by 4 x its index (the byte tables store 4 x index)
template <typename T>
__device__ __forceinline__ T at4(const T *row, int i4) {
    return *(const T *)((const char *)row + (size_t)i4 * (sizeof(T) / 4));
}

template <typename T>
struct TermBuf {
    T *M1, *M2;          // [52*52] dense opponent reach
    T *PC1, *PC2, *PCP;  // [52*52] per-card prefix over strength-sorted partners; PC1 holds G for 2-opponent folds
    T *PG;               // [3][PGS] prefix over strength-sorted combos of M1, M2, M1.*M2
    T *C1, *C2, *CP;     // [52] per-card totals (CP of M1 .* M2)
    T *U1, *U2;          // [52] C1 * M2 and C2 * M1
    T *XD;               // [52] column sums of M1 .* M2 (2-opponent showdowns)
    T *red;              // [64] wave partials
    const uint8_t *clist;  // [52*52] per-card partner lists, ascending strength
    const uint8_t *clen;   // [52]
    const uint8_t *cpos;   // [52*52] 4 x position of partner d in card c's list (255: none)
};
#define PGS 1328   // stride of the PG channels

// A card-removal mass is a sum of non-negative reach products; evaluated by
// inclusion-exclusion it carries rounding noise of order eps * (scale of the
// terms). Values below this floor are set to exactly 0, as an empty sum is in
// the reference's enumeration (an info set with no live opponent hands must
// keep exactly zero regrets).
template <typename T> __device__ __forceinline__ T mass_floor();
template <> __device__ __forceinline__ float mass_floor<float>() { return 1e-5f; }
template <> __device__ __forceinline__ double mass_floor<double>() { return 1e-12; }

// Cross-lane moves by DPP (data-parallel primitives): a lane whose source is
// outside the pattern, or whose row is masked off, reads 0. CTRL 0x110 + n is
// row_shr:n (within 16-lane rows), 0x142 row_bcast:15 (lane 15 of each row
// into the next row), 0x143 row_bcast:31 (lane 31 into rows 2 and 3).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp0(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp0(double x) {
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float lane_value(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lane_value(double x, int l) {
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)v, l), hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// inclusive prefix sum across the 64 lanes of a wavefront (every lane
// active): prefix within each 16-lane row by shifts of 1, 2, 4, 8, then the
// row totals added into the rows above. Fixed order, so every form rounds
// alike.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x, int lane) {
    (void)lane;
    x += dpp0<0x111, 0xf>(x);
    x += dpp0<0x112, 0xf>(x);
    x += dpp0<0x114, 0xf>(x);
    x += dpp0<0x118, 0xf>(x);
    x += dpp0<0x142, 0xa>(x);
    x += dpp0<0x143, 0xc>(x);
    return x;
}

// sum across the 64 lanes of a wavefront (every lane active), the same value
// in every lane: the last lane of the inclusive scan
template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
    return lane_value(wave_incl_scan(x, 0), 63);
}

// sum of a 52-card table, lane c holding card c (every lane active): the same
// value in every wavefront, whatever the block size
template <typename T>
__device__ __forceinline__ T csum52(const T *C, int lane) {
    return wave_sum(lane < 52 ? C[lane] : (T)0);
}

// exclusive scans over the strength-sorted live combos: PG[0..ng] of M1 and,
// with NCH = 3, PG[PGS + ..] of M2 and PG[2 PGS + ..] of M1 .* M2.
// The decomposition is fixed (256 chunks, a wave64 scan of chunk sums, 4
// wave totals) whatever the block size, so every form rounds the same way.
#define SCAN_LANES 256
#define SCAN_CHUNK 5   // ceil(1081 / 256): a river has at most 1081 live combos
// this thread's entries of the strength order (scan_sorted's chunk), loaded
// ahead of the matrices they index
__device__ __forceinline__ void scan_load(const SlotDesc *__restrict__ s, uint32_t *gs, int tid) {
    const int ng = s->ng;
    const int chunk = (ng + SCAN_LANES - 1) / SCAN_LANES;
    const bool on = tid < SCAN_LANES;
    const int j0 = on ? min(ng, tid * chunk) : 0, j1 = on ? min(ng, j0 + chunk) : 0;
#pragma unroll
    for (int u = 0; u < SCAN_CHUNK; ++u) gs[u] = j0 + u < j1 ? gld(s->gsort + j0 + u) : 0u;
}
template <typename T, int NCH>
__device__ void scan_sorted(const SlotDesc *__restrict__ s, const T *M1, const T *M2, T *PG, T *red,
                            const uint32_t *gs, int tid) {
    const int ng = s->ng;
    const int chunk = (ng + SCAN_LANES - 1) / SCAN_LANES;
    const bool on = tid < SCAN_LANES;
    const int j0 = on ? min(ng, tid * chunk) : 0, j1 = on ? min(ng, j0 + chunk) : 0;
    int at[SCAN_CHUNK];
#pragma unroll
    for (int u = 0; u < SCAN_CHUNK; ++u) at[u] = (gs[u] & 0xFF) * DN + (gs[u] >> 8);
    T part[NCH];
#pragma unroll
    for (int q = 0; q < NCH; ++q) part[q] = 0;
#pragma unroll
    for (int u = 0; u < SCAN_CHUNK; ++u) {
        if (j0 + u >= j1) break;
        const T a = M1[at[u]];
        part[0] += a;
        if (NCH == 3) {
            const T b = M2[at[u]];
            part[1] += b;
            part[2] += a * b;
        }
    }
    const int lane = tid & 63, w = tid >> 6;
    T incl[NCH];
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
        incl[q] = wave_incl_scan(part[q], lane);
        if (on && lane == 63) red[q * 4 + w] = incl[q];
    }
    __syncthreads();
    if (on) {
        T run[NCH];
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            run[q] = incl[q] - part[q];
            for (int i = 0; i < w; ++i) run[q] += red[q * 4 + i];
        }
#pragma unroll
        for (int u = 0; u < SCAN_CHUNK; ++u) {
            if (j0 + u >= j1) break;
            const int j = j0 + u;
            const T a = M1[at[u]];
            PG[j] = run[0];
            run[0] += a;
            if (NCH == 3) {
                const T b = M2[at[u]];
                PG[PGS + j] = run[1];
                PG[2 * PGS + j] = run[2];
                run[1] += b;
                run[2] += a * b;
            }
        }
        if (tid == SCAN_LANES - 1) {
#pragma unroll
            for (int q = 0; q < NCH; ++q) PG[q * PGS + ng] = run[q];
        }
    }
    __syncthreads();
}

// reference odometer for 3+ opponents, one traverser hand (street_solve.cu
// :764-823 fold/leaf mass, :936-1002 showdown)
template <typename T>
__device__ T odometer(const SlotDesc *__restrict__ s, int n, int trav, int h, const int *opp, int k,
                      const T *reach, bool showdown, T pot, T bet) {
    const int Mp = s->Mp, N = s->N;
    uint32_t hi = s->hinfo[(size_t)trav * Mp + h];
    int x = hi & 63, y = (hi >> 6) & 63;
    uint32_t hs = showdown ? s->cstr[(hi >> 12) & 2047] : 0;
    int idx[6] = {0, 0, 0, 0, 0, 0};
    T acc = 0;
    for (;;) {
        uint64_t used = (1ull << x) | (1ull << y);
        T prod = 1;
        bool ok = true;
        uint32_t best = hs;
        int ties = 0;
        for (int i = 0; i < k && ok; ++i) {
            int q = opp[i];
            uint32_t oi = s->hinfo[(size_t)q * Mp + idx[i]];
            if (!(oi & SS_LIVE_BIT)) { ok = false; break; }
            uint64_t m = (1ull << (oi & 63)) | (1ull << ((oi >> 6) & 63));
            if (used & m) { ok = false; break; }
            used |= m;
            prod *= reach[((size_t)q * N + s->rrow[(size_t)q * N + n]) * Mp + idx[i]];
            if (showdown) {
                uint32_t os = s->cstr[(oi >> 12) & 2047];
                if (os > best) { best = os; ties = 0; }
                else if (os == best) ++ties;
            }
        }
        if (ok && prod > (T)0) {
            if (!showdown) acc += prod;
            else if (best > hs) acc += prod * (-bet);
            else acc += prod * (pot / (T)(ties + 1) - bet);
        }
        int c = k - 1;
        while (c >= 0) {
            if (++idx[c] < s->nh[opp[c]]) break;
            idx[c] = 0;
            --c;
        }
        if (c < 0) break;
    }
    return acc;
}

// Scatter opponent q's reach at node n into the dense card-pair matrix.
// Each combo owns two cells: with unique hands every cell gets one store
// (bitwise the same as an add onto the zeroed cell); duplicates add.
// Loads are issued four hands at a time, ahead of the adds.
// Opponent q's hands into its 52 x 52 matrix, in two steps so that the
// global loads can be issued before the matrix is cleared: scatter_load
// fetches the first SCN * nt hands (hand record, reach at row rrow),
// scatter_store writes them and any further hands.
template <typename T, int SCN>
__device__ __forceinline__ void scatter_load(const SlotDesc *__restrict__ s, int q, int rrow, const T *reach,
                                             uint32_t *oi, T *r, int tid, int nt) {
    const int Mp = s->Mp, nq = s->nh[q];
    const uint32_t *hq = s->hinfo + (size_t)q * Mp;
    const T *rq = reach + ((size_t)q * s->N + rrow) * Mp;
#pragma unroll
    for (int u = 0; u < SCN; ++u) {
        const int o = tid + u * nt;
        oi[u] = o < nq ? gld(hq + o) : 0u;
        r[u] = o < nq ? gld(rq + o) : (T)0;
    }
}
template <typename T>
__device__ __forceinline__ void scatter_one(const SlotDesc *__restrict__ s, T *M, uint32_t oi, T r) {
    if (!(oi & SS_LIVE_BIT)) return;
    const int a = oi & 63, b = (oi >> 6) & 63;
    if (s->uniq) {   // one writer per cell
        M[a * DN + b] = r;
        M[b * DN + a] = r;
    } else {
        atomicAdd(&M[a * DN + b], r);
        atomicAdd(&M[b * DN + a], r);
    }
}
template <typename T, int SCN>
__device__ __forceinline__ void scatter_store(const SlotDesc *__restrict__ s, int q, int rrow, const T *reach, T *M,
                                              const uint32_t *oi, const T *r, int tid, int nt) {
#pragma unroll
    for (int u = 0; u < SCN; ++u) scatter_one(s, M, oi[u], r[u]);
    const int Mp = s->Mp, nq = s->nh[q];
    const uint32_t *hq = s->hinfo + (size_t)q * Mp;
    const T *rq = reach + ((size_t)q * s->N + rrow) * Mp;
    for (int o = tid + SCN * nt; o < nq; o += nt) scatter_one(s, M, gld(hq + o), gld(rq + o));
}

// per-card totals of M1 (and M2, M1.*M2). The matrices are symmetric, so card
// c's total is column c's sum: lane c of the first wavefront sums its column
// down the rows in order (consecutive lanes read consecutive words).
template <typename T>
__device__ __forceinline__ void card_totals(const T *M1, const T *M2, T *C1, T *C2, T *CP, int tid, int nt) {
    (void)nt;
    if (tid >= 52) return;
    T a = 0, b = 0, p = 0;
#pragma unroll 4
    for (int d = 0; d < 52; ++d) {
        const T m1 = M1[d * DN + tid];
        a += m1;
        if (M2) {
            const T m2 = M2[d * DN + tid];
            b += m2;
            p += m1 * m2;
        }
    }
    C1[tid] = a;
    if (M2) {
        C2[tid] = b;
        CP[tid] = p;
    }
}

// per-card prefix sums of one matrix over strength-sorted partner lists
// (showdown, one opponent): PC1[c*52 + j] = sum of the first j partners of
// card c, C1 = list totals. One lane per card sums its list in order (51
// steps), the list read as 13 words at once (in LDS or global memory).
// Entries past a card's list length are cards too (the lists are
// zero-filled), so every read is in bounds; they add zero. With one table
// this beats a wavefront scan per card.
template <typename T>
__device__ __forceinline__ void card_prefix1(const uint8_t *clist, const uint8_t *clen, const T *M1, T *PC1, T *C1,
                                             int tid) {
    if (tid >= 52) return;
    const int c = tid;
    const uint32_t *cw = (const uint32_t *)(clist + c * DN);   // 52 * c bytes: word aligned
    uint32_t wv[13];
#pragma unroll
    for (int w = 0; w < 13; ++w) wv[w] = cw[w];
    const int len = clen[c];
    const T *A = M1 + c * DN;
    T *P = PC1 + c * DN;
    T run = 0;
    P[0] = run;
#pragma unroll
    for (int j = 0; j < 51; ++j) {
        const int d = (wv[j >> 2] >> (8 * (j & 3))) & 0xFF;
        run += j < len ? A[d] : (T)0;
        P[j + 1] = run;
    }
    C1[c] = run;
}

// per-card prefix sums over strength-sorted partner lists (showdown, two
// opponents): PC[c*52 + j] = sum of the first j partners of card c; C = list
// totals, for M1, M2 and M1.*M2.
// One wavefront per card (lane j = the card's j-th partner: gather, scan and
// write in one pass), two cards in flight per wavefront so that their LDS
// reads and scan steps overlap.
template <typename T>
__device__ __forceinline__ void card_prefix(const uint8_t *clist, const uint8_t *clen, const T *M1, const T *M2,
                                            T *PC1, T *PC2, T *PCP, T *C1, T *C2, T *CP, int tid, int nt) {
    const int lane = tid & 63, nw = nt >> 6, li = min(lane, 50);
    for (int c0 = tid >> 6; c0 < 52; c0 += 2 * nw) {
        const int c1 = c0 + nw < 52 ? c0 + nw : c0;   // uniform: the second card, or c0 again
        const int n0 = clen[c0], n1 = clen[c1];
        const int d0 = clist[c0 * DN + li], d1 = clist[c1 * DN + li];
        T a0 = lane < n0 ? M1[c0 * DN + d0] : (T)0, a1 = lane < n1 ? M1[c1 * DN + d1] : (T)0;
        T b0 = 0, b1 = 0, p0 = 0, p1 = 0;
        if (M2) {
            b0 = lane < n0 ? M2[c0 * DN + d0] : (T)0;
            b1 = lane < n1 ? M2[c1 * DN + d1] : (T)0;
            p0 = a0 * b0;
            p1 = a1 * b1;
        }
        a0 = wave_incl_scan(a0, lane);
        a1 = wave_incl_scan(a1, lane);
        if (M2) {
            b0 = wave_incl_scan(b0, lane);
            b1 = wave_incl_scan(b1, lane);
            p0 = wave_incl_scan(p0, lane);
            p1 = wave_incl_scan(p1, lane);
        }
        const T ta0 = lane_value(a0, 50), ta1 = lane_value(a1, 50);
        if (lane < 51) PC1[c0 * DN + 1 + lane] = a0;
        if (lane == 63) PC1[c0 * DN] = (T)0;
        if (lane == 0) C1[c0] = ta0;
        if (M2) {
            const T tb0 = lane_value(b0, 50), tp0 = lane_value(p0, 50);
            if (lane < 51) { PC2[c0 * DN + 1 + lane] = b0; PCP[c0 * DN + 1 + lane] = p0; }
            if (lane == 63) { PC2[c0 * DN] = (T)0; PCP[c0 * DN] = (T)0; }
            if (lane == 0) { C2[c0] = tb0; CP[c0] = tp0; }
        }
        if (c1 != c0) {
            if (lane < 51) PC1[c1 * DN + 1 + lane] = a1;
            if (lane == 63) PC1[c1 * DN] = (T)0;
            if (lane == 0) C1[c1] = ta1;
            if (M2) {
                const T tb1 = lane_value(b1, 50), tp1 = lane_value(p1, 50);
                if (lane < 51) { PC2[c1 * DN + 1 + lane] = b1; PCP[c1 * DN + 1 + lane] = p1; }
                if (lane == 63) { PC2[c1 * DN] = (T)0; PCP[c1 * DN] = (T)0; }
                if (lane == 0) { C2[c1] = tb1; CP[c1] = tp1; }
            }
        }
    }
}

// Two-opponent fold/leaf precompute: G[p][q] = sum_c M1[c][p] M2[c][q],
// U1 = C1 * M2, U2 = C2 * M1. Makes the per-hand 2-opponent mass O(1):
//   X(x,y) = T - C1[x]C2[x] - C1[y]C2[y] - U1[x] - U1[y] + C1[y]M2[y][x] + C1[x]M2[x][y]
//            - U2[x] - U2[y] + C2[y]M1[y][x] + C2[x]M1[x][y]
//            + CP[x] + CP[y] - 2 M1[x][y]M2[x][y] + G[x][y] + G[y][x]
template <typename T>
__device__ __forceinline__ void pair_products(const T *M1, const T *M2, const T *C1, const T *C2, T *G, T *U1,
                                              T *U2, int tid, int nt) {
    for (int i = tid; i < DN2; i += nt) {
        int p = i / DN, q = i - p * DN;
        T g = 0;
#pragma unroll 4
        for (int c = 0; c < 52; ++c) g += M1[c * DN + p] * M2[c * DN + q];
        G[i] = g;
    }
    for (int z = tid; z < 2 * DN; z += nt) {
        int col = z < DN ? z : z - DN;
        const T *Cv = z < DN ? C1 : C2;
        const T *Mv = z < DN ? M2 : M1;
        T u = 0;
        for (int c = 0; c < 52; ++c) u += Cv[c] * Mv[c * DN + col];
        if (z < DN) U1[col] = u; else U2[col] = u;
    }
}

// FP32 forms: G on the matrix cores, v_mfma_f32_16x16x4_f32 over 16x16 tiles
// of the 64x64-padded product (K = 52 cards in a fixed order). The padding
// carries the vector products too: row 52 of A = M1^T is C1 and column 52 of
// B = M2 is C2, so G'[52][q] = U1[q], G'[p][52] = U2[p], G'[52][52] = C1.C2.
// Lane layout: lane l holds A[l % 16][l / 16], B[l / 16][l % 16],
// D[4 (l / 16) + 0..3][l % 16].
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
__device__ __forceinline__ void pair_products_mfma(const float *M1, const float *M2, const float *C1,
                                                   const float *C2, float *G, float *U1, float *U2, float *TT,
                                                   int tid, int nt) {
    const int l = tid & 63, nw = nt >> 6;
    for (int P = tid >> 6; P < 4; P += nw) {
        f32x4 acc[4];
#pragma unroll
        for (int Q = 0; Q < 4; ++Q) acc[Q] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int p = 16 * P + (l & 15);
#pragma unroll
        for (int k0 = 0; k0 < 52; k0 += 4) {   // unrolled: the next steps' loads overlap these products
            const int c = k0 + (l >> 4);
            const float a = p < 52 ? M1[c * DN + p] : (p == 52 ? C1[c] : 0.f);
#pragma unroll
            for (int Q = 0; Q < 4; ++Q) {
                const int q = 16 * Q + (l & 15);
                const float b = q < 52 ? M2[c * DN + q] : (q == 52 ? C2[c] : 0.f);
                acc[Q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[Q], 0, 0, 0);
            }
        }
#pragma unroll
        for (int Q = 0; Q < 4; ++Q)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 16 * P + 4 * (l >> 4) + i, col = 16 * Q + (l & 15);
                const float g = acc[Q][i];
                if (row < 52 && col < 52) G[row * DN + col] = g;
                else if (row == 52 && col < 52) U1[col] = g;
                else if (col == 52 && row < 52) U2[row] = g;
                else if (row == 52 && col == 52) *TT = g;
            }
    }
}

// Kinds of terminal evaluation
// Two-opponent showdowns: for column x, U1[x] = sum_c C1[c] M2[c][x],
// U2[x] = sum_c C2[c] M1[c][x], XD[x] = sum_c M1[c][x] M2[c][x]. With
// TT = sum_c C1[c] C2[c] they turn the per-card product
//   sum_c (C1[c] - M1[c][x] - M1[c][y]) (C2[c] - M2[c][x] - M2[c][y])
// into TT - U1[x] - U1[y] - U2[x] - U2[y] + XD[x] + XD[y]
//    + sum_c (M1[c][x] M2[c][y] + M1[c][y] M2[c][x]),
// leaving only the last sum to the per-hand loop.
template <typename T>
__device__ __forceinline__ void showdown_pair_vectors(const T *M1, const T *M2, const T *C1, const T *C2, T *U1,
                                                      T *U2, T *XD, int x) {
    T u1 = 0, u2 = 0, dd = 0;
#pragma unroll 4
    for (int c = 0; c < 52; ++c) {
        const T m1 = M1[c * DN + x], m2 = M2[c * DN + x];
        u1 += C1[c] * m2;
        u2 += C2[c] * m1;
        dd += m1 * m2;
    }
    U1[x] = u1;
    U2[x] = u2;
    XD[x] = dd;
}

#define TK_NODE 0      // the node's own type (fold / showdown / leaf)
#define TK_MASS 1      // opponent mass only, all non-traversers, times `scale`

// Closed-form coverage of a terminal sweep: TM_LIGHT handles nodes with at
// most one opponent for the traverser, TM_PAIR at most two, TM_ALL any count
// (three or more by the reference's enumeration). Callers guarantee that no
// node beyond the mode's coverage reaches it.
#define TM_LIGHT 0
#define TM_PAIR 1
#define TM_ALL 2

// Writes out[h] for h < nh[trav] of node n for traverser trav. SDK = 1 / 0
// compiles only the showdown / non-showdown sweep (callers guarantee the node
// type), -1 both.
template <typename T, int MODE, int SDK = -1>
__device__ void terminal_node(const SlotDesc *__restrict__ s, int n, int trav, const T *reach, T *out,
                              TermBuf<T> B, int kind, const float *scale, int tid, int nt, int part = 0,
                              int parts = 1) {
    // the node's fields and the slot's shape, read together up front (not
    // one dependent round trip per field where each is first used)
    const DevNode *ndp = s->nodes + n;
    const int nd_type = gld(&ndp->type), nd_player = gld(&ndp->player), nd_pot = gld(&ndp->pot);
    const int nd_leaf = gld(&ndp->leaf_idx), nd_mask = gld(&ndp->active_mask), nd_bet = gld(&ndp->bets[trav]);
    const int Mp = s->Mp, NP = s->NP, nh = s->nh[trav], is_river = s->is_river, has_leaf = s->has_leaf;
    // every player's reach row at this node, read with the node's fields
    int rrq[SS_MAX_PLAYERS];
#pragma unroll
    for (int q = 0; q < SS_MAX_PLAYERS; ++q) rrq[q] = q < NP ? gld(s->rrow + (size_t)q * s->N + n) : 0;
    const int type = kind == TK_MASS ? SS_NODE_FOLD : nd_type;
    int opp[6], k = 0;
    for (int q = 0; q < NP; ++q) {
        if (q == trav) continue;
        if (type != SS_NODE_FOLD && !((nd_mask >> q) & 1)) continue;   // :866-873, :1037-1044
        opp[k++] = q;
    }
    const bool zero = (type == SS_NODE_SHOWDOWN && (!is_river || !((nd_mask >> trav) & 1))) ||
                      (type == SS_NODE_LEAF && (!has_leaf || nd_leaf < 0));
    if (zero) {
        for (int h = tid; h < nh; h += nt) out[h] = 0;
        return;
    }
    const T pot = (T)nd_pot, bet = (T)nd_bet;
    const T fold_pay = trav == nd_player ? (T)(nd_pot - nd_bet) : -(T)nd_bet;   // :726-733
    auto factor = [&](int h) -> T {
        if (kind == TK_MASS) return (T)scale[h];
        if (type == SS_NODE_FOLD) return fold_pay;
        return (T)s->leaf_values[((size_t)nd_leaf * NP + trav) * s->maxh + h];
    };
    if (MODE == TM_ALL && k >= 3) {
        for (int h = tid; h < nh; h += nt) {
            T v;
            if (type == SS_NODE_SHOWDOWN) {
                uint32_t hi = s->hinfo[(size_t)trav * Mp + h];
                v = (hi & SS_LIVE_BIT) && s->cstr[(hi >> 12) & 2047]
                        ? odometer<T>(s, n, trav, h, opp, k, reach, true, pot, bet) : (T)0;
            } else {
                v = odometer<T>(s, n, trav, h, opp, k, reach, false, pot, bet) * factor(h);
            }
            out[h] = v;
        }
        return;
    }
    // ---- closed forms, k <= 2 ----
    const bool sd = SDK < 0 ? type == SS_NODE_SHOWDOWN : SDK == 1;
    const bool two = MODE != TM_LIGHT && k == 2;
    // the opponents' hands and reach (and a showdown's strength-order
    // entries) are fetched first: their latency overlaps the clearing below.
    // 512-thread workgroups (two-opponent showdowns) need three hands per
    // thread for a street's at most 1128, others five.
    constexpr int SCN = (MODE == TM_PAIR && SDK == 1) ? 3 : 5;
    uint32_t oi1[SCN], oi2[SCN], gs[SCAN_CHUNK];
    T r1[SCN], r2[SCN];
    int rr1 = 0, rr2 = 0;
#pragma unroll
    for (int q = 0; q < SS_MAX_PLAYERS; ++q) {   // static indices: rrq stays in registers
        rr1 = k >= 1 && q == opp[0] ? rrq[q] : rr1;
        rr2 = k >= 2 && q == opp[1] ? rrq[q] : rr2;
    }
    if (k >= 1) scatter_load<T, SCN>(s, opp[0], rr1, reach, oi1, r1, tid, nt);
    if (two) scatter_load<T, SCN>(s, opp[1], rr2, reach, oi2, r2, tid, nt);
    if (sd && k >= 1) scan_load(s, gs, tid);
    // Folds and leaves walk the traverser's hands in order: this thread's
    // first HPT hand records are loaded together now, so their latency
    // overlaps the table setup instead of one round trip per hand.
    constexpr int HPT = 5;
    uint32_t hv[HPT];
#pragma unroll
    for (int u = 0; u < HPT; ++u) {
        const int i = tid + u * nt;
        hv[u] = !sd && i < nh ? gld(s->hinfo + (size_t)trav * Mp + i) : 0u;
    }
    {
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 *m1 = (float4 *)B.M1, *m2 = (float4 *)B.M2;
        for (int i = tid; i < DN2 * (int)sizeof(T) / 16; i += nt) {
            m1[i] = z4;
            if (two) m2[i] = z4;
        }
    }
    __syncthreads();
    if (k >= 1) scatter_store<T, SCN>(s, opp[0], rr1, reach, B.M1, oi1, r1, tid, nt);
    if (two) scatter_store<T, SCN>(s, opp[1], rr2, reach, B.M2, oi2, r2, tid, nt);
    __syncthreads();
    if (sd && two)
        card_prefix<T>(B.clist, B.clen, B.M1, B.M2, B.PC1, B.PC2, B.PCP, B.C1, B.C2, B.CP, tid, nt);
    else if (sd && k >= 1)
        card_prefix1<T>(B.clist, B.clen, B.M1, B.PC1, B.C1, tid);
    else if (k >= 1)
        card_totals<T>(B.M1, two ? B.M2 : nullptr, B.C1, B.C2, B.CP, tid, nt);
    __syncthreads();
    if (sd && k == 1) scan_sorted<T, 1>(s, B.M1, nullptr, B.PG, B.red, gs, tid);
    if (sd && two) {
        // column vectors of the opponent-pair correction, by threads the
        // strength-order scan leaves idle when there are enough
        const bool spare = nt >= SCAN_LANES + 64;
        if (spare && tid >= SCAN_LANES && tid < SCAN_LANES + 52)
            showdown_pair_vectors<T>(B.M1, B.M2, B.C1, B.C2, B.U1, B.U2, B.XD, tid - SCAN_LANES);
        scan_sorted<T, 3>(s, B.M1, B.M2, B.PG, B.red, gs, tid);
        if (!spare) {
            if (tid < 52) showdown_pair_vectors<T>(B.M1, B.M2, B.C1, B.C2, B.U1, B.U2, B.XD, tid);
            __syncthreads();
        }
    }
    if (two && !sd) {
        if constexpr (std::is_same<T, float>::value)
            pair_products_mfma(B.M1, B.M2, B.C1, B.C2, B.PC1, B.U1, B.U2, B.red, tid, nt);
        else
            pair_products<T>(B.M1, B.M2, B.C1, B.C2, B.PC1, B.U1, B.U2, tid, nt);
        __syncthreads();
    }

    T S1 = 0, S2 = 0, SP = 0, TT = 0;
    const int lane = tid & 63;
    if (k >= 1) S1 = (T)0.5 * csum52(B.C1, lane);
    if (two) {
        S2 = (T)0.5 * csum52(B.C2, lane);
        SP = (T)0.5 * csum52(B.CP, lane);
        if (!sd && std::is_same<T, float>::value)
            TT = B.red[0];   // from the MFMA padding
        else
            TT = wave_sum(lane < 52 ? B.C1[lane] * B.C2[lane] : (T)0);
    }

    // Showdowns walk the traverser's hands by strength (shared class
    // boundaries within a wavefront turn prefix reads into broadcasts), one
    // 16-byte record per hand, fetched one hand ahead.
    // A showdown may be split over `parts` workgroups, each taking a range of
    // the strength order (each repeats the setup above in its own LDS).
    const int hlo = sd ? (int)((int64_t)nh * part / parts) : 0;
    const int hhi = sd ? (int)((int64_t)nh * (part + 1) / parts) : nh;
    const uint4 *meta = sd ? s->hmeta + (size_t)trav * Mp : nullptr;
    uint4 mnext = make_uint4(0u, 0u, 0u, 0u);
    if (sd && hlo + tid < hhi) mnext = gld(meta + hlo + tid);
    for (int i = hlo + tid, u = 0; i < hhi; i += nt, ++u) {
        const uint4 mh = mnext;
        if (sd && i + nt < hhi) mnext = gld(meta + i + nt);
        int h, x, y;
        if (sd) {
            h = mh.x & 2047;
            x = (mh.x >> 11) & 63;
            y = (mh.x >> 17) & 63;
        } else {
            h = i;
            uint32_t hi = hv[0];   // u is the same in every lane: a select chain, no indexing
#pragma unroll
            for (int q = 1; q < HPT; ++q) hi = u == q ? hv[q] : hi;
            if (u >= HPT) hi = s->hinfo[(size_t)trav * Mp + h];
            x = hi & 63;
            y = (hi >> 6) & 63;
        }
        const int xy = x * DN + y;
        T v;
        if (!sd) {
            T m;
            if (k == 0) m = 1;
            else {
                T z1 = S1 - B.C1[x] - B.C1[y] + B.M1[xy];
                if (!two) m = z1 < mass_floor<T>() * S1 ? (T)0 : z1;
                else {
                    T m1 = B.M1[xy], m2 = B.M2[xy];
                    T z2 = S2 - B.C2[x] - B.C2[y] + m2;
                    T zp = SP - B.CP[x] - B.CP[y] + m1 * m2;
                    T X = TT - B.C1[x] * B.C2[x] - B.C1[y] * B.C2[y] - B.U1[x] - B.U1[y] + B.C1[y] * m2 +
                          B.C1[x] * m2 - B.U2[x] - B.U2[y] + B.C2[y] * m1 + B.C2[x] * m1 + B.CP[x] + B.CP[y] -
                          (T)2 * m1 * m2 + B.PC1[xy] + B.PC1[y * DN + x];
                    m = z1 * z2 - X + zp;
                    if (m < mass_floor<T>() * S1 * S2) m = 0;
                }
            }
            v = factor(h) * m;
        } else {
            if (!((mh.x >> 23) & 1)) { out[h] = 0; continue; }              // hs = 0, :861
            const int cls = mh.y, jl = mh.z, je = mh.w;
            if (k == 0) {
                v = pot - bet;
            } else if (!two) {                                                 // :886-903
                const uint8_t *rl = s->bnd + (size_t)cls * DN, *re = rl + DN;
                int lx = gld(rl + x) >> 2, ly = gld(rl + y) >> 2, ex = gld(re + x) >> 2, ey = gld(re + y) >> 2;
                T lt = B.PG[jl] - B.PC1[x * DN + lx] - B.PC1[y * DN + ly];
                T le = B.PG[je] - B.PC1[x * DN + ex] - B.PC1[y * DN + ey] + B.M1[xy];
                T z = S1 - B.C1[x] - B.C1[y] + B.M1[xy];
                v = z < mass_floor<T>() * S1 ? (T)0 : pot * lt + pot * (T)0.5 * (le - lt) - bet * z;
            } else {                                                           // :936-1004, n_opp = 2
                // Classes relative to the traverser's strength: "l" (weaker) and
                // "le" (weaker or tied). With A = opponent-1 mass, B = opponent 2,
                // the value is pot * (wll + (wle + wel) / 2 + wee / 3) - bet * z,
                // and wll + (wle + wel)/2 + wee/3 = [Al(2Bl + Ble) + Ale(Bl + 2Ble)] / 6
                // for each bilinear part (pairs sharing a card are removed per card c).
                // Class boundaries of every card: two contiguous 52-byte rows
                // (row kcls: entries ranked <, row kcls+1: ranked <=; stored x4),
                // read a word (4 cards) ahead of use.
                const uint32_t *rowl = (const uint32_t *)(s->bnd + (size_t)cls * DN);
                const uint8_t *rl = (const uint8_t *)rowl;
                const int lx4 = gld(rl + x), ly4 = gld(rl + y), ex4 = gld(rl + DN + x), ey4 = gld(rl + DN + y);
                uint32_t lnext = gld(rowl), enext = gld(rowl + 13);
                // terms of card c: opponent pairs sharing c, with the hands
                // {c,x}, {c,y} (not opponent hands) removed from the classes
                // they fall in, by their position in c's sorted list
                auto card = [&](int c, int l4, int e4, T &xw, T &xaa, bool full) {
                    const T P1l = at4(B.PC1 + c * DN, l4), P2l = at4(B.PC2 + c * DN, l4);
                    const T P1e = at4(B.PC1 + c * DN, e4), P2e = at4(B.PC2 + c * DN, e4);
                    const int px4 = B.cpos[c * DN + x], py4 = B.cpos[c * DN + y];
                    const T m1x = B.M1[c * DN + x], m1y = B.M1[c * DN + y];
                    const T m2x = B.M2[c * DN + x], m2y = B.M2[c * DN + y];
                    const T fxl = px4 < l4 ? (T)1 : (T)0, fxe = px4 < e4 ? (T)1 : (T)0;
                    const T fyl = py4 < l4 ? (T)1 : (T)0, fye = py4 < e4 ? (T)1 : (T)0;
                    const T Al = P1l - fxl * m1x - fyl * m1y;
                    const T Ale = P1e - fxe * m1x - fye * m1y;
                    const T Bl = P2l - fxl * m2x - fyl * m2y;
                    const T Ble = P2e - fxe * m2x - fye * m2y;
                    xw += Al * ((T)2 * Bl + Ble);
                    xw += Ale * (Bl + (T)2 * Ble);
                    if (full) xaa += (B.C1[c] - m1x - m1y) * (B.C2[c] - m2x - m2y);
                    else {
                        xaa = fma(m1x, m2y, xaa);
                        xaa = fma(m1y, m2x, xaa);
                    }
                };
                // every card that can be in a hand (the tables number the cards
                // off the board first: 0..46 on a river, and card 47, a board
                // card, has empty lists: exact zeros), then the traverser's own
                // two cards, which carry no opponent pairs, taken back out
                T xw = 0, xaa = 0;
#pragma unroll 1
                for (int g = 0; g < 12; ++g) {   // cards 0..47: every card off a river board
                    const uint32_t lw = lnext, ew = enext;
                    if (g < 11) {
                        lnext = gld(rowl + g + 1);
                        enext = gld(rowl + 14 + g);
                    }
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        card(4 * g + b, (lw >> (8 * b)) & 0xFF, (ew >> (8 * b)) & 0xFF, xw, xaa, false);
                }
                {
                    T ow = 0, oa = 0;
                    card(x, lx4, ex4, ow, oa, true);
                    card(y, ly4, ey4, ow, oa, true);
                    xw -= ow;
                    xaa += TT - B.U1[x] - B.U1[y] - B.U2[x] - B.U2[y] + B.XD[x] + B.XD[y] - oa;
                }
                const int lx = lx4 >> 2, ly = ly4 >> 2, ex = ex4 >> 2, ey = ey4 >> 2;
                const T m1k = B.M1[xy], m2k = B.M2[xy];
                // class totals over opponent hands disjoint from {x,y}; the
                // traverser's own combo is tied with itself (class "le" only)
                const T Al = B.PG[jl] - B.PC1[x * DN + lx] - B.PC1[y * DN + ly];
                const T Ale = B.PG[je] - B.PC1[x * DN + ex] - B.PC1[y * DN + ey] + m1k;
                const T Bl = B.PG[PGS + jl] - B.PC2[x * DN + lx] - B.PC2[y * DN + ly];
                const T Ble = B.PG[PGS + je] - B.PC2[x * DN + ex] - B.PC2[y * DN + ey] + m2k;
                const T Pl = B.PG[2 * PGS + jl] - B.PCP[x * DN + lx] - B.PCP[y * DN + ly];
                const T Ple = B.PG[2 * PGS + je] - B.PCP[x * DN + ex] - B.PCP[y * DN + ey] + m1k * m2k;
                const T Aa = S1 - B.C1[x] - B.C1[y] + m1k, Ba = S2 - B.C2[x] - B.C2[y] + m2k;
                const T Pa = SP - B.CP[x] - B.CP[y] + m1k * m2k;
                const T W = (Al * ((T)2 * Bl + Ble) + Ale * (Bl + (T)2 * Ble) - xw) * ((T)1 / (T)6) +
                            Pl * ((T)2 / (T)3) + Ple * ((T)1 / (T)3);
                const T z = Aa * Ba - xaa + Pa;
                v = z < mass_floor<T>() * S1 * S2 ? (T)0 : pot * W - bet * z;
            }
        }
        out[h] = v;
    }
    __syncthreads();
}

// LDS of the FP32 terminal sweeps. Heavy (2-opponent nodes, ~76 KB, gfx950
// allows 160 KB per workgroup): both opponents' matrices, three prefix
// tables, the subgame's per-card lists. Light (<= 1 opponent, ~26 KB): one
// matrix, one prefix table; six workgroups fit a CU.
// Rows of 2752 = 43 x 64 floats: the same cell of M1 and M2 is one
// ds_read2st64 away. PC2 sits one float past that spacing from PC1, so the
// two per-card prefix reads of the showdown loop stay single reads sharing
// one address register, each card's row offset in the 16-bit immediate (a
// paired read2st64 cannot encode a 208-byte row step and costs an add per
// card). The byte tables and short arrays come first so their addresses fit
// the 16-bit immediate offsets of LDS instructions.
#define DNP 2752
template <typename T>
struct TermHeavyT {
    alignas(16) uint8_t clist[DN2];
    alignas(16) uint8_t cpos[DN2];
    alignas(16) uint8_t clen[64];
    T C1[52], C2[52], CP[52], U1[52], U2[52], XD[52];
    T red[64];
    T PG[3 * PGS];
    alignas(256) T M1[DNP];
    T M2[DNP];
    T PC1[DNP + 1], PC2[DNP], PCP[DN2];
};
using TermHeavy = TermHeavyT<float>;   // 76 KB; the FP64 one (150 KB) serves exploitability checks

template <typename T>
__device__ __forceinline__ TermBuf<T> heavy_buf(TermHeavyT<T> &t) {
    TermBuf<T> b;
    b.M1 = t.M1; b.M2 = t.M2; b.PC1 = t.PC1; b.PC2 = t.PC2; b.PCP = t.PCP;
    b.PG = t.PG; b.C1 = t.C1; b.C2 = t.C2; b.CP = t.CP; b.U1 = t.U1; b.U2 = t.U2; b.XD = t.XD; b.red = t.red;
    b.clist = t.clist; b.clen = t.clen; b.cpos = t.cpos;
    return b;
}

// Two-opponent fold and leaf nodes (~33 KB in FP32): both matrices and G = M1^T M2.
template <typename T>
struct TermFoldT {
    T M1[DNP], M2[DNP], G[DNP];
    T C1[52], C2[52], CP[52], U1[52], U2[52];
    T red[64];
};
using TermFold = TermFoldT<float>;

template <typename T>
__device__ __forceinline__ TermBuf<T> fold_buf(TermFoldT<T> &t) {
    TermBuf<T> b;
    b.M1 = t.M1; b.M2 = t.M2; b.PC1 = t.G; b.PC2 = nullptr; b.PCP = nullptr;
    b.PG = nullptr; b.C1 = t.C1; b.C2 = t.C2; b.CP = t.CP; b.U1 = t.U1; b.U2 = t.U2; b.XD = nullptr; b.red = t.red;
    b.clist = nullptr; b.clen = nullptr; b.cpos = nullptr;
    return b;
}

template <typename T>
struct TermLightT {
    T M1[DN2];
    T PC1[DN2];
    T PG[1088];   // one channel: a river's at most 1081 live combos + the total (26.4 KB: six per CU)
    T C1[52];
    T red[64];
};
using TermLight = TermLightT<float>;

// copy a river subgame's per-card lists and positions into LDS (caller barriers)
template <typename T>
__device__ __forceinline__ void stage_tables(const SlotDesc *__restrict__ s, TermHeavyT<T> &t, int tid, int nt) {
    if (!s->is_river) return;
    const uint32_t *cl = (const uint32_t *)s->clist, *cp = (const uint32_t *)s->cpos;
    uint32_t *tl = (uint32_t *)t.clist, *tp = (uint32_t *)t.cpos;
    for (int i = tid; i < DN2 / 4; i += nt) {
        tl[i] = gld(cl + i);
        tp[i] = gld(cp + i);
    }
    for (int i = tid; i < 52; i += nt) t.clen[i] = s->clen[i];
}

template <typename T>
__device__ __forceinline__ TermBuf<T> light_buf(TermLightT<T> &t, const SlotDesc *__restrict__ s) {
    TermBuf<T> b;
    b.M1 = t.M1; b.M2 = nullptr; b.PC1 = t.PC1; b.PC2 = nullptr; b.PCP = nullptr;
    b.PG = t.PG; b.C1 = t.C1; b.C2 = nullptr; b.CP = nullptr; b.U1 = nullptr; b.U2 = nullptr; b.XD = nullptr;
    b.red = t.red;
    b.clist = s->clist; b.clen = s->clen; b.cpos = s->cpos;
    return b;
}

// FP64 buffers live in global scratch: one slice per workgroup
#define XSCRATCH_DOUBLES (5 * DN2 + 3 * PGS + 6 * 52 + 64)
__device__ __forceinline__ TermBuf<double> global_buf(double *base, const SlotDesc *__restrict__ s) {
    TermBuf<double> b;
    b.M1 = base; b.M2 = b.M1 + DN2;
    b.PC1 = b.M2 + DN2; b.PC2 = b.PC1 + DN2; b.PCP = b.PC2 + DN2;
    b.PG = b.PCP + DN2;
    b.C1 = b.PG + 3 * PGS; b.C2 = b.C1 + 52; b.CP = b.C2 + 52; b.U1 = b.CP + 52; b.U2 = b.U1 + 52;
    b.red = b.U2 + 52;
    b.XD = b.red + 64;
    b.clist = s->clist;
    b.clen = s->clen;
    b.cpos = s->cpos;
    return b;
}

// ---------------------------------------------------------------------------
// plain form: one subgame, level-synchronous launches
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(SS_BLOCK) k_forward(const SlotDesc *__restrict__ s, int level, int trav) {
    forward_items(s, level, trav, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// skip_sd: showdown values come from the precision form's contraction
__global__ void __launch_bounds__(SS_BLOCK) k_terminals(const SlotDesc *__restrict__ s, int trav, int skip_sd) {
    __shared__ TermHeavy sh;
    TermBuf<float> b = heavy_buf(sh);
    stage_tables(s, sh, threadIdx.x, blockDim.x);
    __syncthreads();
    for (int j = blockIdx.x; j < s->T; j += gridDim.x) {
        int n = s->term[j];
        if (skip_sd && s->nodes[n].type == SS_NODE_SHOWDOWN) continue;
        terminal_node<float, TM_ALL>(s, n, trav, s->reach, s->cfv + (size_t)n * s->Mp, b, TK_NODE, nullptr,
                             threadIdx.x, blockDim.x);
    }
}

__global__ void __launch_bounds__(SS_BLOCK) k_backward(const SlotDesc *__restrict__ s, int level, int trav) {
    int iter = s->iter + 1;
    backward_items(s, level, trav, iter, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__global__ void __launch_bounds__(SS_BLOCK) k_root_sigma(const SlotDesc *__restrict__ s) {
    root_sigma_items(s, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__global__ void k_bump(SlotDesc *s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) s->iter += 1;
}

// ---------------------------------------------------------------------------
// batched form: one subgame per workgroup, `iters` iterations per launch
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(SS_BATCH_BLOCK) k_batched(SlotDesc *__restrict__ slots, int nslots, int iters) {
    __shared__ TermHeavy sh;
    TermBuf<float> b = heavy_buf(sh);
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int si = blockIdx.x; si < nslots; si += gridDim.x) {
        SlotDesc *s = slots + si;
        if (!s->active) continue;
        __syncthreads();
        stage_tables(s, sh, tid, nt);
        __syncthreads();
        int it = s->iter;
        const int end = s->iter_end, NP = s->NP, L = s->L;
        for (int rep = 0; rep < iters && it < end; ++rep) {
            ++it;
            for (int trav = 0; trav < NP; ++trav) {
                if (trav == NP - 1 && (it == s->final_iter || (s->snap_every > 0 && it % s->snap_every == 0))) {
                    root_sigma_items(s, tid, nt);
                    __syncthreads();
                }
                for (int l = 0; l < L; ++l) {
                    forward_items(s, l, trav, tid, nt);
                    __syncthreads();
                }
                for (int j = 0; j < s->T; ++j) {
                    int n = s->term[j];
                    terminal_node<float, TM_ALL>(s, n, trav, s->reach, s->cfv + (size_t)n * s->Mp, b, TK_NODE, nullptr,
                                         tid, nt);
                    __syncthreads();
                }
                for (int l = L - 1; l >= 0; --l) {
                    backward_items(s, l, trav, it, tid, nt);
                    __syncthreads();
                }
            }
            if (tid == 0) s->iter = it;
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 best response / exploitability
// ---------------------------------------------------------------------------

// average strategy at decision node n, hand h, normalised as the host does
// in street_solve.cu:1597-1610
__device__ __forceinline__ void avg_one(const SlotDesc *__restrict__ s, int n, int h) {
    const int Mp = s->Mp;
    const DevNode &nd = s->nodes[n];
    int na = nd.nchild;
    float sum = 0.f;
    for (int a = 0; a < na; ++a) {
        float v = s->SS[(size_t)(nd.act_base + a) * Mp + h];
        sum += v > 0.f ? v : 0.f;
    }
    float inv = sum > 0.f ? 1.0f / sum : 1.0f / (float)na;
    for (int a = 0; a < na; ++a) {
        float v = s->SS[(size_t)(nd.act_base + a) * Mp + h];
        v = v > 0.f ? v : 0.f;
        s->avg[(size_t)(nd.act_base + a) * Mp + h] = sum > 0.f ? v * inv : inv;
    }
}

__device__ __forceinline__ void avg_items(const SlotDesc *__restrict__ s, int begin, int stride) {
    const int Mp = s->Mp;
    const int total = s->D * Mp;
    for (int it = begin; it < total; it += stride) {
        int di = it / Mp, h = it - di * Mp;
        avg_one(s, s->dec_lvl[di], h);
    }
}

// reach of every player under the average strategy through node n, hand h
// reach under the average strategy; as in the CFR forward sweep only the
// actor's child rows are written (the others are read at the row of their
// last action through rrow)
__device__ __forceinline__ void xforward_one(const SlotDesc *__restrict__ s, int n, int h) {
    const int Mp = s->Mp, N = s->N, NP = s->NP;
    const DevNode &nd = s->nodes[n];
    const int q = nd.player;
    if (q < 0 || q >= NP || h >= s->nh[q]) return;
    double *rr = s->xreach + (size_t)q * N * Mp;
    const double pr = rr[(size_t)s->rrow[(size_t)q * N + n] * Mp + h];
    for (int a = 0; a < nd.nchild; ++a) {
        int c = s->children[nd.first_child + a];
        rr[(size_t)c * Mp + h] = pr * (double)s->avg[(size_t)(nd.act_base + a) * Mp + h];
    }
}

__device__ __forceinline__ void xforward_items(const SlotDesc *__restrict__ s, int level, int begin, int stride) {
    const int Mp = s->Mp;
    const int lo = s->dec_off[level];
    const int total = (s->dec_off[level + 1] - lo) * Mp;
    for (int it = begin; it < total; it += stride) {
        int ni = it / Mp, h = it - ni * Mp;
        xforward_one(s, s->dec_lvl[lo + ni], h);
    }
}

// best-response (max) and profile-value (avg-weighted) backup for player p
__device__ __forceinline__ void xbackward_one(const SlotDesc *__restrict__ s, int n, int h, int p) {
    const int Mp = s->Mp;
    if (h >= s->nh[p]) return;
    const DevNode &nd = s->nodes[n];
    double e = 0, b = 0;
    for (int a = 0; a < nd.nchild; ++a) {
        int c = s->children[nd.first_child + a];
        double ce = s->xev[(size_t)c * Mp + h], cb = s->xbr[(size_t)c * Mp + h];
        if (nd.player == p) {
            e += (double)s->avg[(size_t)(nd.act_base + a) * Mp + h] * ce;
            b = (a == 0 || cb > b) ? cb : b;
        } else {
            e += ce;
            b += cb;
        }
    }
    s->xev[(size_t)n * Mp + h] = e;
    s->xbr[(size_t)n * Mp + h] = b;
}

__device__ __forceinline__ void xbackward_items(const SlotDesc *__restrict__ s, int level, int p, int begin,
                                                int stride) {
    const int Mp = s->Mp;
    const int lo = s->dec_off[level];
    const int total = (s->dec_off[level + 1] - lo) * Mp;
    for (int it = begin; it < total; it += stride) {
        int ni = it / Mp, h = it - ni * Mp;
        xbackward_one(s, s->dec_lvl[lo + ni], h, p);
    }
}

// sum_h w_p(h) * (br - ev) and sum_h w_p(h) * ev at the root
__device__ void xreduce_block(const SlotDesc *__restrict__ s, int p, double *red, int tid, int nt) {
    const int Mp = s->Mp;
    double g = 0, v = 0;
    for (int h = tid; h < s->nh[p]; h += nt) {
        double w = (double)s->weights[(size_t)p * Mp + h];
        g += w * (s->xbr[h] - s->xev[h]);
        v += w * s->xev[h];
    }
    red[tid] = g;
    red[nt + tid] = v;
    __syncthreads();
    for (int off = nt / 2; off > 0; off >>= 1) {
        if (tid < off) {
            red[tid] += red[tid + off];
            red[nt + tid] += red[nt + tid + off];
        }
        __syncthreads();
    }
    if (tid == 0) {
        s->xout[p] = red[0];
        s->xout[s->NP + p] = red[nt];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(SS_BLOCK) k_avg(const SlotDesc *__restrict__ s) {
    avg_items(s, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__global__ void __launch_bounds__(SS_BLOCK) k_xforward(const SlotDesc *__restrict__ s, int level) {
    xforward_items(s, level, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__global__ void __launch_bounds__(SS_BLOCK) k_xterminals(const SlotDesc *__restrict__ s, int p, double *scratch) {
    TermBuf<double> b = global_buf(scratch + (size_t)blockIdx.x * XSCRATCH_DOUBLES, s);
    for (int j = blockIdx.x; j < s->T; j += gridDim.x) {
        int n = s->term[j];
        terminal_node<double, TM_ALL>(s, n, p, s->xreach, s->xev + (size_t)n * s->Mp, b, TK_NODE, nullptr, threadIdx.x,
                              blockDim.x);
        for (int h = threadIdx.x; h < s->nh[p]; h += blockDim.x)
            s->xbr[(size_t)n * s->Mp + h] = s->xev[(size_t)n * s->Mp + h];
    }
}

__global__ void __launch_bounds__(SS_BLOCK) k_xbackward(const SlotDesc *__restrict__ s, int level, int p) {
    xbackward_items(s, level, p, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__global__ void __launch_bounds__(SS_BLOCK) k_xreduce(const SlotDesc *__restrict__ s, int p) {
    __shared__ double red[2 * SS_BLOCK];
    xreduce_block(s, p, red, threadIdx.x, blockDim.x);
}

// joint deal mass at the root with raw weights: sum_h w_0(h) * mass_0(h)
__global__ void __launch_bounds__(SS_BLOCK) k_xjoint(const SlotDesc *__restrict__ s, double *scratch) {
    __shared__ double red[SS_BLOCK];
    TermBuf<double> b = global_buf(scratch, s);
    // xreach rows of node 0 hold the raw weights; the per-hand mass lands in xev row 0
    // with scale = weights of player 0 restricted to live hands
    terminal_node<double, TM_ALL>(s, 0, 0, s->xreach, s->xev, b, TK_MASS, s->weights, threadIdx.x, blockDim.x);
    double acc = 0;
    for (int h = threadIdx.x; h < s->nh[0]; h += blockDim.x)
        if (s->hinfo[h] & SS_LIVE_BIT) acc += s->xev[h];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) s->xout[2 * s->NP] = red[0];
}

// whole exploitability evaluation of one slot by one workgroup (batched checks)
__global__ void __launch_bounds__(SS_BLOCK) k_batched_expl(SlotDesc *__restrict__ slots, int nslots,
                                                           double *scratch, const int *want) {
    __shared__ double red[2 * SS_BLOCK];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int si = blockIdx.x; si < nslots; si += gridDim.x) {
        SlotDesc *s = slots + si;
        if (!s->active || !want[si]) continue;
        TermBuf<double> b = global_buf(scratch + (size_t)blockIdx.x * XSCRATCH_DOUBLES, s);
        avg_items(s, tid, nt);
        __syncthreads();
        for (int l = 0; l < s->L; ++l) {
            xforward_items(s, l, tid, nt);
            __syncthreads();
        }
        for (int p = 0; p < s->NP; ++p) {
            for (int j = 0; j < s->T; ++j) {
                int n = s->term[j];
                terminal_node<double, TM_ALL>(s, n, p, s->xreach, s->xev + (size_t)n * s->Mp, b, TK_NODE, nullptr, tid, nt);
                for (int h = tid; h < s->nh[p]; h += nt) s->xbr[(size_t)n * s->Mp + h] = s->xev[(size_t)n * s->Mp + h];
                __syncthreads();
            }
            for (int l = s->L - 1; l >= 0; --l) {
                xbackward_items(s, l, p, tid, nt);
                __syncthreads();
            }
            xreduce_block(s, p, red, tid, nt);
        }
        terminal_node<double, TM_ALL>(s, 0, 0, s->xreach, s->xev, b, TK_MASS, s->weights, tid, nt);
        double acc = 0;
        for (int h = tid; h < s->nh[0]; h += nt)
            if (s->hinfo[h] & SS_LIVE_BIT) acc += s->xev[h];
        red[tid] = acc;
        __syncthreads();
        for (int off = nt / 2; off > 0; off >>= 1) {
            if (tid < off) red[tid] += red[tid + off];
            __syncthreads();
        }
        if (tid == 0) s->xout[2 * s->NP] = red[0];
        __syncthreads();
    }
}

}  // namespace ss

namespace ss {

// ---------------------------------------------------------------------------
// batched form, phased: every launch covers one phase of one traverser pass
// for ALL resident subgames. Tasks are (slot << 20 | node); one workgroup per
// task, so a subgame's nodes spread over every CU. Slots that are retired,
// at their iteration cap, or have fewer players than the pass index skip.
// ---------------------------------------------------------------------------
#define TASK_SLOT(t) ((int)((t) >> 20))
#define TASK_NODE(t) ((int)((t) & 0xFFFFF))

__device__ __forceinline__ bool slot_runs(const SlotDesc *__restrict__ s, int trav) {
    return s->active && s->iter < s->iter_end && trav < s->NP;
}

// Sweeps: one workgroup of SWEEP_BLOCK threads per (subgame, node) task, so a
// 1081-combo node (Mp = 1088) takes two nearly full rounds of hands.
#define SWEEP_BLOCK 576

// regret_match over preloaded regret rows r (same arithmetic, same order);
// A bounds the action count (na <= A)
template <int A>
__device__ __forceinline__ void regret_match_vals(const float *r, int na, int fa, float *sig) {
    float sum = 0.f;
#pragma unroll
    for (int a = 0; a < A; ++a) {
        if (a < na) {
            sig[a] = r[a] > 0.f ? r[a] : 0.f;
            sum += sig[a];
        }
    }
    if (sum > 0.f) {
        float inv = 1.0f / sum;
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (a < na) sig[a] *= inv;
    } else {
        float u = 1.0f / (float)na;
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (a < na) sig[a] = u;
    }
    if (fa >= 0) {
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (a < na) sig[a] = a == fa ? 1.f : 0.f;
    }
}

// Pins loaded values: the compiler may not move the code after it (the
// slot's progress test) above the loads, so the hand data and the slot's
// progress are fetched in one round trip.
#define SWEEP_PIN(x) asm volatile("" ::"v"(x))

// A sweep task's record fields, read into scalars together before anything
// branches on them (one round trip instead of one per use).
struct SweepTask {
    int node, Mp, N, na, ab, fa, arow, nh, actor, NP, slot;
    int ch[NT_DEC];
    float *R, *SS, *cfv, *reach;
};
// nhp: whose hand count (the actor for the forward sweep, the traverser for
// the backward sweep)
__device__ __forceinline__ SweepTask sweep_task(const SweepRec *__restrict__ kr, int trav, bool fwd) {
    SweepTask t;
    t.node = kr->node; t.Mp = kr->Mp; t.N = kr->N; t.na = kr->na; t.ab = kr->act_base; t.fa = kr->fa;
    t.arow = kr->arow; t.actor = kr->actor; t.NP = kr->NP; t.slot = kr->slot;
#pragma unroll
    for (int a = 0; a < NT_DEC; ++a) t.ch[a] = kr->ch[a];
    int nh[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) nh[q] = kr->nh[q];
    t.R = kr->R; t.SS = kr->SS; t.cfv = kr->cfv; t.reach = kr->reach;
    // every field pinned here: the loads are issued together and waited once
    asm volatile("" ::"s"(t.node), "s"(t.Mp), "s"(t.N), "s"(t.na), "s"(t.ab), "s"(t.fa), "s"(t.arow), "s"(t.actor));
    asm volatile("" ::"s"(t.NP), "s"(t.slot), "s"(t.R), "s"(t.SS), "s"(t.cfv), "s"(t.reach));
    asm volatile("" ::"s"(t.ch[0]), "s"(t.ch[1]), "s"(t.ch[2]), "s"(t.ch[3]), "s"(t.ch[4]), "s"(t.ch[5]),
                 "s"(t.ch[6]), "s"(t.ch[7]));
    asm volatile("" ::"s"(nh[0]), "s"(nh[1]), "s"(nh[2]), "s"(nh[3]), "s"(nh[4]), "s"(nh[5]));
    const int w = fwd ? t.actor : trav;   // selects, static indices: nh stays in registers
    int x = nh[0];
#pragma unroll
    for (int q = 1; q < 6; ++q) x = w == q ? nh[q] : x;
    t.nh = x;
    return t;
}

// A workgroup walks several tasks (grid-stride): the record of its next task
// is touched (one word in each of its two 64-byte lines) while the current
// task's loads are in flight, and the words are consumed only at the end of
// the task, so the next task's record reads hit the scalar cache.
__device__ __forceinline__ int2 sweep_warm(const SweepRec *__restrict__ recs, int jn, int ntasks) {
    const SweepRec *r = recs + (jn < ntasks ? jn : 0);
    return make_int2(r->slot, r->nh[0]);
}
#define SWEEP_USE(w) asm volatile("" ::"s"((w).x), "s"((w).y))

// forward_one for a whole task (same results). The task's record carries the
// node: its fields are read into scalars up front (uniform address, so scalar
// loads). H hands per thread (h0 + u * blockDim.x), every load of the H hands
// issued together with the read of the slot's progress; A bounds the action
// count, so a node with few actions issues no repeated loads. Only the actor's
// child rows are written (reach-row table).
template <int A, int H, int NT>
__device__ __forceinline__ void fwd_task(const SweepTask &kr, int actor, int trav, int s_act, int s_it, int s_end,
                                         int s_np, int tid) {
    const int Mp = kr.Mp, N = kr.N, na = kr.na, ab = kr.ab, fa = kr.fa;
    int ch[A];
#pragma unroll
    for (int a = 0; a < A; ++a) ch[a] = kr.ch[a];
    const int nh = kr.nh, nt = NT;
    const float *R = kr.R;
    float *ra = kr.reach + (size_t)actor * N * Mp;
    const float *par = ra + (size_t)kr.arow * Mp;
    bool checked = false;
    for (int h0 = tid; h0 < nh; h0 += H * nt) {
        float r[H][A], pr[H];
#pragma unroll
        for (int u = 0; u < H; ++u) {
            const unsigned o = 4u * (unsigned)(h0 + u * nt < nh ? h0 + u * nt : h0);
#pragma unroll
            for (int a = 0; a < A; ++a) r[u][a] = gldo(R + (size_t)(ab + min(a, na - 1)) * Mp, o);
            pr[u] = gldo(par, o);
        }
        if (!checked) {
#pragma unroll
            for (int u = 0; u < H; ++u) SWEEP_PIN(pr[u]);
            if (!((s_act != 0) & (s_it < s_end) & (trav < s_np))) return;   // slot_runs
            checked = true;
        }
#pragma unroll
        for (int u = 0; u < H; ++u) {
            const int h = h0 + u * nt;
            if (h < nh) {
                float sig[A];
                regret_match_vals<A>(r[u], na, fa, sig);
#pragma unroll
                for (int a = 0; a < A; ++a)
                    if (a < na) gsto(ra + (size_t)ch[a] * Mp, 4u * (unsigned)h, pr[u] * sig[a]);
            }
        }
    }
}

// One task per workgroup, two hands per thread (k_fwd_phased), or, on
// levels with enough tasks to fill the GPU that way, three tasks per
// workgroup (k_fwd_grouped): one per group of three wavefronts (each
// wavefront reads its own task's record into scalars; nothing is shared, so
// the groups run independently), six hands per thread, 192 x 6 = 1152 >=
// 1128, all loads of a task in one round (nodes with more than three actions:
// two rounds of three hands).
#define SWEEP_GROUP 192
#define SWEEP_GH 6
__global__ void __launch_bounds__(SWEEP_BLOCK) k_fwd_phased(SlotDesc *__restrict__ slots,
                                                            const SweepRec *__restrict__ recs, int ntasks,
                                                            int trav) {
    const int G = gridDim.x;
    for (int j = blockIdx.x; j < ntasks; j += G) {
        const SweepTask kr = sweep_task(recs + j, trav, true);
        const int2 warm = sweep_warm(recs, j + G, ntasks);
        const int actor = kr.actor, NP = kr.NP;
        if (actor == trav || actor < 0 || actor >= NP) continue;
        const SlotDesc *s = slots + kr.slot;
        const int s_act = s->active, s_it = s->iter, s_end = s->iter_end, s_np = s->NP;
        if (kr.na <= 3) fwd_task<3, 2, SWEEP_BLOCK>(kr, actor, trav, s_act, s_it, s_end, s_np, threadIdx.x);
        else fwd_task<NT_DEC, 2, SWEEP_BLOCK>(kr, actor, trav, s_act, s_it, s_end, s_np, threadIdx.x);
        SWEEP_USE(warm);
    }
}
// the same, three tasks per workgroup (a kernel of its own: its registers do
// not weigh on the one-task form)
__global__ void __launch_bounds__(SWEEP_BLOCK) __attribute__((amdgpu_waves_per_eu(7)))
k_fwd_grouped(SlotDesc *__restrict__ slots, const SweepRec *__restrict__ recs, int ntasks, int trav) {
    const int G = gridDim.x, nunits = (ntasks + 2) / 3;
    // the group (0..2) is the same in every lane of a wavefront; readfirstlane
    // tells the compiler, so a group's record index and fields stay scalar
    const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x / SWEEP_GROUP);
    const int tg = threadIdx.x - sub * SWEEP_GROUP;
    for (int u = blockIdx.x; u < nunits; u += G) {
        const int j = 3 * u + sub;
        if (j >= ntasks) continue;
        const SweepTask kr = sweep_task(recs + j, trav, true);
        const int2 warm = sweep_warm(recs, j + 3 * G, ntasks);
        const int actor = kr.actor, NP = kr.NP;
        if (actor == trav || actor < 0 || actor >= NP) continue;
        const SlotDesc *s = slots + kr.slot;
        const int s_act = s->active, s_it = s->iter, s_end = s->iter_end, s_np = s->NP;
        if (kr.na <= 3) fwd_task<3, SWEEP_GH, SWEEP_GROUP>(kr, actor, trav, s_act, s_it, s_end, s_np, tg);
        else fwd_task<NT_DEC, 3, SWEEP_GROUP>(kr, actor, trav, s_act, s_it, s_end, s_np, tg);
        SWEEP_USE(warm);
    }
}

// terminal tasks with at most one opponent for every traverser
__global__ void __launch_bounds__(SS_BLOCK) k_term_light(SlotDesc *__restrict__ slots, const uint32_t *tasks,
                                                         int ntasks, int trav) {
    __shared__ TermLight sh;
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        uint32_t tk = tasks[j];
        const SlotDesc *s = slots + TASK_SLOT(tk);
        if (!slot_runs(s, trav)) continue;
        TermBuf<float> b = light_buf(sh, s);
        int n = TASK_NODE(tk);
        terminal_node<float, TM_LIGHT>(s, n, trav, s->reach, s->cfv + (size_t)n * s->Mp, b, TK_NODE, nullptr,
                                    threadIdx.x, blockDim.x);
        __syncthreads();
    }
}

// Terminal tasks with exactly two opponents for some traverser, at most 128
// VGPRs (four waves per SIMD). Showdowns (SDK = 1): 512 threads, two
// workgroups per CU by LDS. Folds and leaves (SDK = 0): 256 threads, four
// workgroups per CU.
#define SS_PAIR_BLOCK 512
template <int SDK>
__global__ void __launch_bounds__(SDK ? SS_PAIR_BLOCK : SS_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_term_pair(SlotDesc *__restrict__ slots, const uint32_t *tasks, int ntasks, int trav, int parts) {
    using Buf = typename std::conditional<SDK == 1, TermHeavy, TermFold>::type;
    __shared__ Buf sh;
    TermBuf<float> b;
    if constexpr (SDK == 1) b = heavy_buf(sh);
    else b = fold_buf(sh);
    // parts > 1 (showdowns, when a pass has fewer tasks than the GPU has
    // workgroup slots): each task's hands split over `parts` workgroups
    for (int jj = blockIdx.x; jj < ntasks * parts; jj += gridDim.x) {
        const int j = jj / parts, part = jj - j * parts;
        uint32_t tk = tasks[j];
        const SlotDesc *s = slots + TASK_SLOT(tk);
        if (!slot_runs(s, trav)) continue;
        if constexpr (SDK == 1) stage_tables(s, sh, threadIdx.x, blockDim.x);
        __syncthreads();
        int n = TASK_NODE(tk);
        terminal_node<float, TM_PAIR, SDK>(s, n, trav, s->reach, s->cfv + (size_t)n * s->Mp, b, TK_NODE, nullptr,
                                           threadIdx.x, blockDim.x, part, parts);
        __syncthreads();
    }
}

// A pass whose closed-form terminal tasks fit in two rounds of resident
// workgroups: the three classes in one launch (the workgroup index picks the
// class; the pass's segment lists showdowns, folds, light tasks in that
// order, so the slowest start first), so their latencies overlap instead of
// adding up. Every class runs
// with 512 threads in the largest LDS layout (occupancy does not matter at
// this size); results equal the separate kernels' (no terminal reduction
// depends on the block size).
union TermAny {
    TermHeavy heavy;
    TermFold fold;
    TermLight light;
};
__global__ void __launch_bounds__(SS_PAIR_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_term_small(SlotDesc *__restrict__ slots, const uint32_t *tasks, int nsd, int parts, int nfold, int nlight,
             int trav) {
    __shared__ TermAny sh;
    const int b = blockIdx.x, nsdw = nsd * parts;
    int j, part = 0;
    if (b < nsdw) {
        j = b / parts;
        part = b - j * parts;
    } else if (b < nsdw + nfold + nlight) {
        j = nsd + (b - nsdw);
    } else {
        return;
    }
    const uint32_t tk = tasks[j];
    const SlotDesc *s = slots + TASK_SLOT(tk);
    if (!slot_runs(s, trav)) return;   // workgroup-uniform
    const int n = TASK_NODE(tk);
    float *out = s->cfv + (size_t)n * s->Mp;
    if (b < nsdw) {
        stage_tables(s, sh.heavy, threadIdx.x, blockDim.x);
        __syncthreads();
        terminal_node<float, TM_PAIR, 1>(s, n, trav, s->reach, out, heavy_buf(sh.heavy), TK_NODE, nullptr,
                                         threadIdx.x, blockDim.x, part, parts);
    } else if (b < nsdw + nfold) {
        terminal_node<float, TM_PAIR, 0>(s, n, trav, s->reach, out, fold_buf(sh.fold), TK_NODE, nullptr,
                                         threadIdx.x, blockDim.x);
    } else {
        terminal_node<float, TM_LIGHT>(s, n, trav, s->reach, out, light_buf(sh.light, s), TK_NODE, nullptr,
                                       threadIdx.x, blockDim.x);
    }
}

// terminal tasks with three or more opponents for some traverser
__global__ void __launch_bounds__(SS_BLOCK) k_term_odo(SlotDesc *__restrict__ slots, const uint32_t *tasks,
                                                       int ntasks, int trav) {
    __shared__ TermHeavy sh;
    TermBuf<float> b = heavy_buf(sh);
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        uint32_t tk = tasks[j];
        const SlotDesc *s = slots + TASK_SLOT(tk);
        if (!slot_runs(s, trav)) continue;
        stage_tables(s, sh, threadIdx.x, blockDim.x);
        __syncthreads();
        int n = TASK_NODE(tk);
        terminal_node<float, TM_ALL>(s, n, trav, s->reach, s->cfv + (size_t)n * s->Mp, b, TK_NODE, nullptr,
                                     threadIdx.x, blockDim.x);
        __syncthreads();
    }
}

// backward_one for a whole task (same results), from the task's record as
// the forward sweep
// one hand of a backward task from its loaded child values (and, at the
// traverser's own node, regret and strategy-sum rows)
template <int A, bool OWN>
__device__ __forceinline__ void bwd_hand(int na, int fa, int ab, int Mp, int h, float d, int iter, const float *cv,
                                         const float *r, const float *sg, float *R, float *SS, float *out) {
    float v = 0.f;
    if (OWN) {
        float sig[A];
        regret_match_vals<A>(r, na, fa, sig);
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (a < na) v += sig[a] * cv[a];
#pragma unroll
        for (int a = 0; a < A; ++a) {
            if (a < na) {
                const size_t row = (size_t)(ab + a) * Mp;
                float x = r[a];
                x += cv[a] - v;
                x *= d;
                gsto(R + row, 4u * (unsigned)h, x);
                gsto(SS + row, 4u * (unsigned)h, sg[a] + (float)iter * sig[a]);
            }
        }
    } else {
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (a < na) v += cv[a];
    }
    gsto(out, 4u * (unsigned)h, v);
}

// backward_one for a whole task (same results): H hands per thread, every
// load of the H hands (child values; at the traverser's own node also the
// regret and strategy-sum rows) issued together with the read of the slot's
// progress. A bounds the action count.
template <int A, int H, bool OWN>
__device__ __forceinline__ void bwd_task(const SweepTask &kr, int trav, int s_act, int s_it, int s_end, int s_np) {
    const int n = kr.node, Mp = kr.Mp, na = kr.na, ab = kr.ab, fa = kr.fa;
    int ch[A];
#pragma unroll
    for (int a = 0; a < A; ++a) ch[a] = kr.ch[a];
    const int nh = kr.nh, nt = SWEEP_BLOCK;
    float *R = kr.R, *SS = kr.SS, *cfv = kr.cfv;
    float *out = cfv + (size_t)n * Mp;
    bool checked = false;
    float d = 0.f;
    int iter = 0;
    for (int h0 = threadIdx.x; h0 < nh; h0 += H * nt) {
        float cv[H][A], r[H][A], sg[H][A];
#pragma unroll
        for (int u = 0; u < H; ++u) {
            const unsigned o = 4u * (unsigned)(h0 + u * nt < nh ? h0 + u * nt : h0);
#pragma unroll
            for (int a = 0; a < A; ++a) {
                cv[u][a] = gldo(cfv + (size_t)ch[a] * Mp, o);
                if (OWN) {
                    const size_t row = (size_t)(ab + min(a, na - 1)) * Mp;
                    r[u][a] = gldo(R + row, o);
                    sg[u][a] = gldo(SS + row, o);
                } else {
                    r[u][a] = sg[u][a] = 0.f;
                }
            }
        }
        if (!checked) {
#pragma unroll
            for (int u = 0; u < H; ++u)
#pragma unroll
                for (int a = 0; a < A; ++a) SWEEP_PIN(cv[u][a]);
            if (!((s_act != 0) & (s_it < s_end) & (trav < s_np))) return;   // slot_runs
            iter = s_it + 1;
            d = (float)iter / ((float)iter + 1.0f);
            checked = true;
        }
#pragma unroll
        for (int u = 0; u < H; ++u) {
            const int h = h0 + u * nt;
            if (h < nh) bwd_hand<A, OWN>(na, fa, ab, Mp, h, d, iter, cv[u], r[u], sg[u], R, SS, out);
        }
    }
}

// A backward task at the traverser's own node: child values, regret and
// strategy-sum rows, one hand per thread per round, the loads issued together
// with the read of the slot's progress. (As a bwd_task instance the row
// addresses are kept live across rounds.)
template <int A>
__device__ __forceinline__ void bwd_own(const SweepTask &kr, int trav, int s_act, int s_it, int s_end, int s_np) {
    const int n = kr.node, Mp = kr.Mp, na = kr.na, ab = kr.ab, fa = kr.fa;
    int ch[A];
#pragma unroll
    for (int a = 0; a < A; ++a) ch[a] = kr.ch[a];
    const int nh = kr.nh;
    float *R = kr.R, *SS = kr.SS, *cfv = kr.cfv;
    bool checked = false;
    float d = 0.f;
    int iter = 0;
    for (int h = threadIdx.x; h < nh; h += SWEEP_BLOCK) {
        float cv[A], r[A], sg[A];
        const unsigned o = 4u * (unsigned)h;
#pragma unroll
        for (int a = 0; a < A; ++a) {
            const size_t row = (size_t)(ab + min(a, na - 1)) * Mp;
            cv[a] = gldo(cfv + (size_t)ch[a] * Mp, o);
            r[a] = gldo(R + row, o);
            sg[a] = gldo(SS + row, o);
        }
        if (!checked) {
#pragma unroll
            for (int a = 0; a < A; ++a) SWEEP_PIN(cv[a]);
            if (!((s_act != 0) & (s_it < s_end) & (trav < s_np))) break;   // slot_runs
            iter = s_it + 1;
            d = (float)iter / ((float)iter + 1.0f);
            checked = true;
        }
        bwd_hand<A, true>(na, fa, ab, Mp, h, d, iter, cv, r, sg, R, SS, cfv + (size_t)n * Mp);
    }
}

// Backward tasks: two hands per thread at other players' nodes; the
// traverser's own nodes (child values, regret and strategy-sum rows: up to 24
// loads per hand) one hand per thread per round. 78 VGPRs (two workgroups per
// CU): capped at 72 for a third the paths spill and measure slower.
__global__ void __launch_bounds__(SWEEP_BLOCK) k_bwd_phased(SlotDesc *__restrict__ slots,
                                                            const SweepRec *__restrict__ recs, int ntasks,
                                                            int trav) {
    const int G = gridDim.x;
    for (int j = blockIdx.x; j < ntasks; j += G) {
        const SweepTask kr = sweep_task(recs + j, trav, false);
        const int2 warm = sweep_warm(recs, j + G, ntasks);
        const SlotDesc *s = slots + kr.slot;
        const int s_act = s->active, s_it = s->iter, s_end = s->iter_end, s_np = s->NP;
        const bool few = kr.na <= 3;
        if (kr.actor == trav) {
            if (few) bwd_own<3>(kr, trav, s_act, s_it, s_end, s_np);
            else bwd_own<NT_DEC>(kr, trav, s_act, s_it, s_end, s_np);
        } else {
            if (few) bwd_task<3, 2, false>(kr, trav, s_act, s_it, s_end, s_np);
            else bwd_task<NT_DEC, 2, false>(kr, trav, s_act, s_it, s_end, s_np);
        }
        SWEEP_USE(warm);
    }
}

// root strategy snapshot before the last pass of a final / check iteration
__global__ void __launch_bounds__(SS_BLOCK) k_rsig_phased(SlotDesc *__restrict__ slots, int nslots, int trav) {
    for (int si = blockIdx.x; si < nslots; si += gridDim.x) {
        const SlotDesc *s = slots + si;
        if (!slot_runs(s, trav) || trav != s->NP - 1) continue;
        int it = s->iter + 1;
        if (it == s->final_iter || (s->snap_every > 0 && it % s->snap_every == 0))
            root_sigma_items(s, threadIdx.x, blockDim.x);
    }
}

__global__ void k_bump_phased(SlotDesc *__restrict__ slots, int nslots) {
    int si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si < nslots && slots[si].active && slots[si].iter < slots[si].iter_end) slots[si].iter += 1;
}

// FP64 exploitability, phased over every slot with want[slot] != 0
__device__ __forceinline__ bool slot_wanted(const SlotDesc *__restrict__ s, const int *want, int si) {
    return s->active && want[si];
}

__global__ void __launch_bounds__(SS_BLOCK) k_xavg_phased(SlotDesc *__restrict__ slots, const uint32_t *tasks,
                                                          int ntasks, const int *want) {
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        uint32_t tk = tasks[j];
        const SlotDesc *s = slots + TASK_SLOT(tk);
        if (!slot_wanted(s, want, TASK_SLOT(tk))) continue;
        for (int h = threadIdx.x; h < s->Mp; h += blockDim.x) avg_one(s, TASK_NODE(tk), h);
    }
}

// FP64 sweeps of the exploitability check over the sweep records of every
// decision task (the first pass's backward lists): the record read into
// scalars in one batch, grid-stride over tasks; same arithmetic as
// xforward_one / xbackward_one.
__global__ void __launch_bounds__(SWEEP_BLOCK) k_xfwd_phased(SlotDesc *__restrict__ slots,
                                                             const SweepRec *__restrict__ recs, int ntasks,
                                                             const int *want) {
    const int G = gridDim.x;
    for (int j = blockIdx.x; j < ntasks; j += G) {
        const SweepTask kr = sweep_task(recs + j, 0, true);
        const SlotDesc *s = slots + kr.slot;
        const int q = kr.actor;
        if (!s->active || !want[kr.slot] || q < 0 || q >= kr.NP) continue;
        const int Mp = kr.Mp, na = kr.na, ab = kr.ab;
        double *rr = s->xreach + (size_t)q * kr.N * Mp;
        const double *par = rr + (size_t)kr.arow * Mp;
        const float *avg = s->avg;
        for (int h = threadIdx.x; h < kr.nh; h += SWEEP_BLOCK) {
            const double pr = par[h];
#pragma unroll
            for (int a = 0; a < NT_DEC; ++a)
                if (a < na) rr[(size_t)kr.ch[a] * Mp + h] = pr * (double)avg[(size_t)(ab + a) * Mp + h];
        }
    }
}
__global__ void __launch_bounds__(SS_BLOCK) k_xterm_phased(SlotDesc *__restrict__ slots, const uint32_t *tasks,
                                                           int ntasks, int p, const int *want, double *scratch) {
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        uint32_t tk = tasks[j];
        const SlotDesc *s = slots + TASK_SLOT(tk);
        if (!slot_wanted(s, want, TASK_SLOT(tk)) || p >= s->NP) continue;
        TermBuf<double> b = global_buf(scratch + (size_t)blockIdx.x * XSCRATCH_DOUBLES, s);
        int n = TASK_NODE(tk);
        terminal_node<double, TM_ALL>(s, n, p, s->xreach, s->xev + (size_t)n * s->Mp, b, TK_NODE, nullptr, threadIdx.x,
                              blockDim.x);
        for (int h = threadIdx.x; h < s->nh[p]; h += blockDim.x)
            s->xbr[(size_t)n * s->Mp + h] = s->xev[(size_t)n * s->Mp + h];
        __syncthreads();
    }
}

// FP64 terminals of the closed-form classes with LDS buffers, split as the
// FP32 ones: KIND 0 at most one opponent, 1 two-opponent showdowns (150 KB,
// one workgroup per CU), 2 two-opponent folds and leaves. k_xterm_phased
// (global scratch) keeps the 3+ opponent tasks.
template <int KIND>
__global__ void __launch_bounds__(KIND == 1 ? 512 : SS_BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
k_xterm_cls(SlotDesc *__restrict__ slots, const uint32_t *tasks, int ntasks, int p, const int *want) {
    using Buf = typename std::conditional<KIND == 0, TermLightT<double>,
                                          typename std::conditional<KIND == 1, TermHeavyT<double>,
                                                                    TermFoldT<double>>::type>::type;
    __shared__ Buf sh;
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        uint32_t tk = tasks[j];
        const SlotDesc *s = slots + TASK_SLOT(tk);
        if (!slot_wanted(s, want, TASK_SLOT(tk)) || p >= s->NP) continue;
        const int n = TASK_NODE(tk);
        double *ev = s->xev + (size_t)n * s->Mp;
        if constexpr (KIND == 0) {
            terminal_node<double, TM_LIGHT>(s, n, p, s->xreach, ev, light_buf(sh, s), TK_NODE, nullptr, threadIdx.x,
                                            blockDim.x);
        } else if constexpr (KIND == 1) {
            stage_tables(s, sh, threadIdx.x, blockDim.x);
            __syncthreads();
            terminal_node<double, TM_PAIR, 1>(s, n, p, s->xreach, ev, heavy_buf(sh), TK_NODE, nullptr, threadIdx.x,
                                              blockDim.x);
        } else {
            terminal_node<double, TM_PAIR, 0>(s, n, p, s->xreach, ev, fold_buf(sh), TK_NODE, nullptr, threadIdx.x,
                                              blockDim.x);
        }
        for (int h = threadIdx.x; h < s->nh[p]; h += blockDim.x) s->xbr[(size_t)n * s->Mp + h] = ev[h];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(SWEEP_BLOCK) k_xbwd_phased(SlotDesc *__restrict__ slots,
                                                             const SweepRec *__restrict__ recs, int ntasks, int p,
                                                             const int *want) {
    const int G = gridDim.x;
    for (int j = blockIdx.x; j < ntasks; j += G) {
        const SweepTask kr = sweep_task(recs + j, p, false);
        const SlotDesc *s = slots + kr.slot;
        if (!s->active || !want[kr.slot] || p >= kr.NP) continue;
        const int Mp = kr.Mp, na = kr.na, ab = kr.ab;
        const bool own = kr.actor == p;
        double *xev = s->xev, *xbr = s->xbr;
        const float *avg = s->avg;
        for (int h = threadIdx.x; h < kr.nh; h += SWEEP_BLOCK) {
            double e = 0, b = 0;
#pragma unroll
            for (int a = 0; a < NT_DEC; ++a) {
                if (a < na) {
                    const double ce = xev[(size_t)kr.ch[a] * Mp + h], cb = xbr[(size_t)kr.ch[a] * Mp + h];
                    if (own) {
                        e += (double)avg[(size_t)(ab + a) * Mp + h] * ce;
                        b = (a == 0 || cb > b) ? cb : b;
                    } else {
                        e += ce;
                        b += cb;
                    }
                }
            }
            xev[(size_t)kr.node * Mp + h] = e;
            xbr[(size_t)kr.node * Mp + h] = b;
        }
    }
}
__global__ void __launch_bounds__(SS_BLOCK) k_xreduce_phased(SlotDesc *__restrict__ slots, int nslots, int p,
                                                             const int *want) {
    __shared__ double red[2 * SS_BLOCK];
    for (int si = blockIdx.x; si < nslots; si += gridDim.x) {
        const SlotDesc *s = slots + si;
        if (!slot_wanted(s, want, si) || p >= s->NP) continue;
        xreduce_block(s, p, red, threadIdx.x, blockDim.x);
    }
}

__global__ void __launch_bounds__(SS_BLOCK) k_xjoint_phased(SlotDesc *__restrict__ slots, int nslots,
                                                            const int *want, double *scratch) {
    __shared__ double red[SS_BLOCK];
    for (int si = blockIdx.x; si < nslots; si += gridDim.x) {
        const SlotDesc *s = slots + si;
        if (!slot_wanted(s, want, si)) continue;
        TermBuf<double> b = global_buf(scratch + (size_t)blockIdx.x * XSCRATCH_DOUBLES, s);
        terminal_node<double, TM_ALL>(s, 0, 0, s->xreach, s->xev, b, TK_MASS, s->weights, threadIdx.x, blockDim.x);
        double acc = 0;
        for (int h = threadIdx.x; h < s->nh[0]; h += blockDim.x)
            if (s->hinfo[h] & SS_LIVE_BIT) acc += s->xev[h];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int off = blockDim.x / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) s->xout[2 * s->NP] = red[0];
        __syncthreads();
    }
}

}  // namespace ss

// ===========================================================================
// host launchers
// ===========================================================================
#include "ss_internal.h"

namespace ss {

static inline int grid_for(int64_t items, int cap = 8192) {
    int64_t g = (items + SS_BLOCK - 1) / SS_BLOCK;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

int launch_combo_ranks(const int *d_board, uint32_t *d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_combo_ranks, dim3(6), dim3(SS_BLOCK), 0, st, d_board, d_out);
    SS_HIP(hipGetLastError());
    return SS_OK;
}

int plain_run(Slot &s, int iters, int final_iter, hipStream_t st, double *kernel_ms) {
    const Plan &pl = s.plan;
    SlotDesc *d = s.d_desc;
    hipEvent_t e0, e1;
    SS_HIP(hipEventCreate(&e0));
    SS_HIP(hipEventCreate(&e1));
    SS_HIP(hipEventRecord(e0, st));
    for (int k = 0; k < iters; ++k) {
        int it = s.desc.iter + 1;
        for (int t = 0; t < pl.NP; ++t) {
            if (it == final_iter && t == pl.NP - 1)
                hipLaunchKernelGGL(k_root_sigma, dim3(grid_for(pl.Mp)), dim3(SS_BLOCK), 0, st, d);
            for (int l = 0; l < pl.L; ++l) {
                int cnt = pl.dec_off[l + 1] - pl.dec_off[l];
                if (cnt > 0)
                    hipLaunchKernelGGL(k_forward, dim3(grid_for((int64_t)cnt * pl.Mp)), dim3(SS_BLOCK), 0, st, d, l, t);
            }
            if (s.prec) {
                int rc = prec_showdowns(s, t, st);
                if (rc) return rc;
            }
            if (pl.T > 0) hipLaunchKernelGGL(k_terminals, dim3(pl.T), dim3(SS_BLOCK), 0, st, d, t, s.prec ? 1 : 0);
            for (int l = pl.L - 1; l >= 0; --l) {
                int cnt = pl.dec_off[l + 1] - pl.dec_off[l];
                if (cnt > 0)
                    hipLaunchKernelGGL(k_backward, dim3(grid_for((int64_t)cnt * pl.Mp)), dim3(SS_BLOCK), 0, st, d, l, t);
            }
        }
        hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, st, d);
        s.desc.iter = it;
    }
    SS_HIP(hipGetLastError());
    SS_HIP(hipEventRecord(e1, st));
    SS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SS_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (kernel_ms) *kernel_ms += ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return SS_OK;
}

int plain_exploitability(Slot &s, hipStream_t st, double *expl_pct, double *gains) {
    const Plan &pl = s.plan;
    int blocks = std::max(1, std::min(pl.T, 64));
    int rc = s.alloc_expl(st, blocks);
    if (rc) return rc;
    SlotDesc *d = s.d_desc;
    hipLaunchKernelGGL(k_avg, dim3(grid_for((int64_t)pl.D * pl.Mp)), dim3(SS_BLOCK), 0, st, d);
    for (int l = 0; l < pl.L; ++l) {
        int cnt = pl.dec_off[l + 1] - pl.dec_off[l];
        if (cnt > 0) hipLaunchKernelGGL(k_xforward, dim3(grid_for((int64_t)cnt * pl.Mp)), dim3(SS_BLOCK), 0, st, d, l);
    }
    for (int p = 0; p < pl.NP; ++p) {
        if (pl.T > 0) hipLaunchKernelGGL(k_xterminals, dim3(blocks), dim3(SS_BLOCK), 0, st, d, p, s.xscratch);
        for (int l = pl.L - 1; l >= 0; --l) {
            int cnt = pl.dec_off[l + 1] - pl.dec_off[l];
            if (cnt > 0)
                hipLaunchKernelGGL(k_xbackward, dim3(grid_for((int64_t)cnt * pl.Mp)), dim3(SS_BLOCK), 0, st, d, l, p);
        }
        hipLaunchKernelGGL(k_xreduce, dim3(1), dim3(SS_BLOCK), 0, st, d, p);
    }
    hipLaunchKernelGGL(k_xjoint, dim3(1), dim3(SS_BLOCK), 0, st, d, s.xscratch);
    SS_HIP(hipGetLastError());
    double xout[2 * SS_MAX_PLAYERS + 1];
    SS_HIP(hipMemcpyAsync(xout, s.desc.xout, sizeof(double) * (2 * pl.NP + 1), hipMemcpyDeviceToHost, st));
    SS_HIP(hipStreamSynchronize(st));
    *expl_pct = expl_from_xout(xout, pl.NP, pl.starting_pot, gains);
    return SS_OK;
}

int launch_batched(SlotDesc *d_slots, int nslots, int iters, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_batched, dim3(grid), dim3(SS_BATCH_BLOCK), 0, st, d_slots, nslots, iters);
    SS_HIP(hipGetLastError());
    return SS_OK;
}

int launch_batched_expl(SlotDesc *d_slots, int nslots, double *scratch, const int *d_want, int grid,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_batched_expl, dim3(grid), dim3(SS_BLOCK), 0, st, d_slots, nslots, scratch, d_want);
    SS_HIP(hipGetLastError());
    return SS_OK;
}

// ---- phased batched form ---------------------------------------------------
PhasedTasks::~PhasedTasks() {
    if (d_tasks) (void)hipFree(d_tasks);
    if (d_recs) (void)hipFree(d_recs);
}

// terminal cost class for scheduling: closed forms with 2 opponents at a
// showdown are the heaviest (a 52-card loop per hand), odometers heavier still
static int term_cost(const Plan &pl, int n, bool prec) {
    const DevNode &nd = pl.nodes[n];
    if (prec && nd.type == SS_NODE_SHOWDOWN) return 1;   // precision form: the contraction, FP64 checks only
    // the most opponents any traverser meets here; at a showdown a traverser
    // who has folded gets 0 with no opponent work (terminal_node's zero case),
    // so only the active players count there
    int kmax = 0;
    for (int t = 0; t < pl.NP; ++t) {
        if (nd.type == SS_NODE_SHOWDOWN && !((nd.active_mask >> t) & 1)) continue;
        int k = 0;
        for (int q = 0; q < pl.NP; ++q)
            if (q != t && (nd.type == SS_NODE_FOLD || ((nd.active_mask >> q) & 1))) ++k;
        kmax = std::max(kmax, k);
    }
    if (kmax >= 3) return 1000;
    if (kmax == 2) return nd.type == SS_NODE_SHOWDOWN ? 100 : 20;   // k_term_pair<1> / <0>
    return 5;
}

// Workgroups of a launch are dealt to the XCDs round-robin (MI355X: 8 XCDs,
// each with its own L2). Every phased launch runs task j on workgroup j, so
// ordering a segment to put subgame s at positions j = s (mod 8) keeps each
// subgame's tables and state in one XCD's L2. Queues that run dry are
// backfilled from the longest remaining one, so the order stays a permutation.
// terminal workgroups of one pass that run as one launch (k_term_small): one
// round of two 76 KB workgroups per CU
static int device_cus() {
    static const int cus = [] {
        int n = 0, dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;   // MI355X
        return n;
    }();
    return cus;
}
static int small_term_limit() { return 2 * device_cus(); }

// launch shapes forced on every task list (tests); -1: not forced
static PhasedTasks::Shape g_shape;
void set_phased_shape(int fwd_group_min, int term_small, int sd_parts) {
    g_shape.fwd_group_min = fwd_group_min;
    g_shape.term_small = term_small;
    g_shape.sd_parts = sd_parts;
}
bool phased_shape_forced() {
    return g_shape.fwd_group_min >= 0 || g_shape.term_small >= 0 || g_shape.sd_parts >= 1;
}
// a task list's shape with the forced fields applied
static PhasedTasks::Shape effective_shape(const PhasedTasks &pt) {
    PhasedTasks::Shape s = pt.shape;
    if (g_shape.fwd_group_min >= 0) s.fwd_group_min = g_shape.fwd_group_min;
    if (g_shape.term_small >= 0) s.term_small = g_shape.term_small;
    if (g_shape.sd_parts >= 1) s.sd_parts = g_shape.sd_parts;
    return s;
}

#define SS_NUM_XCD 8
static void xcd_interleave(std::vector<uint32_t> &v, size_t b, size_t e) {
    std::vector<std::vector<uint32_t>> q(SS_NUM_XCD);
    for (size_t i = b; i < e; ++i) q[(v[i] >> 20) % SS_NUM_XCD].push_back(v[i]);
    std::vector<size_t> at(SS_NUM_XCD, 0);
    for (size_t j = b; j < e; ++j) {
        int x = (int)((j - b) % SS_NUM_XCD);
        if (at[x] == q[x].size()) {
            size_t best = 0;
            for (int y = 0; y < SS_NUM_XCD; ++y)
                if (q[y].size() - at[y] > best) { best = q[y].size() - at[y]; x = y; }
        }
        v[j] = q[x][at[x]++];
    }
}

int build_phased_tasks(const std::vector<Slot *> &slots, PhasedTasks &pt, hipStream_t st, const std::vector<int> *live) {
    const PhasedTasks::Shape shape = effective_shape(pt);
    pt.nslots = (int)slots.size();
    pt.NPmax = 0;
    pt.L = 0;
    auto on = [&](int si) { return slots[si] && (!live || (*live)[si]); };
    for (int si = 0; si < pt.nslots; ++si) {
        const Slot *s = slots[si];
        if (!on(si)) continue;
        if (s->plan.N >= (1 << 20)) return set_error(SS_ERR_UNSUPPORTED, "tree too large for the batched form");
        pt.NPmax = std::max(pt.NPmax, s->plan.NP);
        pt.L = std::max(pt.L, s->plan.L);
    }
    if (pt.nslots >= (1 << 11)) return set_error(SS_ERR_UNSUPPORTED, "at most 2047 subgames per batch");
    std::vector<uint32_t> tasks;
    pt.lvl_off.assign(pt.L + 1, 0);
    for (int l = 0; l < pt.L; ++l) {
        pt.lvl_off[l] = (int)tasks.size();
        for (int si = 0; si < pt.nslots; ++si) {
            if (!on(si)) continue;
            const Plan &pl = slots[si]->plan;
            if (l >= pl.L) continue;
            for (int i = pl.dec_off[l]; i < pl.dec_off[l + 1]; ++i)
                tasks.push_back(((uint32_t)si << 20) | (uint32_t)pl.dec_lvl[i]);
        }
        xcd_interleave(tasks, pt.lvl_off[l], tasks.size());
    }
    pt.lvl_off[pt.L] = (int)tasks.size();
    std::vector<std::pair<int, uint32_t>> term;
    for (int si = 0; si < pt.nslots; ++si) {
        if (!on(si)) continue;
        const Plan &pl = slots[si]->plan;
        const bool prec = slots[si]->prec != nullptr;
        for (int j = 0; j < pl.T; ++j)
            term.push_back({term_cost(pl, pl.term[j], prec), ((uint32_t)si << 20) | (uint32_t)pl.term[j]});
    }
    std::stable_sort(term.begin(), term.end(), [](const std::pair<int, uint32_t> &a, const std::pair<int, uint32_t> &b) {
        return a.first > b.first;
    });
    pt.term_off = (int)tasks.size();
    pt.nterm = (int)term.size();
    pt.nodo = pt.npair_sd = pt.npair = pt.nprec = 0;
    for (auto &t : term) {
        tasks.push_back(t.second);
        if (t.first >= 1000) ++pt.nodo;
        else if (t.first >= 20) ++pt.npair;
        if (t.first == 100) ++pt.npair_sd;
        if (t.first == 1) ++pt.nprec;
    }
    // each kernel's segment in XCD order (equal cost within a segment)
    const size_t t0 = pt.term_off, t1 = t0 + pt.nodo, t2 = t1 + pt.npair_sd, t3 = t0 + pt.nodo + pt.npair;
    xcd_interleave(tasks, t0, t1);
    xcd_interleave(tasks, t1, t2);
    const size_t t4 = tasks.size() - pt.nprec;
    xcd_interleave(tasks, t2, t3);
    xcd_interleave(tasks, t3, t4);
    xcd_interleave(tasks, t4, tasks.size());
    // precision form: split tasks (slot, showdown index), then GEMM tasks
    // (slot, 16-row tile) grouped by column-tile count
    pt.psplit_off = (int)tasks.size();
    for (int si = 0; si < pt.nslots; ++si)
        if (on(si) && slots[si]->prec)
            for (int b = 0; b < slots[si]->prec->S; ++b) tasks.push_back(((uint32_t)si << 20) | (uint32_t)b);
    pt.npsplit = (int)tasks.size() - pt.psplit_off;
    xcd_interleave(tasks, pt.psplit_off, tasks.size());
    pt.pgemm_off = (int)tasks.size();
    pt.pgemm.clear();
    for (int nc = 1; nc <= 16; ++nc) {
        std::vector<int> group, tiles;
        for (int si = 0; si < pt.nslots; ++si)
            if (on(si) && slots[si]->prec && slots[si]->prec->NC == nc) {
                group.push_back(si);
                tiles.push_back(slots[si]->plan.Mp / 16);
            }
        if (group.empty()) continue;
        const int rt = prec_choose_rt(nc, tiles), rows = 64 * rt;
        const size_t g0 = tasks.size();
        for (int si : group)
            for (int r = 0; r < (slots[si]->plan.Mp + rows - 1) / rows; ++r)
                tasks.push_back(((uint32_t)si << 20) | (uint32_t)r);   // row blocks of k_prec_gemm_phased
        xcd_interleave(tasks, g0, tasks.size());
        pt.pgemm.push_back({nc, rt, (int)(tasks.size() - g0)});
    }
    // per traverser pass: terminal tasks of the slots with more than t
    // players, by class, each class in XCD order
    pt.pterm.assign(pt.NPmax, PhasedTasks::PassTerm());
    for (int t = 0; t < pt.NPmax; ++t) {
        PhasedTasks::PassTerm &q = pt.pterm[t];
        q.off = (int)tasks.size();
        // the root snapshot is taken in a subgame's last pass: only passes
        // that are some listed subgame's last launch it
        q.rsig = false;
        for (int si = 0; si < pt.nslots; ++si) q.rsig |= on(si) && slots[si]->plan.NP == t + 1;
        auto seg = [&](int lo, int hi) {
            const size_t b = tasks.size();
            for (auto &e : term)
                if (e.first >= lo && e.first <= hi && t < slots[e.second >> 20]->plan.NP) tasks.push_back(e.second);
            xcd_interleave(tasks, b, tasks.size());
            return (int)(tasks.size() - b);
        };
        q.nodo = seg(1000, 1 << 30);
        q.nsd = seg(100, 100);
        q.nfold = seg(20, 20);
        q.nlight = seg(5, 5);
        // fewer two-opponent showdowns than resident workgroup slots (two per
        // CU): split each one's hands, up to four ways, so the pass's latency
        // is not one whole node's
        const int sd_slots = 2 * device_cus();
        q.sd_parts = q.nsd > 0 ? std::max(1, std::min(4, sd_slots / q.nsd)) : 1;
        if (shape.sd_parts >= 1) q.sd_parts = std::min(4, shape.sd_parts);
        // two-opponent showdowns and the other closed-form classes within two
        // rounds of resident workgroups: one launch (k_term_small), the
        // showdowns first so the lighter classes cycle through the other
        // slots beside them; the showdowns split to fill what the others
        // leave of one round (without showdowns there is nothing to overlap:
        // the light kernel's own shape is faster)
        const int lim = small_term_limit(), nall = q.nsd + q.nfold + q.nlight;
        q.small_parts = 0;
        if (q.nodo == 0 && q.nsd > 0 && nall <= 2 * lim && shape.term_small != 0)
            q.small_parts = std::max(1, std::min(4, (lim - q.nfold - q.nlight) / q.nsd));
    }
    // per traverser pass and level: forward tasks (actor not the traverser),
    // backward tasks, of the slots with more than t players
    // (backward levels: the traverser's own nodes, the heavier tasks, first)
    std::vector<uint32_t> sw;
    pt.sw_off.assign((size_t)std::max(1, pt.NPmax) * 2 * (pt.L + 1), 0);
    for (int t = 0; t < pt.NPmax; ++t)
        for (int dir = 0; dir < 2; ++dir) {
            int *off = &pt.sw_off[((size_t)t * 2 + dir) * (pt.L + 1)];
            for (int l = 0; l < pt.L; ++l) {
                off[l] = (int)sw.size();
                for (int pass = dir == 1 ? 0 : 1; pass < 2; ++pass) {
                    const size_t b = sw.size();
                    for (int j = pt.lvl_off[l]; j < pt.lvl_off[l + 1]; ++j) {
                        const Plan &pl = slots[tasks[j] >> 20]->plan;
                        const int a = pl.nodes[tasks[j] & 0xFFFFF].player;
                        if (t >= pl.NP || (dir == 0 && (a == t || a < 0 || a >= pl.NP))) continue;
                        if (dir == 1 && (pass == 0) != (a == t)) continue;
                        sw.push_back(tasks[j]);
                    }
                    xcd_interleave(sw, b, sw.size());
                }
            }
            off[pt.L] = (int)sw.size();
        }
    const size_t ndec = sw.size();
    std::vector<SweepRec> recs(ndec);
    for (size_t j = 0; j < ndec; ++j) {
        const int si = (int)(sw[j] >> 20), n = (int)(sw[j] & 0xFFFFF);
        const Slot &S = *slots[si];
        const Plan &pl = S.plan;
        const DevNode &nd = pl.nodes[n];
        SweepRec &r = recs[j];   // value-initialised: padding is zero
        r.slot = si;
        r.node = n;
        r.Mp = pl.Mp;
        r.N = pl.N;
        r.NP = pl.NP;
        r.actor = nd.player;
        r.na = nd.nchild;
        r.act_base = nd.act_base;
        r.fa = pl.frozen.empty() ? -1 : pl.frozen[n];
        r.arow = nd.player >= 0 && nd.player < pl.NP ? pl.rrow[(size_t)nd.player * pl.N + n] : 0;
        for (int a = 0; a < 8; ++a) r.ch[a] = pl.children[nd.first_child + std::min(a, nd.nchild - 1)];
        for (int q = 0; q < 6; ++q) r.nh[q] = S.desc.nh[q];
        r.R = S.desc.R;
        r.SS = S.desc.SS;
        r.cfv = S.desc.cfv;
        r.reach = S.desc.reach;
    }
    // device lists: reused while large enough (a compaction only shrinks them)
    if (!pt.d_tasks || pt.cap_tasks < tasks.size()) {
        if (pt.d_tasks) SS_HIP(hipFree(pt.d_tasks));
        pt.d_tasks = nullptr;
        pt.cap_tasks = std::max<size_t>(1, tasks.size());
        SS_HIP(hipMalloc(&pt.d_tasks, sizeof(uint32_t) * pt.cap_tasks));
    }
    if (!pt.d_recs || pt.cap_recs < ndec) {
        if (pt.d_recs) SS_HIP(hipFree(pt.d_recs));
        pt.d_recs = nullptr;
        pt.cap_recs = std::max<size_t>(1, ndec);
        SS_HIP(hipMalloc(&pt.d_recs, sizeof(SweepRec) * pt.cap_recs));
    }
    SS_HIP(hipMemcpyAsync(pt.d_tasks, tasks.data(), sizeof(uint32_t) * tasks.size(), hipMemcpyHostToDevice, st));
    if (ndec) SS_HIP(hipMemcpyAsync(pt.d_recs, recs.data(), sizeof(SweepRec) * ndec, hipMemcpyHostToDevice, st));
    SS_HIP(hipStreamSynchronize(st));
    return SS_OK;
}

// sweep launches: at most this many workgroups (six per CU), each walking
// its share of the level's tasks (SS_SWEEP_GRID overrides)
static const int SWEEP_GRID = [] {
    const char *e = getenv("SS_SWEEP_GRID");
    return e && atoi(e) > 0 ? atoi(e) : 1536;
}();

// sweep tasks per workgroup: three when they still fill three resident
// workgroups per CU that way, else one
static int per_of(int cnt, const PhasedTasks::Shape &s) {
    if (s.fwd_group_min >= 0) return cnt >= s.fwd_group_min ? 3 : 1;
    return cnt >= 3 * 3 * device_cus() ? 3 : 1;
}

int launch_phased_iteration(SlotDesc *d_slots, const PhasedTasks &pt, hipStream_t st) {
    const PhasedTasks::Shape shape = effective_shape(pt);
    const int SWEEP_GRID = shape.sweep_grid > 0 ? shape.sweep_grid : ss::SWEEP_GRID;
    for (int t = 0; t < pt.NPmax; ++t) {
        const int *fo = &pt.sw_off[(size_t)t * 2 * (pt.L + 1)], *bo = fo + pt.L + 1;
        if (pt.pterm[t].rsig)
            hipLaunchKernelGGL(k_rsig_phased, dim3(pt.nslots), dim3(SS_BLOCK), 0, st, d_slots, pt.nslots, t);
        for (int l = 0; l < pt.L; ++l) {
            int cnt = fo[l + 1] - fo[l];
            if (cnt > 0 && per_of(cnt, shape) == 3)
                hipLaunchKernelGGL(k_fwd_grouped, dim3(std::min((cnt + 2) / 3, SWEEP_GRID)), dim3(SWEEP_BLOCK), 0, st,
                                   d_slots, pt.d_recs + fo[l], cnt, t);
            else if (cnt > 0)
                hipLaunchKernelGGL(k_fwd_phased, dim3(std::min(cnt, SWEEP_GRID)), dim3(SWEEP_BLOCK), 0, st, d_slots,
                                   pt.d_recs + fo[l], cnt, t);
        }
        if (pt.npsplit > 0) {
            int rc = launch_prec_phased(d_slots, pt, t, st);
            if (rc) return rc;
        }
        const PhasedTasks::PassTerm &q = pt.pterm[t];
        const uint32_t *tt = pt.d_tasks + q.off;
        if (q.small_parts > 0) {
            hipLaunchKernelGGL(k_term_small, dim3(q.nsd * q.small_parts + q.nfold + q.nlight), dim3(SS_PAIR_BLOCK), 0,
                               st, d_slots, tt, q.nsd, q.small_parts, q.nfold, q.nlight, t);
        } else {
        if (q.nodo > 0)
            hipLaunchKernelGGL(k_term_odo, dim3(q.nodo), dim3(SS_BLOCK), 0, st, d_slots, tt, q.nodo, t);
        if (q.nsd > 0)
            hipLaunchKernelGGL(k_term_pair<1>, dim3(q.nsd * q.sd_parts), dim3(SS_PAIR_BLOCK), 0, st, d_slots,
                               tt + q.nodo, q.nsd, t, q.sd_parts);
        if (q.nfold > 0)
            hipLaunchKernelGGL(k_term_pair<0>, dim3(q.nfold), dim3(SS_BLOCK), 0, st, d_slots, tt + q.nodo + q.nsd,
                               q.nfold, t, 1);
        if (q.nlight > 0)
            hipLaunchKernelGGL(k_term_light, dim3(q.nlight), dim3(SS_BLOCK), 0, st, d_slots,
                               tt + q.nodo + q.nsd + q.nfold, q.nlight, t);
        }
        for (int l = pt.L - 1; l >= 0; --l) {
            int cnt = bo[l + 1] - bo[l];
            if (cnt > 0)
                hipLaunchKernelGGL(k_bwd_phased, dim3(std::min(cnt, SWEEP_GRID)), dim3(SWEEP_BLOCK), 0, st, d_slots,
                                   pt.d_recs + bo[l], cnt, t);
        }
    }
    hipLaunchKernelGGL(k_bump_phased, dim3((pt.nslots + 255) / 256), dim3(256), 0, st, d_slots, pt.nslots);
    SS_HIP(hipGetLastError());
    return SS_OK;
}

int launch_phased_expl(SlotDesc *d_slots, const PhasedTasks &pt, const int *d_want, double *scratch,
                       int scratch_blocks, hipStream_t st, bool joint) {
    const int ndec = pt.lvl_off[pt.L];
    if (ndec > 0) hipLaunchKernelGGL(k_xavg_phased, dim3(ndec), dim3(SS_BLOCK), 0, st, d_slots, pt.d_tasks, ndec, d_want);
    // every decision task: the first pass's backward sweep records
    const int *lv = pt.NPmax > 0 ? &pt.sw_off[(size_t)(pt.L + 1)] : nullptr;
    for (int l = 0; l < pt.L; ++l) {
        int cnt = lv[l + 1] - lv[l];
        if (cnt > 0)
            hipLaunchKernelGGL(k_xfwd_phased, dim3(std::min(cnt, SWEEP_GRID)), dim3(SWEEP_BLOCK), 0, st, d_slots,
                               pt.d_recs + lv[l], cnt, d_want);
    }
    for (int p = 0; p < pt.NPmax; ++p) {
        const uint32_t *tt = pt.d_tasks + pt.term_off;
        const int nsd = pt.npair_sd, nfold = pt.npair - pt.npair_sd, nlight = pt.nterm - pt.nodo - pt.npair;
        if (pt.nodo > 0)
            hipLaunchKernelGGL(k_xterm_phased, dim3(std::min(pt.nodo, scratch_blocks)), dim3(SS_BLOCK), 0, st, d_slots,
                               tt, pt.nodo, p, d_want, scratch);
        if (nsd > 0)
            hipLaunchKernelGGL(k_xterm_cls<1>, dim3(nsd), dim3(512), 0, st, d_slots, tt + pt.nodo, nsd, p, d_want);
        if (nfold > 0)
            hipLaunchKernelGGL(k_xterm_cls<2>, dim3(nfold), dim3(SS_BLOCK), 0, st, d_slots, tt + pt.nodo + nsd, nfold, p,
                               d_want);
        if (nlight > 0)
            hipLaunchKernelGGL(k_xterm_cls<0>, dim3(nlight), dim3(SS_BLOCK), 0, st, d_slots, tt + pt.nodo + pt.npair,
                               nlight, p, d_want);
        for (int l = pt.L - 1; l >= 0; --l) {
            int cnt = lv[l + 1] - lv[l];
            if (cnt > 0)
                hipLaunchKernelGGL(k_xbwd_phased, dim3(std::min(cnt, SWEEP_GRID)), dim3(SWEEP_BLOCK), 0, st, d_slots,
                                   pt.d_recs + lv[l], cnt, p, d_want);
        }
        hipLaunchKernelGGL(k_xreduce_phased, dim3(pt.nslots), dim3(SS_BLOCK), 0, st, d_slots, pt.nslots, p, d_want);
    }
    if (joint)
        hipLaunchKernelGGL(k_xjoint_phased, dim3(std::min(pt.nslots, scratch_blocks)), dim3(SS_BLOCK), 0, st, d_slots,
                           pt.nslots, d_want, scratch);
    SS_HIP(hipGetLastError());
    return SS_OK;
}

// plain streaming copy (HBM counter calibration)
__global__ void __launch_bounds__(SS_BLOCK) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int dbg_copy(int64_t bytes, int repeats, float *ms) {
    if (bytes <= 0 || bytes % 16 || repeats <= 0) return set_error(SS_ERR_ARG, "bytes must be a positive multiple of 16");
    uint4 *a = nullptr, *b = nullptr;
    SS_HIP(hipMalloc(&a, bytes));
    SS_HIP(hipMalloc(&b, bytes));
    SS_HIP(hipMemset(a, 1, bytes));
    hipEvent_t e0, e1;
    SS_HIP(hipEventCreate(&e0));
    SS_HIP(hipEventCreate(&e1));
    const size_t n = (size_t)bytes / 16;
    SS_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(k_copy, dim3(2048), dim3(SS_BLOCK), 0, 0, (const uint4 *)a, b, n);
    SS_HIP(hipGetLastError());
    SS_HIP(hipEventRecord(e1, 0));
    SS_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    SS_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return SS_OK;
}

// ---- test entry points -----------------------------------------------------
int dbg_eval7_device(const int *cards, int n, uint32_t *out) {
    int *dc = nullptr;
    uint32_t *dv = nullptr;
    SS_HIP(hipMalloc(&dc, sizeof(int) * 7 * std::max(1, n)));
    SS_HIP(hipMalloc(&dv, sizeof(uint32_t) * std::max(1, n)));
    SS_HIP(hipMemcpy(dc, cards, sizeof(int) * 7 * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_eval7, dim3(grid_for(n, 1024)), dim3(SS_BLOCK), 0, 0, dc, n, dv);
    SS_HIP(hipGetLastError());
    SS_HIP(hipMemcpy(out, dv, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    (void)hipFree(dc);
    (void)hipFree(dv);
    return SS_OK;
}

int dbg_combo_ranks_device(const int *board, uint32_t *out) {
    int *db = nullptr;
    uint32_t *dv = nullptr;
    SS_HIP(hipMalloc(&db, sizeof(int) * 5));
    SS_HIP(hipMalloc(&dv, sizeof(uint32_t) * SS_NCOMBO));
    SS_HIP(hipMemcpy(db, board, sizeof(int) * 5, hipMemcpyHostToDevice));
    int rc = launch_combo_ranks(db, dv, 0);
    if (rc) return rc;
    SS_HIP(hipMemcpy(out, dv, sizeof(uint32_t) * SS_NCOMBO, hipMemcpyDeviceToHost));
    (void)hipFree(db);
    (void)hipFree(dv);
    return SS_OK;
}

int dbg_eval_census_device(unsigned long long counts[10], unsigned long long *sum) {
    unsigned long long *dc = nullptr;
    SS_HIP(hipMalloc(&dc, sizeof(unsigned long long) * 11));
    SS_HIP(hipMemset(dc, 0, sizeof(unsigned long long) * 11));
    // one block per (c0, c1) pair: 1326 blocks
    hipLaunchKernelGGL(k_eval_census, dim3(SS_NCOMBO), dim3(64), 0, 0, dc, dc + 10);
    SS_HIP(hipGetLastError());
    unsigned long long h[11];
    SS_HIP(hipMemcpy(h, dc, sizeof(h), hipMemcpyDeviceToHost));
    for (int i = 0; i < 10; ++i) counts[i] = h[i];
    *sum = h[10];
    (void)hipFree(dc);
    return SS_OK;
}

}  // namespace ss

// ss_flop.hip -- the 2-player flop-through-river solve with chance nodes
// (include/flop_solve.h): the tree builder on the host and Linear CFR over
// the materialised tree on the GPU.
//
// Semantics follow src/cuda/flop_solve.cu:
//   - a chance node copies the opponent's reach to each dealt-card child and
//     averages the children's counterfactual values (:543-555, :591-607);
//     cards are blocked only at the terminals below it;
//   - fold: starting_pot / 2 plus the loser's bets of the street, times the
//     opponent reach of hands disjoint from the traverser's hand and the
//     node's board (:737-779; the traverser's own board conflicts are not
//     checked);
//   - showdown: +-pot / 2 per opponent hand ranked below / above, 0 for ties,
//     blocked hands (rank 0) and a blocked traverser (:781-833);
//   - each traverser pass regret-matches every decision node, then propagates
//     the opponent's reach down the levels and the values up them, updating
//     the traverser's regrets R <- (R + dv) t / (t + 1) and strategy sums
//     SS += t sigma (:835-891, the iteration loop :1150-1228).
//
// The GPU form. Every per-hand array is [row][Mp] (Mp = the larger hand count
// rounded to 64, one wavefront per 64 hands). One traverser pass is: the
// root's reach, one launch per level down, the fold and showdown launches,
// one launch per level up; the whole iteration is captured once in a hipGraph
// and replayed, with the iteration counter on the device. Regret matching is
// folded into the sweeps: a decision's strategy is computed where it is used
// (the opponent's nodes on the way down, the traverser's on the way up, where
// its regrets are still those of the pass's start), so no strategy array is
// stored. The terminals are (hand x opponent hand) sums in the reference's
// order: at setup every distinct showdown board gets a table of
// 1 + sign(rank difference) codes per (traverser hand, opponent hand), one
// byte each, 0 / 1 card-disjointness codes serve the folds, and each lane of
// a terminal workgroup keeps its hand's column of the table in registers for
// up to 32 nodes, so a hand pair costs a fraction of a byte-to-float
// conversion and one fused multiply-add whose zero terms leave the sum
// unchanged: the results are the reference loop's bit for bit.
#include "ss_internal.h"
#include "ss_eval.h"
#include "../../include/flop_solve.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <vector>

namespace ss {

// ---- tree builder (host) ----------------------------------------------------
namespace {

constexpr int FS_RAISE_CAP = 1;   // raises per street (flop_solve.cu:25)

struct FsBuilder {
    std::vector<FSNode> nodes;
    std::vector<int> children;

    int add(int type, int player, int pot, int b0, int b1, const int *board, int nb) {
        FSNode n;
        memset(&n, 0, sizeof(n));
        n.type = type;
        n.player = player;
        n.pot = pot;
        n.bets[0] = b0;
        n.bets[1] = b1;
        for (int i = 0; i < nb; ++i) n.board_cards[i] = board[i];
        n.num_board = nb;
        nodes.push_back(n);
        return (int)nodes.size() - 1;
    }
    void set_children(int node, const std::vector<int> &kids) {
        nodes[node].first_child = (int)children.size();
        nodes[node].num_children = (int)kids.size();
        children.insert(children.end(), kids.begin(), kids.end());
    }

    // An all-in by `player`: the opponent folds (the all-in player wins) or
    // calls, which ends the street.
    int all_in(bool river, int player, int new_pot, int b0, int b1, int to_call, int amount, const int *board,
               int nb) {
        const int ai = add(FS_NODE_DECISION, 1 - player, new_pot, b0, b1, board, nb);
        const int fold = add(FS_NODE_FOLD, player, new_pot, b0, b1, board, nb);
        const int c0 = player == 0 ? b0 : b1, c1 = c0;   // both matched at the all-in amount
        const int end = add(river ? FS_NODE_SHOWDOWN : FS_NODE_CHANCE, -1, new_pot + (amount - to_call), c0, c1,
                            board, nb);
        set_children(ai, {fold, end});
        return ai;
    }

    // One street's betting from `player` to act (flop_solve.cu:72-265): the
    // street ends with two actions taken and the bets level, or on a call of
    // a bet; node ids are assigned parent first, in the order the actions
    // are listed (fold, check / call, bets by size, all-in), and a node's
    // children are listed once its whole subtree is built.
    int street(bool river, int player, int pot, int stack, int b0, int b1, int raises, int acted,
               const std::vector<float> &sizes, const int *board, int nb) {
        const int to_call = std::max(0, player == 0 ? b1 - b0 : b0 - b1);
        const int end_type = river ? FS_NODE_SHOWDOWN : FS_NODE_CHANCE;
        if (acted >= 2 && b0 == b1) return add(end_type, -1, pot, b0, b1, board, nb);
        const int node = add(FS_NODE_DECISION, player, pot, b0, b1, board, nb);
        std::vector<int> kids;
        if (to_call > 0) kids.push_back(add(FS_NODE_FOLD, 1 - player, pot, b0, b1, board, nb));
        if (to_call == 0) {
            kids.push_back(street(river, 1 - player, pot, stack, b0, b1, raises, acted + 1, sizes, board, nb));
        } else {
            const int m0 = player == 0 ? b1 : b0, m1 = player == 0 ? b1 : b0;   // matched
            if (acted >= 1)
                kids.push_back(add(end_type, -1, pot + to_call, m0, m1, board, nb));
            else
                kids.push_back(street(river, 1 - player, pot + to_call, stack - to_call, m0, m1, raises, acted + 1,
                                      sizes, board, nb));
        }
        if (raises < FS_RAISE_CAP) {
            bool shoved = false;
            for (float f : sizes) {
                int amount = to_call == 0 ? (int)(f * pot) : to_call + (int)(f * (pot + to_call));
                if (amount >= stack) amount = stack;
                if (amount <= to_call) continue;
                const int n0 = player == 0 ? b0 + amount : b0, n1 = player == 0 ? b1 : b1 + amount;
                if (amount >= stack) {
                    if (shoved) continue;
                    shoved = true;
                    kids.push_back(all_in(river, player, pot + amount, n0, n1, to_call, amount, board, nb));
                } else {
                    kids.push_back(street(river, 1 - player, pot + amount, stack - amount + to_call, n0, n1,
                                          raises + 1, acted + 1, sizes, board, nb));
                }
            }
            if (!shoved && stack > to_call && (int)kids.size() < 7) {
                const int n0 = player == 0 ? b0 + stack : b0, n1 = player == 0 ? b1 : b1 + stack;
                kids.push_back(all_in(river, player, pot + stack, n0, n1, to_call, stack, board, nb));
            }
        }
        set_children(node, kids);
        return node;
    }

    // Every chance node not yet expanded (board short of 5 cards) gets one
    // child per card not on its board, ascending, each the next street's
    // betting with the given sizes from player 0, bets level at 0 and a stack
    // of 10000 minus the node's player-0 bet (flop_solve.cu:269-317). The
    // child slots of a chance node are reserved before its subtrees are built.
    void deal(const std::vector<float> &sizes) {
        const int count = (int)nodes.size();
        std::vector<int> todo;
        for (int i = 0; i < count; ++i)
            if (nodes[i].type == FS_NODE_CHANCE && nodes[i].num_board < 5 && nodes[i].num_children == 0)
                todo.push_back(i);
        for (int idx : todo) {
            const FSNode ch = nodes[idx];   // copy: `nodes` grows below
            bool on_board[52] = {};
            for (int b = 0; b < ch.num_board; ++b) on_board[ch.board_cards[b]] = true;
            std::vector<int> cards;
            for (int c = 0; c < 52; ++c)
                if (!on_board[c]) cards.push_back(c);
            const int first = (int)children.size();
            children.resize(children.size() + cards.size());
            const bool river = ch.num_board + 1 == 5;
            for (size_t k = 0; k < cards.size(); ++k) {
                int board[5];
                for (int b = 0; b < ch.num_board; ++b) board[b] = ch.board_cards[b];
                board[ch.num_board] = cards[k];
                const int kid = street(river, 0, ch.pot, std::max(0, 10000 - ch.bets[0]), 0, 0, 0, 0, sizes, board,
                                       ch.num_board + 1);
                children[first + k] = kid;
            }
            nodes[idx].first_child = first;
            nodes[idx].num_children = (int)cards.size();
        }
    }
};

template <typename T> T *copy_out(const std::vector<T> &v) {
    T *p = (T *)malloc(sizeof(T) * std::max<size_t>(1, v.size()));
    if (p && !v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

// ---- device ------------------------------------------------------------------
struct FsDev {
    int N, Mp, nh[2];
    int starting_pot;
    int no16;                  // opponent hands in groups of 16: ceil(max hands / 16)
    const int4 *node;          // [N] x = type | player << 4 (player + 1), y = nchild, z = first_child, w = pot
    const int2 *bets;          // [N]
    const int *children;
    const int *abase;          // [N] first action row of a decision node, -1 otherwise
    const uint32_t *board;     // [N] 5 card slots of 6 bits each, 63 in the slots past the node's board
    const uint32_t *hands;     // [2][Mp] c0 | c1 << 8 (0xFFFF past the hand count)
    const float *weights;      // [2][Mp]
    float *R, *SS;             // [A][Mp]
    float *reach, *cfv;        // [N][Mp]
    float *root_sigma;         // [FS_MAX_ACTIONS][Mp] the root's strategy as last computed
    int *iter;                 // the iteration being run
};

__device__ __forceinline__ int fs_type(int4 n) { return n.x & 15; }
__device__ __forceinline__ int fs_player(int4 n) { return (n.x >> 4) - 1; }
__device__ __forceinline__ bool fs_on_board(uint32_t bw, int c) {
    bool on = false;
    for (int b = 0; b < 5; ++b) on |= (int)((bw >> (6 * b)) & 63) == c;   // an empty slot (63) matches no card
    return on;
}
__device__ __forceinline__ bool fs_disjoint(uint32_t a, uint32_t b) {
    const uint32_t a0 = a & 255, a1 = (a >> 8) & 255, b0 = b & 255, b1 = (b >> 8) & 255;
    return a0 != b0 && a0 != b1 && a1 != b0 && a1 != b1;
}

// Regret matching of one hand at one node (fs_batch_regret_match,
// flop_solve.cu:657-690): positive parts times the reciprocal of their sum,
// uniform when none is positive.
__device__ __forceinline__ void fs_match(const float *R, int Mp, int na, float *sg) {
    float r[FS_MAX_ACTIONS];
    float sum = 0.f;
#pragma unroll
    for (int a = 0; a < FS_MAX_ACTIONS; ++a)
        if (a < na) {
            r[a] = R[(size_t)a * Mp];
            sum += r[a] > 0.f ? r[a] : 0.f;
        }
    if (sum > 0.f) {
        const float inv = 1.0f / sum;
#pragma unroll
        for (int a = 0; a < FS_MAX_ACTIONS; ++a)
            if (a < na) sg[a] = (r[a] > 0.f ? r[a] : 0.f) * inv;
    } else {
        const float u = 1.0f / (float)na;
#pragma unroll
        for (int a = 0; a < FS_MAX_ACTIONS; ++a)
            if (a < na) sg[a] = u;
    }
}

// ranks of both players' hands on every distinct showdown board (0: blocked by it)
__global__ void __launch_bounds__(SS_BLOCK) k_fs_rank(FsDev d, const uint32_t *boards, int nboards,
                                                      uint32_t *rank) {
    __shared__ uint8_t tab[8192];
    fill_straight_table(tab, threadIdx.x, blockDim.x);
    __syncthreads();
    const int total = nboards * 2 * d.Mp;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int h = e % d.Mp, p = (e / d.Mp) % 2, b = e / d.Mp / 2;
        uint32_t r = 0;
        if (h < d.nh[p]) {
            const uint32_t bw = boards[b], hc = d.hands[p * d.Mp + h];
            const int c0 = hc & 255, c1 = (hc >> 8) & 255;
            if (!fs_on_board(bw, c0) && !fs_on_board(bw, c1)) {
                int c[7];
                for (int k = 0; k < 5; ++k) c[k] = (int)((bw >> (6 * k)) & 63);
                c[5] = c0;
                c[6] = c1;
                r = eval7_lut(c, tab);
            }
        }
        rank[e] = r;
    }
}

// Showdown codes of every board and traverser: byte k of word (c, h) holds
// 1 + sign(rank_h - rank_o) for opponent hand o = 16 c + k, 1 (no term) for a
// tie, a blocked hand, shared cards or a hand past the counts.
__global__ void __launch_bounds__(SS_BLOCK) k_fs_wmat(FsDev d, const uint32_t *rank, int nboards, uint4 *W) {
    const int per = d.no16 * d.Mp, total = nboards * 2 * per;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int h = e % d.Mp, c = (e / d.Mp) % d.no16, t = (e / per) % 2, b = e / per / 2, opp = 1 - t;
        uint32_t w[4] = {0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u};
        if (h < d.nh[t]) {
            const uint32_t rt = rank[((size_t)b * 2 + t) * d.Mp + h], hc = d.hands[t * d.Mp + h];
            for (int k = 0; k < 16; ++k) {
                const int o = 16 * c + k;
                if (o >= d.nh[opp] || !rt) continue;
                const uint32_t ro = rank[((size_t)b * 2 + opp) * d.Mp + o];
                if (!ro || !fs_disjoint(hc, d.hands[opp * d.Mp + o])) continue;
                const uint32_t code = 1u + (rt > ro) - (rt < ro);
                w[k >> 2] = (w[k >> 2] & ~(255u << (8 * (k & 3)))) | code << (8 * (k & 3));
            }
        }
        W[e] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Fold codes of every distinct fold board and traverser: byte k of word (c, h)
// is 1 when opponent hand o = 16 c + k shares no card with hand h and none
// with the board (the traverser's own board cards are not checked), else 0.
__global__ void __launch_bounds__(SS_BLOCK) k_fs_fmat(FsDev d, const uint32_t *boards, int nboards, uint4 *F) {
    const int per = d.no16 * d.Mp, total = nboards * 2 * per;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int h = e % d.Mp, c = (e / d.Mp) % d.no16, t = (e / per) % 2, b = e / per / 2, opp = 1 - t;
        uint32_t w[4] = {0, 0, 0, 0};
        if (h < d.nh[t])
            for (int k = 0; k < 16; ++k) {
                const int o = 16 * c + k;
                if (o >= d.nh[opp]) continue;
                const uint32_t oc = d.hands[opp * d.Mp + o];
                if (fs_disjoint(d.hands[t * d.Mp + h], oc) && !fs_on_board(boards[b], oc & 255) &&
                    !fs_on_board(boards[b], (oc >> 8) & 255))
                    w[k >> 2] |= 1u << (8 * (k & 3));
            }
        F[e] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// the opponent's reach at the root: the raw input weights
__global__ void __launch_bounds__(SS_BLOCK) k_fs_root(FsDev d, int opp) {
    for (int h = threadIdx.x; h < d.nh[opp]; h += blockDim.x) d.reach[h] = d.weights[opp * d.Mp + h];
}

// A sweep record: everything one (node, hand segment) item of a level sweep
// needs, in one 64-byte scalar load: node, type | (player + 1) << 4 |
// children << 8, the first action row (decision) or the first child slot
// (chance), and up to six children inline (decision).
constexpr int FS_REC = 16;

// Regret matching from regrets already in registers (the arithmetic of
// fs_match).
__device__ __forceinline__ void fs_match_regs(const float *r, int na, float *sg) {
    float sum = 0.f;
#pragma unroll
    for (int a = 0; a < FS_MAX_ACTIONS; ++a)
        if (a < na) sum += r[a] > 0.f ? r[a] : 0.f;
    if (sum > 0.f) {
        const float inv = 1.0f / sum;
#pragma unroll
        for (int a = 0; a < FS_MAX_ACTIONS; ++a)
            if (a < na) sg[a] = (r[a] > 0.f ? r[a] : 0.f) * inv;
    } else {
        const float u = 1.0f / (float)na;
#pragma unroll
        for (int a = 0; a < FS_MAX_ACTIONS; ++a)
            if (a < na) sg[a] = u;
    }
}

// The opponent's reach down one level (fs_batch_propagate_reach,
// flop_solve.cu:693-735): G wavefronts per node, S segments of 64 hands per
// lane (their loads in flight together); the opponent's own decisions scale
// by its current strategy, regret-matched here.
template <int S, int G>
__global__ void __launch_bounds__(SS_BLOCK) k_fs_down(FsDev d, const int *__restrict__ recs, int nrec, int opp,
                                                      const float *__restrict__ R, float *__restrict__ reach,
                                                      float *__restrict__ root_sigma) {
    const int lane = threadIdx.x & 63, nh = d.nh[opp], Mp = d.Mp;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = blockIdx.x * 4 + w; i < nrec * G; i += gridDim.x * 4) {
        const int ni = G == 1 ? i : i / G, hb = G == 1 ? 0 : (i - ni * G) * S * 64;
        const int *rc = recs + (size_t)ni * FS_REC;
        const int n = rc[0], meta = rc[1], na = meta >> 8;
        float pr[S];
#pragma unroll
        for (int s = 0; s < S; ++s)
            if (hb + s * 64 + lane < nh) pr[s] = reach[(size_t)n * Mp + hb + s * 64 + lane];
        if ((meta & 15) == FS_NODE_DECISION) {
            if (((meta >> 4) & 15) - 1 == opp) {
                float r[S][FS_MAX_ACTIONS];
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int a = 0; a < FS_MAX_ACTIONS; ++a)
                        if (a < na && hb + s * 64 + lane < nh) r[s][a] = R[(size_t)(rc[2] + a) * Mp + hb + s * 64 + lane];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int h = hb + s * 64 + lane;
                    if (h >= nh) continue;
                    float sg[FS_MAX_ACTIONS];
                    fs_match_regs(r[s], na, sg);
#pragma unroll
                    for (int a = 0; a < FS_MAX_ACTIONS; ++a)
                        if (a < na) {
                            reach[(size_t)rc[3 + a] * Mp + h] = pr[s] * sg[a];
                            if (n == 0) root_sigma[a * Mp + h] = sg[a];
                        }
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int a = 0; a < FS_MAX_ACTIONS; ++a)
                        if (a < na && hb + s * 64 + lane < nh) reach[(size_t)rc[3 + a] * Mp + hb + s * 64 + lane] = pr[s];
            }
        } else {
            const int *kid = d.children + rc[2];
            for (int a = 0; a < na; ++a) {
                const size_t row = (size_t)kid[a] * Mp;
#pragma unroll
                for (int s = 0; s < S; ++s)
                    if (hb + s * 64 + lane < nh) reach[row + hb + s * 64 + lane] = pr[s];
            }
        }
    }
}

// Values up one level (fs_batch_propagate_cfv, flop_solve.cu:835-891): at
// the traverser's decisions the strategy of this pass (its regrets are still
// those of the pass's start), the value, the regret and strategy-sum updates;
// chance nodes average their children (their loads eight children at a time,
// the sum in child order).
template <int S, int G>
__global__ void __launch_bounds__(SS_BLOCK) k_fs_up(FsDev d, const int *__restrict__ recs, int nrec, int trav,
                                                    float *__restrict__ R, float *__restrict__ SS,
                                                    float *__restrict__ cfv, float *__restrict__ root_sigma) {
    const int lane = threadIdx.x & 63, nh = d.nh[trav], Mp = d.Mp;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = *d.iter;
    const float disc = (float)t / ((float)t + 1.0f), tw = (float)t;
    for (int i = blockIdx.x * 4 + w; i < nrec * G; i += gridDim.x * 4) {
        const int ni = G == 1 ? i : i / G, hb = G == 1 ? 0 : (i - ni * G) * S * 64;
        const int *rc = recs + (size_t)ni * FS_REC;
        const int n = rc[0], meta = rc[1], na = meta >> 8;
        if ((meta & 15) == FS_NODE_DECISION) {
            const bool mine = ((meta >> 4) & 15) - 1 == trav;
            float cv[S][FS_MAX_ACTIONS], r[S][FS_MAX_ACTIONS], ss[S][FS_MAX_ACTIONS];
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
                for (int a = 0; a < FS_MAX_ACTIONS; ++a)
                    if (a < na && hb + s * 64 + lane < nh) {
                        cv[s][a] = cfv[(size_t)rc[3 + a] * Mp + hb + s * 64 + lane];
                        if (mine) {
                            r[s][a] = R[(size_t)(rc[2] + a) * Mp + hb + s * 64 + lane];
                            ss[s][a] = SS[(size_t)(rc[2] + a) * Mp + hb + s * 64 + lane];
                        }
                    }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int h = hb + s * 64 + lane;
                if (h >= nh) continue;
                float val = 0.f;
                if (mine) {
                    float sg[FS_MAX_ACTIONS];
                    fs_match_regs(r[s], na, sg);
#pragma unroll
                    for (int a = 0; a < FS_MAX_ACTIONS; ++a)
                        if (a < na) val = fmaf(sg[a], cv[s][a], val);
#pragma unroll
                    for (int a = 0; a < FS_MAX_ACTIONS; ++a)
                        if (a < na) {
                            const size_t row = (size_t)(rc[2] + a) * Mp + h;
                            R[row] = (r[s][a] + (cv[s][a] - val)) * disc;
                            SS[row] = fmaf(tw, sg[a], ss[s][a]);
                            if (n == 0) root_sigma[a * Mp + h] = sg[a];
                        }
                } else {
#pragma unroll
                    for (int a = 0; a < FS_MAX_ACTIONS; ++a)
                        if (a < na) val += cv[s][a];
                }
                cfv[(size_t)n * Mp + h] = val;
            }
        } else {
            const int *kid = d.children + rc[2];
            float acc[S];
#pragma unroll
            for (int s = 0; s < S; ++s) acc[s] = 0.f;
            for (int a0 = 0; a0 < na; a0 += 8) {
                float v[8][S];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (a0 + k < na) {
                        const size_t row = (size_t)kid[a0 + k] * Mp;
#pragma unroll
                        for (int s = 0; s < S; ++s)
                            if (hb + s * 64 + lane < nh) v[k][s] = cfv[row + hb + s * 64 + lane];
                    }
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (a0 + k < na)
#pragma unroll
                        for (int s = 0; s < S; ++s) acc[s] += v[k][s];
            }
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (hb + s * 64 + lane < nh) cfv[(size_t)n * Mp + hb + s * 64 + lane] = acc[s] / (float)na;
        }
    }
}

// 64 fused multiply-adds acc[m] = fma(bcast_k(rv[m]), w[k], acc[m]) for
// k = 0..15 (outer) and m = 0..3: rv[m] holds 16 opponent values per 16-lane
// row (lane l: value l % 16), and row_newbcast:k hands every lane of a row
// that row's lane k, so one instruction both broadcasts the opponent value and
// accumulates it. Written as one asm block because the compiler does not fold
// a broadcast into a multiply-add; the leading s_nop covers the DPP read
// hazards after a VALU write of rv or of EXEC. All 64 lanes must be active.
#define FS_DPP_STEP(k, w)  \
    "v_fmac_f32_dpp %0, %4, %" #w " row_newbcast:" #k " row_mask:0xf bank_mask:0xf\n" \
    "v_fmac_f32_dpp %1, %5, %" #w " row_newbcast:" #k " row_mask:0xf bank_mask:0xf\n" \
    "v_fmac_f32_dpp %2, %6, %" #w " row_newbcast:" #k " row_mask:0xf bank_mask:0xf\n" \
    "v_fmac_f32_dpp %3, %7, %" #w " row_newbcast:" #k " row_mask:0xf bank_mask:0xf\n"

__device__ __forceinline__ void fs_dpp_fma16(float (&acc)[4], const float (&rv)[4], const float (&w)[16]) {
    asm volatile("s_nop 4\n" FS_DPP_STEP(0, 8) FS_DPP_STEP(1, 9) FS_DPP_STEP(2, 10) FS_DPP_STEP(3, 11)
                     FS_DPP_STEP(4, 12) FS_DPP_STEP(5, 13) FS_DPP_STEP(6, 14) FS_DPP_STEP(7, 15)
                         FS_DPP_STEP(8, 16) FS_DPP_STEP(9, 17) FS_DPP_STEP(10, 18) FS_DPP_STEP(11, 19)
                             FS_DPP_STEP(12, 20) FS_DPP_STEP(13, 21) FS_DPP_STEP(14, 22) FS_DPP_STEP(15, 23)
                 : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3])
                 : "v"(rv[0]), "v"(rv[1]), "v"(rv[2]), "v"(rv[3]), "v"(w[0]), "v"(w[1]), "v"(w[2]), "v"(w[3]),
                   "v"(w[4]), "v"(w[5]), "v"(w[6]), "v"(w[7]), "v"(w[8]), "v"(w[9]), "v"(w[10]), "v"(w[11]),
                   "v"(w[12]), "v"(w[13]), "v"(w[14]), "v"(w[15]));
}
#undef FS_DPP_STEP

// Terminal values, fold (FOLD) or showdown, for one traverser. The terminal
// nodes of a board are in groups of four (groups[2 g] = the nodes, a short
// group repeating its last; groups[2 g + 1] = {nodes in the group, pot}; the
// four nodes of a showdown group have one pot), and a task is at most eight
// groups of one board (tasks[j] = table, first group, groups), so the tasks
// outnumber the resident workgroups and balance by themselves. A workgroup
// is Mp lanes, one traverser hand each; a lane loads its column of the
// board's codes for this traverser ([no16][Mp] uint4, 16 opponent hands per
// word) once per task and keeps it in registers for every group of the task.
// Per opponent hand o, in the reference's order:
//   fold      acc = fma(r_o, code, acc), code 1 / 0: flop_solve.cu:737-779
//             adds r_o or skips it; the value is acc times the payoff
//   showdown  acc = fma(r_o, c pot/2, acc), c = code - 1 = +1 / -1 / 0:
//             +-r_o pot/2 or nothing, flop_solve.cu:781-833
// The product r_o * code (or r_o * c pot/2) is exact, and a zero term leaves
// acc unchanged, so the sums are the reference's bit for bit.
constexpr int FS_NO16_MAX = (FS_MAX_HANDS + 15) / 16;
constexpr int FS_TASK_GROUPS = 8;
// Four wavefronts per SIMD (128 VGPRs): the register column and the
// unrolled opponent loop use what the cap leaves; 3 or 5 measured slower.
template <bool FOLD>
__global__ void __launch_bounds__(SS_BLOCK, 4) k_fs_term(FsDev d, const int4 *__restrict__ tasks, int ntasks,
                                                        const int4 *__restrict__ groups,
                                                        const uint4 *__restrict__ tables,
                                                        const float *__restrict__ reach, float *__restrict__ cfv,
                                                        int trav) {
    const int n_trav = d.nh[trav], Mp = d.Mp, no16 = d.no16;
    const int h = threadIdx.x, l16 = threadIdx.x & 15;   // blockDim.x == Mp
    const float half_start = d.starting_pot * 0.5f;
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        const int4 tk = tasks[j];
        const uint4 *src = tables + ((size_t)tk.x * 2 + trav) * no16 * Mp + h;
        uint4 q[FS_NO16_MAX];
#pragma unroll
        for (int c = 0; c < FS_NO16_MAX; ++c)
            if (c < no16) q[c] = src[(size_t)c * Mp];
        for (int g = tk.y; g < tk.y + tk.z; ++g) {
            const int4 gn = groups[2 * g], gm = groups[2 * g + 1];
            const int nn[4] = {gn.x, gn.y, gn.z, gn.w};
            float scale[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                if (FOLD) {
                    const int winner = (d.node[nn[m]].x >> 4) - 1, loser = 1 - winner;
                    const int2 bt = d.bets[nn[m]];
                    scale[m] = trav == winner ? half_start + (float)(loser == 0 ? bt.x : bt.y)
                                              : -(half_start + (float)(trav == 0 ? bt.x : bt.y));
                } else {
                    scale[m] = 1.f;
                }
            }
            const float hp = gm.y * 0.5f;
            float acc[4] = {0.f, 0.f, 0.f, 0.f}, rv[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) rv[m] = reach[(size_t)nn[m] * Mp + l16];
#pragma unroll
            for (int c = 0; c < FS_NO16_MAX; ++c)
                if (c < no16) {
                    const int cn = min(c + 1, no16 - 1);
                    float rn[4], w[16];
#pragma unroll
                    for (int m = 0; m < 4; ++m) rn[m] = reach[(size_t)nn[m] * Mp + 16 * cn + l16];
                    const uint32_t qw[4] = {q[c].x, q[c].y, q[c].z, q[c].w};
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        const float code = (float)((qw[k >> 2] >> (8 * (k & 3))) & 255u);
                        w[k] = FOLD ? code : fmaf(code, hp, -hp);
                    }
                    fs_dpp_fma16(acc, rv, w);
#pragma unroll
                    for (int m = 0; m < 4; ++m) rv[m] = rn[m];
                }
            if (h < n_trav)
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    if (m < gm.x) cfv[(size_t)nn[m] * Mp + h] = FOLD ? acc[m] * scale[m] : acc[m];
        }
    }
}

__global__ void k_fs_next_iter(int *iter) {
    if (threadIdx.x == 0) *iter += 1;
}

// ---- host solve --------------------------------------------------------------
namespace {

struct FsRun {
    std::vector<void *> bufs;
    FsDev d{};
    int A = 0;
    std::vector<int> dec;
    // Fold or showdown nodes grouped by board: distinct boards, the nodes
    // sorted by board, one task per board (board, first, count), and on the
    // device the same plus the per-board code tables.
    struct TermSet {
        std::vector<uint32_t> boards;
        std::vector<int4> tasks, groups;
        uint32_t *d_boards = nullptr;
        int4 *d_tasks = nullptr, *d_groups = nullptr;
        uint4 *d_tab = nullptr;
        // by_pot: a group's four nodes share a pot (showdowns)
        // tasks of at most `most` groups each, the largest first
        void split(int most) {
            std::vector<int4> out;
            for (const int4 &t : tasks)
                for (int o = 0; o < t.z; o += most) out.push_back(make_int4(t.x, t.y + o, std::min(most, t.z - o), 0));
            std::stable_sort(out.begin(), out.end(), [](const int4 &a, const int4 &b) { return a.z > b.z; });
            tasks.swap(out);
        }
        void group(const std::vector<int> &nodes, const std::vector<uint32_t> &board, const std::vector<int4> &node,
                   bool by_pot) {
            std::vector<std::tuple<uint32_t, int, int>> keyed;   // board, pot (or 0), node
            for (int n : nodes) keyed.emplace_back(board[n], by_pot ? node[n].w : 0, n);
            std::sort(keyed.begin(), keyed.end());
            for (size_t i = 0; i < keyed.size();) {
                size_t e = i;
                while (e < keyed.size() && std::get<0>(keyed[e]) == std::get<0>(keyed[i])) ++e;
                const int first = (int)groups.size() / 2;
                for (size_t f = i; f < e;) {
                    size_t g = f;
                    while (g < e && g - f < 4 && std::get<1>(keyed[g]) == std::get<1>(keyed[f])) ++g;
                    int nn[4];
                    for (int m = 0; m < 4; ++m) nn[m] = std::get<2>(keyed[std::min(f + m, g - 1)]);
                    groups.push_back(make_int4(nn[0], nn[1], nn[2], nn[3]));
                    groups.push_back(make_int4((int)(g - f), std::get<1>(keyed[f]), 0, 0));
                    f = g;
                }
                tasks.push_back(make_int4((int)boards.size(), first, (int)groups.size() / 2 - first, 0));
                boards.push_back(std::get<0>(keyed[i]));
                i = e;
            }
        }
    } sdset, foldset;
    std::vector<std::vector<int>> lvl;   // decision and chance nodes per depth
    std::vector<int *> d_lvl;
    std::vector<int> abase;
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~FsRun() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *p : bufs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    template <typename T> int alloc(T **p, size_t n, const T *src = nullptr) {
        SS_HIP(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(1, n)));
        bufs.push_back(*p);
        if (src && n) SS_HIP(hipMemcpyAsync(*p, src, sizeof(T) * n, hipMemcpyHostToDevice, st));
        else SS_HIP(hipMemsetAsync(*p, 0, sizeof(T) * std::max<size_t>(1, n), st));
        return SS_OK;
    }

    int setup(const FSTreeData *td) {
        const int N = td->num_nodes;
        if (N <= 0 || !td->nodes || !td->children) return set_error(SS_ERR_ARG, "flop solve: empty tree");
        for (int p = 0; p < 2; ++p)
            if (td->num_hands[p] < 1 || td->num_hands[p] > FS_MAX_HANDS)
                return set_error(SS_ERR_ARG, "flop solve: player %d has %d hands (1..%d)", p, td->num_hands[p],
                                 FS_MAX_HANDS);
        const int maxh = std::max(td->num_hands[0], td->num_hands[1]);
        const int Mp = (maxh + 63) / 64 * 64;
        SS_HIP(hipStreamCreate(&st));
        // host tables
        std::vector<int4> node(N);
        std::vector<int2> bets(N);
        std::vector<uint32_t> board(N);
        std::vector<int> depth(N, -1), sd, fold;
        abase.assign(N, -1);
        const int total_kids = td->num_children_total;
        for (int n = 0; n < N; ++n) {
            const FSNode &f = td->nodes[n];
            if (f.type < 0 || f.type > 3 || f.num_children < 0 || (f.num_children > 0 && (f.first_child < 0 ||
                f.first_child + f.num_children > total_kids)) || f.num_board < 0 || f.num_board > 5)
                return set_error(SS_ERR_ARG, "flop solve: malformed node %d", n);
            if (f.type == FS_NODE_DECISION && (f.num_children < 1 || f.num_children > FS_MAX_ACTIONS ||
                                               f.player < 0 || f.player > 1))
                return set_error(SS_ERR_ARG, "flop solve: decision node %d: %d actions, player %d", n,
                                 f.num_children, f.player);
            if (f.type == FS_NODE_CHANCE && f.num_children < 1)
                return set_error(SS_ERR_ARG, "flop solve: chance node %d has no children", n);
            if ((f.type == FS_NODE_FOLD || f.type == FS_NODE_SHOWDOWN) && f.num_children != 0)
                return set_error(SS_ERR_ARG, "flop solve: terminal node %d has children", n);
            if (f.type == FS_NODE_FOLD && (f.player < 0 || f.player > 1))
                return set_error(SS_ERR_ARG, "flop solve: fold node %d has winner %d", n, f.player);
            if (f.type == FS_NODE_SHOWDOWN && f.num_board != 5)
                return set_error(SS_ERR_ARG, "flop solve: showdown node %d has %d board cards", n, f.num_board);
            node[n] = make_int4(f.type | (f.player + 1) << 4, f.num_children, f.first_child, f.pot);
            bets[n] = make_int2(f.bets[0], f.bets[1]);
            uint32_t bw = 0;
            for (int b = 0; b < 5; ++b) {
                if (b < f.num_board && (f.board_cards[b] < 0 || f.board_cards[b] > 51))
                    return set_error(SS_ERR_ARG, "flop solve: node %d board card %d", n, f.board_cards[b]);
                bw |= (uint32_t)(b < f.num_board ? f.board_cards[b] : 63) << (6 * b);
            }
            board[n] = bw;
            if (f.type == FS_NODE_DECISION) {
                abase[n] = A;
                A += f.num_children;
                dec.push_back(n);
            }
            if (f.type == FS_NODE_SHOWDOWN) sd.push_back(n);
            if (f.type == FS_NODE_FOLD) fold.push_back(n);
        }
        // depths from the root (breadth first); every node must hang off it
        std::vector<int> queue{0};
        depth[0] = 0;
        for (size_t q = 0; q < queue.size(); ++q) {
            const int n = queue[q];
            const FSNode &f = td->nodes[n];
            for (int a = 0; a < f.num_children; ++a) {
                const int k = td->children[f.first_child + a];
                if (k <= 0 || k >= N || depth[k] >= 0)
                    return set_error(SS_ERR_ARG, "flop solve: child %d of node %d is not a tree edge", k, n);
                depth[k] = depth[n] + 1;
                queue.push_back(k);
            }
        }
        if ((int)queue.size() != N)
            return set_error(SS_ERR_ARG, "flop solve: %d nodes off the root's tree", N - (int)queue.size());
        int L = 0;
        for (int n = 0; n < N; ++n) L = std::max(L, depth[n] + 1);
        lvl.assign(L, {});
        for (int n : queue)
            if (td->nodes[n].type == FS_NODE_DECISION || td->nodes[n].type == FS_NODE_CHANCE)
                lvl[depth[n]].push_back(n);
        // terminals grouped by board (the nodes of one river betting tree
        // share it)
        sdset.group(sd, board, node, true);
        foldset.group(fold, board, node, false);
        sdset.split(FS_TASK_GROUPS);
        foldset.split(FS_TASK_GROUPS);
        std::vector<uint32_t> hands(2 * Mp, 0xFFFFu);
        std::vector<float> w(2 * Mp, 0.f);
        for (int p = 0; p < 2; ++p)
            for (int h = 0; h < td->num_hands[p]; ++h) {
                const int c0 = td->hands[p][h][0], c1 = td->hands[p][h][1];
                if (c0 < 0 || c0 > 51 || c1 < 0 || c1 > 51 || c0 == c1)
                    return set_error(SS_ERR_ARG, "flop solve: player %d hand %d has cards %d, %d", p, h, c0, c1);
                hands[p * Mp + h] = (uint32_t)c0 | (uint32_t)c1 << 8;
                w[p * Mp + h] = td->weights[p][h];
            }
        // device
        d.N = N;
        d.Mp = Mp;
        d.nh[0] = td->num_hands[0];
        d.nh[1] = td->num_hands[1];
        d.no16 = (maxh + 15) / 16;
        d.starting_pot = td->nodes[0].pot;
        int4 *p_node;
        int2 *p_bets;
        int *p_kids, *p_abase, *p_iter;
        uint32_t *p_board, *p_hands, *p_rank;
        float *p_w;
        const size_t tab = 2 * (size_t)d.no16 * Mp;   // uint4 per board: both traversers
        int rc;
        if ((rc = alloc(&p_node, N, node.data())) || (rc = alloc(&p_bets, N, bets.data())) ||
            (rc = alloc(&p_kids, (size_t)total_kids, td->children)) || (rc = alloc(&p_abase, N, abase.data())) ||
            (rc = alloc(&p_board, N, board.data())) || (rc = alloc(&p_hands, hands.size(), hands.data())) ||
            (rc = alloc(&p_w, w.size(), w.data())) || (rc = alloc(&p_rank, sdset.boards.size() * 2 * Mp)) ||
            (rc = alloc(&d.R, (size_t)A * Mp)) || (rc = alloc(&d.SS, (size_t)A * Mp)) ||
            (rc = alloc(&d.reach, (size_t)N * Mp)) || (rc = alloc(&d.cfv, (size_t)N * Mp)) ||
            (rc = alloc(&d.root_sigma, (size_t)FS_MAX_ACTIONS * Mp)) || (rc = alloc(&p_iter, 1)))
            return rc;
        for (TermSet *ts : {&sdset, &foldset})
            if ((rc = alloc(&ts->d_boards, ts->boards.size(), ts->boards.data())) ||
                (rc = alloc(&ts->d_groups, ts->groups.size(), ts->groups.data())) ||
                (rc = alloc(&ts->d_tasks, ts->tasks.size(), ts->tasks.data())) ||
                (rc = alloc(&ts->d_tab, ts->boards.size() * tab)))
                return rc;
        d.node = p_node;
        d.bets = p_bets;
        d.children = p_kids;
        d.abase = p_abase;
        d.board = p_board;
        d.hands = p_hands;
        d.weights = p_w;
        d.iter = p_iter;
        // sweep records per level
        d_lvl.assign(L, nullptr);
        for (int l = 0; l < L; ++l) {
            if (lvl[l].empty()) continue;
            std::vector<int> rec(lvl[l].size() * FS_REC, 0);
            for (size_t i = 0; i < lvl[l].size(); ++i) {
                const int n = lvl[l][i];
                const FSNode &f = td->nodes[n];
                int *r = rec.data() + i * FS_REC;
                r[0] = n;
                r[1] = f.type | (f.player + 1) << 4 | f.num_children << 8;
                if (f.type == FS_NODE_DECISION) {
                    r[2] = abase[n];
                    for (int a = 0; a < f.num_children; ++a) r[3 + a] = td->children[f.first_child + a];
                } else {
                    r[2] = f.first_child;
                }
            }
            if ((rc = alloc(&d_lvl[l], rec.size(), rec.data()))) return rc;
        }
        if (const int nb = (int)sdset.boards.size()) {
            hipLaunchKernelGGL(k_fs_rank, dim3(grid_for((size_t)nb * 2 * Mp)), dim3(SS_BLOCK), 0, st, d,
                               sdset.d_boards, nb, p_rank);
            hipLaunchKernelGGL(k_fs_wmat, dim3(grid_for((size_t)nb * tab)), dim3(SS_BLOCK), 0, st, d, p_rank, nb,
                               sdset.d_tab);
        }
        if (const int nb = (int)foldset.boards.size())
            hipLaunchKernelGGL(k_fs_fmat, dim3(grid_for((size_t)nb * tab)), dim3(SS_BLOCK), 0, st, d,
                               foldset.d_boards, nb, foldset.d_tab);
        SS_HIP(hipGetLastError());
        SS_HIP(hipStreamSynchronize(st));
        return SS_OK;
    }

    static unsigned grid_for(size_t work) {
        return (unsigned)std::max<size_t>(1, std::min<size_t>(8192, (work + SS_BLOCK - 1) / SS_BLOCK));
    }
    // level sweeps: G wavefronts per node (four wavefronts per workgroup),
    // S segments of 64 hands in each lane. The reach sweep keeps a node's
    // segments in one lane (S = segments, G = 1); the value sweep, whose
    // traverser rows need three times the registers, gives each segment its
    // own wavefront (S = 1, G = segments) for the occupancy.
    static dim3 sweep_grid(size_t tasks) {
        return dim3((unsigned)std::max<size_t>(1, std::min<size_t>(16384, (tasks + 3) / 4)));
    }
    template <int S, int G> void down_s(size_t l, int opp) {
        hipLaunchKernelGGL((k_fs_down<S, G>), sweep_grid(lvl[l].size() * G), dim3(SS_BLOCK), 0, st, d, d_lvl[l],
                           (int)lvl[l].size(), opp, d.R, d.reach, d.root_sigma);
    }
    template <int S, int G> void up_s(size_t l, int trav) {
        hipLaunchKernelGGL((k_fs_up<S, G>), sweep_grid(lvl[l].size() * G), dim3(SS_BLOCK), 0, st, d, d_lvl[l],
                           (int)lvl[l].size(), trav, d.R, d.SS, d.cfv, d.root_sigma);
    }
    void launch_down(size_t l, int opp) {
        switch ((d.nh[opp] + 63) / 64) {
        case 1: down_s<1, 1>(l, opp); break;
        case 2: down_s<2, 1>(l, opp); break;
        case 3: down_s<3, 1>(l, opp); break;
        default: down_s<4, 1>(l, opp); break;
        }
    }
    void launch_up(size_t l, int trav) {
        switch ((d.nh[trav] + 63) / 64) {
        case 1: up_s<1, 1>(l, trav); break;
        case 2: up_s<1, 2>(l, trav); break;
        case 3: up_s<1, 3>(l, trav); break;
        default: up_s<1, 4>(l, trav); break;
        }
    }
    // terminal values: one workgroup per task, Mp lanes
    template <bool FOLD> void launch_term(const TermSet &ts, int trav) {
        if (!ts.tasks.empty())
            hipLaunchKernelGGL(k_fs_term<FOLD>, dim3((unsigned)ts.tasks.size()), dim3(d.Mp), 0, st, d, ts.d_tasks,
                               (int)ts.tasks.size(), ts.d_groups, ts.d_tab, d.reach, d.cfv, trav);
    }

    // one iteration: traverser 0's pass, then traverser 1's, then t + 1
    void launch_iteration() {
        for (int trav = 0; trav < 2; ++trav) {
            const int opp = 1 - trav;
            hipLaunchKernelGGL(k_fs_root, dim3(1), dim3(SS_BLOCK), 0, st, d, opp);
            for (size_t l = 0; l < lvl.size(); ++l)
                if (!lvl[l].empty()) launch_down(l, opp);
            launch_term<true>(foldset, trav);
            launch_term<false>(sdset, trav);
            for (size_t l = lvl.size(); l-- > 0;)
                if (!lvl[l].empty()) launch_up(l, trav);
        }
        hipLaunchKernelGGL(k_fs_next_iter, dim3(1), dim3(64), 0, st, d.iter);
    }

    int run(int iters) {
        const int one = 1;
        SS_HIP(hipMemcpyAsync(d.iter, &one, sizeof(int), hipMemcpyHostToDevice, st));
        SS_HIP(hipStreamSynchronize(st));
        if (iters <= 0) return SS_OK;
        const bool graphs = !getenv("SS_NO_GRAPH");
        if (graphs) {
            SS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            launch_iteration();
            hipError_t e = hipStreamEndCapture(st, &graph);
            if (e != hipSuccess) return set_error(SS_ERR_HIP, "flop solve: graph capture: %s", hipGetErrorString(e));
            SS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            for (int i = 0; i < iters; ++i) SS_HIP(hipGraphLaunch(exec, st));
        } else {
            for (int i = 0; i < iters; ++i) launch_iteration();
        }
        SS_HIP(hipGetLastError());
        SS_HIP(hipStreamSynchronize(st));
        return SS_OK;
    }
};

int fs_solve_impl(FSTreeData *td, int iters, FSOutput *out, bool all) {
    clear_error();
    if (!td || !out) return set_error(SS_ERR_ARG, "flop solve: null argument");
    if (iters < 0) return set_error(SS_ERR_ARG, "flop solve: %d iterations", iters);
    FsRun r;
    int rc = r.setup(td);
    if (rc) return rc;
    if ((rc = r.run(iters))) return rc;
    const int N = td->num_nodes, Mp = r.d.Mp, maxh = std::max(td->num_hands[0], td->num_hands[1]);
    const FSNode &root = td->nodes[0];
    memset(out, 0, sizeof(*out));
    if (root.type == FS_NODE_DECISION) {
        const int rp = root.player, nh = td->num_hands[rp], na = root.num_children;
        std::vector<float> sg((size_t)na * Mp);
        SS_HIP(hipMemcpy(sg.data(), r.d.root_sigma, sizeof(float) * sg.size(), hipMemcpyDeviceToHost));
        out->root_num_actions = na;
        out->root_player = rp;
        out->root_strategy = (float *)malloc(sizeof(float) * std::max(1, na * nh));
        for (int a = 0; a < na; ++a)
            for (int h = 0; h < nh; ++h) out->root_strategy[a * nh + h] = sg[(size_t)a * Mp + h];
    }
    if (!all) return SS_OK;
    // average strategies at every decision node (negative sums clamped, uniform
    // where nothing accumulated), the last pass's values, the turn roots
    // (flop_solve.cu:1503-1559)
    std::vector<float> ss((size_t)r.A * Mp), cfv((size_t)N * Mp);
    SS_HIP(hipMemcpy(ss.data(), r.d.SS, sizeof(float) * ss.size(), hipMemcpyDeviceToHost));
    SS_HIP(hipMemcpy(cfv.data(), r.d.cfv, sizeof(float) * cfv.size(), hipMemcpyDeviceToHost));
    out->max_hands = maxh;
    out->all_avg_strategies = (float *)calloc((size_t)N * FS_MAX_ACTIONS * maxh, sizeof(float));
    out->all_cfv = (float *)calloc((size_t)N * maxh, sizeof(float));
    if (!out->all_avg_strategies || !out->all_cfv) return set_error(SS_ERR_NOMEM, "flop solve: output allocation");
    const int nlast = td->num_hands[1];   // the last pass is traverser 1's: its hands only, as after a cleared array
    for (int n = 0; n < N; ++n) {
        for (int h = 0; h < nlast; ++h) out->all_cfv[(size_t)n * maxh + h] = cfv[(size_t)n * Mp + h];
        if (td->nodes[n].type != FS_NODE_DECISION) continue;
        const int na = td->nodes[n].num_children, nh = td->num_hands[td->nodes[n].player];
        float *o = out->all_avg_strategies + (size_t)n * FS_MAX_ACTIONS * maxh;
        for (int h = 0; h < nh; ++h) {
            float sum = 0.f;
            for (int a = 0; a < na; ++a) sum += std::max(0.f, ss[(size_t)(r.abase[n] + a) * Mp + h]);
            for (int a = 0; a < na; ++a) {
                const float v = ss[(size_t)(r.abase[n] + a) * Mp + h];
                o[(size_t)a * maxh + h] = sum > 0.f ? (v > 0.f ? v * (1.0f / sum) : 0.f) : 1.0f / na;
            }
        }
    }
    std::vector<int> roots, cards;
    for (int n = 0; n < N; ++n) {
        const FSNode &f = td->nodes[n];
        if (f.type != FS_NODE_CHANCE || f.num_board != 3) continue;
        for (int k = 0; k < f.num_children; ++k) {
            const int c = td->children[f.first_child + k];
            if (td->nodes[c].type != FS_NODE_DECISION) continue;
            roots.push_back(c);
            cards.push_back(td->nodes[c].board_cards[3]);
        }
    }
    out->turn_root_indices = copy_out(roots);
    out->turn_root_cards = copy_out(cards);
    out->num_turn_roots = (int)roots.size();
    return SS_OK;
}

}  // namespace
}  // namespace ss

// ---- C ABI -------------------------------------------------------------------
extern "C" int fs_build_tree(const int *flop_board, int starting_pot, int effective_stack, const float *bet_sizes,
                             int num_bet_sizes, FSTreeData *tree_out) {
    using namespace ss;
    clear_error();
    if (!flop_board || !tree_out || (num_bet_sizes > 0 && !bet_sizes) || num_bet_sizes < 0)
        return set_error(SS_ERR_ARG, "fs_build_tree: null or negative argument");
    memset(tree_out, 0, sizeof(*tree_out));
    FsBuilder b;
    b.nodes.reserve(8192);
    b.children.reserve(32768);
    const std::vector<float> sizes(bet_sizes, bet_sizes + num_bet_sizes), sub{0.75f};
    b.street(false, 0, starting_pot, effective_stack, 0, 0, 0, 0, sizes, flop_board, 3);
    b.deal(sub);   // turn
    b.deal(sub);   // river
    const int N = (int)b.nodes.size();
    std::vector<int> order{0}, depth(N, 0), dec, sd;
    int max_depth = 0;
    for (size_t q = 0; q < order.size(); ++q) {
        const FSNode &f = b.nodes[order[q]];
        for (int a = 0; a < f.num_children; ++a) {
            const int k = b.children[f.first_child + a];
            depth[k] = depth[order[q]] + 1;
            max_depth = std::max(max_depth, depth[k]);
            order.push_back(k);
        }
    }
    for (int n = 0; n < N; ++n) {
        if (b.nodes[n].type == FS_NODE_DECISION) dec.push_back(n);
        if (b.nodes[n].type == FS_NODE_SHOWDOWN) sd.push_back(n);
    }
    tree_out->nodes = copy_out(b.nodes);
    tree_out->children = copy_out(b.children);
    tree_out->num_nodes = N;
    tree_out->num_children_total = (int)b.children.size();
    tree_out->level_order = copy_out(order);
    tree_out->node_depth = copy_out(depth);
    tree_out->max_depth = max_depth;
    tree_out->decision_node_indices = copy_out(dec);
    tree_out->num_decision_nodes = (int)dec.size();
    tree_out->showdown_node_indices = copy_out(sd);
    tree_out->num_showdown_nodes = (int)sd.size();
    return 0;
}

extern "C" int fs_solve_gpu(FSTreeData *tree_data, int max_iterations, FSOutput *output) {
    return ss::fs_solve_impl(tree_data, max_iterations, output, false);
}

extern "C" int fs_solve_gpu_extract_all(FSTreeData *tree_data, int max_iterations, FSOutput *output) {
    return ss::fs_solve_impl(tree_data, max_iterations, output, true);
}

extern "C" void fs_free_tree(FSTreeData *td) {
    if (!td) return;
    free(td->nodes);
    free(td->children);
    free(td->level_order);
    free(td->node_depth);
    free(td->decision_node_indices);
    free(td->showdown_node_indices);
    memset(td, 0, sizeof(*td));
}

extern "C" void fs_free_output(FSOutput *out) {
    if (!out) return;
    free(out->root_strategy);
    free(out->root_ev);
    free(out->all_avg_strategies);
    free(out->all_cfv);
    free(out->turn_root_indices);
    free(out->turn_root_cards);
    memset(out, 0, sizeof(*out));
}

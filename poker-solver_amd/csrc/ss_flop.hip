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
//     node's board (:731-779; the traverser's own board conflicts are not
//     checked);
//   - showdown: +-pot / 2 per opponent hand ranked below / above, 0 for ties,
//     blocked hands (rank 0) and a blocked traverser (:781-833);
//   - each traverser pass re-matches every decision node, then propagates the
//     opponent's reach down the levels and the values up them, updating the
//     traverser's regrets R <- (R + dv) t / (t + 1) and strategy sums
//     SS += t sigma (:835-891, the iteration loop :1150-1228).
//
// The GPU form: every array is [row][Mp] (Mp = the larger hand count rounded
// to 64, one wavefront of hands). One traverser pass is a fixed sequence of
// launches over per-level task lists (decision and chance nodes), the fold
// list and the showdown list; the whole iteration is captured once in a
// hipGraph and replayed, with the iteration counter on the device. Showdown
// ranks are computed once per showdown node at setup. Terminal workgroups
// stage the opponent's reach, cards and ranks in LDS.
#include "ss_internal.h"
#include "ss_eval.h"
#include "../../include/flop_solve.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
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
    const int4 *node;          // [N] x = type | player << 4 (player + 1), y = nchild, z = first_child, w = pot
    const int2 *bets;          // [N]
    const int *children;
    const int *abase;          // [N] first action row of a decision node, -1 otherwise
    const uint32_t *board;     // [N] 5 cards, 6 bits each, and the count << 30
    const uint32_t *hands;     // [2][Mp] c0 | c1 << 8 (0xFFFF past the hand count)
    const float *weights;      // [2][Mp]
    const int *sd_row;         // [N] showdown rank row, -1 otherwise
    uint32_t *rank;            // [num showdowns][2][Mp], written once at setup
    float *R, *sigma, *SS;     // [A][Mp]
    float *reach, *cfv;        // [N][Mp]
    int *iter;                 // the iteration being run
};

__device__ __forceinline__ int fs_type(int4 n) { return n.x & 15; }
__device__ __forceinline__ int fs_player(int4 n) { return (n.x >> 4) - 1; }
__device__ __forceinline__ bool fs_on_board(uint32_t bw, int c) {
    const int nb = (int)(bw >> 30);
    for (int b = 0; b < nb; ++b)
        if ((int)((bw >> (6 * b)) & 63) == c) return true;
    return false;
}

// ranks of both players' hands at every showdown node (0: blocked by the board)
__global__ void __launch_bounds__(SS_BLOCK) k_fs_rank(FsDev d, const int *sd_nodes, int nsd) {
    __shared__ uint8_t tab[8192];
    fill_straight_table(tab, threadIdx.x, blockDim.x);
    __syncthreads();
    const size_t total = (size_t)nsd * 2 * d.Mp;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int h = (int)(e % d.Mp), p = (int)((e / d.Mp) % 2), s = (int)(e / d.Mp / 2);
        uint32_t r = 0;
        if (h < d.nh[p]) {
            const uint32_t bw = d.board[sd_nodes[s]], hc = d.hands[p * d.Mp + h];
            const int c0 = hc & 255, c1 = (hc >> 8) & 255;
            if (!fs_on_board(bw, c0) && !fs_on_board(bw, c1)) {
                int c[7];
                for (int b = 0; b < 5; ++b) c[b] = (int)((bw >> (6 * b)) & 63);
                c[5] = c0;
                c[6] = c1;
                r = eval7_lut(c, tab);
            }
        }
        d.rank[e] = r;
    }
}

// regret matching at every decision node, for its actor's hands
__global__ void __launch_bounds__(SS_BLOCK) k_fs_match(FsDev d, const int *dec, int ndec) {
    const size_t total = (size_t)ndec * d.Mp;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int h = (int)(e % d.Mp), n = dec[e / d.Mp];
        const int4 nd = d.node[n];
        if (h >= d.nh[fs_player(nd)]) continue;
        const int na = nd.y;
        const float *R = d.R + (size_t)d.abase[n] * d.Mp + h;
        float *sg = d.sigma + (size_t)d.abase[n] * d.Mp + h;
        float sum = 0.f;
        for (int a = 0; a < na; ++a) sum += fmaxf(R[(size_t)a * d.Mp], 0.f);
        if (sum > 0.f) {
            const float inv = 1.0f / sum;
            for (int a = 0; a < na; ++a) sg[(size_t)a * d.Mp] = fmaxf(R[(size_t)a * d.Mp], 0.f) * inv;
        } else {
            const float u = 1.0f / na;
            for (int a = 0; a < na; ++a) sg[(size_t)a * d.Mp] = u;
        }
    }
}

// the opponent's reach at the root: the raw input weights
__global__ void __launch_bounds__(SS_BLOCK) k_fs_root(FsDev d, int opp) {
    for (int h = threadIdx.x; h < d.nh[opp]; h += blockDim.x) d.reach[h] = d.weights[opp * d.Mp + h];
}

// the opponent's reach down one level (decision and chance nodes)
__global__ void __launch_bounds__(SS_BLOCK) k_fs_down(FsDev d, const int *lvl, int nlvl, int opp) {
    const size_t total = (size_t)nlvl * d.Mp;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int h = (int)(e % d.Mp), n = lvl[e / d.Mp];
        if (h >= d.nh[opp]) continue;
        const int4 nd = d.node[n];
        const float pr = d.reach[(size_t)n * d.Mp + h];
        const int *kid = d.children + nd.z;
        if (fs_type(nd) == FS_NODE_DECISION && fs_player(nd) == opp) {
            const float *sg = d.sigma + (size_t)d.abase[n] * d.Mp + h;
            for (int a = 0; a < nd.y; ++a) d.reach[(size_t)kid[a] * d.Mp + h] = pr * sg[(size_t)a * d.Mp];
        } else {
            for (int a = 0; a < nd.y; ++a) d.reach[(size_t)kid[a] * d.Mp + h] = pr;
        }
    }
}

// fold and showdown values for the traverser's hands; one workgroup per node,
// the opponent's reach, cards and ranks staged in LDS
__global__ void __launch_bounds__(SS_BLOCK) k_fs_terminal(FsDev d, const int *term, int nterm, int trav) {
    __shared__ float s_reach[256];
    __shared__ uint32_t s_cards[256], s_rank[256];
    const int opp = 1 - trav, n_opp = d.nh[opp], n_trav = d.nh[trav];
    for (int j = blockIdx.x; j < nterm; j += gridDim.x) {
        const int n = term[j];
        const int4 nd = d.node[n];
        const bool fold = fs_type(nd) == FS_NODE_FOLD;
        const int sr = d.sd_row[n];
        const uint32_t bw = d.board[n];
        __syncthreads();   // the previous node's staging is read
        for (int o = threadIdx.x; o < n_opp; o += blockDim.x) {
            const uint32_t oc = d.hands[opp * d.Mp + o];
            s_reach[o] = d.reach[(size_t)n * d.Mp + o];
            s_cards[o] = oc;
            // fold: an opponent hand on the node's board is left out (rank 0 marks it)
            s_rank[o] = fold ? ((fs_on_board(bw, oc & 255) || fs_on_board(bw, (oc >> 8) & 255)) ? 0u : 1u)
                             : d.rank[((size_t)sr * 2 + opp) * d.Mp + o];
        }
        __syncthreads();
        for (int h = threadIdx.x; h < n_trav; h += blockDim.x) {
            const uint32_t hc = d.hands[trav * d.Mp + h];
            const uint32_t a0 = hc & 255, a1 = (hc >> 8) & 255;
            float val = 0.f;
            if (fold) {
                const int winner = fs_player(nd), loser = 1 - winner;
                const float half_start = d.starting_pot * 0.5f;
                const float payoff = trav == winner ? half_start + (float)(loser == 0 ? d.bets[n].x : d.bets[n].y)
                                                    : -(half_start + (float)(trav == 0 ? d.bets[n].x : d.bets[n].y));
                float opp_sum = 0.f;
                for (int o = 0; o < n_opp; ++o) {
                    const uint32_t oc = s_cards[o], b0 = oc & 255, b1 = (oc >> 8) & 255;
                    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1 || !s_rank[o]) continue;
                    opp_sum += s_reach[o];
                }
                val = opp_sum * payoff;
            } else {
                const uint32_t hs = d.rank[((size_t)sr * 2 + trav) * d.Mp + h];
                if (hs) {
                    const float half_pot = nd.w * 0.5f;
                    for (int o = 0; o < n_opp; ++o) {
                        const uint32_t oc = s_cards[o], b0 = oc & 255, b1 = (oc >> 8) & 255;
                        if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) continue;
                        const uint32_t os = s_rank[o];
                        if (!os) continue;
                        const float w = s_reach[o];
                        if (hs > os) val += w * half_pot;
                        else if (hs < os) val -= w * half_pot;
                    }
                }
            }
            d.cfv[(size_t)n * d.Mp + h] = val;
        }
    }
}

// values up one level; the traverser's regrets and strategy sums at its nodes
__global__ void __launch_bounds__(SS_BLOCK) k_fs_up(FsDev d, const int *lvl, int nlvl, int trav) {
    const int t = *d.iter;
    const float disc = (float)t / ((float)t + 1.0f), tw = (float)t;
    const size_t total = (size_t)nlvl * d.Mp;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int h = (int)(e % d.Mp), n = lvl[e / d.Mp];
        if (h >= d.nh[trav]) continue;
        const int4 nd = d.node[n];
        const int *kid = d.children + nd.z;
        const int na = nd.y;
        float val = 0.f;
        if (fs_type(nd) == FS_NODE_DECISION && fs_player(nd) == trav) {
            const size_t row = (size_t)d.abase[n] * d.Mp + h;
            for (int a = 0; a < na; ++a) val += d.sigma[row + (size_t)a * d.Mp] * d.cfv[(size_t)kid[a] * d.Mp + h];
            d.cfv[(size_t)n * d.Mp + h] = val;
            for (int a = 0; a < na; ++a) {
                float *r = d.R + row + (size_t)a * d.Mp;
                *r = (*r + (d.cfv[(size_t)kid[a] * d.Mp + h] - val)) * disc;
                d.SS[row + (size_t)a * d.Mp] += tw * d.sigma[row + (size_t)a * d.Mp];
            }
        } else {
            for (int a = 0; a < na; ++a) val += d.cfv[(size_t)kid[a] * d.Mp + h];
            if (fs_type(nd) == FS_NODE_CHANCE) val = val / (float)na;
            d.cfv[(size_t)n * d.Mp + h] = val;
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
    std::vector<int> dec, fold_sd, sd;
    std::vector<std::vector<int>> lvl;   // decision and chance nodes per depth
    std::vector<int *> d_lvl;
    int *d_dec = nullptr, *d_term = nullptr, *d_sd = nullptr;
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
        std::vector<int> sd_row(N, -1), depth(N, -1);
        abase.assign(N, -1);
        int total_kids = td->num_children_total;
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
            if (f.type == FS_NODE_FOLD && (f.player < 0 || f.player > 1))
                return set_error(SS_ERR_ARG, "flop solve: fold node %d has winner %d", n, f.player);
            if (f.type == FS_NODE_SHOWDOWN && f.num_board != 5)
                return set_error(SS_ERR_ARG, "flop solve: showdown node %d has %d board cards", n, f.num_board);
            node[n] = make_int4(f.type | (f.player + 1) << 4, f.num_children, f.first_child, f.pot);
            bets[n] = make_int2(f.bets[0], f.bets[1]);
            uint32_t bw = (uint32_t)f.num_board << 30;
            for (int b = 0; b < f.num_board; ++b) {
                if (f.board_cards[b] < 0 || f.board_cards[b] > 51)
                    return set_error(SS_ERR_ARG, "flop solve: node %d board card %d", n, f.board_cards[b]);
                bw |= (uint32_t)f.board_cards[b] << (6 * b);
            }
            board[n] = bw;
            if (f.type == FS_NODE_DECISION) {
                abase[n] = A;
                A += f.num_children;
                dec.push_back(n);
            }
            if (f.type == FS_NODE_SHOWDOWN) {
                sd_row[n] = (int)sd.size();
                sd.push_back(n);
            }
            if (f.type == FS_NODE_FOLD || f.type == FS_NODE_SHOWDOWN) fold_sd.push_back(n);
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
        if ((int)queue.size() != N) return set_error(SS_ERR_ARG, "flop solve: %d nodes off the root's tree", N - (int)queue.size());
        int L = 0;
        for (int n = 0; n < N; ++n) L = std::max(L, depth[n] + 1);
        lvl.assign(L, {});
        for (int n : queue)
            if (td->nodes[n].type == FS_NODE_DECISION || td->nodes[n].type == FS_NODE_CHANCE)
                lvl[depth[n]].push_back(n);
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
        d.starting_pot = td->nodes[0].pot;
        int4 *p_node;
        int2 *p_bets;
        int *p_kids, *p_abase, *p_sdrow, *p_iter;
        uint32_t *p_board, *p_hands, *p_rank;
        float *p_w;
        int rc;
        if ((rc = alloc(&p_node, N, node.data())) || (rc = alloc(&p_bets, N, bets.data())) ||
            (rc = alloc(&p_kids, (size_t)total_kids, td->children)) || (rc = alloc(&p_abase, N, abase.data())) ||
            (rc = alloc(&p_board, N, board.data())) || (rc = alloc(&p_hands, hands.size(), hands.data())) ||
            (rc = alloc(&p_w, w.size(), w.data())) || (rc = alloc(&p_sdrow, N, sd_row.data())) ||
            (rc = alloc(&p_rank, sd.size() * 2 * Mp)) || (rc = alloc(&d.R, (size_t)A * Mp)) ||
            (rc = alloc(&d.sigma, (size_t)A * Mp)) || (rc = alloc(&d.SS, (size_t)A * Mp)) ||
            (rc = alloc(&d.reach, (size_t)N * Mp)) || (rc = alloc(&d.cfv, (size_t)N * Mp)) ||
            (rc = alloc(&p_iter, 1)))
            return rc;
        d.node = p_node;
        d.bets = p_bets;
        d.children = p_kids;
        d.abase = p_abase;
        d.board = p_board;
        d.hands = p_hands;
        d.weights = p_w;
        d.sd_row = p_sdrow;
        d.rank = p_rank;
        d.iter = p_iter;
        if ((rc = alloc(&d_dec, dec.size(), dec.data())) || (rc = alloc(&d_term, fold_sd.size(), fold_sd.data())) ||
            (rc = alloc(&d_sd, sd.size(), sd.data())))
            return rc;
        d_lvl.assign(L, nullptr);
        for (int l = 0; l < L; ++l)
            if (!lvl[l].empty() && (rc = alloc(&d_lvl[l], lvl[l].size(), lvl[l].data()))) return rc;
        if (!sd.empty())
            hipLaunchKernelGGL(k_fs_rank, dim3(grid_for((size_t)sd.size() * 2 * Mp)), dim3(SS_BLOCK), 0, st, d, d_sd,
                               (int)sd.size());
        SS_HIP(hipGetLastError());
        SS_HIP(hipStreamSynchronize(st));
        return SS_OK;
    }

    static unsigned grid_for(size_t work) {
        return (unsigned)std::max<size_t>(1, std::min<size_t>(8192, (work + SS_BLOCK - 1) / SS_BLOCK));
    }

    // one iteration: traverser 0's pass, then traverser 1's, then t + 1
    void launch_iteration() {
        for (int trav = 0; trav < 2; ++trav) {
            const int opp = 1 - trav;
            if (!dec.empty())
                hipLaunchKernelGGL(k_fs_match, dim3(grid_for(dec.size() * (size_t)d.Mp)), dim3(SS_BLOCK), 0, st, d,
                                   d_dec, (int)dec.size());
            hipLaunchKernelGGL(k_fs_root, dim3(1), dim3(SS_BLOCK), 0, st, d, opp);
            for (size_t l = 0; l < lvl.size(); ++l)
                if (!lvl[l].empty())
                    hipLaunchKernelGGL(k_fs_down, dim3(grid_for(lvl[l].size() * (size_t)d.Mp)), dim3(SS_BLOCK), 0, st,
                                       d, d_lvl[l], (int)lvl[l].size(), opp);
            if (!fold_sd.empty())
                hipLaunchKernelGGL(k_fs_terminal, dim3((unsigned)std::min<size_t>(fold_sd.size(), 65535)),
                                   dim3(SS_BLOCK), 0, st, d, d_term, (int)fold_sd.size(), trav);
            for (size_t l = lvl.size(); l-- > 0;)
                if (!lvl[l].empty())
                    hipLaunchKernelGGL(k_fs_up, dim3(grid_for(lvl[l].size() * (size_t)d.Mp)), dim3(SS_BLOCK), 0, st,
                                       d, d_lvl[l], (int)lvl[l].size(), trav);
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
        SS_HIP(hipMemcpy(sg.data(), r.d.sigma + (size_t)r.abase[0] * Mp, sizeof(float) * sg.size(),
                         hipMemcpyDeviceToHost));
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

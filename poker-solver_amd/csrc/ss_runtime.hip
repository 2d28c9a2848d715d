// ss_runtime.hip -- host runtime: problem intake, device plan, slot arenas,
// the plain driver, exploitability and output extraction.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <new>
#include <numeric>
#include <thread>
#include <utility>

#include "ss_internal.h"

namespace ss {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    if (log_level() > 0) fprintf(stderr, "[street_solve] error %d: %s\n", code, buf);
    return code;
}

void clear_error() { g_err.clear(); }
const char *last_error_str() { return g_err.c_str(); }

int log_level() {
    static int lvl = [] {
        const char *e = getenv("SS_LOG");
        return e ? atoi(e) : 0;
    }();
    return lvl;
}

// ---------------------------------------------------------------------------
// problem intake
// ---------------------------------------------------------------------------
HostProblem from_legacy(const SSTreeData *td) {
    HostProblem hp;
    hp.tree = td;
    hp.np = td->num_players;
    for (int p = 0; p < hp.np && p < SS_MAX_PLAYERS; ++p) {
        hp.nh[p] = td->num_hands[p];
        hp.hands[p] = &td->hands[p][0][0];
        hp.weights[p] = &td->weights[p][0];
    }
    hp.leaf_values = td->leaf_values;
    hp.frozen = td->frozen_action;
    return hp;
}

int from_ex(const SSProblem *p, HostProblem *out) {
    if (!p || !p->tree) return set_error(SS_ERR_ARG, "problem or problem->tree is NULL");
    HostProblem hp;
    hp.tree = p->tree;
    hp.np = p->tree->num_players;
    if (hp.np < 2 || hp.np > SS_MAX_PLAYERS) return set_error(SS_ERR_ARG, "tree has %d players", hp.np);
    for (int q = 0; q < hp.np; ++q) {
        hp.nh[q] = p->num_hands[q];
        hp.hands[q] = p->hands[q];
        hp.weights[q] = p->weights[q];
        if (hp.nh[q] < 0 || hp.nh[q] > SS_EX_MAX_HANDS)
            return set_error(SS_ERR_ARG, "player %d has %d hands (max %d)", q, hp.nh[q], SS_EX_MAX_HANDS);
        if (hp.nh[q] > 0 && (!hp.hands[q] || !hp.weights[q]))
            return set_error(SS_ERR_ARG, "player %d hands/weights NULL", q);
    }
    hp.leaf_values = p->leaf_values;
    hp.frozen = p->frozen_action;
    *out = hp;
    return SS_OK;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
static inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

int make_plan(const HostProblem &hp, Plan &pl) {
    const SSTreeData *td = hp.tree;
    if (!td || !td->nodes || !td->children || td->num_nodes < 1) return set_error(SS_ERR_ARG, "empty tree");
    const int N = td->num_nodes, NP = hp.np;
    if (NP < 2 || NP > SS_MAX_PLAYERS) return set_error(SS_ERR_ARG, "num_players %d", NP);
    int maxh = 0;
    for (int p = 0; p < NP; ++p) maxh = std::max(maxh, hp.nh[p]);
    if (maxh == 0) return set_error(SS_ERR_GENERIC, "no hands for any player");   // street_solve.cu:1306
    pl = Plan();
    pl.N = N;
    pl.NP = NP;
    pl.maxh = maxh;
    pl.Mp = round_up(maxh, 64);
    pl.is_river = td->is_river ? 1 : 0;
    pl.num_leaves = td->num_leaves;
    pl.num_board = td->num_board;
    for (int i = 0; i < 5; ++i) pl.board[i] = td->board[i];
    pl.starting_pot = td->starting_pot;
    for (int p = 0; p < NP; ++p) pl.nh[p] = hp.nh[p];

    // nodes, depths
    std::vector<int> depth(N, 0), bfs;
    bfs.reserve(N);
    bfs.push_back(0);
    for (size_t i = 0; i < bfs.size(); ++i) {
        const SSNode &nd = td->nodes[bfs[i]];
        if (nd.num_children < 0 || nd.first_child < 0 ||
            nd.first_child + nd.num_children > td->num_children_total)
            return set_error(SS_ERR_ARG, "node %d has a bad child range", bfs[i]);
        for (int a = 0; a < nd.num_children; ++a) {
            int c = td->children[nd.first_child + a];
            if (c <= 0 || c >= N) return set_error(SS_ERR_ARG, "child id %d out of range", c);
            depth[c] = depth[bfs[i]] + 1;
            bfs.push_back(c);
            if ((int)bfs.size() > N) return set_error(SS_ERR_ARG, "tree is not a tree (cycle or shared child)");
        }
    }
    if ((int)bfs.size() != N) return set_error(SS_ERR_ARG, "unreachable nodes in tree");
    // Reach rows: player p's reach changes only below p's own actions, so the
    // row holding p's reach at node c is c where p acted at c's parent, else
    // the parent's row (the root's row holds the range weights). The forward
    // sweep writes only the acting player's child rows; every reader goes
    // through this table, so no reach is copied down the tree.
    pl.rrow.assign((size_t)NP * N, 0);
    for (int i : bfs) {
        const SSNode &nd = td->nodes[i];
        for (int a = 0; a < nd.num_children; ++a) {
            const int c = td->children[nd.first_child + a];
            for (int p = 0; p < NP; ++p)
                pl.rrow[(size_t)p * N + c] =
                    nd.type == SS_NODE_DECISION && nd.player == p ? c : pl.rrow[(size_t)p * N + i];
        }
    }
    int maxd = 0;
    for (int i = 0; i < N; ++i) maxd = std::max(maxd, depth[i]);

    pl.nodes.resize(N);
    int A = 0, D = 0;
    pl.dec_index.assign(N, -1);
    for (int i = 0; i < N; ++i) {
        const SSNode &s = td->nodes[i];
        DevNode &d = pl.nodes[i];
        std::memset(&d, 0, sizeof(d));
        d.type = s.type;
        d.player = s.player;
        d.nchild = s.num_children;
        d.first_child = s.first_child;
        d.pot = s.pot;
        d.leaf_idx = s.leaf_idx;
        d.act_base = -1;
        d.depth = depth[i];
        for (int q = 0; q < SS_MAX_PLAYERS; ++q) {
            d.bets[q] = s.bets[q];
            if (q < NP && s.active_players[q]) d.active_mask |= 1 << q;
        }
        if (s.type == SS_NODE_DECISION) {
            if (s.num_children < 1 || s.num_children > SS_MAX_ACTIONS)
                return set_error(SS_ERR_UNSUPPORTED, "decision node %d has %d children (max %d)", i, s.num_children,
                                 SS_MAX_ACTIONS);
            d.act_base = A;
            A += s.num_children;
            pl.dec_index[i] = D++;
        } else if (s.num_children != 0) {
            return set_error(SS_ERR_ARG, "terminal node %d has children", i);
        }
        if (s.type == SS_NODE_LEAF && s.leaf_idx >= td->num_leaves)
            return set_error(SS_ERR_ARG, "leaf_idx %d >= num_leaves %d", s.leaf_idx, td->num_leaves);
    }
    pl.A = A;
    pl.D = D;
    pl.root_player = td->nodes[0].player;
    pl.children.assign(td->children, td->children + td->num_children_total);

    // decision nodes per level, terminals
    pl.L = maxd + 1;
    pl.dec_off.assign(pl.L + 1, 0);
    for (int i = 0; i < N; ++i)
        if (pl.nodes[i].type == SS_NODE_DECISION) pl.dec_off[depth[i] + 1]++;
    for (int l = 0; l < pl.L; ++l) pl.dec_off[l + 1] += pl.dec_off[l];
    pl.dec_lvl.assign(std::max(D, 1), 0);
    {
        std::vector<int> fill(pl.dec_off.begin(), pl.dec_off.end() - 1);
        for (int i = 0; i < N; ++i)
            if (pl.nodes[i].type == SS_NODE_DECISION) pl.dec_lvl[fill[depth[i]]++] = i;
    }
    for (int i = 0; i < N; ++i)
        if (pl.nodes[i].type != SS_NODE_DECISION) pl.term.push_back(i);
    pl.T = (int)pl.term.size();
    if (pl.term.empty()) pl.term.push_back(0);

    // hands
    const int Mp = pl.Mp;
    pl.hinfo.assign((size_t)NP * Mp, 0u);
    pl.weights.assign((size_t)NP * Mp, 0.f);
    bool board_mask[52] = {false};
    for (int b = 0; b < td->num_board; ++b) {
        if (td->board[b] < 0 || td->board[b] > 51) return set_error(SS_ERR_ARG, "board card out of range");
        board_mask[td->board[b]] = true;
    }
    {
        int nl = 0;
        for (int c = 0; c < 52; ++c)
            if (!board_mask[c]) pl.lab[c] = nl++;
        for (int c = 0; c < 52; ++c)
            if (board_mask[c]) pl.lab[c] = nl++;
    }
    for (int p = 0; p < NP; ++p)
        for (int h = 0; h < hp.nh[p]; ++h) {
            int a = hp.hands[p][2 * h], b = hp.hands[p][2 * h + 1];
            if (a < 0 || a > 51 || b < 0 || b > 51 || a == b)
                return set_error(SS_ERR_ARG, "player %d hand %d has bad cards (%d, %d)", p, h, a, b);
            uint32_t live = (!board_mask[a] && !board_mask[b]) ? SS_LIVE_BIT : 0u;
            pl.hinfo[(size_t)p * Mp + h] =
                (uint32_t)pl.lab[a] | ((uint32_t)pl.lab[b] << 6) | ((uint32_t)combo_id(a, b) << 12) | live;
            pl.weights[(size_t)p * Mp + h] = hp.weights[p][h];
        }
    // live combos unique within every player's list: card-pair scatters are
    // then single writes per cell (duplicates fall back to atomic adds)
    pl.uniq = 1;
    for (int p = 0; p < NP && pl.uniq; ++p) {
        std::vector<char> seen(SS_NCOMBO, 0);
        for (int h = 0; h < hp.nh[p]; ++h) {
            const uint32_t hi = pl.hinfo[(size_t)p * Mp + h];
            if (!(hi & SS_LIVE_BIT)) continue;
            char &f = seen[(hi >> 12) & 2047];
            if (f) { pl.uniq = 0; break; }
            f = 1;
        }
    }

    // terminal opponent counts: 3+ opponents take the odometer path
    int64_t odo_work = 0;
    for (int j = 0; j < pl.T; ++j) {
        const DevNode &nd = pl.nodes[pl.term[j]];
        for (int t = 0; t < NP; ++t) {
            int k = 0;
            int64_t w = 1;
            for (int q = 0; q < NP; ++q) {
                if (q == t) continue;
                if (nd.type != SS_NODE_FOLD && !((nd.active_mask >> q) & 1)) continue;
                ++k;
                w *= std::max(1, hp.nh[q]);
            }
            if (k >= 3) {
                pl.odometer = 1;
                odo_work = std::max(odo_work, w);
            }
        }
    }
    if (pl.odometer && odo_work > (int64_t)1 << 24)
        return set_error(SS_ERR_UNSUPPORTED,
                         "terminals with 3+ opponents need the exact odometer (%lld hand tuples per hand); "
                         "ranges this large are supported for up to 2 opponents per terminal",
                         (long long)odo_work);

    // leaves
    pl.has_leaf = (td->num_leaf_nodes > 0 || pl.num_leaves > 0) && hp.leaf_values ? 1 : 0;
    if (pl.has_leaf) {
        pl.leaf_src = hp.leaf_values;
        pl.leaf_count = (size_t)pl.num_leaves * NP * maxh;
    }
    pl.frozen_src = hp.frozen;
    if (pl.frozen_src) pl.frozen.assign(pl.frozen_src, pl.frozen_src + N);
    if (pl.frozen_src)
        for (int i = 0; i < N; ++i) {
            int fa = pl.frozen_src[i];
            if (fa >= 0 && (pl.nodes[i].type != SS_NODE_DECISION || fa >= pl.nodes[i].nchild))
                return set_error(SS_ERR_ARG, "frozen_action[%d] = %d is not an action of that node", i, fa);
        }

    // algorithmic bytes of one iteration (4-byte state, M = max hands)
    int64_t Ap[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < N; ++i)
        if (pl.nodes[i].type == SS_NODE_DECISION && pl.nodes[i].player >= 0 && pl.nodes[i].player < NP)
            Ap[pl.nodes[i].player] += pl.nodes[i].nchild;
    int64_t nleaf = 0;
    for (int i = 0; i < N; ++i) nleaf += pl.nodes[i].type == SS_NODE_LEAF;
    int64_t rows = 0;
    for (int t = 0; t < NP; ++t) {
        rows += 2 * A;                                    // regret match: read R, write sigma
        for (int rp = 0; rp < NP; ++rp)
            if (rp != t) rows += D + A + Ap[rp];          // reach: parent, children, sigma of rp
        rows += (int64_t)(NP - 1) * pl.T + pl.T + (pl.has_leaf ? nleaf : 0);   // terminals
        rows += A + D + 5 * Ap[t];                        // backup
    }
    pl.algo_bytes_iter = rows * 4 * (int64_t)maxh;
    return SS_OK;
}

// ---------------------------------------------------------------------------
// river tables from combo ranks
// ---------------------------------------------------------------------------
void build_river_tables(const uint32_t *cstr, const int *lab, RiverTables &rt) {
    rt.cstr.assign(cstr, cstr + SS_NCOMBO);
    rt.glt.assign(SS_NCOMBO, 0);
    rt.gle.assign(SS_NCOMBO, 0);
    rt.clist.assign(52 * 52, 0);
    rt.cpos.assign(52 * 52, 255);
    rt.clen.assign(52, 0);
    std::vector<int> live;
    for (int k = 0; k < SS_NCOMBO; ++k)
        if (cstr[k]) live.push_back(k);
    std::stable_sort(live.begin(), live.end(), [&](int a, int b) { return cstr[a] < cstr[b]; });
    rt.ng = (int)live.size();
    rt.gsort.assign(std::max(rt.ng, 1), 0);
    std::vector<uint32_t> sorted(rt.ng);
    for (int i = 0; i < rt.ng; ++i) {
        int a = 0, rem = live[i];
        while (rem >= 51 - a) { rem -= 51 - a; ++a; }
        rt.gsort[i] = (uint16_t)(lab[a] | (lab[a + 1 + rem] << 8));
        sorted[i] = cstr[live[i]];
    }
    // strength classes: distinct live ranks, ascending
    std::vector<uint32_t> cls(sorted);
    cls.erase(std::unique(cls.begin(), cls.end()), cls.end());
    rt.ncls = (int)cls.size();
    rt.kcls.assign(SS_NCOMBO, 0);
    rt.bnd.assign(SS_BND_BYTES, 0);
    for (int k = 0; k < SS_NCOMBO; ++k)
        if (cstr[k]) rt.kcls[k] = (uint16_t)(std::lower_bound(cls.begin(), cls.end(), cstr[k]) - cls.begin());
    for (int k = 0; k < SS_NCOMBO; ++k) {
        uint32_t s = cstr[k];
        rt.glt[k] = (int16_t)(std::lower_bound(sorted.begin(), sorted.end(), s) - sorted.begin());
        rt.gle[k] = (int16_t)(std::upper_bound(sorted.begin(), sorted.end(), s) - sorted.begin());
    }
    for (int c = 0; c < 52; ++c) {
        std::vector<int> part;
        for (int d = 0; d < 52; ++d)
            if (d != c && cstr[combo_id(c, d)]) part.push_back(d);
        std::stable_sort(part.begin(), part.end(),
                         [&](int a, int b) { return cstr[combo_id(c, a)] < cstr[combo_id(c, b)]; });
        const int lc = lab[c];
        rt.clen[lc] = (uint8_t)part.size();
        std::vector<uint32_t> ps(part.size());
        for (size_t j = 0; j < part.size(); ++j) {
            rt.clist[lc * 52 + j] = (uint8_t)lab[part[j]];
            rt.cpos[lc * 52 + lab[part[j]]] = (uint8_t)(4 * j);
            ps[j] = cstr[combo_id(c, part[j])];
        }
        for (size_t t = 0; t < cls.size(); ++t)
            rt.bnd[t * 52 + lc] = (uint8_t)(4 * (std::lower_bound(ps.begin(), ps.end(), cls[t]) - ps.begin()));
        rt.bnd[cls.size() * 52 + lc] = (uint8_t)(4 * ps.size());
    }
}

// ---------------------------------------------------------------------------
// slot arena
// ---------------------------------------------------------------------------
struct ArenaLayout {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    }
};

int launch_combo_ranks(const int *d_board, uint32_t *d_out, hipStream_t st);

Slot::~Slot() {
    if (pool) return;   // the pool's memory
    if (arena) (void)hipFree(arena);
    if (xarena) (void)hipFree(xarena);
    if (d_desc) (void)hipFree(d_desc);
}

int DevicePool::alloc(size_t n, void **out) {
    n = (n + 255) & ~(size_t)255;
    for (auto &c : chunks)
        if (c.cap - c.used >= n) {
            *out = c.p + c.used;
            c.used += n;
            return SS_OK;
        }
    Chunk c;
    c.cap = std::max(n, chunk_bytes);
    SS_HIP(hipMalloc((void **)&c.p, c.cap));
    c.used = n;
    *out = c.p;
    chunks.push_back(c);
    return SS_OK;
}

DevicePool::~DevicePool() {
    for (auto &c : chunks) (void)hipFree(c.p);
}

// device memory of a slot: from its pool when it has one
static int slot_alloc(DevicePool *pool, size_t n, void **out) {
    if (pool) return pool->alloc(n, out);
    SS_HIP(hipMalloc(out, n));
    return SS_OK;
}

int64_t estimate_footprint(const HostProblem &hp) {
    Plan pl;
    int rc = make_plan(hp, pl);
    if (rc != SS_OK) return rc;   // the planner's SS_ERR_* code (message in ss_last_error)
    size_t Mp = pl.Mp;
    int64_t b = 0;
    b += (int64_t)pl.N * sizeof(DevNode) + pl.children.size() * 4 + (pl.D + pl.L + 1 + pl.T) * 4;
    b += (int64_t)pl.NP * Mp * 8 + pl.leaf_count * 4 + pl.N * 4 + (int64_t)pl.NP * pl.N * 4;
    b += (int64_t)2 * pl.A * Mp * 4 + (int64_t)pl.N * Mp * 4 + (int64_t)pl.NP * pl.N * Mp * 4 + 8 * Mp * 4;
    b += pl.is_river ? (SS_NCOMBO * 4 + SS_NCOMBO * 2 + 52 * 53 * 2 + SS_BND_BYTES + (int64_t)pl.NP * Mp * 16) : 0;
    // exploitability buffers
    b += (int64_t)pl.A * Mp * 4 + (int64_t)(pl.NP + 2) * pl.N * Mp * 8;
    return b + 64 * 1024;
}

int PinnedStage::alloc(size_t n, void **out) {
    n = (n + 255) & ~(size_t)255;
    for (auto &c : chunks)
        if (c.cap - c.used >= n) {
            *out = c.p + c.used;
            c.used += n;
            return SS_OK;
        }
    Chunk c;
    c.cap = std::max(n, chunk_bytes);
    SS_HIP(hipHostMalloc((void **)&c.p, c.cap, hipHostMallocDefault));
    c.used = n;
    *out = c.p;
    chunks.push_back(c);
    return SS_OK;
}

int PinnedStage::upload(void *dst, const void *src, size_t n, hipStream_t st) {
    if (n == 0) return SS_OK;
    if (n > chunk_bytes / 2) {   // large (leaf values): straight from the caller's memory
        SS_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st));
        return SS_OK;
    }
    void *pin = nullptr;
    int rc = alloc(n, &pin);
    if (rc) return rc;
    std::memcpy(pin, src, n);
    SS_HIP(hipMemcpyAsync(dst, pin, n, hipMemcpyHostToDevice, st));
    return SS_OK;
}

PinnedStage::~PinnedStage() {
    for (auto &c : chunks) (void)hipHostFree(c.p);
}

// A thread's staging memory for single-subgame setups, kept between calls
// (pinning pages costs more than the copies it serves); each call waits for
// its copies before returning, so the next call can reuse it. Never freed:
// released with the process.
static PinnedStage &thread_stage() {
    static thread_local PinnedStage *stage = new PinnedStage();
    stage->reset();
    return *stage;
}

void PinnedStage::reset() {
    for (auto &c : chunks) c.used = 0;
}

int Slot::setup(const HostProblem &hp, hipStream_t st) {
    PinnedStage &stage = thread_stage();
    int rc = setup_begin(hp, st, stage);
    if (rc) return rc;
    SS_HIP(hipStreamSynchronize(st));
    rc = setup_finish(st, stage);
    if (rc) return rc;
    SS_HIP(hipStreamSynchronize(st));   // the stage's memory is reused by the next call
    return SS_OK;
}

// Plan, device arena, uploads of the plan's arrays and, on a river, the
// device evaluation of every combo's rank, read back into pinned memory
// (h_cstr): all queued on `st`, nothing waited for. setup_finish builds the
// river tables from those ranks once the caller has synchronised the stream.
int Slot::setup_begin(const HostProblem &hp, hipStream_t st, PinnedStage &stage) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc = make_plan(hp, plan);
    plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc) return rc;
    Plan &pl = plan;
    const size_t Mp = pl.Mp;
    ArenaLayout L;
    size_t o_nodes = L.take(sizeof(DevNode) * pl.N);
    size_t o_child = L.take(sizeof(int) * std::max<size_t>(1, pl.children.size()));
    size_t o_declvl = L.take(sizeof(int) * pl.dec_lvl.size());
    size_t o_decoff = L.take(sizeof(int) * pl.dec_off.size());
    size_t o_term = L.take(sizeof(int) * pl.term.size());
    size_t o_hinfo = L.take(sizeof(uint32_t) * pl.NP * Mp);
    size_t o_w = L.take(sizeof(float) * pl.NP * Mp);
    size_t o_leaf = L.take(sizeof(float) * std::max<size_t>(1, pl.leaf_count));
    size_t o_frozen = L.take(sizeof(int) * pl.N);
    size_t o_rrow = L.take(sizeof(int) * pl.rrow.size());
    size_t o_R = L.take(sizeof(float) * (size_t)pl.A * Mp);
    size_t o_SS = L.take(sizeof(float) * (size_t)pl.A * Mp);
    size_t o_cfv = L.take(sizeof(float) * (size_t)pl.N * Mp);
    size_t o_reach = L.take(sizeof(float) * (size_t)pl.NP * pl.N * Mp);
    size_t o_rsig = L.take(sizeof(float) * 8 * Mp);
    size_t o_cstr = L.take(sizeof(uint32_t) * SS_NCOMBO);
    size_t o_gsort = L.take(sizeof(uint16_t) * SS_NCOMBO);
    size_t o_clist = L.take(52 * 52);
    size_t o_clen = L.take(52);
    size_t o_bnd = pl.is_river ? L.take(SS_BND_BYTES) : 0;
    size_t o_cpos = L.take(52 * 52);
    size_t o_hmeta = pl.is_river ? L.take(sizeof(uint4) * pl.NP * Mp) : 0;
    size_t o_board = L.take(sizeof(int) * 8);
    arena_bytes = L.off;
    rc = slot_alloc(pool, arena_bytes, (void **)&arena);
    if (!rc) rc = slot_alloc(pool, sizeof(SlotDesc), (void **)&d_desc);
    if (rc) return rc;
    char *a = arena;
    SS_HIP(hipMemsetAsync(arena, 0, arena_bytes, st));
#define SS_UP(dst, src, n)                                   \
    do {                                                     \
        int rc_ = stage.upload((dst), (src), (n), st);       \
        if (rc_) return rc_;                                 \
    } while (0)
    // the tree and range sections (nodes .. weights) and the frozen actions
    // with the reach-row table (frozen .. rrow) are contiguous stretches of
    // the arena: one staged image each (padding and absent parts zero)
    {
        auto image = [&](size_t lo, size_t hi, std::initializer_list<std::pair<size_t, std::pair<const void *, size_t>>>
                                                    parts) -> int {
            void *pin = nullptr;
            int rc2 = stage.alloc(hi - lo, &pin);
            if (rc2) return rc2;
            char *img = (char *)pin;
            std::memset(img, 0, hi - lo);
            for (const auto &pt : parts)
                if (pt.second.second) std::memcpy(img + (pt.first - lo), pt.second.first, pt.second.second);
            SS_HIP(hipMemcpyAsync(a + lo, img, hi - lo, hipMemcpyHostToDevice, st));
            return SS_OK;
        };
        rc = image(o_nodes, o_leaf,
                   {{o_nodes, {pl.nodes.data(), sizeof(DevNode) * pl.N}},
                    {o_child, {pl.children.data(), sizeof(int) * pl.children.size()}},
                    {o_declvl, {pl.dec_lvl.data(), sizeof(int) * pl.dec_lvl.size()}},
                    {o_decoff, {pl.dec_off.data(), sizeof(int) * pl.dec_off.size()}},
                    {o_term, {pl.term.data(), sizeof(int) * pl.term.size()}},
                    {o_hinfo, {pl.hinfo.data(), sizeof(uint32_t) * pl.NP * Mp}},
                    {o_w, {pl.weights.data(), sizeof(float) * pl.NP * Mp}}});
        if (!rc)
            rc = image(o_frozen, o_R,
                       {{o_frozen, {pl.frozen_src, pl.frozen_src ? sizeof(int) * pl.N : 0}},
                        {o_rrow, {pl.rrow.data(), sizeof(int) * pl.rrow.size()}}});
        if (rc) return rc;
    }
    if (pl.has_leaf) SS_UP(a + o_leaf, pl.leaf_src, sizeof(float) * pl.leaf_count);
    // root reach = raw weights (never overwritten: the forward sweep writes children only)
    for (int p = 0; p < pl.NP; ++p)
        SS_UP(a + o_reach + sizeof(float) * (size_t)p * pl.N * Mp, pl.weights.data() + (size_t)p * Mp,
              sizeof(float) * Mp);
    SlotDesc &d = desc;
    std::memset(&d, 0, sizeof(d));
    d.N = pl.N; d.NP = pl.NP; d.Mp = pl.Mp; d.maxh = pl.maxh;
    d.A = pl.A; d.L = pl.L; d.T = pl.T; d.D = pl.D;
    d.is_river = pl.is_river; d.has_leaf = pl.has_leaf; d.num_leaves = pl.num_leaves;
    d.root_player = pl.root_player;
    for (int p = 0; p < 6; ++p) d.nh[p] = p < pl.NP ? pl.nh[p] : 0;
    d.odometer = pl.odometer;
    d.uniq = pl.uniq;
    d.nodes = (const DevNode *)(a + o_nodes);
    d.children = (const int *)(a + o_child);
    d.dec_lvl = (const int *)(a + o_declvl);
    d.dec_off = (const int *)(a + o_decoff);
    d.term = (const int *)(a + o_term);
    d.hinfo = (const uint32_t *)(a + o_hinfo);
    d.weights = (const float *)(a + o_w);
    d.leaf_values = pl.has_leaf ? (const float *)(a + o_leaf) : nullptr;
    d.frozen = pl.frozen_src ? (const int *)(a + o_frozen) : nullptr;
    d.rrow = (const int *)(a + o_rrow);
    d.R = (float *)(a + o_R);
    d.SS = (float *)(a + o_SS);
    d.cfv = (float *)(a + o_cfv);
    d.reach = (float *)(a + o_reach);
    d.rsig = (float *)(a + o_rsig);
    d.cstr = (const uint32_t *)(a + o_cstr);
    d.gsort = (const uint16_t *)(a + o_gsort);
    d.clist = (const uint8_t *)(a + o_clist);
    d.clen = (const uint8_t *)(a + o_clen);
    d.bnd = pl.is_river ? (const uint8_t *)(a + o_bnd) : nullptr;
    d.cpos = (const uint8_t *)(a + o_cpos);
    d.hmeta = pl.is_river ? (const uint4 *)(a + o_hmeta) : nullptr;
    d.active = 1;

    h_cstr = nullptr;
    off_cstr = o_cstr; off_gsort = o_gsort; off_clist = o_clist; off_clen = o_clen; off_bnd = o_bnd;
    off_cpos = o_cpos; off_hmeta = o_hmeta;
    if (pl.is_river && pl.num_board == 5) {
        // ranks of all combos on this board, evaluated on the device, read
        // back into pinned memory without waiting
        int *dboard = (int *)(a + o_board);
        SS_UP(dboard, pl.board, sizeof(int) * 5);
        rc = launch_combo_ranks(dboard, (uint32_t *)(a + o_cstr), st);
        if (rc) return rc;
        void *pin = nullptr;
        rc = stage.alloc(sizeof(uint32_t) * SS_NCOMBO, &pin);
        if (rc) return rc;
        h_cstr = (uint32_t *)pin;
        SS_HIP(hipMemcpyAsync(h_cstr, a + o_cstr, sizeof(uint32_t) * SS_NCOMBO, hipMemcpyDeviceToHost, st));
    } else {
        d.is_river = 0;   // showdowns off the river carry no ranks: value 0, as street_solve.cu:1520
    }
    return SS_OK;
}

// The host half of setup_finish: the river tables from the ranks setup_begin
// read back (the stream has been synchronised since) and every player's
// hands by ascending strength (dead hands last), one record per hand with
// everything a showdown sweep needs.
int Slot::prepare_tables() {
    if (!h_cstr || h_rt) return SS_OK;
    try {
        const Plan &pl = plan;
        const size_t Mp = pl.Mp;
        std::unique_ptr<RiverTables> rt(new RiverTables);
        build_river_tables(h_cstr, pl.lab, *rt);
        std::vector<uint4> hm((size_t)pl.NP * Mp, make_uint4(0u, 0u, 0u, 0u));
        for (int p = 0; p < pl.NP; ++p) {
            std::vector<int> idx(pl.nh[p]);
            std::iota(idx.begin(), idx.end(), 0);
            auto rank = [&](int h) -> uint32_t {
                uint32_t hi = pl.hinfo[(size_t)p * Mp + h];
                return (hi & SS_LIVE_BIT) ? rt->cstr[(hi >> 12) & 2047] : 0u;
            };
            auto key = [&](int h) -> uint32_t { uint32_t r = rank(h); return r ? r : 0xFFFFFFFFu; };
            std::stable_sort(idx.begin(), idx.end(), [&](int a2, int b2) { return key(a2) < key(b2); });
            for (size_t i = 0; i < idx.size(); ++i) {
                const int h = idx[i];
                const uint32_t hi = pl.hinfo[(size_t)p * Mp + h];
                const int kk = (hi >> 12) & 2047;
                const uint32_t ranked = rank(h) ? 1u : 0u;
                uint4 r;
                r.x = (uint32_t)h | ((hi & 63) << 11) | (((hi >> 6) & 63) << 17) | (ranked << 23);
                r.y = ranked ? rt->kcls[kk] : 0u;
                r.z = ranked ? (uint32_t)rt->glt[kk] : 0u;
                r.w = ranked ? (uint32_t)rt->gle[kk] : 0u;
                hm[(size_t)p * Mp + i] = r;
            }
        }
        h_rt = std::move(rt);
        h_hmeta.swap(hm);
    } catch (const std::bad_alloc &) {
        return set_error(SS_ERR_NOMEM, "out of host memory building the river tables");
    }
    return SS_OK;
}

void prepare_tables_all(const std::vector<Slot *> &slots) {
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i; (i = next.fetch_add(1)) < slots.size();)
            if (slots[i]) (void)slots[i]->prepare_tables();
    };
    const int want = std::min(8, std::max(1, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    try {
        for (int t = 1; t < want && (size_t)t < slots.size(); ++t) th.emplace_back(work);
    } catch (...) {   // no more threads: the calling thread takes the rest
    }
    work();
    for (auto &t : th) t.join();
}

// The river tables (prepare_tables, unless already run) uploaded as one image
// of their contiguous stretch of the arena, then the descriptor.
int Slot::setup_finish(hipStream_t st, PinnedStage &stage) {
    SlotDesc &d = desc;
    char *a = arena;
    if (h_cstr) {
        int rc = prepare_tables();
        if (rc) return rc;
        const RiverTables &rt = *h_rt;
        d.ng = rt.ng;
        // gsort, clist, clen, bnd, cpos, hmeta follow each other in the arena
        const size_t lo = off_gsort, hi = off_hmeta + sizeof(uint4) * h_hmeta.size();
        void *pin = nullptr;
        rc = stage.alloc(hi - lo, &pin);
        if (rc) return rc;
        char *img = (char *)pin;
        std::memset(img, 0, hi - lo);
        std::memcpy(img + (off_gsort - lo), rt.gsort.data(), sizeof(uint16_t) * rt.gsort.size());
        std::memcpy(img + (off_clist - lo), rt.clist.data(), 52 * 52);
        std::memcpy(img + (off_clen - lo), rt.clen.data(), 52);
        std::memcpy(img + (off_bnd - lo), rt.bnd.data(), SS_BND_BYTES);
        std::memcpy(img + (off_cpos - lo), rt.cpos.data(), 52 * 52);
        std::memcpy(img + (off_hmeta - lo), h_hmeta.data(), sizeof(uint4) * h_hmeta.size());
        SS_HIP(hipMemcpyAsync(a + lo, img, hi - lo, hipMemcpyHostToDevice, st));
        h_cstr = nullptr;
        h_rt.reset();
        std::vector<uint4>().swap(h_hmeta);
    }
    SS_UP(d_desc, &d, sizeof(d));
    return SS_OK;
}

int Slot::alloc_expl(hipStream_t st, int blocks) {
    PinnedStage &stage = thread_stage();
    int rc = alloc_expl(st, blocks, stage);
    if (rc) return rc;
    SS_HIP(hipStreamSynchronize(st));   // the stage's memory is reused by the next call
    return SS_OK;
}

int Slot::alloc_expl(hipStream_t st, int blocks, PinnedStage &stage) {
    if (xarena && xscratch_blocks >= blocks) return SS_OK;
    if (xarena && !pool) SS_HIP(hipFree(xarena));   // pool memory stays until the pool goes
    xarena = nullptr;
    const size_t Mp = plan.Mp;
    ArenaLayout L;
    size_t o_avg = L.take(sizeof(float) * (size_t)plan.A * Mp);
    size_t o_xr = L.take(sizeof(double) * (size_t)plan.NP * plan.N * Mp);
    size_t o_xev = L.take(sizeof(double) * (size_t)plan.N * Mp);
    size_t o_xbr = L.take(sizeof(double) * (size_t)plan.N * Mp);
    size_t o_xout = L.take(sizeof(double) * 32);
    size_t o_scr = L.take(sizeof(double) * (size_t)blocks * SS_XSCRATCH_DOUBLES);
    xarena_bytes = L.off;
    {
        int rc = slot_alloc(pool, xarena_bytes, (void **)&xarena);
        if (rc) return rc;
    }
    SS_HIP(hipMemsetAsync(xarena, 0, xarena_bytes, st));
    desc.avg = (float *)(xarena + o_avg);
    desc.xreach = (double *)(xarena + o_xr);
    desc.xev = (double *)(xarena + o_xev);
    desc.xbr = (double *)(xarena + o_xbr);
    desc.xout = (double *)(xarena + o_xout);
    xscratch = (double *)(xarena + o_scr);
    xscratch_blocks = blocks;
    // root rows of xreach = raw weights, as doubles
    std::vector<double> w(Mp);
    for (int p = 0; p < plan.NP; ++p) {
        for (size_t h = 0; h < Mp; ++h) w[h] = plan.weights[p * Mp + h];
        SS_UP(desc.xreach + (size_t)p * plan.N * Mp, w.data(), sizeof(double) * Mp);
    }
    // refresh the device copy of the descriptor, keeping device-side progress
    SS_UP((char *)d_desc + offsetof(SlotDesc, avg), &desc.avg, sizeof(SlotDesc) - offsetof(SlotDesc, avg));
#undef SS_UP
    return SS_OK;
}

double expl_from_xout(const double *xout, int np, int starting_pot, double *gains) {
    double total = 0;
    for (int p = 0; p < np; ++p) {
        total += xout[p];
        if (gains) gains[p] = xout[p];
    }
    double J = xout[2 * np];
    if (J <= 0 || starting_pot <= 0) return -1;
    return 100.0 * (total / np) / J / (double)starting_pot;
}

// ---------------------------------------------------------------------------
// outputs in the reference layout (street_solve.cu:1556-1612)
// ---------------------------------------------------------------------------
int next_check_iter(int t, double e, double target, int every, int step, bool adaptive) {
    if (!adaptive || !(e > target) || target <= 0) return t + every;
    // exploitability of Linear CFR falls roughly as t^-a; over doubling
    // intervals of the batch mix a stays below 2.2 for 99% of (subgame,
    // interval) pairs, so assuming a = 2 rarely schedules a check after the
    // crossing
    const double r = std::sqrt(e / target);
    const double tn = std::min((double)t + 4.0 * every, (double)t * r);
    const int next = (int)(tn / step) * step;
    return std::max(t + step, std::min(t + 4 * every, next));
}

int check_step(int flags, int every, int fine_step, bool *adaptive) {
    *adaptive = (flags & (SS_FLAG_ADAPTIVE_CHECKS | SS_FLAG_SPARSE_CHECKS)) != 0;
    return (flags & SS_FLAG_SPARSE_CHECKS) ? std::max(1, every) : std::max(1, fine_step);
}

int read_output(Slot &s, const SlotDesc &dnow, hipStream_t st, SSOutput *out, int) {
    const Plan &pl = s.plan;
    const size_t Mp = pl.Mp;
    const int maxh = pl.maxh, NP = pl.NP;
    // strategy sums, root snapshot and root values copied into this thread's
    // pinned staging memory (a pageable destination would be staged by the
    // runtime in synchronous chunks)
    PinnedStage &stage = thread_stage();
    const size_t nss = (size_t)pl.A * Mp;
    void *p_ss = nullptr, *p_rsig = nullptr, *p_cfv = nullptr;
    int rc = stage.alloc(sizeof(float) * std::max<size_t>(1, nss), &p_ss);
    if (!rc) rc = stage.alloc(sizeof(float) * 8 * Mp, &p_rsig);
    if (!rc) rc = stage.alloc(sizeof(float) * Mp, &p_cfv);
    if (rc) return rc;
    const float *SS = (const float *)p_ss, *rsig = (const float *)p_rsig, *cfv0 = (const float *)p_cfv;
    if (nss) SS_HIP(hipMemcpyAsync(p_ss, s.desc.SS, sizeof(float) * nss, hipMemcpyDeviceToHost, st));
    SS_HIP(hipMemcpyAsync(p_rsig, s.desc.rsig, sizeof(float) * 8 * Mp, hipMemcpyDeviceToHost, st));
    SS_HIP(hipMemcpyAsync(p_cfv, s.desc.cfv, sizeof(float) * Mp, hipMemcpyDeviceToHost, st));
    SS_HIP(hipStreamSynchronize(st));
    (void)dnow;
    std::memset(out, 0, sizeof(*out));
    const DevNode &root = pl.nodes[0];
    int root_player = root.player;
    int root_na = root.nchild;
    int root_nh = (root_player >= 0 && root_player < NP) ? pl.nh[root_player] : 0;
    out->root_num_actions = root_na;
    out->root_player = root_player;
    out->root_strategy = (float *)std::malloc(sizeof(float) * std::max(1, root_na * root_nh));
    for (int a = 0; a < root_na; ++a)
        for (int h = 0; h < root_nh; ++h) out->root_strategy[a * root_nh + h] = rsig[a * Mp + h];
    out->root_ev = (float *)std::calloc((size_t)NP * maxh, sizeof(float));
    const int last_nh = pl.nh[NP - 1];
    for (int h = 0; h < maxh; ++h) out->root_ev[h] = h < last_nh ? cfv0[h] : 0.f;
    out->num_players = NP;
    out->max_hands = maxh;
    out->num_avg_nodes = pl.D;
    out->avg_strategy_node_ids = (int *)std::malloc(sizeof(int) * std::max(1, pl.D));
    out->avg_strategies = (float *)std::calloc((size_t)std::max(1, pl.D) * SS_MAX_ACTIONS * maxh, sizeof(float));
    int di = 0;
    for (int n = 0; n < pl.N; ++n) {
        const DevNode &nd = pl.nodes[n];
        if (nd.type != SS_NODE_DECISION) continue;
        out->avg_strategy_node_ids[di] = n;
        int q = nd.player;
        if (q >= 0 && q < NP) {
            int na = nd.nchild;
            for (int h = 0; h < pl.nh[q]; ++h) {
                float sum = 0.f;
                for (int a = 0; a < na; ++a) {
                    float v = SS[(size_t)(nd.act_base + a) * Mp + h];
                    sum += v > 0.f ? v : 0.f;
                }
                float inv = sum > 0.f ? 1.0f / sum : 1.0f / (float)na;
                for (int a = 0; a < na; ++a) {
                    float v = SS[(size_t)(nd.act_base + a) * Mp + h];
                    if (v < 0.f) v = 0.f;
                    out->avg_strategies[((size_t)di * SS_MAX_ACTIONS + a) * maxh + h] = sum > 0.f ? v * inv : inv;
                }
            }
        }
        ++di;
    }
    return SS_OK;
}

}  // namespace ss

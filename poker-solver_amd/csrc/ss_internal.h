// ss_internal.h -- host-side declarations shared by the library's sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/street_solve.h"
#include "ss_device.h"

namespace ss {

int set_error(int code, const char *fmt, ...);
void clear_error();
int log_level();

#define SS_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return ::ss::set_error(SS_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                   __FILE__, __LINE__);                                          \
    } while (0)

int build_tree(const int *board, int num_board, int starting_pot, int effective_stack,
               const float *bet_sizes, int num_bet_sizes, int num_players, const int *acting_order,
               int is_river, int use_cont_strats, int max_raises, SSTreeData *out);
void free_tree(SSTreeData *td);

// Input of one solve, independent of which ABI delivered it.
struct HostProblem {
    const SSTreeData *tree = nullptr;
    int np = 0;
    int nh[SS_MAX_PLAYERS] = {0, 0, 0, 0, 0, 0};
    const int *hands[SS_MAX_PLAYERS] = {};     // [nh][2]
    const float *weights[SS_MAX_PLAYERS] = {};
    const float *leaf_values = nullptr;          // [num_leaves][np][maxh]
    const int *frozen = nullptr;                 // [num_nodes]
};

HostProblem from_legacy(const SSTreeData *td);
int from_ex(const SSProblem *p, HostProblem *out);

// Host-side plan of a slot: every device array, filled before upload.
struct Plan {
    int N = 0, NP = 0, Mp = 0, maxh = 0, A = 0, L = 0, T = 0, D = 0;
    int is_river = 0, has_leaf = 0, num_leaves = 0, root_player = 0, ng = 0, odometer = 0, uniq = 1;
    int nh[6] = {0, 0, 0, 0, 0, 0};
    int board[5] = {0, 0, 0, 0, 0};
    int num_board = 0;
    // Card numbering of the device tables: the cards off the board first
    // (ascending), the board cards last, so loops over the cards that can be
    // in a hand stop early. Combo ids (strength lookups) stay real.
    int lab[52] = {0};
    int starting_pot = 0;
    std::vector<DevNode> nodes;
    std::vector<int> children, dec_lvl, dec_off, term, dec_index;
    std::vector<uint32_t> hinfo;
    std::vector<float> weights;
    const float *leaf_src = nullptr;
    size_t leaf_count = 0;
    const int *frozen_src = nullptr;
    std::vector<int> frozen;       // host copy of frozen_src (empty when none)
    std::vector<int> rrow;         // [NP][N] reach row of player p at node n (see make_plan)
    int64_t algo_bytes_iter = 0;   // algorithmic bytes of one iteration
};

int make_plan(const HostProblem &hp, Plan &pl);

// River tables built on the host from device-computed combo ranks.
struct RiverTables {
    std::vector<uint32_t> cstr;   // [1326]
    std::vector<uint16_t> gsort;          // live combos by ascending rank, packed a | b << 8
    std::vector<int16_t> glt, gle;
    std::vector<uint8_t> clist, clen, cpos;
    std::vector<uint16_t> kcls;   // [1326] strength class
    std::vector<uint8_t> bnd;     // [(ncls+1)*52] class boundaries per card
    int ng = 0, ncls = 0;
};
void build_river_tables(const uint32_t *cstr, const int *lab, RiverTables &rt);

// Precision-form buffers of a 2-player river subgame (ss_precision.hip).
struct PrecState {
    int S = 0, NC = 0, MpK = 0;           // showdown nodes; column tiles ceil(2 S / 16); Mp rounded to 256
    char *block = nullptr;                // one device allocation holding every buffer below
    int *sd_nodes = nullptr;              // [S]
    uint32_t *W[2] = {nullptr, nullptr};  // per traverser: W (+1 / -1 / 0) as lane-ordered FP4 fragments
    uint16_t *Bt = nullptr;               // [16 NC][MpK] bf16 opponent reach, hi and lo columns packed
    float *Z = nullptr;                   // [S][Mp] card-removal mass per node and traverser hand
    int2 *zi[2] = {nullptr, nullptr};     // [Mp] per hand: cards and validity, opponent hand with its cards
    float4 *C = nullptr;                  // [2][S] per traverser and node: P/2, traverser bet, node id, opponent reach row
    ~PrecState();
};

// Page-locked host staging for uploads: upload() copies host data into
// pinned memory and queues the copy, so it does not wait for the stream (a
// copy from pageable memory does); large blocks go straight from the
// caller's memory. Keep the stage alive until the stream is synchronised.
struct PinnedStage {
    struct Chunk {
        char *p = nullptr;
        size_t cap = 0, used = 0;
    };
    std::vector<Chunk> chunks;
    size_t chunk_bytes = (size_t)8 << 20;
    PinnedStage() = default;
    PinnedStage(const PinnedStage &) = delete;
    PinnedStage &operator=(const PinnedStage &) = delete;
    ~PinnedStage();
    int alloc(size_t n, void **out);   // n bytes of pinned memory (256-byte aligned)
    void reset();                      // every chunk free again (its copies have completed)
    int upload(void *dst, const void *src, size_t n, hipStream_t st);
};

// Device memory for many slots from a few large allocations: a batch's
// subgames are created and destroyed together, and a hipMalloc / hipFree
// per slot costs tens of microseconds (hipFree also waits for the device).
struct DevicePool {
    struct Chunk {
        char *p = nullptr;
        size_t cap = 0, used = 0;
    };
    std::vector<Chunk> chunks;
    size_t chunk_bytes = (size_t)256 << 20;
    DevicePool() = default;
    DevicePool(const DevicePool &) = delete;
    DevicePool &operator=(const DevicePool &) = delete;
    ~DevicePool();
    int alloc(size_t n, void **out);   // n bytes of device memory (256-byte aligned)
};

// A subgame resident on the device.
struct Slot {
    std::unique_ptr<PrecState> prec;   // precision form only
    Plan plan;
    SlotDesc desc{};              // host copy (pointers are device pointers)
    SlotDesc *d_desc = nullptr;   // device copy (plain form); batched forms copy desc into their array
    char *arena = nullptr;
    size_t arena_bytes = 0;
    char *xarena = nullptr;       // FP64 exploitability buffers
    size_t xarena_bytes = 0;
    double *xscratch = nullptr;   // per-workgroup FP64 terminal scratch
    int xscratch_blocks = 0;
    double plan_ms = 0;           // host time of make_plan in setup
    DevicePool *pool = nullptr;   // when set, the device buffers come from it (and are not freed here)
    uint32_t *h_cstr = nullptr;   // combo ranks read back by setup_begin (pinned, in the stage)
    size_t off_cstr = 0, off_gsort = 0, off_clist = 0, off_clen = 0, off_bnd = 0, off_cpos = 0, off_hmeta = 0;
    // river tables and per-hand showdown records built on the host from
    // h_cstr by prepare_tables, held until setup_finish uploads them
    std::unique_ptr<RiverTables> h_rt;
    std::vector<uint4> h_hmeta;

    ~Slot();
    // setup = setup_begin, a stream synchronisation, setup_finish; a batch
    // runs every subgame's setup_begin before one synchronisation
    int setup(const HostProblem &hp, hipStream_t st);
    int setup_begin(const HostProblem &hp, hipStream_t st, PinnedStage &stage);
    int setup_finish(hipStream_t st, PinnedStage &stage);
    // the host half of setup_finish (no HIP calls, touches only this slot):
    // a batch runs it for many subgames on several host threads at once
    int prepare_tables();
    int alloc_expl(hipStream_t st, int blocks);
    int alloc_expl(hipStream_t st, int blocks, PinnedStage &stage);
    int64_t footprint() const { return (int64_t)arena_bytes + (int64_t)xarena_bytes; }
};
// prepare_tables of every listed slot (null entries skipped) on up to eight
// host threads; a failure is left for setup_finish, which repeats and reports
void prepare_tables_all(const std::vector<Slot *> &slots);

int64_t estimate_footprint(const HostProblem &hp);

// Leaf values evaluated from equity straight into a device leaf buffer
// [num_leaves][np][mh] (ss_leaf.hip); row_of (host, may be null) maps tree
// leaves to profile rows of the ss_leaf_values layout.
int leaf_values_into(float *d_leaf, int num_leaves, int np, int mh, const double *raw, const int *nh,
                     const int *leaf_idx, const int *leaf_pot, int ninfo, const int *row_of, hipStream_t st);

// Range narrowing from a resident solve's strategy sums (ss_range.hip).
int range_update_from_strategy(const float *d_SS, int row0, int na, int action, int Mp, int nsolver,
                               const int *solver_hand, double *weights, int n, double floor_w, double threshold,
                               hipStream_t st);

// plain form
int plain_run(Slot &s, int iters, int final_iter, hipStream_t st, double *kernel_ms);
// Iteration of a subgame's next exploitability check after a check at
// iteration t found e > target (% pot): t + every, or with adaptive checks
// the iteration where e would reach the target if it fell as t^-2, rounded
// down to a multiple of step, at least t + step and at most t + 4 * every.
int next_check_iter(int t, double e, double target, int every, int step, bool adaptive);
// The check grid of a solve: `step` for SS_FLAG_ADAPTIVE_CHECKS (the form's
// iteration step), check_every for SS_FLAG_SPARSE_CHECKS; *adaptive tells
// whether either is set.
int check_step(int flags, int every, int fine_step, bool *adaptive);
int plain_exploitability(Slot &s, hipStream_t st, double *expl_pct, double *gains);
int read_output(Slot &s, const SlotDesc &desc_now, hipStream_t st, SSOutput *out, int legacy_root_ev);

// batched helpers
int launch_batched(SlotDesc *d_slots, int nslots, int iters, int grid, hipStream_t st);
int launch_batched_expl(SlotDesc *d_slots, int nslots, double *scratch, const int *d_want, int grid,
                        hipStream_t st);
double expl_from_xout(const double *xout, int np, int starting_pot, double *gains);

// phased batched form: (slot << 20 | node) task lists on the device
struct PhasedTasks {
    int nslots = 0, NPmax = 0, L = 0;
    std::vector<int> lvl_off;      // level l tasks: [lvl_off[l], lvl_off[l+1]); all decision tasks [0, lvl_off[L])
    int term_off = 0, nterm = 0;   // terminal tasks, most expensive first
    int nodo = 0;                  // terminal tasks [0, nodo): three or more opponents for some traverser
    int npair = 0;                 // [nodo, nodo + npair): two opponents (showdowns first); the rest at most one
    int npair_sd = 0;              // showdowns among the npair
    int nprec = 0;                 // the last nprec terminal tasks: showdowns of precision-form slots
                                   // (FP64 checks only; the CFR pass takes them from the contraction)
    int psplit_off = 0, npsplit = 0;   // precision form: (slot, showdown index) tasks
    struct PrecGroup { int nc, rt, count; };   // column tiles, row tiles per wavefront, tasks
    std::vector<PrecGroup> pgemm;              // groups of (slot, row block) precision GEMM tasks
    int pgemm_off = 0;
    // Per traverser pass t < NPmax, only the tasks that pass has work in (slots
    // with more than t players; the forward sweep also drops nodes where t
    // acts, whose child rows it does not write). Sweep records: forward level
    // l = [sw_off[(2t)(L+1) + l], sw_off[(2t)(L+1) + l + 1]), backward levels
    // at (2t+1)(L+1). Terminal tasks: pterm[t] segments of d_tasks.
    std::vector<int> sw_off;
    struct PassTerm {
        int off = 0, nodo = 0, nsd = 0, nfold = 0, nlight = 0;
        int sd_parts = 1;   // workgroups per two-opponent showdown task
        int small_parts = 0;   // > 0: the pass's closed-form terminals as one launch, showdowns split this many ways
        bool rsig = false;     // some listed subgame has t + 1 players (its root snapshot pass)
    };
    std::vector<PassTerm> pterm;
    // Launch shape of this task list (the batch autotuner's choice): forward
    // levels with >= fwd_group_min tasks run three tasks per workgroup (-1:
    // the automatic threshold), term_small 0 keeps the terminal classes in
    // separate launches (-1: one launch when they fit), sd_parts workgroups
    // per two-opponent showdown (-1: automatic), sweep_grid the sweep
    // launches' workgroup cap (0: the default). Every shape gives the same
    // results bit for bit.
    struct Shape {
        int fwd_group_min = -1, term_small = -1, sd_parts = -1, sweep_grid = 0;
    } shape;
    uint32_t *d_tasks = nullptr;
    SweepRec *d_recs = nullptr;    // the per-pass sweep records
    size_t cap_tasks = 0, cap_recs = 0;   // allocated lengths of d_tasks, d_recs
    ~PhasedTasks();
};
// tasks of the non-null slots with live[slot] != 0 (all of them when live is null)
int dbg_copy(int64_t bytes, int repeats, float *ms);
int prec_setup(Slot &s, hipStream_t st);
int prec_showdowns(Slot &s, int t, hipStream_t st);
struct PhasedTasks;
int launch_prec_phased(SlotDesc *d_slots, const PhasedTasks &pt, int t, hipStream_t st);
int64_t prec_flops_per_pass(const Slot &s);
// Row tiles per wavefront of the batched precision GEMM for subgames of nc
// column tiles with the given row tile counts (the choice that loads the
// most loaded compute unit least); a task covers 64 rt rows.
int prec_choose_rt(int nc, const std::vector<int> &tiles);

int build_phased_tasks(const std::vector<Slot *> &slots, PhasedTasks &pt, hipStream_t st,
                       const std::vector<int> *live = nullptr);
int launch_phased_iteration(SlotDesc *d_slots, const PhasedTasks &pt, hipStream_t st);
// launch shapes of the phased form forced for every task list (-1: not
// forced); see ssdbg_set_phased_shape
void set_phased_shape(int fwd_group_min, int term_small, int sd_parts);
bool phased_shape_forced();
// joint = false skips the joint deal mass (xout[2 NP]), which depends only on
// the raw weights: a caller whose xout persists computes it once
int launch_phased_expl(SlotDesc *d_slots, const PhasedTasks &pt, const int *d_want, double *scratch,
                       int scratch_blocks, hipStream_t st, bool joint = true);
#define SS_XSCRATCH_DOUBLES (5 * 2704 + 3 * 1328 + 6 * 52 + 64)

// debug / test entry points implemented in the kernels unit
int dbg_eval7_device(const int *cards, int n, uint32_t *out);
int dbg_combo_ranks_device(const int *board, uint32_t *out);
int dbg_eval_census_device(unsigned long long counts[10], unsigned long long *sum);

}  // namespace ss

// ss_capi.hip -- extern "C" entry points of libstreet_solve.so (include/street_solve.h).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/street_solve_debug.h"
#include "ss_eval.h"
#include "ss_internal.h"
#include "../../include/street_solve_range.h"

namespace ss {
const char *last_error_str();

struct StreamGuard {
    hipStream_t st = nullptr;
    int rc = SS_OK;
    StreamGuard() {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            rc = set_error(SS_ERR_HIP, "hipStreamCreate failed");
    }
    ~StreamGuard() {
        if (st) (void)hipStreamDestroy(st);
    }
};

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Plain form with optional exploitability-targeted early stop.
static int solve_plain(const HostProblem &hp, const SSSolveConfig &cfg, SSOutput *out, SSReport *rep) {
    StreamGuard g;
    if (g.rc) return g.rc;
    double t0 = now_ms();
    Slot s;
    int rc = s.setup(hp, g.st);
    if (rc) return rc;
    if (cfg.precision == SS_PREC_BF16X2) rc = prec_setup(s, g.st);
    else if (cfg.precision != SS_PREC_FP32) rc = set_error(SS_ERR_ARG, "unknown precision %d", cfg.precision);
    if (rc) return rc;
    const int T = std::max(0, cfg.max_iterations);
    const bool target = cfg.target_exploitability > 0;
    const int chunk = target ? std::max(1, cfg.check_every > 0 ? cfg.check_every : 50) : std::max(1, T);
    double kms = 0, expl = -1, gains[SS_MAX_PLAYERS] = {0, 0, 0, 0, 0, 0};
    int done = 0, converged = 0, expl_at = -1, due = chunk;
    int gstep = std::min(chunk, 10);   // the batched form's graph step: the largest divisor of chunk <= 10
    while (chunk % gstep) --gstep;
    bool adaptive = false;
    const int astep = check_step(cfg.flags, chunk, gstep, &adaptive);
    adaptive = adaptive && target;
    const double t_loop = now_ms();
    while (done < T) {
        int k = std::min(due - done, T - done);
        rc = plain_run(s, k, done + k, g.st, &kms);
        if (rc) return rc;
        done += k;
        if (target) {
            rc = plain_exploitability(s, g.st, &expl, gains);
            if (rc) return rc;
            expl_at = done;
            if (expl >= 0 && expl <= cfg.target_exploitability) {
                converged = 1;
                break;
            }
            due = next_check_iter(done, expl, cfg.target_exploitability, chunk, astep, adaptive);
        } else {
            due = done + chunk;
        }
    }
    if (!(cfg.flags & SS_FLAG_NO_EXPL) && expl_at != done) {
        rc = plain_exploitability(s, g.st, &expl, gains);
        if (rc) return rc;
    }
    const double loop_ms = now_ms() - t_loop;   // every step above ends in a stream sync
    (void)kms;
    if (out) {
        rc = read_output(s, s.desc, g.st, out, 0);
        if (rc) return rc;
    }
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->iterations = done;
        rep->converged = converged;
        rep->exploitability = expl;
        for (int p = 0; p < SS_MAX_PLAYERS; ++p) rep->nashconv[p] = gains[p];
        rep->elapsed_ms = now_ms() - t0;
        rep->kernel_ms = loop_ms;
        rep->algo_bytes = s.plan.algo_bytes_iter * done;
    }
    return SS_OK;
}

// The make-up of a batch, for reusing an autotuned choice: batches drawn from
// one distribution have the same subgame and player counts and task counts
// within a few percent, so a cached choice matches when those agree and
// both task counts are within 25%.
struct Makeup {
    int n = 0, ndec = 0, nterm = 0, npmax = 0, precision = 0;
    bool near(const Makeup &o) const {
        auto close = [](int a, int b) { return 4 * std::abs(a - b) <= std::max(a, b); };
        return n == o.n && npmax == o.npmax && precision == o.precision && close(ndec, o.ndec) &&
               close(nterm, o.nterm);
    }
};
template <typename V> struct MakeupCache {
    std::vector<std::pair<Makeup, V>> entries;
    const V *find(const Makeup &m) const {
        for (auto &e : entries)
            if (e.first.near(m)) return &e.second;
        return nullptr;
    }
    void put(const Makeup &m, const V &v) {
        for (auto &e : entries)
            if (e.first.near(m)) { e.second = v; return; }
        entries.emplace_back(m, v);
    }
};

// Batched form: all subgames advance together, one persistent launch per
// iteration; `gsteps` iterations are captured in a hipGraph and replayed
// chunk / gsteps times between checks.
struct BatchImpl {
    SSSolveConfig cfg{};
    int n = 0, chunk = 1;
    bool target = false, want_final = false;
    std::vector<std::unique_ptr<Slot>> slots;
    std::unique_ptr<DevicePool> pool;   // the slots' device memory, in a few large allocations
    std::vector<SlotDesc> templ;
    SlotDesc *d_slots = nullptr;
    int *d_want = nullptr;
    double *scratch = nullptr;
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    PhasedTasks pt;
    int mode = 1;          // 0: persistent kernel, one workgroup per subgame; 1: phased launches
    int full_tasks = 0;
    bool use_graph = true; // SS_NO_GRAPH=1 issues the same launches directly (profiling)
    int xblocks = 512;     // workgroups (and FP64 scratch slices) of the exploitability sweeps
    int gsteps = 1;        // iterations per captured graph (a divisor of chunk, at most 10)
    double *d_xout = nullptr;     // every subgame's exploitability terms, 32 doubles each
    std::vector<double> h_xout;   // (one read back per check)
    std::vector<char> joint_done; // the joint deal mass (weights only) is in d_xout
    // results of the last run
    std::vector<int> iters, conv, expl_at;
    std::vector<double> expl, gains;

    ~BatchImpl() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        slots.clear();
        pool.reset();   // every slot's device memory
        if (d_slots) (void)hipFree(d_slots);
        if (d_want) (void)hipFree(d_want);
        if (scratch) (void)hipFree(scratch);
        if (d_xout) (void)hipFree(d_xout);
        if (st) (void)hipStreamDestroy(st);
    }

    // every subgame's river tables built on host threads
    void prepare_all() {
        std::vector<Slot *> raw;
        for (auto &sl : slots) raw.push_back(sl.get());
        prepare_tables_all(raw);
    }

    int create(const HostProblem *hps, int count, const SSSolveConfig &c) {
        if (c.precision != SS_PREC_FP32 && c.precision != SS_PREC_BF16X2)
            return set_error(SS_ERR_ARG, "unknown precision %d", c.precision);
        cfg = c;
        const bool prec = c.precision == SS_PREC_BF16X2;
        n = count;
        const int T = std::max(0, cfg.max_iterations);
        target = cfg.target_exploitability > 0;
        want_final = !(cfg.flags & SS_FLAG_NO_EXPL);
        chunk = target ? std::max(1, cfg.check_every > 0 ? cfg.check_every : 50) : std::max(1, std::min(T, 50));
        // a graph of a few iterations replayed chunk / gsteps times: the same
        // launches, with a capture (redone at every task compaction) a
        // fraction of the size
        for (gsteps = std::min(chunk, 10); chunk % gsteps; --gsteps) {}
        SS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        for (auto &e : ev) SS_HIP(hipEventCreate(&e));
        slots.resize(n);
        const double c0 = now_ms();
        double cb = c0, cw = c0, cp = c0, cf = c0;   // setup phases: queued, ranks back, tables built, queued
        {
            // every subgame's plan, arena and uploads queued, one wait for the
            // combo ranks, then the river tables and the FP64 buffers queued
            PinnedStage stage;
            pool.reset(new DevicePool());
            for (int i = 0; i < n; ++i) {
                slots[i].reset(new Slot());
                slots[i]->pool = pool.get();
                int rc = slots[i]->setup_begin(hps[i], st, stage);
                if (rc) return set_error(rc, "subgame %d: %s", i, last_error_str());
            }
            cb = now_ms();
            SS_HIP(hipStreamSynchronize(st));
            cw = now_ms();
            prepare_all();
            cp = now_ms();
            for (int i = 0; i < n; ++i) {
                int rc = slots[i]->setup_finish(st, stage);
                if (!rc && prec) rc = prec_setup(*slots[i], st);
                if (rc) return set_error(rc, "subgame %d: %s", i, last_error_str());
                if (target || want_final) {
                    rc = slots[i]->alloc_expl(st, 1, stage);
                    if (rc) return rc;
                }
            }
            cf = now_ms();
            SS_HIP(hipStreamSynchronize(st));   // before the stage is released
        }
        if (target || want_final) {
            SS_HIP(hipMalloc(&d_xout, sizeof(double) * 32 * (size_t)n));
            SS_HIP(hipMemsetAsync(d_xout, 0, sizeof(double) * 32 * (size_t)n, st));
            for (int i = 0; i < n; ++i) slots[i]->desc.xout = d_xout + 32 * (size_t)i;
            h_xout.assign(32 * (size_t)n, 0.0);
            joint_done.assign(n, 0);
        }
        const double c1 = now_ms();
        templ.resize(n);
        for (int i = 0; i < n; ++i) {
            templ[i] = slots[i]->desc;
            templ[i].iter = 0;
            templ[i].iter_end = T;
            templ[i].final_iter = T;
            // root snapshots on every iteration a check can fall on (sparse checks stay on the chunk grid)
            templ[i].snap_every = target ? ((cfg.flags & SS_FLAG_ADAPTIVE_CHECKS) ? gsteps : chunk) : 0;
            templ[i].active = 1;
        }
        SS_HIP(hipMalloc(&d_slots, sizeof(SlotDesc) * std::max(1, n)));
        SS_HIP(hipMalloc(&d_want, sizeof(int) * std::max(1, n)));
        std::vector<Slot *> raw;
        for (auto &sl : slots) raw.push_back(sl.get());
        int rc = build_phased_tasks(raw, pt, st);
        if (rc) return rc;
        const double c2 = now_ms();
        xblocks = std::max(1, std::min(512, std::max(n, pt.nterm)));
        if (target || want_final)
            SS_HIP(hipMalloc(&scratch, sizeof(double) * (size_t)xblocks * SS_XSCRATCH_DOUBLES));
        rc = choose_mode();
        if (rc) return rc;
        const double c3 = now_ms();
        const char *ng = getenv("SS_NO_GRAPH");
        use_graph = !(ng && atoi(ng) != 0);
        SS_HIP(hipMemcpyAsync(d_slots, templ.data(), sizeof(SlotDesc) * n, hipMemcpyHostToDevice, st));
        SS_HIP(hipStreamSynchronize(st));
        full_tasks = pt.lvl_off.empty() ? 0 : pt.lvl_off[pt.L] + pt.nterm;
        rc = capture();
        if (log_level() > 0) {
            double plan = 0;
            for (auto &sl : slots) plan += sl->plan_ms;
            fprintf(stderr, "[street_solve] batch create: subgames %.1f ms (plans %.1f, uploads queued %.1f, rank "
                            "wait %.1f, tables %.1f on host threads + %.1f queued, final wait %.1f), task lists %.1f, mode "
                            "and shape %.1f, capture %.1f ms\n", c1 - c0, plan, cb - c0, cw - cb, cp - cw, cf - cp,
                    c1 - cf, c2 - c1, c3 - c2,
                    now_ms() - c3);
        }
        return rc;
    }

    // capture `gsteps` iterations of the current task lists into the graph
    int capture() {
        if (!use_graph) return SS_OK;
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
        SS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < gsteps; ++k) {
            int rc = launch_iteration(mode);
            if (rc) {
                (void)hipStreamEndCapture(st, &graph);
                return rc;
            }
        }
        SS_HIP(hipStreamEndCapture(st, &graph));
        SS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return SS_OK;
    }

    // Phased launches are sized by their task lists: when retired subgames
    // hold a third or more of the tasks, rebuild the lists over the live ones
    // and recapture, so the tail of a batch does not dispatch empty workgroups.
    double host_compact_ms = 0, host_capture_ms = 0, host_eval_ms = 0;
    int compact(const std::vector<int> &live) {
        if (mode != 1) return SS_OK;
        int64_t live_tasks = 0;
        for (int i = 0; i < n; ++i)
            if (live[i]) live_tasks += slots[i]->plan.D + slots[i]->plan.T;
        int cur = pt.lvl_off.empty() ? 0 : pt.lvl_off[pt.L] + pt.nterm;
        if (live_tasks == 0 || live_tasks * 3 > (int64_t)cur * 2) return SS_OK;
        std::vector<Slot *> raw;
        for (auto &sl : slots) raw.push_back(sl.get());
        const double t0 = now_ms();
        int rc = build_phased_tasks(raw, pt, st, &live);
        if (rc) return rc;
        const double t1 = now_ms();
        rc = capture();
        host_compact_ms += t1 - t0;
        host_capture_ms += now_ms() - t1;
        return rc;
    }

    int launch_iteration(int m) {
        return m == 0 ? launch_batched(d_slots, n, 1, n, st) : launch_phased_iteration(d_slots, pt, st);
    }

    int reset_state() {
        for (auto &sl : slots) {
            SS_HIP(hipMemsetAsync(sl->desc.R, 0, sizeof(float) * (size_t)sl->plan.A * sl->plan.Mp, st));
            SS_HIP(hipMemsetAsync(sl->desc.SS, 0, sizeof(float) * (size_t)sl->plan.A * sl->plan.Mp, st));
            SS_HIP(hipMemsetAsync(sl->desc.rsig, 0, sizeof(float) * 8 * (size_t)sl->plan.Mp, st));
        }
        SS_HIP(hipMemcpyAsync(d_slots, templ.data(), sizeof(SlotDesc) * n, hipMemcpyHostToDevice, st));
        return SS_OK;
    }

    // The make-up of the batch: subgames, players, and the decision and
    // terminal task counts in quarter-octave buckets (batches drawn from one
    // distribution share a key).
    Makeup makeup() const {
        Makeup m;
        m.n = n;
        m.ndec = pt.lvl_off.empty() ? 0 : pt.lvl_off[pt.L];
        m.nterm = pt.nterm;
        m.npmax = pt.NPmax;
        m.precision = cfg.precision;
        return m;
    }

    // Autotune the batched strategy on this batch: time two iterations of each
    // (after one untimed iteration) and keep the faster; the choice is kept per
    // make-up for later batches. SS_BATCH_MODE=persistent|phased overrides.
    int choose_mode() {
        const char *env = getenv("SS_BATCH_MODE");
        if (env && !strcmp(env, "persistent")) { mode = 0; return SS_OK; }
        if ((env && !strcmp(env, "phased")) || cfg.precision == SS_PREC_BF16X2) {   // the contraction is phased only
            mode = 1;
            int rc = tune_shape();
            return rc ? rc : reset_state();
        }
        if (cfg.max_iterations < 3) { mode = 1; return SS_OK; }
        static std::mutex mu;
        static MakeupCache<int> chosen;
        const Makeup key = makeup();
        {
            std::lock_guard<std::mutex> lk(mu);
            if (const int *hit = chosen.find(key)) {
                mode = *hit;
                int rc = mode == 1 ? tune_shape() : SS_OK;
                return rc ? rc : reset_state();
            }
        }
        float best = 1e30f;
        for (int m = 0; m < 2; ++m) {
            int rc = reset_state();
            if (rc) return rc;
            rc = launch_iteration(m);
            if (rc) return rc;
            SS_HIP(hipEventRecord(ev[0], st));
            for (int k = 0; k < 2; ++k) {
                rc = launch_iteration(m);
                if (rc) return rc;
            }
            SS_HIP(hipEventRecord(ev[1], st));
            SS_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            SS_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            if (log_level() > 0) fprintf(stderr, "[street_solve] batch mode %d: %.3f ms / 2 iterations\n", m, ms);
            if (ms < best) { best = ms; mode = m; }
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            chosen.put(key, mode);
        }
        if (mode == 1) {
            int rc = tune_shape();
            if (rc) return rc;
        }
        return reset_state();
    }

    // Launch-shape autotuner of the phased form: coordinate descent over the
    // forward-sweep grouping, the merged terminal launch, the two-opponent
    // showdown split and the sweep grid, each candidate timed on this batch
    // (one untimed and two timed iterations); a candidate replaces the best
    // shape when it is at least 2% faster. Every shape gives the same results
    // bit for bit, so the choice only moves time. The result is kept per
    // batch signature (subgames, decision and terminal tasks, players) for
    // later batches of the same make-up. SS_NO_AUTOTUNE=1 keeps the
    // automatic rules; shapes forced through ssdbg_set_phased_shape win.
    double tune_ms = 0;
    int tune_shape() {
        const char *no = getenv("SS_NO_AUTOTUNE");
        if ((no && atoi(no) != 0) || phased_shape_forced() || cfg.max_iterations < 3) return SS_OK;
        const double t0 = now_ms();
        std::vector<Slot *> raw;
        for (auto &sl : slots) raw.push_back(sl.get());
        const Makeup key = makeup();
        static std::mutex mu;
        static MakeupCache<PhasedTasks::Shape> cache;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (const PhasedTasks::Shape *hit = cache.find(key)) {
                pt.shape = *hit;
                int rc = build_phased_tasks(raw, pt, st);
                tune_ms = now_ms() - t0;
                return rc;
            }
        }
        auto time_shape = [&](const PhasedTasks::Shape &sh, float *ms) {
            pt.shape = sh;
            int rc = build_phased_tasks(raw, pt, st);
            if (!rc) rc = reset_state();
            if (!rc) rc = launch_iteration(1);
            if (rc) return rc;
            SS_HIP(hipEventRecord(ev[0], st));
            for (int k = 0; k < 2 && !rc; ++k) rc = launch_iteration(1);
            if (rc) return rc;
            SS_HIP(hipEventRecord(ev[1], st));
            SS_HIP(hipEventSynchronize(ev[1]));
            SS_HIP(hipEventElapsedTime(ms, ev[0], ev[1]));
            return SS_OK;
        };
        PhasedTasks::Shape best_shape;
        float best = 0;
        int rc = time_shape(best_shape, &best);
        if (rc) return rc;
        const int knobs[4][4] = {{0, 1 << 30, -2, -2},     // fwd_group_min: always three / always one
                                 {0, -2, -2, -2},          // term_small: separate launches
                                 {1, 2, 4, -2},            // sd_parts
                                 {1024, 2048, -2, -2}};    // sweep_grid
        for (int k = 0; k < 4; ++k)
            for (int v = 0; v < 4 && knobs[k][v] != -2; ++v) {
                PhasedTasks::Shape c = best_shape;
                int *f = k == 0 ? &c.fwd_group_min : k == 1 ? &c.term_small : k == 2 ? &c.sd_parts : &c.sweep_grid;
                if (*f == knobs[k][v]) continue;
                *f = knobs[k][v];
                float ms = 0;
                if ((rc = time_shape(c, &ms))) return rc;
                if (log_level() > 0)
                    fprintf(stderr, "[street_solve] shape fwd %d small %d sd %d grid %d: %.3f ms / 2 iterations\n",
                            c.fwd_group_min, c.term_small, c.sd_parts, c.sweep_grid, ms);
                if (ms < 0.98f * best) {
                    best = ms;
                    best_shape = c;
                }
            }
        pt.shape = best_shape;
        rc = build_phased_tasks(raw, pt, st);
        {
            std::lock_guard<std::mutex> lk(mu);
            cache.put(key, best_shape);
        }
        tune_ms = now_ms() - t0;
        return rc;
    }

    int evaluate(const std::vector<int> &sel, int at, double *check_ms) {
        const double h0 = now_ms();
        SS_HIP(hipEventRecord(ev[2], st));
        SS_HIP(hipMemcpyAsync(d_want, sel.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        bool joint = false;
        for (int i = 0; i < n; ++i) joint |= sel[i] && !joint_done[i];
        int rc = launch_phased_expl(d_slots, pt, d_want, scratch, xblocks, st, joint);
        if (joint)
            for (int i = 0; i < n; ++i) joint_done[i] |= (char)(sel[i] != 0);
        if (rc) return rc;
        SS_HIP(hipEventRecord(ev[3], st));
        SS_HIP(hipEventSynchronize(ev[3]));
        float ms = 0;
        SS_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
        *check_ms += ms;
        if (log_level() > 1) {
            int nw = 0;
            for (int i = 0; i < n; ++i) nw += sel[i] != 0;
            fprintf(stderr, "[street_solve] check at %d: %d subgames, %.3f ms\n", at, nw, ms);
        }
        SS_HIP(hipMemcpy(h_xout.data(), d_xout, sizeof(double) * h_xout.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            if (!sel[i]) continue;
            expl[i] = expl_from_xout(h_xout.data() + 32 * (size_t)i, slots[i]->plan.NP, slots[i]->plan.starting_pot,
                                     &gains[(size_t)i * SS_MAX_PLAYERS]);
            expl_at[i] = at;
            if (log_level() > 2) fprintf(stderr, "[street_solve] exploitability %d %d %.9g\n", i, at, expl[i]);
        }
        host_eval_ms += now_ms() - h0 - ms;
        return SS_OK;
    }

    int run(SSReport *reps, SSBatchStats *stats) {
        const int T = std::max(0, cfg.max_iterations);
        double t0 = now_ms();
        // fresh regrets, strategy sums, root snapshot and descriptors; full task lists
        int rc0 = reset_state();
        if (rc0) return rc0;
        if (mode == 1 && pt.lvl_off[pt.L] + pt.nterm != full_tasks) {
            std::vector<Slot *> raw;
            for (auto &sl : slots) raw.push_back(sl.get());
            rc0 = build_phased_tasks(raw, pt, st);
            if (!rc0) rc0 = capture();
            if (rc0) return rc0;
        }
        iters.assign(n, 0);
        conv.assign(n, 0);
        expl_at.assign(n, -1);
        expl.assign(n, -1.0);
        gains.assign((size_t)n * SS_MAX_PLAYERS, 0.0);
        std::vector<int> live(n, 1), want(n, 0);
        double cfr_ms = 0, check_ms = 0;
        int launches = 0, done = 0;
        int64_t bytes = 0;
        int64_t live_bytes = 0;
        for (int i = 0; i < n; ++i) live_bytes += slots[i]->plan.algo_bytes_iter;
        // iteration of each subgame's next check: every chunk iterations, or
        // adaptive (each on a multiple of gsteps, so a replay count is whole)
        bool adaptive = false;
        const int cstep = check_step(cfg.flags, chunk, gsteps, &adaptive);
        adaptive = adaptive && target;
        std::vector<int> due(n, chunk);
        while (done < T) {
            int next = done + chunk;
            if (target) {
                next = T;
                for (int i = 0; i < n; ++i)
                    if (live[i]) next = std::min(next, due[i]);
                next = std::max(next, done + gsteps);
            }
            const int span = next - done;
            const int reps = (span + gsteps - 1) / gsteps;
            SS_HIP(hipEventRecord(ev[0], st));
            if (use_graph) {
                for (int k = 0; k < reps; ++k) SS_HIP(hipGraphLaunch(exec, st));
            } else {
                for (int k = 0; k < reps * gsteps; ++k) {
                    int rc = launch_iteration(mode);
                    if (rc) return rc;
                }
            }
            SS_HIP(hipEventRecord(ev[1], st));
            SS_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            SS_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            cfr_ms += ms;
            if (log_level() > 1) {
                int nl = 0;
                for (int i = 0; i < n; ++i) nl += live[i];
                fprintf(stderr, "[street_solve] iterations %d..%d: %d live, %.3f ms\n", done, done + span, nl, ms);
            }
            int step = std::min(reps * gsteps, T - done);
            done += step;
            launches += reps * gsteps;
            bytes += live_bytes * step;
            for (int i = 0; i < n; ++i)
                if (live[i]) iters[i] = done;
            if (target) {
                for (int i = 0; i < n; ++i) want[i] = live[i] && (due[i] <= done || done >= T);
                int rc = evaluate(want, done, &check_ms);
                if (rc) return rc;
                int nlive = 0;
                static const int kZero = 0;
                for (int i = 0; i < n; ++i) {
                    if (!live[i]) continue;
                    if (want[i] && expl[i] >= 0 && expl[i] <= cfg.target_exploitability) {
                        conv[i] = 1;
                        live[i] = 0;
                        live_bytes -= slots[i]->plan.algo_bytes_iter;
                        SS_HIP(hipMemcpyAsync(&d_slots[i].active, &kZero, sizeof(int), hipMemcpyHostToDevice, st));
                    } else {
                        if (want[i])
                            due[i] = next_check_iter(done, expl[i], cfg.target_exploitability, chunk, cstep, adaptive);
                        ++nlive;
                    }
                }
                SS_HIP(hipStreamSynchronize(st));
                if (nlive == 0) break;
                int rc2 = compact(live);
                if (rc2) return rc2;
            }
        }
        if (want_final) {
            for (int i = 0; i < n; ++i) want[i] = expl_at[i] != iters[i];
            int rc = evaluate(want, done, &check_ms);
            if (rc) return rc;
        }
        double wall = now_ms() - t0;
        if (log_level() > 0)
            fprintf(stderr, "[street_solve] batch run %.1f ms: cfr %.1f, checks %.1f, host: task rebuilds %.1f, graph captures %.1f, check readback %.1f\n",
                    wall, cfr_ms, check_ms, host_compact_ms, host_capture_ms, host_eval_ms);
        int maxit = 0;
        for (int i = 0; i < n; ++i) maxit = std::max(maxit, iters[i]);
        if (reps)
            for (int i = 0; i < n; ++i) {
                SSReport &r = reps[i];
                std::memset(&r, 0, sizeof(r));
                r.iterations = iters[i];
                r.converged = conv[i];
                r.exploitability = expl[i];
                for (int p = 0; p < SS_MAX_PLAYERS; ++p) r.nashconv[p] = gains[(size_t)i * SS_MAX_PLAYERS + p];
                r.elapsed_ms = wall;
                r.kernel_ms = cfr_ms + check_ms;
                r.algo_bytes = slots[i]->plan.algo_bytes_iter * iters[i];
            }
        if (stats) {
            stats->run_ms = wall;
            stats->cfr_ms = cfr_ms;
            stats->check_ms = check_ms;
            stats->algo_bytes = bytes;
            stats->launches = launches;
            stats->max_iterations = maxit;
            stats->mode = mode;
            stats->shape_fwd_group_min = pt.shape.fwd_group_min;
            stats->shape_term_small = pt.shape.term_small;
            stats->shape_sd_parts = pt.shape.sd_parts;
            stats->shape_sweep_grid = pt.shape.sweep_grid;
            stats->tune_ms = tune_ms;
        }
        return SS_OK;
    }
};

int solve_batch(const HostProblem *hps, int n, const SSSolveConfig &cfg, SSOutput *outs, SSReport *reps) {
    BatchImpl b;
    int rc = b.create(hps, n, cfg);
    if (rc) return rc;
    rc = b.run(reps, nullptr);
    if (rc) return rc;
    if (outs)
        for (int i = 0; i < n; ++i) {
            rc = read_output(*b.slots[i], b.slots[i]->desc, b.st, &outs[i], 0);
            if (rc) return rc;
        }
    return SS_OK;
}

}  // namespace ss

using namespace ss;

static int legacy_check(const SSTreeData *td) {
    if (!td) return set_error(SS_ERR_ARG, "tree_data is NULL");
    if (td->num_players < 2 || td->num_players > SS_MAX_PLAYERS)
        return set_error(SS_ERR_ARG, "num_players %d", td->num_players);
    for (int p = 0; p < td->num_players; ++p)
        if (td->num_hands[p] < 0 || td->num_hands[p] > SS_MAX_HANDS)
            return set_error(SS_ERR_ARG, "player %d has %d hands (legacy ABI max %d)", p, td->num_hands[p],
                             SS_MAX_HANDS);
    return SS_OK;
}

extern "C" {

int ss_build_tree(const int *board, int num_board, int starting_pot, int effective_stack, const float *bet_sizes,
                  int num_bet_sizes, int num_players, const int *acting_order, int is_river, int use_cont_strats,
                  SSTreeData *tree_out) {
    clear_error();
    return build_tree(board, num_board, starting_pot, effective_stack, bet_sizes, num_bet_sizes, num_players,
                      acting_order, is_river, use_cont_strats, SS_MAX_RAISES, tree_out);
}

int ss_build_tree_ex(const int *board, int num_board, int starting_pot, int effective_stack, const float *bet_sizes,
                     int num_bet_sizes, int num_players, const int *acting_order, int is_river, int use_cont_strats,
                     int max_raises, SSTreeData *tree_out) {
    clear_error();
    return build_tree(board, num_board, starting_pot, effective_stack, bet_sizes, num_bet_sizes, num_players,
                      acting_order, is_river, use_cont_strats, max_raises, tree_out);
}

int ss_solve_gpu(SSTreeData *tree_data, int max_iterations, SSOutput *output) {
    clear_error();
    int rc = legacy_check(tree_data);
    if (rc) return rc;
    if (!output) return set_error(SS_ERR_ARG, "output is NULL");
    HostProblem hp = from_legacy(tree_data);
    SSSolveConfig cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.max_iterations = max_iterations;
    cfg.flags = SS_FLAG_NO_EXPL;
    return solve_plain(hp, cfg, output, nullptr);
}

void ss_free_tree(SSTreeData *tree_data) { free_tree(tree_data); }

void ss_free_output(SSOutput *out) {
    if (!out) return;
    std::free(out->root_strategy);
    std::free(out->root_ev);
    std::free(out->avg_strategies);
    std::free(out->avg_strategy_node_ids);
    std::memset(out, 0, sizeof(*out));
}

int ss_solve_gpu_ex(const SSProblem *problem, const SSSolveConfig *config, SSOutput *output, SSReport *report) {
    clear_error();
    if (!config) return set_error(SS_ERR_ARG, "config is NULL");
    HostProblem hp;
    int rc = from_ex(problem, &hp);
    if (rc) return rc;
    return solve_plain(hp, *config, output, report);
}

int ss_solve_batch_ex(const SSProblem *problems, int n, const SSSolveConfig *config, SSOutput *outputs,
                      SSReport *reports) {
    clear_error();
    if (!config || n < 0 || (n > 0 && !problems)) return set_error(SS_ERR_ARG, "bad batch arguments");
    if (n == 0) return SS_OK;
    std::vector<HostProblem> hps(n);
    for (int i = 0; i < n; ++i) {
        int rc = from_ex(&problems[i], &hps[i]);
        if (rc) return rc;
    }
    return solve_batch(hps.data(), n, *config, outputs, reports);
}

}  // extern "C"

struct SSBatch {
    ss::BatchImpl impl;
};

extern "C" {

SSBatch *ss_batch_create(const SSProblem *problems, int n, const SSSolveConfig *config) {
    clear_error();
    if (!config || n <= 0 || !problems) {
        set_error(SS_ERR_ARG, "bad batch arguments");
        return nullptr;
    }
    std::vector<HostProblem> hps(n);
    for (int i = 0; i < n; ++i)
        if (from_ex(&problems[i], &hps[i])) return nullptr;
    SSBatch *b = new SSBatch();
    if (b->impl.create(hps.data(), n, *config) != SS_OK) {
        delete b;
        return nullptr;
    }
    return b;
}

int ss_batch_run(SSBatch *batch, SSReport *reports, SSBatchStats *stats) {
    clear_error();
    if (!batch) return set_error(SS_ERR_ARG, "NULL batch");
    return batch->impl.run(reports, stats);
}

int ss_batch_output(SSBatch *batch, int index, SSOutput *output) {
    clear_error();
    if (!batch || !output || index < 0 || index >= batch->impl.n) return set_error(SS_ERR_ARG, "bad arguments");
    // The batch has finished its work (run returns after it), so the copy
    // may use a stream of the calling thread: callers reading many outputs
    // from several threads keep several copies in flight. Each thread's
    // stream is made once and kept for the life of the process.
    static thread_local hipStream_t ts = nullptr;
    if (!ts && hipStreamCreateWithFlags(&ts, hipStreamNonBlocking) != hipSuccess) ts = nullptr;
    return read_output(*batch->impl.slots[index], batch->impl.slots[index]->desc, ts ? ts : batch->impl.st, output,
                       0);
}

void ss_batch_destroy(SSBatch *batch) { delete batch; }

int ss_batch_leaf_values(SSBatch *batch, int index, const double *raw_equity, const int *leaf_idx,
                         const int *leaf_pot, int num_leaf_infos, int tree_order) {
    clear_error();
    if (!batch || index < 0 || index >= batch->impl.n) return set_error(SS_ERR_ARG, "bad batch or subgame index");
    if (!raw_equity || num_leaf_infos < 0 || (num_leaf_infos > 0 && (!leaf_idx || !leaf_pot)))
        return set_error(SS_ERR_ARG, "bad leaf arguments");
    Slot &s = *batch->impl.slots[index];
    const Plan &pl = s.plan;
    if (!pl.has_leaf || pl.num_leaves <= 0 || !s.desc.leaf_values)
        return set_error(SS_ERR_ARG, "subgame %d has no leaf buffer (create the batch with leaf values)", index);
    const int NP = pl.NP, L = pl.num_leaves;
    const long long rows = (long long)num_leaf_infos << (2 * NP);
    std::vector<int> row_of;
    if (!tree_order) {
        // the reference passes the producer's array as is: its rows must be
        // the tree's leaves (python/street_solver_gpu.py:299-307)
        if (rows != L)
            return set_error(SS_ERR_ARG, "%lld profile rows for %d tree leaves (reference order needs equal counts)",
                             rows, L);
    } else {
        // tree leaf l -> row L0 * 4^NP + sum_p s_p 4^p: climb the continuation
        // chain (one level per active player, the first active player's
        // choice at the top, street_solve.cu:400-455) to its head, the
        // original leaf, numbered in node order; players not in the chain
        // keep profile 0
        const int N = pl.N;
        std::vector<int> parent(N, -1), slot_of(N, 0);
        for (int n = 0; n < N; ++n) {
            const DevNode &d = pl.nodes[n];
            if (d.type != SS_NODE_DECISION) continue;
            for (int a = 0; a < d.nchild; ++a) {
                parent[pl.children[d.first_child + a]] = n;
                slot_of[pl.children[d.first_child + a]] = a;
            }
        }
        std::vector<int> head_of(L, -1), combo_of(L, 0);
        for (int n = 0; n < N; ++n) {
            const DevNode &d = pl.nodes[n];
            if (d.type != SS_NODE_LEAF || d.leaf_idx < 0 || d.leaf_idx >= L) continue;
            int cur = n, combo = 0, ok = 1;
            for (int q = NP - 1; q >= 0; --q) {
                if (!((d.active_mask >> q) & 1)) continue;
                const int par = parent[cur];
                if (par < 0 || pl.nodes[par].player != q || pl.nodes[par].nchild != SS_NUM_CONT_STRATS) {
                    ok = 0;
                    break;
                }
                combo |= slot_of[cur] << (2 * q);
                cur = par;
            }
            if (!ok) return set_error(SS_ERR_ARG, "leaf node %d is not at the end of a continuation chain", n);
            head_of[d.leaf_idx] = cur;
            combo_of[d.leaf_idx] = combo;
        }
        std::vector<int> heads(head_of.begin(), head_of.end());
        std::sort(heads.begin(), heads.end());
        heads.erase(std::unique(heads.begin(), heads.end()), heads.end());
        row_of.resize(L);
        for (int l = 0; l < L; ++l) {
            if (head_of[l] < 0) return set_error(SS_ERR_ARG, "tree leaf %d has no node", l);
            const int L0 = (int)(std::lower_bound(heads.begin(), heads.end(), head_of[l]) - heads.begin());
            row_of[l] = (L0 << (2 * NP)) | combo_of[l];
        }
    }
    return leaf_values_into(const_cast<float *>(s.desc.leaf_values), L, NP, pl.maxh, raw_equity, pl.nh,
                            leaf_idx, leaf_pot, num_leaf_infos, tree_order ? row_of.data() : nullptr,
                            batch->impl.st);
}

int ss_batch_range_update(SSBatch *batch, int index, int node, int action, int num_hands, const int *solver_hand,
                          double *weights, double weight_floor, double p_threshold) {
    clear_error();
    if (!batch || index < 0 || index >= batch->impl.n) return set_error(SS_ERR_ARG, "bad batch or subgame index");
    if (batch->impl.iters.empty()) return set_error(SS_ERR_ARG, "the batch has not been run");
    if (num_hands < 0 || num_hands > SS_RANGE_MAX_HANDS)
        return set_error(SS_ERR_ARG, "num_hands %d (0..%d)", num_hands, SS_RANGE_MAX_HANDS);
    if (num_hands > 0 && (!solver_hand || !weights)) return set_error(SS_ERR_ARG, "NULL solver_hand or weights");
    const Slot &s = *batch->impl.slots[index];
    const Plan &pl = s.plan;
    if (node < 0 || node >= pl.N || pl.nodes[node].type != SS_NODE_DECISION)
        return set_error(SS_ERR_ARG, "node %d is not a decision node", node);
    const DevNode &nd = pl.nodes[node];
    const int p = nd.player, nsolver = pl.nh[p];
    if (action < -1 || action >= nd.nchild) return set_error(SS_ERR_ARG, "action %d (node has %d)", action, nd.nchild);
    for (int h = 0; h < num_hands; ++h)
        if (solver_hand[h] < -1 || solver_hand[h] >= nsolver)
            return set_error(SS_ERR_ARG, "solver_hand[%d] = %d (player %d has %d hands)", h, solver_hand[h], p,
                             nsolver);
    // the probability dict of the host path holds one entry per combo: the
    // solve's hand list must not repeat one
    std::vector<char> seen(SS_NCOMBO, 0);
    for (int h = 0; h < nsolver; ++h) {
        char &f = seen[(pl.hinfo[(size_t)p * pl.Mp + h] >> 12) & 2047];
        if (f) return set_error(SS_ERR_ARG, "player %d's hands repeat a combo", p);
        f = 1;
    }
    if (num_hands == 0) return SS_OK;
    return range_update_from_strategy(s.desc.SS, nd.act_base, nd.nchild, action, pl.Mp, nsolver, solver_hand,
                                      weights, num_hands, weight_floor, p_threshold, batch->impl.st);
}

int64_t ss_problem_footprint(const SSProblem *problem) {
    clear_error();
    HostProblem hp;
    int rc = from_ex(problem, &hp);
    if (rc) return rc;
    return estimate_footprint(hp);
}

const char *ss_last_error(void) { return last_error_str(); }

const char *ss_version(void) { return "street_solve-mi355x 0.1 (gfx950)"; }

// ---- test entry points (include/street_solve_debug.h) ----------------------

int ssdbg_eval7(const int *cards, int n, uint32_t *out, int on_device) {
    clear_error();
    if (n < 0 || (n > 0 && (!cards || !out))) return set_error(SS_ERR_ARG, "bad arguments");
    if (on_device) return dbg_eval7_device(cards, n, out);
    static uint8_t tab[8192];
    static bool ready = false;
    if (!ready) {
        fill_straight_table(tab, 0, 1);
        ready = true;
    }
    for (int i = 0; i < n; ++i) out[i] = eval7_lut(cards + 7 * i, tab);
    return SS_OK;
}

int ssdbg_combo_ranks(const int *board5, uint32_t *out1326) {
    clear_error();
    if (!board5 || !out1326) return set_error(SS_ERR_ARG, "NULL argument");
    return dbg_combo_ranks_device(board5, out1326);
}

int ssdbg_eval_census(unsigned long long counts[10], unsigned long long *rank_sum) {
    clear_error();
    return dbg_eval_census_device(counts, rank_sum);
}

int ssdbg_solve_state(const SSProblem *problem, int iterations, int form, float *regrets, float *strat_sum,
                      float *root_ev) {
    clear_error();
    HostProblem hp;
    int rc = from_ex(problem, &hp);
    if (rc) return rc;
    StreamGuard g;
    if (g.rc) return g.rc;
    Slot s;
    rc = s.setup(hp, g.st);
    if (rc) return rc;
    if (form == 2) {   // plain form with the BF16 precision showdown
        rc = prec_setup(s, g.st);
        if (rc) return rc;
    }
    if (form == 0 || form == 2) {
        double ms = 0;
        rc = plain_run(s, iterations, iterations, g.st, &ms);
        if (rc) return rc;
    } else {
        SlotDesc d = s.desc;
        d.iter = 0;
        d.iter_end = iterations;
        d.final_iter = iterations;
        d.snap_every = 0;
        d.active = 1;
        SS_HIP(hipMemcpyAsync(s.d_desc, &d, sizeof(d), hipMemcpyHostToDevice, g.st));
        rc = launch_batched(s.d_desc, 1, iterations, 1, g.st);
        if (rc) return rc;
    }
    const Plan &pl = s.plan;
    const size_t Mp = pl.Mp;
    std::vector<float> R((size_t)pl.A * Mp), SSv((size_t)pl.A * Mp), cfv0(Mp);
    SS_HIP(hipMemcpyAsync(R.data(), s.desc.R, sizeof(float) * R.size(), hipMemcpyDeviceToHost, g.st));
    SS_HIP(hipMemcpyAsync(SSv.data(), s.desc.SS, sizeof(float) * SSv.size(), hipMemcpyDeviceToHost, g.st));
    SS_HIP(hipMemcpyAsync(cfv0.data(), s.desc.cfv, sizeof(float) * Mp, hipMemcpyDeviceToHost, g.st));
    SS_HIP(hipStreamSynchronize(g.st));
    // reference layout [N][8][maxh]
    for (int n = 0; n < pl.N; ++n)
        for (int a = 0; a < SS_MAX_ACTIONS; ++a)
            for (int h = 0; h < pl.maxh; ++h) {
                size_t o = ((size_t)n * SS_MAX_ACTIONS + a) * pl.maxh + h;
                const DevNode &nd = pl.nodes[n];
                bool has = nd.type == SS_NODE_DECISION && a < nd.nchild;
                if (regrets) regrets[o] = has ? R[(size_t)(nd.act_base + a) * Mp + h] : 0.f;
                if (strat_sum) strat_sum[o] = has ? SSv[(size_t)(nd.act_base + a) * Mp + h] : 0.f;
            }
    if (root_ev)
        for (int h = 0; h < pl.maxh; ++h) root_ev[h] = h < pl.nh[pl.NP - 1] ? cfv0[h] : 0.f;
    return SS_OK;
}

int ssdbg_plan_bytes(const SSProblem *problem, int64_t *algo_bytes_per_iter, int64_t *footprint) {
    clear_error();
    HostProblem hp;
    int rc = from_ex(problem, &hp);
    if (rc) return rc;
    Plan pl;
    rc = make_plan(hp, pl);
    if (rc) return rc;
    if (algo_bytes_per_iter) *algo_bytes_per_iter = pl.algo_bytes_iter;
    if (footprint) *footprint = estimate_footprint(hp);
    return SS_OK;
}

int ssdbg_set_phased_shape(int fwd_group_min, int term_small, int sd_parts) {
    clear_error();
    set_phased_shape(fwd_group_min, term_small, sd_parts);
    return SS_OK;
}

int ssdbg_copy(int64_t bytes, int repeats, float *ms) {
    clear_error();
    return dbg_copy(bytes, repeats, ms);
}

}  // extern "C"

// ss_server.hip -- served form: a continuous-batching request loop on one GPU.
//
// Requests are deep-copied at submit and queued. A worker thread owns the
// device: between steps it admits queued requests into free slots while the
// resident state fits the HBM budget, advances every resident subgame by
// `iters_per_step` iterations of the phased batched form (one launch per phase
// of a traverser pass, over the (subgame, node) tasks of every resident
// subgame), evaluates the FP64 exploitability of subgames that reached a check
// point, and retires those at their target (or at their iteration cap) into
// completed results. The task list is rebuilt when the resident set changes;
// a step over an unchanged set replays a captured graph. Every task is one
// workgroup on one subgame's own state and every reduction has a fixed order,
// so a subgame's result does not depend on what else is resident.
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ss_internal.h"

namespace ss {
const char *last_error_str();

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Owned copy of everything a problem points at.
struct OwnedProblem {
    SSTreeData tree;
    std::vector<SSNode> nodes;
    std::vector<int> children;
    std::vector<int> hands[SS_MAX_PLAYERS];
    std::vector<float> weights[SS_MAX_PLAYERS];
    std::vector<float> leaf;
    std::vector<int> frozen;
    HostProblem hp;

    int copy_from(const SSProblem *p) {
        HostProblem src;
        int rc = from_ex(p, &src);
        if (rc) return rc;
        const SSTreeData *t = src.tree;
        std::memset(&tree, 0, sizeof(tree));
        tree = *t;   // scalars; pointer fields replaced below
        nodes.assign(t->nodes, t->nodes + t->num_nodes);
        children.assign(t->children, t->children + t->num_children_total);
        tree.nodes = nodes.data();
        tree.children = children.data();
        tree.level_order = tree.node_depth = nullptr;
        tree.decision_node_indices = tree.showdown_node_indices = nullptr;
        tree.leaf_node_indices = tree.fold_node_indices = nullptr;
        tree.leaf_values = nullptr;
        tree.frozen_action = nullptr;
        hp = HostProblem();
        hp.tree = &tree;
        hp.np = src.np;
        int maxh = 0;
        for (int q = 0; q < src.np; ++q) {
            hands[q].assign(src.hands[q], src.hands[q] + 2 * src.nh[q]);
            weights[q].assign(src.weights[q], src.weights[q] + src.nh[q]);
            hp.nh[q] = src.nh[q];
            hp.hands[q] = hands[q].data();
            hp.weights[q] = weights[q].data();
            maxh = std::max(maxh, src.nh[q]);
        }
        if (src.leaf_values && t->num_leaves > 0) {
            leaf.assign(src.leaf_values, src.leaf_values + (size_t)t->num_leaves * src.np * maxh);
            hp.leaf_values = leaf.data();
        }
        if (src.frozen) {
            frozen.assign(src.frozen, src.frozen + t->num_nodes);
            hp.frozen = frozen.data();
        }
        return SS_OK;
    }
};

enum ReqState { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_ERROR = 3, RQ_CANCELLED = 4 };

struct Request {
    int64_t ticket = 0;
    OwnedProblem prob;
    SSSolveConfig cfg{};
    int64_t footprint = 0;
    int state = RQ_QUEUED;
    int err = 0;
    std::string errmsg;
    bool cancel = false;
    int slot = -1;
    std::unique_ptr<Slot> dev;
    int iters = 0;
    int next_check = 0;
    double expl = -1;
    double gains[SS_MAX_PLAYERS] = {0, 0, 0, 0, 0, 0};
    int expl_at = -1;
    bool joint_done = false;   // the joint deal mass (weights only) is in the slot's xout
    double t_submit = 0, t_admit = 0, t_done = 0, kernel_ms = 0;
    SSOutput out{};
    SSReport rep{};
};

}  // namespace
}  // namespace ss

struct SSServer {
    SSServerConfig cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_read;
    std::thread worker;
    bool stop = false;
    // Retired requests' outputs are read back by a second host thread on its
    // own stream, so the loop moves on as soon as a request retires: the
    // slot index is free at once, the request's device state (and its share
    // of the HBM budget) when its copy is done.
    std::thread reader;
    std::deque<std::shared_ptr<ss::Request>> read_jobs;
    int n_reading = 0;        // retired requests whose outputs are not read back yet
    bool reader_stop = false;
    hipStream_t rs = nullptr;
    ss::PinnedStage adm_stage;   // admission uploads (worker thread only)
    hipStream_t as = nullptr;     // admissions set up while a step runs
    // requests set up during a step, made resident after it (slot index)
    std::vector<std::pair<std::shared_ptr<ss::Request>, int>> staged;
    int64_t next_ticket = 1;
    std::map<int64_t, std::shared_ptr<ss::Request>> reqs;
    std::deque<int64_t> queue;
    // device side (worker thread only)
    hipStream_t st = nullptr;
    ss::SlotDesc *d_slots = nullptr;
    int *d_want = nullptr;
    double *scratch = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ss::PhasedTasks pt;
    bool dirty = true;        // resident set changed since the tasks were built
    bool have_graph = false;  // exec replays the current tasks
    bool no_graph = false;    // SS_NO_GRAPH set: always launch directly (profilers)
    std::vector<std::shared_ptr<ss::Request>> resident;   // by slot index
    int64_t used_bytes = 0;
    double t_admit = 0, t_step = 0, t_check = 0, t_expl = 0, t_read = 0;   // worker wall time per phase (logged at destroy)
    int n_steps = 0, n_admits = 0;
    int fatal = 0;
    std::string fatal_msg;

    int init();
    void run();
    int admit_locked(std::unique_lock<std::mutex> &lk);
    void setup_group(const std::vector<std::shared_ptr<ss::Request>> &grp, const std::vector<int> &slot_of,
                     std::vector<int> &rcs, std::vector<std::string> &msgs, hipStream_t sst, bool upload_desc);
    void take_admissible(std::vector<std::shared_ptr<ss::Request>> &grp, std::vector<int> &slot_of);
    void fail_request_locked(const std::shared_ptr<ss::Request> &r, int rc, const std::string &msg);
    void make_resident_locked(const std::shared_ptr<ss::Request> &r, int slot);
    void overlap_admit();
    int install_staged();
    int refresh_tasks();
    void drop_graph();
    int step();
    int check_and_retire();
    void release_slot(const std::shared_ptr<ss::Request> &r);
    void finish(const std::shared_ptr<ss::Request> &r, int state);
    void read_loop();
    ~SSServer();
};

using namespace ss;

int SSServer::init() {
    SS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    SS_HIP(hipStreamCreateWithFlags(&rs, hipStreamNonBlocking));
    SS_HIP(hipStreamCreateWithFlags(&as, hipStreamNonBlocking));
    const int n = cfg.max_slots;
    SS_HIP(hipMalloc(&d_slots, sizeof(SlotDesc) * n));
    SS_HIP(hipMemset(d_slots, 0, sizeof(SlotDesc) * n));   // active = 0 everywhere
    SS_HIP(hipMalloc(&d_want, sizeof(int) * n));
    SS_HIP(hipMalloc(&scratch, sizeof(double) * (size_t)n * SS_XSCRATCH_DOUBLES));
    resident.assign(n, nullptr);
    no_graph = std::getenv("SS_NO_GRAPH") != nullptr;
    return SS_OK;
}

void SSServer::drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    have_graph = false;
}

// rebuild the task list over the resident slots
int SSServer::refresh_tasks() {
    drop_graph();
    std::vector<Slot *> sl(cfg.max_slots, nullptr);
    for (int i = 0; i < cfg.max_slots; ++i)
        if (resident[i]) sl[i] = resident[i]->dev.get();
    int rc = build_phased_tasks(sl, pt, st);
    if (rc) return rc;
    dirty = false;
    return SS_OK;
}

SSServer::~SSServer() {
    if (log_level() > 0)
        fprintf(stderr, "[street_solve] server: %d admissions %.1f ms, %d steps %.1f ms, checks and retirement %.1f ms "
                        "(exploitability %.1f, output read-back %.1f)\n",
                n_admits, t_admit, n_steps, t_step, t_check, t_expl, t_read);
    drop_graph();
    for (auto &r : resident) r.reset();
    if (d_slots) (void)hipFree(d_slots);
    if (d_want) (void)hipFree(d_want);
    if (scratch) (void)hipFree(scratch);
    if (st) (void)hipStreamDestroy(st);
    if (rs) (void)hipStreamDestroy(rs);
    if (as) (void)hipStreamDestroy(as);
}

// the request's slot index inactive on the device and free for admission
void SSServer::release_slot(const std::shared_ptr<Request> &r) {
    if (r->slot < 0) return;
    int zero = 0;
    (void)hipMemcpyAsync(&d_slots[r->slot].active, &zero, sizeof(int), hipMemcpyHostToDevice, st);
    (void)hipStreamSynchronize(st);
    resident[r->slot].reset();
    r->slot = -1;
    dirty = true;
}

void SSServer::finish(const std::shared_ptr<Request> &r, int state) {
    r->state = state;
    r->t_done = now_ms();
    release_slot(r);
    if (r->dev) {
        used_bytes -= r->footprint;
        r->dev.reset();
    }
}

// Device state of a group of admitted requests, set up together: every
// request's plan, arena and uploads queued, one wait for the combo ranks,
// the river tables on host threads, the tables, FP64 buffers and slot
// descriptors queued, one wait. A request that fails gets its error in rcs
// / msgs without stopping the others.
void SSServer::setup_group(const std::vector<std::shared_ptr<Request>> &grp, const std::vector<int> &slot_of,
                           std::vector<int> &rcs, std::vector<std::string> &msgs, hipStream_t sst, bool upload_desc) {
    const size_t g = grp.size();
    auto fail = [&](size_t k, int rc) {
        rcs[k] = rc;
        msgs[k] = last_error_str();
    };
    PinnedStage &stage = adm_stage;   // kept between groups: pinning pages costs more than the copies
    stage.reset();
    for (size_t k = 0; k < g; ++k) {
        grp[k]->dev.reset(new Slot());
        int rc = grp[k]->dev->setup_begin(grp[k]->prob.hp, sst, stage);
        if (rc) fail(k, rc);
    }
    if (hipStreamSynchronize(sst) != hipSuccess) {
        const int rc = set_error(SS_ERR_HIP, "admission: stream synchronisation failed");
        for (size_t k = 0; k < g; ++k)
            if (!rcs[k]) fail(k, rc);
        return;
    }
    std::vector<Slot *> raw(g, nullptr);
    for (size_t k = 0; k < g; ++k)
        if (!rcs[k]) raw[k] = grp[k]->dev.get();
    prepare_tables_all(raw);
    for (size_t k = 0; k < g; ++k) {
        if (rcs[k]) continue;
        Request &r = *grp[k];
        int rc = r.dev->setup_finish(sst, stage);
        if (!rc && r.cfg.precision == SS_PREC_BF16X2) rc = prec_setup(*r.dev, sst);   // 2-player rivers only
        if (!rc) rc = r.dev->alloc_expl(sst, 1, stage);
        if (!rc && upload_desc) {
            SlotDesc d = r.dev->desc;
            d.iter = 0;
            d.iter_end = std::max(0, r.cfg.max_iterations);
            d.final_iter = d.iter_end;
            d.snap_every = cfg.iters_per_step;   // every step end: any check or retirement finds its root snapshot
            d.active = 1;
            rc = stage.upload(&d_slots[slot_of[k]], &d, sizeof(d), st);
        }
        if (rc) fail(k, rc);
    }
    if (hipStreamSynchronize(sst) != hipSuccess) {   // before the stage is reused
        const int rc = set_error(SS_ERR_HIP, "admission: slot upload failed");
        for (size_t k = 0; k < g; ++k)
            if (!rcs[k]) fail(k, rc);
    }
}

// Queued requests, in order, while a slot is free and the footprint fits:
// their slots and budget are reserved (caller holds the lock).
void SSServer::take_admissible(std::vector<std::shared_ptr<Request>> &grp, std::vector<int> &slot_of) {
    std::vector<char> taken(cfg.max_slots, 0);
    for (int i = 0; i < cfg.max_slots; ++i) taken[i] = resident[i] != nullptr;
    for (auto &sg : staged) taken[sg.second] = 1;
    while (!queue.empty()) {
        auto it = reqs.find(queue.front());
        if (it == reqs.end() || it->second->cancel) {
            if (it != reqs.end()) it->second->state = RQ_CANCELLED;
            queue.pop_front();
            cv_done.notify_all();
            continue;
        }
        std::shared_ptr<Request> r = it->second;
        int free_slot = -1;
        for (int i = 0; i < cfg.max_slots; ++i)
            if (!taken[i]) { free_slot = i; break; }
        if (free_slot < 0 || used_bytes + r->footprint > cfg.hbm_budget_bytes) break;
        queue.pop_front();
        taken[free_slot] = 1;
        used_bytes += r->footprint;   // held while the request is set up
        grp.push_back(r);
        slot_of.push_back(free_slot);
    }
}

void SSServer::fail_request_locked(const std::shared_ptr<Request> &r, int rc, const std::string &msg) {
    r->err = rc;
    r->errmsg = msg;
    r->dev.reset();
    used_bytes -= r->footprint;
    r->state = RQ_ERROR;
    r->t_done = now_ms();
    cv_done.notify_all();
}

void SSServer::make_resident_locked(const std::shared_ptr<Request> &r, int slot) {
    r->slot = slot;
    r->state = RQ_RUNNING;
    r->t_admit = now_ms();
    r->next_check = r->cfg.check_every;
    resident[slot] = r;
    dirty = true;
}

// Admit what fits now, each group set up on the loop's stream (the loop is
// idle or between steps).
int SSServer::admit_locked(std::unique_lock<std::mutex> &lk) {
    for (;;) {
        std::vector<std::shared_ptr<Request>> grp;
        std::vector<int> slot_of;
        take_admissible(grp, slot_of);
        if (grp.empty()) break;
        std::vector<int> rcs(grp.size(), SS_OK);
        std::vector<std::string> msgs(grp.size());
        lk.unlock();
        const double ta = now_ms();
        setup_group(grp, slot_of, rcs, msgs, st, true);
        t_admit += now_ms() - ta;
        n_admits += (int)grp.size();
        lk.lock();
        for (size_t k = 0; k < grp.size(); ++k) {
            if (rcs[k]) fail_request_locked(grp[k], rcs[k], msgs[k]);
            else make_resident_locked(grp[k], slot_of[k]);
        }
    }
    return SS_OK;
}

// While a step runs on the loop's stream: set up what fits on the admission
// stream (the subgames' own new memory only; their slot descriptors wait for
// install_staged, after the step, since the step's kernels read every slot's
// descriptor).
void SSServer::overlap_admit() {
    std::vector<std::shared_ptr<Request>> grp;
    std::vector<int> slot_of;
    {
        std::lock_guard<std::mutex> lk(mu);
        take_admissible(grp, slot_of);
    }
    if (grp.empty()) return;
    std::vector<int> rcs(grp.size(), SS_OK);
    std::vector<std::string> msgs(grp.size());
    const double ta = now_ms();
    setup_group(grp, slot_of, rcs, msgs, as, false);
    t_admit += now_ms() - ta;
    n_admits += (int)grp.size();
    std::lock_guard<std::mutex> lk(mu);
    for (size_t k = 0; k < grp.size(); ++k) {
        if (rcs[k]) fail_request_locked(grp[k], rcs[k], msgs[k]);
        else staged.emplace_back(grp[k], slot_of[k]);
    }
}

// the requests overlap_admit set up: their descriptors uploaded, resident
int SSServer::install_staged() {
    std::vector<std::pair<std::shared_ptr<Request>, int>> sg;
    {
        std::lock_guard<std::mutex> lk(mu);
        sg.swap(staged);
    }
    if (sg.empty()) return SS_OK;
    PinnedStage &stage = adm_stage;
    stage.reset();
    int rc = SS_OK;
    for (auto &e : sg) {
        SlotDesc d = e.first->dev->desc;
        d.iter = 0;
        d.iter_end = std::max(0, e.first->cfg.max_iterations);
        d.final_iter = d.iter_end;
        d.snap_every = cfg.iters_per_step;   // every step end: any check or retirement finds its root snapshot
        d.active = 1;
        if (!rc) rc = stage.upload(&d_slots[e.second], &d, sizeof(d), st);
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = set_error(SS_ERR_HIP, "slot upload failed");
    const std::string msg = rc ? std::string(last_error_str()) : std::string();
    std::lock_guard<std::mutex> lk(mu);
    for (auto &e : sg) {
        if (rc) fail_request_locked(e.first, rc, msg);
        else make_resident_locked(e.first, e.second);
    }
    return SS_OK;
}

// Advance every resident subgame by iters_per_step iterations: direct launches
// right after the resident set changed, a captured graph of the same launches
// once it has stayed the same for a step.
int SSServer::step() {
    bool capture = false;
    if (dirty) {
        int rc = refresh_tasks();
        if (rc) return rc;
    } else if (!have_graph) {
        capture = !no_graph;
    }
    hipEvent_t e0, e1;
    SS_HIP(hipEventCreate(&e0));
    SS_HIP(hipEventCreate(&e1));
    if (capture) {
        SS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = SS_OK;
        for (int k = 0; k < cfg.iters_per_step && !rc; ++k) rc = launch_phased_iteration(d_slots, pt, st);
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc) return rc;
        SS_HIP(e);
        SS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        have_graph = true;
    }
    SS_HIP(hipEventRecord(e0, st));
    if (have_graph) {
        SS_HIP(hipGraphLaunch(exec, st));
    } else {
        for (int k = 0; k < cfg.iters_per_step; ++k) {
            int rc = launch_phased_iteration(d_slots, pt, st);
            if (rc) return rc;
        }
    }
    SS_HIP(hipEventRecord(e1, st));
    overlap_admit();
    SS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SS_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (auto &r : resident)
        if (r) {
            r->iters = std::min(r->iters + cfg.iters_per_step, std::max(0, r->cfg.max_iterations));
            r->kernel_ms += ms;
        }
    return SS_OK;
}

int SSServer::check_and_retire() {
    const int n = cfg.max_slots;
    std::vector<int> want(n, 0);
    int any = 0;
    for (int i = 0; i < n; ++i) {
        auto &r = resident[i];
        if (!r) continue;
        bool at_cap = r->iters >= r->cfg.max_iterations;
        bool at_check = r->cfg.target_exploitability > 0 && r->iters >= r->next_check;
        bool final_eval = at_cap && !(r->cfg.flags & SS_FLAG_NO_EXPL);
        if ((at_check || final_eval) && r->expl_at != r->iters) { want[i] = 1; any = 1; }
    }
    const double tx = now_ms();
    if (any) {
        SS_HIP(hipMemcpyAsync(d_want, want.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        bool joint = false;
        for (int i = 0; i < n; ++i) joint |= want[i] && !resident[i]->joint_done;
        int rc = launch_phased_expl(d_slots, pt, d_want, scratch, n, st, joint);
        if (rc) return rc;
        SS_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < n; ++i) {
            if (!want[i]) continue;
            auto &r = resident[i];
            double xo[2 * SS_MAX_PLAYERS + 1];
            SS_HIP(hipMemcpy(xo, r->dev->desc.xout, sizeof(double) * (2 * r->dev->plan.NP + 1), hipMemcpyDeviceToHost));
            r->expl = expl_from_xout(xo, r->dev->plan.NP, r->dev->plan.starting_pot, r->gains);
            r->joint_done = true;
            r->expl_at = r->iters;
            bool adaptive = false;
            const int cstep = check_step(r->cfg.flags, std::max(1, r->cfg.check_every), cfg.iters_per_step, &adaptive);
            if (adaptive)
                r->next_check = next_check_iter(r->iters, r->expl, r->cfg.target_exploitability,
                                                std::max(1, r->cfg.check_every), cstep, true);
            else
                while (r->next_check <= r->iters && r->cfg.check_every > 0) r->next_check += r->cfg.check_every;
        }
    }
    t_expl += now_ms() - tx;
    for (int i = 0; i < n; ++i) {
        std::shared_ptr<Request> r = resident[i];
        if (!r) continue;
        bool conv = r->cfg.target_exploitability > 0 && r->expl_at == r->iters && r->expl >= 0 &&
                    r->expl <= r->cfg.target_exploitability;
        bool cap = r->iters >= r->cfg.max_iterations;
        bool cancel;
        {
            std::lock_guard<std::mutex> lk(mu);
            cancel = r->cancel;
        }
        if (!(conv || cap || cancel)) continue;
        const double tr = now_ms();
        std::lock_guard<std::mutex> lk(mu);
        std::memset(&r->rep, 0, sizeof(r->rep));
        r->rep.iterations = r->iters;
        r->rep.converged = conv ? 1 : 0;
        r->rep.exploitability = r->expl_at == r->iters ? r->expl : -1;
        for (int p = 0; p < SS_MAX_PLAYERS; ++p) r->rep.nashconv[p] = r->gains[p];
        r->rep.kernel_ms = r->kernel_ms;
        r->rep.algo_bytes = r->dev->plan.algo_bytes_iter * r->iters;
        if (cancel) {
            finish(r, RQ_CANCELLED);
            r->rep.elapsed_ms = r->t_done - r->t_submit;
            cv_done.notify_all();
        } else {
            // no kernel touches the slot once it is inactive and the task
            // lists are rebuilt (before the next step): the reader copies
            // its outputs while the loop goes on
            release_slot(r);
            read_jobs.push_back(r);
            ++n_reading;
            cv_read.notify_one();
        }
        t_read += now_ms() - tr;
    }
    return SS_OK;
}

void SSServer::read_loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_read.wait(lk, [&] { return reader_stop || !read_jobs.empty(); });
        if (read_jobs.empty()) break;   // reader_stop and nothing left
        std::shared_ptr<Request> r = read_jobs.front();
        read_jobs.pop_front();
        const bool cancel = r->cancel;
        lk.unlock();
        int rc = cancel ? SS_OK : read_output(*r->dev, r->dev->desc, rs, &r->out, 0);
        std::string msg = rc ? std::string(last_error_str()) : std::string();
        r->dev.reset();   // the request's device state
        lk.lock();
        used_bytes -= r->footprint;
        if (rc) {
            r->err = rc;
            r->errmsg = msg;
        }
        r->state = cancel ? RQ_CANCELLED : (rc ? RQ_ERROR : RQ_DONE);
        r->t_done = now_ms();
        r->rep.elapsed_ms = r->t_done - r->t_submit;
        --n_reading;
        cv_done.notify_all();
        cv_work.notify_all();   // budget freed: the worker may admit more
    }
}

void SSServer::run() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        admit_locked(lk);
        bool busy = false;
        for (auto &r : resident)
            if (r) busy = true;
        if (!busy) {
            cv_done.notify_all();
            if (stop) break;
            // a queued request the budget holds back waits for a reader to
            // free its share; an empty queue waits for a submission
            const int reading = n_reading;
            cv_work.wait(lk, [&] { return stop || (!queue.empty() && (reading == 0 || n_reading < reading)); });
            continue;
        }
        lk.unlock();
        const double t0 = now_ms();
        int rc = step();
        const double t1 = now_ms();
        if (!rc) rc = check_and_retire();
        if (!rc) rc = install_staged();
        t_step += t1 - t0;
        t_check += now_ms() - t1;
        ++n_steps;
        lk.lock();
        if (rc) {
            fatal = rc;
            fatal_msg = last_error_str();
            for (auto &r : resident)
                if (r) {
                    r->err = rc;
                    r->errmsg = fatal_msg;
                    finish(r, RQ_ERROR);
                }
            for (auto &e : staged) fail_request_locked(e.first, rc, fatal_msg);
            staged.clear();
            cv_done.notify_all();
        }
    }
}

extern "C" {

SSServer *ss_server_create(const SSServerConfig *config) {
    clear_error();
    SSServerConfig c;
    std::memset(&c, 0, sizeof(c));
    if (config) c = *config;
    if (c.max_slots <= 0) c.max_slots = 256;
    if (c.iters_per_step <= 0) c.iters_per_step = 10;
    if (c.check_every <= 0) c.check_every = 50;
    if (c.hbm_budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            set_error(SS_ERR_HIP, "hipMemGetInfo failed");
            return nullptr;
        }
        c.hbm_budget_bytes = (int64_t)(fr * 0.8);
    }
    SSServer *s = new SSServer();
    s->cfg = c;
    if (s->init() != SS_OK) {
        delete s;
        return nullptr;
    }
    try {
        s->reader = std::thread([s] { s->read_loop(); });
        s->worker = std::thread([s] { s->run(); });
    } catch (...) {
        if (s->reader.joinable()) {
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->reader_stop = true;
                s->cv_read.notify_all();
            }
            s->reader.join();
        }
        delete s;
        set_error(SS_ERR_GENERIC, "could not start the server threads");
        return nullptr;
    }
    return s;
}

int64_t ss_server_submit(SSServer *s, const SSProblem *problem, const SSSolveConfig *config) {
    clear_error();
    if (!s || !config) return set_error(SS_ERR_ARG, "NULL server or config");
    auto r = std::make_shared<Request>();
    int rc = r->prob.copy_from(problem);
    if (rc) return rc;
    r->cfg = *config;
    if (r->cfg.precision != SS_PREC_FP32 && r->cfg.precision != SS_PREC_BF16X2)
        return set_error(SS_ERR_ARG, "unknown precision %d", r->cfg.precision);
    if (r->cfg.check_every <= 0) r->cfg.check_every = s->cfg.check_every;
    r->footprint = estimate_footprint(r->prob.hp);
    if (r->footprint < 0) return set_error(SS_ERR_ARG, "%s", last_error_str());
    // the precision form's sign masks, packed reach and card-removal masses:
    // under 4 MB even for 1326 combos and 128 showdown nodes
    if (r->cfg.precision == SS_PREC_BF16X2) r->footprint += (int64_t)8 << 20;
    if (r->footprint > s->cfg.hbm_budget_bytes)
        return set_error(SS_ERR_NOMEM, "request needs %lld bytes of device state, budget is %lld",
                         (long long)r->footprint, (long long)s->cfg.hbm_budget_bytes);
    r->t_submit = now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal) return set_error(s->fatal, "server failed: %s", s->fatal_msg.c_str());
    r->ticket = s->next_ticket++;
    s->reqs[r->ticket] = r;
    s->queue.push_back(r->ticket);
    s->cv_work.notify_all();
    return r->ticket;
}

int ss_server_poll(SSServer *s, int64_t ticket, SSOutput *output, SSReport *report) {
    clear_error();
    if (!s) return set_error(SS_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_error(SS_ERR_NOTFOUND, "unknown ticket %lld", (long long)ticket);
    auto r = it->second;
    if (r->state == RQ_QUEUED || r->state == RQ_RUNNING) return 0;
    int rc = 1;
    if (r->state == RQ_ERROR) rc = set_error(r->err ? r->err : SS_ERR_GENERIC, "%s", r->errmsg.c_str());
    else if (r->state == RQ_CANCELLED) rc = set_error(SS_ERR_NOTFOUND, "ticket %lld was cancelled", (long long)ticket);
    if (rc == 1) {
        if (output) { *output = r->out; std::memset(&r->out, 0, sizeof(r->out)); }
        else {
            std::free(r->out.root_strategy); std::free(r->out.root_ev);
            std::free(r->out.avg_strategies); std::free(r->out.avg_strategy_node_ids);
        }
        if (report) *report = r->rep;
    }
    s->reqs.erase(it);
    return rc;
}

int ss_server_cancel(SSServer *s, int64_t ticket) {
    clear_error();
    if (!s) return set_error(SS_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_error(SS_ERR_NOTFOUND, "unknown ticket %lld", (long long)ticket);
    it->second->cancel = true;
    if (it->second->state == RQ_QUEUED) it->second->state = RQ_CANCELLED;
    s->cv_work.notify_all();
    return SS_OK;
}

int ss_server_drain(SSServer *s) {
    clear_error();
    if (!s) return set_error(SS_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        if (s->fatal) return true;
        // queued (also while being set up) or running (also while its
        // outputs are read back)
        for (auto &kv : s->reqs)
            if (kv.second->state == RQ_QUEUED || kv.second->state == RQ_RUNNING) return false;
        return true;
    });
    return s->fatal ? set_error(s->fatal, "server failed: %s", s->fatal_msg.c_str()) : SS_OK;
}

void ss_server_destroy(SSServer *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto &kv : s->reqs) kv.second->cancel = true;
        s->queue.clear();
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->reader_stop = true;
        s->cv_read.notify_all();
    }
    if (s->reader.joinable()) s->reader.join();
    for (auto &kv : s->reqs) {
        SSOutput &o = kv.second->out;
        std::free(o.root_strategy); std::free(o.root_ev);
        std::free(o.avg_strategies); std::free(o.avg_strategy_node_ids);
    }
    delete s;
}

}  // extern "C"

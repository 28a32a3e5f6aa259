// serve.hip -- continuous-batching request loop over a resident graph.
//
// No reference counterpart (the reference has no serving mode); built on the
// multi-source BFS engines (msbfs.hip) and the update pipeline (update.hip).
//
//   client threads --submit--> arrival-ordered queue --dispatcher--+--> slot 0 (engine, stream, thread)
//                                                                  +--> slot 1 ...
//                                                                  +--> updates, applied in place
// * Queries (BFS, 2-hop) that arrive between two updates are batched, up to
//   max_batch (<= 64), into one multi-source launch; a free slot takes the
//   batch at once, so launches admitted while others run overlap on the
//   device (one stream per slot).
// * An update is applied once every launch before it has finished, and the
//   queries after it wait for it: each query sees exactly the updates that
//   arrived before it.  Its graph version is reported with the result.
// * Slots are admitted against a device-memory budget: one slot holds a
//   multi-source engine (masks, work lists, load-balancing buffers); batched
//   launches keep only per-source digests, never k x nv distances.  An update
//   is refused (HN_E_OOM result) when its scratch estimate does not fit what
//   the slots leave.
// * Results are digests (reached count, distance sum, eccentricity), so a
//   batch returns 24 B per query instead of its vectors: the batched engine
//   counts them per source on the device; a single-source traversal's come
//   from its level counters (frontier sizes and depths).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "bfs.hpp"
#include "graph.hpp"

// launches of at most this many BFS / 2-hop queries (and nothing else) run
// them as single-source TopDowns (overridable for sweeps)
#ifndef SERVE_SINGLE_MAX
#define SERVE_SINGLE_MAX 3
#endif

namespace hn {
namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct ServeReq {
    int64_t id = 0;
    int32_t kind = 0;
    int32_t source = 0;
    std::vector<int32_t> src, dst;
};

struct ServeSlot {
    MsEngine* eng = nullptr;
    int32_t* dist = nullptr;                 // nv: the single-source BFS's distances
    char* bfs_scratch = nullptr;             // the single-source BFS's level buffers
    int64_t bfs_scratch_bytes = 0;
    unsigned long long* h_digest = nullptr;  // pinned: batched launches' digests, or the BFS level readback
    hipStream_t stream = nullptr;
    std::thread th;
    std::vector<ServeReq*> batch;
    std::vector<uint64_t> qver;              // per-query graph versions of a window launch
    bool window = false;                     // the engine holds a window's side records
    bool has_work = false;
    bool busy = false;
    std::condition_variable cv;
};

}  // namespace hn

struct hn_server {
    hn_graph* g = nullptr;
    hn_serve_config cfg{};
    int64_t slot_bytes = 0;
    int64_t update_budget = 0;
    std::mutex mu;
    std::condition_variable cv_disp, cv_done;
    std::deque<hn::ServeReq*> queue;
    std::vector<std::unique_ptr<hn::ServeReq>> reqs;
    std::vector<hn_result> results;
    std::vector<std::unique_ptr<hn::ServeSlot>> slots;
    std::thread dispatcher;
    hipStream_t upd_stream = nullptr;
    // a window's update entries on the device: src[cap], dst[cap], then one
    // tag byte each (grown on demand, reused by every window)
    int32_t* win_ent = nullptr;
    int64_t win_cap = 0;
    int64_t window_max_entries = 0;   // update entries one window may span
    int inflight = 0;
    int64_t completed = 0;
    bool stop = false;
};

namespace hn {

static void finish_query(hn_server* s, ServeReq* r, const unsigned long long* dg, int32_t k, uint64_t ver,
                         hn_status st, int64_t t_done) {
    hn_result& res = s->results[(size_t)r->id];
    res.status = st;
    res.t_done_ns = t_done;
    res.reached = st == HN_OK ? (int64_t)dg[0] : 0;
    res.dist_sum = st == HN_OK ? (int64_t)dg[1] : 0;
    res.max_depth = st == HN_OK ? (int32_t)dg[2] : 0;
    res.batch = k;
    res.version = ver;
}

static void worker_loop(hn_server* s, ServeSlot* slot) {
    hn_graph* g = s->g;
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        slot->cv.wait(lk, [&] { return s->stop || slot->has_work; });
        if (!slot->has_work) break;
        std::vector<ServeReq*> batch;
        batch.swap(slot->batch);
        std::vector<uint64_t> qver;
        qver.swap(slot->qver);
        const bool window = slot->window;
        slot->window = false;
        slot->has_work = false;
        const uint64_t ver = g->version;   // stable: updates wait for in-flight launches
        lk.unlock();

        const int32_t k = (int32_t)batch.size();
        int32_t srcs[64], limits[64];
        for (int32_t i = 0; i < k; i++) {
            srcs[i] = batch[i]->source;
            limits[i] = batch[i]->kind == HN_REQ_TWO_HOP ? 2 : 0;
        }
        // Up to SERVE_SINGLE_MAX BFS or 2-hop queries run the single-source
        // TopDown one after the other over the slot's own buffers (a 2-hop
        // query stops after two levels; 6.2 ms per BFS on R-MAT s25 against
        // 12.6 ms for a batched launch of one or two sources), each finishing
        // when its own traversal does; a window's queries need the batched
        // engine's version masks.
        BfsBufs bb;
        bool single = !window && k <= SERVE_SINGLE_MAX &&
                      bfs_layout(g->nv, g->ne, nullptr, &bb) <= slot->bfs_scratch_bytes;
        for (int32_t i = 0; i < k; i++)
            single = single && (batch[i]->kind == HN_REQ_BFS || batch[i]->kind == HN_REQ_TWO_HOP);
        if (single) {
            bfs_layout(g->nv, g->ne, slot->bfs_scratch, &bb);
            for (int32_t i = 0; i < k; i++) {
                hn_status st = bfs_run(g, srcs[i], slot->dist, nullptr, slot->stream, bb, false,
                                       reinterpret_cast<uint64_t*>(slot->h_digest), limits[i]);
                // the digest from the level counters bfs_run read back into
                // h_digest: every frontier expanded adds its size and its
                // distance sum; a depth-limited query's last discoveries
                // (never expanded) lie at the limit
                unsigned long long dg[3] = {0, 0, 0};
                if (st == HN_OK) {
                    BfsCtl c;
                    std::memcpy(&c, slot->h_digest, sizeof(BfsCtl));
                    dg[0] = c.visited;
                    dg[1] = c.dsum;
                    dg[2] = c.levels > 0 ? c.levels - 1 : 0;
                    if (limits[i] > 0 && (int64_t)c.levels == (int64_t)limits[i]) {
                        const uint64_t fn = c.fc[(c.levels + 1) & 1] >> 32;
                        dg[0] += fn;
                        dg[1] += fn * c.levels;
                        if (fn) dg[2] = c.levels;
                    }
                }
                const int64_t t_done = now_ns();
                lk.lock();
                finish_query(s, batch[i], dg, 1, ver, st, t_done);
                s->completed += 1;
                s->cv_done.notify_all();
                lk.unlock();
            }
            lk.lock();
            s->inflight--;
            slot->busy = false;
            s->cv_disp.notify_one();
            s->cv_done.notify_all();
            continue;
        }
        // digest mode: the launch never materialises k x nv distances
        hn_status st = ms_run_engine(slot->eng, srcs, k, limits, nullptr, nullptr, slot->stream);
        {
            hipError_t e = st == HN_OK ? hipMemcpyAsync(slot->h_digest, ms_engine_digest(slot->eng),
                                                        sizeof(unsigned long long) * 3 * (size_t)k,
                                                        hipMemcpyDeviceToHost, slot->stream)
                                       : hipSuccess;
            if (e == hipSuccess && st == HN_OK) e = hipStreamSynchronize(slot->stream);
            if (e != hipSuccess) st = HN_E_HIP;
        }
        if (window) {
            const hn_status cs = ms_window_clear(slot->eng, slot->stream);
            if (cs == HN_OK) (void)hipStreamSynchronize(slot->stream);
        }
        const int64_t t_done = now_ns();
        lk.lock();
        for (int32_t i = 0; i < k; i++)
            finish_query(s, batch[i], slot->h_digest + 3 * i, k, qver.empty() ? ver : qver[(size_t)i], st, t_done);
        s->completed += k;
        s->inflight--;
        slot->busy = false;
        s->cv_disp.notify_one();
        s->cv_done.notify_all();
    }
}

static bool is_query(const ServeReq* r) { return r->kind == HN_REQ_BFS || r->kind == HN_REQ_TWO_HOP; }

static ServeSlot* free_slot(hn_server* s) {
    for (auto& sl : s->slots)
        if (!sl->busy) return sl.get();
    return nullptr;
}

// apply one update (barrier and window paths); the lock is not held
static hn_status apply_update(hn_server* s, const ServeReq* r, int64_t* applied) {
    *applied = 0;
    const int64_t n = (int64_t)r->src.size();
    if (n * 160 > s->update_budget) {
        hn::set_error("update batch exceeds the server's memory budget");
        return HN_E_OOM;
    }
    if (r->kind == HN_REQ_INSERT)
        return hn_graph_insert(s->g, r->src.data(), r->dst.data(), n, HN_DEDUP_BATCH | HN_DEDUP_GRAPH, applied,
                               s->upd_stream);
    return hn_graph_erase(s->g, r->src.data(), r->dst.data(), n, HN_DEDUP_BATCH, applied, s->upd_stream);
}

// apply one update of a window whose entries are already on the device
// (validated, and within the budget, when the window was packed)
static hn_status apply_update_dev(hn_server* s, const ServeReq* r, const int32_t* d_src, const int32_t* d_dst,
                                  int64_t* applied) {
    *applied = 0;
    const int64_t n = (int64_t)r->src.size();
    if (r->kind == HN_REQ_INSERT)
        return hn_graph_insert(s->g, d_src, d_dst, n, HN_DEDUP_BATCH | HN_DEDUP_GRAPH | HN_DEVICE_PTRS, applied,
                               s->upd_stream);
    return hn_graph_erase(s->g, d_src, d_dst, n, HN_DEDUP_BATCH | HN_DEVICE_PTRS, applied, s->upd_stream);
}

// room for n window entries in win_ent
static hn_status window_entries(hn_server* s, int64_t n) {
    if (n <= s->win_cap) return HN_OK;
    dfree(s->win_ent);
    s->win_ent = nullptr;
    s->win_cap = 0;
    const int64_t cap = std::max<int64_t>(n, std::min<int64_t>((int64_t)1 << 20, s->window_max_entries));
    HN_TRY(dalloc(&s->win_ent, (size_t)(2 * cap + (cap + 3) / 4)));
    s->win_cap = cap;
    return HN_OK;
}

// lock held
static void record_update(hn_server* s, const ServeReq* r, hn_status st, int64_t applied, uint64_t version,
                          int64_t t_start, int64_t t_done) {
    hn_result& res = s->results[(size_t)r->id];
    res.status = st;
    res.applied = applied;
    res.t_start_ns = t_start;
    res.t_done_ns = t_done;
    res.version = version;
    res.batch = 1;
    s->completed++;
}

// A multi-version window: a queue prefix of up to max_batch queries with the
// updates between them (at most 63, window_max_entries entries), ending at
// its last query.  One launch serves all its queries, each at the graph
// version it arrived at.
//
// Device memory of a window's entries: the server's copy (WINDOW_SERVER_BYTES
// each) and, in the engine of every slot that runs one, its sorted keys,
// segments, existence masks and changed pairs (at most WINDOW_ENGINE_BYTES
// each; an engine keeps them for the largest window it ran).  They are held to
// half the update budget, so updates keep the other half.
constexpr int64_t WINDOW_MAX_ENTRIES = 1 << 23;
constexpr int64_t WINDOW_SERVER_BYTES = 9;
constexpr int64_t WINDOW_ENGINE_BYTES = 96;
struct Window {
    int nq = 0, nu = 0;
    size_t len = 0;
};

static Window scan_window(hn_server* s, int maxb) {
    Window w;
    int nq = 0, nu = 0;
    int64_t ent = 0;
    for (size_t i = 0; i < s->queue.size(); i++) {
        const ServeReq* r = s->queue[i];
        if (is_query(r)) {
            if (nq == maxb) break;
            nq++;
            w.nq = nq;
            w.nu = nu;
            w.len = i + 1;
        } else {
            const int64_t m = (int64_t)r->src.size();
            if (nu == 63 || ent + m > s->window_max_entries) break;
            nu++;
            ent += m;
        }
    }
    return w;
}

// Runs window w (lock held on entry and exit; a slot is free).  Returns false,
// with the queue untouched, when its updates are not symmetric batches.
//
// The window's entries are packed, copied in, sorted and looked up in the
// current graph while launches already in flight still run: that set-up only
// reads the graph, as they do.  The updates apply once every other launch
// has finished.
static bool run_window(hn_server* s, const Window& w, std::unique_lock<std::mutex>& lk) {
    hn_graph* g = s->g;
    ServeSlot* slot = free_slot(s);
    std::vector<ServeReq*> qs, us;
    std::vector<int> qt;   // window version of each query: updates before it
    for (size_t i = 0; i < w.len; i++) {
        ServeReq* r = s->queue[i];
        if (is_query(r)) {
            qs.push_back(r);
            qt.push_back((int)us.size());
        } else {
            us.push_back(r);
        }
    }
    const int nu = (int)us.size();
    const int64_t t0 = now_ns();
    const uint64_t base_version = g->version;
    slot->busy = true;   // the slot's engine holds the window
    s->inflight++;
    lk.unlock();

    // entries of the updates that can apply (bad ids or an over-budget batch
    // fail as a whole and change nothing)
    uint64_t failed = 0;
    std::vector<int32_t> es, ed;
    std::vector<uint8_t> et;
    std::vector<int64_t> uoff((size_t)nu + 1, -1);   // update t's first entry in es (-1: it fails)
    for (int t = 1; t <= nu; t++) {
        const ServeReq* r = us[(size_t)t - 1];
        bool ok = (int64_t)r->src.size() * 160 <= s->update_budget;
        for (size_t j = 0; ok && j < r->src.size(); j++)
            ok = r->src[j] >= 0 && r->src[j] < g->nv && r->dst[j] >= 0 && r->dst[j] < g->nv;
        if (!ok) {
            failed |= 1ull << t;
            continue;
        }
        const uint8_t tag = (uint8_t)((t << 1) | (r->kind == HN_REQ_INSERT ? 1 : 0));
        uoff[(size_t)t] = (int64_t)es.size();
        es.insert(es.end(), r->src.begin(), r->src.end());
        ed.insert(ed.end(), r->dst.begin(), r->dst.end());
        et.insert(et.end(), r->src.size(), tag);
    }
    // the entries go to the device once: the window's sort reads them there,
    // and each update applies from its own range
    const int64_t ne = (int64_t)es.size();
    bool symmetric = false;
    hn_status st = window_entries(s, ne);
    int32_t* d_src = s->win_ent;
    int32_t* d_dst = s->win_ent + s->win_cap;
    uint8_t* d_tt = reinterpret_cast<uint8_t*>(s->win_ent + 2 * s->win_cap);
    if (st == HN_OK && ne > 0 &&
        (hipMemcpyAsync(d_src, es.data(), sizeof(int32_t) * (size_t)ne, hipMemcpyHostToDevice, s->upd_stream) !=
             hipSuccess ||
         hipMemcpyAsync(d_dst, ed.data(), sizeof(int32_t) * (size_t)ne, hipMemcpyHostToDevice, s->upd_stream) !=
             hipSuccess ||
         hipMemcpyAsync(d_tt, et.data(), (size_t)ne, hipMemcpyHostToDevice, s->upd_stream) != hipSuccess)) {
        (void)hipGetLastError();
        set_error("window entries: copy to the device");
        st = HN_E_HIP;
    }
    if (st == HN_OK) st = ms_window_begin(slot->eng, d_src, d_dst, d_tt, ne, s->upd_stream, &symmetric);
    if (st != HN_OK || !symmetric) {
        (void)ms_window_clear(slot->eng, s->upd_stream);
        lk.lock();
        slot->busy = false;
        s->inflight--;
        return false;
    }
    // the graph changes only once no other launch reads it
    lk.lock();
    s->cv_disp.wait(lk, [&] { return s->stop || s->inflight == 1; });
    if (s->stop) {
        (void)ms_window_clear(slot->eng, s->upd_stream);
        slot->busy = false;
        s->inflight--;
        return true;
    }
    lk.unlock();
    // the updates, in arrival order
    std::vector<hn_status> ust((size_t)nu);
    std::vector<int64_t> uapp((size_t)nu), udone((size_t)nu);
    std::vector<uint64_t> uver((size_t)nu);
    for (int t = 1; t <= nu; t++) {
        int64_t applied = 0;
        const int64_t o = uoff[(size_t)t];
        hn_status us_st = o >= 0 ? apply_update_dev(s, us[(size_t)t - 1], d_src + o, d_dst + o, &applied)
                                 : apply_update(s, us[(size_t)t - 1], &applied);
        if (us_st != HN_OK) failed |= 1ull << t;
        ust[(size_t)t - 1] = us_st;
        uapp[(size_t)t - 1] = applied;
        uver[(size_t)t - 1] = g->version;
        udone[(size_t)t - 1] = now_ns();
    }
    bool sym = false;
    st = graph_is_symmetric(g, s->upd_stream, &sym);
    uint64_t sv[64] = {0};
    for (size_t i = 0; i < qs.size(); i++) sv[qt[i]] |= 1ull << i;
    if (st == HN_OK) st = ms_window_end(slot->eng, nu, failed, sv, s->upd_stream);

    lk.lock();
    for (size_t i = 0; i < w.len; i++) s->queue.pop_front();
    for (int t = 0; t < nu; t++) record_update(s, us[(size_t)t], ust[(size_t)t], uapp[(size_t)t], uver[(size_t)t], t0,
                                               udone[(size_t)t]);
    // version a query saw: the base plus the updates before it that applied
    std::vector<uint64_t> vat((size_t)nu + 1, base_version);
    for (int t = 1; t <= nu; t++) vat[(size_t)t] = vat[(size_t)t - 1] + (((failed >> t) & 1ull) ? 0 : 1);
    for (size_t i = 0; i < qs.size(); i++) s->results[(size_t)qs[i]->id].t_start_ns = t0;
    s->cv_done.notify_all();
    if (st != HN_OK) {
        const int64_t t_done = now_ns();
        const unsigned long long zero[3] = {0, 0, 0};
        for (size_t i = 0; i < qs.size(); i++)
            finish_query(s, qs[i], zero, (int32_t)qs.size(), vat[(size_t)qt[i]], st, t_done);
        s->completed += (int64_t)qs.size();
        (void)ms_window_clear(slot->eng, s->upd_stream);
        slot->busy = false;
        s->inflight--;
        s->cv_done.notify_all();
        return true;
    }
    slot->batch = qs;
    slot->qver.clear();
    for (size_t i = 0; i < qs.size(); i++) slot->qver.push_back(vat[(size_t)qt[i]]);
    slot->window = true;
    slot->has_work = true;
    slot->cv.notify_one();
    return true;
}

static void dispatcher_loop(hn_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    const int maxb = s->cfg.plain ? 1 : s->cfg.max_batch;
    const int sbits = std::max(1, ceil_log2_host((int64_t)s->g->nv + 1));
    const bool windows = !s->cfg.plain && 2 * sbits + 7 <= 64;
    // a window pays for its set-up (sorting its update entries, per-pair
    // existence masks, restricted expansion) only when it batches enough
    // queries; with fewer waiting, the updates apply one by one
    const int min_wq = std::max(1, s->cfg.min_window_queries > 0 ? s->cfg.min_window_queries : 8);
    int barrier_updates = 0;   // updates to apply one by one (a window was not symmetric)
    while (true) {
        s->cv_disp.wait(lk, [&] { return s->stop || !s->queue.empty(); });
        if (s->stop) break;
        // queries across updates, in one launch (symmetric graphs)
        if (windows && barrier_updates == 0 && s->g->sym == 1) {
            Window w = scan_window(s, maxb);
            if (w.nq >= min_wq && w.nu > 0) {
                s->cv_disp.wait(lk, [&] { return s->stop || free_slot(s) != nullptr; });
                if (s->stop) break;
                w = scan_window(s, maxb);   // the queue may have grown
                if (run_window(s, w, lk)) continue;
                barrier_updates = w.nu;
                continue;
            }
        }
        ServeReq* head = s->queue.front();
        if (!is_query(head)) {
            // an update waits for every launch before it; the queries after it wait for it
            s->cv_disp.wait(lk, [&] { return s->stop || s->inflight == 0; });
            if (s->stop) break;
            s->queue.pop_front();
            const int64_t t_start = now_ns();
            lk.unlock();
            int64_t applied = 0;
            hn_status st = apply_update(s, head, &applied);
            // settle the adjacency-symmetry state here, while no query runs,
            // so concurrent launches only ever read it
            bool sym = false;
            if (st == HN_OK) st = graph_is_symmetric(s->g, s->upd_stream, &sym);
            const int64_t t_done = now_ns();
            lk.lock();
            // index, never hold a reference: submissions may grow `results`
            // while the lock is released
            record_update(s, head, st, applied, s->g->version, t_start, t_done);
            if (barrier_updates > 0) barrier_updates--;
            s->cv_done.notify_all();
            continue;
        }
        s->cv_disp.wait(lk, [&] { return s->stop || free_slot(s) != nullptr; });
        if (s->stop) break;
        ServeSlot* slot = free_slot(s);
        const int64_t t0 = now_ns();
        while (!s->queue.empty() && is_query(s->queue.front()) && (int)slot->batch.size() < maxb) {
            ServeReq* r = s->queue.front();
            s->queue.pop_front();
            s->results[(size_t)r->id].t_start_ns = t0;
            slot->batch.push_back(r);
        }
        slot->busy = true;
        slot->has_work = true;
        s->inflight++;
        slot->cv.notify_one();
        // in plain mode queries run one at a time, with nothing overlapping
        if (s->cfg.plain) s->cv_disp.wait(lk, [&] { return s->stop || s->inflight == 0; });
    }
}

// the single-source BFS scratch a slot reserves: sized for twice the edges
// the graph has now (or 2^26 more), so that inserts rarely outgrow it; a
// slot whose scratch the graph has outgrown runs lone BFS queries through the
// batched engine instead
static int64_t slot_bfs_scratch(const hn_graph* g) {
    BfsBufs bb;
    return bfs_layout(g->nv, std::max<int64_t>(2 * g->ne, g->ne + ((int64_t)1 << 26)), nullptr, &bb);
}

// device bytes one launch slot holds: an engine and the digests (batched
// launches run in digest mode and never hold k x nv distances), plus one
// distance row and the level buffers of the single-source BFS
static int64_t slot_bytes(const hn_graph* g, bool plain) {
    (void)plain;
    const int64_t nv = g->nv > 0 ? g->nv : 1;
    return ms_engine_bytes(nv, g->ne) + nv * 4 + slot_bfs_scratch(g) + 64 * 24;
}

static void destroy_server(hn_server* s) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_disp.notify_all();
    for (auto& sl : s->slots) sl->cv.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    for (auto& sl : s->slots) {
        if (sl->th.joinable()) sl->th.join();
        ms_engine_destroy(sl->eng);
        dfree(sl->dist);
        dfree(sl->bfs_scratch);
        if (sl->h_digest) (void)hipHostFree(sl->h_digest);
        if (sl->stream) (void)hipStreamDestroy(sl->stream);
    }
    dfree(s->win_ent);
    if (s->upd_stream) (void)hipStreamDestroy(s->upd_stream);
    delete s;
}

}  // namespace hn

using namespace hn;

extern "C" {

hn_status hn_server_slot_bytes(const hn_graph* g, int32_t max_batch, int64_t* bytes) {
    HN_ARG(g && bytes, "null argument");
    HN_ARG(max_batch >= 1 && max_batch <= 64, "max_batch must be in [1, 64]");
    *bytes = slot_bytes(g, false);
    return HN_OK;
}

hn_status hn_server_create(hn_graph* g, const hn_serve_config* cfg, hn_server** out) {
    HN_ARG(g && cfg && out, "null argument");
    *out = nullptr;
    HN_ARG(cfg->max_batch >= 1 && cfg->max_batch <= 64, "max_batch must be in [1, 64]");
    HN_ARG(cfg->slots >= 1, "slots must be >= 1");
    hn_server* s = new hn_server();
    s->g = g;
    s->cfg = *cfg;
    const int64_t nv = g->nv > 0 ? g->nv : 1;
    s->slot_bytes = slot_bytes(g, cfg->plain != 0);
    int64_t budget = cfg->mem_budget_bytes;
    if (budget <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            fr = (size_t)1 << 34;
        }
        budget = (int64_t)(fr / 4);
    }
    int nslots = cfg->plain ? 1 : (int)std::min<int64_t>(cfg->slots, budget / s->slot_bytes);
    if (nslots < 1) {
        delete s;
        set_error("memory budget does not fit one launch slot");
        return HN_E_OOM;
    }
    s->update_budget = std::max<int64_t>(budget - nslots * s->slot_bytes, 0);
    s->window_max_entries = std::min<int64_t>(
        WINDOW_MAX_ENTRIES, s->update_budget / (2 * (nslots * WINDOW_ENGINE_BYTES + WINDOW_SERVER_BYTES)));
    s->cfg.slots = nslots;
    if (hipStreamCreateWithFlags(&s->upd_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        set_error("server: stream creation failed");
        return HN_E_HIP;
    }
    {
        // known before any query runs (the slots only read it)
        bool sym = false;
        const hn_status ss = graph_is_symmetric(g, s->upd_stream, &sym);
        if (ss != HN_OK) {
            (void)hipStreamDestroy(s->upd_stream);
            delete s;
            return ss;
        }
    }
    for (int i = 0; i < nslots; i++) {
        auto sl = std::make_unique<ServeSlot>();
        hn_status st = ms_engine_create(g, &sl->eng);
        if (st == HN_OK) st = dalloc(&sl->dist, (size_t)nv);
        if (st == HN_OK) {
            sl->bfs_scratch_bytes = slot_bfs_scratch(g);
            st = dalloc(&sl->bfs_scratch, (size_t)sl->bfs_scratch_bytes);
        }
        if (st == HN_OK && (hipHostMalloc(reinterpret_cast<void**>(&sl->h_digest), 64 * 3 * 8, 0) != hipSuccess ||
                            hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess)) {
            (void)hipGetLastError();
            set_error("server: slot resources");
            st = HN_E_HIP;
        }
        s->slots.push_back(std::move(sl));
        if (st != HN_OK) {
            destroy_server(s);
            return st;
        }
    }
    for (auto& sl : s->slots) {
        ServeSlot* p = sl.get();
        p->th = std::thread(worker_loop, s, p);
    }
    s->dispatcher = std::thread(dispatcher_loop, s);
    *out = s;
    return HN_OK;
}

hn_status hn_server_submit(hn_server* s, const hn_request* req, int64_t* id) {
    HN_ARG(s && req, "null argument");
    const bool q = req->kind == HN_REQ_BFS || req->kind == HN_REQ_TWO_HOP;
    HN_ARG(q || req->kind == HN_REQ_INSERT || req->kind == HN_REQ_ERASE, "unknown request kind");
    if (q) HN_ARG(req->source >= 0 && req->source < s->g->nv, "query source outside [0, nv)");
    else HN_ARG(req->n >= 0 && (req->n == 0 || (req->src && req->dst)), "bad update batch");
    auto r = std::make_unique<ServeReq>();
    r->kind = req->kind;
    r->source = req->source;
    if (!q) {
        r->src.assign(req->src, req->src + req->n);
        r->dst.assign(req->dst, req->dst + req->n);
    }
    const int64_t t = now_ns();
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = (int64_t)s->reqs.size();
    hn_result res{};
    res.id = r->id;
    res.kind = r->kind;
    res.status = -1;
    res.t_arrive_ns = t;
    s->results.push_back(res);
    s->queue.push_back(r.get());
    if (id) *id = r->id;
    s->reqs.push_back(std::move(r));
    s->cv_disp.notify_one();
    return HN_OK;
}

hn_status hn_server_drain(hn_server* s) {
    HN_ARG(s, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed >= (int64_t)s->reqs.size(); });
    return HN_OK;
}

hn_status hn_server_replay(hn_server* s, const hn_request* reqs, const int64_t* arrival_ns, int64_t n) {
    HN_ARG(s && (n == 0 || (reqs && arrival_ns)), "null argument");
    for (int64_t i = 0; i < n; i++) {
        const hn_request& r = reqs[i];
        const bool q = r.kind == HN_REQ_BFS || r.kind == HN_REQ_TWO_HOP;
        if (q) HN_ARG(r.source >= 0 && r.source < s->g->nv, "query source outside [0, nv)");
        HN_ARG(i == 0 || arrival_ns[i] >= arrival_ns[i - 1], "arrival times must be non-decreasing");
    }
    hn_status rc = HN_OK;
    std::thread client([&] {
        const int64_t t0 = now_ns();
        for (int64_t i = 0; i < n; i++) {
            const int64_t target = t0 + arrival_ns[i];
            int64_t t = now_ns();
            if (target - t > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(target - t - 100000));
            while (now_ns() < target) {
            }
            hn_status st = hn_server_submit(s, &reqs[i], nullptr);
            if (st != HN_OK) rc = st;
        }
    });
    client.join();
    HN_TRY(rc);
    return hn_server_drain(s);
}

hn_status hn_server_results(hn_server* s, hn_result* out, int64_t cap, int64_t* n) {
    HN_ARG(s, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t m = (int64_t)s->results.size();
    if (n) *n = m;
    if (out) {
        if (cap < m) {
            set_error("server results: capacity too small");
            return HN_E_CAPACITY;
        }
        std::memcpy(out, s->results.data(), sizeof(hn_result) * (size_t)m);
    }
    return HN_OK;
}

hn_status hn_server_slots(hn_server* s, int32_t* slots) {
    HN_ARG(s && slots, "null argument");
    *slots = s->cfg.slots;
    return HN_OK;
}

hn_status hn_server_destroy(hn_server* s) {
    if (s) destroy_server(s);
    return HN_OK;
}

}  // extern "C"

// galn_server.cpp -- serving loop: one resident index, requests queued from any thread,
// searched in shared launches by several launch slots at once.
//
// The reference has no serving mode (it is a one-shot CLI, src/galn.cu:146-178); this is the
// request front end for the same search. A request is a galn_hip_reads with its own options.
//
// Launch slots: each slot is a worker thread with its own device read set, stream and stack pool
// (slot 0 those of the index), so up to `slots` launches are on the device together. A launch is bounded by its longest
// search (a read that backtracks millions of entries keeps one wave busy while the rest of the
// GPU has finished), and the next slot's launch fills the device meanwhile: requests that arrive
// during a launch's tail start at once instead of waiting for it.
//
// Each launch takes, in arrival order, up to `quantum_reads` reads from every request that still
// has reads to search, then tops the launch up to `max_launch_reads` from the oldest requests. A
// request's reads may be in several launches at once; each launch's records go to the request
// by read index, so the request's records are in read order whichever launch ends first. A
// read's search depends only on the read, its batch options and the index, so its records do
// not depend on what it shares a launch with.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "galn_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    int n = 0;
    int nb = 0;
    std::vector<uint8_t> query;
    std::vector<uint64_t> qoff;
    std::vector<int32_t> qlen, max_diff, batch, n_stacks;
    std::vector<galn_search_opt_t> bopts;
    uint64_t footprint = 0;
    Clock::time_point t_submit, t_done;
    // scheduling state (under the server lock)
    int next = 0;       // first read not yet given to a launch
    int inflight = 0;   // launches holding reads of this request
    int done_reads = 0; // reads whose launch has completed
    int status = 0;     // 0 pending, 1 finished, < 0 library error of a launch
    bool done = false;  // no launch touches the request any more: wait() and result() may go on
    std::string err;    // text of that error
    // results by read (each read written by the one launch that searched it)
    std::vector<int32_t> n_aln;
    std::vector<uint64_t> pops, trace, bytes;
    std::map<int, std::vector<galn_aln_rec_t>> recs;  // records of the reads from each start index on
};

struct Piece {
    Request *q;
    int start, count;
};

// device bytes a request's reads occupy in a launch: query bytes, the per-read input and
// output arrays, and two record slots per read
uint64_t footprint_of(const galn_hip_reads *r, uint64_t qbytes)
{
    return qbytes + (uint64_t)r->n_reads * (128 + 2 * sizeof(galn_aln_rec_t)) +
           (uint64_t)r->n_batches * (sizeof(galn_search_opt_t) + 4);
}

}  // namespace

// galn_search.hip: this thread's galn_hip_last_error text; the open-handle count of an index;
// the count of servers holding an index, and the mark of their worker threads (only those may
// search an index that a server holds); a slot read set's own stream and stack pool
void galn_hip_set_error(const char *msg);
void galn_hip_index_hold(galn_hip_index *idx, int delta);
void galn_hip_index_serve(galn_hip_index *idx, int delta);
void galn_hip_server_worker_mark();
int galn_hip_readset_make_private(galn_hip_readset *rs);
int galn_hip_readset_reserve(galn_hip_readset *rs, int n, uint64_t qbytes, int nb);
int galn_hip_readset_reserve_search(galn_hip_readset *rs, int n, int max_len, uint32_t flags);

namespace {

int srv_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    galn_hip_set_error(buf);
    return code;
}

}  // namespace

struct galn_hip_server {
    galn_hip_index *idx = nullptr;
    galn_hip_server_cfg_t cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> waiting;                    // requests with reads not yet launched
    std::map<uint64_t, std::unique_ptr<Request>> reqs; // every request not yet collected
    uint64_t next_id = 1;
    uint64_t queued_bytes = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<galn_hip_readset *> slot_rs;
    double stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int busy_slots = 0;
    // launch size: between launch_lo and launch_hi reads (both max_launch_reads when it is set),
    // the reads the server finished in LAUNCH_SECONDS of recent busy time
    int launch_lo = 0, launch_hi = 0, launch_reads = 0;
    double rate_reads = 0, rate_sec = 0;  // reads finished and busy seconds, older launches decayed
    Clock::time_point busy_mark;          // start of the busy time not yet counted
    // GALN_HIP_TEST_SERVER_FAIL=n: the n-th launch (from 1) fails on the host before it reaches
    // the device, so that tests can drive the error path of a request spread over several launches
    long fail_at = 0;
    long n_launch = 0;
};

namespace {

// A launch of the adaptive size holds the reads the server finished in this many seconds of
// recent busy time. Larger launches waste less of the device on a launch's tail where the longest
// searches take a fraction of a second (100 Mbp index, ~4e6 reads/s: peak 4.70e6 reads/s with
// 1M-read launches against 3.51e6 with 262,144), but where single searches take seconds (the
// 3.1 Gbp human-like index, ~2e5 reads/s) a few large launches leave their tails unshared (peak
// 6.7e4 with 1M-read launches against 2.0e5 with 262,144).
constexpr double LAUNCH_SECONDS = 0.5;
constexpr int LAUNCH_LO = 262144, LAUNCH_HI = 1 << 20;

// Choose the reads of the next launch (caller holds the lock).
void schedule(galn_hip_server *S, std::vector<Piece> &pieces)
{
    pieces.clear();
    int room = S->launch_reads;
    for (Request *q : S->waiting) {
        if (room == 0) break;
        const int take = std::min({q->n - q->next, S->cfg.quantum_reads, room});
        pieces.push_back({q, q->next, take});
        room -= take;
    }
    for (Piece &p : pieces) {
        if (room == 0) break;
        const int more = std::min(p.q->n - p.start - p.count, room);
        p.count += more;
        room -= more;
    }
    for (Piece &p : pieces) {
        p.q->next = p.start + p.count;
        p.q->inflight += 1;
    }
    S->waiting.erase(std::remove_if(S->waiting.begin(), S->waiting.end(), [](Request *q) { return q->next == q->n; }),
                     S->waiting.end());
}

// A launch of `pieces` has ended with code r (caller holds the lock).
void complete(galn_hip_server *S, const std::vector<Piece> &pieces, int r, const std::string &err,
              std::vector<std::vector<galn_aln_rec_t>> &piece_recs, Clock::time_point t1)
{
    for (size_t k = 0; k < pieces.size(); ++k) {
        const Piece &p = pieces[k];
        Request &q = *p.q;
        q.inflight -= 1;
        q.done_reads += p.count;
        if (r == 0) {
            q.recs[p.start] = std::move(piece_recs[k]);
        } else if (q.status == 0) {
            q.status = r;
            q.err = err;
        }
        if (q.status < 0 && q.next < q.n) {
            // a failed request gets no further launches
            S->waiting.erase(std::remove(S->waiting.begin(), S->waiting.end(), &q), S->waiting.end());
            q.next = q.n;
        }
        if (q.status == 0 && q.done_reads == q.n) q.status = 1;
        if (!q.done && q.inflight == 0 && (q.status < 0 || q.done_reads == q.n)) {
            q.done = true;
            q.t_done = t1;
            S->stats[2] += 1;
        }
    }
}

void worker_main(galn_hip_server *S, int slot)
{
    galn_hip_server_worker_mark();
    std::vector<Piece> pieces;
    std::vector<uint8_t> query;
    std::vector<uint64_t> qoff, off, pops, trace, bytes;
    std::vector<int32_t> qlen, max_diff, batch, n_stacks, n_aln;
    std::vector<galn_search_opt_t> bopts;
    std::vector<galn_aln_rec_t> rec(1 << 16);
    std::vector<std::vector<galn_aln_rec_t>> piece_recs;
    const bool want_trace = (S->cfg.flags & GALN_HIP_TRACE) != 0;
    galn_hip_readset *&rs = S->slot_rs[slot];
    for (;;) {
        long launch_no;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->waiting.empty(); });
            if (S->waiting.empty()) return;  // stop requested and nothing left
            schedule(S, pieces);
            launch_no = ++S->n_launch;
            if (S->busy_slots == 0) S->busy_mark = Clock::now();
            S->busy_slots += 1;
            S->stats[4] = std::max(S->stats[4], (double)S->busy_slots);
        }
        // the merged read set of this launch (the reads of a scheduled piece are not changed by
        // other threads while they are in flight)
        query.clear(), qoff.clear(), qlen.clear(), max_diff.clear(), batch.clear(), n_stacks.clear(), bopts.clear();
        for (const Piece &p : pieces) {
            // a request keeps its queries back to back, so a piece's are one range
            const Request &q = *p.q;
            const int b0 = (int)bopts.size(), e = p.start + p.count;
            bopts.insert(bopts.end(), q.bopts.begin(), q.bopts.end());
            n_stacks.insert(n_stacks.end(), q.n_stacks.begin(), q.n_stacks.end());
            const uint64_t q0 = q.qoff[p.start], q1 = q.qoff[e - 1] + (uint64_t)q.qlen[e - 1];
            const uint64_t base = query.size();
            query.insert(query.end(), q.query.begin() + q0, q.query.begin() + q1);
            for (int i = p.start; i < e; ++i) {
                qoff.push_back(base + q.qoff[i] - q0);
                batch.push_back(b0 + q.batch[i]);
            }
            qlen.insert(qlen.end(), q.qlen.begin() + p.start, q.qlen.begin() + e);
            max_diff.insert(max_diff.end(), q.max_diff.begin() + p.start, q.max_diff.begin() + e);
        }
        const int m = (int)qlen.size();
        galn_hip_reads R{};
        R.n_reads = m;
        R.query = query.data();
        R.qoff = qoff.data();
        R.qlen = qlen.data();
        R.max_diff = max_diff.data();
        R.batch = batch.data();
        R.n_batches = (int)bopts.size();
        R.bopts = bopts.data();
        R.n_stacks = n_stacks.data();
        const auto t0 = Clock::now();
        int r = 0;
        if (launch_no == S->fail_at)
            r = srv_fail(GALN_HIP_E_DEVICE, "launch %ld: failure injected by GALN_HIP_TEST_SERVER_FAIL", launch_no);
        if (!r) {
            if (rs) {
                r = galn_hip_readset_reload(rs, &R);
            } else if (!(r = galn_hip_readset_upload(S->idx, &R, &rs)) &&
                       !(r = slot == 0 ? 0 : galn_hip_readset_make_private(rs))) {
                // slot 0 keeps the index's own stream and pool (no one else may search the index
                // while the server holds it): one stream fewer, and HIP maps streams onto 4
                // hardware queues by default, so with a stream per slot two slots shared one and
                // their launches ran one after the other (measured: 3.2-3.6e6 reads/s peak on the
                // 100 Mbp index, 4.3e6 with slot 0 on the index's stream)
                // room for a full launch of reads up to 256 bp, then this launch's reads again
                r = galn_hip_readset_reserve(rs, S->cfg.max_launch_reads, (uint64_t)S->cfg.max_launch_reads * 256,
                                             std::max((int)bopts.size(), 1024));
                if (!r) r = galn_hip_readset_reload(rs, &R);
                // and the search's buffers for a full launch, so that they are not regrown (each
                // hipFree waits for the other slots' launches) as launches get larger
                if (!r) r = galn_hip_readset_reserve_search(rs, S->cfg.max_launch_reads, 256, S->cfg.flags);
            }
        }
        if (!r) {
            // the library's tuning, with the configured wave-engine cutoff when one is set
            galn_hip_tune_t t{};
            t.sort_by_cost = 1;
            t.heavy_pops = S->cfg.heavy_pops;
            r = galn_hip_readset_search(S->idx, rs, S->cfg.form, S->cfg.flags, S->cfg.heavy_pops ? &t : nullptr);
        }
        n_aln.resize(m), off.resize(m), pops.resize(m), trace.resize(m), bytes.resize(m);
        galn_hip_out O{};
        O.n_aln = n_aln.data();
        O.rec_off = off.data();
        O.n_pops = pops.data();
        O.trace = want_trace ? trace.data() : nullptr;
        O.occ_bytes = bytes.data();
        if (!r) {
            O.rec = rec.data();
            O.rec_cap = rec.size();
            r = galn_hip_readset_results(rs, &O);
            if (r == GALN_HIP_E_RECCAP) {
                rec.resize(O.n_rec + O.n_rec / 4);
                O.rec = rec.data();
                O.rec_cap = rec.size();
                r = galn_hip_readset_results(rs, &O);
            }
        }
        const auto t1 = Clock::now();
        const std::string err = r ? std::string(galn_hip_last_error()) : std::string();
        // results by read, and each piece's records in read order (the request's arrays are sized
        // at submit; this launch alone writes these reads)
        piece_recs.assign(pieces.size(), {});
        if (!r) {
            int j = 0;
            for (size_t k = 0; k < pieces.size(); ++k) {
                const Piece &p = pieces[k];
                Request &q = *p.q;
                // the launch's records are in read order: a piece's are one range
                const int j0 = j;
                uint64_t r1 = off[j0];
                for (int i = p.start; i < p.start + p.count; ++i, ++j) {
                    q.n_aln[i] = n_aln[j];
                    q.pops[i] = pops[j];
                    q.trace[i] = want_trace ? trace[j] : 0;
                    q.bytes[i] = bytes[j];
                    r1 = off[j] + (uint64_t)std::max(n_aln[j], 0);
                }
                piece_recs[k].assign(rec.begin() + off[j0], rec.begin() + r1);
            }
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            complete(S, pieces, r, err, piece_recs, t1);
            S->busy_slots -= 1;
            S->stats[0] += 1;
            S->stats[1] += m;
            S->stats[3] += std::chrono::duration<double, std::milli>(t1 - t0).count();
            // recent throughput: each finished launch halves the weight of the ones before it
            S->rate_reads = 0.5 * S->rate_reads + m;
            S->rate_sec = 0.5 * S->rate_sec + std::chrono::duration<double>(t1 - S->busy_mark).count();
            S->busy_mark = t1;
            if (S->launch_lo < S->launch_hi && S->rate_sec > 0)
                S->launch_reads = (int)std::clamp(S->rate_reads / S->rate_sec * LAUNCH_SECONDS,
                                                  (double)S->launch_lo, (double)S->launch_hi);
            S->stats[6] = S->launch_reads;
        }
        S->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int galn_hip_server_open(galn_hip_index *idx, const galn_hip_server_cfg_t *cfg, galn_hip_server **out)
{
    if (!idx || !out) return srv_fail(GALN_HIP_E_ARG, "server_open: null argument");
    *out = nullptr;
    galn_hip_server *S = new galn_hip_server;
    S->idx = idx;
    if (cfg) S->cfg = *cfg;
    // launch size: fixed when configured, else adaptive (starting at the smallest)
    if (S->cfg.max_launch_reads > 0) {
        S->launch_lo = S->launch_hi = S->cfg.max_launch_reads;
    } else {
        S->launch_lo = LAUNCH_LO;
        S->launch_hi = S->cfg.max_launch_reads = LAUNCH_HI;
    }
    S->launch_reads = S->launch_lo;
    S->stats[6] = S->launch_reads;
    if (S->cfg.quantum_reads <= 0) S->cfg.quantum_reads = 16384;
    S->cfg.quantum_reads = std::min(S->cfg.quantum_reads, S->launch_lo);
    if (S->cfg.max_queued_bytes == 0) S->cfg.max_queued_bytes = 4ull << 30;
    if (S->cfg.form != 0) S->cfg.form = 1;
    if (S->cfg.slots <= 0) S->cfg.slots = 3;
    S->cfg.slots = std::min(S->cfg.slots, 16);
    if (const char *v = getenv("GALN_HIP_TEST_SERVER_FAIL")) S->fail_at = strtol(v, nullptr, 10);
    galn_hip_index_hold(idx, 1);
    galn_hip_index_serve(idx, 1);
    S->slot_rs.assign(S->cfg.slots, nullptr);
    for (int k = 0; k < S->cfg.slots; ++k) S->workers.emplace_back(worker_main, S, k);
    *out = S;
    return 0;
}

int galn_hip_server_submit(galn_hip_server *S, const galn_hip_reads *r, uint64_t *id)
{
    if (!S || !r || !id) return srv_fail(GALN_HIP_E_ARG, "server_submit: null argument");
    if (r->n_reads < 0 || r->n_batches < 1 || !r->bopts || !r->n_stacks)
        return srv_fail(GALN_HIP_E_ARG, "server_submit: bad read set");
    if (r->n_reads > 0 && (!r->query || !r->qoff || !r->qlen || !r->max_diff || !r->batch))
        return srv_fail(GALN_HIP_E_ARG, "server_submit: null array");
    const int n = r->n_reads;
    auto q = std::make_unique<Request>();
    q->n = n;
    q->nb = r->n_batches;
    uint64_t qbytes = 0;
    q->qoff.resize(n);
    for (int i = 0; i < n; ++i) {
        if (r->batch[i] < 0 || r->batch[i] >= r->n_batches || r->qlen[i] < 0)
            return srv_fail(GALN_HIP_E_ARG, "server_submit: read %d: bad batch or length", i);
        q->qoff[i] = qbytes;
        qbytes += (uint64_t)r->qlen[i];
    }
    q->query.resize(qbytes);
    for (int i = 0; i < n; ++i) memcpy(q->query.data() + q->qoff[i], r->query + r->qoff[i], (size_t)r->qlen[i]);
    q->qlen.assign(r->qlen, r->qlen + n);
    q->max_diff.assign(r->max_diff, r->max_diff + n);
    q->batch.assign(r->batch, r->batch + n);
    q->bopts.assign(r->bopts, r->bopts + r->n_batches);
    q->n_stacks.assign(r->n_stacks, r->n_stacks + r->n_batches);
    q->n_aln.assign(n, 0);
    q->pops.assign(n, 0);
    q->trace.assign(n, 0);
    q->bytes.assign(n, 0);
    q->footprint = footprint_of(r, qbytes);
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->stop) return srv_fail(GALN_HIP_E_ARG, "server_submit: server is closing");
    // admission: a request is always taken into an empty queue, however large
    if (S->queued_bytes > 0 && S->queued_bytes + q->footprint > S->cfg.max_queued_bytes)
        return GALN_HIP_E_BUSY;
    q->id = S->next_id++;
    q->t_submit = Clock::now();
    S->queued_bytes += q->footprint;
    Request *raw = q.get();
    *id = q->id;
    S->reqs.emplace(q->id, std::move(q));
    if (n == 0) {
        raw->status = 1;
        raw->done = true;
        raw->t_done = raw->t_submit;
        S->stats[2] += 1;
        lk.unlock();
        S->cv_done.notify_all();
        return 0;
    }
    S->waiting.push_back(raw);
    lk.unlock();
    S->cv_work.notify_one();
    return 0;
}

int galn_hip_server_wait(galn_hip_server *S, uint64_t id, int timeout_ms)
{
    if (!S) return srv_fail(GALN_HIP_E_ARG, "server_wait: null handle");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return srv_fail(GALN_HIP_E_ARG, "server_wait: unknown request %llu", (unsigned long long)id);
    Request *q = it->second.get();
    auto ready = [&] { return q->done; };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, ready);
    else if (timeout_ms > 0)
        S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    if (!q->done) return 0;
    if (q->status < 0) galn_hip_set_error(q->err.c_str());
    return q->status > 0 ? 1 : q->status;
}

int galn_hip_server_result(galn_hip_server *S, uint64_t id, galn_hip_out *out, double *latency_ms)
{
    if (!S || !out) return srv_fail(GALN_HIP_E_ARG, "server_result: null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return srv_fail(GALN_HIP_E_ARG, "server_result: unknown request %llu", (unsigned long long)id);
    Request *q = it->second.get();
    if (!q->done) return srv_fail(GALN_HIP_E_BUSY, "server_result: request %llu not finished", (unsigned long long)id);
    int r = 0;
    if (q->status < 0) {
        r = q->status;
        galn_hip_set_error(q->err.c_str());
    } else {
        uint64_t total = 0;
        for (const auto &kv : q->recs) total += kv.second.size();
        out->n_rec = total;
        if (q->n > 0 && (!out->n_aln || !out->rec_off)) return srv_fail(GALN_HIP_E_ARG, "server_result: null array");
        if (total > out->rec_cap || (total > 0 && !out->rec))
            return srv_fail(GALN_HIP_E_RECCAP, "server_result: record array holds %llu, %llu needed",
                            (unsigned long long)out->rec_cap, (unsigned long long)total);
        // pieces in read order (they cover reads 0 .. n - 1): the records, and each read's offset
        uint64_t at = 0;
        for (auto pc = q->recs.begin(); pc != q->recs.end(); ++pc) {
            const auto nx = std::next(pc);
            const int end = nx == q->recs.end() ? q->n : nx->first;
            std::copy(pc->second.begin(), pc->second.end(), out->rec + at);
            for (int i = pc->first; i < end; ++i) {
                out->rec_off[i] = at;
                at += (uint64_t)std::max(q->n_aln[i], 0);
            }
        }
        std::copy(q->n_aln.begin(), q->n_aln.end(), out->n_aln);
        if (out->n_pops) std::copy(q->pops.begin(), q->pops.end(), out->n_pops);
        if (out->trace) std::copy(q->trace.begin(), q->trace.end(), out->trace);
        if (out->occ_bytes) std::copy(q->bytes.begin(), q->bytes.end(), out->occ_bytes);
    }
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count();
    S->queued_bytes -= q->footprint;
    S->reqs.erase(it);
    return r;
}

int galn_hip_server_stats(galn_hip_server *S, double out[8])
{
    if (!S || !out) return srv_fail(GALN_HIP_E_ARG, "server_stats: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    std::copy(S->stats, S->stats + 8, out);
    out[5] = S->cfg.slots;
    return 0;
}

int galn_hip_server_close(galn_hip_server *S)
{
    if (!S) return 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &w : S->workers)
        if (w.joinable()) w.join();
    for (galn_hip_readset *rs : S->slot_rs)
        if (rs) galn_hip_readset_free(rs);
    galn_hip_index_serve(S->idx, -1);
    galn_hip_index_hold(S->idx, -1);
    delete S;
    return 0;
}

}  // extern "C"

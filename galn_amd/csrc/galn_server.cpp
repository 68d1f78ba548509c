// galn_server.cpp -- serving loop: one resident index, requests queued from any thread,
// searched by one worker thread in shared launches.
//
// The reference has no serving mode (it is a one-shot CLI, src/galn.cu:146-178); this is the
// request front end for the same search. A request is a galn_hip_reads with its own options.
// Each launch takes, in arrival order, up to `quantum_reads` reads from every request that
// still has reads to search, then tops the launch up to `max_launch_reads` from the oldest
// requests; each request contributes one contiguous range of its reads per launch, so its
// records are appended in read order. A read's search depends only on the read, its batch
// options and the index, so its records do not depend on what it shares a launch with.
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
    int next = 0;       // first read not yet given to a launch
    int finished = 0;   // reads whose launch has completed
    int status = 0;     // 0 pending, 1 finished, < 0 library error of a launch
    std::string err;    // text of that error
    // results, in read order
    std::vector<int32_t> n_aln;
    std::vector<uint64_t> rec_off, pops, trace, bytes;
    std::vector<galn_aln_rec_t> rec;
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
// search an index that a server holds)
void galn_hip_set_error(const char *msg);
void galn_hip_index_hold(galn_hip_index *idx, int delta);
void galn_hip_index_serve(galn_hip_index *idx, int delta);
void galn_hip_server_worker_mark();

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
    std::thread worker;
    galn_hip_readset *rs = nullptr;
    double stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // GALN_HIP_TEST_SERVER_FAIL=n: the n-th launch (from 1) fails on the host before it reaches
    // the device, so that tests can drive the error path of a request spread over several launches
    long fail_at = 0;
    long n_launch = 0;
};

namespace {

// Choose the reads of the next launch (caller holds the lock).
void schedule(galn_hip_server *S, std::vector<Piece> &pieces)
{
    pieces.clear();
    int room = S->cfg.max_launch_reads;
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
    for (Piece &p : pieces) p.q->next = p.start + p.count;
    while (!S->waiting.empty() && S->waiting.front()->next == S->waiting.front()->n) S->waiting.pop_front();
    // requests further back may also be fully given out
    S->waiting.erase(std::remove_if(S->waiting.begin(), S->waiting.end(), [](Request *q) { return q->next == q->n; }),
                     S->waiting.end());
}

void worker_main(galn_hip_server *S)
{
    galn_hip_server_worker_mark();
    std::vector<Piece> pieces;
    std::vector<uint8_t> query;
    std::vector<uint64_t> qoff, off, pops, trace, bytes;
    std::vector<int32_t> qlen, max_diff, batch, n_stacks, n_aln;
    std::vector<galn_search_opt_t> bopts;
    std::vector<galn_aln_rec_t> rec(1 << 16);
    const bool want_trace = (S->cfg.flags & GALN_HIP_TRACE) != 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->waiting.empty(); });
            if (S->waiting.empty()) return;  // stop requested and nothing left
            schedule(S, pieces);
        }
        // the merged read set of this launch (the pieces' requests are not changed by
        // other threads while their reads are in flight)
        query.clear(), qoff.clear(), qlen.clear(), max_diff.clear(), batch.clear(), n_stacks.clear(), bopts.clear();
        for (const Piece &p : pieces) {
            const Request &q = *p.q;
            const int b0 = (int)bopts.size();
            bopts.insert(bopts.end(), q.bopts.begin(), q.bopts.end());
            n_stacks.insert(n_stacks.end(), q.n_stacks.begin(), q.n_stacks.end());
            for (int i = p.start; i < p.start + p.count; ++i) {
                qoff.push_back(query.size());
                query.insert(query.end(), q.query.begin() + q.qoff[i], q.query.begin() + q.qoff[i] + q.qlen[i]);
                qlen.push_back(q.qlen[i]);
                max_diff.push_back(q.max_diff[i]);
                batch.push_back(b0 + q.batch[i]);
            }
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
        if (++S->n_launch == S->fail_at)
            r = srv_fail(GALN_HIP_E_DEVICE, "launch %ld: failure injected by GALN_HIP_TEST_SERVER_FAIL", S->n_launch);
        if (!r) r = S->rs ? galn_hip_readset_reload(S->rs, &R) : galn_hip_readset_upload(S->idx, &R, &S->rs);
        if (!r) r = galn_hip_readset_search(S->idx, S->rs, S->cfg.form, S->cfg.flags, nullptr);
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
            r = galn_hip_readset_results(S->rs, &O);
            if (r == GALN_HIP_E_RECCAP) {
                rec.resize(O.n_rec + O.n_rec / 4);
                O.rec = rec.data();
                O.rec_cap = rec.size();
                r = galn_hip_readset_results(S->rs, &O);
            }
        }
        const auto t1 = Clock::now();
        const std::string err = r ? std::string(galn_hip_last_error()) : std::string();
        // hand the records to their requests, in read order
        int j = 0;
        for (const Piece &p : pieces) {
            Request &q = *p.q;
            for (int i = p.start; i < p.start + p.count; ++i, ++j) {
                if (r) continue;
                q.n_aln[i] = n_aln[j];
                q.pops[i] = pops[j];
                q.trace[i] = want_trace ? trace[j] : 0;
                q.bytes[i] = bytes[j];
                q.rec_off[i] = q.rec.size();
                if (n_aln[j] > 0) q.rec.insert(q.rec.end(), rec.begin() + off[j], rec.begin() + off[j] + n_aln[j]);
            }
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (const Piece &p : pieces) {
                Request &q = *p.q;
                q.finished += p.count;
                if (r && q.status == 0) {
                    q.status = r;
                    q.err = err;
                }
                if (q.status < 0 && q.next < q.n) {
                    // a failed request gets no further launches: out of the queue before its
                    // status is published (result() may release it as soon as the lock drops)
                    S->waiting.erase(std::remove(S->waiting.begin(), S->waiting.end(), &q), S->waiting.end());
                    q.next = q.n;
                }
                if (q.status < 0) q.finished = q.next;
                if (q.finished == q.n && q.status == 0) q.status = 1;
                if (q.status != 0 && q.t_done == Clock::time_point{}) {
                    q.t_done = t1;
                    S->stats[2] += 1;
                }
            }
            S->stats[0] += 1;
            S->stats[1] += m;
            S->stats[3] += std::chrono::duration<double, std::milli>(t1 - t0).count();
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
    if (S->cfg.max_launch_reads <= 0) S->cfg.max_launch_reads = 262144;
    if (S->cfg.quantum_reads <= 0) S->cfg.quantum_reads = 16384;
    S->cfg.quantum_reads = std::min(S->cfg.quantum_reads, S->cfg.max_launch_reads);
    if (S->cfg.max_queued_bytes == 0) S->cfg.max_queued_bytes = 4ull << 30;
    if (S->cfg.form != 0) S->cfg.form = 1;
    if (const char *v = getenv("GALN_HIP_TEST_SERVER_FAIL")) S->fail_at = strtol(v, nullptr, 10);
    galn_hip_index_hold(idx, 1);
    galn_hip_index_serve(idx, 1);
    S->worker = std::thread(worker_main, S);
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
    q->rec_off.assign(n, 0);
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
    auto ready = [&] { return q->status != 0; };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, ready);
    else if (timeout_ms > 0)
        S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
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
    if (q->status == 0 || q->finished < q->next)
        return srv_fail(GALN_HIP_E_BUSY, "server_result: request %llu not finished", (unsigned long long)id);
    int r = 0;
    if (q->status < 0) {
        r = q->status;
        galn_hip_set_error(q->err.c_str());
    } else {
        out->n_rec = q->rec.size();
        if (q->n > 0 && (!out->n_aln || !out->rec_off)) return srv_fail(GALN_HIP_E_ARG, "server_result: null array");
        if (q->rec.size() > out->rec_cap || (!q->rec.empty() && !out->rec))
            return srv_fail(GALN_HIP_E_RECCAP, "server_result: record array holds %llu, %llu needed",
                            (unsigned long long)out->rec_cap, (unsigned long long)q->rec.size());
        std::copy(q->n_aln.begin(), q->n_aln.end(), out->n_aln);
        std::copy(q->rec_off.begin(), q->rec_off.end(), out->rec_off);
        std::copy(q->rec.begin(), q->rec.end(), out->rec);
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
    if (S->worker.joinable()) S->worker.join();
    if (S->rs) galn_hip_readset_free(S->rs);
    galn_hip_index_serve(S->idx, -1);
    galn_hip_index_hold(S->idx, -1);
    delete S;
    return 0;
}

}  // extern "C"

// Serving loop: solves arrive at any time, are admitted by device footprint and stepped
// together with the solves already on the device (continuous batching).
//
// No counterpart in the reference (one mesh per process, app/solver.d:330-428).  A loop
// thread owns a batch (batch.cpp).  Each round it admits queued requests in arrival order
// while their footprint fits the device-byte limit and the solve count allows, advances
// every running solve by `steps_per_round` steps in one batched launch sequence, and
// retires the solves that stopped (converged by extensions.rmsDrop, maxSteps, tEnd or an
// error), freeing their device memory for the next admissions.  Requests that arrive while
// a round runs join at the next round boundary.  Each request's time to converged is
// measured from its submission to its retirement, on the host's steady clock.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "solver_impl.hpp"

namespace {

using Clock = std::chrono::steady_clock;

enum { T_QUEUED = 0, T_RUNNING = 1, T_DONE = 2, T_FAILED = 3 };

struct Ticket {
    const ebb_mesh* m = nullptr;
    const ebb_config* c = nullptr;
    int state = T_QUEUED;
    int64_t member = -1;
    int64_t cells = 0;
    int err = 0;
    double tArrive = 0, tAdmit = -1, tDone = -1;
    double status[10] = {};
    std::vector<double> q;
};

}  // namespace

struct ebb_server {
    ebb_batch* b = nullptr;
    int maxSolves = 0;
    int stepsPerRound = 0;
    bool keepResults = false;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv, doneCv;
    std::deque<int64_t> queue;
    std::map<int64_t, Ticket> tickets;
    std::vector<int64_t> running;
    int64_t nextTicket = 0, nDone = 0, rounds = 0;
    int64_t usedBytes = 0;   // device bytes the batch has admitted, as of the loop's last change
    double gpuSeconds = 0;
    bool stop = false;
    bool dead = false;       // the loop thread ended on an exception: no request is served
    std::thread loop;
    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void serveRounds(ebb_server* sv)
{
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        sv->cv.wait(lk, [&] { return sv->stop || !sv->queue.empty() || !sv->running.empty(); });
        if (sv->stop) break;
        // admission: arrival order, while the footprint fits and a slot is free
        while (!sv->queue.empty() && int(sv->running.size()) < sv->maxSolves) {
            const int64_t id = sv->queue.front();
            Ticket& t = sv->tickets[id];
            Config cfg = t.c->c;
            cfg.stagePath = cfg.stagePath == 3 ? 3 : 1;   // as ebb_batch_add will run it
            cfg.spreadTiles = true;
            const int64_t need = estimateDeviceBytes(t.m->b, cfg);
            int64_t info[6];
            ebb_batch_info(sv->b, info);
            if (need > info[3]) {              // can never fit
                sv->queue.pop_front();
                t.state = T_FAILED;
                t.err = EBB_ERR_ALLOC;
                t.tDone = sv->now();
                sv->nDone++;
                sv->doneCv.notify_all();
                continue;
            }
            if (info[2] + need > info[3]) break;   // wait for running solves to free memory
            sv->queue.pop_front();
            lk.unlock();
            int64_t member = -1;
            const int rc = ebb_batch_add(sv->b, t.m, t.c, &member);
            lk.lock();
            Ticket& t2 = sv->tickets[id];
            if (rc) {
                t2.state = T_FAILED;
                t2.err = rc;
                t2.tDone = sv->now();
                sv->nDone++;
                sv->doneCv.notify_all();
                if (rc == EBB_ERR_ALLOC) break;
                continue;
            }
            t2.member = member;
            t2.state = T_RUNNING;
            t2.tAdmit = sv->now();
            sv->running.push_back(id);
            int64_t inf[6];
            ebb_batch_info(sv->b, inf);   // the loop thread owns the batch
            sv->usedBytes = inf[2];
        }
        if (sv->running.empty()) continue;
        std::vector<int64_t> run = sv->running;
        lk.unlock();
        const auto g0 = Clock::now();
        int64_t left = 0;
        const int rc = ebb_batch_run(sv->b, sv->stepsPerRound, &left);
        const double gs = std::chrono::duration<double>(Clock::now() - g0).count();
        // retire the solves that stopped
        std::vector<std::pair<int64_t, std::vector<double>>> finished;
        for (int64_t id : run) {
            Ticket* tp;
            {
                std::lock_guard<std::mutex> g(sv->mu);
                tp = &sv->tickets[id];
            }
            double st[10];
            ebb_batch_status(sv->b, tp->member, st);
            if (rc == EBB_OK && st[3] != 0.0) continue;
            std::vector<double> q;
            if (sv->keepResults) {
                q.resize(size_t(tp->cells) * 4);
                ebb_batch_state_download(sv->b, tp->member, q.data());
            }
            {
                std::lock_guard<std::mutex> g(sv->mu);
                std::copy(st, st + 10, tp->status);
            }
            ebb_batch_remove(sv->b, tp->member);
            finished.emplace_back(id, std::move(q));
        }
        int64_t inf[6];
        ebb_batch_info(sv->b, inf);
        lk.lock();
        sv->usedBytes = inf[2];
        sv->gpuSeconds += gs;
        sv->rounds++;
        for (auto& f : finished) {
            Ticket& t = sv->tickets[f.first];
            t.q = std::move(f.second);
            t.state = (rc != EBB_OK || t.status[7] != 0.0) ? T_FAILED : T_DONE;
            t.err = rc != EBB_OK ? rc : int(t.status[7]);
            t.tDone = sv->now();
            sv->nDone++;
            sv->running.erase(std::find(sv->running.begin(), sv->running.end(), f.first));
        }
        if (!finished.empty()) sv->doneCv.notify_all();
    }
}

// The loop thread.  An exception (host memory for a result, say) ends the loop and fails
// every request not yet done; their solves stay in the batch until the server is destroyed.
void serveLoop(ebb_server* sv)
{
    try {
        serveRounds(sv);
    } catch (...) {
        std::lock_guard<std::mutex> g(sv->mu);   // the unwound rounds released it
        sv->dead = true;
        for (auto& kv : sv->tickets) {
            Ticket& t = kv.second;
            if (t.state != T_QUEUED && t.state != T_RUNNING) continue;
            t.state = T_FAILED;
            t.err = EBB_ERR_ALLOC;
            t.tDone = sv->now();
            sv->nDone++;
        }
        sv->queue.clear();
        sv->running.clear();
        sv->doneCv.notify_all();
    }
}

double quantile(std::vector<double> v, double p)
{
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double x = p * double(v.size() - 1);
    const size_t i = size_t(x);
    const double f = x - double(i);
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int ebb_server_create(int device, int64_t device_bytes_limit, int max_solves, int steps_per_round, int graph_steps,
                      int keep_results, ebb_server** out)
{
    return guarded([&]() -> int {
        if (!out || max_solves <= 0 || steps_per_round <= 0) return EBB_ERR_ARG;
        auto sv = std::make_unique<ebb_server>();
        int rc = ebb_batch_create(device, device_bytes_limit, graph_steps, 0, &sv->b);
        if (rc) return rc;
        sv->maxSolves = max_solves;
        sv->stepsPerRound = steps_per_round;
        sv->keepResults = keep_results != 0;
        sv->t0 = Clock::now();
        ebb_server* p = sv.get();
        sv->loop = std::thread(serveLoop, p);
        *out = sv.release();
        return EBB_OK;
    });
}

int ebb_server_submit(ebb_server* sv, const ebb_mesh* m, const ebb_config* c, int64_t* ticket)
{
    return guarded([&]() -> int {
        if (!sv || !m || !c || !ticket) return EBB_ERR_ARG;
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->dead) return EBB_ERR_STATE;
        const int64_t id = sv->nextTicket++;
        Ticket& t = sv->tickets[id];
        t.m = m;
        t.c = c;
        t.cells = m->b.nElem;
        t.tArrive = sv->now();
        sv->queue.push_back(id);
        sv->cv.notify_all();
        *ticket = id;
        return EBB_OK;
    });
}

int ebb_server_poll(ebb_server* sv, int64_t ticket, double out[16])
{
    return guarded([&]() -> int {
        if (!sv || !out) return EBB_ERR_ARG;
        std::lock_guard<std::mutex> g(sv->mu);
        auto it = sv->tickets.find(ticket);
        if (it == sv->tickets.end()) return EBB_ERR_ARG;
        const Ticket& t = it->second;
        out[0] = t.state; out[1] = t.err; out[2] = t.tArrive; out[3] = t.tAdmit; out[4] = t.tDone;
        out[5] = double(t.cells);
        std::copy(t.status, t.status + 10, out + 6);
        return EBB_OK;
    });
}

int ebb_server_result(ebb_server* sv, int64_t ticket, double* q)
{
    return guarded([&]() -> int {
        if (!sv || !q) return EBB_ERR_ARG;
        std::lock_guard<std::mutex> g(sv->mu);
        auto it = sv->tickets.find(ticket);
        if (it == sv->tickets.end() || it->second.state != T_DONE || it->second.q.empty()) return EBB_ERR_STATE;
        std::copy(it->second.q.begin(), it->second.q.end(), q);
        return EBB_OK;
    });
}

int ebb_server_wait(ebb_server* sv, int64_t n_done, double timeout_s)
{
    return guarded([&]() -> int {
        if (!sv) return EBB_ERR_ARG;
        std::unique_lock<std::mutex> lk(sv->mu);
        const bool ok = sv->doneCv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                            [&] { return sv->nDone >= n_done; });
        return ok ? EBB_OK : EBB_ERR_STATE;
    });
}

int ebb_server_stats(ebb_server* sv, double out[12])
{
    return guarded([&]() -> int {
        if (!sv || !out) return EBB_ERR_ARG;
        std::lock_guard<std::mutex> g(sv->mu);
        std::vector<double> lat;
        double first = 1e300, last = 0;
        int64_t done = 0, failed = 0;
        for (auto& kv : sv->tickets) {
            const Ticket& t = kv.second;
            if (t.state == T_DONE) {
                lat.push_back(t.tDone - t.tArrive);
                first = std::min(first, t.tArrive);
                last = std::max(last, t.tDone);
                done++;
            } else if (t.state == T_FAILED) {
                failed++;
            }
        }
        // the batch itself belongs to the loop thread: its admitted bytes are read from the
        // copy the loop keeps under the lock
        out[0] = double(sv->tickets.size()); out[1] = double(done); out[2] = double(failed);
        out[3] = double(sv->running.size()); out[4] = double(sv->queue.size());
        out[5] = quantile(lat, 0.5); out[6] = quantile(lat, 0.99);
        out[7] = done && last > first ? double(done) / (last - first) : 0.0;
        out[8] = double(sv->usedBytes); out[9] = sv->gpuSeconds; out[10] = double(sv->rounds); out[11] = sv->now();
        return EBB_OK;
    });
}

void ebb_server_destroy(ebb_server* sv)
{
    if (!sv) return;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
        sv->cv.notify_all();
    }
    if (sv->loop.joinable()) sv->loop.join();
    ebb_batch_destroy(sv->b);
    delete sv;
}

}  // extern "C"

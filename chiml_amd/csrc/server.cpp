// Sweep server: continuous batching of many device-resident simulations on
// one GPU (the C ABI is include/chiml/chiml_server.h).
//
// Three kinds of threads touch a request, one at a time:
//   caller threads   submit / poll / take (under the server mutex);
//   the builders     parse the deck and build the simulation (host tables,
//                    uploads) for admitted requests, off the stepping thread;
//   the stepper      polls every running simulation's stream and enqueues the
//                    next chunk of steps on each one that has drained, so
//                    chunks of different simulations overlap on the device.
// Admission: FIFO with backfill -- a request that does not fit the memory
// budget lets later, smaller ones pass, until it has been passed
// kMaxBackfill times; then admission waits for it.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/chiml/chiml_server.h"

namespace chiml {
void set_sim_error(const std::string& m);  // manager.cpp: chiml_sim_last_error of this thread
}

namespace {

using Clock = std::chrono::steady_clock;
constexpr int kMaxBackfill = 16;

struct Request {
    long long id = 0;
    std::string deck;
    size_t bytes = 0;  // estimate until built, then the measured footprint
    int state = CHIML_REQ_QUEUED;
    int passed = 0;  // times a later request was admitted ahead of it
    Clock::time_point t_submit, t_admit, t_done;
    chiml_sim_t* sim = nullptr;
    int nsteps = 0, done_steps = 0;
    double cells = 0.0;
    std::string err;
};

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

struct chiml_server {
    size_t budget = 0, held = 0;
    int max_active = 1, chunk = 1, graph = 0;
    bool batch = false;  // idle simulations step together through chiml_sim_step_batch
    std::mutex mu;
    std::condition_variable cv_build, cv_idle;
    std::map<long long, std::unique_ptr<Request>> reqs;
    std::deque<Request*> queue;     // submitted, not admitted
    std::deque<Request*> to_build;  // admitted, waiting for the builder
    std::vector<Request*> running;  // built, owned by the stepper
    std::vector<Request*> built;    // handed from builder to stepper
    int admitted = 0;               // building + running
    int peak_active = 0;
    double cell_updates = 0.0;
    long long next_id = 1;
    bool stop = false;
    std::vector<std::thread> builders;  // host-side setup of admitted requests runs here
    std::thread stepper;

    // Moves requests from the queue to the build list while they fit.
    void admit_locked() {
        for (auto it = queue.begin(); it != queue.end();) {
            Request* r = *it;
            if (admitted >= max_active) break;
            if (held + r->bytes <= budget || admitted == 0) {
                // a request still runs when nothing else does: an oversized
                // one, or one that waits only for finished results the caller
                // has not taken yet (so drain() always completes)
                for (auto jt = queue.begin(); jt != it; ++jt) ++(*jt)->passed;
                held += r->bytes;
                ++admitted;
                r->state = CHIML_REQ_BUILDING;
                r->t_admit = Clock::now();
                to_build.push_back(r);
                it = queue.erase(it);
                cv_build.notify_one();
                continue;
            }
            if (r->passed >= kMaxBackfill) break;  // nobody may pass it again
            ++it;
        }
    }

    void build_loop() {
        for (;;) {
            Request* r = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_build.wait(lk, [&] { return stop || !to_build.empty(); });
                if (stop && to_build.empty()) return;
                r = to_build.front();
                to_build.pop_front();
            }
            chiml_sim_t* sim = nullptr;
            std::string err;
            int n3[3], nsteps = 0, prec, nd, nf, nq;
            double dt;
            if (chiml_sim_create(r->deck.c_str(), &sim) != CHIML_OK) err = chiml_sim_last_error();
            else if (chiml_sim_info(sim, n3, &dt, &nsteps, &prec, &nd, &nf, &nq) != CHIML_OK) err = chiml_sim_last_error();
            else if (graph > 0 && chiml_sim_set_graph(sim, graph) != CHIML_OK) err = chiml_sim_last_error();
            size_t fp = 0;
            if (err.empty()) chiml_sim_footprint(sim, &fp);
            std::lock_guard<std::mutex> lk(mu);
            if (!err.empty()) {
                if (sim) chiml_sim_destroy(sim);
                r->state = CHIML_REQ_FAILED;
                r->err = err;
                r->t_done = Clock::now();
                held -= r->bytes;
                r->bytes = 0;
                --admitted;
                admit_locked();
                cv_idle.notify_all();
                continue;
            }
            held = held - r->bytes + fp;
            r->bytes = fp;
            r->sim = sim;
            r->nsteps = nsteps;
            r->cells = static_cast<double>(n3[0]) * n3[1] * n3[2];
            r->state = CHIML_REQ_RUNNING;
            built.push_back(r);
        }
    }

    // Fails every request of a batch whose shared launch could not be enqueued.
    void fail_all(const std::vector<Request*>& rs, std::vector<Request*>& finished) {
        const std::string e = chiml_sim_last_error();
        for (Request* r : rs) {
            r->err = e.empty() ? "batched step failed" : e;
            finished.push_back(r);
        }
    }

    void step_loop() {
        std::vector<Request*> finished, ready, batched;
        std::vector<chiml_sim_t*> sims;
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(mu);
                for (Request* r : built) running.push_back(r);
                built.clear();
                peak_active = std::max(peak_active, static_cast<int>(running.size()));
                if (stop && running.empty() && to_build.empty() && admitted == 0) return;
            }
            bool progressed = false;
            finished.clear();
            ready.clear();
            for (Request* r : running) {
                int idle = 0;
                if (chiml_sim_query(r->sim, &idle) != CHIML_OK) {
                    r->err = chiml_sim_last_error();
                    finished.push_back(r);
                    continue;
                }
                if (!idle) continue;
                if (r->done_steps >= r->nsteps) {
                    finished.push_back(r);
                    continue;
                }
                ready.push_back(r);
            }
            // every drained simulation with a whole chunk left joins one shared
            // batch (batch mode); the others take their next chunk alone
            batched.clear();
            sims.clear();
            for (Request* r : ready) {
                const int k = std::min(chunk, r->nsteps - r->done_steps);
                if (batch && k == chunk) {
                    batched.push_back(r);
                    sims.push_back(r->sim);
                    continue;
                }
                if (chiml_sim_step(r->sim, k) != CHIML_OK) {
                    r->err = chiml_sim_last_error();
                    finished.push_back(r);
                    continue;
                }
                r->done_steps += k;
                progressed = true;
            }
            if (!batched.empty()) {
                if (chiml_sim_step_batch(sims.data(), static_cast<int>(sims.size()), chunk) != CHIML_OK) {
                    fail_all(batched, finished);
                } else {
                    for (Request* r : batched) r->done_steps += chunk;
                    progressed = true;
                }
            }
            if (!finished.empty()) {
                const auto now = Clock::now();
                std::lock_guard<std::mutex> lk(mu);
                for (Request* r : finished) {
                    running.erase(std::find(running.begin(), running.end(), r));
                    r->t_done = now;
                    --admitted;
                    if (r->err.empty()) {
                        r->state = CHIML_REQ_DONE;
                        cell_updates += r->cells * r->nsteps;
                    } else {
                        r->state = CHIML_REQ_FAILED;
                        chiml_sim_destroy(r->sim);
                        r->sim = nullptr;
                        held -= r->bytes;
                        r->bytes = 0;
                    }
                }
                admit_locked();
                cv_idle.notify_all();
                progressed = true;
            }
            if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
};

extern "C" {

static int srv_fail(const char* m, int code = CHIML_E_INVALID) {
    chiml::set_sim_error(m);
    return code;
}

int chiml_server_create(size_t hbm_budget, int max_active, int chunk_steps, int graph_steps, chiml_server_t** out) {
    if (!out || hbm_budget == 0 || max_active < 1 || chunk_steps < 1 || graph_steps < 0) return srv_fail("invalid argument");
    auto s = std::make_unique<chiml_server>();
    s->budget = hbm_budget;
    s->max_active = max_active;
    s->chunk = chunk_steps;
    s->graph = graph_steps;
    chiml_server* p = s.get();
    // a few builders: host-side setup of a large 3D deck takes about a second
    const unsigned hc = std::thread::hardware_concurrency();
    const int nb = static_cast<int>(std::max(1u, std::min(4u, hc / 4)));
    for (int b = 0; b < nb; ++b) p->builders.emplace_back([p] { p->build_loop(); });
    p->stepper = std::thread([p] { p->step_loop(); });
    *out = s.release();
    return CHIML_OK;
}

int chiml_server_set_batch(chiml_server_t* s, int enable) {
    if (!s) return srv_fail("invalid argument");
    std::lock_guard<std::mutex> lk(s->mu);
    s->batch = enable != 0;
    return CHIML_OK;
}

int chiml_server_submit(chiml_server_t* s, const char* deck_json, long long* id) {
    if (!s || !deck_json || !id) return srv_fail("invalid argument");
    size_t est = 0;
    int rc = chiml_sim_estimate_bytes(deck_json, &est);
    if (rc != CHIML_OK) return rc;  // deck rejected: message already set
    auto r = std::make_unique<Request>();
    r->deck = deck_json;
    r->bytes = est;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return srv_fail("server is shutting down", CHIML_E_STATE);
    r->id = s->next_id++;
    *id = r->id;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
    s->admit_locked();
    return CHIML_OK;
}

int chiml_server_poll(chiml_server_t* s, long long id, int* state, double* admitted_s, double* finished_s) {
    if (!s || !state) return srv_fail("invalid argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return srv_fail("unknown request id");
    const Request& r = *it->second;
    *state = r.state;
    if (admitted_s) *admitted_s = r.state == CHIML_REQ_QUEUED ? -1.0 : seconds(r.t_submit, r.t_admit);
    if (finished_s) *finished_s = (r.state == CHIML_REQ_DONE || r.state == CHIML_REQ_FAILED) ? seconds(r.t_submit, r.t_done) : -1.0;
    return CHIML_OK;
}

int chiml_server_take(chiml_server_t* s, long long id, chiml_sim_t** sim) {
    if (!s || !sim) return srv_fail("invalid argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return srv_fail("unknown request id");
    Request& r = *it->second;
    if (r.state != CHIML_REQ_DONE || !r.sim) return srv_fail("request has no finished simulation to take", CHIML_E_STATE);
    *sim = r.sim;
    r.sim = nullptr;
    s->held -= r.bytes;
    r.bytes = 0;
    s->admit_locked();
    return CHIML_OK;
}

int chiml_server_error(chiml_server_t* s, long long id, char* buf, size_t cap) {
    if (!s || !buf || cap == 0) return srv_fail("invalid argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return srv_fail("unknown request id");
    std::strncpy(buf, it->second->err.c_str(), cap - 1);
    buf[cap - 1] = '\0';
    return CHIML_OK;
}

int chiml_server_drain(chiml_server_t* s) {
    if (!s) return srv_fail("invalid argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_idle.wait(lk, [&] { return s->queue.empty() && s->admitted == 0; });
    return CHIML_OK;
}

int chiml_server_stats(chiml_server_t* s, double* cell_updates, size_t* bytes_held, int* active, int* queued, int* peak_active) {
    if (!s) return srv_fail("invalid argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (cell_updates) *cell_updates = s->cell_updates;
    if (bytes_held) *bytes_held = s->held;
    if (active) *active = s->admitted;
    if (queued) *queued = static_cast<int>(s->queue.size());
    if (peak_active) *peak_active = s->peak_active;
    return CHIML_OK;
}

void chiml_server_destroy(chiml_server_t* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        // queued requests are dropped; admitted ones finish
        for (Request* r : s->queue) {
            r->state = CHIML_REQ_FAILED;
            r->err = "server destroyed before admission";
        }
        s->queue.clear();
    }
    s->cv_build.notify_all();
    for (auto& b : s->builders) b.join();
    s->stepper.join();
    for (auto& kv : s->reqs)
        if (kv.second->sim) chiml_sim_destroy(kv.second->sim);
    delete s;
}

}  // extern "C"

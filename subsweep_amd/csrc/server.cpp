// Request loop over libsubsweep_amd handles (host C++, no device code).
//
// A server owns `slots` worker threads. Each admitted request gets its own
// handle, and with it its own HIP stream and device memory, so requests that
// run at the same time overlap on the device. Requests are admitted in arrival
// order, each when a worker is free and the device-memory estimate of the
// request fits in what is left of the server's HBM budget; a request whose
// estimate exceeds the whole budget fails with SSW_ERR_BUDGET. A request runs
// create -> upload grid -> upload state -> `steps` x step -> download, so its
// grid and flux state stay resident on the device for the whole solve.
//
// Resident grids: a request with a non-zero grid_key leaves its handle in an
// idle cache keyed by (grid_key, parameters, transport mode, replay block,
// positions given); a later request with the same key takes it and only
// uploads its state, skipping the grid upload, the Kahn layering and the
// schedule and cone build. Cached handles hold their device memory against
// the budget; when a request does not fit, idle handles are freed least
// recently used first. Other requests free their handle when they complete.
//
// Packed stepping (ssw_server_create_packed): the workers only set requests
// up and finish them; stepper threads step the requests that are set up, each
// taking up to max_batch of them per iteration and stepping them once with
// ssw_step_batch, so their grouped launches are packed into shared launches.
// A request joins at the next iteration after its set-up and leaves after its
// last step (continuous batching at step granularity). There are
// slots / max_batch stepper threads (1 to 4): the device runs launches of up
// to four streams at once, and a packed launch lasts as long as its slowest
// problem, so a few small packs side by side beat one large one.
//
// The reference has no counterpart (it runs one simulation per process); the
// serving loop is how this library is driven with many independent solves.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/subsweep_amd.h"

namespace {

struct Req {
    ssw_request r;
    std::string key;          // cache key; empty for grid_key 0
    std::vector<double> dirs;  // explicit directions, copied at submission
    ssw_sweep* cached = nullptr;  // idle handle taken from the cache at admission
    bool reused = false;
    int64_t ticket = 0;
    int64_t estimate = 0;
    int32_t mode = SSW_TRANSPORT_MODE_GEOMETRIC;
    double t_submit = 0, t_start = 0, t_end = 0;
    int32_t status = SSW_OK;
    std::string error;
    int64_t updates = 0;
    int64_t device_bytes = 0;
    bool done = false;
    // packed stepping: the handle while set up, steps taken, and whether the
    // stepper has finished with it
    ssw_sweep* handle = nullptr;
    int32_t steps_done = 0;
    bool stepped = false;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// An idle handle in the cache and the device memory it holds.
struct Idle {
    ssw_sweep* s = nullptr;
    int64_t bytes = 0;
    uint64_t last_use = 0;
};

template <typename T>
void put(std::string& k, const T& v) {
    k.append(reinterpret_cast<const char*>(&v), sizeof(v));
}

// Everything a handle's device state was built from: the grid (by the
// caller's key), the parameters (explicit directions by value), the
// transport mode, the replay block and whether cell positions were given.
std::string cache_key(const ssw_request& r, int32_t mode) {
    if (r.grid_key == 0) return std::string();
    std::string k;
    put(k, r.grid_key);
    put(k, r.num_cells);
    const ssw_sweep_parameters& p = r.params;
    put(k, p.num_directions);
    if (p.explicit_directions) k.append(reinterpret_cast<const char*>(p.explicit_directions),
                                        sizeof(double) * 3 * (size_t)std::max(0, p.num_directions));
    put(k, p.num_timestep_levels);
    put(k, p.periodic);
    put(k, p.max_timestep);
    put(k, p.significant_rate_threshold);
    put(k, p.timestep_safety_factor);
    put(k, p.chemistry_timestep_safety_factor);
    put(k, p.prevent_cooling);
    put(k, p.transport_precision);
    put(k, p.scale_factor);
    put(k, mode);
    put(k, r.replay_block);
    const int32_t pos = r.cell_position ? 1 : 0;
    put(k, pos);
    return k;
}

}  // namespace

struct ssw_server {
    int slots = 1;
    int64_t budget = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
    int64_t next_ticket = 1;
    int64_t reserved = 0;
    int64_t peak_reserved = 0;
    int running = 0;
    int peak_running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    // packed stepping (max_batch >= 2): requests set up and waiting for steps
    int max_batch = 0;
    std::deque<Req*> ready;
    std::condition_variable cv_ready, cv_stepped;
    std::vector<std::thread> steppers;
    bool stop_stepper = false;
    int64_t batches = 0, batched_steps = 0;
    // idle cached handles by key, and the device memory they hold
    std::unordered_map<std::string, std::vector<Idle>> idle;
    int64_t cached_bytes = 0;
    uint64_t use_clock = 0;
};

namespace {

// Creates or reuses the request's handle and uploads grid and state; returns
// the handle, or nullptr with the request's status set.
ssw_sweep* setup_request(Req* q) {
    ssw_sweep* s = q->cached;
    const ssw_request& r = q->r;
    int rc = SSW_OK;
    if (!s) {
        rc = ssw_create(&r.params, &s);
        if (rc != SSW_OK) {
            q->status = rc;
            q->error = ssw_last_error();
            return nullptr;
        }
        rc = ssw_set_transport_mode(s, q->mode);
        if (rc == SSW_OK) rc = ssw_set_replay_block(s, r.replay_block);
        if (rc == SSW_OK) rc = ssw_set_direction_sweeps(s, r.direction_sweeps ? 1 : 0);
        if (rc == SSW_OK && r.cell_position) rc = ssw_set_cell_positions(s, r.num_cells, r.cell_position);
        if (rc == SSW_OK)
            rc = ssw_upload_grid(s, r.num_cells, r.face_offsets, r.face_neighbour, r.face_kind, r.face_area,
                                 r.face_normal, r.cell_volume, r.cell_size);
    }
    if (rc == SSW_OK) rc = ssw_upload_state(s, r.density, r.ionized_fraction, r.temperature, r.source);
    if (rc != SSW_OK) {
        q->status = rc;
        q->error = ssw_last_error();
        ssw_destroy(s);
        return nullptr;
    }
    return s;
}

// Downloads the results of a stepped request; returns its handle if it should
// go to the cache (else the handle has been freed).
ssw_sweep* finish_request(Req* q, ssw_sweep* s) {
    const ssw_request& r = q->r;
    int rc = q->status;
    ssw_info info;
    if (rc == SSW_OK)
        rc = ssw_download(s, r.out_ionized_fraction, r.out_temperature, r.out_timestep, r.out_photon_rate, r.out_level,
                          r.out_change_timescale);
    if (rc == SSW_OK) rc = ssw_get_info(s, &info);
    if (rc == SSW_OK) q->device_bytes = info.peak_device_bytes;
    if (rc != SSW_OK && q->status == SSW_OK) {
        q->status = rc;
        q->error = ssw_last_error();
    }
    if (rc == SSW_OK && !q->key.empty()) return s;
    ssw_destroy(s);
    return nullptr;
}

// Runs the request on its worker; returns its handle if it should go to the
// cache (else the handle has been freed).
ssw_sweep* run_request(ssw_server* sv, Req* q) {
    ssw_sweep* s = setup_request(q);
    if (!s) return nullptr;
    if (sv->max_batch >= 2 && q->r.steps > 0) {  // the steppers step it, packed with others
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            q->handle = s;
            sv->ready.push_back(q);
        }
        sv->cv_ready.notify_all();
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_stepped.wait(lk, [&] { return q->stepped; });
    } else {
        ssw_info info;
        int rc = SSW_OK;
        for (int32_t k = 0; rc == SSW_OK && k < q->r.steps; k++) {
            double el = 0.0;
            rc = ssw_step(s, &el);
            if (rc == SSW_OK) rc = ssw_get_info(s, &info);
            if (rc == SSW_OK) q->updates += info.updates_last_step;
        }
        if (rc != SSW_OK) {
            q->status = rc;
            q->error = ssw_last_error();
        }
    }
    return finish_request(q, s);
}

// Packed stepping: one step of up to max_batch set-up requests per
// iteration; requests leave after their last step (or a failed one). A
// request is out of the ready queue while a stepper steps it.
void stepper(ssw_server* sv) {
    std::vector<Req*> batch;
    std::vector<ssw_sweep*> hs;
    std::vector<double> el;
    for (;;) {
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_ready.wait(lk, [&] { return sv->stop_stepper || !sv->ready.empty(); });
            if (sv->ready.empty()) return;
            while (!sv->ready.empty() && (int)batch.size() < sv->max_batch) {
                batch.push_back(sv->ready.front());
                sv->ready.pop_front();
            }
        }
        hs.clear();
        for (Req* q : batch) hs.push_back(q->handle);
        el.assign(hs.size(), 0.0);
        const int rc = ssw_step_batch(hs.data(), (int32_t)hs.size(), el.data());
        const std::string err = rc != SSW_OK ? ssw_last_error() : std::string();
        for (Req* q : batch) {
            ssw_info info;
            if (rc == SSW_OK && ssw_get_info(q->handle, &info) == SSW_OK) q->updates += info.updates_last_step;
            q->steps_done++;
        }
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->batches++;
            sv->batched_steps += (int64_t)hs.size();
            for (Req* q : batch) {
                if (rc != SSW_OK) {
                    q->status = rc;
                    q->error = err;
                }
                if (rc != SSW_OK || q->steps_done >= q->r.steps) q->stepped = true;
                else sv->ready.push_back(q);  // to the back: more than max_batch requests take turns
            }
        }
        sv->cv_stepped.notify_all();
        sv->cv_ready.notify_all();
    }
}

// Takes idle handles out of the cache, least recently used first, until
// `need` more bytes fit in the budget or the cache is empty (the caller holds
// sv->mu and frees the returned handles after releasing it).
void evict_for(ssw_server* sv, int64_t need, std::vector<ssw_sweep*>& out) {
    if (sv->reserved + need > sv->budget) return;  // running requests must finish first
    while (sv->reserved + sv->cached_bytes + need > sv->budget && sv->cached_bytes > 0) {
        auto best = sv->idle.end();
        size_t bi = 0;
        uint64_t oldest = UINT64_MAX;
        for (auto it = sv->idle.begin(); it != sv->idle.end(); ++it)
            for (size_t i = 0; i < it->second.size(); i++)
                if (it->second[i].last_use < oldest) {
                    oldest = it->second[i].last_use;
                    best = it;
                    bi = i;
                }
        if (best == sv->idle.end()) break;
        Idle e = best->second[bi];
        best->second.erase(best->second.begin() + (std::ptrdiff_t)bi);
        if (best->second.empty()) sv->idle.erase(best);
        sv->cached_bytes -= e.bytes;
        out.push_back(e.s);
    }
}

void worker(ssw_server* sv) {
    std::vector<ssw_sweep*> evicted;
    for (;;) {
        Req* q = nullptr;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            for (;;) {
                if (sv->stop && sv->queue.empty()) return;
                if (!sv->queue.empty()) {
                    Req* head = sv->queue.front();
                    if (head->estimate > sv->budget) {  // can never fit
                        sv->queue.pop_front();
                        head->status = SSW_ERR_BUDGET;
                        head->error = "request needs an estimated " + std::to_string(head->estimate) +
                                      " bytes of device memory, above the budget of " + std::to_string(sv->budget);
                        head->t_start = head->t_end = now_s(sv->t0);
                        head->done = true;
                        sv->cv_done.notify_all();
                        continue;
                    }
                    // a cached handle of the same key: its memory moves from
                    // the cache into the request's reservation (and back if
                    // the request cannot start yet)
                    Idle hit;
                    if (!head->key.empty()) {
                        auto it = sv->idle.find(head->key);
                        if (it != sv->idle.end() && !it->second.empty()) {
                            hit = it->second.back();
                            sv->cached_bytes -= hit.bytes;
                            it->second.pop_back();
                            if (it->second.empty()) sv->idle.erase(it);
                        }
                    }
                    evict_for(sv, head->estimate, evicted);
                    const bool fits = sv->reserved + sv->cached_bytes + head->estimate <= sv->budget;
                    if (hit.s && !fits) {
                        sv->idle[head->key].push_back(hit);
                        sv->cached_bytes += hit.bytes;
                    } else if (hit.s) {
                        head->cached = hit.s;
                        head->reused = true;
                    }
                    if (fits) {
                        sv->queue.pop_front();
                        q = head;
                        sv->reserved += q->estimate;
                        sv->running++;
                        sv->peak_reserved = std::max(sv->peak_reserved, sv->reserved + sv->cached_bytes);
                        sv->peak_running = std::max(sv->peak_running, sv->running);
                        q->t_start = now_s(sv->t0);
                        break;
                    }
                }
                if (!evicted.empty()) break;  // free them, then look again
                sv->cv_work.wait(lk);
            }
        }
        for (ssw_sweep* e : evicted) ssw_destroy(e);
        if (!q) {
            evicted.clear();
            continue;
        }
        evicted.clear();
        ssw_sweep* keep = run_request(sv, q);
        int64_t held = 0;
        if (keep) {
            ssw_info info;
            if (ssw_get_info(keep, &info) == SSW_OK) held = info.device_bytes;
            else {
                ssw_destroy(keep);
                keep = nullptr;
            }
        }
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            q->t_end = now_s(sv->t0);
            q->done = true;
            sv->reserved -= q->estimate;
            sv->running--;
            if (keep) {
                sv->idle[q->key].push_back(Idle{keep, held, ++sv->use_clock});
                sv->cached_bytes += held;
            }
        }
        sv->cv_done.notify_all();
        sv->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

int ssw_server_create_packed(int32_t slots, int64_t hbm_budget_bytes, int32_t max_batch, ssw_server** out) {
    if (!out || slots < 1 || slots > 64 || hbm_budget_bytes <= 0 || max_batch < 0 || max_batch > 64)
        return SSW_ERR_INVALID;
    ssw_server* sv = new ssw_server();
    sv->slots = slots;
    sv->budget = hbm_budget_bytes;
    sv->max_batch = max_batch >= 2 ? max_batch : 0;
    sv->t0 = std::chrono::steady_clock::now();
    if (sv->max_batch) {
        const int n = std::max(1, std::min(4, slots / sv->max_batch));
        for (int i = 0; i < n; i++) sv->steppers.emplace_back(stepper, sv);
    }
    for (int i = 0; i < slots; i++) sv->workers.emplace_back(worker, sv);
    *out = sv;
    return SSW_OK;
}

int ssw_server_create(int32_t slots, int64_t hbm_budget_bytes, ssw_server** out) {
    return ssw_server_create_packed(slots, hbm_budget_bytes, 0, out);
}

int ssw_server_pack_stats(ssw_server* sv, int64_t* batches, int64_t* batched_steps) {
    if (!sv) return SSW_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (batches) *batches = sv->batches;
    if (batched_steps) *batched_steps = sv->batched_steps;
    return SSW_OK;
}

int ssw_server_submit(ssw_server* sv, const ssw_request* r, int64_t* ticket) {
    if (!sv || !r || !ticket || r->num_cells <= 0 || r->steps < 0) return SSW_ERR_INVALID;
    auto q = std::make_unique<Req>();
    q->r = *r;
    const int64_t F = r->face_offsets[r->num_cells];
    int64_t nper = 0;  // cells with a periodic face (periodic sweeps only)
    if (r->params.periodic)
        for (int64_t c = 0; c < r->num_cells; c++)
            for (int64_t f = r->face_offsets[c]; f < r->face_offsets[c + 1]; f++)
                if (r->face_kind[f] == SSW_FACE_PERIODIC) {
                    nper++;
                    break;
                }
    q->mode = r->transport_mode;
    if (q->mode == SSW_TRANSPORT_MODE_AUTO) {
        // compiled when it fits the budget shared by the slots, else geometric
        const int64_t comp = ssw_estimate_device_bytes(&r->params, r->num_cells, F, nper, SSW_TRANSPORT_MODE_COMPILED);
        q->mode = comp <= sv->budget / sv->slots ? SSW_TRANSPORT_MODE_COMPILED : SSW_TRANSPORT_MODE_GEOMETRIC;
    }
    q->estimate = ssw_estimate_device_bytes(&r->params, r->num_cells, F, nper, q->mode);
    q->key = cache_key(*r, q->mode);
    q->r.params.explicit_directions = nullptr;  // read once, into the key; handles copy them at creation
    if (r->params.explicit_directions) {
        q->dirs.assign(r->params.explicit_directions, r->params.explicit_directions + 3 * r->params.num_directions);
        q->r.params.explicit_directions = q->dirs.data();
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) return SSW_ERR_INVALID;
        q->ticket = sv->next_ticket++;
        q->t_submit = now_s(sv->t0);
        *ticket = q->ticket;
        sv->queue.push_back(q.get());
        sv->reqs[q->ticket] = std::move(q);
    }
    sv->cv_work.notify_all();
    return SSW_OK;
}

int ssw_server_wait(ssw_server* sv, int64_t ticket, ssw_request_result* res) {
    if (!sv || !res) return SSW_ERR_INVALID;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return SSW_ERR_INVALID;
    Req* q = it->second.get();
    sv->cv_done.wait(lk, [&] { return q->done; });
    std::memset(res, 0, sizeof(*res));
    res->status = q->status;
    res->transport_mode = q->mode;
    res->submit_s = q->t_submit;
    res->start_s = q->t_start;
    res->end_s = q->t_end;
    res->updates = q->updates;
    res->device_bytes = q->device_bytes;
    res->estimated_bytes = q->estimate;
    res->reused = q->reused ? 1 : 0;
    std::strncpy(res->error, q->error.c_str(), sizeof(res->error) - 1);
    sv->reqs.erase(it);
    return SSW_OK;
}

int ssw_server_stats(ssw_server* sv, int64_t* peak_reserved_bytes, int32_t* peak_running) {
    if (!sv) return SSW_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (peak_reserved_bytes) *peak_reserved_bytes = sv->peak_reserved;
    if (peak_running) *peak_running = sv->peak_running;
    return SSW_OK;
}

void ssw_server_destroy(ssw_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    for (std::thread& t : sv->workers) t.join();
    if (!sv->steppers.empty()) {
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->stop_stepper = true;
        }
        sv->cv_ready.notify_all();
        for (std::thread& t : sv->steppers) t.join();
    }
    for (auto& kv : sv->idle)
        for (Idle& e : kv.second) ssw_destroy(e.s);
    delete sv;
}

}  // extern "C"

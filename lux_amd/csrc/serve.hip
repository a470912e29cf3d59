// serve.hip -- the continuous-batching render server (host runtime; the GPU
// work is the wavefront engine's batches and the scenes' vertex refits).
#include "serve.hpp"

#include <algorithm>
#include <thread>

namespace lux {
namespace pt {

Server::Server(const ServeConfig& cfg, const Tuning& tuning, hipStream_t stream)
    : cfg_(cfg), stream_(stream), t0_(std::chrono::steady_clock::now()) {
    if (cfg.slice_spp < 0 || cfg.max_inflight < 1 || cfg.max_inflight > 64)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "server config out of range");
    engine_.set_tuning(tuning);
    hip_check(hipGetDevice(&device_), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&update_stream_, hipStreamNonBlocking), "hipStreamCreate");
    engine_.reset_ray_counts(stream_);
    worker_ = std::thread([this] { worker_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();  // finishes the queued tasks first
    for (auto& t : tasks_) {
        if (t->restore) scene_release(t->fresh);  // never adopted
        scene_release(t->dead);
    }
    (void)hipStreamSynchronize(stream_);
    for (Job* j : active_) --j->scene->server_jobs;  // admitted, never retired
    for (Job* j : waiting_) --j->scene->server_jobs;  // admitted, scene work never finished
    for (auto& j : pending_) --j->scene->server_jobs;  // submitted, never admitted
    for (auto& f : inflight_) (void)hipEventDestroy(f.ev);
    for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
    for (auto& kv : jobs_)
        if (kv.second->film) (void)hipFree(kv.second->film);
    for (auto& kv : film_pool_) (void)hipFree(kv.second);
    if (update_stream_) (void)hipStreamDestroy(update_stream_);
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

uint64_t Server::submit(SceneDev* scene, const lux_render_settings& rs, const lux_vec3* vertices, int vertex_count,
                        double arrival) {
    if (!scene || (!scene->soup && !scene->source)) throw Error(LUX_ERR_INVALID_ARGUMENT, "request needs a scene");
    if (rs.mode != LUX_RENDER_WAVEFRONT) throw Error(LUX_ERR_INVALID_ARGUMENT, "the server renders wavefront only");
    if (vertices && scene->instanced)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "instanced scene: vertex updates do not apply");
    if (vertices && vertex_count != scene->vertex_count)
        throw Error(LUX_ERR_MISMATCH, "vertex update must keep the vertex count");
    auto job = std::make_unique<Job>();
    job->id = next_id_++;
    job->scene = scene;
    job->rs = rs;
    job->arrival = arrival;
    if (vertices) {
        job->update = true;
        job->vertices.assign(vertices, vertices + vertex_count);
    }
    job->film_bytes = sizeof(float) * 3 * size_t(rs.width) * size_t(rs.height);
    uint64_t id = job->id;
    ++scene->server_jobs;  // from submission to completion: the scene must stay as it is
    pending_.push_back(std::move(job));
    ++stats_.submitted;
    return id;
}

float* Server::film_alloc(size_t bytes) {
    auto it = film_pool_.find(bytes);
    if (it != film_pool_.end()) {
        float* p = it->second;
        film_pool_.erase(it);
        films_pooled_ -= bytes;
        films_in_use_ += bytes;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw Error(LUX_ERR_OUT_OF_MEMORY, std::string("film hipMalloc: ") + hipGetErrorString(e));
    films_in_use_ += bytes;
    return static_cast<float*>(p);
}

void Server::film_free(float* p, size_t bytes) {
    films_in_use_ -= bytes;
    film_pool_.emplace(bytes, p);
    films_pooled_ += bytes;
}

size_t Server::footprint_in_use() const {
    return resident_scene_bytes_ + freeing_bytes_ + films_in_use_ + films_pooled_ + engine_.planned_bytes();
}

void Server::post(const std::shared_ptr<Task>& t) {
    tasks_.push_back(t);
    {
        std::lock_guard<std::mutex> lk(mu_);
        queue_.push_back(t);
    }
    cv_.notify_one();
}

void Server::worker_loop() {
    (void)hipSetDevice(device_);
    for (;;) {
        std::shared_ptr<Task> t;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return;  // stop_ and nothing left
            t = queue_.front();
            queue_.pop_front();
        }
        const auto t0 = std::chrono::steady_clock::now();
        try {
            if (t->restore) {
                scene_restore(t->fresh, *t->source, update_stream_);
                if (!t->vertices.empty()) {
                    scene_update_vertices(t->fresh, t->vertices.data(), int(t->vertices.size()), update_stream_);
                    t->source->current = t->vertices;
                }
            } else if (!t->vertices.empty()) {
                // no admitted request renders the scene: no queued batch reads it
                scene_update_vertices(*t->scene, t->vertices.data(), int(t->vertices.size()), update_stream_);
            }
            scene_release(t->dead);  // hipFree waits for the device: off the scheduling thread
        } catch (...) {
            t->error = std::current_exception();
        }
        t->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        t->done.store(true, std::memory_order_release);
    }
}

// Adopts finished uploads, releases the budget of finished frees and lets the
// requests that waited on a scene's work into the batches.
void Server::finish_tasks() {
    for (auto it = tasks_.begin(); it != tasks_.end();) {
        Task* t = it->get();
        if (!t->done.load(std::memory_order_acquire)) {
            ++it;
            continue;
        }
        if (t->error) {
            std::exception_ptr e = t->error;
            if (t->scene) preparing_.erase(t->scene);
            tasks_.erase(it);
            std::rethrow_exception(e);
        }
        freeing_bytes_ -= t->dead_bytes;
        if (t->scene) {
            if (t->restore) scene_adopt(*t->scene, std::move(t->fresh));
            preparing_.erase(t->scene);
            stats_.prepare_seconds += t->seconds;
            for (auto w = waiting_.begin(); w != waiting_.end();) {
                if ((*w)->scene == t->scene) {
                    activate(*w);
                    w = waiting_.erase(w);
                } else {
                    ++w;
                }
            }
        }
        it = tasks_.erase(it);
    }
}

// Drops idle resident scenes (kept description, no admitted request, no work
// running), least recently used first, until `need` more bytes fit.
void Server::evict_for(size_t need, const SceneDev* keep) {
    std::vector<std::pair<uint64_t, SceneDev*>> idle;
    for (auto& kv : resident_) {
        SceneDev* sc = kv.first;
        if (sc == keep || !sc->source || sc->instanced || !sc->resident || scene_users_[sc] > 0 ||
            preparing_.count(sc))
            continue;
        idle.emplace_back(last_use_[sc], sc);
    }
    std::sort(idle.begin(), idle.end());
    for (auto& c : idle) {
        if (footprint_in_use() + need <= cfg_.hbm_budget) break;
        SceneDev* sc = c.second;
        const size_t bytes = resident_[sc];
        auto t = std::make_shared<Task>();
        t->dead = scene_detach(*sc);
        t->dead_bytes = bytes;
        resident_.erase(sc);
        resident_scene_bytes_ -= bytes;
        freeing_bytes_ += bytes;
        ++stats_.evictions;
        post(t);
    }
}

void Server::activate(Job* j) {
    j->rq.sv = scene_view(*j->scene);
    j->rq.rs = j->rs;
    j->rq.film = j->film;
    j->rq.total = request_paths(j->rs);
    j->rq.issued = 0;
    j->rq.last_batch = -1;
    j->rq.tag = j->id;
    active_.push_back(j);
    issue_.push_back(j);
}

void Server::admit() {
    finish_tasks();
    std::vector<const SceneDev*> blocked;  // scenes with an earlier request still waiting
    auto is_blocked = [&](const SceneDev* s) { return std::find(blocked.begin(), blocked.end(), s) != blocked.end(); };
    for (auto it = pending_.begin(); it != pending_.end();) {
        Job* j = it->get();
        SceneDev* sc = j->scene;
        if (is_blocked(sc) || preparing_.count(sc)) {  // arrival order within a scene
            blocked.push_back(sc);
            ++it;
            continue;
        }
        if (j->update && scene_users_[sc] > 0) {  // wait until the scene is idle
            blocked.push_back(sc);
            ++it;
            continue;
        }
        const bool cold = !sc->resident;
        const size_t scene_extra = (cold || !resident_.count(sc)) ? sc->device_bytes : 0;
        const size_t extra = j->film_bytes + scene_extra;
        if (footprint_in_use() + extra > cfg_.hbm_budget) {
            // give pooled films back, then drop idle scenes, before refusing
            while (!film_pool_.empty() && footprint_in_use() + extra > cfg_.hbm_budget) {
                auto f = film_pool_.begin();
                (void)hipFree(f->second);
                films_pooled_ -= f->first;
                film_pool_.erase(f);
            }
            if (footprint_in_use() + extra > cfg_.hbm_budget) evict_for(extra, sc);
            if (footprint_in_use() + extra > cfg_.hbm_budget) break;  // arrival order: later ones wait too
        }
        if (!resident_.count(sc)) {
            resident_[sc] = sc->device_bytes;
            resident_scene_bytes_ += sc->device_bytes;
        }
        j->film = film_alloc(j->film_bytes);
        j->admitted = now();
        ++scene_users_[sc];
        last_use_[sc] = ++tick_;
        if (cold || j->update) {
            auto t = std::make_shared<Task>();
            t->scene = sc;
            t->restore = cold;
            if (cold) {
                t->source = sc->source;
                ++stats_.uploads;
            }
            if (j->update) {
                t->vertices = std::move(j->vertices);
                j->vertices = std::vector<lux_vec3>();
                ++stats_.refits;
            }
            preparing_[sc] = t.get();
            waiting_.push_back(j);
            post(t);
        } else {
            activate(j);
        }
        jobs_[j->id] = std::move(*it);
        it = pending_.erase(it);
        ++stats_.admitted;
    }
    stats_.resident_bytes = footprint_in_use();
    stats_.peak_bytes = std::max(stats_.peak_bytes, stats_.resident_bytes);
}

bool Server::enqueue_one() {
    if (issue_.empty() || int(inflight_.size()) >= cfg_.max_inflight) return false;
    std::vector<Request*> order;
    std::vector<uint64_t> before;
    int depth = 0;
    order.reserve(issue_.size());
    before.reserve(issue_.size());
    for (Job* j : issue_) {
        order.push_back(&j->rq);
        before.push_back(j->rq.issued);
        depth = std::max(depth, j->rs.max_depth);
    }
    int64_t batch = -1;
    uint64_t placed = engine_.enqueue_batch(order, cfg_.slice_spp, depth, stream_, &batch);
    if (!placed) return false;
    hipEvent_t ev;
    if (!event_pool_.empty()) {
        ev = event_pool_.back();
        event_pool_.pop_back();
    } else {
        hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    }
    hip_check(hipEventRecord(ev, stream_), "hipEventRecord");
    inflight_.push_back(Inflight{batch, ev});
    ++stats_.batches;
    stats_.paths += placed;
    // round robin: requests that got paths this batch and still have some go
    // to the back; the ones that received nothing keep their place at the front
    std::deque<Job*> keep, served;
    size_t k = 0;
    uint64_t in_batch = 0;
    for (Job* j : issue_) {
        const uint64_t was = before[k++];
        if (j->rq.issued > was) ++in_batch;
        if (j->rq.issued >= j->rq.total) continue;  // fully placed; completes with its last batch
        (j->rq.issued > was ? served : keep).push_back(j);
    }
    if (in_batch >= 2) ++stats_.shared_batches;
    stats_.max_batch_requests = std::max(stats_.max_batch_requests, in_batch);
    issue_.clear();
    for (Job* j : keep) issue_.push_back(j);
    for (Job* j : served) issue_.push_back(j);
    return true;
}

void Server::retire(bool block, std::vector<uint64_t>* done) {
    while (!inflight_.empty()) {
        Inflight f = inflight_.front();
        hipError_t q = block ? hipEventSynchronize(f.ev) : hipEventQuery(f.ev);
        if (q == hipErrorNotReady) break;
        hip_check(q, "batch event");
        block = false;  // waited for one; the rest only if already done
        inflight_.pop_front();
        event_pool_.push_back(f.ev);
        double t = now();
        for (auto it = active_.begin(); it != active_.end();) {
            Job* j = *it;
            if (j->rq.last_batch >= 0 && j->rq.last_batch <= f.batch) {
                j->done = true;
                j->finish = t;
                --scene_users_[j->scene];
                last_use_[j->scene] = ++tick_;
                --j->scene->server_jobs;
                ++stats_.completed;
                if (done) done->push_back(j->id);
                it = active_.erase(it);
            } else {
                ++it;
            }
        }
    }
}

void Server::progress(bool block, std::vector<uint64_t>* done) {
    retire(false, done);
    admit();  // also takes in finished scene work
    bool enqueued = false;
    while (enqueue_one()) enqueued = true;
    if (block && !enqueued && !inflight_.empty()) {
        retire(true, done);
        admit();
        while (enqueue_one()) {
        }
    }
}

void Server::take_result(uint64_t id, float* host_film, double* arrival, double* admitted, double* finish) {
    auto it = jobs_.find(id);
    if (it == jobs_.end() || !it->second->done) throw Error(LUX_ERR_INVALID_ARGUMENT, "request is not complete");
    Job* j = it->second.get();
    if (host_film)
        hip_check(hipMemcpy(host_film, j->film, j->film_bytes, hipMemcpyDeviceToHost), "D2H film");
    if (arrival) *arrival = j->arrival;
    if (admitted) *admitted = j->admitted;
    if (finish) *finish = j->finish;
    film_free(j->film, j->film_bytes);
    jobs_.erase(it);
}

void run_trace(Server& sv, int n, SceneDev* const* scenes, const lux_render_settings* settings, const double* arrival,
               const lux_vec3* const* vertices, const int* vertex_counts, float* const* films, double* latency,
               double* wall) {
    for (int i = 1; i < n; ++i)
        if (arrival[i] < arrival[i - 1]) throw Error(LUX_ERR_INVALID_ARGUMENT, "arrival times must be sorted");
    std::unordered_map<uint64_t, int> index_of;
    std::vector<uint64_t> done;
    const double t0 = sv.now();
    int next = 0, completed = 0;
    while (completed < n) {
        const double t = sv.now() - t0;
        while (next < n && arrival[next] <= t) {
            const lux_vec3* v = vertices ? vertices[next] : nullptr;
            uint64_t id = sv.submit(scenes[next], settings[next], v, v ? vertex_counts[next] : 0, t0 + arrival[next]);
            index_of[id] = next++;
        }
        done.clear();
        sv.progress(sv.has_inflight(), &done);
        for (uint64_t id : done) {
            const int i = index_of.at(id);
            double a = 0, ad = 0, f = 0;
            sv.take_result(id, films ? films[i] : nullptr, &a, &ad, &f);
            latency[i] = f - a;
            ++completed;
        }
        if (!sv.has_inflight() && completed < n) {
            if (next < n) {
                const double wait = arrival[next] - (sv.now() - t0);
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 2e-4)));
            } else if (sv.preparing()) {
                std::this_thread::sleep_for(std::chrono::duration<double>(5e-5));  // scene work on the worker
            } else if (sv.busy()) {
                // nothing on the GPU, no arrival left, and the queue cannot be admitted
                std::vector<uint64_t> again;
                sv.progress(false, &again);
                if (!sv.has_inflight() && again.empty())
                    throw Error(LUX_ERR_OUT_OF_MEMORY, "a request does not fit the HBM budget on its own");
                for (uint64_t id : again) {
                    const int i = index_of.at(id);
                    double a = 0, ad = 0, f = 0;
                    sv.take_result(id, films ? films[i] : nullptr, &a, &ad, &f);
                    latency[i] = f - a;
                    ++completed;
                }
            }
        }
    }
    if (wall) *wall = sv.now() - t0;
}

}  // namespace pt
}  // namespace lux

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
    hip_check(hipStreamCreateWithFlags(&update_stream_, hipStreamNonBlocking), "hipStreamCreate");
    engine_.reset_ray_counts(stream_);
}

Server::~Server() {
    (void)hipStreamSynchronize(stream_);
    for (Job* j : active_) --j->scene->server_jobs;  // admitted, never retired
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
    if (!scene || !scene->soup) throw Error(LUX_ERR_INVALID_ARGUMENT, "request needs a scene");
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
    return resident_scene_bytes_ + films_in_use_ + films_pooled_ + engine_.planned_bytes();
}

void Server::admit() {
    std::vector<const SceneDev*> blocked;  // scenes with an earlier request still waiting
    auto is_blocked = [&](const SceneDev* s) { return std::find(blocked.begin(), blocked.end(), s) != blocked.end(); };
    for (auto it = pending_.begin(); it != pending_.end();) {
        Job* j = it->get();
        if (is_blocked(j->scene)) {
            ++it;
            continue;
        }
        if (j->update && scene_users_[j->scene] > 0) {  // wait until the scene is idle
            blocked.push_back(j->scene);
            ++it;
            continue;
        }
        size_t extra = j->film_bytes + (resident_.count(j->scene) ? 0 : j->scene->device_bytes);
        if (footprint_in_use() + extra > cfg_.hbm_budget) {
            // give pooled films back before refusing
            while (!film_pool_.empty() && footprint_in_use() + extra > cfg_.hbm_budget) {
                auto f = film_pool_.begin();
                (void)hipFree(f->second);
                films_pooled_ -= f->first;
                film_pool_.erase(f);
            }
            if (footprint_in_use() + extra > cfg_.hbm_budget) break;  // arrival order: later ones wait too
        }
        if (!resident_.count(j->scene)) {
            resident_[j->scene] = j->scene->device_bytes;
            resident_scene_bytes_ += j->scene->device_bytes;
        }
        if (j->update) {
            // the scene has no unfinished request, so no queued batch reads it
            scene_update_vertices(*j->scene, j->vertices.data(), int(j->vertices.size()), update_stream_);
            j->vertices.clear();
            j->vertices.shrink_to_fit();
            ++stats_.refits;
        }
        j->film = film_alloc(j->film_bytes);
        j->rq.sv = scene_view(*j->scene);
        j->rq.rs = j->rs;
        j->rq.film = j->film;
        j->rq.total = request_paths(j->rs);
        j->rq.issued = 0;
        j->rq.last_batch = -1;
        j->rq.tag = j->id;
        j->admitted = now();
        ++scene_users_[j->scene];
        active_.push_back(j);
        issue_.push_back(j);
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
    admit();
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

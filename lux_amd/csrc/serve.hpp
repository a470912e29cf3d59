// serve.hpp -- continuous-batching render server over the wavefront engine.
//
// Requests (scene, render settings, optional new vertex positions) arrive at
// any time. They are admitted in arrival order while their HBM footprint fits
// the budget (films come from a pool), and every admitted request's paths join
// the shared wavefront loop in slices of `slice_spp` samples, round-robin, so
// requests that arrive while others are on the GPU start within a batch or two.
//
// The server owns scene residency (PreparedScene::prepare_gpu,
// src/scene/prepared_scene.cu:61-79, on demand): a request for a scene whose
// device state is absent (created deferred, or dropped earlier) admits with
// the scene's bytes counted and has it uploaded and built; when the budget
// binds, idle scenes (no admitted request) are dropped, least recently used
// first, and rebuilt from their kept description when asked for again. A
// request that moves a scene's vertices waits until no admitted request still
// renders that scene, then the scene is refit. Uploads, refits and the frees
// of dropped scenes run on a worker thread and its own stream, never on the
// scheduling thread; the request's paths enter batches once that work is
// done. A request completes when the batch holding its last path has
// finished; its film is then exactly the film of the same render run alone.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pt_runtime.hpp"

namespace lux {
namespace pt {

struct ServeConfig {
    uint64_t hbm_budget = 200ull << 30;
    int slice_spp = 16;
    int max_inflight = 2;  // batches enqueued ahead of the GPU
};

struct ServeStats {
    uint64_t submitted = 0, admitted = 0, completed = 0, refits = 0;
    uint64_t batches = 0, paths = 0;
    uint64_t resident_bytes = 0, peak_bytes = 0;
    uint64_t shared_batches = 0;      // batches holding paths of two or more requests
    uint64_t max_batch_requests = 0;  // most requests with paths in one batch
    uint64_t uploads = 0, evictions = 0;  // scenes made resident / dropped by the server
    double prepare_seconds = 0;           // worker time of uploads and refits
};

class Server {
public:
    Server(const ServeConfig& cfg, const Tuning& tuning, hipStream_t stream);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Seconds on the server's clock (steady, since creation).
    double now() const;

    // Queues a request; `vertices` (nullable) are copied. Returns its id.
    uint64_t submit(SceneDev* scene, const lux_render_settings& rs, const lux_vec3* vertices, int vertex_count,
                    double arrival);

    // One scheduling round: retire finished batches, admit, enqueue batches up
    // to max_inflight. With `block`, waits for the oldest in-flight batch if no
    // new batch could be enqueued. Appends newly completed ids to `done`.
    void progress(bool block, std::vector<uint64_t>* done);

    bool busy() const { return !pending_.empty() || !active_.empty() || !waiting_.empty(); }
    bool has_inflight() const { return !inflight_.empty(); }
    // Uploads, refits or frees still running on the worker.
    bool preparing() const { return !tasks_.empty(); }

    // Completed request: film sums to `host_film` (nullable), times, film freed.
    void take_result(uint64_t id, float* host_film, double* arrival, double* admitted, double* finish);

    ServeStats stats() const { return stats_; }
    void ray_counts(uint64_t* extension, uint64_t* shadow) { engine_.ray_counts(stream_, extension, shadow); }

private:
    struct Job {
        uint64_t id = 0;
        SceneDev* scene = nullptr;
        lux_render_settings rs{};
        std::vector<lux_vec3> vertices;
        bool update = false;
        Request rq;
        float* film = nullptr;
        size_t film_bytes = 0;
        double arrival = 0, admitted = 0, finish = 0;
        bool done = false;
    };

    // Work of the worker thread: a scene made resident (restore) and/or refit
    // (vertices), or the device state of a dropped scene freed (dead).
    struct Task {
        SceneDev* scene = nullptr;
        std::shared_ptr<SceneSource> source;  // restore from this
        bool restore = false;
        std::vector<lux_vec3> vertices;  // refit to these
        SceneDev fresh;                  // restore target, adopted on completion
        SceneDev dead;
        size_t dead_bytes = 0;
        std::atomic<bool> done{false};
        std::exception_ptr error;
        double seconds = 0;
    };

    void retire(bool block, std::vector<uint64_t>* done);
    void admit();
    void activate(Job* j);
    void finish_tasks();
    void evict_for(size_t need, const SceneDev* keep);
    void post(const std::shared_ptr<Task>& t);
    void worker_loop();
    bool enqueue_one();
    float* film_alloc(size_t bytes);
    void film_free(float* p, size_t bytes);
    size_t footprint_in_use() const;

    ServeConfig cfg_;
    hipStream_t stream_;
    hipStream_t update_stream_ = nullptr;
    Renderer engine_;
    std::chrono::steady_clock::time_point t0_;
    uint64_t next_id_ = 1;
    std::deque<std::unique_ptr<Job>> pending_;
    std::vector<Job*> active_;                     // admitted, not yet complete
    std::deque<Job*> issue_;                       // admitted with paths left to place
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs_;  // every job not yet taken
    std::unordered_map<const SceneDev*, int> scene_users_;       // admitted, unfinished jobs per scene
    std::unordered_map<SceneDev*, size_t> resident_;             // scenes resident for the server (bytes)
    size_t resident_scene_bytes_ = 0;
    size_t freeing_bytes_ = 0;                                   // dropped scenes not yet freed
    std::unordered_map<const SceneDev*, uint64_t> last_use_;     // admission tick
    uint64_t tick_ = 0;
    std::vector<Job*> waiting_;                                  // admitted, scene work running
    std::vector<std::shared_ptr<Task>> tasks_;                   // posted, not yet finished
    std::unordered_map<const SceneDev*, Task*> preparing_;       // scene -> its running task
    int device_ = 0;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::shared_ptr<Task>> queue_;
    bool stop_ = false;
    std::thread worker_;
    size_t films_in_use_ = 0, films_pooled_ = 0;
    std::multimap<size_t, float*> film_pool_;
    struct Inflight {
        int64_t batch;
        hipEvent_t ev;
    };
    std::deque<Inflight> inflight_;
    std::vector<hipEvent_t> event_pool_;
    ServeStats stats_;
};

// Replays an arrival trace in real time (arrival[i] seconds after the start,
// sorted): request i renders settings[i] of scenes[i], after moving its
// vertices to vertices[i] when that is non-null. Films (nullable) receive the
// film sums, latency[i] = completion - arrival in seconds, wall = total time.
void run_trace(Server& sv, int n, SceneDev* const* scenes, const lux_render_settings* settings, const double* arrival,
               const lux_vec3* const* vertices, const int* vertex_counts, float* const* films, double* latency,
               double* wall);

}  // namespace pt
}  // namespace lux

// serve.cpp -- render-serving loop over the wavefront path tracer.
//
// Scenes stay resident (their accels, BVHs and shading tables are registered once). Requests
// (scene, resolution, spp, seed, output image) arrive at any time and are admitted in FIFO
// order while the running total of their accumulators fits an HBM arena of the configured budget
// (and while fewer than max_active are admitted). Continuous batching: each scene has its own
// stream and wavefront pipeline; whenever a scene's previous wave has completed, the scheduler
// packs a new wave from all admitted requests of that scene, one sample per request per round
// until the wave's path capacity is reached, so a request that arrives while others are on the
// GPU joins the next wave. Waves of different scenes overlap on the device. A request completes
// with the wave that renders its last sample; its image is the same as a render of the request
// alone (paths depend only on the pixel, the sample index and the request's seed, and samples
// are summed per pixel in sample order).
#include <hip/hip_runtime.h>

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

#include "lc_hip.h"
#include "lc_hip_internal.h"

namespace lchip {

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device allocation (the admission budget).
class Arena {
public:
    hipError_t init(size_t bytes) {
        size_ = bytes;
        if (bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(&base_, bytes);
        if (e != hipSuccess) {
            base_ = nullptr;
            return e;
        }
        free_[0] = bytes;
        return hipSuccess;
    }
    ~Arena() {
        if (base_) hipFree(base_);
    }
    // offset of a block of `bytes` (256-aligned), or SIZE_MAX
    size_t alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return off;
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char *base() const { return static_cast<char *>(base_); }
    size_t peak() const { return peak_; }
    size_t used() const { return used_; }

private:
    void *base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

struct Request {
    uint64_t id = 0;
    uint32_t scene = 0, width = 0, height = 0, spp = 0;
    uint64_t seed = 0, sequence = 0;
    float4 *out = nullptr;
    float4 *acc = nullptr;
    size_t bytes = 0, arena_off = 0;
    uint32_t issued = 0, done = 0; // samples handed to waves / completed
    int state = LC_HIP_REQUEST_QUEUED;
    double t_submit = 0, t_admit = -1, t_done = -1;
    std::string error;
};

struct SceneSlot {
    PtBinding B;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Request *> active;                          // admitted, not finished (admission order)
    std::vector<std::pair<Request *, uint32_t>> in_flight;  // samples of the wave on the GPU
    unsigned long long *counts = nullptr;                    // device: closest, shadow rays
};

} // namespace

struct Server {
    Device *dev = nullptr;
    ServeOptions opt;
    std::mutex mu;
    std::condition_variable cv;      // wakes the scheduler
    std::condition_variable done_cv; // wakes waiters
    std::deque<Request *> pending;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::unique_ptr<SceneSlot>> scenes;
    Arena arena;
    std::thread worker;
    bool stop = false;
    uint64_t next_id = 1;
    Clock::time_point t0 = Clock::now();
    uint64_t waves = 0, finished = 0;
    uint32_t admitted = 0, peak_active = 0;
    std::string fault;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void loop();
    bool step(); // one scheduling pass under the lock; true when it changed something
};

bool Server::step() {
    bool changed = false;
    // 1. retire completed waves
    for (auto &sp : scenes) {
        SceneSlot &sc = *sp;
        if (!sc.busy) continue;
        const hipError_t q = hipEventQuery(sc.ev);
        if (q == hipErrorNotReady) continue;
        sc.busy = false;
        changed = true;
        const double t = now();
        for (auto &f : sc.in_flight) {
            Request *r = f.first;
            if (q != hipSuccess) {
                r->state = LC_HIP_REQUEST_FAILED;
                r->error = hipGetErrorString(q);
            } else {
                r->done += f.second;
                if (r->done == r->spp) r->state = LC_HIP_REQUEST_DONE;
            }
            if (r->state == LC_HIP_REQUEST_DONE || r->state == LC_HIP_REQUEST_FAILED) {
                r->t_done = t;
                arena.release(r->arena_off, r->bytes);
                admitted--;
                finished++;
                sc.active.erase(std::find(sc.active.begin(), sc.active.end(), r));
            }
        }
        sc.in_flight.clear();
        done_cv.notify_all();
    }
    // 2. admit in arrival order while the accumulator budget and the slot limit allow
    while (!pending.empty() && admitted < opt.max_active) {
        Request *r = pending.front();
        const size_t off = arena.alloc(r->bytes);
        if (off == SIZE_MAX) break;
        pending.pop_front();
        r->arena_off = off;
        r->acc = reinterpret_cast<float4 *>(arena.base() + off);
        r->state = LC_HIP_REQUEST_ACTIVE;
        r->t_admit = now();
        scenes[r->scene]->active.push_back(r);
        admitted++;
        peak_active = std::max(peak_active, admitted);
        changed = true;
    }
    // 3. a new wave on every idle scene that has samples left to issue
    for (auto &sp : scenes) {
        SceneSlot &sc = *sp;
        if (sc.busy || sc.active.empty()) continue;
        std::vector<uint32_t> take(sc.active.size(), 0);
        uint64_t paths = 0;
        uint32_t jobs = 0;
        for (bool grew = true; grew;) {
            grew = false;
            for (size_t k = 0; k < sc.active.size(); k++) {
                Request *r = sc.active[k];
                const uint64_t npix = (uint64_t)r->width * r->height;
                if (r->issued + take[k] >= r->spp || paths + npix > opt.wave_paths) continue;
                if (take[k] == 0 && jobs == kMaxWaveJobs) continue;
                if (take[k] == 0) jobs++;
                take[k]++;
                paths += npix;
                grew = true;
            }
        }
        if (paths == 0) continue;
        std::vector<PtJob> table;
        table.reserve(jobs);
        for (size_t k = 0; k < sc.active.size(); k++) {
            if (take[k] == 0) continue;
            Request *r = sc.active[k];
            PtJob J;
            std::memset(&J, 0, sizeof(J));
            J.acc = r->acc;
            J.out = r->out;
            J.width = r->width;
            J.height = r->height;
            J.sample0 = r->issued;
            J.nsamples = take[k];
            J.first = r->issued == 0 ? 1u : 0u;
            J.finish_spp = r->issued + take[k] == r->spp ? r->spp : 0u;
            J.seed = r->seed;
            J.sequence = r->sequence;
            table.push_back(J);
            sc.in_flight.emplace_back(r, take[k]);
            r->issued += take[k];
        }
        std::string err;
        int rc = pt_wave(dev, sc.stream, sc.B, table.data(), (uint32_t)table.size(), opt.wave_paths, opt.pt, sc.counts,
                         err);
        hipError_t e = rc == 0 ? hipEventRecord(sc.ev, sc.stream) : hipErrorUnknown;
        if (rc != 0 || e != hipSuccess) {
            // the wave never reached the GPU: fail its requests now
            fault = rc != 0 ? err : hipGetErrorString(e);
            const double t = now();
            for (auto &f : sc.in_flight) {
                Request *r = f.first;
                if (r->state != LC_HIP_REQUEST_ACTIVE) continue;
                r->state = LC_HIP_REQUEST_FAILED;
                r->error = fault;
                r->t_done = t;
                arena.release(r->arena_off, r->bytes);
                admitted--;
                finished++;
                sc.active.erase(std::find(sc.active.begin(), sc.active.end(), r));
            }
            sc.in_flight.clear();
            done_cv.notify_all();
        } else {
            sc.busy = true;
            waves++;
        }
        changed = true;
    }
    return changed;
}

void Server::loop() {
    hipSetDevice(dev->index);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        const bool changed = step();
        bool busy = false;
        for (auto &sp : scenes) busy = busy || sp->busy;
        if (stop && !busy && pending.empty()) {
            bool idle = true;
            for (auto &sp : scenes) idle = idle && sp->active.empty();
            if (idle) break;
        }
        if (changed) continue;
        // waves on the GPU: poll their events every 50 us; otherwise sleep until a submission
        if (busy) cv.wait_for(lk, std::chrono::microseconds(50));
        else cv.wait_for(lk, std::chrono::milliseconds(5));
    }
}

Server *serve_create(Device *dev, const ServeOptions &opt, std::string &err) {
    auto s = std::make_unique<Server>();
    s->dev = dev;
    s->opt = opt;
    hipError_t e = s->arena.init(opt.hbm_budget);
    if (e != hipSuccess) {
        err = std::string("serving arena: ") + hipGetErrorString(e);
        return nullptr;
    }
    Server *p = s.release();
    p->worker = std::thread([p] { p->loop(); });
    return p;
}

void serve_destroy(Server *s) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        // requests still queued are dropped
        for (Request *r : s->pending) {
            r->state = LC_HIP_REQUEST_FAILED;
            r->error = "server destroyed";
        }
        s->pending.clear();
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto &sp : s->scenes) {
        if (sp->stream) {
            hipStreamSynchronize(sp->stream);
            pt_release_stream(s->dev, sp->stream);
            hipStreamDestroy(sp->stream);
        }
        if (sp->ev) hipEventDestroy(sp->ev);
        if (sp->counts) hipFree(sp->counts);
    }
    delete s;
}

int serve_add_scene(Server *s, const PtBinding &B, std::string &err) {
    auto sc = std::make_unique<SceneSlot>();
    sc->B = B;
    hipError_t e = hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sc->ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&sc->counts, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(sc->counts, 0, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        err = std::string("serving scene: ") + hipGetErrorString(e);
        if (sc->stream) hipStreamDestroy(sc->stream);
        if (sc->ev) hipEventDestroy(sc->ev);
        if (sc->counts) hipFree(sc->counts);
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->scenes.push_back(std::move(sc));
    return (int)s->scenes.size() - 1;
}

int serve_submit(Server *s, uint32_t scene, uint32_t width, uint32_t height, uint32_t spp, uint64_t seed,
                 uint64_t sequence, float4 *out, uint64_t *id, std::string &err) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (scene >= s->scenes.size()) {
        err = "unknown scene";
        return -1;
    }
    const uint64_t npix = (uint64_t)width * height;
    if (npix == 0 || spp == 0) {
        err = "empty image or no samples";
        return -1;
    }
    if (npix > s->opt.wave_paths) {
        err = "one sample of this image exceeds the wave capacity";
        return -1;
    }
    if (npix * sizeof(float4) > s->opt.hbm_budget) {
        err = "the image's accumulator exceeds the HBM budget";
        return -1;
    }
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->scene = scene;
    r->width = width;
    r->height = height;
    r->spp = spp;
    r->seed = seed;
    r->sequence = sequence;
    r->out = out;
    r->bytes = npix * sizeof(float4);
    r->t_submit = s->now();
    *id = r->id;
    s->pending.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->cv.notify_all();
    return 0;
}

int serve_query(Server *s, uint64_t id, ServeStatus *st) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return -1;
    const Request &r = *it->second;
    st->state = r.state;
    st->samples_done = r.done;
    st->t_submit = r.t_submit;
    st->t_admit = r.t_admit;
    st->t_done = r.t_done;
    return 0;
}

int serve_wait(Server *s, uint64_t id, double timeout_s) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return -1;
    Request *r = it->second.get();
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    while (r->state != LC_HIP_REQUEST_DONE && r->state != LC_HIP_REQUEST_FAILED) {
        if (s->done_cv.wait_until(lk, deadline) == std::cv_status::timeout &&
            r->state != LC_HIP_REQUEST_DONE && r->state != LC_HIP_REQUEST_FAILED)
            return 1;
    }
    return r->state == LC_HIP_REQUEST_DONE ? 0 : -2;
}

int serve_drain(Server *s, double timeout_s) {
    std::unique_lock<std::mutex> lk(s->mu);
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    while (s->finished < s->reqs.size()) {
        if (s->done_cv.wait_until(lk, deadline) == std::cv_status::timeout && s->finished < s->reqs.size()) return 1;
    }
    return 0;
}

void serve_stats(Server *s, ServeStats *st) {
    std::vector<unsigned long long> c(2 * s->scenes.size(), 0);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        st->waves = s->waves;
        st->requests_done = s->finished;
        st->peak_active = s->peak_active;
        st->active = s->admitted;
        st->queued = (uint32_t)s->pending.size();
        st->peak_bytes = s->arena.peak();
    }
    st->rays_closest = st->rays_shadow = 0;
    for (size_t k = 0; k < s->scenes.size(); k++) {
        unsigned long long h[2] = {0, 0};
        if (hipMemcpy(h, s->scenes[k]->counts, sizeof(h), hipMemcpyDeviceToHost) == hipSuccess) {
            st->rays_closest += h[0];
            st->rays_shadow += h[1];
        }
    }
}

} // namespace lchip

// A registration server: each request is reg_aladin's affine followed by reg_f3d initialised from
// that affine (reg-apps/reg_aladin.cpp then reg-apps/reg_f3d.cpp -aff), on one device.
//
// Requests are admitted in arrival order as soon as a slot is free and the device memory the
// running requests are expected to hold, plus the new one's, stays within the server's budget.
// Every admitted request runs on a host thread of its own with its own HIP streams (the handles'
// streams), so the kernels of the running requests overlap on the device and a request that
// arrives while others run starts when they leave room, not when they all end.  A request keeps
// its images, grids and optimiser state on the device for its whole run.  Its computation does
// not depend on the others: its result is the same alone and in any mix.  A request that fails
// (a flat or degenerate image, a device error) fails alone with its own status.

#include "common.hpp"
#include "host.hpp"
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace nrh {
namespace {

struct Job {
    uint64_t id = 0;
    nrh_geom ref{}, flo{};
    std::vector<float> refData, floData;
    nrh_aladin_config aladin{};
    nrh_f3d_config f3d{};
    size_t bytes = 0;
    double tSubmit = 0, tStart = 0, tEnd = 0;
    bool done = false;
    int status = NRH_OK;
    std::string error;
    float affine[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    nrh_geom grid{};
    std::vector<float> cpp;
    double similarity = 0;
    int iterations = 0;
};

// Device memory a request holds at its largest stage (the Aladin handles are freed before F3D
// starts): the images, warped image and block lists of Aladin; for F3D with NMI the fused path's
// images (with their float16 copies), exact-pass list, voxel gradient and projection buffers;
// with LNCC the per-operation chain's field, warped image and gradients, convolution scratch and
// two sets of six statistics images; plus a fixed allowance for grids and small buffers.
size_t estimate_bytes(const nrh_geom& r, const nrh_geom& f, const nrh_f3d_config& c) {
    const size_t N = (size_t)r.nx * r.ny * r.nz, Nf = (size_t)f.nx * f.ny * f.nz;
    const size_t aladin = 12 * N + 4 * Nf;
    const size_t f3d = c.measure == NRH_MEASURE_LNCC ? 124 * N + 4 * Nf : 44 * N + 8 * Nf;
    return std::max(aladin, f3d) + ((size_t)64 << 20);
}

}  // namespace
}  // namespace nrh

struct nrh_server {
    int device = 0;
    size_t budget = 0;
    int maxActive = 1;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<nrh::Job>> pending;
    std::map<uint64_t, std::shared_ptr<nrh::Job>> jobs;
    std::map<uint64_t, std::thread> running;  // worker threads, joined once their request is done
    std::vector<uint64_t> finished;           // requests whose worker has returned its result
    std::thread dispatcher;
    size_t inUse = 0, peakBytes = 0;
    int active = 0, peakActive = 0, completed = 0, failed = 0;
    uint64_t nextId = 1;
    bool stopping = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace nrh {
namespace {

void run_job(nrh_server *s, std::shared_ptr<Job> job) {
    int rc = hipSetDevice(s->device) == hipSuccess ? NRH_OK : NRH_ERR_HIP;
    std::string err = rc == NRH_OK ? "" : "hipSetDevice failed";
    if (rc == NRH_OK) {
        rc = nrh_aladin_register(&job->ref, job->refData.data(), &job->flo, job->floData.data(), &job->aladin,
                                 job->affine);
        if (rc != NRH_OK) err = std::string("Aladin: ") + nrh_last_error();
    }
    if (rc == NRH_OK) {
        nrh_f3d_config cfg = job->f3d;
        cfg.affine = job->affine;
        const float sp = std::fabs(cfg.spacing);
        size_t cap = 3;
        const int dims[3] = { job->ref.nx, job->ref.ny, job->ref.nz };
        const float pix[3] = { job->ref.dx, job->ref.dy, job->ref.dz };
        for (int a = 0; a < 3; ++a) cap *= (size_t)std::ceil(dims[a] * pix[a] / (cfg.spacing < 0 ? sp * pix[a] : sp) + 4);
        job->cpp.resize(cap);
        rc = nrh_f3d_register(&job->ref, job->refData.data(), &job->flo, job->floData.data(), &cfg, &job->grid,
                              job->cpp.data(), cap, &job->similarity, &job->iterations);
        if (rc != NRH_OK) err = std::string("F3D: ") + nrh_last_error();
        else job->cpp.resize(3 * (size_t)job->grid.nx * job->grid.ny * job->grid.nz);
    }
    // the images are no longer needed
    std::vector<float>().swap(job->refData);
    std::vector<float>().swap(job->floData);
    std::lock_guard<std::mutex> lk(s->mu);
    job->status = rc;
    job->error = err;
    job->tEnd = s->now();
    job->done = true;
    s->inUse -= job->bytes;
    s->active -= 1;
    if (rc == NRH_OK) s->completed += 1;
    else s->failed += 1;
    s->finished.push_back(job->id);
    s->cv.notify_all();
}

// Joins the workers that have finished (called with the lock held; they no longer need it).
void reap(nrh_server *s) {
    for (uint64_t id : s->finished) {
        auto it = s->running.find(id);
        if (it != s->running.end()) {
            if (it->second.joinable()) it->second.join();
            s->running.erase(it);
        }
    }
    s->finished.clear();
}

// Admits the oldest pending request when a slot is free and its memory fits beside the running
// requests' (a request alone is always admitted: its size was checked against the budget).
void dispatch(nrh_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] {
            if (!s->finished.empty()) return true;
            if (s->pending.empty()) return s->stopping;
            const auto& j = s->pending.front();
            return s->active == 0 || (s->active < s->maxActive && s->inUse + j->bytes <= s->budget);
        });
        reap(s);
        if (s->pending.empty()) {
            if (s->stopping) return;  // nothing left to admit
            continue;
        }
        const auto& head = s->pending.front();
        if (!(s->active == 0 || (s->active < s->maxActive && s->inUse + head->bytes <= s->budget))) continue;
        std::shared_ptr<Job> job = s->pending.front();
        s->pending.pop_front();
        s->active += 1;
        s->inUse += job->bytes;
        s->peakActive = std::max(s->peakActive, s->active);
        s->peakBytes = std::max(s->peakBytes, s->inUse);
        job->tStart = s->now();
        s->running[job->id] = std::thread(run_job, s, job);
    }
}

}  // namespace
}  // namespace nrh

using namespace nrh;

extern "C" {

int nrh_server_create(int device, size_t hbm_budget_bytes, int max_active, nrh_server **out) {
    return guard([&] {
        NRH_REQUIRE(out && max_active >= 1 && hbm_budget_bytes > 0, NRH_ERR_ARG, "invalid server arguments");
        int count = 0;
        NRH_HIP(hipGetDeviceCount(&count));
        NRH_REQUIRE(device >= 0 && device < count, NRH_ERR_ARG, "no such device");
        std::unique_ptr<nrh_server> s(new nrh_server);
        s->device = device;
        s->budget = hbm_budget_bytes;
        s->maxActive = max_active;
        s->dispatcher = std::thread(dispatch, s.get());
        *out = s.release();
    });
}

int nrh_server_submit(nrh_server *s, const nrh_request *req, uint64_t *id) {
    return guard([&] {
        NRH_REQUIRE(s && req && id && req->ref_data && req->flo_data, NRH_ERR_ARG, "null argument");
        check_geom(&req->ref, "reference");
        check_geom(&req->flo, "floating");
        auto job = std::make_shared<Job>();
        job->ref = req->ref;
        job->flo = req->flo;
        job->refData.assign(req->ref_data, req->ref_data + voxels(req->ref));
        job->floData.assign(req->flo_data, req->flo_data + voxels(req->flo));
        job->aladin = req->aladin;
        job->f3d = req->f3d;
        job->f3d.affine = nullptr;
        job->bytes = estimate_bytes(req->ref, req->flo, req->f3d);
        NRH_REQUIRE(job->bytes <= s->budget, NRH_ERR_NOMEM, "the request alone exceeds the server's memory budget");
        std::lock_guard<std::mutex> lk(s->mu);
        NRH_REQUIRE(!s->stopping, NRH_ERR_STATE, "the server is shutting down");
        job->id = s->nextId++;
        job->tSubmit = s->now();
        s->jobs[job->id] = job;
        s->pending.push_back(job);
        *id = job->id;
        s->cv.notify_all();
    });
}

int nrh_server_wait(nrh_server *s, uint64_t id, nrh_result *res, float *cpp_out, size_t capacity) {
    return guard([&] {
        NRH_REQUIRE(s && res, NRH_ERR_ARG, "null argument");
        std::shared_ptr<Job> job;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            auto it = s->jobs.find(id);
            NRH_REQUIRE(it != s->jobs.end(), NRH_ERR_ARG, "unknown request id (or its result was already taken)");
            job = it->second;
            s->cv.wait(lk, [&] { return job->done; });
            s->jobs.erase(it);
        }
        *res = nrh_result{};
        res->status = job->status;
        std::snprintf(res->error, sizeof res->error, "%s", job->error.c_str());
        std::memcpy(res->affine, job->affine, sizeof res->affine);
        res->grid = job->grid;
        res->similarity = job->similarity;
        res->iterations = job->iterations;
        res->submitted = job->tSubmit;
        res->started = job->tStart;
        res->finished = job->tEnd;
        res->device_bytes = job->bytes;
        if (cpp_out && job->status == NRH_OK) {
            NRH_REQUIRE(job->cpp.size() <= capacity, NRH_ERR_ARG, "cpp_out too small for the request's grid");
            std::memcpy(cpp_out, job->cpp.data(), job->cpp.size() * sizeof(float));
        }
    });
}

int nrh_server_get_stats(nrh_server *s, nrh_server_stats *out) {
    return guard([&] {
        NRH_REQUIRE(s && out, NRH_ERR_ARG, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        out->pending = (int)s->pending.size();
        out->active = s->active;
        out->completed = s->completed;
        out->failed = s->failed;
        out->peak_active = s->peakActive;
        out->peak_bytes = s->peakBytes;
        out->budget_bytes = s->budget;
    });
}

int nrh_server_destroy(nrh_server *s) {
    return guard([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stopping = true;
            s->cv.notify_all();
        }
        if (s->dispatcher.joinable()) s->dispatcher.join();
        // the dispatcher has started every pending request; wait for them to finish
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->active == 0; });
            reap(s);
        }
        for (auto& w : s->running)
            if (w.second.joinable()) w.second.join();
        delete s;
    });
}

}  // extern "C"

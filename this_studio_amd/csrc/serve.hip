// Batched serving of independent scenes on one GPU.
//
// The reference runs one scene per process on one stream with host
// synchronisation inside its step loop (interaction.h:164-175); it has no
// serving path.  This server keeps a FIFO of submitted scene jobs (a staged
// WCSPH scene, DEM scene, or both coupled), admits the head job once it has
// arrived, a slot is free and its estimated device memory fits the budget,
// initialises it on admission (the memory is allocated then), advances it on
// its own stream in blocks of steps, copies its final state to the host and
// destroys it, which returns its memory and slot.  Each slot is a host
// thread, so a host synchronisation inside one job's step block (the DEM
// pair-list resize) stalls only that slot while the other slots' jobs keep
// the device busy.  With graphs on, WCSPH jobs replay captured two-step
// graphs and DEM jobs the four steps between pair rebuilds.  Every step ends with a device timestamp, from which the per-job
// step latencies (intervals between consecutive step completions) come.
//
// One slot with graphs off is the plain path: the same jobs served one at a
// time, one kernel launch per stage from the host.

#include "common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

using namespace demsph;

namespace {

void ok(int rc)
{
    if (rc == DEMSPH_OK) return;
    const std::string msg = demsph_last_error();
    if (rc == DEMSPH_ERR_CAPACITY) throw CapacityError(msg);
    if (rc == DEMSPH_ERR_HIP) throw HipFailure(msg);
    throw InvalidArg(msg);
}

}  // namespace

struct demsph_server {
    struct Job {
        int id = 0;
        int kind = 0;  // 0 WCSPH, 1 DEM, 2 coupled
        demsph_sph* fluid = nullptr;
        demsph_dem* dem = nullptr;
        demsph_couple* couple = nullptr;
        double fmin[3] = {}, fmax[3] = {}, dmin[3] = {}, dmax[3] = {}, g[3] = {};
        double dt = 0.0;
        long steps = 0;
        double arrival = 0.0;
        size_t bytes = 0;
        long particles = 0;
        // outcome
        int status = DEMSPH_OK;
        std::string error;
        double admitted = -1.0, started = -1.0, stepped = -1.0, finished = -1.0;
        long done = 0;
        std::vector<unsigned long long> stamps;
        std::vector<double> fluid_pos, fluid_vel, ball_pos, ball_vel;

        void destroy_scenes()
        {
            if (couple) (void)demsph_couple_destroy(couple);
            if (dem) (void)demsph_dem_destroy(dem);
            if (fluid) (void)demsph_sph_destroy(fluid);
            couple = nullptr;
            dem = nullptr;
            fluid = nullptr;
        }
    };

    int slots = 1;
    size_t budget = 0;
    bool graphs = true;
    long block = 100;  // steps per step() call
    int autotune = -1;   // WCSPH autotuner mode on admission (-1: off)
    // Jobs shorter than this many steps are not tuned on admission: a size
    // class's trials cost some 40-50 steps of the scene, which a short job
    // does not earn back.
    long tune_min_steps = 2000;
    int tune_steps = 3;  // timed steps per autotuner trial
    int device = 0;
    double clock_hz = 1.0e8;
    std::vector<std::unique_ptr<Job>> jobs;
    bool ran = false;

    // run state
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Job*> order;
    size_t next = 0;
    size_t in_use = 0, peak_bytes = 0;
    int active = 0, peak_active = 0;
    int finished_jobs = 0;
    demsph_server_progress_fn progress = nullptr;
    void* progress_user = nullptr;
    std::chrono::steady_clock::time_point t0;
    double wall = 0.0;

    ~demsph_server()
    {
        for (auto& j : jobs) j->destroy_scenes();
    }

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    void submit(demsph_sph* fluid, demsph_dem* dem, const double* fmin, const double* fmax, const double* dmin,
                const double* dmax, const double* g, double dt, long steps, double arrival, int* id)
    {
        if (ran) throw InvalidArg("the server has already run");
        if (!fluid && !dem) throw InvalidArg("a job needs a fluid scene, a DEM scene or both");
        if (!g) throw InvalidArg("null gravity");
        if (fluid && (!fmin || !fmax)) throw InvalidArg("fluid bounds missing");
        if (dem && (!dmin || !dmax)) throw InvalidArg("DEM bounds missing");
        if (!(dt > 0.0)) throw InvalidArg("time step must be positive");
        if (steps < 1) throw InvalidArg("steps must be positive");
        if (!(arrival >= 0.0)) throw InvalidArg("arrival must be non-negative");
        auto j = std::make_unique<Job>();
        j->id = (int)jobs.size();
        j->kind = fluid && dem ? 2 : (fluid ? 0 : 1);
        j->fluid = fluid;
        j->dem = dem;
        for (int q = 0; q < 3; q++) {
            j->g[q] = g[q];
            if (fluid) { j->fmin[q] = fmin[q]; j->fmax[q] = fmax[q]; }
            if (dem) { j->dmin[q] = dmin[q]; j->dmax[q] = dmax[q]; }
        }
        j->dt = dt;
        j->steps = steps;
        j->arrival = arrival;
        if (fluid) {
            j->bytes += sph_bytes_estimate(fluid, fmin, fmax);
            j->particles += sph_num_fluid(fluid);
        }
        if (dem) {
            j->bytes += dem_bytes_estimate(dem, dmin, dmax);
            j->particles += dem_num_balls(dem);
        }
        if (fluid && dem) j->bytes += couple_bytes_estimate(fluid, dem, fmin, fmax);
        j->bytes += sizeof(unsigned long long) * (size_t)(steps + 1) + 256;  // step stamps
        if (id) *id = j->id;
        jobs.push_back(std::move(j));
    }

    // Initialise, step, read back and destroy one admitted job.
    void run_job(Job& j)
    {
        unsigned long long* sbuf = nullptr;
        int* scount = nullptr;
        hipStream_t own = nullptr;
        try {
            // A WCSPH scene creates its own non-blocking stream; a DEM-only
            // job gets one from the server, a coupled job runs both scenes
            // on the fluid's.
            hipStream_t st = nullptr;
            if (j.fluid) {
                ok(demsph_sph_set_graphs(j.fluid, graphs ? 1 : 0));
                ok(demsph_sph_initialize(j.fluid, j.fmin, j.fmax, j.g));
                // trials restore the state; a cached size class costs nothing
                if (tune_job(j)) ok(demsph_sph_autotune(j.fluid, j.dt, tune_steps, autotune, 1, nullptr));
                st = sph_stream(j.fluid);
            } else {
                DEMSPH_HIP(stream_create(&own));
                st = own;
            }
            if (j.dem) {
                ok(demsph_dem_set_graphs(j.dem, graphs ? 1 : 0));
                dem_set_stream(j.dem, st);
                ok(demsph_dem_initialize(j.dem, j.dmin, j.dmax));
                ok(demsph_dem_solve_begin(j.dem));
                // pair lanes build the same lists in the same order: every mode keeps results
                if (tune_job(j)) dem_autotune_cached(j.dem, j.g, j.dt, std::max(5, tune_steps));
            }
            if (j.kind == 2) ok(demsph_couple_create(&j.couple, j.fluid, j.dem));
            DEMSPH_HIP(dev_malloc((void**)&sbuf, sizeof(unsigned long long) * (size_t)(j.steps + 1)));
            DEMSPH_HIP(dev_malloc((void**)&scount, sizeof(int)));
            fill_i32(scount, 1, 0, st);
            StepStamps ss;
            ss.buf = sbuf;
            ss.counter = scount;
            ss.cap = (int)(j.steps + 1);
            if (j.fluid) sph_set_stamps(j.fluid, ss);  // the coupled step stamps through the fluid scene
            else dem_set_stamps(j.dem, ss);
            DEMSPH_HIP(hipStreamSynchronize(st));
            j.started = now();
            while (j.done < j.steps) {
                const int n = (int)std::min(block, j.steps - j.done);
                if (j.kind == 0) ok(demsph_sph_step(j.fluid, j.dt, n));
                else if (j.kind == 1) ok(demsph_dem_step(j.dem, j.g, j.dt, n));
                else ok(demsph_couple_step(j.couple, j.g, j.dt, n));
                j.done += n;
            }
            DEMSPH_HIP(hipStreamSynchronize(st));
            j.stepped = now();
            int cnt = 0;
            copy_sync(&cnt, scount, sizeof(int), hipMemcpyDeviceToHost, st);
            cnt = std::min<int>(cnt, (int)(j.steps + 1));
            j.stamps.resize((size_t)cnt);
            copy_sync(j.stamps.data(), sbuf, sizeof(unsigned long long) * (size_t)cnt, hipMemcpyDeviceToHost, st);
            if (j.fluid) sph_set_stamps(j.fluid, StepStamps());
            if (j.dem) dem_set_stamps(j.dem, StepStamps());
            auto field = [](auto get, auto* h, const char* f, std::vector<double>& out) {
                size_t n = 0;
                ok(get(h, f, nullptr, 0, &n));
                out.resize(n);
                ok(get(h, f, out.data(), n, &n));
            };
            if (j.fluid) {
                field(demsph_sph_get, j.fluid, "position", j.fluid_pos);
                field(demsph_sph_get, j.fluid, "velocity", j.fluid_vel);
            }
            if (j.dem) {
                field(demsph_dem_get, j.dem, "position", j.ball_pos);
                field(demsph_dem_get, j.dem, "velocity", j.ball_vel);
            }
        } catch (const CapacityError& e) {
            j.status = DEMSPH_ERR_CAPACITY;
            j.error = e.what();
        } catch (const HipFailure& e) {
            j.status = DEMSPH_ERR_HIP;
            j.error = e.what();
        } catch (const std::exception& e) {
            j.status = DEMSPH_ERR_INVALID;
            j.error = e.what();
        }
        j.destroy_scenes();
        dev_free(sbuf);
        dev_free(scount);
        stream_destroy(own);
        j.finished = now();
    }

    bool tune_job(const Job& j) const { return autotune >= 0 && j.steps >= tune_min_steps; }

    void worker()
    {
        (void)hipSetDevice(device);
        for (;;) {
            Job* j = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (next >= order.size()) return;
                    Job* h = order[next];
                    if (h->bytes > budget) {  // can never fit: fail it, keep the queue moving
                        h->status = DEMSPH_ERR_CAPACITY;
                        h->error = "job needs " + std::to_string(h->bytes) + " bytes, above the memory budget";
                        h->destroy_scenes();
                        next++;
                        finished_jobs++;
                        if (progress) progress(progress_user, finished_jobs, (int)order.size(), now());
                        continue;
                    }
                    const double t = now();
                    if (t < h->arrival) {
                        cv.wait_for(lk, std::chrono::duration<double>(h->arrival - t));
                        continue;
                    }
                    if (in_use + h->bytes > budget) {
                        cv.wait(lk);
                        continue;
                    }
                    next++;
                    in_use += h->bytes;
                    peak_bytes = std::max(peak_bytes, in_use);
                    active++;
                    peak_active = std::max(peak_active, active);
                    h->admitted = t;
                    j = h;
                    break;
                }
            }
            cv.notify_all();  // the next head may be admissible by another slot
            run_job(*j);
            {
                std::lock_guard<std::mutex> lk(mu);
                in_use -= j->bytes;
                active--;
                finished_jobs++;
                // called under the lock: one call at a time, in finishing order
                if (progress) progress(progress_user, finished_jobs, (int)order.size(), now());
            }
            cv.notify_all();
        }
    }

    void run()
    {
        if (ran) throw InvalidArg("the server has already run");
        ran = true;
        order.clear();
        for (auto& j : jobs) order.push_back(j.get());
        std::stable_sort(order.begin(), order.end(), [](const Job* a, const Job* b) { return a->arrival < b->arrival; });
        next = 0;
        // Jobs' device blocks are recycled through the block cache while the
        // server runs (no device-wide waits in hipFree between jobs).
        struct CacheUse {
            explicit CacheUse(size_t cap) { block_cache_begin(cap); }
            ~CacheUse() { block_cache_end(); }
        } cache_use(budget);
        t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> pool;
        for (int s = 0; s < slots; s++) pool.emplace_back([this] { worker(); });
        for (auto& t : pool) t.join();
        wall = now();
    }

    std::vector<double> latencies(const Job& j) const
    {
        std::vector<double> out;
        for (size_t k = 1; k < j.stamps.size(); k++)
            out.push_back((double)(j.stamps[k] - j.stamps[k - 1]) / clock_hz);
        return out;
    }
};

namespace {

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

demsph_server::Job& job_of(demsph_server* v, int id)
{
    if (!v) throw InvalidArg("null server");
    if (id < 0 || (size_t)id >= v->jobs.size()) throw InvalidArg("job id out of range");
    return *v->jobs[(size_t)id];
}

int emit(const std::vector<double>& v, double* out, size_t capacity, size_t* count)
{
    if (count) *count = v.size();
    if (out) {
        if (capacity < v.size()) throw InvalidArg("output capacity too small");
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    }
    return DEMSPH_OK;
}

}  // namespace

extern "C" int demsph_server_create(demsph_server** out, int slots, double memoryBudgetBytes, int graphs)
{
    return guarded([&] {
        if (!out) throw InvalidArg("null argument");
        if (slots < 1 || slots > 64) throw InvalidArg("slots must be in [1, 64]");
        if (!(memoryBudgetBytes > 0.0)) throw InvalidArg("memory budget must be positive");
        auto* v = new demsph_server();
        v->slots = slots;
        v->budget = (size_t)memoryBudgetBytes;
        v->graphs = graphs != 0;
        int dev = 0, khz = 0;
        try {
            DEMSPH_HIP(hipGetDevice(&dev));
            DEMSPH_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        } catch (...) {
            delete v;
            throw;
        }
        v->device = dev;
        v->clock_hz = khz > 0 ? 1.0e3 * khz : 1.0e8;
        *out = v;
    });
}

extern "C" int demsph_server_destroy(demsph_server* v)
{
    return guarded([&] { delete v; });
}

extern "C" int demsph_server_set_block(demsph_server* v, long steps)
{
    return guarded([&] {
        if (!v) throw InvalidArg("null server");
        if (steps < 1) throw InvalidArg("block must be positive");
        v->block = steps;
    });
}

extern "C" int demsph_server_set_autotune(demsph_server* v, int mode, int trialSteps)
{
    return guarded([&] {
        if (!v) throw InvalidArg("null server");
        if (mode < -1 || mode > 1) throw InvalidArg("autotune mode must be -1 (off), 0 or 1");
        if (trialSteps < 1 || trialSteps > 1000) throw InvalidArg("trial steps must be in 1..1000");
        if (v->ran) throw InvalidArg("the server has already run");
        v->autotune = mode;
        v->tune_steps = trialSteps;
    });
}

extern "C" int demsph_server_set_autotune_min_steps(demsph_server* v, long steps)
{
    return guarded([&] {
        if (!v) throw InvalidArg("null server");
        if (steps < 0) throw InvalidArg("the step count must not be negative");
        v->tune_min_steps = steps;
    });
}

extern "C" int demsph_server_submit(demsph_server* v, demsph_sph* fluid, demsph_dem* dem, const double* fluidMin,
                                    const double* fluidMax, const double* demMin, const double* demMax,
                                    const double* gravity, double timeStep, long steps, double arrival, int* jobId)
{
    return guarded([&] {
        if (!v) throw InvalidArg("null server");
        v->submit(fluid, dem, fluidMin, fluidMax, demMin, demMax, gravity, timeStep, steps, arrival, jobId);
    });
}

extern "C" int demsph_server_run(demsph_server* v)
{
    return guarded([&] {
        if (!v) throw InvalidArg("null server");
        v->run();
    });
}

extern "C" int demsph_server_set_progress(demsph_server* v, demsph_server_progress_fn fn, void* user)
{
    return guarded([&] {
        if (!v) throw InvalidArg("null server");
        if (v->ran) throw InvalidArg("set the progress callback before demsph_server_run");
        v->progress = fn;
        v->progress_user = user;
    });
}

extern "C" int demsph_server_job_info(demsph_server* v, int jobId, demsph_job_info* info, char* error, size_t errorCap)
{
    return guarded([&] {
        const auto& j = job_of(v, jobId);
        if (info) {
            info->kind = j.kind;
            info->status = j.status;
            info->particles = j.particles;
            info->steps = j.steps;
            info->stepsDone = j.done;
            info->bytes = j.bytes;
            info->arrival = j.arrival;
            info->admitted = j.admitted;
            info->started = j.started;
            info->stepped = j.stepped;
            info->finished = j.finished;
        }
        if (error && errorCap) {
            std::strncpy(error, j.error.c_str(), errorCap - 1);
            error[errorCap - 1] = 0;
        }
    });
}

extern "C" int demsph_server_job_latencies(demsph_server* v, int jobId, double* out, size_t capacity, size_t* count)
{
    return guarded([&] { emit(v->latencies(job_of(v, jobId)), out, capacity, count); });
}

extern "C" int demsph_server_job_result(demsph_server* v, int jobId, const char* field, double* out, size_t capacity,
                                        size_t* count)
{
    return guarded([&] {
        const auto& j = job_of(v, jobId);
        if (!field) throw InvalidArg("null field");
        const std::string f = field;
        if (f == "position") return (void)emit(j.fluid_pos, out, capacity, count);
        if (f == "velocity") return (void)emit(j.fluid_vel, out, capacity, count);
        if (f == "ballPosition") return (void)emit(j.ball_pos, out, capacity, count);
        if (f == "ballVelocity") return (void)emit(j.ball_vel, out, capacity, count);
        throw InvalidArg("unknown field '" + f + "'");
    });
}

extern "C" int demsph_server_stats(demsph_server* v, demsph_server_summary* s)
{
    return guarded([&] {
        if (!v || !s) throw InvalidArg("null argument");
        std::vector<double> all;
        double ps = 0.0;
        int done = 0, failed = 0;
        for (const auto& j : v->jobs) {
            if (j->status != DEMSPH_OK) {
                failed++;
                continue;
            }
            if (j->done < j->steps) continue;
            done++;
            ps += (double)j->particles * (double)j->done;
            auto l = v->latencies(*j);
            all.insert(all.end(), l.begin(), l.end());
        }
        s->wall = v->wall;
        s->particleSteps = ps;
        s->particleStepsPerSecond = v->wall > 0.0 ? ps / v->wall : 0.0;
        s->p50 = percentile(all, 0.50);
        s->p99 = percentile(all, 0.99);
        s->jobsDone = done;
        s->jobsFailed = failed;
        s->peakBytes = v->peak_bytes;
        s->peakActive = v->peak_active;
    });
}

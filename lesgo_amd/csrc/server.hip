// Serving loop: LES jobs of different grids admitted under an HBM budget,
// kept device-resident and stepped together (continuous batching).
//
// Jobs of one grid class (every lg_params field except the per-member z0 and
// mean pressure gradient) share pools: batches of fixed slot capacity, each
// on its own stream, so pools of different grids overlap on the device.  A
// tick admits queued jobs into free slots, replays one captured step of
// every non-empty pool, waits for all of them and retires finished jobs (the
// last member moves into the freed slot).  A job joins a pool only when the
// pool's step counter is a multiple of cs_count, so its dynamic-model steps
// fall where they would alone: every job computes exactly what it computes
// run alone.  Pools are created when no slot is free and the budget allows,
// and empty pools are released when the budget is needed elsewhere.
#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <vector>

#include "case.hpp"

namespace lg {

struct Job {
    lg_params prm;
    long seed;
    double noise;
    int nsteps;
    std::vector<double> disks;      // 8 per turbine
    bool keep;
    int state = 0;                  // 0 queued, 1 running, 2 done
    int pool = -1, slot = -1;
    int steps_done = 0;
    double t_arrive = 0, t_admit = -1, t_done = -1;
    std::vector<std::vector<double>> kept;   // u, v, w, p at completion
};

struct Pool {
    lg_params key;
    std::unique_ptr<Case<double>> c64;
    std::unique_ptr<Case<float>> c32;
    std::vector<int> slot_job;      // job of each stepped slot
    size_t bytes = 0;
    long pts = 0;                   // grid points per member
    template <class F>
    void with(F&& f) {
        if (c64) f(*c64);
        else f(*c32);
    }
    int M() const { return c64 ? c64->M : c32->M; }
    int Ma() const { return c64 ? c64->Ma : c32->Ma; }
};

static bool same_class(const lg_params& a, const lg_params& b) {
    return a.nx == b.nx && a.ny == b.ny && a.Nz == b.Nz && a.lx == b.lx && a.ly == b.ly &&
           a.lz == b.lz && a.dt == b.dt && a.vonk == b.vonk && a.cs_count == b.cs_count &&
           a.use_mean_p_force == b.use_mean_p_force && a.nu == b.nu &&
           a.precision == b.precision && a.initu == b.initu && a.dyn_init == b.dyn_init &&
           a.use_cfl_dt == b.use_cfl_dt && (!a.use_cfl_dt || a.cfl == b.cfl);
}

struct Server {
    double budget;
    int pool_members;
    size_t used = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::vector<Job> jobs;
    std::deque<int> queue;
    std::vector<std::unique_ptr<Pool>> pools;
    double pts_steps = 0;
    long ticks = 0;

    Server(double b, int pm) : budget(b), pool_members(pm) {}

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void release_empty_pools() {
        for (auto& p : pools)
            if (p && p->Ma() == 0) {
                used -= p->bytes;
                p.reset();
            }
    }

    Pool* new_pool(const lg_params& prm) {
        const int ts = prm.precision == LG_FP64 ? 8 : 4;
        const size_t per = case_bytes(prm, 2, ts) - case_bytes(prm, 1, ts);
        const size_t base = case_bytes(prm, 1, ts) - per;
        auto fits = [&](int m) { return used + base + per * (size_t)m <= (size_t)budget; };
        if (!fits(1)) {
            release_empty_pools();
            if (!fits(1)) return nullptr;
        }
        int cap = pool_members;
        while (cap > 1 && !fits(cap)) --cap;
        auto p = std::make_unique<Pool>();
        p->key = prm;
        if (prm.precision == LG_FP64) p->c64 = std::make_unique<Case<double>>(prm, cap);
        else p->c32 = std::make_unique<Case<float>>(prm, cap);
        p->with([](auto& c) { c.Ma = 0; });
        p->bytes = case_bytes(prm, cap, ts);
        p->pts = (long)prm.nx * prm.ny * prm.Nz;
        used += p->bytes;
        for (auto& q : pools)
            if (!q) {
                q = std::move(p);
                return q.get();
            }
        pools.push_back(std::move(p));
        return pools.back().get();
    }

    // Admit job j if a slot of its class is free and aligned; returns true
    // when admitted.
    bool admit(int j) {
        Job& job = jobs[j];
        Pool* target = nullptr;
        bool class_has_room = false;
        for (auto& p : pools) {
            if (!p || !same_class(p->key, job.prm) || p->Ma() >= p->M()) continue;
            class_has_room = true;
            bool aligned = false;
            p->with([&](auto& c) {
                aligned = c.Ma == 0 || c.jt_total % c.d.prm.cs_count == 0;
            });
            if (aligned) {
                target = p.get();
                break;
            }
        }
        if (!target) {
            if (class_has_room) return false;     // wait for the dynamic-model boundary
            target = new_pool(job.prm);
            if (!target) return false;
        }
        const int pi = (int)(std::find_if(pools.begin(), pools.end(),
                                          [&](const std::unique_ptr<Pool>& q) {
                                              return q.get() == target;
                                          }) -
                             pools.begin());
        target->with([&](auto& c) {
            if (c.Ma == 0) {
                c.jt = 0;
                c.jt_total = 0;
            }
            const int m = c.Ma;
            c.set_member_params(m, job.prm.zo, job.prm.mean_p_force_x, job.prm.mean_p_force_y);
            c.seed_member(m, job.seed, job.noise);
            if (!job.disks.empty() || c.mturb[m].n() > 0)
                turbines_init(c, m, (int)job.disks.size() / 8, job.disks.data(), 1.5, 1.5, 1e-2,
                              1);
            c.Ma = m + 1;
            job.slot = m;
        });
        target->slot_job.resize(target->Ma());
        target->slot_job[job.slot] = j;
        job.pool = pi;
        job.state = 1;
        job.t_admit = now();
        return true;
    }

    void retire(Pool& p, int slot) {
        const int j = p.slot_job[slot];
        Job& job = jobs[j];
        p.with([&](auto& c) {
            if (job.keep) {
                c.touch(false);
                using T = typename std::remove_reference<decltype(*c.fld[0])>::type;
                std::vector<T> tmp((size_t)c.d.field);
                for (int id : {LG_U, LG_V, LG_W, LG_P}) {
                    LG_HIP(hipMemcpy(tmp.data(), c.mem(id, slot), c.d.field * sizeof(T),
                                     hipMemcpyDeviceToHost));
                    job.kept.emplace_back(tmp.begin(), tmp.end());
                }
            }
            const int last = c.Ma - 1;
            c.retire_member(slot);
            if (slot != last) {
                p.slot_job[slot] = p.slot_job[last];
                jobs[p.slot_job[slot]].slot = slot;
            }
        });
        p.slot_job.resize(p.Ma());
        job.state = 2;
        job.slot = -1;
        job.t_done = now();
    }

    void tick(double& secs, int& running, int& queued) {
        for (size_t q = queue.size(); q > 0; --q) {
            const int j = queue.front();
            queue.pop_front();
            if (!admit(j)) queue.push_back(j);
        }
        const double a = now();
        for (auto& p : pools)
            if (p && p->Ma() > 0) p->with([](auto& c) { c.step_graph(); });
        for (auto& p : pools)
            if (p && p->Ma() > 0)
                p->with([](auto& c) {
                    LG_HIP(hipStreamSynchronize(c.stream));
                    c.check_device_error();
                });
        secs = now() - a;
        running = 0;
        for (auto& p : pools) {
            if (!p || p->Ma() == 0) continue;
            pts_steps += (double)p->pts * p->Ma();
            for (int s = 0; s < p->Ma(); ++s) ++jobs[p->slot_job[s]].steps_done;
            for (int s = p->Ma() - 1; s >= 0; --s)
                if (jobs[p->slot_job[s]].steps_done >= jobs[p->slot_job[s]].nsteps) retire(*p, s);
            running += p->Ma();
        }
        queued = (int)queue.size();
        ++ticks;
    }
};

}  // namespace lg

struct lg_server {
    lg::Server s;
    lg_server(double b, int pm) : s(b, pm) {}
};

namespace {
template <typename F>
int sguard(F&& f) {
    try {
        f();
        return LG_OK;
    } catch (const lg::ArgError&) {
        return LG_ERR_ARG;
    } catch (const lg::HipError&) {
        return LG_ERR_HIP;
    } catch (const lg::PivotError&) {
        return LG_ERR_PIVOT;
    } catch (const std::bad_alloc&) {
        lg::set_error("host allocation failed");
        return LG_ERR_ALLOC;
    } catch (...) {
        lg::set_error("unknown error");
        return LG_ERR_STATE;
    }
}

#define SREQ(cond, msg)             \
    do {                            \
        if (!(cond)) {              \
            lg::set_error(msg);     \
            throw lg::ArgError();   \
        }                           \
    } while (0)
}  // namespace

extern "C" {

int lg_case_bytes(const lg_params* p, int members, double* bytes) {
    return sguard([&] {
        SREQ(p && bytes && members >= 1, "bad argument");
        *bytes = (double)lg::case_bytes(*p, members, p->precision == LG_FP64 ? 8 : 4);
    });
}

int lg_mem_info(double* free_bytes, double* total_bytes) {
    return sguard([&] {
        size_t f = 0, t = 0;
        LG_HIP(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = (double)f;
        if (total_bytes) *total_bytes = (double)t;
    });
}

int lg_server_create(double budget_bytes, int pool_members, lg_server** out) {
    return sguard([&] {
        SREQ(out && budget_bytes > 0 && pool_members >= 1, "bad argument");
        *out = new lg_server(budget_bytes, pool_members);
    });
}

int lg_server_destroy(lg_server* s) {
    if (!s) return LG_OK;
    const int rc = sguard([&] { s->s.pools.clear(); });
    delete s;
    return rc;
}

int lg_server_now(lg_server* s, double* t) {
    if (!s || !t) return LG_ERR_ARG;
    *t = s->s.now();
    return LG_OK;
}

int lg_server_submit(lg_server* s, const lg_params* prm, long seed, double noise, int nsteps,
                     int nturb, const double* spec8, double t_arrive, int keep_fields,
                     int* job) {
    return sguard([&] {
        SREQ(s && prm && job, "null argument");
        SREQ(nsteps >= 1, "nsteps must be >= 1");
        SREQ(nturb >= 0 && (nturb == 0 || spec8), "bad turbine list");
        SREQ(prm->nx >= 4 && prm->ny >= 4 && prm->nx % 4 == 0 && prm->ny % 4 == 0 &&
                 prm->Nz >= 3 && prm->zo > 0 && prm->dt > 0 && prm->cs_count >= 1 &&
                 (!prm->use_cfl_dt || prm->cfl > 0),
             "bad job parameters");
        lg::Job j;
        j.prm = *prm;
        j.seed = seed;
        j.noise = noise;
        j.nsteps = nsteps;
        j.disks.assign(spec8, spec8 + 8 * nturb);
        j.keep = keep_fields != 0;
        j.t_arrive = t_arrive;
        s->s.jobs.push_back(std::move(j));
        *job = (int)s->s.jobs.size() - 1;
        s->s.queue.push_back(*job);
    });
}

int lg_server_tick(lg_server* s, double* tick_seconds, int* running, int* queued) {
    return sguard([&] {
        SREQ(s, "null server");
        double t = 0;
        int r = 0, q = 0;
        s->s.tick(t, r, q);
        if (tick_seconds) *tick_seconds = t;
        if (running) *running = r;
        if (queued) *queued = q;
    });
}

int lg_server_job(lg_server* s, int job, int* state, int* steps_done, double* t_arrive,
                  double* t_admit, double* t_done) {
    return sguard([&] {
        SREQ(s && job >= 0 && job < (int)s->s.jobs.size(), "job out of range");
        const lg::Job& j = s->s.jobs[job];
        if (state) *state = j.state;
        if (steps_done) *steps_done = j.steps_done;
        if (t_arrive) *t_arrive = j.t_arrive;
        if (t_admit) *t_admit = j.t_admit;
        if (t_done) *t_done = j.t_done;
    });
}

int lg_server_result(lg_server* s, int job, int field, double* host) {
    return sguard([&] {
        SREQ(s && host && job >= 0 && job < (int)s->s.jobs.size(), "bad argument");
        const lg::Job& j = s->s.jobs[job];
        SREQ(j.state == 2 && j.keep, "job not finished or its fields were not kept");
        int q = -1;
        if (field == LG_U) q = 0;
        if (field == LG_V) q = 1;
        if (field == LG_W) q = 2;
        if (field == LG_P) q = 3;
        SREQ(q >= 0, "kept fields are u, v, w and p");
        std::copy(j.kept[q].begin(), j.kept[q].end(), host);
    });
}

int lg_server_stats(lg_server* s, double* pts_steps, long* ticks, double* bytes_used,
                    int* npools) {
    return sguard([&] {
        SREQ(s, "null server");
        if (pts_steps) *pts_steps = s->s.pts_steps;
        if (ticks) *ticks = s->s.ticks;
        if (bytes_used) *bytes_used = (double)s->s.used;
        int n = 0;
        for (auto& p : s->s.pools) n += p ? 1 : 0;
        if (npools) *npools = n;
    });
}

}  // extern "C"

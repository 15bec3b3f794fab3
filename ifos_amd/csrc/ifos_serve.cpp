// ifos_serve.cpp -- request server over the native engine (include/ifos2d_serve.h).
//
// Host-only C++: a queue, a pool of resident engines counted against the
// HBM budget, and worker threads.  Each worker takes the oldest admissible
// request, reserves (or reuses) an engine for it, and runs its shots in
// groups of up to `slots` per launch on that engine's own stream.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ifos2d_serve.h"

extern "C" void ifos_set_error(const char *fmt, ...);

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    long long id = 0;
    ifos_config cfg;
    int kind = 0, snap_prec = 0, nshots = 0;
    std::vector<int> nsrc, ntr, recpos, kill;
    std::vector<float> srcpos, vp, vs, rho, taus, taup, obs_vx, obs_vy;
    long long model_id = 0;
    int slots = 1;
    size_t bytes = 0;
    int overtaken = 0;
    // outputs
    std::vector<float> seis_vx, seis_vy, g;
    std::vector<double> misfit;
    double t_submit = 0, t_start = 0, t_end = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int reused = 0;
    std::string error;
};

struct Resident {
    ifos_engine *eng = nullptr;
    ifos_config cfg;
    int slots = 0, snap_prec = 0, snapshots = 0, ntr_cap = 0, nsrc_cap = 0;
    long long model_id = 0;
    size_t bytes = 0;
    bool busy = false;
    long long last_use = 0;
};

bool same_config(const ifos_config &a, const ifos_config &b) { return memcmp(&a, &b, sizeof a) == 0; }

}  // namespace

struct ifos_server {
    int device = 0, max_slots = 1;
    size_t budget = 0, in_use = 0, peak = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Req>> queue;
    std::map<long long, std::shared_ptr<Req>> all;
    std::vector<std::unique_ptr<Resident>> pool;
    std::vector<std::thread> workers;
    bool stop = false;
    long long next_id = 1, tick = 0;
    ifos_server_stats st{};
    Clock::time_point t0 = Clock::now();

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    // engine bytes of request r at its slot count
    size_t bytes_of(const Req &r) const
    {
        int ntr = 0, nsrc = 0;
        // traces and sources of the largest group of `slots` consecutive shots
        for (int k = 0; k < r.nshots; k += r.slots) {
            int t = 0, s = 0;
            for (int q = k; q < std::min(r.nshots, k + r.slots); q++) {
                t += r.ntr[q];
                s += r.nsrc[q];
            }
            ntr = std::max(ntr, t);
            nsrc = std::max(nsrc, s);
        }
        return ifos_engine_bytes(&r.cfg, r.slots, r.snap_prec, r.kind == IFOS_REQ_GRADIENT, ntr, nsrc);
    }

    // under mu: a resident engine for r (reused or newly reserved), or null
    Resident *admit(Req &r)
    {
        const int snapshots = r.kind == IFOS_REQ_GRADIENT;
        for (auto &p : pool)
            if (!p->busy && p->eng && same_config(p->cfg, r.cfg) && p->slots == r.slots && p->snapshots >= snapshots &&
                (!snapshots || p->snap_prec == r.snap_prec) && p->bytes >= r.bytes) {
                p->busy = true;
                p->last_use = ++tick;
                r.reused = 1;
                st.engines_reused++;
                return p.get();
            }
        // evict idle engines, least recently used first, until r fits; when
        // even evicting every idle engine would not make room, keep them all
        size_t idle = 0;
        for (auto &p : pool)
            if (!p->busy && p->eng) idle += p->bytes;
        if (in_use - idle + r.bytes > budget) return nullptr;
        while (in_use + r.bytes > budget) {
            Resident *lru = nullptr;
            for (auto &p : pool)
                if (!p->busy && p->eng && (!lru || p->last_use < lru->last_use)) lru = p.get();
            if (!lru) return nullptr;
            ifos_engine_destroy(lru->eng);
            lru->eng = nullptr;
            in_use -= lru->bytes;
            st.engines_evicted++;
            pool.erase(std::find_if(pool.begin(), pool.end(), [&](auto &p) { return p.get() == lru; }));
        }
        auto res = std::make_unique<Resident>();
        res->cfg = r.cfg;
        res->slots = r.slots;
        res->snap_prec = r.snap_prec;
        res->snapshots = snapshots;
        res->bytes = r.bytes;
        res->busy = true;
        res->last_use = ++tick;
        in_use += r.bytes;
        peak = std::max(peak, in_use);
        pool.push_back(std::move(res));
        return pool.back().get();
    }

    // under mu: the next request to run and its engine slot, or null
    std::shared_ptr<Req> pick(Resident **where)
    {
        for (size_t q = 0; q < queue.size(); q++) {
            auto r = queue[q];
            Resident *res = admit(*r);
            if (res) {
                for (size_t k = 0; k < q; k++) queue[k]->overtaken++;
                queue.erase(queue.begin() + q);
                *where = res;
                return r;
            }
            if (r->overtaken >= 8) break;  // the oldest waiter goes next
        }
        return nullptr;
    }

    int run(Req &r, Resident &res)
    {
        // a fresh engine for a newly reserved slot
        if (!res.eng) {
            const int rc = res.slots > 1 ? ifos_engine_create_batch(&r.cfg, device, res.slots, &res.eng)
                                         : ifos_engine_create(&r.cfg, device, &res.eng);
            if (rc) return rc;
            std::lock_guard<std::mutex> lk(mu);
            st.engines_created++;
        }
        ifos_engine *E = res.eng;
        if (r.kind == IFOS_REQ_GRADIENT) {
            int rc = ifos_engine_set_snapshot_precision(E, r.snap_prec);
            if (rc) return rc;
        }
        if (!(r.model_id && res.model_id == r.model_id)) {
            res.model_id = 0;
            const int rc = ifos_engine_set_model(E, r.vp.data(), r.vs.data(), r.rho.data(),
                                                 r.taus.empty() ? nullptr : r.taus.data(),
                                                 r.taup.empty() ? nullptr : r.taup.data());
            if (rc) return rc;
            res.model_id = r.model_id;
        }
        const int NT = r.cfg.NT;
        int ntr_all = 0;
        for (int k = 0; k < r.nshots; k++) ntr_all += r.ntr[k];
        if (r.kind == IFOS_REQ_FORWARD) {
            r.seis_vx.assign((size_t)ntr_all * NT, 0.0f);
            r.seis_vy.assign((size_t)ntr_all * NT, 0.0f);
        } else {
            r.g.assign(3 * (size_t)r.cfg.NX * r.cfg.NY, 0.0f);
            r.misfit.assign(2 * (size_t)r.nshots, 0.0);
            if (int rc = ifos_engine_reset_gradient(E)) return rc;
        }
        int src0 = 0, tr0 = 0;
        for (int k = 0; k < r.nshots; k += r.slots) {
            const int n = std::min(r.slots, r.nshots - k);
            int ns = 0, nt = 0;
            for (int q = k; q < k + n; q++) {
                ns += r.nsrc[q];
                nt += r.ntr[q];
            }
            int rc = ifos_engine_set_shots(E, n, r.nsrc.data() + k, r.srcpos.data() + (size_t)src0 * IFOS_SRC_STRIDE,
                                           r.ntr.data() + k, r.recpos.data() + 2 * (size_t)tr0);
            if (rc) return rc;
            if (r.kind == IFOS_REQ_FORWARD) {
                if ((rc = ifos_engine_forward(E, 0))) return rc;
                if ((rc = ifos_engine_seismograms(E, r.seis_vx.data() + (size_t)tr0 * NT,
                                                  r.seis_vy.data() + (size_t)tr0 * NT)))
                    return rc;
            } else {
                double mis[2];
                rc = ifos_engine_shot_gradient(E, r.obs_vx.data() + (size_t)tr0 * NT,
                                               r.obs_vy.data() + (size_t)tr0 * NT,
                                               r.kill.empty() ? nullptr : r.kill.data() + tr0, 0, mis);
                if (rc) return rc;
                if ((rc = ifos_engine_shot_misfit(E, r.misfit.data() + 2 * (size_t)k))) return rc;
            }
            src0 += ns;
            tr0 += nt;
        }
        if (r.kind == IFOS_REQ_GRADIENT) {
            const size_t plane = (size_t)r.cfg.NX * r.cfg.NY;
            if (int rc = ifos_engine_gradient(E, 1, r.g.data(), r.g.data() + plane, r.g.data() + 2 * plane)) return rc;
        }
        return 0;
    }

    void worker()
    {
        for (;;) {
            std::shared_ptr<Req> r;
            Resident *res = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    r = pick(&res);
                    return r != nullptr;
                });
                if (!r) return;  // stopping with nothing admitted
                r->state = 1;
                r->t_start = now();
            }
            const int rc = run(*r, *res);
            std::string err = rc ? ifos_last_error() : "";
            {
                std::lock_guard<std::mutex> lk(mu);
                res->busy = false;
                if (rc) {
                    // a failed engine is not reused
                    if (res->eng) ifos_engine_destroy(res->eng);
                    res->eng = nullptr;
                    in_use -= res->bytes;
                    pool.erase(std::find_if(pool.begin(), pool.end(), [&](auto &p) { return p.get() == res; }));
                }
                r->t_end = now();
                r->state = rc ? 3 : 2;
                r->error = err;
                if (rc) st.failed++;
                else st.finished++;
                // its inputs are no longer needed
                std::vector<float>().swap(r->vp);
                std::vector<float>().swap(r->vs);
                std::vector<float>().swap(r->rho);
                std::vector<float>().swap(r->taus);
                std::vector<float>().swap(r->taup);
                std::vector<float>().swap(r->obs_vx);
                std::vector<float>().swap(r->obs_vy);
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" int ifos_server_create(int device, size_t hbm_budget, int workers, int max_slots, ifos_server **out)
{
    *out = nullptr;
    if (workers < 1 || workers > 64) return ifos_set_error("workers must be 1..64"), -2;
    if (max_slots < 1 || max_slots > 256) return ifos_set_error("max_slots must be 1..256"), -2;
    if (hbm_budget == 0) return ifos_set_error("the HBM budget must be positive"), -2;
    auto *s = new ifos_server();
    s->device = device;
    s->budget = hbm_budget;
    s->max_slots = max_slots;
    s->st.budget = hbm_budget;
    for (int k = 0; k < workers; k++) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return 0;
}

extern "C" void ifos_server_destroy(ifos_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto &r : s->queue) {
            r->state = 3;
            r->error = "server stopped";
        }
        s->queue.clear();
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &p : s->pool)
        if (p->eng) ifos_engine_destroy(p->eng);
    delete s;
}

extern "C" int ifos_server_submit(ifos_server *s, const ifos_request *q, long long *id)
{
    if (q->nshots < 1) return ifos_set_error("a request needs at least one shot"), -2;
    if (q->kind != IFOS_REQ_FORWARD && q->kind != IFOS_REQ_GRADIENT) return ifos_set_error("unknown request kind"), -2;
    if (q->kind == IFOS_REQ_GRADIENT && (!q->obs_vx || !q->obs_vy))
        return ifos_set_error("a gradient request needs observed traces"), -2;
    if (!q->vp || !q->vs || !q->rho || (q->cfg.L && (!q->taus || !q->taup)))
        return ifos_set_error("the request's model is incomplete"), -2;
    auto r = std::make_shared<Req>();
    r->cfg = q->cfg;
    r->kind = q->kind;
    r->snap_prec = q->snap_prec;
    r->nshots = q->nshots;
    r->model_id = q->model_id;
    r->nsrc.assign(q->nsrc, q->nsrc + q->nshots);
    r->ntr.assign(q->ntr, q->ntr + q->nshots);
    int ns = 0, nt = 0;
    for (int k = 0; k < q->nshots; k++) {
        if (r->nsrc[k] < 0 || r->ntr[k] < 0) return ifos_set_error("negative source or receiver count"), -2;
        ns += r->nsrc[k];
        nt += r->ntr[k];
    }
    r->srcpos.assign(q->srcpos, q->srcpos + (size_t)ns * IFOS_SRC_STRIDE);
    r->recpos.assign(q->recpos, q->recpos + 2 * (size_t)nt);
    const size_t plane = (size_t)q->cfg.NX * q->cfg.NY;
    r->vp.assign(q->vp, q->vp + plane);
    r->vs.assign(q->vs, q->vs + plane);
    r->rho.assign(q->rho, q->rho + plane);
    if (q->cfg.L) {
        r->taus.assign(q->taus, q->taus + plane);
        r->taup.assign(q->taup, q->taup + plane);
    }
    if (q->kind == IFOS_REQ_GRADIENT) {
        r->obs_vx.assign(q->obs_vx, q->obs_vx + (size_t)nt * q->cfg.NT);
        r->obs_vy.assign(q->obs_vy, q->obs_vy + (size_t)nt * q->cfg.NT);
        if (q->kill) r->kill.assign(q->kill, q->kill + nt);
    }
    r->slots = std::min(s->max_slots, q->nshots);
    // fewer slots when the request's engine would not fit in the budget
    for (;;) {
        r->bytes = s->bytes_of(*r);
        if (r->bytes == 0) return ifos_set_error("invalid configuration in the request"), -2;
        if (r->bytes <= s->budget || r->slots == 1) break;
        r->slots = (r->slots + 1) / 2;
    }
    if (r->bytes > s->budget)
        return ifos_set_error("request needs %.2f GB, more than the %.2f GB budget", r->bytes / 1e9, s->budget / 1e9), -3;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        r->t_submit = s->now();
        s->queue.push_back(r);
        s->all[r->id] = r;
        s->st.submitted++;
        *id = r->id;
    }
    s->cv_work.notify_all();
    return 0;
}

extern "C" int ifos_server_wait(ifos_server *s, long long id, double timeout_s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return ifos_set_error("unknown request %lld", id), -2;
    auto r = it->second;
    auto done = [&] { return r->state >= 2; };
    if (timeout_s < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return 1;
    if (r->state == 3) return ifos_set_error("request %lld failed: %s", id, r->error.c_str()), -5;
    return 0;
}

extern "C" int ifos_server_result(ifos_server *s, long long id, ifos_result *out)
{
    std::shared_ptr<Req> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->all.find(id);
        if (it == s->all.end()) return ifos_set_error("unknown request %lld", id), -2;
        r = it->second;
        if (r->state < 2) return ifos_set_error("request %lld has not finished", id), -2;
        s->all.erase(it);
    }
    if (r->state == 3) return ifos_set_error("request %lld failed: %s", id, r->error.c_str()), -5;
    auto put = [](float *dst, const std::vector<float> &v) {
        if (dst && !v.empty()) memcpy(dst, v.data(), sizeof(float) * v.size());
    };
    put(out->seis_vx, r->seis_vx);
    put(out->seis_vy, r->seis_vy);
    const size_t plane = (size_t)r->cfg.NX * r->cfg.NY;
    if (!r->g.empty()) {
        if (out->g_vp) memcpy(out->g_vp, r->g.data(), sizeof(float) * plane);
        if (out->g_vs) memcpy(out->g_vs, r->g.data() + plane, sizeof(float) * plane);
        if (out->g_rho) memcpy(out->g_rho, r->g.data() + 2 * plane, sizeof(float) * plane);
    }
    if (out->misfit && !r->misfit.empty()) memcpy(out->misfit, r->misfit.data(), sizeof(double) * r->misfit.size());
    out->times[0] = r->t_submit;
    out->times[1] = r->t_start;
    out->times[2] = r->t_end;
    out->slots = r->slots;
    out->reused = r->reused;
    return 0;
}

extern "C" int ifos_server_stats_get(ifos_server *s, ifos_server_stats *out)
{
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->bytes_in_use = s->in_use;
    out->peak_bytes = s->peak;
    return 0;
}

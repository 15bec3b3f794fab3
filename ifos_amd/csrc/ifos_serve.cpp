// ifos_serve.cpp -- request server over the native engine (include/ifos2d_serve.h).
//
// Host-only C++: a queue of requests, a pool of resident engines counted
// against the HBM budget, and worker threads.  The unit of work is a launch:
// up to `slots` shots advanced together on one engine.  A worker takes the
// oldest request with shots left whose launch fits, fills the launch with
// that request's next shots and then with the next shots of later queued
// requests of the same configuration, kind, snapshot precision and model,
// and runs it on its engine's own stream.  Several launches (of one request
// or of different ones) run at once on different workers.
//
// A request's result does not depend on how its shots were grouped: forward
// traces are per shot, and gradients are kept per shot
// (ifos_engine_slot_gradient) and added to the request's sum in shot order,
// in FP32, exactly as an engine running the request alone adds them.
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
    std::vector<int> src0, tr0;  // first source row and trace of each shot
    std::vector<float> srcpos, vp, vs, rho, taus, taup, obs_vx, obs_vy;
    long long model_id = 0;
    int cap = 1;                 // shots per launch: max_slots, at most max(2, MALL-sized count)
    int slots = 1;               // shots per launch with which it fits the budget alone
    int overtaken = 0;
    int next_shot = 0, done_shots = 0, folded = 0;
    int inflight = 0;            // launches holding shots of this request
    // outputs
    std::vector<float> seis_vx, seis_vy, g;
    std::vector<double> misfit;
    std::map<int, std::vector<float>> pending;  // finished shots' gradients not yet in the sum
    double t_submit = 0, t_start = 0, t_end = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int reused = 0, max_launch = 0, shared = 0;
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

// shots (request, shot index) advanced together on one engine
struct Launch {
    std::vector<std::pair<std::shared_ptr<Req>, int>> shots;
    int ntr = 0, nsrc = 0;
    bool mixed = false;
};

bool same_config(const ifos_config &a, const ifos_config &b) { return memcmp(&a, &b, sizeof a) == 0; }

// requests whose shots may share a launch: same grid and parameters, kind,
// snapshot precision and model (a model id of 0 is never shared)
bool shareable(const Req &a, const Req &b)
{
    return a.model_id != 0 && a.model_id == b.model_id && a.kind == b.kind &&
           (a.kind != IFOS_REQ_GRADIENT || a.snap_prec == b.snap_prec) && same_config(a.cfg, b.cfg);
}

}  // namespace

struct ifos_server {
    int device = 0, max_slots = 1;
    size_t budget = 0, in_use = 0, peak = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Req>> queue;  // requests with shots not yet taken
    std::map<long long, std::shared_ptr<Req>> all;
    std::vector<std::unique_ptr<Resident>> pool;
    std::vector<std::thread> workers;
    bool stop = false;
    long long next_id = 1, tick = 0;
    ifos_server_stats st{};
    Clock::time_point t0 = Clock::now();

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    static size_t engine_bytes(const Req &r, int slots, int ntr, int nsrc)
    {
        return ifos_engine_bytes(&r.cfg, slots, r.snap_prec, r.kind == IFOS_REQ_GRADIENT, ntr, nsrc);
    }

    // bytes of the largest launch of `slots` consecutive shots of r alone
    static size_t alone_bytes(const Req &r, int slots)
    {
        int ntr = 0, nsrc = 0;
        for (int k = 0; k < r.nshots; k += slots) {
            int t = 0, s = 0;
            for (int q = k; q < std::min(r.nshots, k + slots); q++) {
                t += r.ntr[q];
                s += r.nsrc[q];
            }
            ntr = std::max(ntr, t);
            nsrc = std::max(nsrc, s);
        }
        return engine_bytes(r, slots, ntr, nsrc);
    }

    // under mu: an engine for launch L led by request r (an idle resident one
    // of the same kind with room, possibly grown, or a new one), or null
    Resident *admit(const Req &r, const Launch &L)
    {
        const int n = (int)L.shots.size();
        const int snapshots = r.kind == IFOS_REQ_GRADIENT;
        Resident *best = nullptr;
        size_t best_grow = 0;
        for (auto &p : pool) {
            if (p->busy || !p->eng || !same_config(p->cfg, r.cfg) || p->slots < n || p->snapshots < snapshots ||
                (snapshots && p->snap_prec != r.snap_prec))
                continue;
            const int ntr = std::max(p->ntr_cap, L.ntr), nsrc = std::max(p->nsrc_cap, L.nsrc);
            const size_t need = ifos_engine_bytes(&p->cfg, p->slots, p->snap_prec, p->snapshots, ntr, nsrc);
            const size_t grow = need > p->bytes ? need - p->bytes : 0;
            if (in_use + grow > budget) continue;
            // prefer the engine closest in size, then one that needs no growth
            if (!best || p->slots < best->slots || (p->slots == best->slots && grow < best_grow)) {
                best = p.get();
                best_grow = grow;
            }
        }
        if (best) {
            best->ntr_cap = std::max(best->ntr_cap, L.ntr);
            best->nsrc_cap = std::max(best->nsrc_cap, L.nsrc);
            best->bytes += best_grow;
            in_use += best_grow;
            peak = std::max(peak, in_use);
            best->busy = true;
            best->last_use = ++tick;
            st.engines_reused++;
            return best;
        }
        const size_t need = engine_bytes(r, n, L.ntr, L.nsrc);
        // evict idle engines, least recently used first, until the launch
        // fits; when even evicting every idle engine would not make room,
        // keep them all
        size_t idle = 0;
        for (auto &p : pool)
            if (!p->busy && p->eng) idle += p->bytes;
        if (in_use - idle + need > budget) return nullptr;
        while (in_use + need > budget) {
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
        res->slots = n;
        res->snap_prec = r.snap_prec;
        res->snapshots = snapshots;
        res->ntr_cap = L.ntr;
        res->nsrc_cap = L.nsrc;
        res->bytes = need;
        res->busy = true;
        res->last_use = ++tick;
        in_use += need;
        peak = std::max(peak, in_use);
        pool.push_back(std::move(res));
        return pool.back().get();
    }

    // under mu: the launch led by queued request q, up to max_slots shots
    // and shrunk until it fits (down to one shot), with its engine; false
    // when nothing fits
    bool plan(size_t q, Launch &L, Resident **where)
    {
        const auto &r = queue[q];
        for (int cap = r->cap;; cap = (cap + 1) / 2) {
            L = Launch();
            // the lead request's next shots, then later requests' of the same kind
            for (size_t k = q; k < queue.size() && (int)L.shots.size() < cap; k++) {
                const auto &o = queue[k];
                if (k != q && !shareable(*r, *o)) continue;
                for (int s = o->next_shot; s < o->nshots && (int)L.shots.size() < cap; s++) {
                    L.shots.emplace_back(o, s);
                    L.ntr += o->ntr[s];
                    L.nsrc += o->nsrc[s];
                    if (k != q) L.mixed = true;
                }
            }
            if ((*where = admit(*r, L))) return true;
            if (cap == 1) return false;
        }
    }

    // under mu: the next launch to run and its engine, or false
    bool pick(Launch &L, Resident **where)
    {
        for (size_t q = 0; q < queue.size(); q++) {
            if (plan(q, L, where)) {
                for (size_t k = 0; k < q; k++) queue[k]->overtaken++;
                // take the shots; requests with no shots left leave the queue
                for (auto &[r, s] : L.shots) {
                    r->next_shot = s + 1;
                    if (r->state == 0) {
                        r->state = 1;
                        r->t_start = now();
                        r->reused = (*where)->eng != nullptr;
                    }
                    int mine = 0;
                    for (auto &x : L.shots) mine += x.first == r;
                    r->max_launch = std::max(r->max_launch, mine);
                }
                std::shared_ptr<Req> last;
                for (auto &[r, s] : L.shots)
                    if (r != last) {
                        if (L.mixed) r->shared++;
                        r->inflight++;
                        last = r;
                    }
                queue.erase(std::remove_if(queue.begin(), queue.end(),
                                           [](const std::shared_ptr<Req> &r) { return r->next_shot >= r->nshots; }),
                            queue.end());
                st.launches++;
                if (L.mixed) st.shared_launches++;
                return true;
            }
            if (queue[q]->overtaken >= 8) break;  // the oldest waiter goes next
        }
        return false;
    }

    // outside mu: run launch L on engine res; per-shot outputs go to `grads`
    int run(const Launch &L, Resident &res, std::vector<std::vector<float>> &grads, std::vector<double> &mis,
            std::vector<float> &svx, std::vector<float> &svy)
    {
        const Req &lead = *L.shots[0].first;
        const int n = (int)L.shots.size();
        if (!res.eng) {
            const int rc = res.slots > 1 ? ifos_engine_create_batch(&lead.cfg, device, res.slots, &res.eng)
                                         : ifos_engine_create(&lead.cfg, device, &res.eng);
            if (rc) return rc;
            std::lock_guard<std::mutex> lk(mu);
            st.engines_created++;
        }
        ifos_engine *E = res.eng;
        if (lead.kind == IFOS_REQ_GRADIENT)
            if (int rc = ifos_engine_set_snapshot_precision(E, lead.snap_prec)) return rc;
        if (!(lead.model_id && res.model_id == lead.model_id)) {
            res.model_id = 0;
            const int rc = ifos_engine_set_model(E, lead.vp.data(), lead.vs.data(), lead.rho.data(),
                                                 lead.taus.empty() ? nullptr : lead.taus.data(),
                                                 lead.taup.empty() ? nullptr : lead.taup.data());
            if (rc) return rc;
            res.model_id = lead.model_id;
        }
        const int NT = lead.cfg.NT;
        std::vector<int> nsrc(n), ntr(n), recpos, kill;
        std::vector<float> srcpos, ovx, ovy;
        bool any_kill = false;
        for (int k = 0; k < n; k++) {
            const Req &r = *L.shots[k].first;
            const int s = L.shots[k].second;
            nsrc[k] = r.nsrc[s];
            ntr[k] = r.ntr[s];
            srcpos.insert(srcpos.end(), r.srcpos.begin() + (size_t)r.src0[s] * IFOS_SRC_STRIDE,
                          r.srcpos.begin() + (size_t)(r.src0[s] + r.nsrc[s]) * IFOS_SRC_STRIDE);
            recpos.insert(recpos.end(), r.recpos.begin() + 2 * (size_t)r.tr0[s],
                          r.recpos.begin() + 2 * (size_t)(r.tr0[s] + r.ntr[s]));
            if (r.kind == IFOS_REQ_GRADIENT) {
                const size_t a = (size_t)r.tr0[s] * NT, b = (size_t)(r.tr0[s] + r.ntr[s]) * NT;
                ovx.insert(ovx.end(), r.obs_vx.begin() + a, r.obs_vx.begin() + b);
                ovy.insert(ovy.end(), r.obs_vy.begin() + a, r.obs_vy.begin() + b);
                if (r.kill.empty()) kill.insert(kill.end(), r.ntr[s], 0);
                else {
                    kill.insert(kill.end(), r.kill.begin() + r.tr0[s], r.kill.begin() + r.tr0[s] + r.ntr[s]);
                    any_kill = true;
                }
            }
        }
        int rc = ifos_engine_set_shots(E, n, nsrc.data(), srcpos.data(), ntr.data(), recpos.data());
        if (rc) return rc;
        if (lead.kind == IFOS_REQ_FORWARD) {
            svx.assign((size_t)L.ntr * NT, 0.0f);
            svy.assign((size_t)L.ntr * NT, 0.0f);
            if ((rc = ifos_engine_forward(E, 0))) return rc;
            return ifos_engine_seismograms(E, svx.data(), svy.data());
        }
        double m2[2];
        if ((rc = ifos_engine_shot_gradient(E, ovx.data(), ovy.data(), any_kill ? kill.data() : nullptr, 0, m2)))
            return rc;
        mis.assign(2 * (size_t)n, 0.0);
        if ((rc = ifos_engine_shot_misfit(E, mis.data()))) return rc;
        const size_t plane = (size_t)lead.cfg.NX * lead.cfg.NY;
        grads.assign(n, std::vector<float>());
        for (int k = 0; k < n; k++) {
            grads[k].resize(3 * plane);
            if ((rc = ifos_engine_slot_gradient(E, k, grads[k].data(), grads[k].data() + plane,
                                                grads[k].data() + 2 * plane)))
                return rc;
        }
        return 0;
    }

    // under mu: a request has all its shots in, or has failed
    void finish(Req &r, int state)
    {
        r.t_end = now();
        r.state = state;
        if (state == 3) st.failed++;
        else st.finished++;
    }

    // under mu: inputs of a finished request that no launch reads any more
    static void release(Req &r)
    {
        std::vector<float>().swap(r.vp);
        std::vector<float>().swap(r.vs);
        std::vector<float>().swap(r.rho);
        std::vector<float>().swap(r.taus);
        std::vector<float>().swap(r.taup);
        std::vector<float>().swap(r.obs_vx);
        std::vector<float>().swap(r.obs_vy);
        r.pending.clear();
    }

    // under mu: store launch L's per-shot outputs in their requests
    void deliver(const Launch &L, int rc, const std::string &err, std::vector<std::vector<float>> &grads,
                 const std::vector<double> &mis, const std::vector<float> &svx, const std::vector<float> &svy)
    {
        size_t t = 0;
        for (size_t k = 0; k < L.shots.size(); k++) {
            Req &r = *L.shots[k].first;
            const int s = L.shots[k].second;
            const int NT = r.cfg.NT;
            if (r.state == 3) continue;
            if (rc) {
                r.error = err;
                // shots of a failed request still queued are dropped
                r.next_shot = r.nshots;
                queue.erase(std::remove_if(queue.begin(), queue.end(),
                                           [&](const std::shared_ptr<Req> &q) { return q.get() == &r; }),
                            queue.end());
                finish(r, 3);
                continue;
            }
            if (r.kind == IFOS_REQ_FORWARD) {
                const size_t nv = (size_t)r.ntr[s] * NT;
                memcpy(r.seis_vx.data() + (size_t)r.tr0[s] * NT, svx.data() + t * NT, sizeof(float) * nv);
                memcpy(r.seis_vy.data() + (size_t)r.tr0[s] * NT, svy.data() + t * NT, sizeof(float) * nv);
                t += r.ntr[s];
            } else {
                r.misfit[2 * (size_t)s] = mis[2 * k];
                r.misfit[2 * (size_t)s + 1] = mis[2 * k + 1];
                r.pending[s] = std::move(grads[k]);
                // the sum adds shots in order: 0 + g0, then + g1, ... in FP32
                for (auto it = r.pending.find(r.folded); it != r.pending.end(); it = r.pending.find(r.folded)) {
                    const std::vector<float> &g = it->second;
                    for (size_t e = 0; e < r.g.size(); e++) r.g[e] = r.g[e] + g[e];
                    r.pending.erase(it);
                    r.folded++;
                }
            }
            if (++r.done_shots == r.nshots) finish(r, 2);
        }
        // each request of the launch once: one launch fewer reads its inputs
        std::shared_ptr<Req> last;
        for (auto &sh : L.shots)
            if (sh.first != last) {
                last = sh.first;
                if (--last->inflight == 0 && last->state >= 2) release(*last);
            }
    }

    void worker()
    {
        for (;;) {
            Launch L;
            Resident *res = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                bool got = false;
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    got = pick(L, &res);
                    return got;
                });
                if (!got) return;  // stopping with nothing admitted
            }
            std::vector<std::vector<float>> grads;
            std::vector<double> mis;
            std::vector<float> svx, svy;
            const int rc = run(L, *res, grads, mis, svx, svy);
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
                deliver(L, rc, err, grads, mis, svx, svy);
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
            if (r->state >= 2) continue;
            r->error = "server stopped";
            r->state = 3;
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
        r->src0.push_back(ns);
        r->tr0.push_back(nt);
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
        r->g.assign(3 * plane, 0.0f);
        r->misfit.assign(2 * (size_t)q->nshots, 0.0);
    } else {
        r->seis_vx.assign((size_t)nt * q->cfg.NT, 0.0f);
        r->seis_vy.assign((size_t)nt * q->cfg.NT, 0.0f);
    }
    // launches wider than the MALL holds turn DRAM-bound and slower per shot
    // (on a 2000x800 grid: 2 shots per launch 2.57, 1 shot 2.38, 4 shots 2.13
    // shot-gradients/s); two are allowed where one already exceeds it
    r->cap = std::min(s->max_slots, std::max(2, ifos_engine_mall_slots(&q->cfg)));
    r->slots = std::min(r->cap, q->nshots);
    // fewer shots per launch when the request's engine would not fit in the budget
    size_t bytes = 0;
    for (;;) {
        bytes = s->alone_bytes(*r, r->slots);
        if (bytes == 0) return ifos_set_error("invalid configuration in the request"), -2;
        if (bytes <= s->budget || r->slots == 1) break;
        r->slots = (r->slots + 1) / 2;
    }
    if (bytes > s->budget)
        return ifos_set_error("request needs %.2f GB, more than the %.2f GB budget", bytes / 1e9, s->budget / 1e9), -3;
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
    out->slots = r->max_launch;
    out->reused = r->reused;
    out->shared = r->shared;
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

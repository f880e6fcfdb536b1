// Native serving loop of libgsx: continuous batching of single-patch
// Poisson requests on one GPU (host code only, on top of the C ABI).
//
// The reference has no serving layer: it assembles and solves one problem at
// a time (src/gsAssembler/gsExprAssembler.h:754-833,
// src/gsSolver/gsConjugateGradient.hpp:20-93).  Here requests arrive at any
// time and share the device:
//
//   assembler thread (its own context): admits requests in arrival order while
//     their estimated device bytes fit the budget; a request for the
//     multigrid lane is handed on, any other is assembled (its structured
//     operator included) with its right-hand sides uploaded;
//   solver thread (its own context): adds one box member per right-hand
//     side to the running gsx_batch (tile shape from the rules given with
//     gsx_server_set_tile), steps the batch, and completes a request when all
//     its members have finished: solutions to the caller's host buffers,
//     operator and workspace released, budget returned;
//   multigrid threads (mg_workers of them, each with its own context): each
//     takes a request, assembles it with the element matrices kept, builds
//     the hierarchy once and solves every right-hand side by MG-preconditioned
//     CG with it resident.
//
// A member's iterates depend only on its own rows and tile shape, so a
// request's solutions are bitwise the same alone and in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "gsx.h"

namespace {

struct TileRule {
    int d, p;
    int64_t lo, hi;   // rows in [lo, hi)
    int lanes, half;
};

struct Req {
    int64_t rid = 0;
    // deep copy of the patch
    gsx_patch_desc desc{};
    std::vector<double> knots[3], ctrl, kappa, felem;
    std::vector<std::vector<double>> rhs;   // empty: the assembled load vector
    std::vector<double *> xout;
    int32_t *iters_out = nullptr, *status_out = nullptr;
    double *relres_out = nullptr;
    double tol = 1e-10;
    int32_t maxit = 20000;
    int d = 2, p = 2;
    int nb[3] = {1, 1, 1}, nel[3] = {1, 1, 1}, m[3] = {1, 1, 1};
    int64_t nbasis = 0, nfree = 0, nelem = 0, bytes = 0;
    int levels = 0;
    double t_submit = 0, t_admit = 0, t_asm = 0;
    gsx_problem *prob = nullptr;
    gsx_matrix *A = nullptr;
    gsx_ctx *owner = nullptr;   // context of prob, A and dev
    double *dev = nullptr;      // nrhs right-hand sides, then nrhs solutions
    std::vector<int32_t> ids;
    int nrhs() const { return (int)rhs.size(); }
    double *b(int k) const { return dev + (size_t)k * nfree; }
    double *x(int k) const { return dev + (size_t)(nrhs() + k) * nfree; }
};

std::string status_text(gsx_ctx *c, gsx_status st, const char *what)
{
    const char *msg = c ? gsx_last_error(c) : nullptr;
    return std::string(what) + ": status " + std::to_string((int)st) + (msg && *msg ? std::string(" (") + msg + ")" : "");
}

}  // namespace

struct gsx_server {
    gsx_server_opts o{};
    gsx_ctx *ca = nullptr, *cb = nullptr;
    std::vector<gsx_ctx *> cms;      // one context per multigrid worker
    gsx_batch *batch = nullptr;
    std::vector<TileRule> tiles;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req *> incoming, ready, mgq;
    std::deque<gsx_server_result> done;
    std::vector<Req *> active;       // solver thread only
    int64_t used = 0, outstanding = 0, next_rid = 0;
    int members = 0;                 // solver thread only
    bool stop = false;
    std::string fatal;
    std::thread tasm, tsol;
    std::vector<std::thread> tmg;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // device bytes a request keeps resident: CSR (+ FP32 copy), structured
    // operator, rhs/solution vectors and the CG workspace per right-hand side
    int64_t estimate(const Req &q) const
    {
        const bool mixed = o.precision == GSX_MIXED;
        int64_t slots = 1;
        for (int k = 0; k < q.d; ++k) slots *= 2 * q.p + 1;
        const int64_t nnz = q.nfree * slots, tiles_rows = (q.nfree + 255) / 256 * 256;
        const int64_t stencil = tiles_rows * slots * 8 + (mixed ? tiles_rows * (slots - slots / 2) * 4 : 0);
        const int64_t per_rhs = q.nfree * (8 * 8 + (mixed ? 4 * 6 : 0));
        return nnz * (4 + 8 + (mixed ? 4 : 0)) + stencil + q.nfree * 8 * 3 + q.nrhs() * per_rhs + (64 << 20);
    }

    int64_t element_bytes(const Req &q) const
    {
        int64_t nb = 1;
        for (int k = 0; k < q.d; ++k) nb *= q.p + 1;
        return q.nelem * nb * nb * 8;
    }

    // levels of a dyadic hierarchy (coarsest at most mg_max_coarse free
    // DOFs, element-level coarsening for the degree), or 0
    int mg_levels(const Req &q) const
    {
        if (o.mg_min_free <= 0 || q.nfree < o.mg_min_free) return 0;
        if (!((q.d == 2 && q.p <= 4) || (q.d == 3 && q.p <= 3))) return 0;
        for (int k = 1; k < q.d; ++k)
            if (q.nel[k] != q.nel[0]) return 0;
        if (element_bytes(q) > o.elem_budget) return 0;
        auto coarse = [&](int n) {
            int64_t v = 1;
            for (int k = 0; k < q.d; ++k) v *= std::max(0, n + q.p - 2);
            return v;
        };
        int n = q.nel[0], levels = 1;
        while (n % 2 == 0 && n / 2 >= 2) {
            if (coarse(n) <= o.mg_max_coarse) break;
            n /= 2;
            ++levels;
        }
        if (levels < 2 || coarse(n) > o.mg_max_coarse) return 0;
        return levels;
    }

    void tile_of(const Req &q, int *lanes, int *half) const
    {
        *lanes = 1;
        *half = 1;
        for (const TileRule &t : tiles)
            if (t.d == q.d && t.p == q.p && q.nfree >= t.lo && q.nfree < t.hi) {
                *lanes = t.lanes;
                *half = t.half;
                return;
            }
    }

    void release(Req *q)
    {
        if (q->A) gsx_matrix_destroy(q->A);
        if (q->prob) gsx_problem_destroy(q->prob);
        if (q->dev) gsx_device_free(q->owner, q->dev);
        q->A = nullptr;
        q->prob = nullptr;
        q->dev = nullptr;
    }

    // the request is complete (or failed): record its result, return its
    // budget; the caller has released its device objects
    void finish(Req *q, int32_t status, int32_t lane, const std::vector<int32_t> &its,
                const std::vector<double> &rels, const std::string &err, const std::vector<int32_t> &sts = {})
    {
        for (int k = 0; k < q->nrhs(); ++k) {
            const bool have = k < (int)its.size() && status != GSX_SERVER_FAILED;
            if (q->iters_out) q->iters_out[k] = have ? its[k] : 0;
            if (q->relres_out) q->relres_out[k] = have ? rels[k] : NAN;
            if (q->status_out)
                q->status_out[k] = !have ? GSX_SERVER_FAILED : (k < (int)sts.size() ? sts[k] : status);
        }
        gsx_server_result r{};
        r.rid = q->rid;
        r.status = status;
        r.lane = lane;
        r.nrhs = q->nrhs();
        r.nbasis = q->nbasis;
        r.nfree = q->nfree;
        r.t_submit = q->t_submit;
        r.t_admit = q->t_admit;
        r.t_assembled = q->t_asm;
        r.t_done = now();
        r.iters_max = 0;
        r.relres_max = 0.0;
        for (int32_t it : its) r.iters_max = std::max(r.iters_max, it);
        for (double v : rels) r.relres_max = std::max(r.relres_max, v);
        std::snprintf(r.error, sizeof(r.error), "%s", err.c_str());
        {
            std::lock_guard<std::mutex> lk(mu);
            done.push_back(r);
            used -= q->bytes;
            --outstanding;
        }
        cv.notify_all();
        delete q;
    }

    // create the problem, the device vectors and the operator on context c
    gsx_status build(Req *q, gsx_ctx *c, int32_t mode, std::string &err)
    {
        q->owner = c;
        int32_t dir[6][2];
        for (int s = 0; s < 2 * q->d; ++s) {
            dir[s][0] = 0;
            dir[s][1] = s + 1;
        }
        gsx_status st = gsx_problem_create(c, 1, &q->desc, 2 * q->d, &dir[0][0], 0, nullptr, &q->prob);
        if (st != GSX_OK) {
            err = status_text(c, st, "problem");
            return st;
        }
        int64_t nb, nf, ne, ng;
        gsx_problem_info(q->prob, &nb, &nf, &ne, &ng);
        if (nf != q->nfree) {
            err = "problem: free DOFs differ from the patch's interior";
            return GSX_ERR_ARG;
        }
        st = gsx_device_alloc(c, (int64_t)sizeof(double) * 2 * q->nrhs() * q->nfree, (void **)&q->dev);
        if (st != GSX_OK) {
            err = status_text(c, st, "device vectors");
            return st;
        }
        st = gsx_assemble(q->prob, mode, o.precision, &q->A, q->b(0));
        if (st != GSX_OK) {
            err = status_text(c, st, "assembly");
            return st;
        }
        // right-hand side 0 is the assembled load vector unless given; the
        // others are always given (checked at submission); x = 0
        const std::vector<double> zero(q->nfree, 0.0);
        for (int k = 0; k < q->nrhs() && st == GSX_OK; ++k) {
            if (!q->rhs[k].empty())
                st = gsx_copy_to_device(c, q->b(k), q->rhs[k].data(), (int64_t)sizeof(double) * q->nfree);
            if (st == GSX_OK)
                st = gsx_copy_to_device(c, q->x(k), zero.data(), (int64_t)sizeof(double) * q->nfree);
        }
        if (st == GSX_OK) st = gsx_ctx_synchronize(c);
        if (st != GSX_OK) err = status_text(c, st, "right-hand sides");
        return st;
    }

    // device memory still free after admitting `need` bytes stays above the
    // reserve (kernels with private arrays need the runtime to allocate
    // scratch at dispatch: a device without free memory aborts the queue);
    // the contexts' cached blocks are released first when that helps
    bool device_room(int64_t need)
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return true;
        if ((int64_t)fr - need >= o.reserve_bytes) return true;
        gsx_ctx_release_cache(ca);
        gsx_ctx_release_cache(cb);
        for (gsx_ctx *c : cms) gsx_ctx_release_cache(c);
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return true;
        return (int64_t)fr - need >= o.reserve_bytes;
    }

    void assembler()
    {
        (void)hipSetDevice(o.device);
        while (true) {
            Req *q = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] {
                    if (stop) return true;
                    if (incoming.empty()) return false;
                    const Req *h = incoming.front();
                    return used == 0 || used + h->bytes <= o.budget_bytes;
                });
                if (stop) return;
                const int64_t need = incoming.front()->bytes;
                if (used > 0) {
                    lk.unlock();
                    const bool room = device_room(need);
                    lk.lock();
                    if (!room) {   // wait for a request to finish, then look again
                        cv.wait_for(lk, std::chrono::milliseconds(20));
                        continue;
                    }
                }
                if (stop) return;
                q = incoming.front();
                incoming.pop_front();
                used += q->bytes;
            }
            q->t_admit = now();
            if (q->levels) {
                {
                    std::lock_guard<std::mutex> lk(mu);
                    mgq.push_back(q);
                }
                cv.notify_all();
                continue;
            }
            std::string err;
            if (build(q, ca, GSX_PATTERN_EXPR, err) != GSX_OK) {
                release(q);
                finish(q, GSX_SERVER_FAILED, 0, {}, {}, err);
                continue;
            }
            q->t_asm = now();
            {
                std::lock_guard<std::mutex> lk(mu);
                ready.push_back(q);
            }
            cv.notify_all();
        }
    }

    void solver()
    {
        const int iters = std::max(1, o.iters_per_step);
        while (true) {
            std::vector<Req *> add;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (active.empty())
                    cv.wait(lk, [&] { return stop || !ready.empty(); });
                if (stop) return;
                // a request larger than the capacity still goes in alone
                // (its members are refused and the request fails)
                while (!ready.empty() && (members + ready.front()->nrhs() <= o.capacity ||
                                          (active.empty() && add.empty()))) {
                    add.push_back(ready.front());
                    members += ready.front()->nrhs();
                    ready.pop_front();
                }
            }
            for (Req *q : add) {
                int lanes, half;
                tile_of(*q, &lanes, &half);
                gsx_status st = GSX_OK;
                for (int k = 0; k < q->nrhs() && st == GSX_OK; ++k) {
                    int32_t id = -1;
                    st = gsx_batch_add_box(batch, q->A, 0, (int32_t)q->nfree, q->d, q->p, q->m[0], q->m[1],
                                           q->d == 3 ? q->m[2] : 1, lanes, half, GSX_PREC_JACOBI, 1.0, q->b(k),
                                           q->x(k), q->tol, q->maxit, &id);
                    if (st == GSX_OK) q->ids.push_back(id);
                }
                if (st != GSX_OK) {
                    const std::string err = status_text(cb, st, "batch member");
                    for (int32_t id : q->ids) gsx_batch_remove(batch, id);
                    members -= q->nrhs();
                    release(q);
                    finish(q, GSX_SERVER_FAILED, 0, {}, {}, err);
                    continue;
                }
                active.push_back(q);
            }
            if (active.empty()) continue;
            int32_t running = 0;
            gsx_status st = gsx_batch_step(batch, iters, &running);
            if (st != GSX_OK) {
                std::lock_guard<std::mutex> lk(mu);
                fatal = status_text(cb, st, "batch step");
                stop = true;
                cv.notify_all();
                return;
            }
            for (size_t i = 0; i < active.size();) {
                Req *q = active[i];
                std::vector<int32_t> its, sts;
                std::vector<double> rels;
                bool finished = true;
                for (int32_t id : q->ids) {
                    int32_t s = 0, it = 0;
                    double rr = 0;
                    gsx_batch_result(batch, id, &s, &it, &rr);
                    if (s == 0) finished = false;
                    if (s == 3) s = 1;   // zero right-hand side: x = 0 is the solution
                    sts.push_back(s);
                    its.push_back(it);
                    rels.push_back(rr);
                }
                if (!finished) {
                    ++i;
                    continue;
                }
                gsx_status cs = GSX_OK;
                for (int k = 0; k < q->nrhs(); ++k)
                    if (q->xout[k] && cs == GSX_OK)
                        cs = gsx_copy_to_host(cb, q->xout[k], q->x(k), (int64_t)sizeof(double) * q->nfree);
                for (int32_t id : q->ids) gsx_batch_remove(batch, id);
                members -= q->nrhs();
                active.erase(active.begin() + (ptrdiff_t)i);
                release(q);
                int32_t status = 1;
                for (int32_t s : sts)
                    if (s == 2) status = 2;
                finish(q, cs == GSX_OK ? status : GSX_SERVER_FAILED, 0, its, rels,
                       cs == GSX_OK ? std::string() : status_text(cb, cs, "solution download"), sts);
            }
        }
    }

    void mg_worker(gsx_ctx *cm)
    {
        while (true) {
            Req *q = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !mgq.empty(); });
                if (stop) return;
                q = mgq.front();
                mgq.pop_front();
            }
            std::string err;
            gsx_mg *mg = nullptr;
            std::vector<int32_t> its, sts;
            std::vector<double> rels;
            int32_t status = 1;
            gsx_status st = build(q, cm, GSX_PATTERN_EXPR | GSX_ASM_KEEP_ELEMENTS, err);
            if (st == GSX_OK) {
                st = gsx_mg_create(q->prob, q->A, q->levels, 2, 2, 1, 0.6, 1.0, o.precision, &mg);
                if (st != GSX_OK) err = status_text(cm, st, "multigrid set-up");
            }
            if (st == GSX_OK) q->t_asm = now();
            for (int k = 0; k < q->nrhs() && st == GSX_OK; ++k) {
                int32_t it = 0;
                double rel = 0;
                const gsx_status cs = gsx_cg_solve(q->A, GSX_PREC_MG, 1.0, mg, o.precision, q->b(k), q->x(k),
                                                   q->tol, q->maxit, &it, &rel);
                if (cs == GSX_ERR_NOT_CONVERGED) status = 2;
                else if (cs != GSX_OK) {
                    st = cs;
                    err = status_text(cm, cs, "multigrid CG");
                    break;
                }
                its.push_back(it);
                rels.push_back(rel);
                sts.push_back(cs == GSX_OK ? 1 : 2);
                if (q->xout[k]) {
                    st = gsx_copy_to_host(cm, q->xout[k], q->x(k), (int64_t)sizeof(double) * q->nfree);
                    if (st != GSX_OK) err = status_text(cm, st, "solution download");
                }
            }
            if (mg) gsx_mg_destroy(mg);
            release(q);
            finish(q, st == GSX_OK ? status : GSX_SERVER_FAILED, 1, its, rels, err, sts);
        }
    }
};

extern "C" {

gsx_status gsx_server_create(const gsx_server_opts *opts, gsx_server **out)
{
    if (!opts || !out || (opts->precision != GSX_FP64 && opts->precision != GSX_MIXED) || opts->capacity < 1 ||
        opts->budget_bytes <= 0)
        return GSX_ERR_ARG;
    gsx_server *s = new gsx_server();
    s->o = *opts;
    if (s->o.elem_budget <= 0) s->o.elem_budget = (int64_t)16 << 30;
    if (s->o.mg_max_coarse <= 0) s->o.mg_max_coarse = 4096;
    if (s->o.reserve_bytes <= 0) s->o.reserve_bytes = (int64_t)16 << 30;
    gsx_status st = gsx_ctx_create(opts->device, &s->ca);
    if (st == GSX_OK) st = gsx_ctx_create(opts->device, &s->cb);
    const int nmg = s->o.mg_min_free > 0 ? std::max(1, s->o.mg_workers) : 0;
    for (int k = 0; k < nmg && st == GSX_OK; ++k) {
        gsx_ctx *c = nullptr;
        st = gsx_ctx_create(opts->device, &c);
        if (c) s->cms.push_back(c);
    }
    if (st == GSX_OK) st = gsx_batch_create(s->cb, opts->precision, opts->capacity, &s->batch);
    if (st == GSX_OK) st = gsx_batch_set_group(s->batch, opts->group_max > 0 ? opts->group_max : 8);
    if (st == GSX_OK && opts->group_rows > 0) st = gsx_batch_set_group_rows(s->batch, opts->group_rows);
    if (st != GSX_OK) {
        if (s->batch) gsx_batch_destroy(s->batch);
        for (gsx_ctx *c : {s->ca, s->cb})
            if (c) gsx_ctx_destroy(c);
        for (gsx_ctx *c : s->cms) gsx_ctx_destroy(c);
        delete s;
        return st;
    }
    s->t0 = std::chrono::steady_clock::now();
    s->tasm = std::thread([s] { s->assembler(); });
    s->tsol = std::thread([s] { s->solver(); });
    for (gsx_ctx *c : s->cms) s->tmg.emplace_back([s, c] { s->mg_worker(c); });
    *out = s;
    return GSX_OK;
}

gsx_status gsx_server_set_tile(gsx_server *s, int32_t d, int32_t p, int64_t rows_lo, int64_t rows_hi,
                               int32_t lanes, int32_t half)
{
    if (!s || (lanes != 1 && lanes != 2 && lanes != 4) || (half != 0 && half != 1) || rows_hi <= rows_lo)
        return GSX_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->tiles.push_back({d, p, rows_lo, rows_hi, lanes, half});
    return GSX_OK;
}

gsx_status gsx_server_submit(gsx_server *s, const gsx_server_request *r, int64_t *rid)
{
    if (!s || !r || !rid || !r->patch || r->nrhs < 1 || r->nrhs > 64 || r->maxit < 1) return GSX_ERR_ARG;
    const gsx_patch_desc &pd = *r->patch;
    if ((pd.d != 2 && pd.d != 3) || !pd.ctrl_host) return GSX_ERR_ARG;
    Req *q = new Req();
    q->d = pd.d;
    q->p = pd.p[0];
    q->desc = pd;
    int64_t nbasis = 1, nfree = 1, nelem = 1;
    for (int k = 0; k < pd.d; ++k) {
        if (pd.p[k] != q->p || !pd.knots_host[k] || pd.nknots[k] < 2 * (q->p + 1)) {
            delete q;
            return GSX_ERR_ARG;
        }
        q->knots[k].assign(pd.knots_host[k], pd.knots_host[k] + pd.nknots[k]);
        q->desc.knots_host[k] = q->knots[k].data();
        q->nb[k] = pd.nknots[k] - q->p - 1;
        int spans = 0;
        for (int i = 1; i < pd.nknots[k]; ++i) spans += q->knots[k][i] > q->knots[k][i - 1];
        q->nel[k] = spans;
        q->m[k] = q->nb[k] - 2;
        if (q->m[k] < 1) {
            delete q;
            return GSX_ERR_ARG;
        }
        nbasis *= q->nb[k];
        nfree *= q->m[k];
        nelem *= spans;
    }
    q->nbasis = nbasis;
    q->nfree = nfree;
    q->nelem = nelem;
    q->ctrl.assign(pd.ctrl_host, pd.ctrl_host + nbasis * pd.d);
    q->desc.ctrl_host = q->ctrl.data();
    if (pd.kappa_host) {
        q->kappa.assign(pd.kappa_host, pd.kappa_host + nelem);
        q->desc.kappa_host = q->kappa.data();
    }
    if (pd.felem_host) {
        q->felem.assign(pd.felem_host, pd.felem_host + nelem);
        q->desc.felem_host = q->felem.data();
    }
    q->rhs.resize(r->nrhs);
    q->xout.assign(r->nrhs, nullptr);
    for (int k = 0; k < r->nrhs; ++k) {
        if (r->rhs_host && r->rhs_host[k]) q->rhs[k].assign(r->rhs_host[k], r->rhs_host[k] + nfree);
        else if (k > 0) {   // only the first may default to the load vector
            delete q;
            return GSX_ERR_ARG;
        }
        if (r->x_host) q->xout[k] = r->x_host[k];
    }
    q->tol = r->tol;
    q->maxit = r->maxit;
    q->iters_out = r->iters_host;
    q->relres_out = r->relres_host;
    q->status_out = r->status_host;
    q->levels = s->mg_levels(*q);
    q->bytes = s->estimate(*q) + (q->levels ? s->element_bytes(*q) + s->estimate(*q) / 2 : 0);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete q;
            return GSX_ERR_ARG;
        }
        q->rid = s->next_rid++;
        q->t_submit = s->now();
        ++s->outstanding;
        s->incoming.push_back(q);
    }
    s->cv.notify_all();
    *rid = q->rid;
    return GSX_OK;
}

gsx_status gsx_server_poll(gsx_server *s, int32_t max, gsx_server_result *out, int32_t *n)
{
    if (!s || !n || max < 0 || (max > 0 && !out)) return GSX_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    int32_t k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    return s->fatal.empty() ? GSX_OK : GSX_ERR_HIP;
}

gsx_status gsx_server_status(gsx_server *s, int64_t *outstanding, int64_t *used_bytes, double *clock)
{
    if (!s) return GSX_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (outstanding) *outstanding = s->outstanding;
    if (used_bytes) *used_bytes = s->used;
    if (clock) *clock = s->now();
    return s->fatal.empty() ? GSX_OK : GSX_ERR_HIP;
}

const char *gsx_server_error(const gsx_server *s) { return s ? s->fatal.c_str() : ""; }

gsx_status gsx_server_drain(gsx_server *s, double timeout_s)
{
    if (!s) return GSX_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                   [&] { return s->outstanding == 0 || !s->fatal.empty(); });
    if (!s->fatal.empty()) return GSX_ERR_HIP;
    return ok ? GSX_OK : GSX_ERR_NOT_CONVERGED;
}

gsx_status gsx_server_destroy(gsx_server *s)
{
    if (!s) return GSX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    s->tasm.join();
    s->tsol.join();
    for (std::thread &th : s->tmg) th.join();
    // requests still held: free their device objects
    for (Req *q : s->active) {
        s->release(q);
        delete q;
    }
    for (auto *dq : {&s->incoming, &s->ready, &s->mgq})
        for (Req *q : *dq) {
            s->release(q);
            delete q;
        }
    gsx_batch_destroy(s->batch);
    gsx_ctx_destroy(s->ca);
    gsx_ctx_destroy(s->cb);
    for (gsx_ctx *c : s->cms) gsx_ctx_destroy(c);
    delete s;
    return GSX_OK;
}

}  // extern "C"

// Request server of libgrmp_hip.so: continuous batching of many small and
// large PDE solves on one GPU.
//
// A request names a problem (Poisson or P1 elasticity on an N x N unit
// square, P1 or P2, seeded log-normal cell coefficient).  Served mode:
//   * admission: a request is admitted when a problem slot is free and its
//     HBM footprint fits the memory budget; otherwise it waits in FIFO order;
//   * resident state: mesh, dofmap and sparsity pattern (with the element-
//     to-slot map) are built once per problem shape and kept in HBM (LRU);
//     per request only the coefficient, matrix values, right-hand side and
//     Krylov vectors are allocated;
//   * overlap: assembly of newly admitted requests runs on its own stream
//     while the Krylov engine steps the problems already in flight on
//     another; an assembled problem joins the engine at the next step;
//   * the engine runs FP32 CG sweeps inside FP64 iterative refinement for all
//     problems in flight at once (block-diagonal batch), replayed as a
//     hipGraph.
// Plain mode (the baseline): requests are served one at a time, each built
// from scratch and solved by the host-driven FP64 PCG.
// Multigrid mode: up to mg_slots requests at once, each on its own stream,
// solved in FP64 by CG preconditioned with a geometric V-cycle whose levels
// reuse the resident patterns of the coarser shapes; the host enqueues a few
// iterations per request and polls a status copy between them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "grmp_engine.h"

static const int64_t kTileRows = ENGINE_TILE_ROWS;

static inline int64_t req_nnodes(const grmp_request &r) { return (int64_t)(r.N + 1) * (r.N + 1); }
static inline int64_t req_ncells(const grmp_request &r) { return 2ll * r.N * r.N; }

extern "C" int64_t grmp_request_ndofs(const grmp_request *r)
{
    if (!r || r->N < 1) return -1;
    const int64_t N = r->N, nn = req_nnodes(*r);
    if (r->kind == GRMP_REQ_ELASTICITY) return 2 * nn;
    if (r->fe == GRMP_FE_H1P2) return nn + 3 * N * N + 2 * N;
    return nn;
}

static int64_t req_nnz_bound(const grmp_request &r)
{
    const int64_t n = grmp_request_ndofs(&r);
    if (r.kind == GRMP_REQ_ELASTICITY) return 14 * n;
    if (r.fe == GRMP_FE_H1P2) return 12 * n;
    return 7 * n;
}

// bytes of the per-request device buffer (see Job::layout)
static inline size_t al(size_t x) { return (x + 255) & ~(size_t)255; }
extern "C" int64_t grmp_request_footprint(const grmp_request *r)
{
    if (!r || r->N < 1) return -1;
    const int64_t n = grmp_request_ndofs(r), nnz = req_nnz_bound(*r), nc = req_ncells(*r);
    size_t b = al(nc * 8) + al(nnz * 8) + al(nnz * 4) + al(n * 8) + al(n * 4) + 3 * al(n * 8) + al(4 * n * 4) +
               al(sizeof(DevForm) * 4);
    return (int64_t)b;
}

struct Template {
    int kind = 0, fe = 0, N = 0;
    grmp_mesh *mesh = nullptr;
    grmp_space *space = nullptr;
    grmp_matrix *A = nullptr;
    grmp_vector *kappa0 = nullptr;
    int refs = 0;
    int64_t bytes = 0;
    uint64_t last_use = 0;
    grmp_lform lf;
    bool ready = true;   // false while a multigrid setup thread builds it
    int build_rc = 0;
};

struct Job {
    int64_t ticket = 0;
    grmp_request req;
    grmp_result res;
    int state = 0;  // 0 queued, 1 assembling, 2 solving, 3 done
    Template *tpl = nullptr;
    char *dbuf = nullptr;
    size_t dbytes = 0;
    double *kappa = nullptr, *val64 = nullptr, *dinv64 = nullptr, *b = nullptr, *x = nullptr, *r64 = nullptr;
    float *val32 = nullptr, *dinv32 = nullptr, *f4 = nullptr;
    DevForm *dforms = nullptr;
    DevForm hforms[4];
    hipEvent_t ev = nullptr;
    int slot = -1;
    int64_t footprint = 0;
    int64_t n = 0;
    // multigrid mode
    MgJob *mg = nullptr;
    std::vector<Template *> mtpl;
    int mslot = -1;
    int mg_it = 0;         // iteration count and residual at the previous status
    double mg_rel = 0.0;
    hipEvent_t mg_ready = nullptr;  // lockstep batches: end of the job's setup
};

struct grmp_server {
    grmp_ctx *user = nullptr;
    grmp_ctx *sctx = nullptr;  // server context; its stream assembles
    hipStream_t st_solve = nullptr;
    grmp_server_config cfg;
    KEngine *eng = nullptr;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<Job *> queue;
    std::unordered_map<int64_t, Job *> jobs;
    std::list<Job *> assembling, solving;
    bool stop = false;
    int64_t next_ticket = 1;
    double t0 = 0.0;
    int64_t mem_used = 0;
    int64_t tpl_bytes = 0;
    uint64_t use_clock = 0;
    std::map<std::tuple<int, int, int>, Template *> templates;
    int fatal = GRMP_OK;
    std::string fatal_msg;
    int64_t ntemplates = 0;
    double template_s = 0.0;
    // multigrid mode: one context (stream) and one pinned status per slot;
    // slots are split over mg_threads host threads (thread t drives the slots
    // s with s % mg_threads == t); tpl_mu guards the template cache for them
    std::vector<grmp_ctx *> mg_ctx;
    std::vector<std::thread> mg_workers;
    std::mutex tpl_mu;
    // lockstep batches: setup threads move requests from queue to ready
    std::deque<Job *> mg_readyq;
    int mg_setting_up = 0, mg_in_batch = 0;
    std::condition_variable cv_ready;
    std::condition_variable cv_tpl;  // a shape under construction became ready
    MgStatus *mg_host = nullptr;
    double prof_s = 0.0, prof_bytes = 0.0;
    int64_t prof_launches = 0;
};

static double now(const grmp_server *s) { return wall_seconds() - s->t0; }

// ------------------------------------------------------------ templates
static void template_free(Template *t)
{
    grmp_matrix_destroy(t->A);
    grmp_vector_destroy(t->kappa0);
    grmp_space_destroy(t->space);
    grmp_mesh_destroy(t->mesh);
    delete t;
}

static int template_make(grmp_server *S, grmp_ctx *c, const grmp_request &r, Template **out)
{
    Template *t = new Template();
    t->kind = r.kind;
    t->fe = r.fe;
    t->N = r.N;
    const int ncomp = r.kind == GRMP_REQ_ELASTICITY ? 2 : 1;
    const int fe = r.kind == GRMP_REQ_ELASTICITY ? GRMP_FE_H1P1 : r.fe;
    int rc;
    if ((rc = grmp_mesh_unitsquare(c, r.N, &t->mesh)) || (rc = grmp_space_create(c, t->mesh, fe, ncomp, &t->space)) ||
        (rc = grmp_vector_create(c, t->mesh->ncells, &t->kappa0)) ||
        (rc = grmp_vector_lognormal(t->kappa0, r.seed, 0.0, r.sigma))) {
        template_free(t);
        return rc;
    }
    const grmp_space *sp[1] = {t->space};
    if ((rc = grmp_matrix_create(c, 1, sp, &t->A))) { template_free(t); return rc; }
    grmp_form f;
    memset(&f, 0, sizeof(f));
    f.coef_mode = GRMP_COEF_CELL;
    f.cell_coef = t->kappa0;
    f.factor = 1.0;
    memset(&t->lf, 0, sizeof(t->lf));
    t->lf.kind = GRMP_LF_CONST;
    t->lf.factor = 1.0;
    if (r.kind == GRMP_REQ_ELASTICITY) {
        const double E = r.E > 0 ? r.E : 1.0, nu = r.nu > 0 ? r.nu : 0.3;
        f.kind = GRMP_BLF_HOOKE2D;
        f.quadorder = 0;
        f.params[0] = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        f.params[1] = E / (2.0 * (1.0 + nu));
        t->lf.quadorder = 1;
        t->lf.f[0] = 0.0;
        t->lf.f[1] = -1.0;
    } else {
        f.kind = GRMP_BLF_LAPLACE;
        f.quadorder = fe == GRMP_FE_H1P2 ? 2 : 0;
        t->lf.quadorder = fe == GRMP_FE_H1P2 ? 2 : 1;
        t->lf.f[0] = 1.0;
    }
    int64_t nb = 0;
    if ((rc = grmp_matrix_add_form(t->A, &f)) || (rc = grmp_space_boundary_dofs(t->space, &nb, nullptr))) {
        template_free(t);
        return rc;
    }
    std::vector<int64_t> bd(nb);
    if ((rc = grmp_space_boundary_dofs(t->space, &nb, bd.data())) || (rc = grmp_matrix_set_fixed(t->A, nb, bd.data(), 1e60)) ||
        (rc = grmp_matrix_build_pattern(t->A))) {
        template_free(t);
        return rc;
    }
    const int64_t nc = t->mesh->ncells, nn = t->mesh->nnodes, n = t->A->n;
    t->bytes = nn * 16 + nc * 12 + nc * t->space->nloc * 8 + n * 4 + t->A->nnz * 4 + n * 6 + nc * 8 +
               nc * t->space->nloc * 4 * ((t->space->nloc + 3) / 4);
    *out = t;
    return GRMP_OK;
}

static void templates_trim(grmp_server *S, int64_t limit)
{
    while (S->tpl_bytes > limit) {
        Template *victim = nullptr;
        for (auto &kv : S->templates)
            if (kv.second->refs == 0 && kv.second->ready && (!victim || kv.second->last_use < victim->last_use))
                victim = kv.second;
        if (!victim) return;
        S->templates.erase(std::make_tuple(victim->kind, victim->fe, victim->N));
        S->tpl_bytes -= victim->bytes;
        template_free(victim);
    }
}

static int template_get(grmp_server *S, const grmp_request &r, Template **out)
{
    auto key = std::make_tuple(r.kind, r.kind == GRMP_REQ_ELASTICITY ? GRMP_FE_H1P1 : r.fe, r.N);
    auto it = S->templates.find(key);
    if (it == S->templates.end()) {
        Template *t = nullptr;
        double t0 = wall_seconds();
        GRMP_TRY(template_make(S, S->sctx, r, &t));
        S->template_s += wall_seconds() - t0;
        S->ntemplates++;
        it = S->templates.emplace(key, t).first;
        S->tpl_bytes += t->bytes;
    }
    it->second->last_use = ++S->use_clock;
    *out = it->second;
    return GRMP_OK;
}

// ----------------------------------------------------------------- jobs
static int job_start(grmp_server *S, Job *j)
{
    grmp_ctx *c = S->sctx;
    hipStream_t st = c->stream;
    j->res.t_admit = now(S);
    GRMP_TRY(template_get(S, j->req, &j->tpl));
    Template *t = j->tpl;
    t->refs++;
    const grmp_matrix *A = t->A;
    const int64_t n = A->n, nnz = A->nnz, nc = t->mesh->ncells;
    j->n = n;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += al(bytes); return o; };
    size_t o_k = take(nc * 8), o_v64 = take(nnz * 8), o_v32 = take(nnz * 4), o_d64 = take(n * 8), o_d32 = take(n * 4);
    size_t o_b = take(n * 8), o_x = take(n * 8), o_r = take(n * 8), o_f = take(4 * n * 4), o_df = take(sizeof(DevForm) * 4);
    j->dbytes = off;
    GRMP_HIP(c, hipMallocAsync((void **)&j->dbuf, off, st));
    j->kappa = (double *)(j->dbuf + o_k);
    j->val64 = (double *)(j->dbuf + o_v64);
    j->val32 = (float *)(j->dbuf + o_v32);
    j->dinv64 = (double *)(j->dbuf + o_d64);
    j->dinv32 = (float *)(j->dbuf + o_d32);
    j->b = (double *)(j->dbuf + o_b);
    j->x = (double *)(j->dbuf + o_x);
    j->r64 = (double *)(j->dbuf + o_r);
    j->f4 = (float *)(j->dbuf + o_f);
    j->dforms = (DevForm *)(j->dbuf + o_df);
    // coefficient field, matrix values (FP64 + FP32 + Jacobi), rhs, x0
    grmp_vector kv{c, nc, j->kappa};
    GRMP_TRY(grmp_vector_lognormal(&kv, j->req.seed, 0.0, j->req.sigma));
    GRMP_TRY(assemble_values(A, j->kappa, j->dforms, j->hforms, j->val64, j->val32, j->dinv64, j->dinv32, st));
    GRMP_HIP(c, hipMemsetAsync(j->b, 0, n * 8, st));
    GRMP_TRY(linearform_launch(c, t->space, &t->lf, j->b, st));
    GRMP_TRY(zero_fixed_launch(A, j->b, st));
    GRMP_TRY(init_x_launch(A, j->val64, j->b, j->x, st));
    GRMP_HIP(c, hipEventCreateWithFlags(&j->ev, hipEventDisableTiming));
    GRMP_HIP(c, hipEventRecord(j->ev, st));
    j->state = 1;
    return GRMP_OK;
}

static int job_activate(grmp_server *S, Job *j)
{
    const grmp_matrix *A = j->tpl->A;
    KProb p;
    memset(&p, 0, sizeof(p));
    p.rowptr = A->rowptr;
    p.cols = A->cols;
    p.val64 = j->val64;
    p.val32 = j->val32;
    p.dinv32 = j->dinv32;
    p.dinv64 = j->dinv64;
    p.fixed = A->fixed;
    p.b = j->b;
    p.x = j->x;
    p.r64 = j->r64;
    p.r = j->f4;
    p.p = j->f4 + j->n;
    p.q = j->f4 + 2 * j->n;
    p.dx = j->f4 + 3 * j->n;
    p.p64 = (double *)j->f4;
    p.q64 = (double *)(j->f4 + 2 * j->n);
    p.n = (int32_t)j->n;
    p.max_outer = 200;
    p.maxit = 4000000;
    p.rtol = j->req.rtol > 0 ? j->req.rtol : 1e-10;
    p.delta = 0.1;  // reliable update when the FP32 residual dropped 10x
    p.fp64 = j->req.precision == GRMP_PRECISION_FP64;
    p.fp64_after = S->cfg.fp64_first_sweep;
    GRMP_TRY(S->eng->add(p, &j->slot));
    j->state = 2;
    return GRMP_OK;
}

// Frees the job's device buffer on the stream that used it last: the solve
// stream once the job has been in the engine (after its slot went idle),
// the assembly stream before that.
static void job_release(grmp_server *S, Job *j)
{
    hipStream_t st = j->slot >= 0 ? S->st_solve : S->sctx->stream;
    if (j->dbuf) (void)hipFreeAsync(j->dbuf, st);
    j->dbuf = nullptr;
    if (j->ev) (void)hipEventDestroy(j->ev);
    j->ev = nullptr;
    if (j->tpl) j->tpl->refs--;
    S->mem_used -= j->footprint;
}

static void job_finish(grmp_server *S, Job *j, const KProb &q)
{
    j->res.status = q.phase == PH_DONE ? GRMP_OK : (q.status ? q.status : GRMP_ERR_NOCONV);
    j->res.iterations = q.iters;
    j->res.outer = q.outer;
    j->res.relres = q.relres;
    j->res.fp64_fallback = q.fp64;
    j->res.ndofs = j->n;
    if (j->req.x_out && (hipMemcpyAsync(j->req.x_out, j->x, j->n * 8, hipMemcpyDeviceToHost, S->st_solve) != hipSuccess ||
                         hipStreamSynchronize(S->st_solve) != hipSuccess) && j->res.status == GRMP_OK)
        j->res.status = GRMP_ERR_HIP;
    S->eng->remove(j->slot);
    job_release(S, j);
    j->res.t_done = now(S);
}

static void job_fail(grmp_server *S, Job *j, int rc)
{
    j->res.status = rc;
    j->res.t_done = now(S);
    job_release(S, j);
    j->state = 3;
}

// A device error ends the server: every job still pending fails with the
// error, engine slots are released first (their records go idle before any
// buffer is freed), and later submissions are refused.
static void server_fatal(grmp_server *S, int rc)
{
    (void)hipStreamSynchronize(S->st_solve);
    (void)hipStreamSynchronize(S->sctx->stream);
    std::lock_guard<std::mutex> g(S->mu);
    S->fatal = rc;
    S->fatal_msg = S->sctx->err;
    for (Job *j : S->solving) {
        S->eng->remove(j->slot);
        job_fail(S, j, rc);
    }
    S->solving.clear();
    for (Job *j : S->assembling) job_fail(S, j, rc);
    S->assembling.clear();
    for (Job *j : S->queue) {
        j->res.status = rc;
        j->res.t_done = now(S);
        j->state = 3;
    }
    S->queue.clear();
    (void)hipStreamSynchronize(S->st_solve);
    S->cv_done.notify_all();
}

// ------------------------------------------------------------ served loop
static void served_loop(grmp_server *S)
{
    std::vector<KProb> st;
    const int steps = 16;
    for (;;) {
        std::vector<Job *> admit;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->queue.empty()) break;
                Job *j = S->queue.front();
                const int inflight = (int)(S->assembling.size() + S->solving.size() + admit.size());
                if (inflight >= S->cfg.max_active) break;
                if (inflight > 0 && S->mem_used + j->footprint > S->cfg.memory_budget) break;
                S->queue.pop_front();
                S->mem_used += j->footprint;
                admit.push_back(j);
            }
            if (admit.empty() && S->assembling.empty() && S->solving.empty()) {
                if (S->stop) return;
                S->cv.wait(lk, [&] { return S->stop || !S->queue.empty(); });
                continue;
            }
        }
        for (Job *j : admit) {
            int rc = job_start(S, j);
            std::lock_guard<std::mutex> g(S->mu);
            if (rc) {
                job_fail(S, j, rc);
                S->cv_done.notify_all();
            } else {
                S->assembling.push_back(j);
            }
        }
        // assembled problems join the engine
        for (auto it = S->assembling.begin(); it != S->assembling.end();) {
            Job *j = *it;
            if (hipEventQuery(j->ev) != hipSuccess) { ++it; continue; }
            int rc = job_activate(S, j);
            std::lock_guard<std::mutex> g(S->mu);
            it = S->assembling.erase(it);
            if (rc) {
                job_fail(S, j, rc);
                S->cv_done.notify_all();
            } else {
                S->solving.push_back(j);
            }
        }
        if (S->eng->nactive() == 0) {
            if (!S->assembling.empty()) (void)hipEventSynchronize(S->assembling.front()->ev);
            continue;
        }
        int rc = S->eng->run(steps, S->cfg.graph != 0);
        if (!rc) rc = S->eng->poll(st);
        if (rc) {
            server_fatal(S, rc);
            return;
        }
        for (auto it = S->solving.begin(); it != S->solving.end();) {
            Job *j = *it;
            const KProb &q = st[j->slot];
            if (q.phase != PH_DONE && q.phase != PH_FAIL) { ++it; continue; }
            job_finish(S, j, q);
            std::lock_guard<std::mutex> g(S->mu);
            j->state = 3;
            it = S->solving.erase(it);
            S->cv_done.notify_all();
        }
        {
            std::lock_guard<std::mutex> g(S->mu);
            templates_trim(S, S->cfg.memory_budget / 2);
        }
    }
}

// ---------------------------------------------------------- multigrid loop
static const int kMgChunk = 4;     // PCG iterations enqueued between status polls
static const int kMgMaxIt = 500;
static const int kMgSetupGroup = 16;  // requests a setup thread sets up together (one assembly launch)
static const int kMgProfileEvery = 4;

// (caller holds S->mu; lock order S->mu before tpl_mu)
static void mg_release(grmp_server *S, Job *j)
{
    mg_job_destroy(j->mg);
    j->mg = nullptr;
    if (j->mg_ready) (void)hipEventDestroy(j->mg_ready);
    j->mg_ready = nullptr;
    {
        std::lock_guard<std::mutex> g(S->tpl_mu);
        for (Template *t : j->mtpl) t->refs--;
    }
    j->mtpl.clear();
    S->mem_used -= j->footprint;
}

// A shape for a multigrid job, with a reference taken.  A missing shape is
// built by the calling thread on its own context, outside the cache lock, and
// made complete on the device before others use it; threads that need the
// same shape meanwhile wait for it.
static int mg_template(grmp_server *S, grmp_ctx *c, const grmp_request &r, Template **out)
{
    auto key = std::make_tuple(r.kind, r.kind == GRMP_REQ_ELASTICITY ? GRMP_FE_H1P1 : r.fe, r.N);
    Template *t = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->tpl_mu);
        auto it = S->templates.find(key);
        if (it != S->templates.end()) {
            t = it->second;
            t->refs++;
            t->last_use = ++S->use_clock;
            S->cv_tpl.wait(lk, [&] { return t->ready; });
            if (t->build_rc) {  // the build failed; the entry already left the cache
                const int rc = t->build_rc;
                if (--t->refs == 0) delete t;
                return rc;
            }
            *out = t;
            return GRMP_OK;
        }
        t = new Template();  // placeholder: found by others, not ready
        t->ready = false;
        t->refs = 1;
        t->last_use = ++S->use_clock;
        S->templates.emplace(key, t);
    }
    const double t0 = wall_seconds();
    Template *b = nullptr;
    int rc = template_make(S, c, r, &b);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = grmp_fail(c, GRMP_ERR_HIP, "multigrid shape build");
    std::lock_guard<std::mutex> g(S->tpl_mu);
    if (rc) {
        // failed: waiting threads see the code; the entry leaves the cache
        t->build_rc = rc;
        t->ready = true;
        t->refs--;
        S->templates.erase(key);
        S->cv_tpl.notify_all();
        if (t->refs == 0) delete t;  // else the last waiter deletes it
        if (b) template_free(b);
        return rc;
    }
    // move the built objects into the placeholder others already hold
    t->kind = b->kind;
    t->fe = b->fe;
    t->N = b->N;
    t->mesh = b->mesh;
    t->space = b->space;
    t->A = b->A;
    t->kappa0 = b->kappa0;
    t->bytes = b->bytes;
    t->lf = b->lf;
    b->mesh = nullptr;
    b->space = nullptr;
    b->A = nullptr;
    b->kappa0 = nullptr;
    delete b;
    t->ready = true;
    S->tpl_bytes += t->bytes;
    S->ntemplates++;
    S->template_s += wall_seconds() - t0;
    S->cv_tpl.notify_all();
    *out = t;
    return GRMP_OK;
}

// Levels: the request's shape, P1 on the same square when it is P2, then P1
// on ceil(N/2), ceil(N/4), ... down to N <= 4.  mg_build makes the shapes,
// the job, its coefficient and right-hand side; the level assembly and the
// first V-cycle follow (mg_prepare: for this job alone; the setup threads:
// for a group of jobs with one assembly launch).
static int mg_build(grmp_server *S, Job *j, grmp_ctx *c, bool prof)
{
    hipStream_t st = c->stream;
    j->res.t_admit = now(S);
    const grmp_request &r = j->req;
    const int ncomp = r.kind == GRMP_REQ_ELASTICITY ? 2 : 1;
    const int fe0 = r.kind == GRMP_REQ_ELASTICITY ? GRMP_FE_H1P1 : r.fe;
    std::vector<std::pair<int, int>> shape;  // (fe, N)
    shape.push_back({fe0, r.N});
    if (fe0 == GRMP_FE_H1P2) shape.push_back({GRMP_FE_H1P1, r.N});
    for (int N = r.N; N > 4;) {
        N = (N + 1) / 2;
        shape.push_back({GRMP_FE_H1P1, N});
    }
    std::vector<MgLevelIn> in(shape.size());
    for (size_t l = 0; l < shape.size(); ++l) {
        grmp_request rl = r;
        rl.fe = shape[l].first;
        rl.N = shape[l].second;
        Template *t = nullptr;
        GRMP_TRY(mg_template(S, c, rl, &t));
        j->mtpl.push_back(t);
        in[l] = MgLevelIn{t->A, t->space, rl.N, shape[l].first, ncomp};
    }
    GRMP_TRY(mg_job_create(c, st, in.data(), (int)in.size(), prof, S->cfg.mg_block_rows, S->cfg.mg_coarse_rows,
                           S->cfg.graph != 0 && S->cfg.mg_batch == 0, &j->mg));
    Template *t0 = j->mtpl[0];
    j->n = t0->A->n;
    grmp_vector kv{c, t0->mesh->ncells, mg_job_kappa0(j->mg)};
    GRMP_TRY(grmp_vector_lognormal(&kv, r.seed, 0.0, r.sigma));
    double *rhs = mg_job_rhs(j->mg);
    GRMP_HIP(c, hipMemsetAsync(rhs, 0, j->n * 8, st));
    GRMP_TRY(linearform_launch(c, t0->space, &t0->lf, rhs, st));
    GRMP_TRY(zero_fixed_launch(t0->A, rhs, st));
    return GRMP_OK;
}

static int mg_prepare(grmp_server *S, Job *j, grmp_ctx *c, bool prof)
{
    GRMP_TRY(mg_build(S, j, c, prof));
    GRMP_TRY(mg_job_setup(j->mg, j->req.rtol > 0 ? j->req.rtol : 1e-10, kMgMaxIt));
    j->state = 2;
    return GRMP_OK;
}

static int mg_start(grmp_server *S, Job *j, int slot)
{
    j->mslot = slot;
    // Profiling times every grid residual launch with an event pair.  Events
    // cannot be recorded inside a capture, so with graphs on only every
    // kMgProfileEvery-th request (tickets 1, 1 + k, ...) is profiled, and it
    // launches its kernels one by one.
    const bool prof = S->cfg.profile != 0 && (!S->cfg.graph || (j->ticket - 1) % kMgProfileEvery == 0);
    GRMP_TRY(mg_prepare(S, j, S->mg_ctx[slot], prof));
    GRMP_TRY(mg_job_iterate(j->mg, kMgChunk));
    GRMP_TRY(mg_job_status_async(j->mg, &S->mg_host[slot]));
    return GRMP_OK;
}

// Iterations to enqueue before the next status: with replayed graphs the
// captured chunk; otherwise what the last chunk's residual reduction rate
// predicts to reach rtol, between 1 and 8, so few launches follow convergence.
static int mg_chunk(const grmp_server *S, const Job *j, const MgStatus &h)
{
    if (S->cfg.graph) return kMgChunk;
    const double rtol = j->req.rtol > 0 ? j->req.rtol : 1e-10;
    if (j->mg_it <= 0 || h.it <= j->mg_it || !(h.relres > 0.0) || !(h.relres < j->mg_rel)) return kMgChunk;
    const double rate = pow(h.relres / j->mg_rel, 1.0 / (h.it - j->mg_it));
    const double need = ceil(log(rtol / h.relres) / log(rate));
    if (!(need >= 1.0)) return 1;
    return need > 8.0 ? 8 : (int)need;
}

static void mg_finish(grmp_server *S, Job *j, const MgStatus h, grmp_ctx *c)
{
    j->res.status = h.status;
    j->res.iterations = h.it;
    j->res.outer = 1;
    j->res.relres = h.relres;
    j->res.fp64_fallback = 0;
    j->res.ndofs = j->n;
    // the solution's copy to the host is part of the request: a failed copy fails it
    if (j->req.x_out &&
        hipMemcpyAsync(j->req.x_out, mg_job_solution(j->mg), j->n * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess &&
        j->res.status == GRMP_OK)
        j->res.status = GRMP_ERR_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess && j->res.status == GRMP_OK) j->res.status = GRMP_ERR_HIP;
    double sec = 0.0, bytes = 0.0;
    int64_t nl = 0;
    mg_job_profile(j->mg, &sec, &bytes, &nl);
    std::lock_guard<std::mutex> g(S->mu);
    S->prof_s += sec;
    S->prof_bytes += bytes;
    S->prof_launches += nl;
    mg_release(S, j);
    j->res.t_done = now(S);
}

static void mg_fatal(grmp_server *S, std::vector<Job *> &act, int rc)
{
    for (grmp_ctx *c : S->mg_ctx) (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(S->sctx->stream);
    std::lock_guard<std::mutex> g(S->mu);
    S->fatal = rc;
    S->fatal_msg = S->sctx->err;
    for (Job *&j : act) {
        if (!j) continue;
        mg_release(S, j);
        j->res.status = rc;
        j->res.t_done = now(S);
        j->state = 3;
        j = nullptr;
    }
    for (Job *j : S->queue) {
        j->res.status = rc;
        j->res.t_done = now(S);
        j->state = 3;
    }
    S->queue.clear();
    S->cv_done.notify_all();
}

static void mg_loop(grmp_server *S, int tid, int nthreads)
{
    const int M = (int)S->mg_ctx.size();
    std::vector<Job *> act(M, nullptr);  // this thread uses the slots s % nthreads == tid
    for (;;) {
        int nact = 0;
        for (Job *j : act) nact += j != nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (nact == 0 && S->queue.empty()) {
                if (S->stop) return;
                S->cv.wait(lk, [&] { return S->stop || !S->queue.empty(); });
                continue;
            }
        }
        // admit into free slots, in FIFO order, within the memory budget
        for (int s = tid; s < M; s += nthreads) {
            if (act[s]) continue;
            Job *j = nullptr;
            {
                std::lock_guard<std::mutex> g(S->mu);
                if (S->queue.empty()) break;
                Job *f = S->queue.front();
                if (S->mem_used > 0 && S->mem_used + f->footprint > S->cfg.memory_budget) break;
                S->queue.pop_front();
                S->mem_used += f->footprint;
                j = f;
            }
            const int rc = mg_start(S, j, s);
            if (rc == GRMP_ERR_HIP) {
                act[s] = j;
                mg_fatal(S, act, rc);
                return;
            }
            if (rc) {
                if (j->mg) (void)hipStreamSynchronize(S->mg_ctx[s]->stream);
                std::lock_guard<std::mutex> g(S->mu);
                mg_release(S, j);
                j->res.status = rc;
                j->res.t_done = now(S);
                j->state = 3;
                S->cv_done.notify_all();
                continue;
            }
            act[s] = j;
        }
        // finished chunks: enqueue the next one or hand out the result
        bool progress = false;
        for (int s = tid; s < M; s += nthreads) {
            Job *j = act[s];
            if (!j || !mg_job_status_ready(j->mg)) continue;
            progress = true;
            mg_job_collect(j->mg);
            const MgStatus h = S->mg_host[s];
            if (h.done != 1) {
                int rc = GRMP_OK;
                if (h.done == 2) {
                    // the true-residual check; a request checked in place by its
                    // one-workgroup kernel does it at the start of its next launch
                    rc = mg_job_verify(j->mg);
                    if (!rc && mg_job_block(j->mg)) rc = mg_job_iterate(j->mg, 1);
                } else {
                    rc = mg_job_iterate(j->mg, mg_chunk(S, j, h));
                }
                j->mg_it = h.it;
                j->mg_rel = h.relres;
                if (!rc) rc = mg_job_status_async(j->mg, &S->mg_host[s]);
                if (rc) {
                    mg_fatal(S, act, rc);
                    return;
                }
                continue;
            }
            mg_finish(S, j, S->mg_host[s], S->mg_ctx[s]);
            std::lock_guard<std::mutex> g(S->mu);
            j->state = 3;
            act[s] = nullptr;
            S->cv_done.notify_all();
        }
        if (!progress)
            for (Job *j : act)
                if (j) {
                    (void)mg_job_wait(j->mg);
                    break;
                }
        std::lock_guard<std::mutex> g(S->tpl_mu);
        templates_trim(S, S->cfg.memory_budget / 2);
    }
}

// Lockstep batches (mg_batch > 0).  mg_threads setup threads, each with its
// own stream, take requests from the queue while the requests set up, ready
// and in the batch stay within mg_batch; they build the shapes and levels and
// hand the request on with an event.  The batch thread (one stream) adds
// ready requests to the batch, which advances kMgChunk iterations per run;
// requests leave between runs when done.
static void mg_setup_fail(grmp_server *S, Job *j, int rc)
{
    // with S->mu held
    mg_release(S, j);
    j->res.status = rc;
    j->res.t_done = now(S);
    j->state = 3;
}

static void mg_setup_loop(grmp_server *S, int tid)
{
    grmp_ctx *c = S->mg_ctx[tid];
    auto room = [&](Job *f) {
        return S->mg_in_batch + (int)S->mg_readyq.size() + S->mg_setting_up < S->cfg.mg_batch &&
               !(S->mem_used > 0 && S->mem_used + f->footprint > S->cfg.memory_budget);
    };
    for (;;) {
        // a group of queued requests, set up together: their levels are
        // assembled by one launch (mg_jobs_setup_batched)
        std::vector<Job *> grp;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] {
                if (S->stop && S->queue.empty()) return true;
                if (S->queue.empty() || S->fatal) return S->stop;
                return room(S->queue.front());
            });
            if (S->queue.empty() || S->fatal) return;
            while (!S->queue.empty() && (int)grp.size() < kMgSetupGroup && room(S->queue.front())) {
                Job *j = S->queue.front();
                S->queue.pop_front();
                S->mem_used += j->footprint;
                S->mg_setting_up++;
                grp.push_back(j);
            }
        }
        std::vector<int> rcs(grp.size(), GRMP_OK);
        std::vector<MgJob *> ok;
        std::vector<double> rtol;
        for (size_t k = 0; k < grp.size(); ++k) {
            rcs[k] = mg_build(S, grp[k], c, false);
            if (!rcs[k]) {
                ok.push_back(grp[k]->mg);
                rtol.push_back(grp[k]->req.rtol > 0 ? grp[k]->req.rtol : 1e-10);
            }
        }
        int grc = mg_jobs_setup_batched(ok.data(), (int)ok.size(), c->stream, rtol.data(), kMgMaxIt);
        for (size_t k = 0; k < grp.size(); ++k) {
            Job *j = grp[k];
            if (rcs[k]) continue;
            if (grc) {
                rcs[k] = grc;
                continue;
            }
            j->state = 2;
            if (hipEventCreateWithFlags(&j->mg_ready, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(j->mg_ready, c->stream) != hipSuccess)
                rcs[k] = GRMP_ERR_HIP;
        }
        bool any_fail = false;
        for (int rc : rcs) any_fail |= rc != GRMP_OK;
        if (any_fail) (void)hipStreamSynchronize(c->stream);
        if (S->fatal) {  // the batch ended on a device error meanwhile
            (void)hipStreamSynchronize(c->stream);
            for (int &rc : rcs)
                if (!rc) rc = S->fatal;
        }
        std::lock_guard<std::mutex> g(S->mu);
        for (size_t k = 0; k < grp.size(); ++k) {
            Job *j = grp[k];
            S->mg_setting_up--;
            if (rcs[k]) {
                mg_setup_fail(S, j, rcs[k]);
                continue;
            }
            S->mg_readyq.push_back(j);
        }
        S->cv_done.notify_all();
        S->cv.notify_all();
        S->cv_ready.notify_all();
    }
}

static void mg_batch_loop(grmp_server *S)
{
    grmp_ctx *c = S->mg_ctx[0];
    MgBatch *B = nullptr;
    std::vector<Job *> act;
    int rc = mg_batch_create(c, c->stream, S->cfg.profile != 0, &B);
    if (rc) {
        mg_fatal(S, act, rc);
        S->cv.notify_all();
        return;
    }
    for (;;) {
        // without setup threads the batch thread sets requests up itself, between runs
        while (S->cfg.mg_threads == 0 && (int)act.size() < S->cfg.mg_batch) {
            Job *j = nullptr;
            {
                std::lock_guard<std::mutex> g(S->mu);
                if (S->queue.empty() || S->fatal) break;
                Job *f = S->queue.front();
                if (S->mem_used > 0 && S->mem_used + f->footprint > S->cfg.memory_budget) break;
                S->queue.pop_front();
                S->mem_used += f->footprint;
                S->mg_setting_up++;
                j = f;
            }
            rc = mg_prepare(S, j, c, false);
            std::lock_guard<std::mutex> g(S->mu);
            S->mg_setting_up--;
            if (rc == GRMP_ERR_HIP) {
                act.push_back(j);
                S->mg_in_batch++;
                break;
            }
            if (rc) {
                (void)hipStreamSynchronize(c->stream);
                mg_release(S, j);
                j->res.status = rc;
                j->res.t_done = now(S);
                j->state = 3;
                S->cv_done.notify_all();
                rc = GRMP_OK;
                continue;
            }
            act.push_back(j);
            S->mg_in_batch++;
        }
        if (rc) break;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (act.empty() && S->cfg.mg_threads == 0) {
                if (S->stop && S->queue.empty()) break;
                if (S->queue.empty()) S->cv.wait(lk, [&] { return S->stop || !S->queue.empty(); });
                continue;
            }
            if (act.empty())
                S->cv_ready.wait(lk, [&] {
                    return !S->mg_readyq.empty() || (S->stop && S->queue.empty() && S->mg_setting_up == 0);
                });
            if (act.empty() && S->mg_readyq.empty()) break;
            while (!S->mg_readyq.empty()) {
                Job *j = S->mg_readyq.front();
                S->mg_readyq.pop_front();
                act.push_back(j);
                S->mg_in_batch++;
                // the batch stream waits for the job's setup, then owns it
                (void)hipStreamWaitEvent(c->stream, j->mg_ready, 0);
                mg_job_set_stream(j->mg, c->stream);
            }
        }
        std::vector<MgJob *> jobs(act.size());
        for (size_t k = 0; k < act.size(); ++k) jobs[k] = act[k]->mg;
        rc = mg_batch_run(B, jobs.data(), (int)jobs.size(), kMgChunk);
        if (!rc) rc = mg_batch_wait(B);
        if (rc) break;
        if (S->cfg.profile) {  // the batch is the only writer of the totals in this mode
            double sec = 0.0, bytes = 0.0;
            int64_t nl = 0;
            mg_batch_profile(B, &sec, &bytes, &nl);
            std::lock_guard<std::mutex> g(S->mu);
            S->prof_s = sec;
            S->prof_bytes = bytes;
            S->prof_launches = nl;
        }
        const MgStatus *st = mg_batch_status(B);
        std::vector<Job *> keep;
        for (size_t k = 0; k < act.size(); ++k) {
            Job *j = act[k];
            const MgStatus h = st[k];
            if (h.done == 1) {
                mg_finish(S, j, h, c);
                std::lock_guard<std::mutex> g(S->mu);
                j->state = 3;
                S->mg_in_batch--;
                S->cv_done.notify_all();
                S->cv.notify_all();
                continue;
            }
            if (h.done == 2 && (rc = mg_job_verify(j->mg))) break;
            keep.push_back(j);
        }
        if (rc) break;
        act.swap(keep);
        std::lock_guard<std::mutex> g(S->tpl_mu);
        templates_trim(S, S->cfg.memory_budget / 2);
    }
    if (rc) {
        {
            std::lock_guard<std::mutex> g(S->mu);
            S->mg_in_batch -= (int)act.size();
            for (Job *j : S->mg_readyq) act.push_back(j);  // set up, not yet in the batch
            S->mg_readyq.clear();
        }
        mg_fatal(S, act, rc);
        S->cv.notify_all();
    }
    mg_batch_destroy(B);
}

// ------------------------------------------------------------- plain loop
static int plain_one(grmp_server *S, Job *j)
{
    grmp_ctx *c = S->sctx;
    j->res.t_admit = now(S);
    Template *t = nullptr;
    GRMP_TRY(template_make(S, c, j->req, &t));
    // the pattern above used this request's own coefficient field (same seed)
    grmp_matrix *A = t->A;
    int rc = grmp_matrix_assemble(A, GRMP_VALUES_FP64);
    grmp_vector *b = nullptr, *x = nullptr;
    if (!rc) rc = grmp_vector_create(c, A->n, &b);
    if (!rc) rc = grmp_vector_create(c, A->n, &x);
    if (!rc) rc = linearform_launch(c, t->space, &t->lf, b->d, c->stream);
    if (!rc) rc = zero_fixed_launch(A, b->d, c->stream);
    grmp_solve_opts o;
    grmp_solve_opts_default(&o);
    o.rtol = j->req.rtol > 0 ? j->req.rtol : 1e-10;
    grmp_solve_info info;
    memset(&info, 0, sizeof(info));
    if (!rc) rc = plain_cg_solve(A, b->d, x->d, o, info);
    if (!rc && j->req.x_out) rc = grmp_vector_download(x, j->req.x_out);
    j->res.status = rc ? rc : info.status;
    j->res.iterations = info.iterations;
    j->res.outer = 1;
    j->res.relres = info.relres;
    j->res.ndofs = A->n;
    grmp_vector_destroy(b);
    grmp_vector_destroy(x);
    template_free(t);
    (void)hipStreamSynchronize(c->stream);
    j->res.t_done = now(S);
    return GRMP_OK;
}

static void plain_loop(grmp_server *S)
{
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->queue.empty()) return;
            j = S->queue.front();
            S->queue.pop_front();
        }
        int rc = plain_one(S, j);
        std::lock_guard<std::mutex> g(S->mu);
        if (rc) j->res.status = rc;
        j->state = 3;
        S->cv_done.notify_all();
    }
}

// ------------------------------------------------------------------ API
extern "C" int grmp_server_create(grmp_ctx *ctx, const grmp_server_config *cfg, grmp_server **out)
{
    GRMP_CHECK(ctx, ctx && out, GRMP_ERR_ARG, "grmp_server_create: null argument");
    grmp_server *S = new grmp_server();
    S->user = ctx;
    S->cfg.mode = 1;
    S->cfg.memory_budget = 64ll << 30;
    S->cfg.max_active = 256;
    S->cfg.graph = 1;
    S->cfg.profile = 0;
    S->cfg.tile_rows = (int)kTileRows;
    S->cfg.fp64_first_sweep = 0;
    S->cfg.mg_slots = 8;
    S->cfg.mg_threads = 4;
    S->cfg.mg_block_rows = 2048;
    S->cfg.mg_coarse_rows = 2048;
    S->cfg.mg_batch = 256;
    if (cfg) {
        if (cfg->mode) S->cfg.mode = cfg->mode;
        if (cfg->memory_budget > 0) S->cfg.memory_budget = cfg->memory_budget;
        if (cfg->max_active > 0) S->cfg.max_active = std::min(cfg->max_active, 4096);
        S->cfg.graph = cfg->graph;
        S->cfg.profile = cfg->profile;
        if (cfg->tile_rows > 0) S->cfg.tile_rows = std::max(64, std::min(cfg->tile_rows, 8192));
        if (cfg->fp64_first_sweep >= 0) S->cfg.fp64_first_sweep = cfg->fp64_first_sweep;
        if (cfg->mg_slots > 0) S->cfg.mg_slots = std::min(cfg->mg_slots, 64);
        if (cfg->mg_threads > 0) S->cfg.mg_threads = std::min(cfg->mg_threads, 16);
        if (cfg->mg_threads < 0) S->cfg.mg_threads = 0;  // batches: set up inline
        if (cfg->mg_block_rows > 0) S->cfg.mg_block_rows = cfg->mg_block_rows;
        if (cfg->mg_coarse_rows > 0) S->cfg.mg_coarse_rows = cfg->mg_coarse_rows;
        if (cfg->mg_batch >= 0) S->cfg.mg_batch = std::min(cfg->mg_batch, 4096);
    }
    // lockstep batches: two setup threads by default.  On the 4096-request
    // stream they served 92.9M DoF-solves/s against 91.8M with one and 85.7M
    // with four, whose HIP calls and kernels contend with the batch's.
    if (S->cfg.mg_batch > 0 && (!cfg || cfg->mg_threads == 0)) S->cfg.mg_threads = 2;
    if (S->cfg.mode < 1 || S->cfg.mode > 3) {
        const int mode = S->cfg.mode;
        delete S;
        return grmp_fail(ctx, GRMP_ERR_ARG, "grmp_server_create: mode %d", mode);
    }
    int rc = grmp_ctx_create(ctx->device, &S->sctx);
    if (rc) { delete S; return grmp_fail(ctx, rc, "grmp_server_create: no server context"); }
    if (hipStreamCreateWithFlags(&S->st_solve, hipStreamNonBlocking) != hipSuccess) {
        grmp_ctx_destroy(S->sctx);
        delete S;
        return grmp_fail(ctx, GRMP_ERR_HIP, "grmp_server_create: stream");
    }
    if (S->cfg.mode == 1) {
        S->eng = new KEngine(S->sctx, S->st_solve, S->cfg.max_active, 1 << 20, S->cfg.tile_rows);
        S->eng->set_profile(S->cfg.profile != 0);
        if ((rc = S->eng->init())) {
            delete S->eng;
            (void)hipStreamDestroy(S->st_solve);
            grmp_ctx_destroy(S->sctx);
            delete S;
            return grmp_fail(ctx, rc, "grmp_server_create: engine init failed");
        }
    }
    if (S->cfg.mode == 3) {
        // lockstep batches use slot 0 for the batch and slots 1 .. mg_threads for setup
        if (S->cfg.mg_batch > 0) S->cfg.mg_slots = std::max(S->cfg.mg_slots, S->cfg.mg_threads + 1);
        for (int k = 0; k < S->cfg.mg_slots && !rc; ++k) {
            grmp_ctx *c = nullptr;
            rc = grmp_ctx_create(ctx->device, &c);
            if (!rc) S->mg_ctx.push_back(c);
        }
        if (!rc && hipHostMalloc((void **)&S->mg_host, sizeof(MgStatus) * S->cfg.mg_slots, hipHostMallocDefault) !=
                       hipSuccess)
            rc = GRMP_ERR_HIP;
        if (rc) {
            for (grmp_ctx *c : S->mg_ctx) grmp_ctx_destroy(c);
            (void)hipStreamDestroy(S->st_solve);
            grmp_ctx_destroy(S->sctx);
            delete S;
            return grmp_fail(ctx, rc, "grmp_server_create: multigrid streams");
        }
        memset(S->mg_host, 0, sizeof(MgStatus) * S->cfg.mg_slots);
        S->cfg.mg_threads = S->cfg.mg_batch > 0 ? std::min(S->cfg.mg_threads, S->cfg.mg_slots - 1)
                                                 : std::max(1, std::min(S->cfg.mg_threads, S->cfg.mg_slots));
    }
    S->t0 = wall_seconds();
    S->worker = std::thread([S] {
        (void)hipSetDevice(S->sctx->device);
        if (S->cfg.mode == 1) served_loop(S);
        else if (S->cfg.mode == 3 && S->cfg.mg_batch > 0) mg_batch_loop(S);
        else if (S->cfg.mode == 3) mg_loop(S, 0, S->cfg.mg_threads);
        else plain_loop(S);
    });
    for (int t = 1; S->cfg.mode == 3 && S->cfg.mg_batch == 0 && t < S->cfg.mg_threads; ++t)
        S->mg_workers.emplace_back([S, t] {
            (void)hipSetDevice(S->sctx->device);
            mg_loop(S, t, S->cfg.mg_threads);
        });
    // lockstep batches: setup threads on slots 1 .. mg_threads (slot 0 is the batch's)
    for (int t = 1; S->cfg.mode == 3 && S->cfg.mg_batch > 0 && t <= S->cfg.mg_threads; ++t)
        S->mg_workers.emplace_back([S, t] {
            (void)hipSetDevice(S->sctx->device);
            mg_setup_loop(S, t);
        });
    *out = S;
    return GRMP_OK;
}

extern "C" int grmp_server_submit(grmp_server *S, const grmp_request *req, int64_t *ticket)
{
    if (!S || !req || !ticket) return GRMP_ERR_ARG;
    GRMP_CHECK(S->user, req->N >= 1 && req->N <= 4096 && (req->kind == GRMP_REQ_POISSON || req->kind == GRMP_REQ_ELASTICITY) &&
                            (req->fe == GRMP_FE_H1P1 || req->fe == GRMP_FE_H1P2),
               GRMP_ERR_ARG, "grmp_server_submit: malformed request");
    Job *j = new Job();
    j->req = *req;
    memset(&j->res, 0, sizeof(j->res));
    // multigrid jobs hold the hierarchy's values and vectors besides the fine level's
    j->footprint = grmp_request_footprint(req) * (S->cfg.mode == 3 ? 2 : 1);
    if (S->cfg.mode != 2 && j->footprint > S->cfg.memory_budget) {
        delete j;
        return grmp_fail(S->user, GRMP_ERR_ADMISSION, "grmp_server_submit: footprint %lld B exceeds the budget %lld B",
                         (long long)grmp_request_footprint(req), (long long)S->cfg.memory_budget);
    }
    std::lock_guard<std::mutex> g(S->mu);
    if (S->fatal) {
        delete j;
        return grmp_fail(S->user, S->fatal, "grmp_server_submit: server stopped after a device error: %s",
                         S->fatal_msg.c_str());
    }
    j->ticket = S->next_ticket++;
    j->res.t_submit = now(S);
    S->jobs[j->ticket] = j;
    S->queue.push_back(j);
    *ticket = j->ticket;
    S->cv.notify_all();
    return GRMP_OK;
}

extern "C" int grmp_server_poll(grmp_server *S, int64_t ticket, grmp_result *res)
{
    if (!S) return GRMP_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->jobs.find(ticket);
    if (it == S->jobs.end()) return GRMP_ERR_ARG;
    if (it->second->state != 3) return GRMP_ERR_BUSY;
    if (res) *res = it->second->res;
    return GRMP_OK;
}

extern "C" int grmp_server_wait(grmp_server *S, int64_t ticket, grmp_result *res)
{
    if (!S) return GRMP_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(ticket);
    if (it == S->jobs.end()) return GRMP_ERR_ARG;
    Job *j = it->second;
    S->cv_done.wait(lk, [&] { return j->state == 3; });
    if (res) *res = j->res;
    return GRMP_OK;
}

extern "C" double grmp_server_clock(const grmp_server *S) { return S ? now(S) : 0.0; }

extern "C" int grmp_server_stats(grmp_server *S, grmp_kernel_stats *out)
{
    if (!S || !out) return GRMP_ERR_ARG;
    memset(out, 0, sizeof(*out));
    std::lock_guard<std::mutex> g(S->mu);
    std::lock_guard<std::mutex> gt(S->tpl_mu);
    out->templates = S->ntemplates;
    out->template_seconds = S->template_s;
    if (S->cfg.mode == 3) {
        out->spmv_seconds = S->prof_s;
        out->spmv_launches = S->prof_launches;
        out->spmv_bytes = S->prof_bytes;
        return GRMP_OK;
    }
    if (!S->eng) return GRMP_OK;
    return S->eng->stats(&out->spmv_seconds, &out->spmv_launches, &out->spmv_bytes, &out->steps);
}

extern "C" int grmp_server_destroy(grmp_server *S)
{
    if (!S) return GRMP_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->cv.notify_all();
        S->cv_ready.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    for (std::thread &w : S->mg_workers)
        if (w.joinable()) w.join();
    (void)hipStreamSynchronize(S->st_solve);
    delete S->eng;
    for (auto &kv : S->templates) template_free(kv.second);
    for (auto &kv : S->jobs) delete kv.second;
    for (grmp_ctx *c : S->mg_ctx) grmp_ctx_destroy(c);
    if (S->mg_host) (void)hipHostFree(S->mg_host);
    (void)hipStreamDestroy(S->st_solve);
    grmp_ctx_destroy(S->sctx);
    delete S;
    return GRMP_OK;
}

// Uneven batches: one request list over several shared patterns, solved by
// the mixed-precision path (FP32 inner, FP64 refinement).  Included at the
// end of solver.hip (uses its internals).
//
// A plan groups the requests by pattern once (host), uploads the per-pattern
// job lists (system index in the pattern's arrays, request index for the
// results), allocates the per-pattern workspaces and one stream per pattern.
// A solve launches, per pattern, the persistent factor kernel and the
// persistent mixed solve kernel over that pattern's jobs; the per-pattern
// chains are independent branches (forked from and joined back into the
// caller's stream), so the persistent kernels of all patterns share the
// machine, each taking work from its own counter until its jobs run out.
// The whole launch sequence is captured into a hipGraph on the first solve
// and replayed while the arguments stay the same.
//
// Per pattern the plan runs either the segregated chain (factor kernel, FP32
// factors / ISAI values through HBM, solve kernel) or the merged solve kernel
// alone (each wave factors its system in LDS and solves it: the factors never
// leave the CU, at the cost of the larger per-wave LDS region).  The
// autotuner times the candidates of each pattern -- segregated with {2, 4, 8}
// waves per factor workgroup x {4, 8, 12} per solve workgroup, merged with
// {4, 8, 12} (capped by the kernel's register budget), and for the ISAI
// templates of 48 and 72 slots the solve kernel's slot records in LDS or in
// global memory (the LDS they free makes room for more waves) -- first alone on the pattern's requests, then, pattern by
// pattern from the slowest, inside the whole launch with every other pattern
// running its current choice, and keeps the fastest.  Results do not depend on
// the choice (every system is processed by one wave with the same arithmetic).

struct bilu_uneven_s {
    int npat = 0, nreq = 0, device = 0;
    std::vector<bilu_pattern_s *> pats;
    std::vector<int> count;          // jobs per pattern
    int *d_jobs = nullptr;           // per pattern: sysidx[count] then outidx[count]
    std::vector<size_t> job_off;
    bilu_opts o{};
    std::vector<int> fwaves, swaves; // tuned waves per workgroup (factor, solve; merged: solve only)
    std::vector<int> merged;         // per pattern: 1 = the merged kernel
    std::vector<int> recg;           // per pattern: 1 = the solve kernel's slot records in global memory
    struct Cand { int merged, fw, sw, recg; float alone_ms, corun_ms; };
    std::vector<std::vector<Cand>> table;   // per pattern: the candidates of the last tuning
    std::vector<void *> ws;          // per pattern: FP32 factors / ISAI [count * nnz], then flags [count]
    std::vector<size_t> fbytes;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;  // [0] fork, [1 + p] join of branch p, [npat + 1 / + 2] caller links
    hipStream_t cap = nullptr;       // origin stream of the graph (capture is not allowed on the null stream)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void *> captured;   // arguments of the captured graph
};

namespace {

void free_uneven(bilu_uneven_s *u)
{
    if (!u) return;
    if (u->exec) (void)hipGraphExecDestroy(u->exec);
    if (u->graph) (void)hipGraphDestroy(u->graph);
    for (auto s : u->streams) (void)hipStreamDestroy(s);
    if (u->cap) (void)hipStreamDestroy(u->cap);
    for (auto e : u->events) (void)hipEventDestroy(e);
    for (auto w : u->ws)
        if (w) (void)hipFree(w);
    if (u->d_jobs) (void)hipFree(u->d_jobs);
    delete u;
}

// the factor + solve chain of pattern p on stream st (mg: the merged kernel alone)
int uneven_branch(bilu_uneven_s &u, int p, const double *A, const double *b, double *x, float *it, int *cv,
                  double *ir, double *tr, int *rf, hipStream_t st, int fw, int sw, int mg, int rg)
{
    bilu_pattern_s &h = *u.pats[p];
    const int cnt = u.count[p];
    if (cnt == 0) return BILU_OK;
    const bool isai = u.o.precond == BILU_PRECOND_ILU0_ISAI || u.o.precond == BILU_PRECOND_PARILU0_ISAI;
    const bool par = u.o.precond == BILU_PRECOND_PARILU0 || u.o.precond == BILU_PRECOND_PARILU0_ISAI;
    Jobs jobs;
    jobs.sysidx = u.d_jobs + u.job_off[p];
    jobs.outidx = jobs.sysidx + cnt;
    MixedParams mp{u.o.tol, u.o.tol_mode, (float)u.o.inner_tol, u.o.max_iter, u.o.max_refine, u.o.inner_tol,
                   u.o.inner_adapt};
    if (mg) {
        FactorArgs fa{par ? FS_PARILU : FS_ILU, u.o.parilu_sweeps,
                      par && u.o.parilu_order == BILU_PARILU_INPLACE ? 1 : 0, par ? u.o.parilu_stop_tol : 0.0,
                      nullptr, nullptr, 0, nullptr};
        return launch_mixed(h, cnt, A, b, x, nullptr, nullptr, isai ? PCK_ISAI : PCK_LU, mp, it, cv, ir, tr, rf, st,
                            jobs, sw, &fa, rg);
    }
    float *F = static_cast<float *>(u.ws[p]);
    int *Z = reinterpret_cast<int *>(static_cast<char *>(u.ws[p]) + u.fbytes[p]);
    int rc = launch_factor_fast<float>(h, cnt, A, nullptr, F, par ? FS_PARILU : FS_ILU, u.o.parilu_sweeps,
                                       par ? u.o.parilu_stop_tol : 0.0, nullptr, st, isai ? 1 : 0, F, Z,
                                       jobs.sysidx, fw, nullptr, 0,
                                       par && u.o.parilu_order == BILU_PARILU_INPLACE ? 1 : 0);
    if (rc != BILU_OK) return rc;
    return launch_mixed(h, cnt, A, b, x, F, Z, isai ? PCK_ISAI : PCK_LU, mp, it, cv, ir, tr, rf, st, jobs, sw,
                        nullptr, rg);
}

// all branches: fork from st, one chain per pattern on its stream, join into st
int uneven_launch(bilu_uneven_s &u, const double *const *A, const double *const *b, double *const *x, float *it,
                  int *cv, double *ir, double *tr, int *rf, hipStream_t st)
{
    HIP_TRY(hipEventRecord(u.events[0], st));
    for (int p = 0; p < u.npat; p++) {
        HIP_TRY(hipStreamWaitEvent(u.streams[p], u.events[0], 0));
        const int rc = uneven_branch(u, p, A[p], b[p], x[p], it, cv, ir, tr, rf, u.streams[p], u.fwaves[p],
                                     u.swaves[p], u.merged[p], u.recg[p]);
        if (rc != BILU_OK) return rc;
        HIP_TRY(hipEventRecord(u.events[1 + p], u.streams[p]));
        HIP_TRY(hipStreamWaitEvent(st, u.events[1 + p], 0));
    }
    return BILU_OK;
}

std::vector<const void *> uneven_args(const bilu_uneven_s &u, const double *const *A, const double *const *b,
                                      double *const *x, float *it, int *cv, double *ir, double *tr, int *rf,
                                      hipStream_t st)
{
    std::vector<const void *> a;
    for (int p = 0; p < u.npat; p++) { a.push_back(A[p]); a.push_back(b[p]); a.push_back(x[p]); }
    for (const void *q : {(const void *)it, (const void *)cv, (const void *)ir, (const void *)tr, (const void *)rf})
        a.push_back(q);
    (void)st;
    for (int p = 0; p < u.npat; p++) {
        a.push_back((const void *)(intptr_t)u.fwaves[p]);
        a.push_back((const void *)(intptr_t)u.swaves[p]);
        a.push_back((const void *)(intptr_t)u.merged[p]);
        a.push_back((const void *)(intptr_t)u.recg[p]);
    }
    return a;
}

}  // namespace

extern "C" {

int bilu_uneven_create(int npat, const bilu_pattern_t *patterns, int nreq, const int *req_pattern,
                       const int *req_index, const bilu_opts *opts, bilu_uneven_t *out)
{
    if (npat <= 0 || !patterns || nreq < 0 || (nreq > 0 && (!req_pattern || !req_index)) || !out)
        return BILU_ERR_INVALID_ARG;
    bilu_opts o;
    if (opts) o = *opts; else bilu_opts_default(&o);
    if (o.precision != BILU_PRECISION_FP32_IR ||
        !(o.precond == BILU_PRECOND_ILU0 || o.precond == BILU_PRECOND_PARILU0 ||
          o.precond == BILU_PRECOND_ILU0_ISAI || o.precond == BILU_PRECOND_PARILU0_ISAI) ||
        !(o.inner_tol > 0.0 && o.inner_tol < 1.0) || o.max_refine < 1 || !(o.tol >= 0.0) || o.max_iter < 0 ||
        o.isai_power != 1 || (o.parilu_order != BILU_PARILU_JACOBI && o.parilu_order != BILU_PARILU_INPLACE) ||
        (o.merge != BILU_MERGED && o.merge != BILU_SEGREGATED && o.merge != BILU_MERGE_AUTO) ||
        !(o.inner_adapt >= 0.0 && o.inner_adapt < 1.0))
        return BILU_ERR_INVALID_ARG;
    for (int p = 0; p < npat; p++) {
        if (!patterns[p]) return BILU_ERR_INVALID_ARG;
        for (int q = 0; q < p; q++)
            if (patterns[q] == patterns[p]) return BILU_ERR_INVALID_ARG;   // a handle's counters serve one branch
        const auto &pa = patterns[p]->pa;
        const int nr = (pa.n + WAVE - 1) / WAVE;
        if (!pa.fast.ok || nr > MAX_NR || slot_template_mixed(pa.fast.S) < 0) return BILU_ERR_TOO_LARGE;
    }
    std::vector<std::vector<int>> sys(npat), outi(npat);
    for (int r = 0; r < nreq; r++) {
        const int p = req_pattern[r];
        if (p < 0 || p >= npat || req_index[r] < 0) return BILU_ERR_INVALID_ARG;
        sys[p].push_back(req_index[r]);
        outi[p].push_back(r);
    }
    auto *u = new bilu_uneven_s();
    u->npat = npat;
    u->nreq = nreq;
    u->o = o;
    u->device = patterns[0]->device;
    if (hipSetDevice(u->device) != hipSuccess) { free_uneven(u); return BILU_ERR_HIP; }
    std::vector<int> flat;
    for (int p = 0; p < npat; p++) {
        u->pats.push_back(patterns[p]);
        u->count.push_back(int(sys[p].size()));
        u->job_off.push_back(flat.size());
        flat.insert(flat.end(), sys[p].begin(), sys[p].end());
        flat.insert(flat.end(), outi[p].begin(), outi[p].end());
        u->fwaves.push_back(KF_MAX_WAVES);
        u->swaves.push_back(64);
        u->merged.push_back(o.merge == BILU_MERGED ? 1 : 0);
        u->recg.push_back(0);
    }
    if (hipMalloc((void **)&u->d_jobs, std::max<size_t>(1, flat.size()) * sizeof(int)) != hipSuccess ||
        (!flat.empty() && hipMemcpy(u->d_jobs, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)) {
        free_uneven(u);
        return BILU_ERR_HIP;
    }
    for (int p = 0; p < npat; p++) {
        const size_t fb = ((size_t)std::max(1, u->count[p]) * patterns[p]->pa.nnz * sizeof(float) + 255) / 256 * 256;
        void *w = nullptr;
        if (hipMalloc(&w, fb + (size_t)std::max(1, u->count[p]) * sizeof(int)) != hipSuccess) { free_uneven(u); return BILU_ERR_HIP; }
        u->ws.push_back(w);
        u->fbytes.push_back(fb);
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) { free_uneven(u); return BILU_ERR_HIP; }
        u->streams.push_back(s);
    }
    if (hipStreamCreateWithFlags(&u->cap, hipStreamNonBlocking) != hipSuccess) { free_uneven(u); return BILU_ERR_HIP; }
    for (int e = 0; e <= npat + 2; e++) {
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) { free_uneven(u); return BILU_ERR_HIP; }
        u->events.push_back(ev);
    }
    *out = u;
    return BILU_OK;
}

int bilu_uneven_destroy(bilu_uneven_t u)
{
    if (!u) return BILU_ERR_INVALID_ARG;
    (void)hipSetDevice(u->device);
    (void)hipDeviceSynchronize();
    free_uneven(u);
    return BILU_OK;
}

int bilu_uneven_tune(bilu_uneven_t u, const double *const *d_A, const double *const *d_b, double *const *d_x,
                     float *d_iter, int *d_conv, double *d_iter_res, double *d_true_res, int *d_refine, void *stream)
{
    if (!u || !d_A || !d_b || !d_x || !d_iter || !d_conv) return BILU_ERR_INVALID_ARG;
    HIP_TRY(hipSetDevice(u->device));
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // candidates of a pattern: segregated (factor waves, solve waves) and merged
    // (solve waves), each with the solve kernel's slot records in LDS and, for
    // the ISAI templates of 48 and 72 slots, in global memory
    const bool isai = u->o.precond == BILU_PRECOND_ILU0_ISAI || u->o.precond == BILU_PRECOND_PARILU0_ISAI;
    auto cands_of = [&](int p) {
        std::vector<bilu_uneven_s::Cand> cands;
        const auto &pl = u->pats[p]->pa.fast;
        int sm = 0, slu = 0;
        const bool place = isai && u->pats[p]->d_rec32 && isai_tier(pl.S, pl.SL, pl.SU, sm, slu) && sm >= 48;
        for (int rg = 0; rg <= (place ? 1 : 0); rg++) {
            if (u->o.merge != BILU_MERGED)
                for (int fw : {2, 4, 8})
                    for (int sw : {4, 8, 12}) cands.push_back({0, fw, sw, rg, -1.0f, -1.0f});
            if (u->o.merge != BILU_SEGREGATED)
                for (int sw : {4, 8, 12}) cands.push_back({1, 0, sw, rg, -1.0f, -1.0f});
        }
        return cands;
    };
    // one untimed run (first-use costs), then the timed one
    auto timed = [&](auto &&launch, float &ms) -> int {
        for (int rep = 0; rep < 2; rep++) {
            if (rep == 1 && hipEventRecord(e0, st) != hipSuccess) return BILU_ERR_HIP;
            const int rc = launch();
            if (rc != BILU_OK) return rc;
        }
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            return BILU_ERR_HIP;
        return BILU_OK;
    };
    int rc = BILU_OK;
    u->table.assign(u->npat, {});
    // 1: every pattern alone on its own requests
    std::vector<std::pair<float, int>> order;
    for (int p = 0; p < u->npat && rc == BILU_OK; p++) {
        if (u->count[p] == 0) continue;
        auto &tab = u->table[p];
        tab = cands_of(p);
        float best = 1e30f;
        for (size_t c = 0; c < tab.size() && rc == BILU_OK; c++) {
            float ms = 0.0f;
            rc = timed([&]() {
                return uneven_branch(*u, p, d_A[p], d_b[p], d_x[p], d_iter, d_conv, d_iter_res, d_true_res,
                                     d_refine, st, tab[c].fw, tab[c].sw, tab[c].merged, tab[c].recg);
            }, ms);
            tab[c].alone_ms = ms;
            if (rc == BILU_OK && ms < best) {
                best = ms;
                u->fwaves[p] = tab[c].fw ? tab[c].fw : u->fwaves[p];
                u->swaves[p] = tab[c].sw;
                u->merged[p] = tab[c].merged;
                u->recg[p] = tab[c].recg;
            }
        }
        order.push_back({best, p});
    }
    // 2: pattern by pattern, slowest first, inside the whole launch (the other
    // patterns' branches running concurrently with their current choice)
    std::sort(order.begin(), order.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    for (const auto &op : order) {
        if (rc != BILU_OK || u->npat < 2) break;
        const int p = op.second;
        auto &tab = u->table[p];
        float best = 1e30f;
        bilu_uneven_s::Cand keep{u->merged[p], u->fwaves[p], u->swaves[p], u->recg[p], 0.0f, 0.0f};
        for (size_t c = 0; c < tab.size() && rc == BILU_OK; c++) {
            u->fwaves[p] = tab[c].fw ? tab[c].fw : keep.fw;
            u->swaves[p] = tab[c].sw;
            u->merged[p] = tab[c].merged;
            u->recg[p] = tab[c].recg;
            float ms = 0.0f;
            rc = timed([&]() {
                return uneven_launch(*u, d_A, d_b, d_x, d_iter, d_conv, d_iter_res, d_true_res, d_refine, st);
            }, ms);
            tab[c].corun_ms = ms;
            if (rc == BILU_OK && ms < best) {
                best = ms;
                keep = {tab[c].merged, u->fwaves[p], tab[c].sw, tab[c].recg, 0.0f, 0.0f};
            }
        }
        u->fwaves[p] = keep.fw;
        u->swaves[p] = keep.sw;
        u->merged[p] = keep.merged;
        u->recg[p] = keep.recg;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int bilu_uneven_config(bilu_uneven_t u, int p, int *factor_waves, int *solve_waves, int *jobs)
{
    if (!u || p < 0 || p >= u->npat) return BILU_ERR_INVALID_ARG;
    if (factor_waves) *factor_waves = u->merged[p] ? 0 : u->fwaves[p];
    if (solve_waves) *solve_waves = u->swaves[p];
    if (jobs) *jobs = u->count[p];
    return BILU_OK;
}

int bilu_uneven_merge_point(bilu_uneven_t u, int p, int *merged)
{
    if (!u || p < 0 || p >= u->npat || !merged) return BILU_ERR_INVALID_ARG;
    *merged = u->merged[p];
    return BILU_OK;
}

int bilu_uneven_configure(bilu_uneven_t u, int p, int merged, int factor_waves, int solve_waves, int records_global)
{
    if (!u || p < 0 || p >= u->npat || (merged != 0 && merged != 1) || solve_waves < 1 ||
        (!merged && factor_waves < 1) || (records_global != 0 && records_global != 1))
        return BILU_ERR_INVALID_ARG;
    if ((merged && u->o.merge == BILU_SEGREGATED) || (!merged && u->o.merge == BILU_MERGED)) return BILU_ERR_INVALID_ARG;
    u->merged[p] = merged;
    if (!merged) u->fwaves[p] = factor_waves;
    u->swaves[p] = solve_waves;
    u->recg[p] = records_global;
    return BILU_OK;
}

int bilu_uneven_placement(bilu_uneven_t u, int p, int *records_global)
{
    if (!u || p < 0 || p >= u->npat || !records_global) return BILU_ERR_INVALID_ARG;
    *records_global = u->recg[p];
    return BILU_OK;
}

int bilu_uneven_tune_table(bilu_uneven_t u, int p, int cap, int *merged, int *factor_waves, int *solve_waves,
                           int *records_global, float *alone_ms, float *corun_ms, int *count)
{
    if (!u || p < 0 || p >= u->npat || !count || cap < 0) return BILU_ERR_INVALID_ARG;
    const auto &tab = p < (int)u->table.size() ? u->table[p] : std::vector<bilu_uneven_s::Cand>();
    *count = (int)tab.size();
    for (int c = 0; c < (int)tab.size() && c < cap; c++) {
        if (merged) merged[c] = tab[c].merged;
        if (factor_waves) factor_waves[c] = tab[c].fw;
        if (solve_waves) solve_waves[c] = tab[c].sw;
        if (records_global) records_global[c] = tab[c].recg;
        if (alone_ms) alone_ms[c] = tab[c].alone_ms;
        if (corun_ms) corun_ms[c] = tab[c].corun_ms;
    }
    return BILU_OK;
}

int bilu_uneven_solve(bilu_uneven_t u, const double *const *d_A, const double *const *d_b, double *const *d_x,
                      float *d_iter, int *d_conv, double *d_iter_res, double *d_true_res, int *d_refine, void *stream)
{
    if (!u || !d_A || !d_b || !d_x || (u->nreq > 0 && (!d_iter || !d_conv))) return BILU_ERR_INVALID_ARG;
    for (int p = 0; p < u->npat; p++)
        if (u->count[p] > 0 && (!d_A[p] || !d_b[p] || !d_x[p])) return BILU_ERR_INVALID_ARG;
    if (u->nreq == 0) return BILU_OK;
    HIP_TRY(hipSetDevice(u->device));
    hipStream_t st = (hipStream_t)stream;
    const auto args = uneven_args(*u, d_A, d_b, d_x, d_iter, d_conv, d_iter_res, d_true_res, d_refine, st);
    if (!u->exec || args != u->captured) {
        if (u->exec) { (void)hipGraphExecDestroy(u->exec); u->exec = nullptr; }
        if (u->graph) { (void)hipGraphDestroy(u->graph); u->graph = nullptr; }
        HIP_TRY(hipStreamSynchronize(u->cap));
        HIP_TRY(hipStreamBeginCapture(u->cap, hipStreamCaptureModeRelaxed));
        const int rc = uneven_launch(*u, d_A, d_b, d_x, d_iter, d_conv, d_iter_res, d_true_res, d_refine, u->cap);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(u->cap, &g);
        if (rc != BILU_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (ec != hipSuccess || !g) return BILU_ERR_HIP;
        u->graph = g;
        if (hipGraphInstantiate(&u->exec, u->graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(u->graph);
            u->graph = nullptr;
            u->exec = nullptr;
            return BILU_ERR_HIP;
        }
        u->captured = args;
    }
    // ordered after the caller's earlier work on st, and st after the graph
    HIP_TRY(hipEventRecord(u->events[u->npat + 1], st));
    HIP_TRY(hipStreamWaitEvent(u->cap, u->events[u->npat + 1], 0));
    HIP_TRY(hipGraphLaunch(u->exec, u->cap));
    HIP_TRY(hipEventRecord(u->events[u->npat + 2], u->cap));
    HIP_TRY(hipStreamWaitEvent(st, u->events[u->npat + 2], 0));
    return BILU_OK;
}

}  // extern "C"

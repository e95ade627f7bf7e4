// server.cpp -- the request loop: log-det tiles, Foldy-Lax solves with a psi grid and
// resonance searches of mixed N and d, served with continuous batching.
//
// * Admission.  Requests wait in arrival order and are admitted while the device
//   bytes they hold (16 N^2 per complex128 matrix, 8 N^2 per complex64 matrix, times
//   the matrices the request keeps live at once) fit the budget; a request is always
//   admitted when nothing else is admitted, so a request larger than the budget still
//   runs (alone).  A later request that fits may pass one that does not, but only
//   kMaxBypass times: then nothing behind the passed-over request is admitted until it
//   fits (at the latest once the device has drained), so a large request is not starved
//   by a steady stream of small ones.
// * Residency.  Positions live on the device once per medium content: requests over the
//   same medium share that copy, which stays resident after they finish (the kResidentMedia
//   most recently used media).  A complex128 solve keeps its factored M(k) on the device,
//   keyed by medium content, scattering model and k; a later solve at the same point (a
//   different source, another psi grid) solves with it instead of assembling and factoring
//   M(k) again, with the same kernels, so its amplitudes are the same bit for bit.  Kept
//   factors count against the budget (at most half of it) and are dropped, least recently
//   used first, when a request needs the room.
// * Continuous batching.  The batch loop repeatedly forms one step from every
//   admitted log-det and resonance request: each contributes the matrices of its
//   current stage (all tile cells; the FP64 re-evaluation cells of a mixed tile; the
//   background-fit samples or the live roots of an Aberth sweep).  The step runs as at
//   most four concurrent multi-medium batches (trace-log and tr(M^-1 M'), complex64 and
//   complex128), each grouped by padded size on the engine's streams.  Requests that
//   arrive meanwhile join the next step; finished requests leave after it.
// * Solves run on their own thread, overlapping the batch steps on the device.
// Each matrix is computed on its own by the kernels, so a request's results do not
// depend on what else shares its steps.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "aberth.h"
#include "capi_internal.h"

namespace msg {
namespace {

using cd = std::complex<double>;
using clk = std::chrono::steady_clock;

inline cx<double>* as_cx(msgpu_cplx* p) { return reinterpret_cast<cx<double>*>(p); }
inline const cx<double>* as_cx(const msgpu_cplx* p) { return reinterpret_cast<const cx<double>*>(p); }
inline double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// The device copy of one medium's positions, shared by every request over the same content.
struct ResidentMedium {
    uint64_t id = 0;                    // never reused, keys the factors computed over it
    uint64_t hash = 0;
    int d = 0, n = 0;
    std::vector<double> pos;
    const double* dpos = nullptr;
    size_t tok = 0;
    int refs = 0;                       // admitted requests and kept factors using it
    uint64_t last = 0;
};

// A kept complex128 factor of M(k) over a resident medium.
struct KeptFactor {
    ResidentMedium* med = nullptr;
    ScModel m{};
    cd k;
    std::shared_ptr<ResidentFactor> f;
    size_t bytes = 0;
    uint64_t last = 0;
};

// FNV-1a over 64-bit words: a quick first test; equal hashes are confirmed byte for byte.
uint64_t content_hash(int d, int n, const double* pos) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t w) { h ^= w; h *= 1099511628211ull; };
    mix((uint64_t)d);
    mix((uint64_t)n);
    for (size_t i = 0; i < (size_t)n * d; i++) {
        uint64_t w;
        memcpy(&w, pos + i, sizeof w);
        mix(w);
    }
    return h;
}

constexpr int kResidentMedia = 64;

struct Req {
    long long id = 0;
    msgpu_request r{};                  // the caller's request; output pointers stay the caller's
    MediumView mv;
    std::vector<double> hpos;           // host copy of the positions
    ResidentMedium* rmed = nullptr;     // its resident device copy while admitted
    std::vector<cd> ks;                 // LOGDET: copy of the k points
    std::vector<double> stats;          // LOGDET, MSGPU_MIXED_GATED: the complex64 pivot statistics
    int stage = 0;                      // LOGDET: 0 first pass, 1 FP64 re-evaluation
    std::vector<int> reidx;
    std::unique_ptr<AberthRun> run;     // RESONANCE
    size_t footprint = 0;
    int bypassed = 0;                   // later requests admitted while this one waited
    int status = 0, count = 0;
    std::string err;
    bool done = false;
    clk::time_point t_submit, t_admit, t_done;
};

}  // namespace

struct Server {
    std::mutex mu;
    std::condition_variable cv;         // work for the loops, and completions
    std::deque<Req*> pending;
    std::vector<Req*> active;           // admitted LOGDET / RESONANCE
    std::deque<Req*> solveq;            // admitted SOLVE
    std::map<long long, Req*> all;
    size_t budget = 0, in_use = 0, peak = 0;
    int max_active = 64, n_admitted = 0;
    long long next_id = 0, steps = 0, items = 0, max_step_reqs = 0;
    bool stop = false;
    std::thread batch_thread, solve_thread;
    std::vector<std::unique_ptr<ResidentMedium>> media;
    std::vector<KeptFactor> kept;
    size_t kept_bytes = 0;
    uint64_t tick = 0, next_medium = 0;
    long long factor_hits = 0, factor_misses = 0;

    void admit_locked();
    void batch_loop();
    void solve_loop();
    void finish_locked(Req* q);
    ResidentMedium* acquire_medium_locked(const MediumView& mv);
    void release_medium_locked(ResidentMedium* rm);
    void drop_kept_locked(size_t i);
    bool make_room_locked(size_t need);
    std::shared_ptr<ResidentFactor> find_kept_locked(const Req* q);
    void keep_factor_locked(const Req* q, std::shared_ptr<ResidentFactor> f);
    ~Server();
};

namespace {

size_t footprint_of(const msgpu_request& r, int n) {
    const double m2 = (double)n * n;
    const double e = (r.mode == MSGPU_FP64) ? 16.0 : 8.0;
    switch (r.kind) {
        case MSGPU_REQ_LOGDET: return (size_t)(r.nk * m2 * e);
        case MSGPU_REQ_SOLVE: return (size_t)(m2 * (16.0 + (r.mode == MSGPU_FP64 ? 0.0 : 8.0)));
        default: return (size_t)(2.0 * r.nz * m2 * e);
    }
}

}  // namespace

ResidentMedium* Server::acquire_medium_locked(const MediumView& mv) {
    const uint64_t h = content_hash(mv.d, mv.n, mv.pos);
    const size_t cnt = (size_t)mv.n * mv.d;
    for (auto& rm : media)
        if (rm->hash == h && rm->d == mv.d && rm->n == mv.n && memcmp(rm->pos.data(), mv.pos, sizeof(double) * cnt) == 0) {
            rm->refs++;
            rm->last = ++tick;
            return rm.get();
        }
    // drop unused media beyond the kept number, least recently used first
    while ((int)media.size() >= kResidentMedia) {
        int victim = -1;
        for (int i = 0; i < (int)media.size(); i++)
            if (media[i]->refs == 0 && (victim < 0 || media[i]->last < media[victim]->last)) victim = i;
        if (victim < 0) break;
        release_positions(media[victim]->dpos, media[victim]->tok);
        media.erase(media.begin() + victim);
    }
    std::unique_ptr<ResidentMedium> rm(new ResidentMedium());
    rm->id = ++next_medium;
    rm->hash = h;
    rm->d = mv.d;
    rm->n = mv.n;
    rm->pos.assign(mv.pos, mv.pos + cnt);
    MediumView v = mv;
    v.pos = rm->pos.data();
    if (upload_positions(v, &rm->dpos, &rm->tok) < 0) return nullptr;
    rm->refs = 1;
    rm->last = ++tick;
    media.push_back(std::move(rm));
    return media.back().get();
}

void Server::release_medium_locked(ResidentMedium* rm) {
    if (!rm) return;
    rm->refs--;
    rm->last = ++tick;
}

void Server::drop_kept_locked(size_t i) {
    kept_bytes -= kept[i].bytes;
    in_use -= kept[i].bytes;
    release_medium_locked(kept[i].med);
    kept.erase(kept.begin() + (long)i);         // the factor's memory goes when its last user lets go
}

// Drops kept factors, least recently used first, until need more bytes fit the budget.
bool Server::make_room_locked(size_t need) {
    while (in_use + need > budget && !kept.empty()) {
        size_t v = 0;
        for (size_t i = 1; i < kept.size(); i++)
            if (kept[i].last < kept[v].last) v = i;
        drop_kept_locked(v);
    }
    return in_use + need <= budget;
}

std::shared_ptr<ResidentFactor> Server::find_kept_locked(const Req* q) {
    const cd k = q->r.k;
    for (KeptFactor& e : kept)
        if (e.med == q->rmed && e.k == k && e.m.type == q->mv.m.type && e.m.p0 == q->mv.m.p0 &&
            e.m.p1 == q->mv.m.p1) {
            e.last = ++tick;
            return e.f;
        }
    return nullptr;
}

void Server::keep_factor_locked(const Req* q, std::shared_ptr<ResidentFactor> f) {
    const size_t bytes = resident_factor_bytes(*f);
    if (bytes > budget / 2) return;
    while (kept_bytes + bytes > budget / 2 && !kept.empty()) {
        size_t v = 0;
        for (size_t i = 1; i < kept.size(); i++)
            if (kept[i].last < kept[v].last) v = i;
        drop_kept_locked(v);
    }
    if (!make_room_locked(bytes)) return;
    KeptFactor e;
    e.med = q->rmed;
    e.med->refs++;
    e.m = q->mv.m;
    e.k = q->r.k;
    e.f = std::move(f);
    e.bytes = bytes;
    e.last = ++tick;
    kept.push_back(std::move(e));
    kept_bytes += bytes;
    in_use += bytes;
    peak = std::max(peak, in_use);
}

Server::~Server() {
    kept.clear();
    for (auto& rm : media) release_positions(rm->dpos, rm->tok);
}

void Server::finish_locked(Req* q) {
    q->done = true;
    q->t_done = clk::now();
    in_use -= q->footprint;
    n_admitted--;
    release_medium_locked(q->rmed);
    q->hpos.clear();
    q->hpos.shrink_to_fit();
    cv.notify_all();
}

// Arrival order; a request that does not fit is skipped for now and later, smaller
// requests may pass it, at most kMaxBypass times; after that it holds the queue.
constexpr int kMaxBypass = 4;

void Server::admit_locked() {
    std::vector<Req*> passed;           // waiting requests ahead of the current one
    for (auto it = pending.begin(); it != pending.end();) {
        Req* q = *it;
        bool fits = in_use + q->footprint <= budget;
        if (!fits && n_admitted < max_active) fits = make_room_locked(q->footprint);
        if (!fits && n_admitted == 0) fits = true;     // alone, a request larger than the budget still runs
        if (!fits || n_admitted >= max_active) {
            if (q->bypassed >= kMaxBypass) break;      // it has waited long enough: hold the rest
            passed.push_back(q);
            ++it;
            continue;
        }
        q->rmed = acquire_medium_locked(q->mv);
        if (!q->rmed) {
            q->status = -2;
            q->err = last_error();
            q->done = true;
            q->t_admit = q->t_done = clk::now();
            it = pending.erase(it);
            cv.notify_all();
            continue;
        }
        for (Req* w : passed) w->bypassed++;
        q->mv.dpos = q->rmed->dpos;
        q->t_admit = clk::now();
        in_use += q->footprint;
        peak = std::max(peak, in_use);
        n_admitted++;
        if (q->r.kind == MSGPU_REQ_SOLVE) solveq.push_back(q);
        else active.push_back(q);
        it = pending.erase(it);
        cv.notify_all();
    }
}

void Server::solve_loop() {
    for (;;) {
        Req* q = nullptr;
        {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&] { return !solveq.empty() || (stop && pending.empty()); });
            if (solveq.empty()) return;      // stop requested and nothing left to admit
            q = solveq.front();
            solveq.pop_front();
        }
        const msgpu_request& r = q->r;
        int rc;
        if (r.mode == MSGPU_FP64) {            // through the kept factors
            std::shared_ptr<ResidentFactor> f, made;
            {
                std::lock_guard<std::mutex> g(mu);
                f = find_kept_locked(q);
                (f ? factor_hits : factor_misses)++;
            }
            rc = solve_resident(q->mv, cx<double>{r.k.real(), r.k.imag()}, r.source, as_cx(r.a), f,
                                f ? nullptr : &made);
            if (rc >= 0 && made) {
                std::lock_guard<std::mutex> g(mu);
                in_use -= q->footprint;        // the request's matrix becomes the kept factor
                q->footprint = 0;
                keep_factor_locked(q, std::move(made));
            }
        } else {
            rc = solve_amplitudes(q->mv, cx<double>{r.k.real(), r.k.imag()}, r.source, as_cx(r.a), 1, r.tol);
        }
        std::string err = rc < 0 ? last_error() : "";
        if (rc >= 0 && r.psi && r.nx > 0 && r.ny > 0) {
            const int rc2 = wave_grid(q->mv, cx<double>{r.k.real(), r.k.imag()}, r.source, as_cx(r.a), r.nx, r.ny,
                                      r.xmin, r.xmax, r.ymin, r.ymax, as_cx(r.psi));
            if (rc2 < 0) { rc = rc2; err = last_error(); }
        }
        std::lock_guard<std::mutex> g(mu);
        q->status = rc < 0 ? rc : 0;
        q->count = rc < 0 ? 0 : rc;
        q->err = err;
        finish_locked(q);
        admit_locked();
    }
}

void Server::batch_loop() {
    for (;;) {
        std::vector<Req*> step;
        {
            std::unique_lock<std::mutex> g(mu);
            admit_locked();
            cv.wait(g, [&] { return !active.empty() || (stop && pending.empty()); });
            if (active.empty()) return;
            step = active;
        }
        // The step: (op, precision) -> items; op 0 trace-log, 1 tr(M^-1 M').
        struct Batch {
            std::vector<msgpu_cplx> z, f;
            std::vector<double> st;          // pivot statistics (complex64 trace-logs)
            std::vector<int> m;
            std::vector<Req*> owner;
            std::vector<int> slot;
        } bt[2][2];
        std::vector<MediumView> views;
        for (size_t j = 0; j < step.size(); j++) {
            Req* q = step[j];
            views.push_back(q->mv);
            auto add = [&](int op, int prec, const cd& z, int slot) {
                Batch& b = bt[op][prec];
                b.z.push_back(z);
                b.m.push_back((int)j);
                b.owner.push_back(q);
                b.slot.push_back(slot);
            };
            if (q->r.kind == MSGPU_REQ_LOGDET) {
                if (q->stage == 0) {
                    const int prec = q->r.mode == MSGPU_FP64 ? 0 : 1;
                    for (size_t l = 0; l < q->ks.size(); l++) add(0, prec, q->ks[l], (int)l);
                } else {
                    for (int l : q->reidx) add(0, 0, q->ks[l], l);
                }
            } else {
                std::vector<cd> fz = q->run->fit_points();
                if (!fz.empty()) {
                    for (size_t i = 0; i < fz.size(); i++) add(1, 0, fz[i], (int)i);
                } else {
                    int prec = 0;
                    std::vector<cd> sz = q->run->sweep_points(&prec);
                    for (size_t i = 0; i < sz.size(); i++) add(1, prec, sz[i], (int)i);
                }
            }
        }
        int rc[2][2] = {{0, 0}, {0, 0}};
        std::string errs[2][2];
        std::vector<std::future<void>> futs;
        long long nitems = 0;
        for (int op = 0; op < 2; op++)
            for (int p = 0; p < 2; p++) {
                Batch& b = bt[op][p];
                if (b.z.empty()) continue;
                nitems += (long long)b.z.size();
                b.f.resize(b.z.size());
                futs.push_back(std::async(std::launch::async, [&, op, p] {
                    Batch& bb = bt[op][p];
                    if (op == 0 && p == 1) bb.st.resize(2 * bb.z.size());
                    rc[op][p] = op == 0 ? tracelog_multi(views, (int)bb.z.size(), bb.m.data(), as_cx(bb.z.data()),
                                                         as_cx(bb.f.data()), p == 1 ? bb.st.data() : nullptr, p)
                                        : dk_log_det_multi(views, (int)bb.z.size(), bb.m.data(), as_cx(bb.z.data()),
                                                           as_cx(bb.f.data()), p);
                    if (rc[op][p] < 0) errs[op][p] = last_error();
                }));
            }
        for (auto& f : futs) f.get();
        // hand the results back, advance every request of the step
        std::map<Req*, std::vector<cd>> fvals;
        std::map<Req*, int> failed;
        for (int op = 0; op < 2; op++)
            for (int p = 0; p < 2; p++) {
                Batch& b = bt[op][p];
                for (size_t i = 0; i < b.z.size(); i++) {
                    Req* q = b.owner[i];
                    if (rc[op][p] < 0) { failed[q] = rc[op][p]; q->err = errs[op][p]; continue; }
                    if (q->r.kind == MSGPU_REQ_LOGDET) {
                        q->r.out[b.slot[i]] = b.f[i];
                        if (op == 0 && p == 1 && q->r.mode == MSGPU_MIXED_GATED && q->stage == 0) {
                            q->stats.resize(2 * q->ks.size());
                            q->stats[2 * b.slot[i]] = b.st[2 * i];
                            q->stats[2 * b.slot[i] + 1] = b.st[2 * i + 1];
                        }
                    } else {
                        fvals[q].push_back(b.f[i]);
                    }
                }
            }
        std::lock_guard<std::mutex> g(mu);
        steps++;
        items += nitems;
        max_step_reqs = std::max<long long>(max_step_reqs, (long long)step.size());
        for (Req* q : step) {
            bool fin = false;
            if (failed.count(q)) {
                q->status = failed[q];
                fin = true;
            } else if (q->r.kind == MSGPU_REQ_LOGDET) {
                if (q->stage == 0 && (q->r.mode == MSGPU_MIXED || q->r.mode == MSGPU_MIXED_GATED)) {
                    q->reidx = q->r.mode == MSGPU_MIXED
                                   ? reeval_cells(q->mv.n, q->r.nk, q->r.out, q->r.reeval_rel)
                                   : gated_cells(q->r.nk, q->stats.data(), q->r.out, q->r.reeval_rel);
                    q->count = (int)q->reidx.size();
                    q->stage = 1;
                    fin = q->reidx.empty();
                } else {
                    fin = true;
                }
            } else {
                if (!q->run->fit_points().empty()) q->run->take_fit(fvals[q]);
                else if (fvals.count(q)) q->run->take_sweep(fvals[q]);
                if (q->run->finished()) {
                    const std::vector<cd> roots = q->run->roots();
                    *q->r.nroots = (int)roots.size();
                    for (size_t i = 0; i < roots.size(); i++) q->r.roots[i] = roots[i];
                    q->count = q->run->sweeps();
                    fin = true;
                }
            }
            if (fin) {
                active.erase(std::find(active.begin(), active.end(), q));
                finish_locked(q);
            }
        }
        admit_locked();
    }
}

}  // namespace msg

using namespace msg;

extern "C" {

MSGPU_API msgpu_server* msgpu_server_create(double hbm_gb, int max_active) {
    if (engine_init() != 0) return nullptr;
    Server* s = new Server();
    s->budget = (size_t)(std::max(hbm_gb, 0.0) * (double)(1ull << 30));
    s->max_active = max_active > 0 ? max_active : 64;
    s->batch_thread = std::thread([s] { s->batch_loop(); });
    s->solve_thread = std::thread([s] { s->solve_loop(); });
    return reinterpret_cast<msgpu_server*>(s);
}

MSGPU_API long long msgpu_server_submit(msgpu_server* sv, const msgpu_request* r) {
    Server* s = reinterpret_cast<Server*>(sv);
    if (!s || !r || !r->med) { set_error("msgpu_server_submit: null server, request or medium"); return -2; }
    const MediumView mv = view_of(r->med);
    if (mv.d < 1 || mv.d > 4 || mv.n < 1 || mv.n > 8192 || !mv.pos) {
        set_error("msgpu_server_submit: the device path supports d = 1..4 and 1 <= N <= 8192");
        return -3;
    }
    if ((r->mode != MSGPU_FP64 && r->mode != MSGPU_MIXED && r->mode != MSGPU_FP32 &&
         r->mode != MSGPU_MIXED_GATED) ||
        (r->mode == MSGPU_MIXED_GATED && r->kind != MSGPU_REQ_LOGDET)) {
        set_error("msgpu_server_submit: unknown mode (MSGPU_MIXED_GATED is for log-det tiles)");
        return -4;
    }
    const bool ok = (r->kind == MSGPU_REQ_LOGDET && r->nk > 0 && r->ks && r->out) ||
                    (r->kind == MSGPU_REQ_SOLVE && r->a && (r->nx * r->ny == 0 || r->psi)) ||
                    (r->kind == MSGPU_REQ_RESONANCE && r->nz > 0 && r->z && r->roots && r->nroots);
    if (!ok) { set_error("msgpu_server_submit: incomplete request"); return -2; }
    Req* q = new Req();
    q->r = *r;
    q->hpos.assign(mv.pos, mv.pos + (size_t)mv.n * mv.d);
    q->mv = mv;
    q->mv.pos = q->hpos.data();
    if (r->kind == MSGPU_REQ_LOGDET) q->ks.assign(r->ks, r->ks + r->nk);
    if (r->kind == MSGPU_REQ_RESONANCE)
        q->run.reset(new AberthRun(r->z, r->nz, r->dom[0], r->dom[1], r->dom[2], r->dom[3], r->nexpsup, r->maxit,
                                   r->toler, r->mode != MSGPU_FP64));
    q->footprint = footprint_of(*r, mv.n);
    q->t_submit = clk::now();
    std::lock_guard<std::mutex> g(s->mu);
    q->id = s->next_id++;
    s->all[q->id] = q;
    s->pending.push_back(q);
    s->admit_locked();
    s->cv.notify_all();
    return q->id;
}

MSGPU_API int msgpu_server_wait(msgpu_server* sv, long long id, msgpu_result* res) {
    Server* s = reinterpret_cast<Server*>(sv);
    Req* q = nullptr;
    {
        std::unique_lock<std::mutex> g(s->mu);
        auto it = s->all.find(id);
        if (it == s->all.end()) { set_error("msgpu_server_wait: unknown request id"); return -2; }
        q = it->second;
        s->cv.wait(g, [&] { return q->done; });
        s->all.erase(it);
    }
    if (q->status < 0) set_error(q->err);
    if (res) {
        res->status = q->status;
        res->count = q->count;
        res->queue_ms = ms_between(q->t_submit, q->t_admit);
        res->total_ms = ms_between(q->t_submit, q->t_done);
    }
    const int st = q->status;
    delete q;
    return st;
}

MSGPU_API int msgpu_server_stats(msgpu_server* sv, double* stats, int nstats) {
    Server* s = reinterpret_cast<Server*>(sv);
    std::lock_guard<std::mutex> g(s->mu);
    const double v[9] = {(double)s->steps, (double)s->items, (double)s->max_step_reqs, (double)s->peak,
                         (double)s->budget, (double)s->factor_hits, (double)s->factor_misses,
                         (double)s->kept_bytes, (double)s->media.size()};
    for (int i = 0; i < nstats && i < 9; i++) stats[i] = v[i];
    return 0;
}

MSGPU_API void msgpu_server_destroy(msgpu_server* sv) {
    Server* s = reinterpret_cast<Server*>(sv);
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    s->batch_thread.join();
    s->solve_thread.join();
    for (auto& kv : s->all) delete kv.second;   // requests never waited for
    delete s;
}

}  // extern "C"

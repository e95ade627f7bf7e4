// engine.cpp -- host driver of the device path: device buffer pool, stream
// leases for concurrent callers, and the batched assembly / LU / solve / psi
// pipelines built from the kernels of kern_*.hip.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "capi_internal.h"
#include "kernels.h"

namespace msg {

// ------------------------------------------------------------------ errors
static thread_local std::string t_err;
void set_error(const std::string& m) { t_err = m; }
const char* last_error() { return t_err.c_str(); }

struct HipFail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            throw HipFail(std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" +   \
                          std::to_string(__LINE__) + ")");                                           \
    } while (0)

// ------------------------------------------------------------------ device pool
// Buffers are recycled by size class (next power of two above 1 MiB granularity)
// and never returned to the driver, so steady-state calls allocate nothing.  The same
// pool type holds the pinned host buffers of the host-buffer services.
class Pool {
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    const bool pinned_;
    hipError_t raw_alloc(void** p, size_t b) { return pinned_ ? hipHostMalloc(p, b, hipHostMallocDefault) : hipMalloc(p, b); }
    void raw_free(void* p) { (void)(pinned_ ? hipHostFree(p) : hipFree(p)); }

public:
    explicit Pool(bool pinned = false) : pinned_(pinned) {}
    static size_t klass(size_t b) {
        size_t c = (size_t)1 << 16;
        while (c < b) c <<= 1;
        return c;
    }
    void* get(size_t bytes, size_t& cls) {
        cls = klass(bytes == 0 ? 1 : bytes);
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = free_.find(cls);
            if (it != free_.end()) {
                void* p = it->second;
                free_.erase(it);
                return p;
            }
        }
        void* p = nullptr;
        hipError_t e = raw_alloc(&p, cls);
        if (e != hipSuccess) {
            // release cached buffers and retry once
            std::lock_guard<std::mutex> g(mu_);
            for (auto& kv : free_) raw_free(kv.second);
            free_.clear();
            HIPCHK(raw_alloc(&p, cls));
        }
        return p;
    }
    void put(void* p, size_t cls) {
        std::lock_guard<std::mutex> g(mu_);
        free_.emplace(cls, p);
    }
};

static Pool& pool() {
    static Pool* p = new Pool();
    return *p;
}
static Pool& pinned_pool() {
    static Pool* p = new Pool(true);
    return *p;
}

// When a buffer is released while an exception unwinds the stack (a HIP call failed
// after kernels were queued), work still in flight may write it: the device is drained
// before the buffer goes back to the pool that other threads draw from.
static void drain_if_unwinding() {
    if (std::uncaught_exceptions() > 0) (void)hipDeviceSynchronize();
}

template <class T> struct DBuf {
    T* p = nullptr;
    size_t cls = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        if (p) pool().put(p, cls);
        p = (T*)pool().get(sizeof(T) * count, cls);
    }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() {
        if (p) {
            drain_if_unwinding();
            pool().put(p, cls);
        }
    }
};

// A pinned host buffer of count complex128 entries from the pinned pool.
struct HostStage {
    cx<double>* p = nullptr;
    size_t cls = 0;
    explicit HostStage(size_t count) { p = (cx<double>*)pinned_pool().get(sizeof(cx<double>) * count, cls); }
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    ~HostStage() {
        drain_if_unwinding();
        pinned_pool().put(p, cls);
    }
};

// ------------------------------------------------------------------ streams
class StreamPool {
    std::mutex mu_;
    std::vector<hipStream_t> free_;

public:
    hipStream_t get() {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (!free_.empty()) {
                hipStream_t s = free_.back();
                free_.pop_back();
                return s;
            }
        }
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        return s;
    }
    void put(hipStream_t s) {
        std::lock_guard<std::mutex> g(mu_);
        free_.push_back(s);
    }
};
static StreamPool& streams() {
    static StreamPool* p = new StreamPool();
    return *p;
}
struct Lease {
    hipStream_t st;
    Lease() : st(streams().get()) {}
    explicit Lease(bool take) : st(take ? streams().get() : nullptr) {}
    Lease(const Lease&) = delete;
    Lease& operator=(const Lease&) = delete;
    Lease(Lease&& o) noexcept : st(o.st) { o.st = nullptr; }
    ~Lease() { if (st) streams().put(st); }
    void sync() { HIPCHK(hipStreamSynchronize(st)); }
};

static std::once_flag g_once;
static int g_init_rc = -1;

// The tuning table shipped beside the library (tuning_gfx950.txt, written by
// msgpu_autotune + msgpu_tune_save on an MI355X) is loaded at initialisation unless
// MSGPU_TUNE_FILE names another one; MSGPU_TUNE_FILE set to an empty string loads none.
static void load_initial_tuning() {
    if (const char* env = getenv("MSGPU_TUNE_FILE")) {
        if (*env) (void)load_tuning(env);
        return;
    }
    Dl_info info;
    if (!dladdr((void*)&load_initial_tuning, &info) || !info.dli_fname) return;
    std::string path(info.dli_fname);
    const size_t slash = path.rfind('/');
    path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning_gfx950.txt";
    if (FILE* f = fopen(path.c_str(), "r")) {
        fclose(f);
        (void)load_tuning(path.c_str());
    }
}

int engine_init() {
    std::call_once(g_once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
            g_init_rc = -1;
            return;
        }
        if (hipSetDevice(0) != hipSuccess) {
            g_init_rc = -2;
            return;
        }
        g_init_rc = 0;
        load_initial_tuning();
        (void)noalloc_setup();      // a failure there is reported by the allocation-free entry point
    });
    if (g_init_rc != 0) set_error("msgpu: no usable HIP device (hipGetDeviceCount/hipSetDevice failed)");
    return g_init_rc;
}

template <class F>
static int guarded(F&& f) {
    if (engine_init() != 0) return -1;
    try {
        return f();
    } catch (const std::exception& ex) {
        set_error(std::string("msgpu: ") + ex.what());
        return -2;
    }
}

// ------------------------------------------------------------------ kernel timing
// When enabled, every launch of the batched pipelines is bracketed by HIP events on
// its own stream; categories: 0 panel, 1 swap+trsm, 2 gemm, 3 assembly, 4 tracelog,
// 5 psi grid.  flops/bytes are the algorithmic counts of the launch.
struct ProfRec {
    int cat;
    hipEvent_t a, b;
    double flops, bytes;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;

// Launches of the allocation-free entry point are never timed (no events are created).
static thread_local bool t_untimed = false;

template <class F>
static void timed(int cat, double flops, double bytes, hipStream_t st, F&& launch) {
    if (!g_prof_on || t_untimed) { launch(); return; }
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, st));
    launch();
    HIPCHK(hipEventRecord(b, st));
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof.push_back({cat, a, b, flops, bytes});
}

int profile_begin() {
    return guarded([&]() -> int {
        std::lock_guard<std::mutex> g(g_prof_mu);
        for (auto& r : g_prof) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        g_prof.clear();
        g_prof_on = true;
        return 0;
    });
}

int profile_end(double* ms, double* flops, double* bytes, int* count, int ncat) {
    return guarded([&]() -> int {
        std::lock_guard<std::mutex> g(g_prof_mu);
        g_prof_on = false;
        for (int c = 0; c < ncat; c++) { ms[c] = 0; flops[c] = 0; bytes[c] = 0; count[c] = 0; }
        for (auto& r : g_prof) {
            HIPCHK(hipEventSynchronize(r.b));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
            if (r.cat < ncat) { ms[r.cat] += t; flops[r.cat] += r.flops; bytes[r.cat] += r.bytes; count[r.cat]++; }
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        g_prof.clear();
        return 0;
    });
}

MediumView view_of(const msgpu_medium* med) {
    MediumView v;
    v.d = med->d;
    v.n = med->n;
    v.np = padded(med->n);
    v.m = scmodel_of(med);
    v.pos = med->pos;
    return v;
}

static constexpr int NB = 64;    // panel width == triangular block of the solves

// Workspace budget for a batch of matrices (bytes), MSGPU_WORKSPACE_GB overrides.
static size_t batch_budget() {
    static size_t b = [] {
        const char* e = getenv("MSGPU_WORKSPACE_GB");
        double gb = e ? atof(e) : 16.0;
        return (size_t)(gb * (1ull << 30));
    }();
    return b;
}
static int max_batch(int np, size_t elem, int mats_per_item) {
    size_t per = (size_t)np * np * elem * mats_per_item;
    size_t b = batch_budget() / per;
    if (b < 1) b = 1;
    if (b > 2048) b = 2048;
    return (int)b;
}

// ------------------------------------------------------------------ tuning table
// Per (padded size, precision): the swap + TRSM tile shape, the chunk size, the
// stream count, the storage order of the trace-log factorisation (rm: 1 row-major,
// 0 column-major), the panel shape, whether trace-log chunks replay a captured HIP
// graph, and the blocking of the LU (blk2: 128-column blocks, see factor2).  Filled by autotune() or load_tuning(); an entry of 0 keeps the built-in
// choice (variant 0, MSGPU_CHUNK, MSGPU_STREAMS, column-major, default panel, eager
// launches).
struct Tune {
    int swap = 0, chunk = 0, streams = 0, rm = 0, panel = 0, graph = 0, blk2 = 0;
};
static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, Tune> g_tune;

// A size without an entry takes that of the nearest size of the same precision (by
// ratio; the larger on a tie), without graph replay (a capture per size would not pay
// off for sizes met a few times); with no entry of that precision, the built-in shapes.
static Tune tune_for(int np, int prec) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find({np, prec});
    if (it != g_tune.end()) return it->second;
    const Tune* best = nullptr;
    double bd = 0.0;
    for (const auto& kv : g_tune) {
        if (kv.first.second != prec) continue;
        const double d = std::fabs(std::log((double)kv.first.first / np));
        if (!best || d < bd || (d == bd && kv.first.first > np)) {
            best = &kv.second;
            bd = d;
        }
    }
    if (!best) return Tune{};
    Tune t = *best;
    t.graph = 0;
    // panel variants name different shapes in each size range (launch_panel): the panel
    // is that of the nearest entry in np's range, else the range's default
    auto range = [](int n) { return n <= 1024 ? 0 : n <= 2048 ? 1 : n <= 4096 ? 2 : 3; };
    const Tune* bp = nullptr;
    double bpd = 0.0;
    for (const auto& kv : g_tune) {
        if (kv.first.second != prec || range(kv.first.first) != range(np)) continue;
        const double d = std::fabs(std::log((double)kv.first.first / np));
        if (!bp || d < bpd || (d == bpd && kv.first.first > np)) {
            bp = &kv.second;
            bpd = d;
        }
    }
    t.panel = bp ? bp->panel : 0;
    return t;
}
template <class T> static int prec_of() { return sizeof(T) == 8 ? 0 : 1; }

// ------------------------------------------------------------------ LU core
// rm: the matrices are stored row-major.  The trailing update is the same GEMM call in
// both orders: read column-major, the row-major blocks L21, U12 and A22 are U12^T,
// L21^T and A22^T at the addresses of the column-major L21, U12 and A22, and
// A22^T -= U12^T L21^T is the transposed update.  The pipelines of symmetric M(k)
// use row-major storage, where the row interchanges move contiguous rows.
// left: apply the interchanges to the columns left of the panel too (the full
// P^T M = L U that the solves need); the trace-log needs only U's diagonal.
// after_first_panel (optional) is recorded once the first panel is queued, so that a
// concurrent stream can start its own factorization one panel phase later and the
// latency-bound panels of one stream run under the GEMMs of the other.
template <class T>
static void factor2(cx<T>* mats, int np, int nbatch, int* ipiv, int* info, int* pmap, bool left, bool rm,
                    hipStream_t st, hipEvent_t after_first_panel);

template <class T>
static void factor1(cx<T>* mats, int np, int nbatch, int* ipiv, int* info, int* pmap, bool left, bool rm,
                    hipStream_t st, hipEvent_t after_first_panel);

// Blocked LU of a batch (see factor1 / factor2); the tuning table picks the blocking.
template <class T>
static void factor(cx<T>* mats, int np, int nbatch, int* ipiv, int* info, int* pmap, bool left, bool rm,
                   hipStream_t st, hipEvent_t after_first_panel = nullptr) {
    if (tune_for(np, prec_of<T>()).blk2) factor2<T>(mats, np, nbatch, ipiv, info, pmap, left, rm, st, after_first_panel);
    else factor1<T>(mats, np, nbatch, ipiv, info, pmap, left, rm, st, after_first_panel);
}

// The blocked LU of a batch for the allocation-free entry point (noalloc.hip): the
// tuning table's blocking, storage order and launch shapes for np, no timing events.
// Returns the storage order used (1 row-major).
template <class T>
int factor_untimed(cx<T>* mats, int np, int nbatch, int* ipiv, int* info, int* ws, hipStream_t st) {
    const bool rm = tune_for(np, prec_of<T>()).rm == 1;
    t_untimed = true;
    try {
        factor<T>(mats, np, nbatch, ipiv, info, ws, false, rm, st);
    } catch (...) {
        t_untimed = false;
        throw;
    }
    t_untimed = false;
    return rm ? 1 : 0;
}
template int factor_untimed<double>(cx<double>*, int, int, int*, int*, int*, hipStream_t);
template int factor_untimed<float>(cx<float>*, int, int, int*, int*, int*, hipStream_t);

template <class T>
static void factor1(cx<T>* mats, int np, int nbatch, int* ipiv, int* info, int* pmap, bool left, bool rm,
                    hipStream_t st, hipEvent_t after_first_panel) {
    const long long ms = (long long)np * np;
    // the panel kernel holds at most 8192 rows of a panel (1024 threads x 8 rows)
    if (np > 8192) throw std::runtime_error("matrices above N = 8192 are not supported by the device LU");
    launch_fill_int(info, 0, nbatch, st);
    const double es = sizeof(cx<T>);
    const Tune tn = tune_for(np, prec_of<T>());
    const int swapv = tn.swap, panelv = tn.panel;
    for (int k0 = 0; k0 < np; k0 += NB) {
        const int kend = k0 + NB, rest = np - kend;
        timed(0, 8.0 * nbatch * ((double)(np - k0) * NB * NB / 2), es * nbatch * 2.0 * (np - k0) * NB, st,
              [&] { launch_panel<T>(mats, np, ms, ipiv, info, pmap, k0, NB, nbatch, st, rm, panelv); });
        if (k0 == 0 && after_first_panel) HIPCHK(hipEventRecord(after_first_panel, st));
        timed(1, 8.0 * nbatch * (double)NB * NB / 2 * rest, es * nbatch * 2.0 * NB * rest, st,
              [&] { launch_swap_trsm<T>(mats, np, ms, pmap, k0, NB, left ? 1 : 0, nbatch, st, swapv, rm); });
        if (kend < np)
            timed(2, 8.0 * nbatch * (double)rest * rest * NB, es * nbatch * (2.0 * rest * rest + 2.0 * rest * NB), st,
                  [&] {
                      launch_gemm_sub<T>(mats + kend + (size_t)kend * np, mats + kend + (size_t)k0 * np,
                                         mats + k0 + (size_t)kend * np, rest, rest, NB, np, np, np, ms, ms, ms, nbatch,
                                         st);
                  });
    }
    HIPCHK(hipGetLastError());
}

// Two-level blocking: blocks of 128 columns factored as two 64-column panels, with one
// trailing update of K = 128 per block instead of two of K = 64 (half the passes over
// the trailing matrix, whose read and write bound the K = 64 update).  Per block K0
// (second panel at k1 = K0 + 64, block end ke = K0 + 128):
//   panel K0; its interchanges and U12 on [k1, ke) (and on [0, K0) with left);
//   A(k1:, k1:ke) -= L(k1:, K0:k1) U(K0:k1, k1:ke);  panel k1;
//   its interchanges on [K0, k1) (on [0, k1) with left);
//   trailing columns [ke, np): interchanges of K0 and of k1 (one composed row map;
//   those of k1 move rows >= k1 only and so commute with U12 of K0), U12 of K0, then
//   U12 of k1 = L11(k1)^-1 (A(k1:ke, ke:) - L(k1:ke, K0:k1) U(K0:k1, ke:)) as one update
//   of K = 128 with the 64 x 128 block Wb = [L11^-1 L21, I - L11^-1] (block_w_kernel);
//   A(ke:, ke:) -= L(ke:, K0:ke) U(K0:ke, ke:)   (K = 128).
// Row-major storage applies each update to the transposed block (see gemm below).
template <class T>
static void factor2(cx<T>* mats, int np, int nbatch, int* ipiv, int* info, int* pmap, bool left, bool rm,
                    hipStream_t st, hipEvent_t after_first_panel) {
    const long long ms = (long long)np * np;
    if (np > 8192) throw std::runtime_error("matrices above N = 8192 are not supported by the device LU");
    launch_fill_int(info, 0, nbatch, st);
    const double es = sizeof(cx<T>);
    const Tune tn = tune_for(np, prec_of<T>());
    // the update blocks Wb follow the row maps in the factorisation workspace
    cx<T>* wblk = reinterpret_cast<cx<T>*>(pmap + wblock_offset(np, NB, nbatch));
    // M(r0:r0+m, c0:c0+n) -= M(r0:, p0:p0+k) M(p0:, c0:).  Row-major storage read column-major
    // is S = M^T, where this is S(c0:, r0:) -= S(c0:, p0:) S(p0:, r0:): the transposed block.
    auto gemm = [&](int r0, int c0, int p0, int m, int n, int k) {
        if (m <= 0 || n <= 0) return;
        const int R = rm ? c0 : r0, Cc = rm ? r0 : c0;
        timed(2, 8.0 * nbatch * (double)m * n * k, es * nbatch * (2.0 * m * n + (double)(m + n) * k), st, [&] {
            launch_gemm_sub<T>(mats + R + (size_t)Cc * np, mats + R + (size_t)p0 * np, mats + p0 + (size_t)Cc * np,
                               rm ? n : m, rm ? m : n, k, np, np, np, ms, ms, ms, nbatch, st);
        });
    };
    auto panel = [&](int k0, int xa = 0, int xb = 0) {
        timed(0, 8.0 * nbatch * ((double)(np - k0) * NB * NB / 2), es * nbatch * 2.0 * (np - k0 + (xb - xa)) * NB, st,
              [&] { launch_panel<T>(mats, np, ms, ipiv, info, pmap, k0, NB, nbatch, st, rm, tn.panel, xa, xb); });
    };
    auto swap = [&](int k0, int lb, int rb, int re, int do_swap, int do_trsm, int k0b = -1) {
        const double cols = (double)std::max(0, re - rb) + std::max(0, k0 - lb);
        timed(1, do_trsm ? 8.0 * nbatch * (double)NB * NB / 2 * std::max(0, re - rb) : 0.0,
              es * nbatch * 2.0 * NB * cols, st, [&] {
                  launch_swap_range<T>(mats, np, ms, pmap, k0, NB, lb, rb, re, do_swap, do_trsm, nbatch, st, tn.swap,
                                       rm, k0b);
              });
    };
    for (int K0 = 0; K0 < np; K0 += 2 * NB) {
        const int k1 = K0 + NB, ke = K0 + 2 * NB;
        panel(K0);
        if (K0 == 0 && after_first_panel) HIPCHK(hipEventRecord(after_first_panel, st));
        if (k1 >= np) {                               // a last, single panel
            swap(K0, left ? 0 : K0, np, np, 1, 0);
            break;
        }
        swap(K0, left ? 0 : K0, k1, ke, 1, 1);
        gemm(k1, k1, K0, np - k1, NB, NB);
        if (left) {
            panel(k1);
            swap(k1, 0, 0, 0, 1, 0);
        } else {
            panel(k1, K0, k1);                    // its interchanges on [K0, k1) in the same launch
        }
        if (ke < np) {
            swap(K0, K0, ke, np, 1, 1, k1);       // interchanges of K0 and k1 composed, U12 of K0
            // U12 of k1 = L11^-1 (A - L21 U12(K0)) as one K = 128 update with Wb
            launch_block_w<T>(mats, np, ms, K0, wblk, nbatch, st, rm);
            // algorithmic flops: L21 U12(K0) (NB per column and row) and the unit-lower
            // solve (NB / 2); the other NB / 2 multiply the zero upper part of I - L11^-1
            timed(2, 8.0 * nbatch * (double)NB * (np - ke) * 1.5 * NB,
                  es * nbatch * (2.0 * NB * (np - ke) + (NB + np - ke) * 2.0 * NB), st, [&] {
                if (rm)
                    launch_gemm_sub<T>(mats + ke + (size_t)k1 * np, mats + ke + (size_t)K0 * np, wblk, np - ke, NB,
                                       2 * NB, np, 2 * NB, np, ms, ms, 64 * 128, nbatch, st);
                else
                    launch_gemm_sub<T>(mats + k1 + (size_t)ke * np, wblk, mats + K0 + (size_t)ke * np, NB, np - ke,
                                       2 * NB, NB, np, np, ms, 64 * 128, ms, nbatch, st);
            });
            gemm(ke, ke, K0, np - ke, np - ke, 2 * NB);
        }
    }
    HIPCHK(hipGetLastError());
}

// X <- M^-1 X for X (np x ncol, ld np) given the factor of factor() without left swaps.
template <class T>
static void lu_solve(const cx<T>* lu, cx<T>* X, int np, int ncol, int nbatch, const int* pmap, long long sx,
                     hipStream_t st) {
    const long long ms = (long long)np * np;
    const int nq = np / NB;
    for (int q = 0; q < nq; q++) {
        const int k0 = q * NB;
        launch_apply_map<T>(X, np, ncol, pmap, NB, q, sx, nbatch, st);
        launch_trsm_block<T>(lu, X, np, ncol, k0, 0, ms, sx, nbatch, st);
        if (k0 + NB < np)
            launch_gemm_sub<T>(X + k0 + NB, lu + (k0 + NB) + (size_t)k0 * np, X + k0, np - k0 - NB, ncol, NB, np, np,
                               np, sx, ms, sx, nbatch, st);
    }
    for (int q = nq - 1; q >= 0; q--) {
        const int k0 = q * NB;
        launch_trsm_block<T>(lu, X, np, ncol, k0, 1, ms, sx, nbatch, st);
        if (k0 > 0)
            launch_gemm_sub<T>(X, lu + (size_t)k0 * np, X + k0, k0, ncol, NB, np, np, np, sx, ms, sx, nbatch, st);
    }
    HIPCHK(hipGetLastError());
}

// X <- V with M V' = X solved from the row-major factor (left interchanges applied)
// of a symmetric M, read column-major as S = M^T = U^T L^T P^T: first U^T W = X
// (non-unit lower), then L^T V = W (unit upper).  The solution of M Y = X is then
// Y = P V, i.e. Y[perm[i]] = V[i] with perm from launch_build_perm.
// Blocks of 128 rows (two 64-row triangular blocks and the 64-row update between them),
// so that the updates of the rest of X have K = 128 (one pass over X per 128 rows).
// identity_rhs: X is the identity (np columns), whose column c is zero above row c: the
// forward solve then leaves every column c >= the end of the current block untouched,
// so each forward step runs on the columns [0, block end) only (a third of the forward
// work; the inverse then costs 16 N^3 / 3 real flops instead of 8 N^3).
template <class T>
static void lu_solve_rm(const cx<T>* lu, cx<T>* X, int np, int ncol, int nbatch, long long sx, hipStream_t st,
                        bool identity_rhs = false) {
    const long long ms = (long long)np * np;
    // X[r0:r0+m, 0:n] -= S[r0:r0+m, c0:c0+k] X[c0:c0+k, 0:n]
    auto upd = [&](int r0, int c0, int m, int k, int n) {
        if (m > 0 && k > 0 && n > 0)
            launch_gemm_sub<T>(X + r0, lu + r0 + (size_t)c0 * np, X + c0, m, n, k, np, np, np, sx, ms, sx, nbatch,
                               st);
    };
    auto fwd_cols = [&](int end) { return identity_rhs ? std::min(ncol, end) : ncol; };
    for (int K0 = 0; K0 < np; K0 += 2 * NB) {
        const int k1 = K0 + NB, ke = std::min(K0 + 2 * NB, np);
        launch_trsm_block<T>(lu, X, np, fwd_cols(k1), K0, 0, ms, sx, nbatch, st, true);
        if (k1 < np) {
            upd(k1, K0, ke - k1, NB, fwd_cols(k1));
            launch_trsm_block<T>(lu, X, np, fwd_cols(ke), k1, 0, ms, sx, nbatch, st, true);
        }
        upd(ke, K0, np - ke, ke - K0, fwd_cols(ke));
    }
    for (int K0 = ((np - 1) / (2 * NB)) * (2 * NB); K0 >= 0; K0 -= 2 * NB) {
        const int k1 = K0 + NB, ke = std::min(K0 + 2 * NB, np);
        if (k1 < np) {
            launch_trsm_block<T>(lu, X, np, ncol, k1, 1, ms, sx, nbatch, st, true);
            upd(K0, k1, NB, ke - k1, ncol);
        }
        launch_trsm_block<T>(lu, X, np, ncol, K0, 1, ms, sx, nbatch, st, true);
        upd(0, K0, K0, ke - K0, ncol);
    }
    HIPCHK(hipGetLastError());
}

// ------------------------------------------------------------------ batched trace-log
// Chunking of a batch of nk items over NSTREAM streams: chunks of at most bmax items
// (rounded to a multiple of 8 so the GEMM's XCD mapping applies), at least two chunks
// when there is enough work, so that one chunk's latency-bound panel phases overlap
// the other chunk's GEMMs.
// NSTREAM streams (MSGPU_STREAMS overrides; the box exposes 4 hardware queues) and
// chunks of at most MSGPU_CHUNK matrices.
static constexpr int NSTREAM_MAX = 8;
static int default_streams() {
    static int v = [] {
        const char* e = getenv("MSGPU_STREAMS");
        int x = e ? atoi(e) : 4;
        return std::max(1, std::min(x, NSTREAM_MAX));
    }();
    return v;
}
static int streams_for(int np, int prec) {
    const int t = tune_for(np, prec).streams;
    return t > 0 ? std::min(t, NSTREAM_MAX) : default_streams();
}
static int chunk_cap() {
    static int v = [] {
        const char* e = getenv("MSGPU_CHUNK");
        int x = e ? atoi(e) : 256;
        return std::max(8, std::min(x, 4096));
    }();
    return v;
}
static int chunk_size(int nk, int bmax, int np, int prec, int ns) {
    const int tuned = tune_for(np, prec).chunk;
    bmax = std::min(bmax, tuned > 0 ? tuned : chunk_cap());
    int nch = (nk + bmax - 1) / bmax;
    if (nch < ns && nk >= 32 * ns) nch = ns;
    else if (nch < 2 && nk >= 64) nch = 2;
    int c = (nk + nch - 1) / nch;
    if (c >= 8) c = (c + 7) / 8 * 8;
    return std::max(1, std::min(c, bmax));
}

struct Lanes {   // ns leased streams; lane 0 owns the inputs, the others wait for them
    int ns;
    std::vector<Lease> l;
    hipEvent_t ev[NSTREAM_MAX];
    explicit Lanes(int n) : ns(std::max(1, std::min(n, NSTREAM_MAX))) {
        l.reserve(ns);
        for (int i = 0; i < ns; i++) l.emplace_back(true);
        for (int i = 0; i < ns; i++) HIPCHK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
    ~Lanes() { for (int i = 0; i < ns; i++) (void)hipEventDestroy(ev[i]); }
    hipStream_t operator[](int i) const { return l[i].st; }
    void fork() {
        HIPCHK(hipEventRecord(ev[0], l[0].st));
        for (int i = 1; i < ns; i++) HIPCHK(hipStreamWaitEvent(l[i].st, ev[0], 0));
    }
    void join() {
        for (int i = 1; i < ns; i++) {
            HIPCHK(hipEventRecord(ev[i], l[i].st));
            HIPCHK(hipStreamWaitEvent(l[0].st, ev[i], 0));
        }
    }
};

// ------------------------------------------------------------------ multi-medium batches
// A batch is a list of (medium, k) items.  Items are grouped by padded size np
// (largest first); each group is cut into chunks (chunk_size) that run round-robin
// on the leased streams.  Every matrix of a chunk has the same np but may belong to
// another medium: the assembly, the diagonal and the final reductions read the
// item's own positions, N, d and scattering model, the LU kernels only see np.
struct Plan {
    struct Chunk { int np, start, count; };
    std::vector<int> order;            // position in the device item list -> input index
    std::vector<Chunk> chunks;
    size_t mat_elems = 0, piv_ints = 0, map_ints = 0;
    int max_b = 0;
};

static int max_np(const std::vector<MediumView>& meds, int nk, const int* midx) {
    int m = 0;
    for (int i = 0; i < nk; i++) m = std::max(m, meds[midx ? midx[i] : 0].np);
    return m;
}

static Plan make_plan(const std::vector<MediumView>& meds, int nk, const int* midx, size_t elem, int mats_per_item,
                      int ns) {
    Plan P;
    P.order.resize(nk);
    for (int i = 0; i < nk; i++) P.order[i] = i;
    auto np_of = [&](int i) { return meds[midx ? midx[i] : 0].np; };
    std::stable_sort(P.order.begin(), P.order.end(), [&](int a, int b) { return np_of(a) > np_of(b); });
    // Groups of fewer than kFill matrices leave most CUs idle in the latency-bound panel
    // phases: such a group takes in the next smaller sizes down to 3/4 of its own, padded
    // to its size.  The identity padding changes no result bit: padded rows never win a
    // pivot, every product with a padded row or column is an exact zero, and the blocks
    // over the real rows stay the same, provided that
    //  - the storage order and the 64/128-column blocking (tuning table) are the same for
    //    both sizes, and
    //  - for the N-column solves of tr(M^-1 M') in complex128, the smaller size is a
    //    multiple of 128 (the solves' 128-row blocks then end at its last row; otherwise
    //    its last 64 rows would join a 128-row block, and complex128 updates of K = 128
    //    round differently from K = 64 ones).
    constexpr int kFill = 256;
    const int prec = elem == 16 ? 0 : 1;
    auto numerics = [&](int np) { const Tune t = tune_for(np, prec); return std::make_pair(t.rm, t.blk2); };
    auto may_pad = [&](int np2, int np) {
        return 4 * np2 >= 3 * np && numerics(np2) == numerics(np) &&
               (mats_per_item == 1 || prec == 1 || np2 % 128 == 0);
    };
    for (int g0 = 0; g0 < nk;) {
        const int np = np_of(P.order[g0]);
        int g1 = g0;
        while (g1 < nk && np_of(P.order[g1]) == np) g1++;
        while (g1 < nk && g1 - g0 < kFill) {
            const int np2 = np_of(P.order[g1]);
            if (!may_pad(np2, np)) break;
            while (g1 < nk && np_of(P.order[g1]) == np2) g1++;
        }
        const int cnt = g1 - g0;
        const int bsz = chunk_size(cnt, max_batch(np, elem, mats_per_item * ns), np, elem == 16 ? 0 : 1, ns);
        for (int s0 = g0; s0 < g1; s0 += bsz) {
            const int B = std::min(bsz, g1 - s0);
            P.chunks.push_back({np, s0, B});
            P.mat_elems = std::max(P.mat_elems, (size_t)B * np * np);
            P.piv_ints = std::max(P.piv_ints, (size_t)B * np);
            P.map_ints = std::max(P.map_ints, factor_ws_ints(np, NB, B));
            P.max_b = std::max(P.max_b, B);
        }
        g0 = g1;
    }
    return P;
}

// Positions of every medium in one device buffer, and the device item list in plan order.
struct DeviceItems {
    DBuf<double> pos;
    DBuf<MatItem> items;
    std::vector<double> hpos;
    std::vector<MatItem> hitems;
    DeviceItems(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                const std::vector<int>& order, hipStream_t st) {
        std::vector<size_t> off(meds.size());
        size_t tot = 0;      // media without a resident device copy are uploaded together
        for (size_t m = 0; m < meds.size(); m++) {
            off[m] = tot;
            if (!meds[m].dpos) tot += (size_t)meds[m].n * meds[m].d;
        }
        hpos.resize(std::max<size_t>(tot, 1));
        for (size_t m = 0; m < meds.size(); m++)
            if (!meds[m].dpos)
                std::copy(meds[m].pos, meds[m].pos + (size_t)meds[m].n * meds[m].d, hpos.begin() + off[m]);
        pos.alloc(hpos.size());
        if (tot) HIPCHK(hipMemcpyAsync(pos.p, hpos.data(), sizeof(double) * tot, hipMemcpyHostToDevice, st));
        hitems.resize(nk);
        for (int q = 0; q < nk; q++) {
            const int i = order[q], m = midx ? midx[i] : 0;
            MatItem& it = hitems[q];
            it.pos = meds[m].dpos ? meds[m].dpos : pos.p + off[m];
            it.n = meds[m].n;
            it.d = meds[m].d;
            it.m = meds[m].m;
            it.k = ks[i];
            it.diag0 = it.diag1 = cx<double>{0.0, 0.0};
        }
        items.alloc(nk);
        HIPCHK(hipMemcpyAsync(items.p, hitems.data(), sizeof(MatItem) * nk, hipMemcpyHostToDevice, st));
        launch_item_diag(items.p, nk, st);
    }
    // the dimension d shared by the items [s, s + B) in plan order, 0 when they differ
    int dim(int s, int B) const {
        const int d = hitems[s].d;
        for (int q = s + 1; q < s + B; q++)
            if (hitems[q].d != d) return 0;
        return d;
    }
};

// ------------------------------------------------------------------ HIP graphs
// A trace-log chunk (assembly, factorisation, reduction: 3 + 3 Np/64 launches) can be
// captured once per (stream, shape) and replayed.  The graph owns its workspace: the
// chunk's items are copied into it before each replay and the results copied out
// after, so the captured launches always see the same addresses.  Workspaces are
// kept for the process lifetime, within a total of kGraphBytes; beyond that, chunks
// are launched eagerly.  The only memory operations inside a graph are kernels.
template <class T>
struct ChunkGraph {
    DBuf<MatItem> items;
    DBuf<cx<T>> mats;
    DBuf<int> ipiv, info, pmap;
    DBuf<cx<double>> out;
    DBuf<double> stats;
    hipGraphExec_t exec = nullptr;
};
static constexpr size_t kGraphBytes = (size_t)8 << 30;
static std::mutex g_graph_mu;
static size_t g_graph_bytes = 0;

template <class T>
static ChunkGraph<T>* chunk_graph(hipStream_t st, int np, int B, bool rm, const Tune& tn, int dall) {
    using Key = std::tuple<hipStream_t, int, int, int, int, int, int, int>;
    static std::map<Key, std::unique_ptr<ChunkGraph<T>>> cache;
    const Key key{st, np, B, rm ? 1 : 0, tn.swap, tn.panel, tn.blk2, dall == 3 ? 3 : 0};
    std::lock_guard<std::mutex> g(g_graph_mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second.get();
    const long long ms = (long long)np * np;
    const size_t bytes = (size_t)B * ms * sizeof(cx<T>);
    if (np > 8192 || g_graph_bytes + bytes > kGraphBytes) return nullptr;
    std::unique_ptr<ChunkGraph<T>> cg(new ChunkGraph<T>());
    cg->items.alloc(B);
    cg->mats.alloc((size_t)B * ms);
    cg->ipiv.alloc((size_t)B * np);
    cg->info.alloc(B);
    cg->pmap.alloc(factor_ws_ints(np, NB, B));
    cg->out.alloc(B);
    cg->stats.alloc(2 * (size_t)B);
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    launch_assemble_items<T>(cg->items.p, np, 0, cg->mats.p, ms, B, st, dall);
    factor<T>(cg->mats.p, np, B, cg->ipiv.p, cg->info.p, cg->pmap.p, false, rm, st);
    launch_tracelog<T>(cg->mats.p, 0, cg->items.p, np, ms, cg->ipiv.p, cg->out.p, cg->stats.p, B, st);
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamEndCapture(st, &graph));
    const hipError_t e = hipGraphInstantiate(&cg->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    g_graph_bytes += bytes;
    ChunkGraph<T>* raw = cg.get();
    cache.emplace(key, std::move(cg));
    return raw;
}

// Pivot-free LDL^T of nbatch symmetric matrices (kern_ldlt.hip) in 128-column blocks:
// the block's first 64 columns (diagonal block, rows), their update of its second 64
// columns, the second 64 columns, then one K = 128 update of the lower tiles right of
// the block (the complex128 GEMM's 3M product).  growth[b] receives the largest |L| of
// matrix b (as the bits of a non-negative double), rho[b] its error estimate
// sum_r sum_{k<r} |L_rk|^2 |D_k| / |D_r| (acc: np doubles per matrix of workspace).
static void factor_ldlt(cx<double>* mats, int np, int nbatch, unsigned long long* growth, double* acc, double* rho,
                        hipStream_t st) {
    const long long ms = (long long)np * np;
    auto at = [&](int r, int c) { return mats + r + (size_t)c * np; };
    auto panel = [&](int k0) {
        timed(0, 0.0, 0.0, st, [&] { launch_ldlt_diag<double>(mats, np, ms, k0, growth, acc, rho, nbatch, st); });
        if (np - k0 - NB > 0)
            timed(1, 0.0, 0.0, st, [&] { launch_ldlt_rows<double>(mats, np, ms, k0, growth, acc, nbatch, st); });
    };
    HIPCHK(hipMemsetAsync(growth, 0, sizeof(unsigned long long) * nbatch, st));
    HIPCHK(hipMemsetAsync(acc, 0, sizeof(double) * nbatch * np, st));
    HIPCHK(hipMemsetAsync(rho, 0, sizeof(double) * nbatch, st));
    for (int k0 = 0; k0 < np; k0 += 2 * NB) {
        panel(k0);
        const int W1 = np - k0 - NB;
        if (W1 <= 0) break;
        // algorithmic bytes: C read and written once, A and B read once
        const double es = sizeof(cx<double>);
        timed(2, 8.0 * W1 * NB * NB * nbatch, es * nbatch * (2.0 * W1 * NB + (double)(W1 + NB) * NB), st, [&] {
            launch_gemm_sub<double>(at(k0 + NB, k0 + NB), at(k0 + NB, k0), at(k0, k0 + NB), W1, NB, NB, np, np, np, ms,
                                    ms, ms, nbatch, st);
        });
        panel(k0 + NB);
        const int W2 = np - k0 - 2 * NB;
        if (W2 <= 0) break;
        timed(2, 8.0 * W2 * (W2 + (double)NB) / 2.0 * (2 * NB) * nbatch,
              es * nbatch * ((double)W2 * (W2 + NB) + 2.0 * W2 * (2 * NB)), st, [&] {
            launch_gemm_sub_lower<double>(at(k0 + 2 * NB, k0 + 2 * NB), at(k0 + 2 * NB, k0), at(k0, k0 + 2 * NB), W2,
                                          2 * NB, np, np, np, ms, ms, ms, nbatch, st);
        });
    }
}

// The LDL^T trace-log of every item in one persistent launch (kern_persist.hip): the
// items in order of decreasing size (the largest start first), as many workgroups as stay
// resident on the GPU (within the workspace budget, one Np x Np matrix each).  Used for
// batches whose largest padded size is at most MSGPU_LDLT_PERSIST_NP (default 0: never).
// Its results equal the per-phase launches' bit for bit, but it measured slower on the
// MI355X (d = 3 maps: 0.61x at N = 200, 0.73x at N = 500, 0.80x at N = 1000): one
// workgroup carries an item's latency-bound phases and its GEMM tiles one after the
// other, where the per-phase launches spread each phase over all CUs.
static int ldlt_persist_np() {
    static const int v = [] {
        const char* e = getenv("MSGPU_LDLT_PERSIST_NP");
        return e ? atoi(e) : 0;
    }();
    return v;
}

static void tracelog_ldlt_persist(const std::vector<MediumView>& meds, int nk, const int* midx,
                                  const cx<double>* ks, cx<double>* out, double* stats) {
    Lease L;
    std::vector<int> order(nk);
    for (int i = 0; i < nk; i++) order[i] = i;
    auto np_of = [&](int i) { return meds[midx ? midx[i] : 0].np; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return np_of(a) > np_of(b); });
    DeviceItems D(meds, nk, midx, ks, order, L.st);
    const int npmax = np_of(order[0]);
    const bool d3 = D.dim(0, nk) == 3;
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const long long ms = (long long)npmax * npmax;
    const int per_cu = std::max(1, persist_ldlt_blocks_per_cu(d3));
    const int nwg = std::max(1, std::min({nk, ncu * per_cu,
                                          (int)std::min<size_t>(1 << 20, batch_budget() / (ms * sizeof(cx<double>)))}));
    DBuf<cx<double>> ws((size_t)nwg * ms), dout(nk);
    DBuf<double> accws((size_t)nwg * npmax), drho(nk);
    DBuf<unsigned long long> dgrow(nk);
    DBuf<int> counter(1);
    HIPCHK(hipMemsetAsync(counter.p, 0, sizeof(int), L.st));
    double flops = 0.0;        // algorithmic: 4 N^3 / 3 per complex LDL^T
    for (int i = 0; i < nk; i++) {
        const double n = meds[midx ? midx[i] : 0].n;
        flops += 4.0 * n * n * n / 3.0;
    }
    timed(6, flops, 0.0, L.st,
          [&] {
              launch_ldlt_persist(D.items.p, nk, d3, ws.p, ms, accws.p, npmax, nwg, counter.p, dout.p, dgrow.p, drho.p,
                                  L.st);
          });
    HIPCHK(hipGetLastError());
    std::vector<cx<double>> hout(nk);
    std::vector<unsigned long long> hgrow(nk);
    std::vector<double> hrho(nk);
    HIPCHK(hipMemcpyAsync(hout.data(), dout.p, sizeof(cx<double>) * nk, hipMemcpyDeviceToHost, L.st));
    HIPCHK(hipMemcpyAsync(hgrow.data(), dgrow.p, sizeof(unsigned long long) * nk, hipMemcpyDeviceToHost, L.st));
    HIPCHK(hipMemcpyAsync(hrho.data(), drho.p, sizeof(double) * nk, hipMemcpyDeviceToHost, L.st));
    L.sync();
    for (int q = 0; q < nk; q++) {
        out[order[q]] = hout[q];
        if (stats) {
            memcpy(&stats[2 * order[q]], &hgrow[q], sizeof(double));
            stats[2 * order[q] + 1] = hrho[q];
        }
    }
}

// ldlt (complex128 only): the pivot-free LDL^T instead of the LU; stats[2 l] receives the
// growth max |L| of cell l, stats[2 l + 1] its error estimate
// sum_r sum_{k<r} |L_rk|^2 |D_k| / |D_r| (eps times it estimates |d log det|).
template <class T>
static void tracelog_impl(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                          cx<double>* out, double* stats, bool ldlt = false) {
    if constexpr (std::is_same<T, double>::value) {
        if (ldlt && max_np(meds, nk, midx) <= ldlt_persist_np()) {
            tracelog_ldlt_persist(meds, nk, midx, ks, out, stats);
            return;
        }
    }
    Lanes S(streams_for(max_np(meds, nk, midx), prec_of<T>()));
    const Plan P = make_plan(meds, nk, midx, sizeof(cx<T>), 1, S.ns);
    DeviceItems D(meds, nk, midx, ks, P.order, S[0]);
    DBuf<cx<double>> dout(nk);
    DBuf<double> dstats(2 * (size_t)nk);
    DBuf<unsigned long long> dgrow(ldlt ? nk : 1);
    DBuf<double> drho(ldlt ? nk : 1);
    size_t acc_elems = 1;      // the LDL^T's row sums: np doubles per matrix of a chunk
    for (const Plan::Chunk& ch : P.chunks) acc_elems = std::max(acc_elems, (size_t)ch.count * ch.np);
    S.fork();
    const int nlane = std::min<int>(S.ns, (int)P.chunks.size());
    std::vector<DBuf<cx<T>>> mats(nlane);
    std::vector<DBuf<int>> ipiv(nlane), info(nlane), pmap(nlane);
    std::vector<DBuf<double>> acc(ldlt ? nlane : 0);
    for (int i = 0; i < nlane; i++) {
        if (ldlt) acc[i].alloc(acc_elems);
        mats[i].alloc(P.mat_elems);
        ipiv[i].alloc(P.piv_ints);
        info[i].alloc(P.max_b);
        pmap[i].alloc(P.map_ints);
    }
    // stagger: the first chunk of lane i starts after lane i-1 has finished its first panel
    std::vector<hipEvent_t> stag(nlane, nullptr);
    for (int i = 0; i + 1 < nlane; i++) HIPCHK(hipEventCreateWithFlags(&stag[i], hipEventDisableTiming));
    for (size_t c = 0; c < P.chunks.size(); c++) {
        const Plan::Chunk& ch = P.chunks[c];
        const int B = ch.count, np = ch.np, ln = (int)(c % nlane);
        const long long ms = (long long)np * np;
        hipStream_t st = S[ln];
        if (c > 0 && (int)c < nlane) HIPCHK(hipStreamWaitEvent(st, stag[c - 1], 0));
        const Tune tn = tune_for(np, prec_of<T>());
        if constexpr (std::is_same<T, double>::value) {
            if (ldlt) {
                if ((int)c + 1 < nlane) HIPCHK(hipEventRecord(stag[c], st));
                // the lower triangle and the 32 x 32 diagonal tiles: np (np + 32) / 2 elements
                timed(3, 0.0, (double)B * 0.5 * np * (np + 32) * sizeof(cx<T>), st, [&] {
                    launch_assemble_items<T>(D.items.p + ch.start, np, 0, mats[ln].p, ms, B, st, D.dim(ch.start, B),
                                             true);
                });
                factor_ldlt(mats[ln].p, np, B, dgrow.p + ch.start, acc[ln].p, drho.p + ch.start, st);
                timed(4, 0.0, (double)B * np * sizeof(cx<T>), st, [&] {
                    launch_tracelog<T>(mats[ln].p, 0, D.items.p + ch.start, np, ms, nullptr, dout.p + ch.start,
                                       nullptr, B, st);
                });
                continue;
            }
        }
        const int dall = D.dim(ch.start, B);
        ChunkGraph<T>* cg = (tn.graph && !g_prof_on) ? chunk_graph<T>(st, np, B, tn.rm == 1, tn, dall) : nullptr;
        if (cg) {
            HIPCHK(hipMemcpyAsync(cg->items.p, D.items.p + ch.start, sizeof(MatItem) * B, hipMemcpyDeviceToDevice, st));
            if ((int)c + 1 < nlane) HIPCHK(hipEventRecord(stag[c], st));
            HIPCHK(hipGraphLaunch(cg->exec, st));
            HIPCHK(hipMemcpyAsync(dout.p + ch.start, cg->out.p, sizeof(cx<double>) * B, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(dstats.p + 2 * (size_t)ch.start, cg->stats.p, sizeof(double) * 2 * B,
                                  hipMemcpyDeviceToDevice, st));
            continue;
        }
        timed(3, 0.0, (double)B * ms * sizeof(cx<T>), st,
              [&] { launch_assemble_items<T>(D.items.p + ch.start, np, 0, mats[ln].p, ms, B, st, dall); });
        factor<T>(mats[ln].p, np, B, ipiv[ln].p, info[ln].p, pmap[ln].p, false, tn.rm == 1, st,
                  ((int)c + 1 < nlane) ? stag[c] : nullptr);
        timed(4, 0.0, (double)B * np * sizeof(cx<T>), st, [&] {
            launch_tracelog<T>(mats[ln].p, 0, D.items.p + ch.start, np, ms, ipiv[ln].p, dout.p + ch.start,
                               dstats.p + 2 * (size_t)ch.start, B, st);
        });
    }
    S.join();
    std::vector<cx<double>> hout(nk);
    std::vector<double> hst(stats ? 2 * (size_t)nk : 0);
    std::vector<unsigned long long> hgrow(ldlt && stats ? nk : 0);
    std::vector<double> hrho(ldlt && stats ? nk : 0);
    HIPCHK(hipMemcpyAsync(hout.data(), dout.p, sizeof(cx<double>) * nk, hipMemcpyDeviceToHost, S[0]));
    if (stats && !ldlt)
        HIPCHK(hipMemcpyAsync(hst.data(), dstats.p, sizeof(double) * 2 * nk, hipMemcpyDeviceToHost, S[0]));
    if (stats && ldlt) {
        HIPCHK(hipMemcpyAsync(hgrow.data(), dgrow.p, sizeof(unsigned long long) * nk, hipMemcpyDeviceToHost, S[0]));
        HIPCHK(hipMemcpyAsync(hrho.data(), drho.p, sizeof(double) * nk, hipMemcpyDeviceToHost, S[0]));
    }
    S.l[0].sync();
    for (size_t q = 0; q < hgrow.size(); q++) {
        double g;
        memcpy(&g, &hgrow[q], sizeof g);
        hst[2 * q] = g;
        hst[2 * q + 1] = hrho[q];
    }
    for (hipEvent_t e : stag)
        if (e) (void)hipEventDestroy(e);
    for (int q = 0; q < nk; q++) {
        out[P.order[q]] = hout[q];
        if (stats) { stats[2 * P.order[q]] = hst[2 * q]; stats[2 * P.order[q] + 1] = hst[2 * q + 1]; }
    }
}

static int check_media(const std::vector<MediumView>& meds, int nk, const int* midx) {
    if (meds.empty()) { set_error("msgpu: no medium given"); return -2; }
    for (const MediumView& mv : meds)
        if (mv.d < 1 || mv.d > 4 || mv.n < 1 || !mv.pos) {
            set_error("msgpu: the device path supports d = 1..4 and N >= 1 with positions filled");
            return -3;
        }
    if (midx)
        for (int i = 0; i < nk; i++)
            if (midx[i] < 0 || midx[i] >= (int)meds.size()) { set_error("msgpu: medium index out of range"); return -2; }
    return 0;
}

int tracelog_multi(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                   cx<double>* out, double* stats, int precision) {
    if (nk <= 0) return 0;
    if (int rc = check_media(meds, nk, midx)) return rc;
    return guarded([&]() -> int {
        if (precision == 1) tracelog_impl<float>(meds, nk, midx, ks, out, stats);
        else tracelog_impl<double>(meds, nk, midx, ks, out, stats, precision == 2);
        return 0;
    });
}

int tracelog_points(const MediumView& mv, int nk, const cx<double>* ks, cx<double>* out, double* stats,
                    int precision) {
    return tracelog_multi(std::vector<MediumView>{mv}, nk, nullptr, ks, out, stats, precision);
}

// ------------------------------------------------------------------ batched d/dk log det
template <class T>
static void dk_impl(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                    cx<double>* out, bool trace_inv = false) {
    Lanes S(streams_for(max_np(meds, nk, midx), prec_of<T>()));
    const Plan P = make_plan(meds, nk, midx, sizeof(cx<T>), 2, S.ns);
    DeviceItems D(meds, nk, midx, ks, P.order, S[0]);
    DBuf<cx<double>> dout(nk);
    S.fork();
    const int nlane = std::min<int>(S.ns, (int)P.chunks.size());
    std::vector<DBuf<cx<T>>> mats(nlane), rhs(nlane);
    std::vector<DBuf<int>> ipiv(nlane), info(nlane), pmap(nlane), perm(nlane);
    std::vector<DBuf<cx<double>>> partial(nlane);
    std::vector<DBuf<int>> invp(nlane);
    size_t nparts = 1;          // partial sums of the contraction per chunk
    for (const Plan::Chunk& ch : P.chunks) nparts = std::max(nparts, (size_t)ch.count * dk_contract_partials(ch.np));
    for (int i = 0; i < nlane; i++) {
        mats[i].alloc(P.mat_elems);
        rhs[i].alloc(P.mat_elems);
        ipiv[i].alloc(P.piv_ints);
        perm[i].alloc(P.piv_ints);
        info[i].alloc(P.max_b);
        pmap[i].alloc(P.map_ints);
        partial[i].alloc(nparts);
        invp[i].alloc(P.piv_ints);
    }
    for (size_t c = 0; c < P.chunks.size(); c++) {
        const Plan::Chunk& ch = P.chunks[c];
        const int B = ch.count, np = ch.np, ln = (int)(c % nlane);
        const long long ms = (long long)np * np;
        hipStream_t st = S[ln];
        const MatItem* it = D.items.p + ch.start;
        // M^-1 = P V with V the solve of the row-major factor against the identity; then
        //   tr(M^-1)    = sum_i V[i, perm[i]]                 (traceinv)
        //   tr(M^-1 M') = sum_i sum_c V[i, c] M'[perm[i], c]  (M' generated on the fly)
        launch_assemble_items<T>(it, np, 0, mats[ln].p, ms, B, st, D.dim(ch.start, B));
        HIPCHK(launch_set_identity<T>(rhs[ln].p, np, ms, B, st));
        factor<T>(mats[ln].p, np, B, ipiv[ln].p, info[ln].p, pmap[ln].p, true, true, st);
        lu_solve_rm<T>(mats[ln].p, rhs[ln].p, np, np, B, ms, st, true);
        launch_build_perm(pmap[ln].p, np, NB, perm[ln].p, B, st);
        if (trace_inv) launch_diag_sum<T>(rhs[ln].p, 0, it, np, ms, dout.p + ch.start, B, st, perm[ln].p);
        else launch_dk_contract<T>(rhs[ln].p, ms, np, it, perm[ln].p, invp[ln].p, partial[ln].p, dout.p + ch.start,
                                   B, st);
    }
    S.join();
    std::vector<cx<double>> hout(nk);
    HIPCHK(hipMemcpyAsync(hout.data(), dout.p, sizeof(cx<double>) * nk, hipMemcpyDeviceToHost, S[0]));
    S.l[0].sync();
    for (int q = 0; q < nk; q++) out[P.order[q]] = hout[q];
}

int dk_log_det_multi(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                     cx<double>* out, int precision) {
    if (nk <= 0) return 0;
    if (int rc = check_media(meds, nk, midx)) return rc;
    return guarded([&]() -> int {
        if (precision == 1) dk_impl<float>(meds, nk, midx, ks, out);
        else dk_impl<double>(meds, nk, midx, ks, out);
        return 0;
    });
}

// tr(M(k)^-1) (traceinv, ms_matrix_util.c:319-335): the tr(M^-1 M') pipeline with the
// identity in place of M'.
int trace_inv_multi(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                    cx<double>* out, int precision) {
    if (nk <= 0) return 0;
    if (int rc = check_media(meds, nk, midx)) return rc;
    return guarded([&]() -> int {
        if (precision == 1) dk_impl<float>(meds, nk, midx, ks, out, true);
        else dk_impl<double>(meds, nk, midx, ks, out, true);
        return 0;
    });
}

int dk_log_det_points(const MediumView& mv, int nk, const cx<double>* ks, cx<double>* out, int precision) {
    return dk_log_det_multi(std::vector<MediumView>{mv}, nk, nullptr, ks, out, precision);
}

// ------------------------------------------------------------------ autotuner
// Times the batched trace-log of nitems matrices of padded size np on a synthetic
// medium (d = 3, N = np scatterers uniform in the unit-density ball, hard spheres
// alpha = 0.05, k on a grid over Re k in [5, 30] x Im k in [-2, 0]) for each
// candidate and keeps the fastest, one parameter at a time: the swap + TRSM shape,
// then the chunk size, then the stream count.  Every trial runs once untimed first.
int autotune(int np, int prec, int nitems, double* report) {
    if (np < 64 || np > 8192 || np % 64) { set_error("msgpu_autotune: N must give 64 <= Np <= 8192"); return -2; }
    nitems = std::max(8, nitems);
    // synthetic medium
    std::vector<double> pos(3 * (size_t)np);
    const double R = std::cbrt(np / (4.0 * kPi / 3.0));
    uint64_t st = 88172645463325252ull;
    auto uni = [&] { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (st >> 11) * (1.0 / 9007199254740992.0); };
    for (int i = 0; i < np;) {
        double x[3], r2 = 0;
        for (int c = 0; c < 3; c++) { x[c] = R * (2 * uni() - 1); r2 += x[c] * x[c]; }
        if (r2 > R * R) continue;
        for (int c = 0; c < 3; c++) pos[3 * (size_t)i + c] = x[c];
        i++;
    }
    MediumView mv;
    mv.d = 3; mv.n = np; mv.np = np; mv.m = ScModel{0, 0.05, 0.0}; mv.pos = pos.data();
    const int side = (int)std::ceil(std::sqrt((double)nitems));
    std::vector<cx<double>> ks(nitems), out(nitems);
    for (int l = 0; l < nitems; l++)
        ks[l] = cx<double>{5.0 + 25.0 * (l % side) / std::max(1, side - 1), -2.0 * (l / side) / std::max(1, side - 1)};
    const std::vector<MediumView> meds{mv};
    auto trial = [&](Tune t) -> double {
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            g_tune[{np, prec}] = t;
        }
        if (tracelog_multi(meds, nitems, nullptr, ks.data(), out.data(), nullptr, prec) < 0) return INFINITY;
        auto t0 = std::chrono::steady_clock::now();
        if (tracelog_multi(meds, nitems, nullptr, ks.data(), out.data(), nullptr, prec) < 0) return INFINITY;
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    Tune best;
    best.chunk = chunk_cap();
    best.streams = default_streams();
    double best_ms = INFINITY;
    for (int b2 = 0; b2 < 2; b2++)
        for (int rm = 0; rm < 2; rm++)
            for (int v = 0; v < swap_variants(); v++) {
                Tune t = best;
                t.swap = v;
                t.rm = rm;
                t.blk2 = b2;
                const double ms = trial(t);
                if (ms < best_ms) { best_ms = ms; best = t; }
            }
    {   // panel shape and chunk size together: smaller panel workgroups pay off with more
        // matrices per launch (several per CU)
        const Tune base = best;
        for (int pv = 0; pv < kPanelVariants; pv++)
            for (int c : {128, 256, 512, 1024, 2048}) {
                if (c > 512 && (c > nitems / 2 || c > max_batch(np, prec == 0 ? 16 : 8, 1))) continue;
                Tune t = base;
                t.panel = pv;
                t.chunk = c;
                if (pv == base.panel && c == base.chunk) continue;
                const double ms = trial(t);
                if (ms < best_ms) { best_ms = ms; best = t; }
            }
    }
    for (int ns : {1, 2, 4}) {
        Tune t = best;
        t.streams = ns;
        const double ms = trial(t);
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    {
        Tune t = best;
        t.graph = 1;
        const double ms = trial(t);
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune[{np, prec}] = best;
    }
    if (!std::isfinite(best_ms)) return -2;
    if (report) {
        report[0] = best.swap; report[1] = best.chunk; report[2] = best.streams; report[3] = best_ms;
        report[4] = best.rm;
        report[5] = best.panel;
        report[6] = best.graph;
        report[7] = best.blk2;
    }
    return 0;
}

int save_tuning(const char* path) {
    FILE* f = fopen(path, "w");
    if (!f) { set_error(std::string("msgpu_tune_save: cannot write ") + path); return -2; }
    std::lock_guard<std::mutex> g(g_tune_mu);
    fprintf(f, "# np precision swap_variant chunk streams row_major panel_variant graph block128\n");
    for (auto& kv : g_tune)
        fprintf(f, "%d %d %d %d %d %d %d %d %d\n", kv.first.first, kv.first.second, kv.second.swap, kv.second.chunk,
                kv.second.streams, kv.second.rm, kv.second.panel, kv.second.graph, kv.second.blk2);
    fclose(f);
    return 0;
}

int load_tuning(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) { set_error(std::string("msgpu_tune_load: cannot read ") + path); return -2; }
    char line[256];
    int n = 0;
    std::lock_guard<std::mutex> g(g_tune_mu);
    while (fgets(line, sizeof line, f)) {
        int a, b, c, d, e = 0, f6 = 0, g7 = 0, h8 = 0, i9 = 0;
        if (line[0] == '#') continue;
        const int got = sscanf(line, "%d %d %d %d %d %d %d %d %d", &a, &b, &c, &d, &e, &f6, &g7, &h8, &i9);
        if (got >= 4 && c >= 0 && c < swap_variants() && d >= 0 && e >= 0 && e <= NSTREAM_MAX && (f6 == 0 || f6 == 1) &&
            g7 >= 0 && g7 < kPanelVariants && (h8 == 0 || h8 == 1) && (i9 == 0 || i9 == 1)) {
            g_tune[{a, b}] = Tune{c, d, e, f6, g7, h8, i9};
            n++;
        }
    }
    fclose(f);
    return n;
}

int upload_positions(const MediumView& mv, const double** dpos, size_t* token) {
    return guarded([&]() -> int {
        const size_t bytes = sizeof(double) * (size_t)mv.n * mv.d;
        void* p = pool().get(bytes, *token);
        HIPCHK(hipMemcpy(p, mv.pos, bytes, hipMemcpyHostToDevice));
        *dpos = (const double*)p;
        return 0;
    });
}

void release_positions(const double* dpos, size_t token) {
    if (dpos) pool().put((void*)dpos, token);
}

// ------------------------------------------------------------------ amplitude solve
static double host_norm(const std::vector<cx<double>>& v, int n) {
    double s = 0;
    for (int i = 0; i < n; i++) s += v[i].x * v[i].x + v[i].y * v[i].y;
    return std::sqrt(s);
}

int solve_amplitudes(const MediumView& mv, cx<double> k, int source, cx<double>* a, int mixed, double tol) {
    if (mv.d < 1 || mv.d > 4) { set_error("msgpu: the device path supports d = 1..4"); return -3; }
    return guarded([&]() -> int {
        Lease L;
        const int np = mv.np, n = mv.n;
        const long long ms = (long long)np * np;
        DBuf<double> dpos((size_t)n * mv.d);
        DBuf<cx<double>> dk(1), ddiag(1), phi(np), x(np), r(np);
        DBuf<cx<double>> m64(ms);
        DBuf<int> ipiv(np), info(1), pmap(factor_ws_ints(np, NB, 1)), perm(np);
        HIPCHK(hipMemcpyAsync(dpos.p, mv.pos, sizeof(double) * n * mv.d, hipMemcpyHostToDevice, L.st));
        HIPCHK(hipMemcpyAsync(dk.p, &k, sizeof(k), hipMemcpyHostToDevice, L.st));
        launch_diag(dk.p, 1, mv.m, mv.d, 0, ddiag.p, nullptr, L.st);
        launch_incident(dpos.p, n, mv.d, np, k, source, 1.0, 0.0, phi.p, L.st);
        launch_assemble<double>(dpos.p, n, mv.d, np, dk.p, ddiag.p, nullptr, 0, m64.p, ms, 1, L.st);
        int steps = 0;
        // row-major LU of the symmetric M (left interchanges applied); M^-1 v = P V
        if (!mixed) {
            factor<double>(m64.p, np, 1, ipiv.p, info.p, pmap.p, true, true, L.st);
            launch_build_perm(pmap.p, np, NB, perm.p, 1, L.st);
            lu_solve_rm<double>(m64.p, phi.p, np, 1, 1, np, L.st);
            launch_scatter_rows<double>(x.p, phi.p, perm.p, np, 1, np, np, 1, L.st);
        } else {
            // complex64 LU, FP64 residuals: x <- x + M32^-1 (phi - M64 x)
            if (tol <= 0) tol = 1e-12;
            DBuf<cx<float>> m32(ms), d32(np), v32(np);
            launch_convert<double, float>(m64.p, m32.p, ms, L.st);
            factor<float>(m32.p, np, 1, ipiv.p, info.p, pmap.p, true, true, L.st);
            launch_build_perm(pmap.p, np, NB, perm.p, 1, L.st);
            HIPCHK(hipMemsetAsync(x.p, 0, sizeof(cx<double>) * np, L.st));
            std::vector<cx<double>> hphi(np), hr(np);
            HIPCHK(hipMemcpyAsync(hphi.data(), phi.p, sizeof(cx<double>) * np, hipMemcpyDeviceToHost, L.st));
            L.sync();
            const double nphi = host_norm(hphi, n);
            // Refinement converges when cond(M) eps32 < 1; near a resonance it stalls or
            // diverges.  It stops when the residual meets tol, or when it has failed to
            // halve in two consecutive steps, or after kMaxSteps steps; in the last two
            // cases the amplitudes are recomputed with the complex128 LU, so the result
            // always meets the complex128 solve's accuracy.
            constexpr int kMaxSteps = 30;
            double prev = INFINITY;
            int slow = 0;
            bool met = false;
            for (steps = 0; steps < kMaxSteps; steps++) {
                launch_residual(m64.p, x.p, phi.p, np, r.p, L.st);
                HIPCHK(hipMemcpyAsync(hr.data(), r.p, sizeof(cx<double>) * np, hipMemcpyDeviceToHost, L.st));
                L.sync();
                const double rn = host_norm(hr, n);
                if (rn <= tol * nphi) { met = true; break; }
                slow = (rn > 0.5 * prev) ? slow + 1 : 0;
                if (slow >= 2 || !std::isfinite(rn)) break;
                prev = rn;
                launch_convert<double, float>(r.p, v32.p, np, L.st);
                lu_solve_rm<float>(m32.p, v32.p, np, 1, 1, np, L.st);
                launch_scatter_rows<float>(d32.p, v32.p, perm.p, np, 1, np, np, 1, L.st);
                launch_axpy<float>(x.p, d32.p, np, L.st);
            }
            if (!met) {
                factor<double>(m64.p, np, 1, ipiv.p, info.p, pmap.p, true, true, L.st);
                launch_build_perm(pmap.p, np, NB, perm.p, 1, L.st);
                lu_solve_rm<double>(m64.p, phi.p, np, 1, 1, np, L.st);
                launch_scatter_rows<double>(x.p, phi.p, perm.p, np, 1, np, np, 1, L.st);
                steps++;         // the complex128 solve counts as one more step
            }
        }
        int hinfo = 0;
        HIPCHK(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, L.st));
        HIPCHK(hipMemcpyAsync(a, x.p, sizeof(cx<double>) * n, hipMemcpyDeviceToHost, L.st));
        L.sync();
        if (hinfo > 0)
            printf("[LAPACK] Element U(%d,%d) is exactly zero, the solution could not be computed.\n", hinfo, hinfo);
        return steps;
    });
}

// ------------------------------------------------------------------ resident factors
struct ResidentFactor {
    int np = 0, info = 0;
    DBuf<cx<double>> lu;
    DBuf<int> perm;
};
size_t resident_factor_bytes(const ResidentFactor& f) {
    return sizeof(cx<double>) * (size_t)f.np * f.np + sizeof(int) * (size_t)f.np;
}

int solve_resident(const MediumView& mv, cx<double> k, int source, cx<double>* a,
                   const std::shared_ptr<ResidentFactor>& reuse, std::shared_ptr<ResidentFactor>* keep) {
    if (mv.d < 1 || mv.d > 4) { set_error("msgpu: the device path supports d = 1..4"); return -3; }
    if (reuse && reuse->np != mv.np) { set_error("msgpu: resident factor of another size"); return -2; }
    return guarded([&]() -> int {
        Lease L;
        const int np = mv.np, n = mv.n;
        const long long ms = (long long)np * np;
        DBuf<double> own;
        const double* dpos = mv.dpos;
        if (!dpos) {
            own.alloc((size_t)n * mv.d);
            HIPCHK(hipMemcpyAsync(own.p, mv.pos, sizeof(double) * n * mv.d, hipMemcpyHostToDevice, L.st));
            dpos = own.p;
        }
        DBuf<cx<double>> phi(np), x(np);
        std::shared_ptr<ResidentFactor> f = reuse;
        if (!f) {
            f = std::make_shared<ResidentFactor>();
            f->np = np;
            f->lu.alloc(ms);
            f->perm.alloc(np);
            DBuf<cx<double>> dk(1), ddiag(1);
            DBuf<int> ipiv(np), info(1), pmap(factor_ws_ints(np, NB, 1));
            HIPCHK(hipMemcpyAsync(dk.p, &k, sizeof(k), hipMemcpyHostToDevice, L.st));
            launch_diag(dk.p, 1, mv.m, mv.d, 0, ddiag.p, nullptr, L.st);
            launch_incident(dpos, n, mv.d, np, k, source, 1.0, 0.0, phi.p, L.st);
            launch_assemble<double>(dpos, n, mv.d, np, dk.p, ddiag.p, nullptr, 0, f->lu.p, ms, 1, L.st);
            factor<double>(f->lu.p, np, 1, ipiv.p, info.p, pmap.p, true, true, L.st);
            launch_build_perm(pmap.p, np, NB, f->perm.p, 1, L.st);
            HIPCHK(hipMemcpyAsync(&f->info, info.p, sizeof(int), hipMemcpyDeviceToHost, L.st));
        } else {
            launch_incident(dpos, n, mv.d, np, k, source, 1.0, 0.0, phi.p, L.st);
        }
        lu_solve_rm<double>(f->lu.p, phi.p, np, 1, 1, np, L.st);
        launch_scatter_rows<double>(x.p, phi.p, f->perm.p, np, 1, np, np, 1, L.st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(a, x.p, sizeof(cx<double>) * n, hipMemcpyDeviceToHost, L.st));
        L.sync();
        if (f->info > 0)
            printf("[LAPACK] Element U(%d,%d) is exactly zero, the solution could not be computed.\n", f->info, f->info);
        if (keep) *keep = f;
        return 0;
    });
}

// ------------------------------------------------------------------ amplitudes of many items
// a = M(k)^-1 phi for items (medium, k) with phi the plane wave along 1x: grouped and
// chunked like the trace-log batches; each chunk assembled, factored row-major with left
// interchanges, solved (128-row blocks, one right-hand side per matrix) and permuted
// (M^-1 phi = P V).  Item q's N_q amplitudes go to out + sum_{q' < q} N_q'.
int amplitudes_multi(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                     cx<double>* out) {
    if (nk <= 0) return 0;
    if (int rc = check_media(meds, nk, midx)) return rc;
    return guarded([&]() -> int {
        Lease L;
        const Plan P = make_plan(meds, nk, midx, sizeof(cx<double>), 1, 1);
        DeviceItems D(meds, nk, midx, ks, P.order, L.st);
        DBuf<cx<double>> mats(P.mat_elems), X(P.piv_ints), Y(P.piv_ints);
        DBuf<int> ipiv(P.piv_ints), info(P.max_b), pmap(P.map_ints), perm(P.piv_ints);
        std::vector<size_t> off(nk + 1, 0);
        for (int i = 0; i < nk; i++) off[i + 1] = off[i] + meds[midx ? midx[i] : 0].n;
        std::vector<cx<double>> hy;
        std::vector<int> hinfo;
        for (const Plan::Chunk& ch : P.chunks) {
            const int B = ch.count, np = ch.np;
            const long long ms = (long long)np * np;
            launch_assemble_items<double>(D.items.p + ch.start, np, 0, mats.p, ms, B, L.st, D.dim(ch.start, B));
            factor<double>(mats.p, np, B, ipiv.p, info.p, pmap.p, true, true, L.st);
            launch_build_perm(pmap.p, np, NB, perm.p, B, L.st);
            launch_plane_items(D.items.p + ch.start, np, X.p, B, L.st);
            lu_solve_rm<double>(mats.p, X.p, np, 1, B, np, L.st);
            launch_scatter_rows<double>(Y.p, X.p, perm.p, np, 1, np, np, B, L.st);
            hy.resize((size_t)B * np);
            hinfo.resize(B);
            HIPCHK(hipMemcpyAsync(hy.data(), Y.p, sizeof(cx<double>) * B * np, hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpyAsync(hinfo.data(), info.p, sizeof(int) * B, hipMemcpyDeviceToHost, L.st));
            L.sync();
            for (int q = 0; q < B; q++) {
                const int i = P.order[ch.start + q];
                const int n = meds[midx ? midx[i] : 0].n;
                if (hinfo[q] > 0)
                    printf("[LAPACK] Element U(%d,%d) is exactly zero, the solution could not be computed.\n",
                           hinfo[q], hinfo[q]);
                std::copy(hy.begin() + (size_t)q * np, hy.begin() + (size_t)q * np + n, out + off[i]);
            }
        }
        return 0;
    });
}

// ------------------------------------------------------------------ eigenstates
// Eigenvectors at resonances (compute_eigenstates, ms_core.c:296-313, with
// inverse_iteration, ms_matrix_util.c:279-314): M(k_r) of all roots assembled and
// factored in one batch (row-major LU with left interchanges: M^-1 v = P V), then the
// inverse power iteration on every live vector together, from the constant unit vector
// (constant_unit_cvector): v <- normalize(M^-1 v) until |v - v_old| < toler or maxit
// steps; mu = 1/fac of the last step.  One host check per step.
int eigenstates(const MediumView& mv, int nr, const cx<double>* kr, int maxit, double toler, cx<double>* vecs,
                cx<double>* mu, int* conv, const cx<double>* v_init) {
    if (mv.d < 1 || mv.d > 4) { set_error("msgpu: the device path supports d = 1..4"); return -3; }
    if (nr <= 0) return 0;
    return guarded([&]() -> int {
        Lease L;
        const int np = mv.np, n = mv.n;
        const long long ms = (long long)np * np;
        const int bmax = std::max(1, std::min(nr, max_batch(np, sizeof(cx<double>), 1)));
        DBuf<double> dpos((size_t)n * mv.d);
        DBuf<cx<double>> dk(bmax), ddiag(bmax), mats((size_t)bmax * ms), V((size_t)bmax * np), X((size_t)bmax * np),
            Y((size_t)bmax * np), dfac(bmax);
        DBuf<double> ddelta(bmax);
        DBuf<int> ipiv((size_t)bmax * np), info(bmax), pmap(factor_ws_ints(np, NB, bmax)), perm((size_t)bmax * np),
            done(bmax);
        HIPCHK(hipMemcpyAsync(dpos.p, mv.pos, sizeof(double) * n * mv.d, hipMemcpyHostToDevice, L.st));
        // constant_unit_cvector (entries 1/(sqrt(n) + 0i)), normalised once more on entry
        // to inverse_iteration (ms_matrix_util.c:286); zero padding
        cx<double> u = cx<double>{1.0, 0.0} / cx<double>{std::sqrt((double)n), 0.0};
        {
            double r2 = 0.0;
            for (int i = 0; i < n; i++) r2 += u.x * u.x + u.y * u.y;
            const cx<double> f = u * (std::sqrt(r2) / std::hypot(u.x, u.y));
            u = u / f;
        }
        std::vector<cx<double>> v0((size_t)bmax * np, cx<double>{0.0, 0.0});
        for (int b = 0; b < bmax; b++)
            for (int i = 0; i < n; i++) v0[(size_t)b * np + i] = u;
        std::vector<cx<double>> hfac(bmax);
        std::vector<int> hdone(bmax), hinfo(bmax);
        for (int r0 = 0; r0 < nr; r0 += bmax) {
            const int B = std::min(bmax, nr - r0);
            HIPCHK(hipMemcpyAsync(dk.p, kr + r0, sizeof(cx<double>) * B, hipMemcpyHostToDevice, L.st));
            launch_diag(dk.p, B, mv.m, mv.d, 0, ddiag.p, nullptr, L.st);
            launch_assemble<double>(dpos.p, n, mv.d, np, dk.p, ddiag.p, nullptr, 0, mats.p, ms, B, L.st);
            factor<double>(mats.p, np, B, ipiv.p, info.p, pmap.p, true, true, L.st);
            launch_build_perm(pmap.p, np, NB, perm.p, B, L.st);
            if (v_init) {       // warm starts, normalised on entry as inverse_iteration does
                for (int b = 0; b < B; b++) {
                    const cx<double>* w = v_init + (size_t)(r0 + b) * n;
                    cx<double>* t = v0.data() + (size_t)b * np;
                    double r2 = 0.0;
                    int i0 = n;
                    for (int i = 0; i < n; i++) {
                        r2 += w[i].x * w[i].x + w[i].y * w[i].y;
                        if (i0 == n && (w[i].x != 0.0 || w[i].y != 0.0)) i0 = i;
                    }
                    const cx<double> f = i0 < n ? w[i0] * (std::sqrt(r2) / std::hypot(w[i0].x, w[i0].y))
                                                : cx<double>{1.0, 0.0};
                    for (int i = 0; i < n; i++) t[i] = i0 < n ? w[i] / f : w[i];
                }
            }
            HIPCHK(hipMemcpyAsync(V.p, v0.data(), sizeof(cx<double>) * B * np, hipMemcpyHostToDevice, L.st));
            launch_fill_int(done.p, 0, B, L.st);
            std::fill(hdone.begin(), hdone.end(), 0);
            for (int s = 0; s < maxit; s++) {
                HIPCHK(hipMemcpyAsync(X.p, V.p, sizeof(cx<double>) * B * np, hipMemcpyDeviceToDevice, L.st));
                lu_solve_rm<double>(mats.p, X.p, np, 1, B, np, L.st);
                launch_scatter_rows<double>(Y.p, X.p, perm.p, np, 1, np, np, B, L.st);
                launch_power_step(V.p, Y.p, n, np, toler, done.p, dfac.p, ddelta.p, B, L.st);
                HIPCHK(hipMemcpyAsync(hdone.data(), done.p, sizeof(int) * B, hipMemcpyDeviceToHost, L.st));
                L.sync();
                if (std::all_of(hdone.begin(), hdone.begin() + B, [](int x) { return x != 0; })) break;
            }
            HIPCHK(hipMemcpyAsync(hfac.data(), dfac.p, sizeof(cx<double>) * B, hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpyAsync(hinfo.data(), info.p, sizeof(int) * B, hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpy2DAsync(vecs + (size_t)r0 * n, sizeof(cx<double>) * n, V.p, sizeof(cx<double>) * np,
                                    sizeof(cx<double>) * n, B, hipMemcpyDeviceToHost, L.st));
            L.sync();
            for (int b = 0; b < B; b++) {
                if (hinfo[b] > 0)
                    printf("[LAPACK] Element U(%d,%d) is exactly zero, the inverse power iteration will fail.\n",
                           hinfo[b], hinfo[b]);
                mu[r0 + b] = cx<double>{1.0, 0.0} / hfac[b];
                conv[r0 + b] = hdone[b];
            }
        }
        return 0;
    });
}

// ------------------------------------------------------------------ psi
// The grid is computed in chunks of rows (about 4 MiB of psi each) on one stream; a second
// stream copies each finished chunk into one of kWaveSlots pinned buffers while later
// chunks compute, and the host moves it from there into the caller's memory, so the
// download and the host-side copy overlap the kernels instead of following them.
namespace {
constexpr int kWaveSlots = 4;
constexpr size_t kWaveChunk = (size_t)1 << 18;          // points per chunk (4 MiB of psi, 1024 workgroups)
// memcpy split over up to 8 threads for large blocks: a copy into fresh pageable memory
// runs at the rate its page faults are served, and several threads take them in parallel.
void parallel_copy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kPiece = (size_t)1 << 20;
    const int nt = (int)std::min<size_t>(8, bytes / kPiece);
    if (nt <= 1) { memcpy(dst, src, bytes); return; }
    const size_t per = (bytes + nt - 1) / nt;
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) {
        const size_t o = per * t, len = std::min(per, bytes - o);
        th.emplace_back([=] { memcpy((char*)dst + o, (const char*)src + o, len); });
    }
    memcpy(dst, src, per);
    for (auto& x : th) x.join();
}
struct EventSet {
    std::vector<hipEvent_t> ev;
    explicit EventSet(int count) : ev(count, nullptr) {
        for (auto& e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    EventSet(const EventSet&) = delete;
    EventSet& operator=(const EventSet&) = delete;
    ~EventSet() {
        drain_if_unwinding();
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};
}  // namespace

int wave_grid(const MediumView& mv, cx<double> k, int source, const cx<double>* a, int nx, int ny, double xmin,
              double xmax, double ymin, double ymax, cx<double>* psi) {
    if (mv.d < 1 || mv.d > 4) { set_error("msgpu: the device path supports d = 1..4"); return -3; }
    if (nx <= 0 || ny <= 0) return 0;
    return guarded([&]() -> int {
        Lease L, LC;
        const int n = mv.n;
        const size_t tot = (size_t)nx * ny;
        DBuf<double> own;
        const double* dpos = mv.dpos;          // a resident copy of the positions, if the caller keeps one
        if (!dpos) {
            own.alloc((size_t)n * mv.d);
            HIPCHK(hipMemcpyAsync(own.p, mv.pos, sizeof(double) * n * mv.d, hipMemcpyHostToDevice, L.st));
            dpos = own.p;
        }
        DBuf<cx<double>> da(n), dpsi(tot);
        HIPCHK(hipMemcpyAsync(da.p, a, sizeof(cx<double>) * n, hipMemcpyHostToDevice, L.st));
        const double dx = nx > 1 ? (xmax - xmin) / (nx - 1) : 0.0, dy = ny > 1 ? (ymax - ymin) / (ny - 1) : 0.0;
        const int rows = (int)std::max<size_t>(1, std::min<size_t>((size_t)ny, kWaveChunk / (size_t)nx));
        const int nch = (ny + rows - 1) / rows;
        EventSet done(nch), copied(nch);
        for (int c = 0; c < nch; c++) {
            const int r0 = c * rows, rc = std::min(rows, ny - r0);
            timed(5, 0.0, sizeof(cx<double>) * (double)rc * nx, L.st, [&] {
                launch_wave(dpos, n, mv.d, da.p, k, source, nx, rc, r0, xmin, dx, ymax, dy, nullptr,
                            dpsi.p + (size_t)r0 * nx, L.st);
            });
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(done.ev[c], L.st));
        }
        const int ns = std::min(kWaveSlots, nch);
        std::vector<std::unique_ptr<HostStage>> slot;
        for (int i = 0; i < ns; i++) slot.emplace_back(new HostStage((size_t)rows * nx));
        auto chunk_bytes = [&](int c) { return sizeof(cx<double>) * (size_t)std::min(rows, ny - c * rows) * nx; };
        auto enqueue_copy = [&](int c) {
            HIPCHK(hipStreamWaitEvent(LC.st, done.ev[c], 0));
            HIPCHK(hipMemcpyAsync(slot[c % ns]->p, dpsi.p + (size_t)c * rows * nx, chunk_bytes(c),
                                  hipMemcpyDeviceToHost, LC.st));
            HIPCHK(hipEventRecord(copied.ev[c], LC.st));
        };
        for (int c = 0; c < ns; c++) enqueue_copy(c);
        for (int c = 0; c < nch; c++) {
            HIPCHK(hipEventSynchronize(copied.ev[c]));
            parallel_copy(psi + (size_t)c * rows * nx, slot[c % ns]->p, chunk_bytes(c));
            if (c + ns < nch) enqueue_copy(c + ns);
        }
        L.sync();
        LC.sync();
        return 0;
    });
}

int wave_points(const MediumView& mv, cx<double> k, int source, const cx<double>* a, int npt, const double* pts,
                cx<double>* psi) {
    if (mv.d < 1 || mv.d > 4) { set_error("msgpu: the device path supports d = 1..4"); return -3; }
    return guarded([&]() -> int {
        Lease L;
        const int n = mv.n;
        DBuf<double> dpos((size_t)n * mv.d), dpts((size_t)npt * mv.d);
        DBuf<cx<double>> da(n), dpsi(npt);
        HIPCHK(hipMemcpyAsync(dpos.p, mv.pos, sizeof(double) * n * mv.d, hipMemcpyHostToDevice, L.st));
        HIPCHK(hipMemcpyAsync(dpts.p, pts, sizeof(double) * npt * mv.d, hipMemcpyHostToDevice, L.st));
        HIPCHK(hipMemcpyAsync(da.p, a, sizeof(cx<double>) * n, hipMemcpyHostToDevice, L.st));
        launch_wave(dpos.p, n, mv.d, da.p, k, source, npt, 1, 0, 0.0, 0.0, 0.0, 0.0, dpts.p, dpsi.p, L.st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(psi, dpsi.p, sizeof(cx<double>) * npt, hipMemcpyDeviceToHost, L.st));
        L.sync();
        return 0;
    });
}

// ------------------------------------------------------------------ host-buffer matrix services
// Host (pageable) column-major blocks move through a pinned buffer of the call: the
// calling thread packs or unpacks the block with memcpy, and the device copies the
// packed block at the link's rate.  (A pitched copy straight from pageable memory is
// staged by the runtime and serializes the threads that call concurrently.)
// stage <- host src (ld lds), rows x cols, packed (ld rows)
static void pack_block(const HostStage& hs, const cx<double>* src, int lds, int rows, int cols) {
    const size_t rb = sizeof(cx<double>) * rows;
    if (lds == rows) memcpy(hs.p, src, rb * cols);
    else for (int j = 0; j < cols; j++) memcpy(hs.p + (size_t)j * rows, src + (size_t)j * lds, rb);
}
// device dst (ld ldd) <- the packed stage
static void h2d_packed(cx<double>* dst, int ldd, int rows, int cols, const HostStage& hs, hipStream_t st) {
    const size_t rb = sizeof(cx<double>) * rows;
    HIPCHK(hipMemcpy2DAsync(dst, sizeof(cx<double>) * ldd, hs.p, rb, rb, cols, hipMemcpyHostToDevice, st));
}
// device dst (ld ldd) <- host src (ld lds), rows x cols
static void h2d_block(cx<double>* dst, int ldd, const cx<double>* src, int lds, int rows, int cols,
                      const HostStage& hs, hipStream_t st) {
    pack_block(hs, src, lds, rows, cols);
    h2d_packed(dst, ldd, rows, cols, hs, st);
}
// host dst (ld ldd) <- device src (ld lds): the copy into the stage, then (after the
// stream is synchronized) its unpacking
static void d2h_block_copy(const cx<double>* src, int lds, int rows, int cols, const HostStage& hs, hipStream_t st) {
    const size_t rb = sizeof(cx<double>) * rows;
    HIPCHK(hipMemcpy2DAsync(hs.p, rb, src, sizeof(cx<double>) * lds, rb, cols, hipMemcpyDeviceToHost, st));
}
static void unpack_block(cx<double>* dst, int ldd, int rows, int cols, const HostStage& hs) {
    const size_t rb = sizeof(cx<double>) * rows;
    if (ldd == rows) memcpy(dst, hs.p, rb * cols);
    else for (int j = 0; j < cols; j++) memcpy(dst + (size_t)j * ldd, hs.p + (size_t)j * rows, rb);
}

int build_matrix_host(const MediumView& mv, cx<double> k, int kind, cx<double>* out) {
    if (mv.d < 1 || mv.d > 4) { set_error("msgpu: the device path supports d = 1..4"); return -3; }
    return guarded([&]() -> int {
        Lease L;
        const int np = mv.np, n = mv.n;
        DBuf<double> dpos((size_t)n * mv.d);
        DBuf<cx<double>> dk(1), ddiag(1), diz(1), m((size_t)np * np);
        HostStage hs((size_t)n * n);
        HIPCHK(hipMemcpyAsync(dpos.p, mv.pos, sizeof(double) * n * mv.d, hipMemcpyHostToDevice, L.st));
        HIPCHK(hipMemcpyAsync(dk.p, &k, sizeof(k), hipMemcpyHostToDevice, L.st));
        launch_diag(dk.p, 1, mv.m, mv.d, kind, ddiag.p, diz.p, L.st);
        launch_assemble<double>(dpos.p, n, mv.d, np, dk.p, ddiag.p, diz.p, kind, m.p, (long long)np * np, 1, L.st);
        HIPCHK(hipGetLastError());
        d2h_block_copy(m.p, np, n, n, hs, L.st);
        L.sync();
        unpack_block(out, n, n, n, hs);
        return 0;
    });
}

int incident_host(const MediumView& mv, cx<double> k, int source, cx<double>* out) {
    return guarded([&]() -> int {
        Lease L;
        const int n = mv.n;
        DBuf<double> dpos((size_t)n * mv.d);
        DBuf<cx<double>> v(mv.np);
        HIPCHK(hipMemcpyAsync(dpos.p, mv.pos, sizeof(double) * n * mv.d, hipMemcpyHostToDevice, L.st));
        launch_incident(dpos.p, n, mv.d, mv.np, k, source, 1.0, 0.0, v.p, L.st);
        HIPCHK(hipMemcpyAsync(out, v.p, sizeof(cx<double>) * n, hipMemcpyDeviceToHost, L.st));
        L.sync();
        return 0;
    });
}

int plane_wave_host(const MediumView& mv, cx<double> k, double theta_deg, cx<double>* out) {
    return guarded([&]() -> int {
        Lease L;
        const int n = mv.n;
        const double th = theta_deg * (3.1415926535897932385 / 180);
        double ct = std::cos(th), stt = std::sin(th);
        if (mv.d == 1) ct = ct > 0. ? 1. : -1.;
        DBuf<double> dpos((size_t)n * mv.d);
        DBuf<cx<double>> v(mv.np);
        HIPCHK(hipMemcpyAsync(dpos.p, mv.pos, sizeof(double) * n * mv.d, hipMemcpyHostToDevice, L.st));
        launch_incident(dpos.p, n, mv.d, mv.np, k, 2, ct, stt, v.p, L.st);
        HIPCHK(hipMemcpyAsync(out, v.p, sizeof(cx<double>) * n, hipMemcpyDeviceToHost, L.st));
        L.sync();
        return 0;
    });
}

// Host matrix -> padded device matrix (identity padding).
static void upload_padded(cx<double>* dm, int n, int np, const cx<double>* a, int lda, const HostStage& hs,
                          hipStream_t st) {
    HIPCHK(hipMemsetAsync(dm, 0, sizeof(cx<double>) * (size_t)np * np, st));
    h2d_block(dm, np, a, lda, n, n, hs, st);
    launch_pad_identity<double>(dm, n, np, 1, st);
}

int lu_host(int n, cx<double>* a, int lda, int* ipiv, int nrhs, cx<double>* b, int ldb, cx<double>* tracelog,
            int* info) {
    return guarded([&]() -> int {
        Lease L;
        const int np = padded(n);
        DBuf<cx<double>> m((size_t)np * np), x((size_t)np * (nrhs > 0 ? nrhs : 1)), tl(1);
        DBuf<int> dpiv(np), dinfo(1), pmap(factor_ws_ints(np, NB, 1));
        HostStage ha((size_t)n * n), hb((size_t)n * (nrhs > 0 ? nrhs : 1));
        upload_padded(m.p, n, np, a, lda, ha, L.st);
        if (nrhs > 0) {
            HIPCHK(hipMemsetAsync(x.p, 0, sizeof(cx<double>) * (size_t)np * nrhs, L.st));
            h2d_block(x.p, np, b, ldb, n, nrhs, hb, L.st);
        }
        // factor without left interchanges, solve, then apply them for the LAPACK-format output
        factor<double>(m.p, np, 1, dpiv.p, dinfo.p, pmap.p, false, false, L.st);
        if (tracelog) launch_tracelog<double>(m.p, n, nullptr, np, (long long)np * np, dpiv.p, tl.p, nullptr, 1, L.st);
        if (nrhs > 0) lu_solve<double>(m.p, x.p, np, nrhs, 1, pmap.p, np, L.st);
        for (int k0 = NB; k0 < np; k0 += NB) launch_swap_left<double>(m.p, np, (long long)np * np, pmap.p, k0, NB, 1, L.st);
        HIPCHK(hipGetLastError());
        d2h_block_copy(m.p, np, n, n, ha, L.st);      // the upload from ha precedes it on the stream
        HIPCHK(hipMemcpyAsync(ipiv, dpiv.p, sizeof(int) * n, hipMemcpyDeviceToHost, L.st));
        HIPCHK(hipMemcpyAsync(info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost, L.st));
        if (nrhs > 0) d2h_block_copy(x.p, np, n, nrhs, hb, L.st);
        if (tracelog) HIPCHK(hipMemcpyAsync(tracelog, tl.p, sizeof(cx<double>), hipMemcpyDeviceToHost, L.st));
        L.sync();
        unpack_block(a, lda, n, n, ha);
        if (nrhs > 0) unpack_block(b, ldb, n, nrhs, hb);
        return 0;
    });
}

// ------------------------------------------------------------------ coalesced trace-logs
// Concurrent one-matrix trace-logs (the drop-in tracelog_lu, called from the caller's
// host threads) run as one batched factorization.  A call packs its matrix into a
// pinned buffer and queues a request; a waiting caller, while fewer than
// kTlBatches batches run, takes every queued request of the first one's size, runs
// them as one batch (one upload per matrix, the batched LU, trace-logs, left
// interchanges, one download per matrix) and marks them done; each caller unpacks its
// own factor.  Requests that arrive while batches run form the next ones; two batches
// at a time let one's copies overlap the other's factorization.
namespace {
// MSGPU_TRACELOG_KEEP_INPUT=1: the drop-in tracelog_lu leaves the caller's matrix as given
// instead of overwriting it with zgetrf_'s LAPACK-format factor (which the reference's
// callers never read: ms_core.c:74 and 126 rebuild M before every call).  That saves the
// factor's download and its unpacking: a third of the host-link bytes of a drop-in log-det.
static bool tracelog_keeps_input() {
    static const bool v = [] {
        const char* e = getenv("MSGPU_TRACELOG_KEEP_INPUT");
        return e && *e && *e != '0';
    }();
    return v;
}
struct TlReq {
    int n;
    const HostStage* hs;       // packed n x n matrix in, LAPACK-format factor out
    int* ipiv;                  // n pivots out (host)
    cx<double> tl{0.0, 0.0};
    int info = 0;
    bool done = false;
    std::string err;
};
std::mutex g_tl_mu;
std::condition_variable g_tl_cv;
std::vector<TlReq*> g_tl_queue;
int g_tl_active = 0;
constexpr int kTlBatches = 2;

void run_tl_batch(const std::vector<TlReq*>& B) {
    Lease L;
    const int n = B[0]->n, np = padded(n), nb = (int)B.size();
    const long long ms = (long long)np * np;
    DBuf<cx<double>> m((size_t)nb * ms), tl(nb);
    DBuf<int> dpiv((size_t)nb * np), dinfo(nb), pmap(factor_ws_ints(np, NB, nb));
    HIPCHK(hipMemsetAsync(m.p, 0, sizeof(cx<double>) * (size_t)nb * ms, L.st));
    for (int i = 0; i < nb; i++) h2d_packed(m.p + (size_t)i * ms, np, n, n, *B[i]->hs, L.st);
    launch_pad_identity<double>(m.p, n, np, nb, L.st);
    factor<double>(m.p, np, nb, dpiv.p, dinfo.p, pmap.p, false, false, L.st);
    launch_tracelog<double>(m.p, n, nullptr, np, ms, dpiv.p, tl.p, nullptr, nb, L.st);
    const bool back = !tracelog_keeps_input();
    if (back)
        for (int k0 = NB; k0 < np; k0 += NB) launch_swap_left<double>(m.p, np, ms, pmap.p, k0, NB, nb, L.st);
    HIPCHK(hipGetLastError());
    std::vector<int> hpiv((size_t)nb * np), hinfo(nb);
    std::vector<cx<double>> htl(nb);
    if (back)
        for (int i = 0; i < nb; i++) d2h_block_copy(m.p + (size_t)i * ms, np, n, n, *B[i]->hs, L.st);
    HIPCHK(hipMemcpyAsync(hpiv.data(), dpiv.p, sizeof(int) * (size_t)nb * np, hipMemcpyDeviceToHost, L.st));
    HIPCHK(hipMemcpyAsync(hinfo.data(), dinfo.p, sizeof(int) * nb, hipMemcpyDeviceToHost, L.st));
    HIPCHK(hipMemcpyAsync(htl.data(), tl.p, sizeof(cx<double>) * nb, hipMemcpyDeviceToHost, L.st));
    L.sync();
    for (int i = 0; i < nb; i++) {
        memcpy(B[i]->ipiv, hpiv.data() + (size_t)i * np, sizeof(int) * n);
        B[i]->info = hinfo[i];
        B[i]->tl = htl[i];
    }
}
}  // namespace

int tracelog_host_coalesced(int n, cx<double>* a, int lda, int* ipiv, cx<double>* tracelog, int* info) {
    return guarded([&]() -> int {
        if (n < 0) throw std::runtime_error("tracelog_lu: negative matrix size");
        if (n > 8192) throw std::runtime_error("matrices above N = 8192 are not supported by the device LU");
        if (n == 0) {                   // the empty sum divided by N = 0
            *tracelog = cx<double>{NAN, NAN};
            *info = 0;
            return 0;
        }
        HostStage hs((size_t)n * n);
        pack_block(hs, a, lda, n, n);
        TlReq req;
        req.n = n;
        req.hs = &hs;
        req.ipiv = ipiv;
        std::unique_lock<std::mutex> lk(g_tl_mu);
        g_tl_queue.push_back(&req);
        while (!req.done) {
            if (g_tl_active >= kTlBatches || g_tl_queue.empty()) {
                g_tl_cv.wait(lk);
                continue;
            }
            // lead one batch: the queued requests of the first request's size
            g_tl_active++;
            std::vector<TlReq*> batch;
            const int bn = g_tl_queue.front()->n;
            // at most 256 matrices, and no more than the workspace budget holds (one at least)
            const size_t cap = std::max<size_t>(
                1, std::min<size_t>(256, batch_budget() / (sizeof(cx<double>) * (size_t)padded(bn) * padded(bn))));
            for (size_t i = 0; i < g_tl_queue.size();) {
                if (g_tl_queue[i]->n == bn && batch.size() < cap) {
                    batch.push_back(g_tl_queue[i]);
                    g_tl_queue.erase(g_tl_queue.begin() + (long)i);
                } else {
                    i++;
                }
            }
            lk.unlock();
            std::string err;
            try {
                run_tl_batch(batch);
            } catch (const std::exception& ex) {
                err = ex.what();
            }
            lk.lock();
            for (TlReq* r : batch) {
                r->err = err;
                r->done = true;
            }
            g_tl_active--;
            g_tl_cv.notify_all();
        }
        lk.unlock();
        if (!req.err.empty()) throw std::runtime_error(req.err);
        if (!tracelog_keeps_input()) unpack_block(a, lda, n, n, hs);
        *tracelog = req.tl;
        *info = req.info;
        return 0;
    });
}

// ------------------------------------------------------------------ coalesced psi points
// Concurrent one-point evaluations of psi (the drop-in eval_total_wave, called once per
// pixel by every thread of the caller's pixel loop) run as one launch.  The positions,
// amplitudes, k and source of the calls are a "context": its device copy is made once
// per content and reused while the calls keep passing the same content (each call
// compares its inputs with the context's host copy, outside the lock).  A call queues its
// point; a waiting caller, while fewer than kWvBatches batches run, takes every queued
// point of the first one's context (at most kWvMax), runs them as one batch (one launch
// that reads the points from pinned host memory and writes psi there) and marks them done.
namespace {
struct WaveCtx {
    int d = 0, n = 0, source = 0;
    cx<double> k{0.0, 0.0};
    std::vector<double> pos;
    std::vector<cx<double>> a;
    DBuf<double> dpos;
    DBuf<cx<double>> da;
    bool same(const MediumView& mv, cx<double> k_, int source_, const cx<double>* a_) const {
        return d == mv.d && n == mv.n && source == source_ && k.x == k_.x && k.y == k_.y &&
               memcmp(pos.data(), mv.pos, sizeof(double) * (size_t)n * d) == 0 &&
               memcmp(a.data(), a_, sizeof(cx<double>) * (size_t)n) == 0;
    }
};
struct WaveReq {
    std::shared_ptr<WaveCtx> ctx;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    cx<double> psi{0.0, 0.0};
    std::string err;
    std::atomic<bool> done{false};      // set last, with release order, by the batch's runner
};
// The queue is guarded by g_wv_mu; the counters are read without it by the waiting callers,
// which spin (yielding) on their own request's flag rather than sleeping on a condition
// variable: a batch takes about as long as a thread wake-up, so sleeping callers would
// spend most of the time being woken.
std::mutex g_wv_mu;
std::vector<WaveReq*> g_wv_queue;
std::shared_ptr<WaveCtx> g_wv_ctx;
std::atomic<int> g_wv_queued{0};       // g_wv_queue.size()
std::atomic<int> g_wv_active{0};       // running batches
std::atomic<int> g_wv_callers{0};      // threads inside wave_point_coalesced
std::atomic<int> g_wv_hwm{0};          // the most callers seen at once since the last time none was inside
std::atomic<int> g_wv_inflight{0};     // points in the running batches
// Up to kWvBatches batches run at once, and a leader gathers until a kWvSplit-th of the
// caller's threads (those not in a running batch, at most) have queued: the batches then
// overlap the threads' return trips (best of 1-4 batches at once, split 1-4, measured with
// 1, 4 and 16 calling threads at N = 1000 on an MI355X).
constexpr int kWvBatches = 4;
constexpr int kWvSplit = 4;
constexpr int kWvMax = 4096;
// A leader waits at most this long for the other callers to queue their points.
constexpr auto kWvGather = std::chrono::microseconds(50);

void run_wave_batch(const std::vector<WaveReq*>& B) {
    Lease L;
    const WaveCtx& c = *B[0]->ctx;
    const int npt = (int)B.size(), d = c.d;
    HostStage hp(((size_t)npt * d + 1) / 2), hq(npt);
    double* hx = reinterpret_cast<double*>(hp.p);
    for (int i = 0; i < npt; i++)
        for (int q = 0; q < d; q++) hx[(size_t)i * d + q] = B[i]->x[q];
    // the kernel reads the points from, and writes psi into, the pinned stages themselves
    // (a few bytes per workgroup over the host link): no copy commands, so a batch costs one
    // launch and one synchronisation
    launch_wave_points(c.dpos.p, c.n, d, c.da.p, c.k, c.source, npt, hx, hq.p, L.st);
    HIPCHK(hipGetLastError());
    L.sync();
    for (int i = 0; i < npt; i++) B[i]->psi = hq.p[i];
}

// The resident psi kernel (launch_wave_resident): callers take consecutive request
// indices, fill slot index % kRing of the ring in coherent pinned host memory and
// publish it; the kernel's workgroup index % kResWG serves it and writes psi back into
// the slot.  A caller spins on its slot's done word, relaunching the kernel when its
// launch has ended (idle for kResIdleMs or older than kResLifeMs): each workgroup
// resumes at its next request index, so no request is lost or served twice.  A slot is
// reused by request index + kRing only after the caller of index has read its psi.
// The ring and the stop word are never freed (a launch may still read them at exit);
// the destructor sets the stop word, which ends a running launch at its next poll.
constexpr int kRing = 4096;
constexpr int kResWG = 64;
constexpr double kResIdleMs = 20.0;
constexpr double kResLifeMs = 2000.0;

struct WaveResident {
    WaveSlot* ring = nullptr;           // host addresses
    unsigned* stop = nullptr;
    WaveSlot* dring = nullptr;          // the same memory as the device addresses it
    unsigned* dstop = nullptr;
    DBuf<unsigned long long> resume, last;
    DBuf<unsigned> closing;
    std::unique_ptr<std::atomic<unsigned long long>[]> consumed;
    std::atomic<unsigned long long> next{0};
    std::atomic<bool> failed{false};
    std::string error;
    std::mutex mu;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool launched = false;
    long long idle_ticks = 0, life_ticks = 0;

    WaveResident() {
        int dev = 0, khz = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) throw std::runtime_error("msgpu: the device reports no wall clock rate");
        idle_ticks = (long long)(kResIdleMs * khz);
        life_ticks = (long long)(kResLifeMs * khz);
        void* h = nullptr;
        HIPCHK(hipHostMalloc(&h, sizeof(WaveSlot) * kRing + 64, hipHostMallocCoherent | hipHostMallocMapped));
        memset(h, 0, sizeof(WaveSlot) * kRing + 64);
        ring = static_cast<WaveSlot*>(h);
        stop = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(h) + sizeof(WaveSlot) * kRing);
        void* dh = nullptr;
        HIPCHK(hipHostGetDevicePointer(&dh, h, 0));
        dring = static_cast<WaveSlot*>(dh);
        dstop = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(dh) + sizeof(WaveSlot) * kRing);
        consumed.reset(new std::atomic<unsigned long long>[kRing]);
        for (int q = 0; q < kRing; q++) consumed[q].store(0);
        resume.alloc(kResWG);
        last.alloc(1);
        closing.alloc(1);
        std::vector<unsigned long long> r0(kResWG);
        for (int w = 0; w < kResWG; w++) r0[w] = (unsigned long long)w;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIPCHK(hipMemcpyAsync(resume.p, r0.data(), sizeof(unsigned long long) * kResWG, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    // At process exit: the stop word ends a running launch at its next poll; wait for it
    // briefly so that the process does not end with the launch still reading the ring.
    void shutdown() {
        __atomic_store_n(stop, 1u, __ATOMIC_RELEASE);
        std::lock_guard<std::mutex> g(mu);
        for (int t = 0; launched && t < 1000 && hipEventQuery(ev) == hipErrorNotReady; t++)
            std::this_thread::sleep_for(std::chrono::microseconds(100));
    }

    // Launches the kernel unless a launch is running; false (and the error kept) if it
    // cannot, after which every waiting caller gives up.
    bool ensure_running() {
        std::lock_guard<std::mutex> g(mu);
        if (failed.load()) return false;
        try {
            if (launched) {
                const hipError_t q = hipEventQuery(ev);
                if (q == hipErrorNotReady) return true;
                HIPCHK(q);
            }
            HIPCHK(hipMemsetAsync(closing.p, 0, sizeof(unsigned), st));
            HIPCHK(hipMemsetAsync(last.p, 0, sizeof(unsigned long long), st));
            launch_wave_resident(dring, kRing, dstop, resume.p, closing.p, last.p, idle_ticks, life_ticks, kResWG, st);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(ev, st));
            launched = true;
            return true;
        } catch (const std::exception& ex) {
            error = ex.what();
            failed.store(true);
            return false;
        }
    }

    cx<double> request(const WaveCtx& c, const double* x) {
        const unsigned long long i = next.fetch_add(1);
        const int s = (int)(i % kRing);
        if (i >= (unsigned long long)kRing)
            while (consumed[s].load(std::memory_order_acquire) != i - kRing + 1) std::this_thread::yield();
        WaveSlot& sl = ring[s];
        sl.pos = c.dpos.p;
        sl.a = c.da.p;
        for (int q = 0; q < 4; q++) sl.x[q] = q < c.d ? x[q] : 0.0;
        sl.k = c.k;
        sl.n = c.n;
        sl.d = c.d;
        sl.source = c.source;
        __atomic_store_n(&sl.seq, i + 1, __ATOMIC_RELEASE);
        bool ok = ensure_running();
        const auto t0 = std::chrono::steady_clock::now();
        for (unsigned spins = 1; ok && __atomic_load_n(&sl.done, __ATOMIC_ACQUIRE) != i + 1; spins++) {
            if ((spins & 4095) != 0) {
                __builtin_ia32_pause();
                continue;
            }
            ok = ensure_running();      // the launch may have ended before serving i
            if (ok && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10)) {
                // a request left unserved would hold up every later one of its workgroup:
                // stop the kernel and fail every call from here on
                std::lock_guard<std::mutex> g(mu);
                error = "request not served within 10 s";
                failed.store(true);
                __atomic_store_n(stop, 1u, __ATOMIC_RELEASE);
                ok = false;
            }
        }
        const cx<double> v = sl.psi;
        consumed[s].store(i + 1, std::memory_order_release);
        if (!ok) throw std::runtime_error("resident psi kernel: " + error);
        return v;
    }
};

static bool wave_resident_enabled() {
    static const bool on = [] {
        const char* e = getenv("MSGPU_WAVE_RESIDENT");
        return !(e && e[0] == '0');
    }();
    return on;
}

static WaveResident* g_wave_resident = nullptr;

static WaveResident& wave_resident() {
    // never destroyed (a launch may outlive static destruction); stopped at exit
    static WaveResident* r = [] {
        WaveResident* w = new WaveResident();
        g_wave_resident = w;
        atexit([] { g_wave_resident->shutdown(); });
        return w;
    }();
    return *r;
}
}  // namespace

int wave_point_coalesced(const MediumView& mv, cx<double> k, int source, const cx<double>* a, const double* x,
                         cx<double>* psi) {
    if (mv.d < 1 || mv.d > 4) { set_error("msgpu: the device path supports d = 1..4"); return -3; }
    return guarded([&]() -> int {
        // The calling thread's last content, kept by the thread itself: the common case (the
        // same content as its previous call) takes no lock and touches no shared counter.
        thread_local std::shared_ptr<WaveCtx> t_ctx;
        if (t_ctx && t_ctx->same(mv, k, source, a) && wave_resident_enabled()) {
            *psi = wave_resident().request(*t_ctx, x);
            return 0;
        }
        std::shared_ptr<WaveCtx> ctx;
        {
            std::lock_guard<std::mutex> g(g_wv_mu);
            ctx = g_wv_ctx;
        }
        if (!ctx || !ctx->same(mv, k, source, a)) {        // a new content: its device copy
            auto nc = std::make_shared<WaveCtx>();
            nc->d = mv.d;
            nc->n = mv.n;
            nc->source = source;
            nc->k = k;
            nc->pos.assign(mv.pos, mv.pos + (size_t)mv.n * mv.d);
            nc->a.assign(a, a + mv.n);
            nc->dpos.alloc((size_t)mv.n * mv.d);
            nc->da.alloc(mv.n);
            Lease L;
            HIPCHK(hipMemcpyAsync(nc->dpos.p, nc->pos.data(), sizeof(double) * nc->pos.size(), hipMemcpyHostToDevice,
                                  L.st));
            HIPCHK(hipMemcpyAsync(nc->da.p, nc->a.data(), sizeof(cx<double>) * mv.n, hipMemcpyHostToDevice, L.st));
            L.sync();
            std::lock_guard<std::mutex> g(g_wv_mu);
            if (g_wv_ctx && g_wv_ctx->same(mv, k, source, a)) ctx = g_wv_ctx;   // another thread was first
            else g_wv_ctx = ctx = nc;
        }
        if (wave_resident_enabled()) {
            t_ctx = ctx;
            *psi = wave_resident().request(*ctx, x);
            return 0;
        }
        WaveReq req;
        req.ctx = ctx;
        for (int q = 0; q < mv.d; q++) req.x[q] = x[q];
        {
            std::lock_guard<std::mutex> g(g_wv_mu);
            const int c = ++g_wv_callers;
            if (c > g_wv_hwm.load()) g_wv_hwm.store(c);
            g_wv_queue.push_back(&req);
            g_wv_queued.store((int)g_wv_queue.size());
        }
        while (!req.done.load(std::memory_order_acquire)) {
            if (g_wv_active.load() >= kWvBatches || g_wv_queued.load() == 0) {
                std::this_thread::yield();
                continue;
            }
            // gather: the threads of the caller's loop that are not in a running batch are
            // about to queue a point (the loops call again at once); wait for a share of
            // them, briefly
            const auto until = std::chrono::steady_clock::now() + kWvGather;
            while (!req.done.load(std::memory_order_acquire) &&
                   g_wv_queued.load() <
                       std::min(g_wv_hwm.load() - g_wv_inflight.load(), (g_wv_hwm.load() + kWvSplit - 1) / kWvSplit) &&
                   std::chrono::steady_clock::now() < until)
                std::this_thread::yield();
            std::vector<WaveReq*> batch;
            {
                std::lock_guard<std::mutex> g(g_wv_mu);
                if (req.done.load(std::memory_order_acquire) || g_wv_active.load() >= kWvBatches ||
                    g_wv_queue.empty())
                    continue;
                g_wv_active++;
                const WaveCtx* bc = g_wv_queue.front()->ctx.get();
                for (size_t i = 0; i < g_wv_queue.size();) {
                    if (g_wv_queue[i]->ctx.get() == bc && (int)batch.size() < kWvMax) {
                        batch.push_back(g_wv_queue[i]);
                        g_wv_queue.erase(g_wv_queue.begin() + (long)i);
                    } else {
                        i++;
                    }
                }
                g_wv_queued.store((int)g_wv_queue.size());
                g_wv_inflight += (int)batch.size();
            }
            std::string err;
            try {
                run_wave_batch(batch);
            } catch (const std::exception& ex) {
                err = ex.what();
            }
            g_wv_inflight -= (int)batch.size();
            g_wv_active--;
            for (WaveReq* r : batch) {           // r belongs to its caller once done is set
                r->err = err;
                r->done.store(true, std::memory_order_release);
            }
        }
        {
            std::lock_guard<std::mutex> g(g_wv_mu);
            if (--g_wv_callers == 0) g_wv_hwm.store(0);
        }
        if (!req.err.empty()) throw std::runtime_error(req.err);
        *psi = req.psi;
        return 0;
    });
}

int lu_solve_host(int n, const cx<double>* lu, int lda, const int* ipiv, int nrhs, cx<double>* b, int ldb) {
    return guarded([&]() -> int {
        Lease L;
        const int np = padded(n);
        // LAPACK-format factor: apply ipiv to b sequentially on the host-side copy, then two triangular solves
        std::vector<int> perm(np);
        for (int i = 0; i < np; i++) perm[i] = i;
        for (int i = 0; i < n; i++) std::swap(perm[i], perm[ipiv[i] - 1]);
        DBuf<cx<double>> m((size_t)np * np), x((size_t)np * nrhs), y((size_t)np * nrhs);
        DBuf<int> dperm(np);
        HostStage ha((size_t)n * n), hb((size_t)n * nrhs);
        upload_padded(m.p, n, np, lu, lda, ha, L.st);
        HIPCHK(hipMemsetAsync(x.p, 0, sizeof(cx<double>) * (size_t)np * nrhs, L.st));
        h2d_block(x.p, np, b, ldb, n, nrhs, hb, L.st);
        HIPCHK(hipMemcpyAsync(dperm.p, perm.data(), sizeof(int) * np, hipMemcpyHostToDevice, L.st));
        launch_gather_rows<double>(y.p, x.p, dperm.p, np, nrhs, 0, 0, 1, L.st);
        const long long ms = (long long)np * np;
        for (int k0 = 0; k0 < np; k0 += NB) {
            launch_trsm_block<double>(m.p, y.p, np, nrhs, k0, 0, ms, 0, 1, L.st);
            if (k0 + NB < np)
                launch_gemm_sub<double>(y.p + k0 + NB, m.p + (k0 + NB) + (size_t)k0 * np, y.p + k0, np - k0 - NB, nrhs,
                                        NB, np, np, np, 0, 0, 0, 1, L.st);
        }
        for (int k0 = np - NB; k0 >= 0; k0 -= NB) {
            launch_trsm_block<double>(m.p, y.p, np, nrhs, k0, 1, ms, 0, 1, L.st);
            if (k0 > 0)
                launch_gemm_sub<double>(y.p, m.p + (size_t)k0 * np, y.p + k0, k0, nrhs, NB, np, np, np, 0, 0, 0, 1, L.st);
        }
        HIPCHK(hipGetLastError());
        d2h_block_copy(y.p, np, n, nrhs, hb, L.st);
        L.sync();
        unpack_block(b, ldb, n, nrhs, hb);
        return 0;
    });
}

}  // namespace msg

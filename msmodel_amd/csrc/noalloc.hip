// noalloc.hip -- msgpu_tracelog_batch_noalloc: the trace-log map of msgpu_tracelog_batch
// (FP64 or FP32) computed without allocating device memory, for callers that cannot
// call hipMalloc (code under HIP-graph capture, tight serving loops, harnesses that
// forbid allocation).
//
// Every device buffer of the computation -- the positions, the chunk's wave numbers,
// items, matrices, pivots, row maps and 128-column update blocks, and its results --
// lives in `scratch`, one file-scope device array of kScratchBytes per GPU, held from
// the library's load until the process ends.  Its device address is fetched once per
// device with hipGetSymbolAddress (at initialisation for device 0, on the first call
// for another current device); host code only ever passes that address, never the
// array's name.  The map runs in chunks of as many matrices as the scratch holds, on
// the null stream, with the same kernels and launch shapes as msgpu_tracelog_batch, so
// each result is bit for bit the one the allocating entry point returns.
//
// Calls on one process share each GPU's array: they are serialised by a lock, and each
// chunk writes every scratch word it later reads (assembly fills whole matrices, the
// factorisation sets its info words and row maps), so nothing stale from a previous
// call is read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

#include "capi_internal.h"
#include "kernels.h"

namespace msg {

template <class T>
int factor_untimed(cx<T>* mats, int np, int nbatch, int* ipiv, int* info, int* ws, hipStream_t st);

constexpr size_t kScratchBytes = (size_t)256 << 20;     // 256 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static unsigned char* g_scratch[kMaxDevices];           // device address of `scratch`, per device
static std::mutex g_na_mu;

// With device dev current: fetch the address of this device's copy of `scratch`.
static int scratch_setup_locked(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return -1;
    if (g_scratch[dev]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        set_error(std::string("msgpu: hipGetSymbolAddress of the allocation-free scratch failed: ") +
                  hipGetErrorString(e));
        return -1;
    }
    g_scratch[dev] = static_cast<unsigned char*>(p);
    return 0;
}

// Called by the library's initialisation with device 0 current.
int noalloc_setup() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    std::lock_guard<std::mutex> g(g_na_mu);
    return scratch_setup_locked(dev);
}

namespace {

__global__ void noalloc_items_kernel(MatItem* __restrict__ items, const cx<double>* __restrict__ ks,
                                     const double* pos, int n, int d, ScModel m, int cnt) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= cnt) return;
    MatItem it;
    it.pos = pos;
    it.n = n;
    it.d = d;
    it.m = m;
    it.k = ks[l];
    it.diag0 = it.diag1 = cx<double>{0.0, 0.0};
    items[l] = it;
}

// Byte offsets of a chunk of B matrices of padded size np inside the scratch.
struct Layout {
    size_t pos, ks, items, out, info, ipiv, ws, mats, total;
};
inline size_t up256(size_t x) { return (x + 255) / 256 * 256; }
Layout layout_of(int n, int d, int np, int B, size_t elem) {
    Layout L;
    size_t o = 0;
    L.pos = o;   o = up256(o + sizeof(double) * (size_t)n * d);
    L.ks = o;    o = up256(o + sizeof(cx<double>) * B);
    L.items = o; o = up256(o + sizeof(MatItem) * B);
    L.out = o;   o = up256(o + sizeof(cx<double>) * B);
    L.info = o;  o = up256(o + sizeof(int) * B);
    L.ipiv = o;  o = up256(o + sizeof(int) * (size_t)B * np);
    L.ws = o;    o = up256(o + sizeof(int) * factor_ws_ints(np, 64, B));
    L.mats = o;  o = up256(o + elem * (size_t)np * np * B);
    L.total = o;
    return L;
}

constexpr int kMaxChunk = 64;

#define NACHK(x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
void run_map(unsigned char* base, const MediumView& mv, int np, int B, const Layout& L, int nk,
             const cx<double>* ks, cx<double>* out) {
    hipStream_t st = nullptr;          // the null stream: no stream is created
    double* dpos = reinterpret_cast<double*>(base + L.pos);
    cx<double>* dks = reinterpret_cast<cx<double>*>(base + L.ks);
    MatItem* items = reinterpret_cast<MatItem*>(base + L.items);
    cx<double>* dout = reinterpret_cast<cx<double>*>(base + L.out);
    int* info = reinterpret_cast<int*>(base + L.info);
    int* ipiv = reinterpret_cast<int*>(base + L.ipiv);
    int* ws = reinterpret_cast<int*>(base + L.ws);
    cx<T>* mats = reinterpret_cast<cx<T>*>(base + L.mats);
    const long long ms = (long long)np * np;
    NACHK(hipMemcpyAsync(dpos, mv.pos, sizeof(double) * (size_t)mv.n * mv.d, hipMemcpyHostToDevice, st));
    for (int l0 = 0; l0 < nk; l0 += B) {
        const int b = std::min(B, nk - l0);
        NACHK(hipMemcpyAsync(dks, ks + l0, sizeof(cx<double>) * b, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(noalloc_items_kernel, dim3((b + 63) / 64), dim3(64), 0, st, items, dks, dpos, mv.n,
                           mv.d, mv.m, b);
        launch_item_diag(items, b, st);
        launch_assemble_items<T>(items, np, 0, mats, ms, b, st, mv.d);
        factor_untimed<T>(mats, np, b, ipiv, info, ws, st);
        launch_tracelog<T>(mats, 0, items, np, ms, ipiv, dout, nullptr, b, st);
        NACHK(hipGetLastError());
        NACHK(hipMemcpyAsync(out + l0, dout, sizeof(cx<double>) * b, hipMemcpyDeviceToHost, st));
    }
    NACHK(hipStreamSynchronize(st));
}

}  // namespace
}  // namespace msg

using namespace msg;

extern "C" MSGPU_API int msgpu_tracelog_batch_noalloc(const msgpu_medium* med, int nk, const msgpu_cplx* ks,
                                                      msgpu_cplx* out, int mode) {
    if (!med || !med->pos || nk < 0 || (nk > 0 && (!ks || !out))) {
        set_error("msgpu_tracelog_batch_noalloc: bad arguments");
        return -2;
    }
    if (mode != MSGPU_FP64 && mode != MSGPU_FP32) {
        set_error("msgpu_tracelog_batch_noalloc: mode must be MSGPU_FP64 or MSGPU_FP32 (MSGPU_MIXED re-evaluates "
                  "cells in buffers of the batch's size)");
        return -4;
    }
    const MediumView mv = view_of(med);
    if (mv.d < 1 || mv.d > 4 || mv.n < 1) {
        set_error("msgpu_tracelog_batch_noalloc: the device path supports d = 1..4 and N >= 1");
        return -3;
    }
    const size_t elem = mode == MSGPU_FP64 ? sizeof(cx<double>) : sizeof(cx<float>);
    const int np = padded(mv.n);
    if (layout_of(mv.n, mv.d, np, 1, elem).total > kScratchBytes) {
        char msg[256];
        snprintf(msg, sizeof msg,
                 "msgpu_tracelog_batch_noalloc: N = %d needs %zu bytes of scratch for one %s matrix; the "
                 "allocation-free scratch holds %zu bytes (N <= 4032 in complex128, N <= 5760 in complex64)",
                 mv.n, layout_of(mv.n, mv.d, np, 1, elem).total, mode == MSGPU_FP64 ? "complex128" : "complex64",
                 kScratchBytes);
        set_error(msg);
        return -5;
    }
    if (nk == 0) return 0;
    if (engine_init() != 0) return -1;
    std::lock_guard<std::mutex> g(g_na_mu);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || scratch_setup_locked(dev) != 0) {
        if (last_error()[0] == '\0') set_error("msgpu_tracelog_batch_noalloc: no current device");
        return -1;
    }
    int B = std::min(kMaxChunk, nk);
    while (B > 1 && layout_of(mv.n, mv.d, np, B, elem).total > kScratchBytes) B--;
    const Layout L = layout_of(mv.n, mv.d, np, B, elem);
    try {
        const cx<double>* kc = reinterpret_cast<const cx<double>*>(ks);
        cx<double>* oc = reinterpret_cast<cx<double>*>(out);
        if (mode == MSGPU_FP64) run_map<double>(g_scratch[dev], mv, np, B, L, nk, kc, oc);
        else run_map<float>(g_scratch[dev], mv, np, B, L, nk, kc, oc);
    } catch (const std::exception& ex) {
        (void)hipDeviceSynchronize();       // nothing of this call is in flight when the lock is released
        set_error(std::string("msgpu_tracelog_batch_noalloc: ") + ex.what());
        return -2;
    }
    return 0;
}

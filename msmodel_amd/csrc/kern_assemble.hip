// kern_assemble.hip -- device assembly of the multiple-scattering matrices, the
// per-k diagonal, incident vectors and the psi grid.
//
// Layout: every matrix of a batch is Np x Np, column-major, interleaved complex
// (re, im), leading dimension Np, matrices back to back (stride Np*Np).  Np is N
// rounded up to a multiple of 64; the padding is the identity (diagonal 1,
// zeros elsewhere), which leaves the determinant, the pivot sequence of the
// first N columns and the solution of the first N unknowns unchanged.
#include "kernels.h"
#include "assemble_dev.h"

namespace msg {

// ------------------------------------------------------------------ diagonal
// kind 0: F(k)^-1; kind 1: (F^-1(k+h) - F^-1(k-h))/(2h), h = 1e-4
// (ms_matrix_util.c:68-81); kind 2: normalised matrix (diag i, izero = I_d(k,0)).
__global__ void diag_kernel(const cx<double>* ks, int nk, ScModel m, int d, int kind, cx<double>* diag,
                            cx<double>* izero) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nk) return;
    cx<double> k = ks[l];
    if (kind == 0) {
        diag[l] = invf(m, d, k);
    } else if (kind == 1) {
        const cx<double> h{1e-4, 0.0};
        diag[l] = (invf(m, d, k + h) - invf(m, d, k - h)) / (2.0 * h);
    } else {
        diag[l] = cx<double>{0.0, 1.0};
        izero[l] = green_imag(d, k, 0.0);
    }
}

void launch_diag(const cx<double>* ks, int nk, ScModel m, int d, int kind, cx<double>* diag, cx<double>* izero,
                 hipStream_t st) {
    hipLaunchKernelGGL(diag_kernel, dim3((nk + 63) / 64), dim3(64), 0, st, ks, nk, m, d, kind, diag, izero);
}

// Both diagonals of every item, each with its own scattering model and dimension.
__global__ void item_diag_kernel(MatItem* items, int cnt) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= cnt) return;
    MatItem& it = items[l];
    const cx<double> h{1e-4, 0.0};
    it.diag0 = invf(it.m, it.d, it.k);
    it.diag1 = (invf(it.m, it.d, it.k + h) - invf(it.m, it.d, it.k - h)) / (2.0 * h);
}

void launch_item_diag(MatItem* items, int cnt, hipStream_t st) {
    if (cnt > 0) hipLaunchKernelGGL(item_diag_kernel, dim3((cnt + 63) / 64), dim3(64), 0, st, items, cnt);
}

// ------------------------------------------------------------------ assembly
// One workgroup per pair of 32x32 tiles (I <= J) of one matrix (assemble_pair).
template <class T, bool D3>
__global__ __launch_bounds__(256) void assemble_kernel(const double* __restrict__ pos, int n, int d, int np,
                                                       const cx<double>* __restrict__ ks,
                                                       const cx<double>* __restrict__ diag,
                                                       const cx<double>* __restrict__ izero, int kind,
                                                       cx<T>* __restrict__ out, long long mstride,
                                                       const MatItem* __restrict__ items, int lower) {
    const int b = blockIdx.y;
    cx<double> kd, dg;
    if (items) {            // per-matrix medium (a batch over several media of one padded size)
        const MatItem& it = items[b];
        pos = it.pos;
        n = it.n;
        d = it.d;
        kd = it.k;
        dg = kind == 1 ? it.diag1 : it.diag0;
    } else {
        kd = ks[b];
        dg = diag[b];
    }
    int I, J;
    tile_pair(blockIdx.x, np / AT, I, J);
    __shared__ double sx[2][AT][4];
    __shared__ cx<T> tile[AT][AT + 1];
    cx<double> iz{1.0, 0.0};
    if (kind == 2) iz = izero[b];
    assemble_pair<T, D3>(pos, n, d, np, kd, dg, iz, kind, out + (size_t)b * mstride, I, J, sx, tile, lower != 0);
}

template <class T>
void launch_assemble(const double* pos, int n, int d, int np, const cx<double>* ks, const cx<double>* diag,
                     const cx<double>* izero, int kind, cx<T>* out, long long mstride, int nbatch, hipStream_t st) {
    const int nt = np / AT;
    dim3 grid(nt * (nt + 1) / 2, nbatch);
    if (d == 3 && kind == 0)
        hipLaunchKernelGGL((assemble_kernel<T, true>), grid, dim3(256), 0, st, pos, n, d, np, ks, diag, izero, kind,
                           out, mstride, (const MatItem*)nullptr, 0);
    else
        hipLaunchKernelGGL((assemble_kernel<T, false>), grid, dim3(256), 0, st, pos, n, d, np, ks, diag, izero, kind,
                           out, mstride, (const MatItem*)nullptr, 0);
}

template <class T>
void launch_assemble_items(const MatItem* items, int np, int kind, cx<T>* out, long long mstride, int nbatch,
                           hipStream_t st, int dall, bool lower) {
    const int nt = np / AT;
    dim3 grid(nt * (nt + 1) / 2, nbatch);
    if (dall == 3 && kind == 0)
        hipLaunchKernelGGL((assemble_kernel<T, true>), grid, dim3(256), 0, st, (const double*)nullptr, 0, 0, np,
                           (const cx<double>*)nullptr, (const cx<double>*)nullptr, (const cx<double>*)nullptr, kind,
                           out, mstride, items, lower ? 1 : 0);
    else
        hipLaunchKernelGGL((assemble_kernel<T, false>), grid, dim3(256), 0, st, (const double*)nullptr, 0, 0, np,
                           (const cx<double>*)nullptr, (const cx<double>*)nullptr, (const cx<double>*)nullptr, kind,
                           out, mstride, items, lower ? 1 : 0);
}
template void launch_assemble_items<double>(const MatItem*, int, int, cx<double>*, long long, int, hipStream_t, int,
                                            bool);
template void launch_assemble_items<float>(const MatItem*, int, int, cx<float>*, long long, int, hipStream_t, int,
                                           bool);
template void launch_assemble<double>(const double*, int, int, int, const cx<double>*, const cx<double>*,
                                      const cx<double>*, int, cx<double>*, long long, int, hipStream_t);
template void launch_assemble<float>(const double*, int, int, int, const cx<double>*, const cx<double>*,
                                     const cx<double>*, int, cx<float>*, long long, int, hipStream_t);

// ------------------------------------------------------------------ tr(M^-1 M') by contraction
// For each matrix of a batch, from V = the solution of the row-major factorisation
// against the identity (M^-1 = P V: row a of M^-1 is row inv[a] of V, inv the inverse
// of the row permutation; column-major, ld np):
//   tr(M^-1 M') = sum_{a,c} (M^-1)_{ac} M'_{ca}
//              = sum_a (M^-1)_{aa} M'_{aa} + sum_{a<c} M'_{ac} ((M^-1)_{ac} + (M^-1)_{ca})
// since M' = dM/dk is symmetric.  M' is generated pair by pair as build_dk_ms_matrix does
// (ms_matrix_util.c:68-81): -G'_d off the diagonal, the finite-difference derivative of
// F^-1 on it; no M' matrix is stored.  One workgroup per pair of 32 x 32 atom tiles
// (I <= J) of one matrix, as the assembly: one Green evaluation per unordered pair; the
// two V blocks (rows inv[I], columns J and rows inv[J], columns I) are read column by
// column (coalesced), the second through an LDS transpose.  Each workgroup writes one
// FP64 partial sum; dk_reduce_kernel adds a matrix's partial sums in a fixed order.
constexpr int DKT = 32;

template <class T>
__global__ __launch_bounds__(256) void dk_contract_kernel(const cx<T>* __restrict__ V, long long sv, int np,
                                                          const MatItem* __restrict__ items,
                                                          const int* __restrict__ inv_all,
                                                          cx<double>* __restrict__ partial) {
    __shared__ double sx[2][DKT][4];
    __shared__ int sinv[2][DKT];
    __shared__ cx<T> vt[DKT][DKT + 1];       // V(inv[J], I) transposed: vt[a - I*DKT][c - J*DKT]
    __shared__ double sre[256], sim[256];
    const int b = blockIdx.y, tid = threadIdx.x;
    const MatItem& it = items[b];
    const int n = it.n, d = it.d;
    const int nt = np / DKT;
    int t = blockIdx.x, I = 0;
    while (t >= nt - I) { t -= nt - I; I++; }
    const int J = I + t;
    const cx<T>* Vb = V + (size_t)b * sv;
    const int* inv = inv_all + (size_t)b * np;
    if (tid < 2 * DKT) {
        const int side = tid / DKT, q = tid % DKT;
        const int atom = (side == 0 ? I : J) * DKT + q;
        for (int c = 0; c < 4; c++) sx[side][q][c] = (atom < n && c < d) ? it.pos[(size_t)atom * d + c] : 0.0;
        sinv[side][q] = atom < n ? inv[atom] : 0;
    }
    __syncthreads();
    const int ii = tid % DKT, j0 = tid / DKT;          // thread: row ii of the tile, columns j0 + 8 q
    // V(inv[c], a) for c in tile J (rows), a in tile I (columns), stored transposed
#pragma unroll
    for (int q = 0; q < DKT / 8; q++) {
        const int jj = j0 + 8 * q;
        const int c = J * DKT + ii, a = I * DKT + jj;
        vt[jj][ii] = (c < n && a < n) ? Vb[(size_t)sinv[1][ii] + (size_t)a * np] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    const cx<T> k{(T)it.k.x, (T)it.k.y};
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int q = 0; q < DKT / 8; q++) {
        const int jj = j0 + 8 * q;
        const int a = I * DKT + ii, c = J * DKT + jj;
        if (a >= n || c >= n || (I == J && a > c)) continue;
        const cx<T> v1 = Vb[(size_t)sinv[0][ii] + (size_t)c * np];      // (M^-1)_{ac}
        cx<T> w, g;
        if (a == c) {
            w = v1;
            g = cx<T>{(T)it.diag1.x, (T)it.diag1.y};
        } else {
            w = v1 + vt[ii][jj];                                           // + (M^-1)_{ca}
            double s = 0.0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const double dd = sx[0][ii][e] - sx[1][jj][e];
                s += dd * dd;
            }
            g = neg_dk_green<T>(d, k, (T)sqrt(s));
        }
        re += (double)w.x * (double)g.x - (double)w.y * (double)g.y;
        im += (double)w.x * (double)g.y + (double)w.y * (double)g.x;
    }
    sre[tid] = re;
    sim[tid] = im;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) { sre[tid] += sre[tid + h]; sim[tid] += sim[tid + h]; }
        __syncthreads();
    }
    if (tid == 0) partial[(size_t)b * gridDim.x + blockIdx.x] = cx<double>{sre[0], sim[0]};
}

__global__ void dk_reduce_kernel(const cx<double>* __restrict__ partial, int nblk, int nbatch,
                                 cx<double>* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbatch) return;
    cx<double> s{0.0, 0.0};
    for (int j = 0; j < nblk; j++) s = s + partial[(size_t)b * nblk + j];
    out[b] = s;
}

// inv[b][perm[b][i]] = i
__global__ void invert_perm_kernel(const int* __restrict__ perm, int np, int* __restrict__ inv) {
    const int b = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) inv[(size_t)b * np + perm[(size_t)b * np + i]] = i;
}

size_t dk_contract_partials(int np) {
    const size_t nt = (size_t)np / DKT;
    return nt * (nt + 1) / 2;
}

template <class T>
void launch_dk_contract(const cx<T>* V, long long sv, int np, const MatItem* items, const int* perm, int* inv,
                        cx<double>* partial, cx<double>* out, int nbatch, hipStream_t st) {
    if (nbatch <= 0) return;
    hipLaunchKernelGGL(invert_perm_kernel, dim3((np + 255) / 256, nbatch), dim3(256), 0, st, perm, np, inv);
    const int nblk = (int)dk_contract_partials(np);
    hipLaunchKernelGGL(dk_contract_kernel<T>, dim3(nblk, nbatch), dim3(256), 0, st, V, sv, np, items, inv,
                       partial);
    hipLaunchKernelGGL(dk_reduce_kernel, dim3((nbatch + 63) / 64), dim3(64), 0, st, partial, nblk, nbatch, out);
}
template void launch_dk_contract<double>(const cx<double>*, long long, int, const MatItem*, const int*, int*,
                                         cx<double>*, cx<double>*, int, hipStream_t);
template void launch_dk_contract<float>(const cx<float>*, long long, int, const MatItem*, const int*, int*,
                                        cx<double>*, cx<double>*, int, hipStream_t);

// ------------------------------------------------------------------ incident vectors
// source 0: plane wave e^{ik x_0} (incident_wave_util.c:76); 1: spherical G_d(k,|x|);
// 2: build_plane_wave with angle theta (ms_matrix_util.c:88-106).  Padding rows get 0.
__global__ void incident_kernel(const double* __restrict__ pos, int n, int d, int np, cx<double> k, int source,
                                double ct, double stt, cx<double>* __restrict__ phi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    if (i >= n) { phi[i] = cx<double>{0.0, 0.0}; return; }
    const double* x = pos + (size_t)i * d;
    cx<double> v;
    if (source == 0) {
        v = cexp(cx<double>{-k.y * x[0], k.x * x[0]});
    } else if (source == 1) {
        double s = 0.0;
        for (int c = 0; c < d; c++) s += x[c] * x[c];
        v = green<double>(d, k, sqrt(s));
    } else {
        double proj = (d == 1) ? ct * x[0] : ct * x[0] + stt * x[1];
        v = cexp(cx<double>{-k.y * proj, k.x * proj});
    }
    phi[i] = v;
}

void launch_incident(const double* pos, int n, int d, int np, cx<double> k, int source, double ct, double st_,
                     cx<double>* phi, hipStream_t st) {
    hipLaunchKernelGGL(incident_kernel, dim3((np + 255) / 256), dim3(256), 0, st, pos, n, d, np, k, source, ct, st_,
                       phi);
}

// Per item of a batch: the unit plane wave along the 1x direction, e^{i k x_0}
// (ms_matrix_util.c:349-352), padding rows 0; item b's vector at phi + b np.
__global__ void plane_items_kernel(const MatItem* __restrict__ items, int np, cx<double>* __restrict__ phi) {
    const int b = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const MatItem& it = items[b];
    cx<double> v{0.0, 0.0};
    if (i < it.n) {
        const double x = it.pos[(size_t)i * it.d];
        v = cexp(cx<double>{-it.k.y * x, it.k.x * x});
    }
    phi[(size_t)b * np + i] = v;
}

void launch_plane_items(const MatItem* items, int np, cx<double>* phi, int nbatch, hipStream_t st) {
    if (nbatch > 0)
        hipLaunchKernelGGL(plane_items_kernel, dim3((np + 255) / 256, nbatch), dim3(256), 0, st, items, np, phi);
}

// ------------------------------------------------------------------ psi grid
// psi(r) = phi(r) + sum_i a_i G_d(k, |x_i - r|) at r = (xmin + ix dx, ymax - iy dy, 0, ...)
// (incident_wave_util.c:87-96, wave_function_util.c:76-81).  One thread per
// grid point; scatterers and amplitudes are staged through LDS in chunks.
constexpr int WCH = 256;

// D3: d = 3 through neg_green3_s (REALK: Im k = 0); otherwise the general Green function.
template <bool D3, bool REALK>
__global__ __launch_bounds__(256) void wave_kernel(const double* __restrict__ pos, int n, int d,
                                                   const cx<double>* __restrict__ a, cx<double> k, int source,
                                                   int nx, int ny, int row0, double xmin, double dx, double ymax,
                                                   double dy, const double* __restrict__ pts,
                                                   cx<double>* __restrict__ psi) {
    __shared__ double sp[WCH][4];
    __shared__ cx<double> sa[WCH];
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = pts ? (long long)nx : (long long)nx * ny;
    const bool live = l < tot;
    const int ix = live ? (int)(l % nx) : 0, iy = live ? row0 + (int)(l / nx) : 0;
    double pt[4] = {xmin + ix * dx, (d > 1) ? ymax - iy * dy : 0.0, 0.0, 0.0};
    if (pts && live)
        for (int c = 0; c < d; c++) pt[c] = pts[(size_t)l * d + c];
    cx<double> acc{0.0, 0.0};
    for (int base = 0; base < n; base += WCH) {
        __syncthreads();
        const int cnt = min(WCH, n - base);
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            for (int c = 0; c < 4; c++) sp[q][c] = (c < d) ? pos[(size_t)(base + q) * d + c] : 0.0;
            sa[q] = a[base + q];
        }
        __syncthreads();
        if (live) {
            if constexpr (D3) {
                for (int q = 0; q < cnt; q++) {
                    const double e0 = sp[q][0] - pt[0], e1 = sp[q][1] - pt[1], e2 = sp[q][2] - pt[2];
                    acc = acc + sa[q] * neg_green3_s<double, REALK>(k, e0 * e0 + e1 * e1 + e2 * e2);
                }
            } else {
                for (int q = 0; q < cnt; q++) {
                    double s = 0.0;
                    for (int c = 0; c < d; c++) { double e = sp[q][c] - pt[c]; s += e * e; }
                    acc = acc + sa[q] * green<double>(d, k, sqrt(s));
                }
            }
        }
    }
    if (!live) return;
    if constexpr (D3) acc = cx<double>{-acc.x, -acc.y};      // G_3 = -(e^{ikr} / (4 pi r))
    cx<double> inc;
    if (source == 0) {
        inc = cexp(cx<double>{-k.y * pt[0], k.x * pt[0]});
    } else {
        double s = 0.0;
        for (int c = 0; c < d; c++) s += pt[c] * pt[c];
        inc = green<double>(d, k, sqrt(s));
    }
    psi[l] = acc + inc;
}

// psi at a few points (the coalesced one-point calls): one workgroup per point, its
// threads splitting the scatterers, then a block reduction, so that a small batch costs a
// few Green evaluations per thread instead of N in one thread.
__global__ __launch_bounds__(256) void wave_points_kernel(const double* __restrict__ pos, int n, int d,
                                                          const cx<double>* __restrict__ a, cx<double> k,
                                                          int source, const double* __restrict__ pts,
                                                          cx<double>* __restrict__ psi) {
    __shared__ double sre[256], sim[256];
    const int p = blockIdx.x, tid = threadIdx.x;
    double pt[4];
#pragma unroll
    for (int c = 0; c < 4; c++) pt[c] = c < d ? pts[(size_t)p * d + c] : 0.0;
    cx<double> acc{0.0, 0.0};
    for (int q = tid; q < n; q += 256) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const double e = (c < d ? pos[(size_t)q * d + c] : 0.0) - pt[c];
            s += e * e;
        }
        acc = acc + a[q] * (d == 3 ? -neg_green3_s<double, false>(k, s) : green<double>(d, k, sqrt(s)));
    }
    sre[tid] = acc.x;
    sim[tid] = acc.y;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) { sre[tid] += sre[tid + h]; sim[tid] += sim[tid + h]; }
        __syncthreads();
    }
    if (tid != 0) return;
    cx<double> inc;
    if (source == 0) {
        inc = cexp(cx<double>{-k.y * pt[0], k.x * pt[0]});
    } else {
        double s = 0.0;
        for (int c = 0; c < d; c++) s += pt[c] * pt[c];
        inc = green<double>(d, k, sqrt(s));
    }
    psi[p] = cx<double>{sre[0], sim[0]} + inc;
}

void launch_wave_points(const double* pos, int n, int d, const cx<double>* a, cx<double> k, int source, int npt,
                        const double* pts, cx<double>* psi, hipStream_t st) {
    if (npt > 0)
        hipLaunchKernelGGL(wave_points_kernel, dim3(npt), dim3(256), 0, st, pos, n, d, a, k, source, pts, psi);
}

// ------------------------------------------------------------------ resident psi kernel
// Serves one-point psi requests from a ring of slots in coherent pinned host memory, so
// that a drop-in eval_total_wave call costs a round trip over the host link instead of
// a kernel launch and a stream synchronisation.  Workgroup w serves the request indices
// i = w, w + G, w + 2G, ... (G workgroups) in order, starting from resume[w]: thread 0
// polls slot i % nring until the host publishes request i (seq = i + 1), the workgroup
// sums psi at its point over the scatterers (one workgroup per point, as
// wave_points_kernel), and thread 0 writes psi, then done = i + 1 (release, system scope).
// Every workgroup leaves the loop, recording its next index in resume[w], when
//  - the host sets *stop, or
//  - no request has been served for idle_ticks (of the constant wall clock) since the
//    later of the last one served by any workgroup and this workgroup's start, or
//  - the launch has run for life_ticks;
// the first workgroup to find the launch idle or too old sets *closing, which ends the
// others at their next poll, so the launch drains as a whole and the host relaunches it
// for the requests that arrive after.  *closing and *last are reset by the host before
// each launch.
__device__ __forceinline__ unsigned long long ld_sys(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(256) void wave_resident_kernel(WaveSlot* ring, int nring, const unsigned* stop,
                                                            unsigned long long* resume, unsigned* closing,
                                                            unsigned long long* last, long long idle_ticks,
                                                            long long life_ticks) {
    __shared__ double sre[256], sim[256];
    __shared__ const double* s_pos;
    __shared__ const cx<double>* s_a;
    __shared__ double s_x[4];
    __shared__ cx<double> s_k;
    __shared__ int s_n, s_d, s_src, s_cmd;
    const int w = blockIdx.x, G = gridDim.x, tid = threadIdx.x;
    unsigned long long i = resume[w];
    const long long t0 = wall_clock64();
    for (;;) {
        if (tid == 0) {
            int cmd = 0;                                  // 1: serve request i, 2: leave
            WaveSlot* sl = ring + (i % (unsigned long long)nring);
            while (cmd == 0) {
                if (__hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) ||
                    __hip_atomic_load(closing, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    cmd = 2;
                } else if (ld_sys(&sl->seq) == i + 1) {
                    cmd = 1;
                } else {
                    const long long now = wall_clock64();
                    const long long ref = max(t0, (long long)__hip_atomic_load(last, __ATOMIC_RELAXED,
                                                                                 __HIP_MEMORY_SCOPE_AGENT));
                    if (now - ref > idle_ticks || now - t0 > life_ticks) {
                        __hip_atomic_store(closing, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        cmd = 2;
                    } else {
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
            }
            if (cmd == 1) {
                s_pos = sl->pos;
                s_a = sl->a;
                for (int c = 0; c < 4; c++) s_x[c] = sl->x[c];
                s_k = sl->k;
                s_n = sl->n;
                s_d = sl->d;
                s_src = sl->source;
            }
            s_cmd = cmd;
        }
        __syncthreads();
        if (s_cmd == 2) break;
        const int n = s_n, d = s_d;
        const cx<double> k = s_k;
        const double* __restrict__ pos = s_pos;
        const cx<double>* __restrict__ a = s_a;
        double pt[4];
#pragma unroll
        for (int c = 0; c < 4; c++) pt[c] = c < d ? s_x[c] : 0.0;
        cx<double> acc{0.0, 0.0};
        for (int q = tid; q < n; q += 256) {
            double s2 = 0.0;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const double e = (c < d ? pos[(size_t)q * d + c] : 0.0) - pt[c];
                s2 += e * e;
            }
            acc = acc + a[q] * (d == 3 ? -neg_green3_s<double, false>(k, s2) : green<double>(d, k, sqrt(s2)));
        }
        sre[tid] = acc.x;
        sim[tid] = acc.y;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h) { sre[tid] += sre[tid + h]; sim[tid] += sim[tid + h]; }
            __syncthreads();
        }
        if (tid == 0) {
            cx<double> inc;
            if (s_src == 0) {
                inc = cexp(cx<double>{-k.y * pt[0], k.x * pt[0]});
            } else {
                double r2 = 0.0;
                for (int c = 0; c < d; c++) r2 += pt[c] * pt[c];
                inc = green<double>(d, k, sqrt(r2));
            }
            WaveSlot* sl = ring + (i % (unsigned long long)nring);
            sl->psi = cx<double>{sre[0], sim[0]} + inc;
            __hip_atomic_store(&sl->done, i + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_fetch_max(last, (unsigned long long)wall_clock64(), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        i += (unsigned long long)G;
        __syncthreads();
    }
    if (tid == 0) resume[w] = i;
}

void launch_wave_resident(WaveSlot* ring, int nring, const unsigned* stop, unsigned long long* resume,
                          unsigned* closing, unsigned long long* last, long long idle_ticks, long long life_ticks,
                          int nwg, hipStream_t st) {
    hipLaunchKernelGGL(wave_resident_kernel, dim3(nwg), dim3(256), 0, st, ring, nring, stop, resume, closing, last,
                       idle_ticks, life_ticks);
}

void launch_wave(const double* pos, int n, int d, const cx<double>* a, cx<double> k, int source, int nx, int ny,
                 int row0, double xmin, double dx, double ymax, double dy, const double* pts, cx<double>* psi,
                 hipStream_t st) {
    long long tot = pts ? (long long)nx : (long long)nx * ny;
    auto kern = d != 3 ? wave_kernel<false, false> : k.y == 0.0 ? wave_kernel<true, true> : wave_kernel<true, false>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, pos, n, d, a, k, source, nx, ny,
                       row0, xmin, dx, ymax, dy, pts, psi);
}

}  // namespace msg

// lu_device.h -- device helpers shared by the LU kernels (kern_panel*.hip,
// kern_swap.hip, kern_lu.hip): MFMA operand conventions, the storage order of
// the factorisations, DPP/readlane arg-max reductions of the pivot search, and the
// in-register triangular solves (four lanes per right-hand side).
#pragma once
#include <utility>

#include "kernels.h"

namespace msg {


typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
template <class T> struct mfma;
// Operands (both types): lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Result: lane l, register r holds D[i][l%16] with i = 4*(l/16)+r for f32 and
// i = l/16 + 4*r for f64 (checked on gfx950 with an in-bounds probe).  For f64 the
// GEMM feeds MFMA row i with real row arow(i) = 4*(i%4) + i/4 of the tile, so that
// for both types the four registers of a lane are real rows 4*(l/16)..+3, i.e. two
// consecutive complex elements of one column (one 32- or 16-byte access).
template <> struct mfma<double> {
    using acc = d4v;
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int arow(int i) { return 4 * (i & 3) + (i >> 2); }
};
template <> struct mfma<float> {
    using acc = f4v;
    static __device__ __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int arow(int i) { return i; }
};

// Element (r, c) of an np x np matrix: column-major (RM = false) or row-major (RM =
// true).  The row-major factorisation of a symmetric M is the column-pivoted one of
// the same buffer read column-major: its row interchanges move contiguous rows.
template <bool RM>
__device__ __forceinline__ size_t at(int r, int c, int np) {
    return RM ? (size_t)r * np + (size_t)c : (size_t)r + (size_t)c * np;
}

// ------------------------------------------------------------------ panel
// pmap layout per matrix and panel q: [cnt, dst[2*nb], src[2*nb]] (1 + 4*nb ints): the
// panel's interchanges as one move of row src[e] to row dst[e] (rows that end where
// they started are left out).

// DPP lane exchange (all rows and banks enabled; the controls used have a source lane
// in the same row for every lane).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) { return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false); }
template <int CTRL>
__device__ __forceinline__ float dpp_v(float v) { return __int_as_float(dpp_i<CTRL>(__float_as_int(v))); }
template <int CTRL>
__device__ __forceinline__ double dpp_v(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)b), hi = dpp_i<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rdlane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double rdlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Pivot order of LAPACK izamax: the larger |Re|+|Im| first, the smaller row on ties.
template <class T>
__device__ __forceinline__ void key_max(T& v, int& i, T ov, int oi) {
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
}
template <int CTRL, class T>
__device__ __forceinline__ void key_dpp(T& v, int& i) { key_max(v, i, dpp_v<CTRL>(v), dpp_i<CTRL>(i)); }
// Largest key of each row of 16 lanes, in every lane of the row: quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror.  Called with all lanes active.
template <class T>
__device__ __forceinline__ void key_row16(T& v, int& i) {
    key_dpp<0xB1>(v, i);
    key_dpp<0x4E>(v, i);
    key_dpp<0x141>(v, i);
    key_dpp<0x140>(v, i);
}
// Largest key of the wave (wave-uniform result).
template <class T>
__device__ __forceinline__ void key_wave(T& v, int& i) {
    key_row16(v, i);
    T bv = rdlane(v, 0);
    int bi = rdlane(i, 0);
    key_max(bv, bi, rdlane(v, 16), rdlane(i, 16));
    key_max(bv, bi, rdlane(v, 32), rdlane(i, 32));
    key_max(bv, bi, rdlane(v, 48), rdlane(i, 48));
    v = bv;
    i = bi;
}

// ------------------------------------------------------------------ trace-log of a factor
// *out = (sum_i log U_ii + i pi [odd interchanges]) / n  (ms_matrix_util.c:185-202) for
// one factor M (ld np) by a 256-thread workgroup, accumulated in FP64; stats (may be
// null) = {max_i log|U_ii|, min_i log|U_ii|}; ipiv null: no interchanges (LDL^T: the
// diagonal is D).  s*: 256 entries each of shared memory.
template <class T>
__device__ __forceinline__ void tracelog_block(const cx<T>* __restrict__ M, int n, int np,
                                               const int* __restrict__ ipiv, double* sre, double* sim,
                                               double* smax, double* smin, int* sodd, cx<double>* out,
                                               double* stats) {
    const int tid = threadIdx.x;
    double re = 0.0, im = 0.0, mx = -INFINITY, mn = INFINITY;
    int odd = 0;
    for (int i = tid; i < n; i += 256) {
        const cx<T> u = M[(size_t)i + (size_t)i * np];
        const double la = log(hypot((double)u.x, (double)u.y));
        re += la;
        im += atan2((double)u.y, (double)u.x);
        mx = fmax(mx, la);
        mn = fmin(mn, la);
        if (ipiv) odd ^= (ipiv[i] != i + 1);
    }
    sre[tid] = re; sim[tid] = im; smax[tid] = mx; smin[tid] = mn; sodd[tid] = odd;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            sre[tid] += sre[tid + s];
            sim[tid] += sim[tid + s];
            smax[tid] = fmax(smax[tid], smax[tid + s]);
            smin[tid] = fmin(smin[tid], smin[tid + s]);
            sodd[tid] ^= sodd[tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        double imt = sim[0] + (sodd[0] ? kPi : 0.0);
        *out = cx<double>{sre[0] / n, imt / n};
        if (stats) { stats[0] = smax[0]; stats[1] = smin[0]; }
    }
}

// ------------------------------------------------------------------ in-register triangular solves
// Value of lane q of each quad (DPP quad_perm [q,q,q,q]); q is a constant after unrolling.
template <class T>
__device__ __forceinline__ T quad_bcast(T v, int q) {
    switch (q) {
        case 0: return dpp_v<0x00>(v);
        case 1: return dpp_v<0x55>(v);
        case 2: return dpp_v<0xAA>(v);
        default: return dpp_v<0xFF>(v);
    }
}

// L(row, i) from packed strictly-lower storage (row r at offset r (r - 1) / 2; unit
// diagonal only).
template <class T>
struct PackedL {
    const cx<T>* s;
    __device__ __forceinline__ cx<T> operator()(int r, int c) const { return s[(r * (r - 1) >> 1) + c]; }
};
// Packed triangles with their diagonal: lower (c <= r) by rows, upper (r <= c) by columns.
template <class T>
struct PackedLD {
    const cx<T>* s;
    __device__ __forceinline__ cx<T> operator()(int r, int c) const { return s[(r * (r + 1) >> 1) + c]; }
};
template <class T>
struct PackedUD {
    const cx<T>* s;
    __device__ __forceinline__ cx<T> operator()(int r, int c) const { return s[(c * (c + 1) >> 1) + r]; }
};

// Four lanes per right-hand side (a quad): lane qq of a quad holds rows 4m+qq in xv[m],
// and each solved row reaches the other lanes of its quad by a DPP broadcast.  The
// register file is rotated by one after every 4 rows, so the row being eliminated is
// always xv[0] (forward) or xv[15] (backward) and every register index is a
// compile-time constant while the 16-step loop stays rolled.  Called by whole waves.
template <class T, bool UNIT = true, class LA>
__device__ __forceinline__ void lower_unit_solve(cx<T> (&xv)[16], LA sL, int nb, int qq, int /*lane*/) {
#pragma unroll 1
    for (int ib = 0; ib < 16; ib++) {
#pragma unroll
        for (int iq = 0; iq < 4; iq++) {
            const int i = 4 * ib + iq;
            if (!UNIT && qq == iq && i < nb) xv[0] = xv[0] / sL(i, i);
            const cx<T> xi{quad_bcast(xv[0].x, iq), quad_bcast(xv[0].y, iq)};
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const int row = 4 * (m + ib) + qq;
                if (m + ib < 16 && row > i && row < nb) xv[m] = xv[m] - sL(row, i) * xi;
            }
        }
        const cx<T> t = xv[0];
#pragma unroll
        for (int m = 0; m < 15; m++) xv[m] = xv[m + 1];
        xv[15] = t;
    }
}

template <class T, bool UNIT = false, class LA>
__device__ __forceinline__ void upper_solve(cx<T> (&xv)[16], LA sL, int qq, int /*lane*/) {
#pragma unroll 1
    for (int ib = 15; ib >= 0; ib--) {
#pragma unroll
        for (int iq = 3; iq >= 0; iq--) {
            const int i = 4 * ib + iq;
            if (!UNIT && qq == iq) xv[15] = xv[15] / sL(i, i);
            const cx<T> xi{quad_bcast(xv[15].x, iq), quad_bcast(xv[15].y, iq)};
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const int row = 4 * (m - (15 - ib)) + qq;
                if (m >= 15 - ib && row < i) xv[m] = xv[m] - sL(row, i) * xi;
            }
        }
        const cx<T> t = xv[15];
#pragma unroll
        for (int m = 15; m > 0; m--) xv[m] = xv[m - 1];
        xv[0] = t;
    }
}

// x - l y in four fused multiply-adds
template <class T>
__device__ __forceinline__ cx<T> fms4(cx<T> x, cx<T> l, cx<T> y) {
    return {fma(-l.x, y.x, fma(l.y, y.y, x.x)), fma(-l.x, y.y, fma(-l.y, y.x, x.y))};
}

// The unit-diagonal forward solve of a full 64 x 64 triangle with the 64 steps unrolled
// (expanded from a parameter pack): same lanes and registers as the rolled form above.
// Every register index, step I and row offset is a compile-time constant, so an element
// sL(4m+qq, I) of the packed triangle is read at a per-lane base of register m plus a
// constant offset, with no address arithmetic per access (the rolled loop, whose register
// file rotates, issues several times the multiply-adds in instructions, mostly for those
// addresses).  Branch-free: each step updates every register's rows below I with x - l y
// in four fused multiply-adds, and the quad that straddles I is masked by l = 0 (a branch
// there, or a runtime row limit, made the compiler keep the triangle's loads in scratch).
// SB: a scheduling barrier every SB steps (1: 168 registers, three waves per SIMD, in the
// kernels that use it).  Kernels that run several solves over the same triangle keep the
// rolled form (unrolled, they spill); so do the solves with a stored diagonal.  Called by
// whole waves.
template <class T, int I, int SB, class LA>
__device__ __forceinline__ void lower_step64(cx<T> (&xv)[16], LA sL, int qq) {
    constexpr int mi = I >> 2, iq = I & 3;
    const cx<T> xi{quad_bcast(xv[mi].x, iq), quad_bcast(xv[mi].y, iq)};
    {   // the quad's rows at or above I take l = 0: x - 0 y = x, without a branch
        const cx<T> l = sL(4 * mi + qq, I);
        const bool below = qq > iq;
        xv[mi] = fms4(xv[mi], cx<T>{below ? l.x : T(0), below ? l.y : T(0)}, xi);
    }
#pragma unroll
    for (int m = mi + 1; m < 16; m++) xv[m] = fms4(xv[m], sL(4 * m + qq, I), xi);
    if ((I % SB) == SB - 1) __builtin_amdgcn_sched_barrier(0);
}

template <class T, int SB, class LA, int... Is>
__device__ __forceinline__ void lower_steps64(cx<T> (&xv)[16], LA sL, int qq, std::integer_sequence<int, Is...>) {
    (lower_step64<T, Is, SB>(xv, sL, qq), ...);
}

template <class T, int SB = 4, class LA>
__device__ __forceinline__ void lower_unit_solve64(cx<T> (&xv)[16], LA sL, int qq) {
    lower_steps64<T, SB>(xv, sL, qq, std::make_integer_sequence<int, 64>{});
}

}  // namespace msg

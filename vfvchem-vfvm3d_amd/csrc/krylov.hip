// Restarted GMRES on the device: CSR SpMV, Jacobi and multicolour ILU(0) right
// preconditioning, classical Gram-Schmidt (CGS, or CGS2 with one
// re-orthogonalisation pass), Givens least squares, FP64 and FP32-inner /
// FP64-refined variants.
//
// Replaces LinearGmres::solve_matrix (src/linear/linear_gmres.rs:76-84), which
// calls faer_gmres::restarted_gmres(A, b, x, restart, max_outer, tol, Jacobi).
// faer_gmres 0.5.0 is not vendored with the reference; this solver implements
// the textbook algorithm (Saad, Iterative Methods for Sparse Linear Systems,
// 2nd ed., Alg. 6.10 GMRES(m) with right preconditioning; Alg. 10.4 ILU(0);
// section 12.4 multicolour orderings) and stops on the FP64 relative residual
// ||b - A x||_2 / ||b||_2 <= tol, recomputed from x at every restart.
//
// Device-resident state: the Krylov basis V (column-major, n x (m+1)), the
// Hessenberg matrix, Givens rotations and the control word live in HBM; the
// host reads the control word every few steps.  Reductions are two-stage and
// fixed-order, so results are bitwise reproducible run to run.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "order.hpp"
#include "xcd.hpp"


namespace vfvm {

namespace {

constexpr int BLK = 256;
constexpr int WAVE = 64;
constexpr int MAX_RESTART = 256;

// ---------------------------------------------------------------- state layout (doubles)
struct StateLayout {
    int m;
    __host__ __device__ int64_t H() const { return 0; }                    // (m+1) x m column-major
    __host__ __device__ int64_t cs() const { return (int64_t)(m + 1) * m; }
    __host__ __device__ int64_t sn() const { return cs() + m; }
    __host__ __device__ int64_t g() const { return sn() + m; }              // m+1
    __host__ __device__ int64_t y() const { return g() + m + 1; }           // m
    __host__ __device__ int64_t h1() const { return y() + m; }              // m+1
    __host__ __device__ int64_t h2() const { return h1() + m + 1; }         // m+1
    __host__ __device__ int64_t misc() const { return h2() + m + 1; }       // [0]=beta [1]=inv_h [2]=target [3]=res_est [4]=1.0 [5]=||b||
    __host__ __device__ int64_t sig() const { return misc() + 8; }          // m+1: basis scale factors
    __host__ __device__ int64_t total() const { return sig() + m + 1; }
};
// The basis is kept unnormalised: v_k = sig_k V[k], sig_0 = 1 (V[0] = r / beta is
// normalised explicitly) and sig_(j+1) = 1 / h_(j+1,j); every consumer folds sig into its
// coefficients, which saves a scaling pass per Arnoldi step.
// control word (ints): [0] active, [1] steps done in this cycle, [2] non-finite flag
constexpr int CTRL_INTS = 4;

__device__ inline double wave_sum(double v) {
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, WAVE);
    return v;
}

// fixed-order block sum over BLK threads; result valid in thread 0
__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < BLK / WAVE; ++k) r += sh[k];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- SpMV
// LPR lanes per row (scalar rows carry ~17 entries, coupled flow rows ~100-480)
// the same sum with two entries per trip, both gathers issued before the first product
// (the lane's entries are still added in ascending order)
template <class T>
__device__ inline T row_dot2(const int32_t* __restrict__ col, const T* __restrict__ val, const T* __restrict__ x, int32_t k,
                             int32_t e1, int step) {
    T s = 0;
    for (; k + step < e1; k += 2 * step) {
        const int32_t c0 = col[k], c1 = col[k + step];
        const T a0 = val[k], a1 = val[k + step];
        const T x0 = x[c0], x1 = x[c1];
        s += a0 * x0;
        s += a1 * x1;
    }
    if (k < e1) s += val[k] * x[col[k]];
    return s;
}
// Stop test of a row kernel after convergence (ctrl[0] == 0).  The control word is loaded
// first (ctrl_go) and tested together with a condition on the row's first loads that always
// holds (e.g. rp[t] <= rp[t + 1]): the control word and every load that depends only on the
// row index are then one round trip to memory, and the column / value loads the next
// (the test alone would be waited for before the row's loads issue).
__device__ const int kCtrlGo = 1;
__device__ const double kUnitScale = 1.0;
// pointer selects instead of null-test branches, so that the loads issue without a wait
__device__ inline int ctrl_go(const int* __restrict__ ctrl) { return *(ctrl ? ctrl : &kCtrlGo); }
// rhs scale factor (1 when absent: x * 1 == x exactly)
__device__ inline double scale_of(const double* __restrict__ scale) { return *(scale ? scale : &kUnitScale); }
__device__ inline bool ctrl_stop(int go, bool loaded) { return (go == 0) & loaded; }
template <class T, int LPR>
__global__ void __launch_bounds__(BLK) k_spmv(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                              const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y,
                                              const double* __restrict__ scale, const int* __restrict__ ctrl) {
    const int64_t tid = (int64_t)xcd_block() * BLK + threadIdx.x;
    const int64_t row = tid / LPR;
    const int lane = (int)(tid % LPR);
    if (row >= n) return;
    const int go = ctrl_go(ctrl);
    const int32_t k0 = rp[row], k1 = rp[row + 1];
    const T sc = (T)scale_of(scale);
    if (ctrl_stop(go, (k0 <= k1) | (sc != sc))) return;
    T s = row_dot2<T>(col, val, x, k0 + lane, k1, LPR);
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LPR);
    if (lane == 0) y[row] = (T)(s * sc);
}

// SpMV over chunks of SPMV_ROWS consecutive rows, SPMV_CHUNKS consecutive chunks per workgroup.
// A chunk's entries are read with STAGE independent column / value loads per thread (the
// whole chunk in one trip: STAGE * SPMV_BLK >= emax), the x entries gathered, the products
// kept in LDS and each row then sums its products in column order.  The next chunk's column
// and value loads are issued before the current chunk's products are stored and summed, so
// the HBM stream of the workgroup does not stop during the gathers, the LDS phase and the
// barriers (one chunk per workgroup leaves it idle there: 22% slower FP32, 17% FP64 at
// 16 entries per row).  emax: most entries of a chunk (the LDS capacity).
constexpr int SPMV_ROWS = 128;
constexpr int SPMV_BLK = 256;
constexpr int SPMV_CHUNKS = 2;
template <class T, int STAGE>
__global__ void __launch_bounds__(SPMV_BLK) k_spmv_chunk(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                         const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y,
                                                         const double* __restrict__ scale, const int* __restrict__ ctrl,
                                                         int32_t emax) {
    extern __shared__ __align__(16) unsigned char lds[];
    T* prod = reinterpret_cast<T*>(lds);
    int32_t* lrp = reinterpret_cast<int32_t*>(prod + emax);
    const int64_t nchunk = (n + SPMV_ROWS - 1) / SPMV_ROWS;
    const int64_t c0 = (int64_t)xcd_block() * SPMV_CHUNKS;
    const int go = ctrl_go(ctrl);
    const T sc = (T)scale_of(scale);
    int32_t j[STAGE];
    T a[STAGE];
    int32_t base = 0, ne = 0, nr = 0, myrp = 0;  // myrp: rp of this thread's row of the chunk
    auto load = [&](int64_t c) {
        const int64_t r0 = c * SPMV_ROWS, r1 = r0 + SPMV_ROWS < n ? r0 + SPMV_ROWS : n;
        nr = (int)(r1 - r0);
        base = rp[r0];
        ne = rp[r1] - base;
        if ((int)threadIdx.x <= nr) myrp = rp[r0 + threadIdx.x];
#pragma unroll
        for (int u = 0; u < STAGE; ++u) {
            const int e = u * SPMV_BLK + threadIdx.x;
            j[u] = e < ne ? col[base + e] : 0;
            a[u] = e < ne ? val[base + e] : (T)0;
        }
    };
    load(c0);
    if (ctrl_stop(go, (base <= base + ne) | (sc != sc))) return;
    for (int k = 0; k < SPMV_CHUNKS; ++k) {
        const int64_t c = c0 + k;
        if (c >= nchunk) break;
        T xv[STAGE];
#pragma unroll
        for (int u = 0; u < STAGE; ++u) xv[u] = x[j[u]];
        T pr[STAGE];
#pragma unroll
        for (int u = 0; u < STAGE; ++u) pr[u] = a[u] * xv[u];
        const int cne = ne, cnr = nr;
        const int32_t crp = myrp - base;
        if (k + 1 < SPMV_CHUNKS && c + 1 < nchunk) load(c + 1);
        if ((int)threadIdx.x <= cnr) lrp[threadIdx.x] = crp;
#pragma unroll
        for (int u = 0; u < STAGE; ++u) {
            const int e = u * SPMV_BLK + threadIdx.x;
            if (e < cne) prod[e] = pr[u];
        }
        __syncthreads();
        if ((int)threadIdx.x < cnr) {
            const int q1 = lrp[threadIdx.x + 1];
            T s = 0;
#pragma unroll 4
            for (int q = lrp[threadIdx.x]; q < q1; ++q) s += prod[q];
            y[c * SPMV_ROWS + threadIdx.x] = (T)(s * sc);
        }
        __syncthreads();
    }
}
// most entries of a chunk of SPMV_ROWS consecutive rows of the CSR row pointers rp
inline int32_t spmv_block_emax(const std::vector<int32_t>& rp) {
    const int64_t n = (int64_t)rp.size() - 1;
    int32_t m = 0;
    for (int64_t r0 = 0; r0 < n; r0 += SPMV_ROWS) {
        const int64_t r1 = std::min<int64_t>(n, r0 + SPMV_ROWS);
        m = std::max(m, rp[r1] - rp[r0]);
    }
    return m;
}
// LDS of a k_spmv_chunk block (at most 12 * SPMV_BLK products: 25 KB in FP64)
inline size_t spmv_chunk_lds(int vb, int64_t emax) { return (size_t)(vb * emax + 4 * (SPMV_ROWS + 1)); }

// SpMV of the coupled flow system [u, v, w, p] (rows and columns comp-major), one wave per
// vertex v computing its four rows together.  The three velocity rows of a point share their
// column sets ({v} u ngh(v) in their own component, grad_coeff(v) in p) and the continuity
// row reads the velocity components at grad_coeff(v) and p on the 3-ring, so one pass over
// grad_coeff(v) gathers p[m] for the three velocity rows and u[m], v[m], w[m] for the
// continuity row: the column indices are read once per vertex instead of once per entry
// (the u row's own block, its p block and the continuity row's 3-ring), and every gathered
// x entry feeds up to three products.  Rows of length 1 (velocity Dirichlet rows, the
// reference-pressure row) are diagonal.
struct FlowSpmvGeom {
    int64_t N, NP;
    const int32_t* rp;      // vertex adjacency (degree of v)
    const int32_t* gc_ptr;  // grad_coeff list lengths
};
template <class T, int L = WAVE>
__device__ inline T wave_sum(T s) {
    for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
    return s;
}

// SpMV of a generic CSR numbered vertex-interleaved (rows S v .. S v + S - 1 of vertex v
// contiguous in the CSR, e.g. the transposed coupled flow system): one wave per vertex streams
// the vertex's whole entry range [rp[S v], rp[S v + S]) with coalesced column / value loads
// (each lane every 64th entry, two in flight) and adds each product to the accumulator of its
// row; the S sums are then reduced across the wave in a fixed butterfly order.  Rows of very
// different lengths (velocity, pressure and scalar rows of one vertex) share one wave instead
// of each leaving lanes idle.
template <class T, int S>
__global__ void __launch_bounds__(BLK) k_spmv_vblock(int64_t nv, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                     const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y,
                                                     const double* __restrict__ scale, const int* __restrict__ ctrl) {
    const int64_t tid = (int64_t)xcd_block() * BLK + threadIdx.x;
    const int64_t v = tid / WAVE;
    const int lane = (int)(tid % WAVE);
    if (v >= nv) return;
    const int go = ctrl_go(ctrl);
    int32_t b[S + 1];
#pragma unroll
    for (int c = 0; c <= S; ++c) b[c] = rp[S * v + c];
    const T sc = (T)scale_of(scale);
    if (ctrl_stop(go, (b[0] <= b[S]) | (sc != sc))) return;
    T acc[S];
#pragma unroll
    for (int c = 0; c < S; ++c) acc[c] = 0;
    int32_t k = b[0] + lane;
    // eight entries in flight per lane (all column, value and x loads issued before the first
    // product), then two; each lane still adds its entries in ascending order
    constexpr int U = 8;
    for (; k + (U - 1) * WAVE < b[S]; k += U * WAVE) {
        int32_t cu[U];
        T au[U], xu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            cu[u] = col[k + u * WAVE];
            au[u] = val[k + u * WAVE];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xu[u] = x[cu[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const T pu = au[u] * xu[u];
            const int32_t ku = k + u * WAVE;
#pragma unroll
            for (int c = 0; c < S; ++c) acc[c] += (ku >= b[c] && ku < b[c + 1]) ? pu : (T)0;
        }
    }
    for (; k + WAVE < b[S]; k += 2 * WAVE) {
        const int32_t c0 = col[k], c1 = col[k + WAVE];
        const T a0 = val[k], a1 = val[k + WAVE];
        const T p0 = a0 * x[c0], p1 = a1 * x[c1];
#pragma unroll
        for (int c = 0; c < S; ++c) {
            acc[c] += (k >= b[c] && k < b[c + 1]) ? p0 : (T)0;
            acc[c] += (k + WAVE >= b[c] && k + WAVE < b[c + 1]) ? p1 : (T)0;
        }
    }
    if (k < b[S]) {
        const T p0 = val[k] * x[col[k]];
#pragma unroll
        for (int c = 0; c < S; ++c) acc[c] += (k >= b[c] && k < b[c + 1]) ? p0 : (T)0;
    }
#pragma unroll
    for (int c = 0; c < S; ++c) acc[c] = wave_sum<T>(acc[c]);
    T out = acc[0];
#pragma unroll
    for (int c = 1; c < S; ++c) out = lane == c ? acc[c] : out;
    if (lane < S) y[S * v + lane] = (T)(out * sc);
}

// position of unknown (component c, vertex m) in a system of S unknowns per vertex:
// comp-major c N + m, or vertex-interleaved S m + c
template <bool IL, int S>
__device__ inline int64_t fidx(int64_t N, int c, int64_t m) { return IL ? (int64_t)S * m + c : c * N + m; }

// the row sums of vertex v (partials of its L lanes, not yet reduced): the four flow rows, and
// the rows of the NS coupled scalar unknowns (rows (4 + k) N + v, columns in their own block:
// {v} u ngh(v), or {v} alone)
template <class T, bool IL, int L = WAVE, int NS = 0, int UR = 4, bool UG = true>
__device__ inline void flow_rows(const FlowSpmvGeom& fg, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                 const T* __restrict__ val, const T* __restrict__ x, int64_t v, int lane, T& su, T& sv,
                                 T& sw, T& sq, T* st) {
    constexpr int S = 4 + NS;
    const int64_t N = fg.N;
    const int32_t pN3 = (int32_t)(3 * N);
    const int32_t ru = rp[v], rv = rp[N + v], rw = rp[2 * N + v], rq = rp[3 * N + v];
    const int32_t lu = rp[v + 1] - ru, lq = rp[3 * N + v + 1] - rq;
    su = 0; sv = 0; sw = 0; sq = 0;
    const int32_t G = fg.gc_ptr[v + 1] - fg.gc_ptr[v];
    if (lu > 1) {
        // own-component part of the velocity rows: {v} u ngh(v)
        const int32_t d1 = fg.rp[v + 1] - fg.rp[v] + 1;
        for (int32_t k = lane; k < d1; k += L) {
            const int32_t m = col[ru + k];
            su += val[ru + k] * x[fidx<IL, S>(N, 0, m)];
            sv += val[rv + k] * x[fidx<IL, S>(N, 1, m)];
            sw += val[rw + k] * x[fidx<IL, S>(N, 2, m)];
        }
        // p part of the velocity rows and velocity part of the continuity row: grad_coeff(v)
        // (a point's continuity row is always full: the reference-pressure row is a node's).
        // Two entries per lane and trip, every load issued before the first product; each
        // lane still sums its entries in ascending order.
        int32_t q = lane;
        if (UG) for (; q + L < G; q += 2 * L) {
            const int32_t m0 = col[ru + d1 + q] - pN3, m1 = col[ru + d1 + q + L] - pN3;
            const T a0 = val[ru + d1 + q], a1 = val[ru + d1 + q + L];
            const T b0 = val[rv + d1 + q], b1 = val[rv + d1 + q + L];
            const T c0 = val[rw + d1 + q], c1 = val[rw + d1 + q + L];
            const T e0 = val[rq + q], e1 = val[rq + q + L];
            const T f0 = val[rq + G + q], f1 = val[rq + G + q + L];
            const T g0 = val[rq + 2 * G + q], g1 = val[rq + 2 * G + q + L];
            const T p0 = x[fidx<IL, S>(N, 3, m0)], p1 = x[fidx<IL, S>(N, 3, m1)];
            const T u0 = x[fidx<IL, S>(N, 0, m0)], u1 = x[fidx<IL, S>(N, 0, m1)];
            const T v0 = x[fidx<IL, S>(N, 1, m0)], v1 = x[fidx<IL, S>(N, 1, m1)];
            const T w0 = x[fidx<IL, S>(N, 2, m0)], w1 = x[fidx<IL, S>(N, 2, m1)];
            su += a0 * p0;
            sv += b0 * p0;
            sw += c0 * p0;
            sq += e0 * u0 + f0 * v0 + g0 * w0;
            su += a1 * p1;
            sv += b1 * p1;
            sw += c1 * p1;
            sq += e1 * u1 + f1 * v1 + g1 * w1;
        }
        for (; q < G; q += L) {
            const int32_t m = col[ru + d1 + q] - pN3;
            const T pm = x[fidx<IL, S>(N, 3, m)];
            su += val[ru + d1 + q] * pm;
            sv += val[rv + d1 + q] * pm;
            sw += val[rw + d1 + q] * pm;
            sq += val[rq + q] * x[fidx<IL, S>(N, 0, m)] + val[rq + G + q] * x[fidx<IL, S>(N, 1, m)] +
                  val[rq + 2 * G + q] * x[fidx<IL, S>(N, 2, m)];
        }
    } else {
        if (lane == 0) {
            su = val[ru] * x[fidx<IL, S>(N, 0, v)];
            sv = val[rv] * x[fidx<IL, S>(N, 1, v)];
            sw = val[rw] * x[fidx<IL, S>(N, 2, v)];
        }
        if (lq > 1)
            for (int32_t q = lane; q < G; q += L) {
                const int32_t m = col[rq + q];
                sq += val[rq + q] * x[fidx<IL, S>(N, 0, m)] + val[rq + G + q] * x[fidx<IL, S>(N, 1, m)] +
                      val[rq + 2 * G + q] * x[fidx<IL, S>(N, 2, m)];
            }
    }
    if (lq > 1) {
        // p part of the continuity row (the 3-ring): four entries per lane and trip, loads
        // first, products summed in ascending order
        int32_t i = 3 * G + lane;
        if (UR > 1) for (; i + (UR - 1) * L < lq; i += UR * L) {
            int32_t cc[UR];
            T vv[UR], xx[UR];
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                cc[u] = col[rq + i + u * L];
                vv[u] = val[rq + i + u * L];
            }
#pragma unroll
            for (int u = 0; u < UR; ++u) xx[u] = x[fidx<IL, S>(N, 3, cc[u] - pN3)];
#pragma unroll
            for (int u = 0; u < UR; ++u) sq += vv[u] * xx[u];
        }
        for (; i < lq; i += L) sq += val[rq + i] * x[fidx<IL, S>(N, 3, col[rq + i] - pN3)];
    } else if (lane == 0) {
        sq = val[rq] * x[fidx<IL, S>(N, 3, v)];
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int64_t row = (int64_t)(4 + k) * N + v;
        const int32_t rs = rp[row], ls = rp[row + 1] - rs;
        const int32_t off = (int32_t)((4 + k) * N);
        T s = 0;
        for (int32_t i = lane; i < ls; i += L) s += val[rs + i] * x[fidx<IL, S>(N, 4 + k, col[rs + i] - off)];
        st[k] = s;
    }
}

// SpMV of the coupled flow system [u, v, w, p], one wave per vertex v computing its four rows
// together.  The three velocity rows of a point share their column sets ({v} u ngh(v) in
// their own component, grad_coeff(v) in p) and the continuity row reads the velocity
// components at grad_coeff(v) and p on the 3-ring, so one pass over grad_coeff(v) gathers
// p[m] for the three velocity rows and u[m], v[m], w[m] for the continuity row: the column
// indices are read once per vertex instead of once per entry (the u row's own block, its p
// block and the continuity row's 3-ring), and every gathered x entry feeds up to three
// products.  With IL the vectors are vertex-interleaved, so the four unknowns of a gathered
// vertex share one 32-byte (FP64) segment.  Rows of length 1 (velocity Dirichlet rows, the
// reference-pressure row) are diagonal.
template <class T, bool IL, int L, int NS, int UR = 4, bool UG = true>
__global__ void __launch_bounds__(BLK) k_spmv_flow(FlowSpmvGeom fg, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                   const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y,
                                                   const double* __restrict__ scale, const int* __restrict__ ctrl) {
    constexpr int S = 4 + NS;
    if (ctrl && ctrl[0] == 0) return;
    const int64_t N = fg.N;
    const int64_t v = ((int64_t)xcd_block() * BLK + threadIdx.x) / L;
    const int lane = threadIdx.x % L;
    if (v >= N) return;
    T su, sv, sw, sq, st[NS + 1];
    flow_rows<T, IL, L, NS, UR, UG>(fg, rp, col, val, x, v, lane, su, sv, sw, sq, st);
    su = wave_sum<T, L>(su);
    sv = wave_sum<T, L>(sv);
    sw = wave_sum<T, L>(sw);
    sq = wave_sum<T, L>(sq);
#pragma unroll
    for (int k = 0; k < NS; ++k) st[k] = wave_sum<T, L>(st[k]);
    if (lane == 0) {
        const T sc = scale ? (T)scale[0] : (T)1;
        y[fidx<IL, S>(N, 0, v)] = su * sc;
        y[fidx<IL, S>(N, 1, v)] = sv * sc;
        y[fidx<IL, S>(N, 2, v)] = sw * sc;
        y[fidx<IL, S>(N, 3, v)] = sq * sc;
#pragma unroll
        for (int k = 0; k < NS; ++k) y[fidx<IL, S>(N, 4 + k, v)] = st[k] * sc;
    }
}

// r = b - A x (FP64) of the flow system, waves striding over the vertices, and per-block
// partial sums of r^2 and b^2
template <bool IL, int NS>
__global__ void __launch_bounds__(BLK) k_residual_flow(FlowSpmvGeom fg, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ col, const double* __restrict__ val,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       double* __restrict__ r, double* __restrict__ part, int nblk) {
    constexpr int S = 4 + NS;
    __shared__ double sh[BLK / WAVE];
    const int64_t N = fg.N;
    const int lane = threadIdx.x % WAVE;
    const int64_t waves = (int64_t)gridDim.x * (BLK / WAVE);
    double rr = 0.0, bb = 0.0;
    for (int64_t v = ((int64_t)xcd_block() * BLK + threadIdx.x) / WAVE; v < N; v += waves) {
        double s[S];
        flow_rows<double, IL, WAVE, NS>(fg, rp, col, val, x, v, lane, s[0], s[1], s[2], s[3], s + 4);
#pragma unroll
        for (int c = 0; c < S; ++c) s[c] = wave_sum(s[c]);
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < S; ++c) {
                const int64_t i = fidx<IL, S>(N, c, v);
                const double bi = b[i];
                const double ri = bi - s[c];
                r[i] = ri;
                rr += ri * ri;
                bb += bi * bi;
            }
    }
    double t1 = block_sum(rr, sh);
    double t2 = block_sum(bb, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = t1; part[nblk + blockIdx.x] = t2; }
}

// comp-major <-> vertex-interleaved copies of an n = 4N vector
template <class T>
__global__ void __launch_bounds__(BLK) k_to_interleaved(int64_t N, int S, const T* __restrict__ in, T* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < S * N) out[S * (i % N) + i / N] = in[i];
}
template <class T>
__global__ void __launch_bounds__(BLK) k_from_interleaved(int64_t N, int S, const T* __restrict__ in, T* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < S * N) out[i] = in[S * (i % N) + i / N];
}

// r = b - A x (FP64) and per-block partial sums of r^2 and b^2 (grid-stride over rows)
template <int LPR>
__global__ void __launch_bounds__(BLK) k_residual(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                  const double* __restrict__ val, const double* __restrict__ x,
                                                  const double* __restrict__ b, double* __restrict__ r,
                                                  double* __restrict__ part, int nblk) {
    __shared__ double sh[BLK / WAVE];
    const int lane = (int)(threadIdx.x % LPR);
    const int64_t rows_per_pass = (int64_t)gridDim.x * (BLK / LPR);
    double rr = 0.0, bb = 0.0;
    for (int64_t row = ((int64_t)xcd_block() * BLK + threadIdx.x) / LPR; row < n; row += rows_per_pass) {
        double s = row_dot2<double>(col, val, x, rp[row] + lane, rp[row + 1], LPR);
        for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LPR);
        if (lane == 0) {
            const double bi = b[row];
            const double ri = bi - s;
            r[row] = ri;
            rr += ri * ri;
            bb += bi * bi;
        }
    }
    double t1 = block_sum(rr, sh);
    double t2 = block_sum(bb, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = t1; part[nblk + blockIdx.x] = t2; }
}

// r = b (the residual of x = 0) and the partial sums of r^2 and b^2 in k_residual's layout
__global__ void __launch_bounds__(BLK) k_residual_zero(int64_t n, const double* __restrict__ b, double* __restrict__ r,
                                                       double* __restrict__ part, int nblk) {
    __shared__ double sh[BLK / WAVE];
    double rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
        const double bi = b[i];
        r[i] = bi;
        rr += bi * bi;
    }
    const double t = block_sum(rr, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = t; part[nblk + blockIdx.x] = t; }
}

// 16-byte vector of T (4 floats or 2 doubles) for the basis-vector streams
template <class T> struct Vec16;
template <> struct Vec16<float> { typedef float4 type; static constexpr int N = 4; };
template <> struct Vec16<double> { typedef double2 type; static constexpr int N = 2; };
template <class T> __device__ inline T vget(const typename Vec16<T>::type& v, int e);
template <> __device__ inline float vget<float>(const float4& v, int e) { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }
template <> __device__ inline double vget<double>(const double2& v, int e) { return e == 0 ? v.x : v.y; }

// ---------------------------------------------------------------- multi-dot: part[k][blk] = V[k] . w over a block's rows
// 2-D grid: x = row ranges (BLK / WAVE waves per block, one range per wave, a multiple of
// the vector width), y = groups of CH basis vectors.  Each lane streams 16-byte vectors
// of w and of CH basis vectors (ldv a multiple of the vector width) and keeps CH dot
// products in registers; the block's waves are summed in LDS into one partial per vector.
template <class T, int CH>
__global__ void __launch_bounds__(BLK) k_mdot(int64_t n, int nv, const T* __restrict__ V, int64_t ldv,
                                              const T* __restrict__ w, double* __restrict__ part, int64_t rows_per_wave,
                                              const int* __restrict__ ctrl) {
    if (ctrl[0] == 0) return;
    typedef typename Vec16<T>::type VT;
    constexpr int VW = Vec16<T>::N;
    __shared__ double wsum[BLK / WAVE][CH];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    const int nbx = gridDim.x;
    const int wv = blockIdx.x * (BLK / WAVE) + wid;
    const int k0 = blockIdx.y * CH;
    const int64_t r0 = (int64_t)wv * rows_per_wave;
    const int64_t r1 = min(n, r0 + rows_per_wave);
    const int64_t rv = r1 > r0 ? r0 + (r1 - r0) / VW * VW : r0;  // end of the vector part
    const int kq = min(CH, nv - k0);
    const T* Vk = V + (int64_t)k0 * ldv;
    double acc[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) acc[q] = 0.0;
    for (int64_t i = r0 + (int64_t)lane * VW; i < rv; i += (int64_t)WAVE * VW) {
        const VT w4 = *reinterpret_cast<const VT*>(w + i);
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            if (q < kq) {
                const VT v4 = *reinterpret_cast<const VT*>(Vk + (int64_t)q * ldv + i);
#pragma unroll
                for (int e = 0; e < VW; ++e) acc[q] += (double)vget<T>(v4, e) * (double)vget<T>(w4, e);
            }
        }
    }
    for (int64_t i = rv + lane; i < r1; i += WAVE) {
        const double wi = (double)w[i];
#pragma unroll
        for (int q = 0; q < CH; ++q)
            if (q < kq) acc[q] += (double)Vk[(int64_t)q * ldv + i] * wi;
    }
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) wsum[wid][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < kq) {
        double v = 0.0;
        for (int k = 0; k < BLK / WAVE; ++k) v += wsum[k][threadIdx.x];
        part[(int64_t)(k0 + threadIdx.x) * nbx + blockIdx.x] = v;
    }
}

// out[k] = sig_k sum_blk part[k][blk]; one block per k
__global__ void __launch_bounds__(BLK) k_reduce(int nblk, const double* __restrict__ part, const double* __restrict__ sig,
                                                double* __restrict__ out, const int* __restrict__ ctrl) {
    __shared__ double sh[BLK / WAVE];
    const int k = blockIdx.x;
    const double sk = sig[k];  // issued with the partials, not after the block sum
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += BLK) s += part[(int64_t)k * nblk + i];
    if (ctrl[0] == 0) return;  // tested after the partials' loads issue: one round trip, not two
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[k] = s * sk;
}

// Arnoldi step j (one block): h = h1 (+ h2), h_(j+1,j) = ||w|| from the update's partials,
// Givens update of the column and of g, convergence test, sig_(j+1) = 1 / h_(j+1,j).  The
// column and the rotations are staged in LDS by all threads; one thread applies the j
// rotations in order.
__global__ void __launch_bounds__(BLK) k_arnoldi_step(int j, int reorth, int nblk, const double* __restrict__ part_nrm,
                                                      double* __restrict__ S, StateLayout L, int* __restrict__ ctrl) {
    __shared__ double sh[BLK / WAVE];
    __shared__ double hc[MAX_RESTART + 1], cs_s[MAX_RESTART], sn_s[MAX_RESTART];
    const int m = L.m;
    double* H = S + L.H();
    const int64_t cj = (int64_t)j * (m + 1);
    for (int i = threadIdx.x; i <= j; i += BLK) {
        hc[i] = reorth ? S[L.h1() + i] + S[L.h2() + i] : S[L.h1() + i];
        if (i < j) { cs_s[i] = S[L.cs() + i]; sn_s[i] = S[L.sn() + i]; }
    }
    double* g = S + L.g();
    // thread 0's scalars loaded with the partials: after the state stores below they could
    // not be hoisted, and each would add a memory round trip to the serial chain
    const double gj = g[j], target = S[L.misc() + 2];
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += BLK) s += part_nrm[i];
    if (ctrl[0] == 0) return;  // tested after the loads above issue (nothing is written before)
    s = block_sum(s, sh);  // contains the barrier that publishes hc/cs_s/sn_s
    if (threadIdx.x == 0) {
        const double hn = sqrt(s);
        hc[j + 1] = hn;
        // the rotated entry i + 1 is carried in a register to the next rotation (the same
        // operations as rotating hc in place)
        double a = hc[0];
        for (int i = 0; i < j; ++i) {
            const double b = hc[i + 1];
            hc[i] = cs_s[i] * a + sn_s[i] * b;
            a = -sn_s[i] * a + cs_s[i] * b;
        }
        const double b = hc[j + 1];
        const double r = hypot(a, b);
        const double c = r > 0.0 ? a / r : 1.0;
        const double sgn = r > 0.0 ? b / r : 0.0;
        S[L.cs() + j] = c;
        S[L.sn() + j] = sgn;
        hc[j] = r;
        hc[j + 1] = 0.0;
        g[j + 1] = -sgn * gj;
        g[j] = c * gj;
        const double res = fabs(sgn * gj);
        S[L.misc() + 3] = res;
        S[L.misc() + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
        S[L.sig() + j + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
        ctrl[1] = j + 1;
        const bool finite = isfinite(res) && isfinite(hn) && isfinite(r);
        if (!finite) ctrl[2] = 1;
        if (!finite || res <= target || hn == 0.0 || r == 0.0) ctrl[0] = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= j + 1; i += BLK) H[cj + i] = hc[i];
}

// wout = win - sum_k V[k] (sig_k h[k]) (wout may alias win), in 16-byte vectors over the
// block's rows (a multiple of the vector width); optional per-block partial of ||wout||^2
template <class T>
__global__ void __launch_bounds__(BLK) k_update(int64_t n, int nv, const T* __restrict__ V, int64_t ldv,
                                                const double* __restrict__ h, const double* __restrict__ sig, const T* win,
                                                T* wout, double* __restrict__ part_nrm, int64_t rows_per_blk,
                                                const int* __restrict__ ctrl) {
    typedef typename Vec16<T>::type VT;
    constexpr int VW = Vec16<T>::N;
    __shared__ double hs[MAX_RESTART + 1];
    __shared__ double sh[BLK / WAVE];
    const int go = ctrl[0];  // waited for with the coefficients' loads
    for (int k = threadIdx.x; k < nv; k += BLK) hs[k] = h[k] * sig[k];
    if (go == 0) return;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_blk;
    const int64_t r1 = min(n, r0 + rows_per_blk);
    const int64_t rv = r1 > r0 ? r0 + (r1 - r0) / VW * VW : r0;
    double nrm = 0.0;
    for (int64_t i = r0 + (int64_t)threadIdx.x * VW; i < rv; i += (int64_t)BLK * VW) {
        const VT w4 = *reinterpret_cast<const VT*>(win + i);
        double sv[VW];
#pragma unroll
        for (int e = 0; e < VW; ++e) sv[e] = (double)vget<T>(w4, e);
        for (int k = 0; k < nv; ++k) {
            const VT v4 = *reinterpret_cast<const VT*>(V + (int64_t)k * ldv + i);
            const double hk = hs[k];
#pragma unroll
            for (int e = 0; e < VW; ++e) sv[e] -= (double)vget<T>(v4, e) * hk;
        }
        T out[VW];
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            out[e] = (T)sv[e];
            nrm += (double)out[e] * (double)out[e];
        }
        *reinterpret_cast<VT*>(wout + i) = *reinterpret_cast<const VT*>(out);
    }
    for (int64_t i = rv + threadIdx.x; i < r1; i += BLK) {
        double sv = (double)win[i];
        for (int k = 0; k < nv; ++k) sv -= (double)V[(int64_t)k * ldv + i] * hs[k];
        const T st = (T)sv;
        wout[i] = st;
        nrm += (double)st * (double)st;
    }
    if (part_nrm) {
        double t = block_sum(nrm, sh);
        if (threadIdx.x == 0) part_nrm[blockIdx.x] = t;
    }
}

// dst = src * S[idx]
template <class TS, class TD>
__global__ void __launch_bounds__(BLK) k_scale(int64_t n, const TS* __restrict__ src, TD* __restrict__ dst,
                                               const double* __restrict__ S, int64_t idx, const int* __restrict__ ctrl) {
    if (ctrl && ctrl[0] == 0) return;
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    dst[i] = (TD)((double)src[i] * S[idx]);
}

// start of a cycle: beta = ||r|| from residual partials (or the stored beta);
// g = [beta, 0, ...]; inv = 1/beta; target of the Arnoldi estimate
__global__ void __launch_bounds__(BLK) k_cycle_start(int nblk, const double* __restrict__ part, double* __restrict__ S,
                                                     StateLayout L, int* __restrict__ ctrl, double target_abs,
                                                     int use_given_beta) {
    __shared__ double sh[BLK / WAVE];
    double s = 0.0, sb = 0.0;
    if (!use_given_beta) {
        for (int i = threadIdx.x; i < nblk; i += BLK) { s += part[i]; sb += part[nblk + i]; }
    }
    s = block_sum(s, sh);
    sb = block_sum(sb, sh);
    if (threadIdx.x == 0) {
        double beta = use_given_beta ? S[L.misc() + 0] : sqrt(s);
        if (!use_given_beta) S[L.misc() + 5] = sqrt(sb);  // ||b||
        S[L.misc() + 0] = beta;
        S[L.misc() + 1] = beta > 0.0 ? 1.0 / beta : 0.0;
        S[L.misc() + 2] = target_abs;
        S[L.misc() + 3] = beta;
        S[L.misc() + 4] = 1.0;
        for (int i = 0; i <= L.m; ++i) S[L.g() + i] = 0.0;
        S[L.g()] = beta;
        S[L.sig()] = 1.0;
        ctrl[0] = (beta > 0.0 && isfinite(beta)) ? 1 : 0;
        ctrl[1] = 0;
        ctrl[2] = isfinite(beta) ? 0 : 1;
    }
}

// y = H(1:k,1:k) \ g(1:k), k = steps done: one wave, H staged in LDS, each row's inner
// product summed across the lanes in a fixed order
__global__ void __launch_bounds__(WAVE) k_backsolve(double* __restrict__ S, StateLayout L, const int* __restrict__ ctrl) {
    __shared__ double Hs[64 * 64];
    __shared__ double ys[MAX_RESTART], gs[MAX_RESTART], ds[MAX_RESTART];
    const int k = ctrl[1];
    const int m = L.m;
    const int lane = threadIdx.x;
    const double* H = S + L.H();
    const double* g = S + L.g();
    // columns 0..k-1, rows 0..k-1 (k <= 64 staged; longer cycles read H directly)
    const bool staged = k <= 64;
    if (staged)
        for (int idx = lane; idx < k * k; idx += WAVE) {
            const int l = idx / k, i = idx % k;
            Hs[l * 64 + i] = H[(int64_t)l * (m + 1) + i];
        }
    // g and the diagonal staged too: a global load per row would sit on the serial chain
    for (int i = lane; i < k; i += WAVE) {
        gs[i] = g[i];
        ds[i] = H[(int64_t)i * (m + 1) + i];
    }
    __syncthreads();
    for (int i = k - 1; i >= 0; --i) {
        double p = 0.0;
        for (int l = i + 1 + lane; l < k; l += WAVE)
            p += (staged ? Hs[l * 64 + i] : H[(int64_t)l * (m + 1) + i]) * ys[l];
        p = wave_sum(p);
        if (lane == 0) {
            const double d = ds[i];
            ys[i] = d != 0.0 ? (gs[i] - p) / d : 0.0;
        }
        __syncthreads();
    }
    for (int i = lane; i < k; i += WAVE) S[L.y() + i] = ys[i];
}

// z = sum_{l<k} v_l y[l] = sum_l V[l] (sig_l y[l]) (z in FP64)
template <class T>
__global__ void __launch_bounds__(BLK) k_combine(int64_t n, const T* __restrict__ V, int64_t ldv,
                                                 const double* __restrict__ S, StateLayout L, const int* __restrict__ ctrl,
                                                 double* __restrict__ z) {
    __shared__ double ys[MAX_RESTART];
    const int k = ctrl[1];
    for (int l = threadIdx.x; l < k; l += BLK) ys[l] = S[L.y() + l] * S[L.sig() + l];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int l = 0; l < k; ++l) s += (double)V[(int64_t)l * ldv + i] * ys[l];
    z[i] = s;
}

// x += dinv .* z (dinv == nullptr: x += z)
__global__ void __launch_bounds__(BLK) k_xupdate(int64_t n, const double* __restrict__ z, const double* __restrict__ dinv,
                                                 double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    x[i] += dinv ? dinv[i] * z[i] : z[i];
}

// Jacobi: dinv = 1 / diag(A) (1 where the diagonal is absent or zero)
__global__ void __launch_bounds__(BLK) k_dinv(int64_t n, const int32_t* __restrict__ dslot, const double* __restrict__ val,
                                              double* __restrict__ dinv, int jacobi) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    double d = (jacobi && dslot[i] >= 0) ? val[dslot[i]] : 1.0;
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
}

// right-scaled matrix A D^-1 in the working precision
template <class T>
__global__ void __launch_bounds__(BLK) k_scale_cols(int64_t nnz, const int32_t* __restrict__ col, const double* __restrict__ val,
                                                    const double* __restrict__ dinv, T* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k >= nnz) return;
    out[k] = (T)(val[k] * dinv[col[k]]);
}

// ---------------------------------------------------------------- multicolour ILU(0)
// Rows are ordered by sigma(i) = (colour(i), i) and stored colour-major: permuted row t
// is row perm[t] of A, its entries are [L | diag | U] with permuted column indices in
// ascending order.  No two rows of one colour share an entry, so the rows of a colour
// are independent in the factorisation and in both triangular sweeps.
template <class T>
__global__ void __launch_bounds__(BLK) k_ilu_gather(int64_t nnz, const int32_t* __restrict__ src, const double* __restrict__ val,
                                                    T* __restrict__ lu) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k >= nnz) return;
    lu[k] = (T)val[src[k]];
}

// IKJ ILU(0) (Saad Alg. 10.4) for the permuted rows [t0, t1) of one colour
template <class T>
__global__ void __launch_bounds__(BLK) k_ilu_factor(int32_t t0, int32_t t1, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ col, const int32_t* __restrict__ diag,
                                                    T* __restrict__ lu) {
    const int32_t t = t0 + blockIdx.x * BLK + threadIdx.x;
    if (t >= t1) return;
    const int32_t e = rp[t + 1], dt = diag[t];
    for (int32_t kk = rp[t]; kk < dt; ++kk) {
        const int32_t k = col[kk];
        const T lik = lu[kk] / lu[diag[k]];
        lu[kk] = lik;
        // a_tj -= l_tk u_kj for j > k, when (t, j) is in the pattern
        int32_t p = kk + 1;
        for (int32_t q = diag[k] + 1; q < rp[k + 1]; ++q) {
            const int32_t j = col[q];
            while (p < e && col[p] < j) ++p;
            if (p < e && col[p] == j) lu[p] -= lik * lu[q];
        }
    }
}

// The factor in the split layout [L (nnz_l) | U (nnz_u) | diag (n)] from the colour-major
// [L | diag | U] rows it was computed in.
template <class T>
__global__ void __launch_bounds__(BLK) k_ilu_split(int64_t nnz_l, int64_t nnz_u, int64_t n, const int32_t* __restrict__ lpos,
                                                   const int32_t* __restrict__ upos, const int32_t* __restrict__ diag,
                                                   const T* __restrict__ full, T* __restrict__ split) {
    const int64_t k = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (k < nnz_l) split[k] = full[lpos[k]];
    else if (k < nnz_l + nnz_u) split[k] = full[upos[k - nnz_l]];
    else if (k < nnz_l + nnz_u + n) split[k] = full[diag[k - nnz_l - nnz_u]];
}

// Forward sweep over the rows [t0, t1) of one colour of the colour-major system (L has a
// unit diagonal), in place in z:  z_t = c rhs_t - sum_L l_tj z_j  (c = *rscale or 1).  The L
// entries of row t are [lptr[t], lptr[t+1]) of (lcol, lval): the sweep streams only its own
// triangle.  The last colour has no U entries: with FINAL its backward sweep is folded in
// (division by the diagonal u_tt).
template <class T, int LPR, bool FINAL>
__global__ void __launch_bounds__(BLK) k_ilu_lower(int32_t t0, int32_t t1, const int32_t* __restrict__ lptr,
                                                   const int32_t* __restrict__ lcol, const T* __restrict__ lval,
                                                   const T* __restrict__ dval, const T* __restrict__ rhs,
                                                   const double* __restrict__ rscale, T* __restrict__ z,
                                                   const int* __restrict__ ctrl) {
    const int64_t tid = (int64_t)xcd_block() * BLK + threadIdx.x;
    const int32_t t = t0 + (int32_t)(tid / LPR);
    const int lane = (int)(tid % LPR);
    if (t >= t1) return;
    const int go = ctrl_go(ctrl);
    const int32_t k0 = lptr[t], k1 = lptr[t + 1];
    const T r = (T)(rhs[t] * (T)scale_of(rscale));
    if (ctrl_stop(go, (k0 <= k1) | (r != r))) return;
    const T dd = FINAL ? dval[t] : (T)1;  // issued with the row's entries
    T s = row_dot2<T>(lcol, lval, z, k0 + lane, k1, LPR);
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LPR);
    if (lane == 0) z[t] = FINAL ? (r - s) / dd : r - s;
}
// The first two colours' forward sweeps in one launch.  Colour 0 = rows [0, t0) has no L
// entries (z_t = c rhs_t), and every L entry of colour 1 = rows [t0, t1) lies in colour 0,
// so colour 1 reads c rhs_j where the colour-by-colour sweep reads z_j: the same value, so
// the result is bitwise that of the two launches.  The colour-0 rows are written by a
// grid-stride copy over the launch's threads.  (c = 1 without a scale: x * 1 == x.)
template <class T, int LPR, bool FINAL>
__global__ void __launch_bounds__(BLK) k_ilu_lower01(int32_t t0, int32_t t1, const int32_t* __restrict__ lptr,
                                                     const int32_t* __restrict__ lcol, const T* __restrict__ lval,
                                                     const T* __restrict__ dval, const T* __restrict__ rhs,
                                                     const double* __restrict__ rscale, T* __restrict__ z,
                                                     const int* __restrict__ ctrl) {
    const int64_t tid = (int64_t)xcd_block() * BLK + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * BLK;
    const T c = (T)scale_of(rscale);
    const int32_t t = t0 + (int32_t)(tid / LPR);
    const int lane = (int)(tid % LPR);
    const bool has_row = t < t1;
    const int32_t tc = has_row ? t : t0;  // threads past the colour read row t0 (in bounds: t0 < t1)
    const int go = ctrl_go(ctrl);
    const int32_t k0 = lptr[tc], k1 = lptr[tc + 1];
    const T r = (T)(rhs[tc] * c);
    const bool cp = tid < t0;  // this thread's first colour-0 row, loaded with the row bounds
    const T cv = (T)(rhs[cp ? tid : 0] * c);
    if (ctrl_stop(go, (k0 <= k1) | (r != r) | (cv != cv))) return;
    if (has_row) {
        const T dd = FINAL ? dval[t] : (T)1;
        T s = 0;
        int32_t k = k0 + lane;
        for (; k + LPR < k1; k += 2 * LPR) {
            const int32_t c0 = lcol[k], c1 = lcol[k + LPR];
            const T a0 = lval[k], a1 = lval[k + LPR];
            const T x0 = (T)(rhs[c0] * c), x1 = (T)(rhs[c1] * c);
            s += a0 * x0;
            s += a1 * x1;
        }
        for (; k < k1; k += LPR) s += lval[k] * (T)(rhs[lcol[k]] * c);
        for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LPR);
        if (lane == 0) z[t] = FINAL ? (r - s) / dd : r - s;
    }
    if (cp) z[tid] = cv;
    for (int64_t i = tid + nthr; i < t0; i += nthr) z[i] = (T)(rhs[i] * c);
}
// backward sweep over one colour below the last, in place: z_t = (z_t - sum_U u_tj z_j) / u_tt
// (the U entries of row t are [uptr[t], uptr[t+1]) of (ucol, uval))
template <class T, int LPR>
__global__ void __launch_bounds__(BLK) k_ilu_upper(int32_t t0, int32_t t1, const int32_t* __restrict__ uptr,
                                                   const int32_t* __restrict__ ucol, const T* __restrict__ uval,
                                                   const T* __restrict__ dval, T* __restrict__ z, const int* __restrict__ ctrl) {
    const int64_t tid = (int64_t)xcd_block() * BLK + threadIdx.x;
    const int32_t t = t0 + (int32_t)(tid / LPR);
    const int lane = (int)(tid % LPR);
    if (t >= t1) return;
    const int go = ctrl_go(ctrl);
    const T zt = z[t];  // row t's own entry: written below by this row only
    const int32_t k0 = uptr[t], k1 = uptr[t + 1];
    if (ctrl_stop(go, (k0 <= k1) | (zt != zt))) return;
    const T dd = dval[t];  // issued with the row's entries
    T s = row_dot2<T>(ucol, uval, z, k0 + lane, k1, LPR);
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LPR);
    if (lane == 0) z[t] = (zt - s) / dd;
}

// ---- tiled order: one workgroup per tile, the tile's rows in LDS, one launch per application
// A tile's rows are contiguous (r0 .. r1) and sorted by level; tiles of one tile colour do not
// touch.  A tile's sweep stages its triangle entries in LDS (an entry whose column lies outside
// the tile, in a finished tile, is multiplied by that z at once; an entry inside keeps its value
// and local column), then walks the levels with a barrier between them (the rows of one level
// do not touch), each row summing its entries in column order, so only the tile's own z
// round-trips, through LDS; then it writes the tile's z.  LDS: T sv[emax], T zl[rmax], int16
// sc[emax], int16 lp[rmax + 1].  One thread per row (a tile holds at most TBLK rows); a thread
// loads up to TSTAGE entries, all issued before their gathers.
//
// All tile sweeps of one application run in one launch of nrole workgroups.  A workgroup takes
// a ticket from a counter that is never reset (ticket t = app * nrole + i: role i of the app-th
// launch), so tickets follow the order in which workgroups start.  Role i is the forward sweep
// of a tile (tile colours ascending), both sweeps of a tile of the last tile colour (its U
// entries all lie inside the tile), or the backward sweep of a tile (tile colours descending).
// A role waits for the completion flags of the roles it reads: the forward sweeps of adjacent
// tiles of lower colours; the backward sweeps of adjacent tiles of higher colours and its own
// tile's forward sweep (which also covers the reads of its z by the forward sweeps of those
// higher tiles).  Every such role comes earlier in the list, so its workgroup started before:
// the earliest unfinished role waits for nothing unfinished and every wait ends.  The flag of
// tile k reads 2 app + 1 after its forward sweep and 2 app + 2 after its backward sweep; a
// role polls a flag at most wait_limit times (2^17 unless set), then writes NaN into its z (the Krylov solve
// stops with a non-finite result) and counts the lapse.  The XCDs' L2 caches are not coherent
// with each other, and device-scope release / acquire fences write back / invalidate a whole
// L2 (which costs a thousand times the sweep), so the sweeps' z and the flags are written and
// read as device-scope atomics instead, which reach memory past the L2: a tile's z stores are
// complete (s_waitcnt before the workgroup barrier) before its flag is stored, and a role issues
// its z loads only after it has seen its flags.
constexpr int TBLK = 512;
constexpr int TSTAGE = 10;
constexpr int TILE_FWD = 1, TILE_BWD = 2;
static_assert(ILU_TILE_RMAX <= TBLK, "a tile's rows take one thread each");
template <class T>
struct TileChain {
    const int32_t* role;  // nrole records: r0, r1, L begin, L end, U begin, U end, tile, kind, deps begin, deps end
    const int32_t* lev;   // per tile the first row of each of the ILU_MAXLEV levels and the tile's end
    const int32_t* dep;   // (tile << 1) | 1 when only that tile's forward sweep is needed
    const int32_t *lptr, *lcol, *uptr, *ucol;
    const T *lval, *uval, *dval, *rhs;
    const double* rscale;
    T* z;
    const int* ctrl;
    unsigned long long* sync;  // [0] tickets taken, [1] lapsed waits, [2 + k] completion of tile k
    int32_t nrole, nlev, emax, rmax, wait_limit;
};
// the entries [base, base + ne) of a triangle, TSTAGE per thread (ne <= TBLK * TSTAGE)
template <class T>
__device__ inline void tile_load(const int32_t* __restrict__ col, const T* __restrict__ val, int32_t base, int32_t ne,
                                 int32_t r0, int32_t (&j)[TSTAGE], T (&v)[TSTAGE]) {
#pragma unroll
    for (int u = 0; u < TSTAGE; ++u) {
        const int e = u * TBLK + threadIdx.x;
        const bool in = e < ne;
        j[u] = in ? col[base + e] : r0;
        v[u] = in ? val[base + e] : (T)0;
    }
}
// z of another workgroup's tile, read at device scope (past this XCD's L2)
template <class T>
__device__ inline T z_load(T* z) { return __hip_atomic_load(z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// the loaded entries into LDS: an entry outside the tile as its product with z (sc = -1)
template <class T>
__device__ inline void tile_place(T* z, int32_t r0, int32_t r1, int32_t ne, bool lower, const int32_t (&j)[TSTAGE],
                                  const T (&v)[TSTAGE], T* sv, int16_t* sc) {
    T zx[TSTAGE];
#pragma unroll
    for (int u = 0; u < TSTAGE; ++u) {
        const bool outside = lower ? j[u] < r0 : j[u] >= r1;
        zx[u] = z_load<T>(z + (outside ? j[u] : r0));
    }
#pragma unroll
    for (int u = 0; u < TSTAGE; ++u) {
        const int e = u * TBLK + threadIdx.x;
        if (e < ne) {
            const bool outside = lower ? j[u] < r0 : j[u] >= r1;
            sv[e] = outside ? v[u] * zx[u] : v[u];
            sc[e] = outside ? (int16_t)-1 : (int16_t)(j[u] - r0);
        }
    }
}
// Phase 2 of a row: its staged entries outside the tile summed (a prefix of the row for the
// forward sweep, a suffix for the backward one: columns ascending) and the range of the entries
// inside returned, packed as (begin << 16) | end
template <class T>
__device__ inline int tile_row_outside(const T* sv, const int16_t* sc, const int16_t* lp, int i, bool lower, T& acc) {
    const int q0 = lp[i], q1 = lp[i + 1];
    T s = 0;
    int nin = 0;
#pragma unroll 4
    for (int q = q0; q < q1; ++q) {
        const int16_t j = sc[q];
        const T v = sv[q];
        s += j < 0 ? v : (T)0;
        nin += j >= 0;
    }
    acc = s;
    return lower ? ((q1 - nin) << 16) | q1 : (q0 << 16) | (q0 + nin);
}
// the entries inside the tile of a row (range packed as above) against the tile's z
template <class T>
__device__ inline T tile_row_inside(const T* sv, const int16_t* sc, const T* zl, int range) {
    const int q0 = range >> 16, q1 = range & 0xffff;
    T s = 0;
    int q = q0;
    for (; q + 1 < q1; q += 2) {
        const T a0 = sv[q], a1 = sv[q + 1];
        const T z0 = zl[sc[q]], z1 = zl[sc[q + 1]];
        s += a0 * z0;
        s += a1 * z1;
    }
    if (q < q1) s += sv[q] * zl[sc[q]];
    return s;
}
// wait for the completion flags of deps [d0, d1) (one thread each, d1 - d0 <= TBLK); true when
// a wait lapsed
__device__ inline bool tile_wait(const int32_t* __restrict__ dep, int32_t d0, int32_t d1, unsigned long long* sync,
                                 unsigned long long app, int32_t limit, int* fail) {
    const int32_t i = d0 + (int32_t)threadIdx.x;
    if (i < d1) {
        const int32_t d = dep[i];
        const unsigned long long want = 2 * app + 2 - (unsigned long long)(d & 1);
        unsigned long long* f = sync + 2 + (d >> 1);
        for (int spins = 0; limit < 0 || __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want; ++spins) {
            if (spins >= limit) {
                *fail = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    return *fail != 0;
}
template <class T>
__global__ void __launch_bounds__(TBLK) k_ilu_tile_chain(TileChain<T> a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int32_t lv[ILU_MAXLEV + 1];
    __shared__ unsigned long long s_ticket;
    __shared__ int s_fail;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_ticket = __hip_atomic_fetch_add(a.sync, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_fail = 0;
    }
    const int go = ctrl_go(a.ctrl);
    __syncthreads();
    if (go == 0) return;  // after convergence: every workgroup of the launch still took its ticket
    const unsigned long long t = s_ticket;
    const unsigned long long app = t / (unsigned long long)a.nrole;
    const int32_t i = __builtin_amdgcn_readfirstlane((int32_t)(t - app * (unsigned long long)a.nrole));
    const int32_t* R = a.role + (int64_t)TILE_ROLE_INTS * i;
    const int32_t r0 = R[0], r1 = R[1], lb = R[2], le = R[3], ub = R[4], ue = R[5], k = R[6], kind = R[7], d0 = R[8],
                  d1 = R[9];
    T* sv = reinterpret_cast<T*>(lds);
    T* zl = sv + a.emax;
    int16_t* sc = reinterpret_cast<int16_t*>(zl + a.rmax);
    int16_t* lp = sc + a.emax;
    const int nr = r1 - r0;
    if (tid <= ILU_MAXLEV) lv[tid] = a.lev[(int64_t)k * (ILU_MAXLEV + 1) + tid] - r0;
    int32_t j[TSTAGE];
    T v[TSTAGE];
    bool fail = false;
    if (kind & TILE_FWD) {  // z_t = c rhs_t - sum_L l_tj z_j
        const T c = (T)scale_of(a.rscale);
        T rv = 0;
        if (tid < nr) {
            rv = (T)(a.rhs[r0 + tid] * c);
            lp[tid] = (int16_t)(a.lptr[r0 + tid] - lb);
        }
        if (tid == 0) lp[nr] = (int16_t)(le - lb);
        tile_load<T>(a.lcol, a.lval, lb, le - lb, r0, j, v);
        fail = tile_wait(a.dep, d0, d1, a.sync, app, a.wait_limit, &s_fail);
        tile_place<T>(a.z, r0, r1, le - lb, true, j, v, sv, sc);
        __syncthreads();
        int range = 0;
        if (tid < nr) {
            T acc;
            range = tile_row_outside<T>(sv, sc, lp, tid, true, acc);
            zl[tid] = rv - acc;
        }
        __syncthreads();
        // both sweeps: the U entries load while the levels are walked
        if (kind & TILE_BWD) tile_load<T>(a.ucol, a.uval, ub, ue - ub, r0, j, v);
        for (int l = 0; l < a.nlev; ++l) {
            if (tid >= lv[l] && tid < lv[l + 1]) zl[tid] -= tile_row_inside<T>(sv, sc, zl, range);
            __syncthreads();
        }
    }
    if (kind & TILE_BWD) {  // z_t = (z_t - sum_U u_tj z_j) / u_tt
        T dv = 1;
        int range = 0;
        if (kind & TILE_FWD) {
            // every U entry inside the tile: they replace the L entries in LDS
            if (tid < nr) {
                dv = a.dval[r0 + tid];
                range = ((a.uptr[r0 + tid] - ub) << 16) | (a.uptr[r0 + tid + 1] - ub);
            }
#pragma unroll
            for (int u = 0; u < TSTAGE; ++u) {
                const int e = u * TBLK + tid;
                if (e < ue - ub) {
                    sv[e] = v[u];
                    sc[e] = (int16_t)(j[u] - r0);
                }
            }
            __syncthreads();
        } else {
            T zv = 0;
            if (tid < nr) {
                dv = a.dval[r0 + tid];
                lp[tid] = (int16_t)(a.uptr[r0 + tid] - ub);
            }
            if (tid == 0) lp[nr] = (int16_t)(ue - ub);
            tile_load<T>(a.ucol, a.uval, ub, ue - ub, r0, j, v);
            fail = tile_wait(a.dep, d0, d1, a.sync, app, a.wait_limit, &s_fail);
            if (tid < nr) zv = z_load<T>(a.z + r0 + tid);
            tile_place<T>(a.z, r0, r1, ue - ub, false, j, v, sv, sc);
            __syncthreads();
            if (tid < nr) {
                T acc;
                range = tile_row_outside<T>(sv, sc, lp, tid, false, acc);
                zl[tid] = zv - acc;
            }
            __syncthreads();
        }
        for (int l = a.nlev - 1; l >= 0; --l) {
            if (tid >= lv[l] && tid < lv[l + 1]) zl[tid] = (zl[tid] - tile_row_inside<T>(sv, sc, zl, range)) / dv;
            __syncthreads();
        }
    }
    if (tid < nr)
        __hip_atomic_store(a.z + r0 + tid, fail ? (T)__builtin_nan("") : zl[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (fail && tid == 0) __hip_atomic_fetch_add(a.sync + 1, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // publish: every wave's z stores complete before the barrier, then the flag
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0)
        __hip_atomic_store(a.sync + 2 + k, 2 * app + ((kind & TILE_BWD) ? 2ull : 1ull), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// IKJ ILU(0) rows of one (tile colour, level) group: the same update as k_ilu_factor
template <class T>
__global__ void __launch_bounds__(BLK) k_ilu_factor_rows(const int32_t* __restrict__ rows, int32_t cnt,
                                                         const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                         const int32_t* __restrict__ diag, T* __restrict__ lu) {
    const int32_t q = blockIdx.x * BLK + threadIdx.x;
    if (q >= cnt) return;
    const int32_t t = rows[q];
    const int32_t e = rp[t + 1], dt = diag[t];
    for (int32_t kk = rp[t]; kk < dt; ++kk) {
        const int32_t k = col[kk];
        const T lik = lu[kk] / lu[diag[k]];
        lu[kk] = lik;
        int32_t p = kk + 1;
        for (int32_t qq = diag[k] + 1; qq < rp[k + 1]; ++qq) {
            const int32_t j = col[qq];
            while (p < e && col[p] < j) ++p;
            if (p < e && col[p] == j) lu[p] -= lik * lu[qq];
        }
    }
}

// out[t] = in[perm[t]]
template <class T>
__global__ void __launch_bounds__(BLK) k_gather_perm(int64_t n, const int32_t* __restrict__ perm, const T* __restrict__ in,
                                                     T* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (t < n) out[t] = in[perm[t]];
}
// out[perm[t]] = in[t]
template <class T>
__global__ void __launch_bounds__(BLK) k_scatter_perm(int64_t n, const int32_t* __restrict__ perm, const T* __restrict__ in,
                                                      T* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (t < n) out[perm[t]] = in[t];
}

template <class TS, class TD>
__global__ void __launch_bounds__(BLK) k_convert(int64_t n, const TS* __restrict__ x, TD* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) y[i] = (TD)x[i];
}

// z = dinv * r
__global__ void __launch_bounds__(BLK) k_xupdate_set(int64_t n, const double* __restrict__ r, const double* __restrict__ dinv,
                                                     double* __restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) z[i] = r[i] * dinv[i];
}

// ---------------------------------------------------------------- block Jacobi
// Inverse of the NB x NB block of vertex v (unknowns v, N + v, ..., (NB-1) N + v): the block is
// gathered from the CSR rows by binary search (columns ascending), inverted by Gauss-Jordan
// elimination with partial pivoting in registers (compile-time indices only), and stored as
// inv[(i NB + j) N + v]; a singular or non-finite block becomes the identity.
template <int NB, bool IL>
__global__ void __launch_bounds__(BLK) k_bjac_setup(int64_t N, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                    const double* __restrict__ val, double* __restrict__ inv,
                                                    float* __restrict__ inv32) {
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= N) return;
    double a[NB][2 * NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int64_t row = IL ? (int64_t)NB * v + i : (int64_t)i * N + v;
        const int32_t s = rp[row], e = rp[row + 1];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int32_t target = (int32_t)(IL ? (int64_t)NB * v + j : (int64_t)j * N + v);
            int32_t lo = s, hi = e;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (col[mid] < target) lo = mid + 1;
                else hi = mid;
            }
            a[i][j] = (lo < e && col[lo] == target) ? val[lo] : 0.0;
            a[i][NB + j] = i == j ? 1.0 : 0.0;
        }
    }
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        // pivot row p >= k of largest magnitude, swapped into row k
        int p = k;
        double best = fabs(a[k][k]);
#pragma unroll
        for (int i = k + 1; i < NB; ++i)
            if (fabs(a[i][k]) > best) { best = fabs(a[i][k]); p = i; }
#pragma unroll
        for (int i = k + 1; i < NB; ++i) {
            if (i == p) {
#pragma unroll
                for (int j = 0; j < 2 * NB; ++j) { const double t = a[k][j]; a[k][j] = a[i][j]; a[i][j] = t; }
            }
        }
        if (!(best > 0.0) || !isfinite(best)) ok = false;
        const double piv = 1.0 / a[k][k];
#pragma unroll
        for (int j = 0; j < 2 * NB; ++j) a[k][j] *= piv;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            if (i == k) continue;
            const double f = a[i][k];
#pragma unroll
            for (int j = 0; j < 2 * NB; ++j) a[i][j] -= f * a[k][j];
        }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            double x = ok ? a[i][NB + j] : (i == j ? 1.0 : 0.0);
            if (!isfinite(x)) x = i == j ? 1.0 : 0.0;
            inv[((int64_t)i * NB + j) * N + v] = x;
            if (inv32) inv32[((int64_t)i * NB + j) * N + v] = (float)x;
        }
}
// z = B^-1 (c r) per vertex block (c = *rscale or 1); z must not alias r
// (IL: the vectors hold the NB unknowns of a vertex contiguously, entry (j, v) at NB v + j)
template <class T, int NB, bool IL>
__global__ void __launch_bounds__(BLK) k_bjac_apply(int64_t N, const T* __restrict__ inv, const T* __restrict__ r,
                                                    const double* __restrict__ rscale, T* __restrict__ z,
                                                    const int* __restrict__ ctrl) {
    if (ctrl && ctrl[0] == 0) return;
    const int64_t v = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (v >= N) return;
    T rr[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int64_t k = IL ? NB * v + j : (int64_t)j * N + v;
        rr[j] = rscale ? (T)(r[k] * (T)rscale[0]) : r[k];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        T acc = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) acc += inv[((int64_t)i * NB + j) * N + v] * rr[j];
        z[IL ? NB * v + i : (int64_t)i * N + v] = acc;
    }
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

struct HipEvents {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    hipError_t create() {
        for (auto& x : e) {
            hipError_t r = hipEventCreate(&x);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
    ~HipEvents() {
        for (auto& x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

}  // namespace

// ---------------------------------------------------------------- timing helpers
void ktime_begin(Ctx& c, int kclass, hipEvent_t* a) {
    *a = nullptr;
    if (!(c.timing_mask & (1u << kclass))) return;
    if (c.ev_next + 2 > c.ev_pool.size()) {
        for (int k = 0; k < 256; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            c.ev_pool.push_back(e);
        }
    }
    *a = c.ev_pool[c.ev_next++];
    (void)hipEventRecord(*a, c.stream);
}
void ktime_end(Ctx& c, int kclass, double bytes, hipEvent_t a) {
    if (!a) return;
    hipEvent_t b = c.ev_pool[c.ev_next++];
    (void)hipEventRecord(b, c.stream);
    c.pending.push_back({kclass, bytes, a, b, c.cur_step});
}
int ktime_flush(Ctx& c) {
    if (c.pending.empty()) {
        c.ev_next = 0;
        return VFVM_OK;
    }
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    for (auto& p : c.pending) {
        float ms = 0.0f;
        VFVM_HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
        if (p.bytes >= 0.0) {
            c.kt[p.kclass].launches += 1;
            c.kt[p.kclass].ms += ms;
            c.kt[p.kclass].bytes += p.bytes;
        }
    }
    c.pending.clear();
    c.ev_next = 0;
    return VFVM_OK;
}

// ---------------------------------------------------------------- solver
namespace {

// Launches `kernel`; when its class is timed, the launch carries start/stop events
// (hipExtLaunchKernelGGL records them in the dispatch itself, no extra queue commands).
template <typename K, typename... Args>
void launch_timed(Ctx& c, int kclass, double bytes, K kernel, dim3 grid, dim3 block, Args... args) {
    hipEvent_t a = nullptr, b = nullptr;
    if (c.timing_mask & (1u << kclass)) {
        if (c.ev_next + 2 > c.ev_pool.size()) {
            for (int k = 0; k < 256; ++k) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) break;
                c.ev_pool.push_back(e);
            }
        }
        if (c.ev_next + 2 <= c.ev_pool.size()) {
            a = c.ev_pool[c.ev_next++];
            b = c.ev_pool[c.ev_next++];
        }
    }
    if (a) {
        hipExtLaunchKernelGGL(kernel, grid, block, 0, c.stream, a, b, 0, args...);
        c.pending.push_back({kclass, bytes, a, b, c.cur_step});
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, c.stream, args...);
    }
}

// the same with `shmem` bytes of dynamic LDS
template <typename K, typename... Args>
void launch_timed_lds(Ctx& c, int kclass, double bytes, K kernel, dim3 grid, dim3 block, size_t shmem, Args... args) {
    hipEvent_t a = nullptr, b = nullptr;
    if (c.timing_mask & (1u << kclass)) {
        if (c.ev_next + 2 > c.ev_pool.size()) {
            for (int k = 0; k < 256; ++k) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) break;
                c.ev_pool.push_back(e);
            }
        }
        if (c.ev_next + 2 <= c.ev_pool.size()) {
            a = c.ev_pool[c.ev_next++];
            b = c.ev_pool[c.ev_next++];
        }
    }
    if (a) {
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, c.stream, a, b, 0, args...);
        c.pending.push_back({kclass, bytes, a, b, c.cur_step});
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, c.stream, args...);
    }
}

struct Work {
    int64_t n, nnz;
    int64_t ldv;  // basis stride: n rounded up to 64 entries (16-byte aligned columns)
    int m;
    int nblk;  // blocks of the update kernel (one norm partial each)
    int64_t rows_per_blk;
    int nwav;  // waves of the multi-dot (one partial per wave and basis vector)
    int64_t rows_per_wave;
    StateLayout L;
    int* ctrl;    // device control word
    int* h_ctrl;  // pinned
    double* S;    // device state
    double* part; // reduction partials
};

// lanes per CSR row from the average row length (scalar rows carry ~17 entries, the generic
// coupled rows ~100-480)
inline int pick_lpr(int64_t n, int64_t nnz) {
    double avg = n > 0 ? (double)nnz / (double)n : 1.0;
    return avg > 96.0 ? 32 : (avg > 24.0 ? 16 : (avg > 10.0 ? 8 : 4));
}

template <class T>
int spmv_launch(Ctx& c, const SysView& sy, int lpr, int64_t n, const T* val, const T* x, T* y, const double* scale,
                const int* ctrl, double bytes) {
    unsigned grid = grid_for(n * lpr);
    const int32_t* rp = sy.row;
    const int32_t* cl = sy.col;
    if (sy.flow_layout) {
        // the coupled flow system: vertex-interleaved vectors take 32 lanes per vertex (four
        // 3-ring and two grad_coeff entries per lane and trip), comp-major vectors one wave
        FlowSpmvGeom fg{c.g.N, c.g.NP, c.g.row_ptr.p, c.g.gc_ptr.p};
#define VFVM_FLOW_SPMV(NS)                                                                                                   \
    if (sy.flow_layout == 2)                                                                                                 \
        launch_timed(c, VFVM_K_SPMV, bytes, k_spmv_flow<T, true, 32, NS>, dim3(grid_for(c.g.N * 32)), dim3(BLK), fg, rp, cl, \
                     val, x, y, scale, ctrl);                                                                                \
    else                                                                                                                     \
        launch_timed(c, VFVM_K_SPMV, bytes, k_spmv_flow<T, false, 64, NS>, dim3(grid_for(c.g.N * WAVE)), dim3(BLK), fg, rp,  \
                     cl, val, x, y, scale, ctrl);
        switch (sy.flow_ns) {
            case 0: VFVM_FLOW_SPMV(0) break;
            case 1: VFVM_FLOW_SPMV(1) break;
            default: VFVM_FLOW_SPMV(2) break;
        }
#undef VFVM_FLOW_SPMV
        VFVM_HIP_TRY(hipGetLastError());
        return VFVM_OK;
    }
    if (sy.interleaved && c.g.N > 0 && n % c.g.N == 0) {
        const int64_t nv = c.g.N;
        const dim3 vg(grid_for(nv * WAVE));
        switch (n / nv) {
            case 1: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv_vblock<T, 1>, vg, dim3(BLK), nv, rp, cl, val, x, y, scale, ctrl); break;
            case 2: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv_vblock<T, 2>, vg, dim3(BLK), nv, rp, cl, val, x, y, scale, ctrl); break;
            case 3: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv_vblock<T, 3>, vg, dim3(BLK), nv, rp, cl, val, x, y, scale, ctrl); break;
            case 4: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv_vblock<T, 4>, vg, dim3(BLK), nv, rp, cl, val, x, y, scale, ctrl); break;
            case 5: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv_vblock<T, 5>, vg, dim3(BLK), nv, rp, cl, val, x, y, scale, ctrl); break;
            case 6: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv_vblock<T, 6>, vg, dim3(BLK), nv, rp, cl, val, x, y, scale, ctrl); break;
            default: return set_error(VFVM_ERR_INVALID_ARGUMENT, "SpMV: 1 to 6 unknowns per vertex");
        }
        VFVM_HIP_TRY(hipGetLastError());
        return VFVM_OK;
    }
    if (sy.spmv_emax > 0 && lpr <= 8 && sy.spmv_emax <= SPMV_BLK * 12) {
        // one trip per chunk: STAGE * SPMV_BLK >= emax
        const int64_t nchunk = (n + SPMV_ROWS - 1) / SPMV_ROWS;
        const dim3 g((unsigned)((nchunk + SPMV_CHUNKS - 1) / SPMV_CHUNKS));
        const size_t lds = spmv_chunk_lds(sizeof(T), sy.spmv_emax);
        const int32_t em = sy.spmv_emax;
        if (em <= SPMV_BLK * 8)
            launch_timed_lds(c, VFVM_K_SPMV, bytes, k_spmv_chunk<T, 8>, g, dim3(SPMV_BLK), lds, n, rp, cl, val, x, y, scale, ctrl, em);
        else if (em <= SPMV_BLK * 9)
            launch_timed_lds(c, VFVM_K_SPMV, bytes, k_spmv_chunk<T, 9>, g, dim3(SPMV_BLK), lds, n, rp, cl, val, x, y, scale, ctrl, em);
        else if (em <= SPMV_BLK * 10)
            launch_timed_lds(c, VFVM_K_SPMV, bytes, k_spmv_chunk<T, 10>, g, dim3(SPMV_BLK), lds, n, rp, cl, val, x, y, scale, ctrl, em);
        else
            launch_timed_lds(c, VFVM_K_SPMV, bytes, k_spmv_chunk<T, 12>, g, dim3(SPMV_BLK), lds, n, rp, cl, val, x, y, scale, ctrl, em);
        VFVM_HIP_TRY(hipGetLastError());
        return VFVM_OK;
    }
    switch (lpr) {
        case 4: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv<T, 4>, dim3(grid), dim3(BLK), n, rp, cl, val, x, y, scale, ctrl); break;
        case 8: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv<T, 8>, dim3(grid), dim3(BLK), n, rp, cl, val, x, y, scale, ctrl); break;
        case 32: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv<T, 32>, dim3(grid), dim3(BLK), n, rp, cl, val, x, y, scale, ctrl); break;
        default: launch_timed(c, VFVM_K_SPMV, bytes, k_spmv<T, 16>, dim3(grid), dim3(BLK), n, rp, cl, val, x, y, scale, ctrl); break;
    }
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

int residual_launch(Ctx& c, const SysView& sy, int lpr, const double* x, double* r, double* part, int nblk_r) {
    const int64_t n = sy.n;
    unsigned grid = (unsigned)nblk_r;
    if (sy.flow_layout) {
        FlowSpmvGeom fg{c.g.N, c.g.NP, c.g.row_ptr.p, c.g.gc_ptr.p};
#define VFVM_FLOW_RES(NS)                                                                                                      \
    if (sy.flow_layout == 2)                                                                                                   \
        hipLaunchKernelGGL((k_residual_flow<true, NS>), dim3(grid), dim3(BLK), 0, c.stream, fg, sy.row, sy.col, sy.vals, x, sy.b, r, \
                           part, nblk_r);                                                                                      \
    else                                                                                                                       \
        hipLaunchKernelGGL((k_residual_flow<false, NS>), dim3(grid), dim3(BLK), 0, c.stream, fg, sy.row, sy.col, sy.vals, x, sy.b,   \
                           r, part, nblk_r);
        switch (sy.flow_ns) {
            case 0: VFVM_FLOW_RES(0) break;
            case 1: VFVM_FLOW_RES(1) break;
            default: VFVM_FLOW_RES(2) break;
        }
#undef VFVM_FLOW_RES
        VFVM_HIP_TRY(hipGetLastError());
        return VFVM_OK;
    }
    switch (lpr) {
        case 4: hipLaunchKernelGGL((k_residual<4>), dim3(grid), dim3(BLK), 0, c.stream, n, sy.row, sy.col, sy.vals, x, sy.b, r, part, nblk_r); break;
        case 8: hipLaunchKernelGGL((k_residual<8>), dim3(grid), dim3(BLK), 0, c.stream, n, sy.row, sy.col, sy.vals, x, sy.b, r, part, nblk_r); break;
        case 32: hipLaunchKernelGGL((k_residual<32>), dim3(grid), dim3(BLK), 0, c.stream, n, sy.row, sy.col, sy.vals, x, sy.b, r, part, nblk_r); break;
        default: hipLaunchKernelGGL((k_residual<16>), dim3(grid), dim3(BLK), 0, c.stream, n, sy.row, sy.col, sy.vals, x, sy.b, r, part, nblk_r); break;
    }
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

// hout[k] = sig_k V[k] . wv over the basis vectors [0, nv): per-block partials, then one
// reduction block per vector.  Two basis vectors per pass: at 100^3 the average multi-dot
// takes FP32 21.4 us and FP64 40.3 us against 24.4 / 47.4 us for eight per pass (fewer
// accumulators per lane keep more waves in flight, and the re-reads of w hit L2)
constexpr int MDOT_CHUNK = 2;
template <class T>
void mdot_launch(Ctx& c, Work& w, int nv, const T* V, const T* wv, double* part, double* hout) {
    const int nbx = w.nwav / (BLK / WAVE);
    // algorithmic bytes: the nv basis vectors and w once
    const double bytes = (double)sizeof(T) * w.n * (nv + 1);
    const dim3 grid(nbx, (nv + MDOT_CHUNK - 1) / MDOT_CHUNK), block(BLK);
    launch_timed(c, VFVM_K_MDOT, bytes, k_mdot<T, MDOT_CHUNK>, grid, block, w.n, nv, V, w.ldv, wv, part, w.rows_per_wave,
                 (const int*)w.ctrl);
    hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(BLK), 0, c.stream, nbx, (const double*)part, (const double*)(w.S + w.L.sig()),
                       hout, (const int*)w.ctrl);
}

// LDS of a tile sweep workgroup (value size vb): staged entries and their 16-bit local columns,
// the tile's z, 16-bit row offsets
constexpr int64_t ILU_TILE_LDS = 64 * 1024 - 256;
inline int64_t tile_lds_bytes(int vb, int64_t emax, int64_t rmax) { return vb * (emax + rmax) + 2 * (emax + rmax + 1); }

// Row order of the matrix pattern (host, ilu_order.cpp: tiled, or multicolour), the permuted
// CSR and the split triangle index arrays, uploaded once per pattern.
int ilu_setup(Ctx& c, const SysView& sy) {
    IluData& I = *sy.ilu;
    if (I.ready) return VFVM_OK;
    IluOrder o;
    std::string err;
    const double* pos = c.g.h_pos.empty() ? nullptr : c.g.h_pos.data();
    if (ilu_order_build(sy.n, *sy.h_row, *sy.h_col, pos, c.g.N, o, err)) return set_error(VFVM_ERR_INVALID_ARGUMENT, err);
    // the tiled order only when a tile's FP64 entries of both triangles fit the LDS of one
    // workgroup (16-bit offsets inside a tile) and one load round of its threads, and no tile
    // sweep waits for more tiles than it has threads; else the multicolour order
    std::vector<int32_t> role, dep;
    const int64_t emax = std::max(o.emax_l, o.emax_u);
    if (!o.tile_row.empty() && (tile_lds_bytes(8, emax, o.rmax) > ILU_TILE_LDS || emax > TBLK * TSTAGE ||
                                !tile_roles_build(o, TBLK, role, dep)))
        if (ilu_order_build(sy.n, *sy.h_row, *sy.h_col, pos, c.g.N, o, err, false)) return set_error(VFVM_ERR_INVALID_ARGUMENT, err);
    I.tiled = !o.tile_row.empty();
    if (I.tiled) {
        const int64_t nt = (int64_t)o.tile_row.size() - 1;
        VFVM_HIP_TRY(I.lev_row.upload(o.lev_row.data(), o.lev_row.size(), c.stream));
        VFVM_HIP_TRY(I.fac_rows.upload(o.fac_rows.data(), o.fac_rows.size(), c.stream));
        VFVM_HIP_TRY(I.trole.upload(role.data(), role.size(), c.stream));
        VFVM_HIP_TRY(I.tdep.upload(dep.data(), std::max<size_t>(dep.size(), 1), c.stream));
        // the ticket counter starts at 0 and every flag below the first application's values
        const std::vector<unsigned long long> zero(2 + nt, 0ull);
        VFVM_HIP_TRY(I.tsync.upload(zero.data(), zero.size(), c.stream));
        I.nrole = (int32_t)(role.size() / TILE_ROLE_INTS);
        I.h_fac_ptr = o.fac_ptr;
        I.h_tile_row = o.tile_row;
        I.h_lev_row = o.lev_row;
        I.sweep_rows = I.sweep_lent = I.sweep_uent = I.fused_rows = 0.0;
        for (int32_t i = 0; i < I.nrole; ++i) {
            const int32_t* R = role.data() + (size_t)TILE_ROLE_INTS * i;
            if (R[7] & 1) { I.sweep_rows += R[1] - R[0]; I.sweep_lent += R[3] - R[2]; }
            if (R[7] & 2) I.sweep_uent += R[5] - R[4];
            if (R[7] == 3) I.fused_rows += R[1] - R[0];
        }
        I.nlev = o.nlev; I.rmax = o.rmax; I.emax_l = o.emax_l; I.emax_u = o.emax_u;
    }
    VFVM_HIP_TRY(I.perm.upload(o.perm.data(), o.perm.size(), c.stream));
    VFVM_HIP_TRY(I.prp.upload(o.prp.data(), o.prp.size(), c.stream));
    I.h_prp = o.prp;
    VFVM_HIP_TRY(I.src.upload(o.src.data(), o.src.size(), c.stream));
    VFVM_HIP_TRY(I.pcol.upload(o.pcol.data(), o.pcol.size(), c.stream));
    VFVM_HIP_TRY(I.pdiag.upload(o.pdiag.data(), o.pdiag.size(), c.stream));
    VFVM_HIP_TRY(I.lptr.upload(o.lptr.data(), o.lptr.size(), c.stream));
    VFVM_HIP_TRY(I.lcol.upload(o.lcol.data(), std::max<size_t>(o.lcol.size(), 1), c.stream));
    VFVM_HIP_TRY(I.lpos.upload(o.lpos.data(), std::max<size_t>(o.lpos.size(), 1), c.stream));
    VFVM_HIP_TRY(I.uptr.upload(o.uptr.data(), o.uptr.size(), c.stream));
    VFVM_HIP_TRY(I.ucol.upload(o.ucol.data(), std::max<size_t>(o.ucol.size(), 1), c.stream));
    VFVM_HIP_TRY(I.upos.upload(o.upos.data(), std::max<size_t>(o.upos.size(), 1), c.stream));
    I.nnz_l = (int64_t)o.lpos.size();
    I.nnz_u = (int64_t)o.upos.size();
    ++I.generation;
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    I.h_color_ptr = o.color_ptr;
    I.ready = true;
    return VFVM_OK;
}

// ILU(0) of A in multicolour order (IKJ, Saad Alg. 10.4, one launch per colour) in the
// colour-major [L | diag | U] rows of `full` (nnz entries), then split into lu = [L | U | diag]
template <class T>
int ilu_factor(Ctx& c, const SysView& sy, T* lu, T* full, const double* val) {
    const IluData& I = *sy.ilu;
    hipLaunchKernelGGL((k_ilu_gather<T>), dim3(grid_for(sy.nnz)), dim3(BLK), 0, c.stream, sy.nnz, I.src.p, val, full);
    const int ncol = I.tiled ? 0 : (int)I.h_color_ptr.size() - 1;
    // tiled order: the rows of every (tile colour, level) group are independent
    for (size_t g = 0; I.tiled && g + 1 < I.h_fac_ptr.size(); ++g) {
        const int32_t s = I.h_fac_ptr[g], e = I.h_fac_ptr[g + 1];
        if (e > s)
            hipLaunchKernelGGL((k_ilu_factor_rows<T>), dim3(grid_for(e - s)), dim3(BLK), 0, c.stream, I.fac_rows.p + s, e - s,
                               I.prp.p, I.pcol.p, I.pdiag.p, full);
    }
    for (int q = 0; q < ncol; ++q) {
        int32_t s = I.h_color_ptr[q], e = I.h_color_ptr[q + 1];
        if (e > s)
            hipLaunchKernelGGL((k_ilu_factor<T>), dim3(grid_for(e - s)), dim3(BLK), 0, c.stream, s, e, I.prp.p,
                               I.pcol.p, I.pdiag.p, full);
    }
    hipLaunchKernelGGL((k_ilu_split<T>), dim3(grid_for(sy.nnz)), dim3(BLK), 0, c.stream, I.nnz_l, I.nnz_u, sy.n, I.lpos.p,
                       I.upos.p, I.pdiag.p, (const T*)full, lu);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

// Lanes per row of the sweeps: four (a row holds ~8 entries of each triangle)
constexpr int ILU_LPR = 4;

// z = (LU)^-1 rhs in the permuted space, in place in z (z must not alias rhs); lu is the split
// factor [L | U | diag].  Tiled order: one launch (k_ilu_tile_chain).  Multicolour order: forward
// sweep over the colours ascending (the last one also finishing its backward sweep; colours 0
// and 1 share a launch), backward sweep over the other colours descending: 2 colours - 2
// launches.
template <class T>
int ilu_apply(Ctx& c, const SysView& sy, const T* lu, const T* rhs, const double* rscale, T* z, const int* ctrl) {
    constexpr int LPR = ILU_LPR;
    const IluData& I = *sy.ilu;
    const int ncol = (int)I.h_color_ptr.size() - 1;
    const T* lval = lu;
    const T* uval = lu + I.nnz_l;
    const T* dval = lu + I.nnz_l + I.nnz_u;
    // algorithmic bytes of one sweep launch over the rows [s, e): the triangle's entries (value
    // + column), its row pointers (8 B per row; the last colour also reads the diagonal), the
    // row's rhs / z read and z write
    auto sweep_bytes = [&](int32_t s, int32_t e, bool lower, bool final) -> double {
        const double ent = lower ? (double)I.nnz_l : (double)I.nnz_u;
        // entries of the colour in proportion to its rows (the exact count needs the host
        // pointers; rows of one colour carry the same share of a triangle on average)
        const double per_row = ent / (double)std::max<int64_t>(sy.n, 1);
        return (double)(e - s) * (per_row * (sizeof(T) + 4.0) + 8.0 + 2.0 * sizeof(T) + ((final || !lower) ? sizeof(T) : 0.0));
    };
    if (I.tiled) {
        // every tile sweep of the application in one launch; algorithmic bytes: the triangles'
        // entries (value + column), per row its L and U row pointers, rhs, diagonal and z written
        // twice and read once between the sweeps (once, for the tiles doing both sweeps)
        const double vb = (double)sizeof(T);
        const double bytes = (I.sweep_lent + I.sweep_uent) * (vb + 4.0) + I.sweep_rows * (8.0 + 5.0 * vb) -
                             I.fused_rows * 2.0 * vb;
        TileChain<T> a{I.trole.p, I.lev_row.p, I.tdep.p, I.lptr.p, I.lcol.p, I.uptr.p, I.ucol.p, lval, uval, dval, rhs,
                       rscale, z, ctrl, I.tsync.p, I.nrole, I.nlev, std::max(I.emax_l, I.emax_u), I.rmax, I.wait_limit};
        const size_t lds = (size_t)tile_lds_bytes(sizeof(T), std::max(I.emax_l, I.emax_u), I.rmax);
        launch_timed_lds(c, VFVM_K_PRECOND, bytes, k_ilu_tile_chain<T>, dim3(I.nrole), dim3(TBLK), lds, a);
        VFVM_HIP_TRY(hipGetLastError());
        return VFVM_OK;
    }
    int q0 = 0;
    if (ncol >= 2 && I.h_color_ptr[0] == 0 && I.h_color_ptr[2] > I.h_color_ptr[1]) {
        // colours 0 and 1 in one launch (colour 0 is a scaled copy)
        const int32_t s = I.h_color_ptr[1], e = I.h_color_ptr[2];
        const dim3 grid(grid_for((int64_t)(e - s) * LPR));
        const double bytes = (double)s * 2.0 * sizeof(T) + sweep_bytes(s, e, true, ncol == 2);
        if (ncol == 2)
            launch_timed(c, VFVM_K_PRECOND, bytes, k_ilu_lower01<T, LPR, true>, grid, dim3(BLK), s, e, I.lptr.p, I.lcol.p, lval,
                         dval, rhs, rscale, z, ctrl);
        else
            launch_timed(c, VFVM_K_PRECOND, bytes, k_ilu_lower01<T, LPR, false>, grid, dim3(BLK), s, e, I.lptr.p, I.lcol.p, lval,
                         dval, rhs, rscale, z, ctrl);
        q0 = 2;
    }
    for (int q = q0; q < ncol; ++q) {
        int32_t s = I.h_color_ptr[q], e = I.h_color_ptr[q + 1];
        if (e == s) continue;
        const dim3 grid(grid_for((int64_t)(e - s) * LPR));
        const double bytes = sweep_bytes(s, e, true, q == ncol - 1);
        if (q == ncol - 1)
            launch_timed(c, VFVM_K_PRECOND, bytes, k_ilu_lower<T, LPR, true>, grid, dim3(BLK), s, e, I.lptr.p, I.lcol.p, lval,
                         dval, rhs, rscale, z, ctrl);
        else
            launch_timed(c, VFVM_K_PRECOND, bytes, k_ilu_lower<T, LPR, false>, grid, dim3(BLK), s, e, I.lptr.p, I.lcol.p, lval,
                         dval, rhs, rscale, z, ctrl);
    }
    for (int q = ncol - 2; q >= 0; --q) {
        int32_t s = I.h_color_ptr[q], e = I.h_color_ptr[q + 1];
        if (e == s) continue;
        launch_timed(c, VFVM_K_PRECOND, sweep_bytes(s, e, false, false), k_ilu_upper<T, LPR>,
                     dim3(grid_for((int64_t)(e - s) * LPR)), dim3(BLK), s, e, I.uptr.p, I.ucol.p, uval, dval, z, ctrl);
    }
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

// Explicitly applied preconditioner of a cycle: the ILU(0) factor or the block-Jacobi
// inverse blocks (Jacobi enters through the column-scaled matrix instead).
template <class T>
struct Pc {
    const T* lu = nullptr;
    const T* bj = nullptr;
    int nb = 0;
    int64_t nv = 0;   // vertices (block-Jacobi blocks)
    bool il = false;  // vectors vertex-interleaved
};

template <class T>
int bjac_apply(Ctx& c, const Pc<T>& pc, const T* r, const double* rscale, T* z, const int* ctrl) {
    const double bytes = (double)sizeof(T) * pc.nv * (pc.nb * pc.nb + 2 * pc.nb);
    const dim3 grid(grid_for(pc.nv));
#define VFVM_BJ(NB)                                                                                                         \
    case NB:                                                                                                                \
        if (pc.il) launch_timed(c, VFVM_K_PRECOND, bytes, k_bjac_apply<T, NB, true>, grid, dim3(BLK), pc.nv, pc.bj, r, rscale, z, ctrl); \
        else launch_timed(c, VFVM_K_PRECOND, bytes, k_bjac_apply<T, NB, false>, grid, dim3(BLK), pc.nv, pc.bj, r, rscale, z, ctrl); \
        break;
    switch (pc.nb) {
        VFVM_BJ(1) VFVM_BJ(2) VFVM_BJ(3) VFVM_BJ(4) VFVM_BJ(5) VFVM_BJ(6)
        default: return set_error(VFVM_ERR_INVALID_ARGUMENT, "block Jacobi: block size must be 1..6");
    }
#undef VFVM_BJ
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

// block-Jacobi set-up: blocks of nb = n / N unknowns per vertex
int bjac_setup(Ctx& c, const SysView& sy, bool with32) {
    const int64_t N = c.g.N;
    if (N <= 0 || sy.n % N != 0 || sy.n / N < 1 || sy.n / N > 6)
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "block Jacobi: the system size must be b x (vertices) with b in 1..6");
    const int nb = (int)(sy.n / N);
    c.bjac.nb = nb;
    VFVM_HIP_TRY(c.bjac.inv64.alloc((size_t)nb * nb * N));
    float* i32 = nullptr;
    if (with32) {
        VFVM_HIP_TRY(c.bjac.inv32.alloc((size_t)nb * nb * N));
        i32 = c.bjac.inv32.p;
    }
    const dim3 grid(grid_for(N));
#define VFVM_BJS(NB)                                                                                                             \
    case NB:                                                                                                                     \
        if (sy.interleaved) hipLaunchKernelGGL((k_bjac_setup<NB, true>), grid, dim3(BLK), 0, c.stream, N, sy.row, sy.col, sy.vals, \
                                               c.bjac.inv64.p, i32);                                                             \
        else hipLaunchKernelGGL((k_bjac_setup<NB, false>), grid, dim3(BLK), 0, c.stream, N, sy.row, sy.col, sy.vals, c.bjac.inv64.p, \
                                i32);                                                                                            \
        break;
    switch (nb) {
        VFVM_BJS(1) VFVM_BJS(2) VFVM_BJS(3) VFVM_BJS(4) VFVM_BJS(5)
        default: VFVM_BJS(6)
    }
#undef VFVM_BJS
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

// The launches of Arnoldi step j of the cycle on A M^-1 (M = D through the scaled matrix for
// Jacobi, M = LU for ILU(0), M = block diagonal for block Jacobi).
template <class T>
int issue_step(Ctx& c, const SysView& sy, Work& w, T* V, T* wv, T* zv, const T* aval, const Pc<T>& pc, bool reorth, int j) {
    const T* lu = pc.lu;
    const int64_t n = w.n;
    const int lpr = pick_lpr(n, w.nnz);
    double* part = w.part;
    double* part_nrm = part + (int64_t)(MAX_RESTART + 1) * w.nwav;
    // algorithmic bytes: values, the column indices the kernel must read, row pointers, x, y
    const bool flow_spmv = sy.flow_layout != 0;
    const double bytes_spmv = flow_spmv ? (double)w.nnz * sizeof(T) + 4.0 * c.fs.spmv_idx + 4.0 * (n + 1) + 8.0 * c.g.N +
                                              2.0 * sizeof(T) * n
                                        : (double)w.nnz * (sizeof(T) + 4) + 4.0 * (n + 1) + 2.0 * sizeof(T) * n;
    // v_j = sig_j V[j]: the scale enters the preconditioner's right-hand side, or the SpMV
    // output when there is no ILU
    T* vj = V + (int64_t)j * w.ldv;
    const double* sigj = w.S + w.L.sig() + j;
    const T* xin = vj;
    const double* yscale = sigj;
    if (lu) {
        if (int st = ilu_apply<T>(c, sy, lu, vj, sigj, zv, w.ctrl)) return st;
        xin = zv;
        yscale = nullptr;
    } else if (pc.bj) {
        if (int st = bjac_apply<T>(c, pc, vj, sigj, zv, w.ctrl)) return st;
        xin = zv;
        yscale = nullptr;
    }
    if (int st = spmv_launch<T>(c, sy, lpr, n, aval, xin, wv, yscale, w.ctrl, bytes_spmv)) return st;
    // V[j+1] = w - sum_k v_k h_k (CGS; CGS2 repeats the pass on V[j+1]), then the Arnoldi step
    const int nv = j + 1;
    T* vnext = V + (int64_t)(j + 1) * w.ldv;
    const double bytes_upd = (double)sizeof(T) * n * (nv + 2);
    for (int pass = 0; pass < (reorth ? 2 : 1); ++pass) {
        double* hout = w.S + (pass == 0 ? w.L.h1() : w.L.h2());
        const bool last = pass == (reorth ? 1 : 0);
        const T* src = pass == 0 ? (const T*)wv : (const T*)vnext;
        mdot_launch<T>(c, w, nv, V, src, part, hout);
        launch_timed(c, VFVM_K_UPDATE, bytes_upd, k_update<T>, dim3(w.nblk), dim3(BLK), n, nv, (const T*)V, w.ldv,
                     (const double*)hout, (const double*)(w.S + w.L.sig()), src, vnext,
                     last ? part_nrm : (double*)nullptr, w.rows_per_blk, (const int*)w.ctrl);
    }
    hipLaunchKernelGGL(k_arnoldi_step, dim3(1), dim3(BLK), 0, c.stream, j, reorth ? 1 : 0, w.nblk, (const double*)part_nrm,
                       w.S, w.L, w.ctrl);
    VFVM_HIP_TRY(hipGetLastError());
    return VFVM_OK;
}

// Cycle graphs: the launches of steps [CHECK * q, CHECK * q + CHECK) of a cycle captured as
// one HIP graph each, keyed by every pointer and size baked into them; replayed for every
// cycle of every solve with the same key.  Launches after convergence return at once (the
// kernels test the control word).
constexpr int CHECK = 8;

template <class T>
CycleGraphs* cycle_graphs(Ctx& c, const SysView& sy, Work& w, T* V, T* wv, T* zv, const T* aval, const Pc<T>& pc, bool reorth,
                          int* status) {
    const T* lu = pc.lu;
    const IluData* I = lu ? sy.ilu : nullptr;
    std::array<uint64_t, 22> key = {(uint64_t)(uintptr_t)pc.bj, (uint64_t)pc.nb + (pc.il ? 256u : 0u) + 512u * (uint64_t)(sy.flow_layout + 4 * sy.flow_ns),
                                    (uint64_t)(uintptr_t)V,
                                    (uint64_t)(uintptr_t)wv, (uint64_t)(uintptr_t)zv,
                                    (uint64_t)(uintptr_t)aval, (uint64_t)(uintptr_t)lu, (uint64_t)(uintptr_t)w.S,
                                    (uint64_t)(uintptr_t)w.ctrl, (uint64_t)(uintptr_t)w.part,
                                    (uint64_t)(uintptr_t)sy.row, (uint64_t)(uintptr_t)sy.col, (uint64_t)w.n,
                                    (uint64_t)w.nnz, (uint64_t)w.m, (uint64_t)reorth, (uint64_t)sizeof(T),
                                    (uint64_t)(uintptr_t)I, I ? I->generation : 0, (uint64_t)w.nblk, (uint64_t)w.nwav, (uint64_t)w.ldv};
    for (auto& g : c.graphs)
        if (g.key == key) { g.last_use = ++c.graph_clock; return &g; }
    if (c.graphs.size() >= 4) {
        auto old = std::min_element(c.graphs.begin(), c.graphs.end(),
                                    [](const CycleGraphs& a, const CycleGraphs& b) { return a.last_use < b.last_use; });
        for (auto e : old->chunks) (void)hipGraphExecDestroy(e);
        c.graphs.erase(old);
    }
    CycleGraphs g;
    g.key = key;
    g.last_use = ++c.graph_clock;
    for (int j0 = 0; j0 < w.m; j0 += CHECK) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) { *status = set_error(VFVM_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e)); break; }
        int st = VFVM_OK;
        for (int j = j0; j < std::min(w.m, j0 + CHECK) && st == VFVM_OK; ++j)
            st = issue_step<T>(c, sy, w, V, wv, zv, aval, pc, reorth, j);
        e = hipStreamEndCapture(c.stream, &graph);
        if (st != VFVM_OK) { *status = st; if (graph) (void)hipGraphDestroy(graph); break; }
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) { *status = set_error(VFVM_ERR_HIP, std::string("cycle graph: ") + hipGetErrorString(e)); break; }
        g.chunks.push_back(exec);
    }
    if ((int)g.chunks.size() * CHECK < w.m) {
        for (auto e : g.chunks) (void)hipGraphExecDestroy(e);
        return nullptr;
    }
    c.graphs.push_back(std::move(g));
    return &c.graphs.back();
}

// One restart cycle from V[0] (normalised): leaves ctrl[1] = k steps, y = H^-1 g.  Runs the
// cycle graphs, or the launches one by one when kernel timing is on (the per-launch events
// need the eager path) or graphs are off (VFVM_GRAPHS=0).
template <class T>
int arnoldi_cycle(Ctx& c, const SysView& sy, Work& w, T* V, T* wv, T* zv, const T* aval, const Pc<T>& pc, bool reorth,
                  bool graphs, int* steps_out) {
    int steps = 0;
    CycleGraphs* cg = nullptr;
    if (graphs && c.timing_mask == 0 && c.graphs_enabled) {
        int status = VFVM_OK;
        cg = cycle_graphs<T>(c, sy, w, V, wv, zv, aval, pc, reorth, &status);
        if (status != VFVM_OK) return status;
    }
    if (cg) {
        // The next chunk is launched before the host waits for the control word of the
        // current one, so the GPU does not idle while the host reads it back and launches;
        // after convergence the queued chunk's kernels return at once.  Each chunk's control
        // word is read back into its own pinned slot (slot q % 2, 256 bytes apart).
        for (auto& e : c.ctrl_ev)
            if (!e) VFVM_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        const int nchunk = (w.m + CHECK - 1) / CHECK;
        auto issue = [&](int q) -> int {
            VFVM_HIP_TRY(hipGraphLaunch(cg->chunks[q], c.stream));
            VFVM_HIP_TRY(hipMemcpyAsync(w.h_ctrl + 64 * (q % 2), w.ctrl, CTRL_INTS * sizeof(int), hipMemcpyDeviceToHost, c.stream));
            VFVM_HIP_TRY(hipEventRecord(c.ctrl_ev[q % 2], c.stream));
            return VFVM_OK;
        };
        if (int st = issue(0)) return st;
        const int* hc = w.h_ctrl;
        for (int q = 0; q < nchunk; ++q) {
            if (q + 1 < nchunk)
                if (int st = issue(q + 1)) return st;
            VFVM_HIP_TRY(hipEventSynchronize(c.ctrl_ev[q % 2]));
            hc = w.h_ctrl + 64 * (q % 2);
            steps = hc[1];
            if (hc[0] == 0) break;
        }
        // the caller reads the last control word from slot 0 (a chunk still queued after
        // convergence leaves the control word as it is, so its readback writes the same ints)
        if (hc != w.h_ctrl) std::memcpy(w.h_ctrl, hc, CTRL_INTS * sizeof(int));
    }
    for (int j0 = 0; !cg && j0 < w.m; j0 += CHECK) {
        for (int j = j0; j < std::min(w.m, j0 + CHECK); ++j) {
            c.cur_step = j;
            if (int st = issue_step<T>(c, sy, w, V, wv, zv, aval, pc, reorth, j)) return st;
        }
        c.cur_step = -1;
        VFVM_HIP_TRY(hipMemcpyAsync(w.h_ctrl, w.ctrl, CTRL_INTS * sizeof(int), hipMemcpyDeviceToHost, c.stream));
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        steps = w.h_ctrl[1];
        if (!c.pending.empty()) {
            for (auto& p : c.pending)
                if (p.step >= steps) p.bytes = -1.0;  // launches after convergence returned at once
            if (int st = ktime_flush(c)) return st;
        }
        if (w.h_ctrl[0] == 0) break;
    }
    hipLaunchKernelGGL(k_backsolve, dim3(1), dim3(WAVE), 0, c.stream, w.S, w.L, w.ctrl);
    VFVM_HIP_TRY(hipGetLastError());
    *steps_out = steps;
    return VFVM_OK;
}

}  // namespace

SysView primal_view(Ctx& c) {
    SysView s;
    s.n = c.n; s.nnz = c.nnz;
    s.row = c.mrow.p; s.col = c.mcol.p; s.dslot = c.dslot.p;
    s.vals = c.vals.p; s.b = c.b.p;
    s.h_row = &c.h_mrow; s.h_col = &c.h_mcol;
    s.ilu = &c.ilu;
    s.flow_layout = c.flow ? 1 : 0;
    s.flow_ns = c.flow ? (int)c.fs.blk_unk.size() : 0;
    return s;
}

SysView transpose_view(Ctx& c) {
    SysView s;
    s.n = c.n; s.nnz = c.nnz;
    s.row = c.trow.p; s.col = c.tcol.p; s.dslot = c.tdslot.p;
    s.vals = c.tvals.p; s.b = c.tb.p;
    s.h_row = &c.h_trow; s.h_col = &c.h_tcol;
    s.ilu = &c.tilu;
    s.interleaved = c.t_interleaved;
    return s;
}

int solver_setup(Ctx& c, const SysView& sy, const vfvm_solve_opts& o) {
    if (o.precond == VFVM_PRECOND_ILU0) return ilu_setup(c, sy);
    return VFVM_OK;
}

int precond_apply(Ctx& c, const vfvm_solve_opts& o, const double* r_host, double* z_host) {
    const SysView sy = primal_view(c);
    const int64_t n = sy.n;
    if (o.precond < VFVM_PRECOND_NONE || o.precond > VFVM_PRECOND_BJACOBI) return set_error(VFVM_ERR_INVALID_ARGUMENT, "unknown preconditioner");
    if (o.precision != VFVM_FP64 && o.precision != VFVM_FP32_REFINED) return set_error(VFVM_ERR_INVALID_ARGUMENT, "unknown precision");
    DevBuf<double> r, z;
    VFVM_HIP_TRY(r.alloc(n));
    VFVM_HIP_TRY(z.alloc(n));
    VFVM_HIP_TRY(hipMemcpyAsync(r.p, r_host, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (o.precond == VFVM_PRECOND_ILU0) {
        if (int s = ilu_setup(c, sy)) return s;
        if (o.precision == VFVM_FP64) {
            VFVM_HIP_TRY(c.ilu64.alloc(sy.nnz));
            VFVM_HIP_TRY(c.iluf64.alloc(sy.nnz));
            if (int s = ilu_factor<double>(c, sy, c.ilu64.p, c.iluf64.p, sy.vals)) return s;
            DevBuf<double> rp_, zp_;
            VFVM_HIP_TRY(rp_.alloc(n));
            VFVM_HIP_TRY(zp_.alloc(n));
            hipLaunchKernelGGL((k_gather_perm<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, sy.ilu->perm.p, (const double*)r.p, rp_.p);
            if (int s = ilu_apply<double>(c, sy, c.ilu64.p, rp_.p, nullptr, zp_.p, nullptr)) return s;
            hipLaunchKernelGGL((k_scatter_perm<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, sy.ilu->perm.p, (const double*)zp_.p, z.p);
        } else {
            DevBuf<float> r32, z32;
            VFVM_HIP_TRY(r32.alloc(n));
            VFVM_HIP_TRY(z32.alloc(n));
            VFVM_HIP_TRY(c.ilu32.alloc(sy.nnz));
            VFVM_HIP_TRY(c.iluf32.alloc(sy.nnz));
            if (int s = ilu_factor<float>(c, sy, c.ilu32.p, c.iluf32.p, sy.vals)) return s;
            DevBuf<double> zp_;
            VFVM_HIP_TRY(zp_.alloc(n));
            hipLaunchKernelGGL((k_gather_perm<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, sy.ilu->perm.p, (const double*)r.p, zp_.p);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, (const double*)zp_.p, r32.p);
            if (int s = ilu_apply<float>(c, sy, c.ilu32.p, r32.p, nullptr, z32.p, nullptr)) return s;
            hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, (const float*)z32.p, zp_.p);
            hipLaunchKernelGGL((k_scatter_perm<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, sy.ilu->perm.p, (const double*)zp_.p, z.p);
            VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        }
    } else if (o.precond == VFVM_PRECOND_BJACOBI) {
        const bool fp32 = o.precision != VFVM_FP64;
        if (int s = bjac_setup(c, sy, fp32)) return s;
        if (!fp32) {
            Pc<double> pc;
            pc.bj = c.bjac.inv64.p; pc.nb = c.bjac.nb; pc.nv = c.g.N;
            if (int s = bjac_apply<double>(c, pc, (const double*)r.p, nullptr, z.p, nullptr)) return s;
        } else {
            DevBuf<float> r32, z32;
            VFVM_HIP_TRY(r32.alloc(n));
            VFVM_HIP_TRY(z32.alloc(n));
            Pc<float> pc;
            pc.bj = c.bjac.inv32.p; pc.nb = c.bjac.nb; pc.nv = c.g.N;
            hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, (const double*)r.p, r32.p);
            if (int s = bjac_apply<float>(c, pc, (const float*)r32.p, nullptr, z32.p, nullptr)) return s;
            hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, (const float*)z32.p, z.p);
            VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        }
    } else {
        VFVM_HIP_TRY(c.dinv64.alloc(n));
        hipLaunchKernelGGL(k_dinv, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, sy.dslot, sy.vals, c.dinv64.p,
                           o.precond == VFVM_PRECOND_JACOBI ? 1 : 0);
        hipLaunchKernelGGL(k_xupdate_set, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, r.p, c.dinv64.p, z.p);
    }
    VFVM_HIP_TRY(hipGetLastError());
    VFVM_HIP_TRY(hipMemcpyAsync(z_host, z.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    return VFVM_OK;
}

int ilu_order(Ctx& c, int32_t* perm, int32_t* color_ptr, int32_t* ncolors) {
    const SysView sy = primal_view(c);
    if (int s = ilu_setup(c, sy)) return s;
    const IluData& I = *sy.ilu;
    const int32_t nc = (int32_t)I.h_color_ptr.size() - 1;
    if (ncolors) *ncolors = nc;
    if (color_ptr) std::memcpy(color_ptr, I.h_color_ptr.data(), I.h_color_ptr.size() * sizeof(int32_t));
    if (perm) VFVM_HIP_TRY(hipMemcpy(perm, I.perm.p, sy.n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return VFVM_OK;
}

int ilu_wait_limit(Ctx& c, int32_t polls, uint64_t* lapsed) {
    IluData& I = c.ilu;
    if (lapsed) {
        *lapsed = 0;
        if (I.ready && I.tiled) {
            VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
            VFVM_HIP_TRY(hipMemcpy(lapsed, I.tsync.p + 1, sizeof(uint64_t), hipMemcpyDeviceToHost));
        }
    }
    if (polls != I.wait_limit) {
        I.wait_limit = polls;
        ++I.generation;  // captured cycles bake the limit in
    }
    return VFVM_OK;
}

int ilu_tiles(Ctx& c, int32_t* tile_ptr, int32_t* ntiles, int32_t* level) {
    const SysView sy = primal_view(c);
    if (int s = ilu_setup(c, sy)) return s;
    const IluData& I = *sy.ilu;
    const int32_t nt = I.tiled ? (int32_t)I.h_tile_row.size() - 1 : 0;
    if (ntiles) *ntiles = nt;
    if (tile_ptr && nt > 0) std::memcpy(tile_ptr, I.h_tile_row.data(), I.h_tile_row.size() * sizeof(int32_t));
    if (level) {
        for (int64_t t = 0; t < sy.n; ++t) level[t] = 0;
        for (int32_t k = 0; k < nt; ++k)
            for (int l = 0; l < ILU_MAXLEV; ++l)
                for (int32_t r = I.h_lev_row[(size_t)k * (ILU_MAXLEV + 1) + l]; r < I.h_lev_row[(size_t)k * (ILU_MAXLEV + 1) + l + 1]; ++r)
                    level[r] = l;
    }
    return VFVM_OK;
}

int spmv_apply(Ctx& c, const SysView& sy, const double* x, double* y) {
    const int lpr = pick_lpr(sy.n, sy.nnz);
    const double bytes = (double)sy.nnz * 12.0 + 4.0 * (sy.n + 1) + 16.0 * sy.n;
    return spmv_launch<double>(c, sy, lpr, sy.n, sy.vals, x, y, nullptr, nullptr, bytes);
}

int residual_norm(Ctx& c, const SysView& sy, const double* x_dev, double* rel, double* bnorm) {
    const int64_t n = sy.n;
    const int lpr = pick_lpr(n, sy.nnz);
    const int nblk_r = (int)std::min<int64_t>(1024, grid_for(n * lpr));
    VFVM_HIP_TRY(c.r64.alloc(n));
    DevBuf<double> part;
    VFVM_HIP_TRY(part.alloc(2 * (size_t)nblk_r));
    if (int st = residual_launch(c, sy, lpr, x_dev, c.r64.p, part.p, nblk_r)) return st;
    std::vector<double> h(2 * (size_t)nblk_r);
    VFVM_HIP_TRY(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
    double rr = 0.0, bb = 0.0;
    for (int k = 0; k < nblk_r; ++k) { rr += h[k]; bb += h[nblk_r + k]; }
    *bnorm = std::sqrt(bb);
    *rel = bb > 0.0 ? std::sqrt(rr) / std::sqrt(bb) : std::sqrt(rr);
    return VFVM_OK;
}

namespace {

// Device workspace of one solve: pointers into the context's buffers, or into a caller-provided
// region (the no-allocation entry point).  In the latter mode the solve launches eagerly (no
// graph capture or instantiation, no events) and supports FP64 with Jacobi or no preconditioner.
struct GmresBufs {
    bool noalloc = false;
    // an inner solve of an outer iteration: x0 = 0 (the first residual is b), the FP32 matrix
    // and the preconditioner kept from the previous solve of the same values, and no residual
    // after the last allowed cycle (the caller takes the iterate as it is)
    bool x0_zero = false, reuse_setup = false, no_final_residual = false;
    double* S = nullptr;
    int* ctrl = nullptr;
    double *part = nullptr, *r64 = nullptr, *dinv64 = nullptr, *w64 = nullptr, *z64 = nullptr, *V64 = nullptr,
           *avals64 = nullptr, *bp64 = nullptr, *xp64 = nullptr;
    float *V32 = nullptr, *w32 = nullptr, *avals32 = nullptr;
};

// sizes of the workspace (entries) for a solve of n rows, nnz entries, restart m
struct GmresSizes {
    int64_t ldv, nwav, nblk, nblk_r, state, part;
};
GmresSizes gmres_sizes(int64_t n, int64_t nnz, int m) {
    GmresSizes z;
    z.nblk = std::min<int64_t>(1024, std::max<int64_t>(1, (n + 4 * BLK - 1) / (4 * BLK)));
    // multi-dot: up to 1024 waves per group of basis vectors, 8+ rows per lane
    constexpr int64_t mdot_waves = 1024;
    z.nwav = std::min<int64_t>(mdot_waves, std::max<int64_t>(1, (n + WAVE * 8 - 1) / (WAVE * 8)));
    z.nwav = (z.nwav + BLK / WAVE - 1) / (BLK / WAVE) * (BLK / WAVE);
    z.ldv = (n + 63) / 64 * 64;
    const int lpr = pick_lpr(n, nnz);
    z.nblk_r = std::min<int64_t>(1024, grid_for(n * lpr));
    StateLayout L;
    L.m = m;
    z.state = L.total() + 64;
    z.part = (int64_t)(MAX_RESTART + 1) * z.nwav + z.nblk + 2 * z.nblk_r + 16;
    return z;
}

int gmres_run(Ctx& c, const SysView& sy0, const vfvm_solve_opts& o, double* x0, vfvm_solve_stats& st, GmresBufs& B);

}  // namespace

int gmres_solve(Ctx& c, const SysView& sy0, const vfvm_solve_opts& o, double* x0, vfvm_solve_stats& st) {
    std::memset(&st, 0, sizeof(st));
    if (o.restart <= 0 || o.restart > MAX_RESTART) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_RESTART, "GMRES: restart must be in [1, 256]");
    const bool mixed = o.precision == VFVM_FP32_REFINED;
    const int64_t n = sy0.n;
    const GmresSizes z = gmres_sizes(n, sy0.nnz, o.restart);
    GmresBufs B;
    VFVM_HIP_TRY(c.scal.alloc((size_t)z.state));
    B.S = c.scal.p;
    B.ctrl = reinterpret_cast<int*>(c.scal.p + (z.state - 64));
    VFVM_HIP_TRY(c.part.alloc((size_t)z.part));
    B.part = c.part.p;
    VFVM_HIP_TRY(c.r64.alloc(n));
    VFVM_HIP_TRY(c.dinv64.alloc(n));
    VFVM_HIP_TRY(c.w64.alloc(n));
    VFVM_HIP_TRY(c.z64.alloc(n));
    B.r64 = c.r64.p; B.dinv64 = c.dinv64.p; B.w64 = c.w64.p; B.z64 = c.z64.p;
    if (!mixed) {
        VFVM_HIP_TRY(c.V64.alloc((size_t)z.ldv * (o.restart + 1)));
        B.V64 = c.V64.p;
    } else {
        VFVM_HIP_TRY(c.V32.alloc((size_t)z.ldv * (o.restart + 1)));
        VFVM_HIP_TRY(c.w32.alloc(2 * (size_t)n));
        B.V32 = c.V32.p; B.w32 = c.w32.p;
    }
    return gmres_run(c, sy0, o, x0, st, B);
}

int gmres_solve_inner(Ctx& c, const SysView& sy0, const vfvm_solve_opts& o, double* x0, vfvm_solve_stats& st, bool reuse_setup) {
    std::memset(&st, 0, sizeof(st));
    if (o.restart <= 0 || o.restart > MAX_RESTART) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_RESTART, "GMRES: restart must be in [1, 256]");
    const bool mixed = o.precision == VFVM_FP32_REFINED;
    const int64_t n = sy0.n;
    const GmresSizes z = gmres_sizes(n, sy0.nnz, o.restart);
    GmresBufs B;
    B.x0_zero = true;
    B.reuse_setup = reuse_setup;
    B.no_final_residual = true;
    VFVM_HIP_TRY(c.scal.alloc((size_t)z.state));
    B.S = c.scal.p;
    B.ctrl = reinterpret_cast<int*>(c.scal.p + (z.state - 64));
    VFVM_HIP_TRY(c.part.alloc((size_t)z.part));
    B.part = c.part.p;
    VFVM_HIP_TRY(c.r64.alloc(n));
    VFVM_HIP_TRY(c.dinv64.alloc(n));
    VFVM_HIP_TRY(c.w64.alloc(n));
    VFVM_HIP_TRY(c.z64.alloc(n));
    B.r64 = c.r64.p; B.dinv64 = c.dinv64.p; B.w64 = c.w64.p; B.z64 = c.z64.p;
    if (!mixed) {
        VFVM_HIP_TRY(c.V64.alloc((size_t)z.ldv * (o.restart + 1)));
        B.V64 = c.V64.p;
    } else {
        VFVM_HIP_TRY(c.V32.alloc((size_t)z.ldv * (o.restart + 1)));
        VFVM_HIP_TRY(c.w32.alloc(2 * (size_t)n));
        B.V32 = c.V32.p; B.w32 = c.w32.p;
    }
    return gmres_run(c, sy0, o, x0, st, B);
}

// ---------------------------------------------------------------- no-allocation solve
int64_t gmres_noalloc_entries(int64_t n, int64_t nnz, int m, bool jacobi, bool interleaved) {
    const GmresSizes z = gmres_sizes(n, nnz, m);
    auto up = [](int64_t k) { return (k + 31) / 32 * 32; };  // 256-byte aligned pieces
    int64_t e = up(z.state) + up(z.part) + 4 * up(n) + up(z.ldv * (m + 1));
    if (jacobi) e += up(nnz);
    if (interleaved) e += 2 * up(n);
    return e;
}

int gmres_solve_noalloc(Ctx& c, const SysView& sy0, const vfvm_solve_opts& o, double* x0, vfvm_solve_stats& st, double* scratch,
                        int64_t scratch_entries) {
    std::memset(&st, 0, sizeof(st));
    if (o.restart <= 0 || o.restart > MAX_RESTART) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_RESTART, "GMRES: restart must be in [1, 256]");
    if (o.precision != VFVM_FP64 || (o.precond != VFVM_PRECOND_NONE && o.precond != VFVM_PRECOND_JACOBI))
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "no-allocation solve: FP64 with the Jacobi preconditioner or none only");
    const int64_t n = sy0.n;
    const bool jac = o.precond == VFVM_PRECOND_JACOBI;
    const bool il = sy0.flow_layout != 0 && o.precond == VFVM_PRECOND_NONE;
    const int64_t need = gmres_noalloc_entries(n, sy0.nnz, o.restart, jac, il);
    if (need > scratch_entries)
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "no-allocation solve: the system needs " + std::to_string(need * 8) +
                                                        " bytes of workspace, the scratch array holds " + std::to_string(scratch_entries * 8) +
                                                        " (n (restart + 5) + nnz doubles and a few MB of partials)");
    const GmresSizes z = gmres_sizes(n, sy0.nnz, o.restart);
    auto up = [](int64_t k) { return (k + 31) / 32 * 32; };
    double* p = scratch;
    auto take = [&](int64_t k) { double* q = p; p += up(k); return q; };
    GmresBufs B;
    B.noalloc = true;
    B.S = take(z.state);
    B.ctrl = reinterpret_cast<int*>(B.S + (z.state - 64));
    B.part = take(z.part);
    B.r64 = take(n); B.dinv64 = take(n); B.w64 = take(n); B.z64 = take(n);
    B.V64 = take(z.ldv * (o.restart + 1));
    if (jac) B.avals64 = take(sy0.nnz);
    if (il) { B.bp64 = take(n); B.xp64 = take(n); }
    // no events: kernel timing is off for the call
    const uint32_t mask = c.timing_mask;
    c.timing_mask = 0;
    int r = gmres_run(c, sy0, o, x0, st, B);
    c.timing_mask = mask;
    return r;
}

namespace {

int gmres_run(Ctx& c, const SysView& sy0, const vfvm_solve_opts& o, double* x0, vfvm_solve_stats& st, GmresBufs& B) {
    std::memset(&st, 0, sizeof(st));
    if (!(o.tol > 0.0)) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_TOLERANCE, "GMRES: tolerance must be positive");
    if (o.restart <= 0 || o.restart > MAX_RESTART) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_RESTART, "GMRES: restart must be in [1, 256]");
    if (o.max_outer <= 0) return set_error(VFVM_ERR_LINEAR_NON_POSITIVE_MAX_OUTER, "GMRES: max_outer must be positive");
    if (o.precond < VFVM_PRECOND_NONE || o.precond > VFVM_PRECOND_BJACOBI) return set_error(VFVM_ERR_INVALID_ARGUMENT, "GMRES: unknown preconditioner");
    if (o.precision != VFVM_FP64 && o.precision != VFVM_FP32_REFINED) return set_error(VFVM_ERR_INVALID_ARGUMENT, "GMRES: unknown precision");
    if (o.ortho < VFVM_ORTHO_AUTO || o.ortho > VFVM_ORTHO_CGS2) return set_error(VFVM_ERR_INVALID_ARGUMENT, "GMRES: unknown orthogonalisation");
    const bool mixed = o.precision == VFVM_FP32_REFINED;
    const bool reorth = o.ortho == VFVM_ORTHO_CGS2 || (o.ortho == VFVM_ORTHO_AUTO && !mixed);
    const bool ilu = o.precond == VFVM_PRECOND_ILU0;
    const bool jac = o.precond == VFVM_PRECOND_JACOBI;
    const bool bjac = o.precond == VFVM_PRECOND_BJACOBI;
    // With ILU(0) the whole solve runs in the colour-major space of the factor: A' = P A P^T
    // shares the factor's index arrays, b' = P b, x' = P x, so the triangular sweeps read and
    // write their rows in place, with no indirection.
    SysView sy = sy0;
    double* x = x0;
    // the coupled flow system with block Jacobi (or none) runs with vertex-interleaved
    // vectors: P A P^T with the four unknowns of a vertex adjacent
    const bool il = sy0.flow_layout != 0 && (bjac || o.precond == VFVM_PRECOND_NONE);
    if (il) {
        if (!B.noalloc) {
            VFVM_HIP_TRY(c.bp64.alloc(sy0.n));
            VFVM_HIP_TRY(c.xp64.alloc(sy0.n));
            B.bp64 = c.bp64.p;
            B.xp64 = c.xp64.p;
        }
        sy.flow_layout = 2;
        sy.b = B.bp64;
        x = B.xp64;
    } else if (sy0.flow_layout && ilu) {
        sy.flow_layout = 0;
    }
    if (ilu) {
        if (int s = ilu_setup(c, sy0)) return s;
        const IluData& I = *sy0.ilu;
        VFVM_HIP_TRY(c.ap64.alloc(sy0.nnz));
        VFVM_HIP_TRY(c.bp64.alloc(sy0.n));
        VFVM_HIP_TRY(c.xp64.alloc(sy0.n));
        sy.row = I.prp.p;
        sy.col = I.pcol.p;
        sy.dslot = I.pdiag.p;
        sy.vals = c.ap64.p;
        sy.b = c.bp64.p;
        x = c.xp64.p;
    }
    // the block SpMV of a generic CSR: its blocks' largest entry count, from the host row
    // pointers of the rows the solve multiplies by
    if (!sy.flow_layout && !sy.interleaved) {
        const std::vector<int32_t>* hr = ilu ? &sy0.ilu->h_prp : sy0.h_row;
        sy.spmv_emax = (hr && (int64_t)hr->size() == sy.n + 1) ? spmv_block_emax(*hr) : 0;
    }
    const int64_t n = sy.n;
    const int m = o.restart;
    const GmresSizes z = gmres_sizes(n, sy.nnz, m);
    Work w;
    w.n = n;
    w.nnz = sy.nnz;
    w.m = m;
    w.nblk = (int)z.nblk;
    w.rows_per_blk = ((n + w.nblk - 1) / w.nblk + 3) / 4 * 4;
    w.nwav = (int)z.nwav;
    w.rows_per_wave = ((n + w.nwav - 1) / w.nwav + 3) / 4 * 4;
    w.ldv = z.ldv;
    w.L.m = m;
    const int lpr = pick_lpr(n, sy.nnz);
    const int nblk_r = (int)z.nblk_r;
    w.S = B.S;
    w.ctrl = B.ctrl;
    w.part = B.part;
    if (!c.h_pinned) return set_error(VFVM_ERR_HIP, "GMRES: the context has no pinned readback buffer");
    w.h_ctrl = reinterpret_cast<int*>(c.h_pinned);
    double* const part = B.part;
    double* const r64 = B.r64;
    double* const dinv64 = B.dinv64;
    double* const w64 = B.w64;
    double* const z64 = B.z64;

    HipEvents ev;
    const auto host_t0 = std::chrono::steady_clock::now();
    auto host_t1 = host_t0;
    if (!B.noalloc) {
        VFVM_HIP_TRY(ev.create());
        VFVM_HIP_TRY(hipEventRecord(ev.e[0], c.stream));
    }

    // permuted system and preconditioner set-up (inside the timed solve)
    if (il) {
        const int S = 4 + sy0.flow_ns;
        hipLaunchKernelGGL((k_to_interleaved<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, c.g.N, S, sy0.b, B.bp64);
        hipLaunchKernelGGL((k_to_interleaved<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, c.g.N, S, (const double*)x0, B.xp64);
    }
    if (ilu) {
        const IluData& I = *sy0.ilu;
        hipLaunchKernelGGL((k_ilu_gather<double>), dim3(grid_for(sy0.nnz)), dim3(BLK), 0, c.stream, sy0.nnz, I.src.p, sy0.vals, c.ap64.p);
        hipLaunchKernelGGL((k_gather_perm<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, I.perm.p, sy0.b, c.bp64.p);
        hipLaunchKernelGGL((k_gather_perm<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, I.perm.p, (const double*)x0, c.xp64.p);
    }
    hipLaunchKernelGGL(k_dinv, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, sy.dslot, sy.vals, dinv64, jac ? 1 : 0);
    const double* dinv_x = jac ? dinv64 : nullptr;
    const void* aval = nullptr;
    const void* lu = nullptr;
    if (!mixed) {
        if (jac) {
            if (!B.noalloc) {
                VFVM_HIP_TRY(c.avals64.alloc(sy.nnz));
                B.avals64 = c.avals64.p;
            }
            if (!B.reuse_setup)
                hipLaunchKernelGGL((k_scale_cols<double>), dim3(grid_for(sy.nnz)), dim3(BLK), 0, c.stream, sy.nnz, sy.col, sy.vals, dinv64, B.avals64);
            aval = B.avals64;
        } else {
            aval = sy.vals;
        }
        if (ilu) {
            VFVM_HIP_TRY(c.ilu64.alloc(sy.nnz));
            VFVM_HIP_TRY(c.iluf64.alloc(sy.nnz));
            if (!B.reuse_setup)
                if (int s = ilu_factor<double>(c, sy0, c.ilu64.p, c.iluf64.p, sy0.vals)) return s;
            lu = c.ilu64.p;
        }
    } else {
        // FP32 copy of A D^-1 (Jacobi) or of A (ILU(0), none); dinv = 1 unless Jacobi
        VFVM_HIP_TRY(c.avals32.alloc(sy.nnz));
        B.avals32 = c.avals32.p;
        if (!B.reuse_setup)
            hipLaunchKernelGGL((k_scale_cols<float>), dim3(grid_for(sy.nnz)), dim3(BLK), 0, c.stream, sy.nnz, sy.col, sy.vals, dinv64, B.avals32);
        aval = B.avals32;
        if (ilu) {
            VFVM_HIP_TRY(c.ilu32.alloc(sy.nnz));
            VFVM_HIP_TRY(c.iluf32.alloc(sy.nnz));
            if (!B.reuse_setup)
                if (int s = ilu_factor<float>(c, sy0, c.ilu32.p, c.iluf32.p, sy0.vals)) return s;
            lu = c.ilu32.p;
        }
    }
    if (bjac && !B.reuse_setup)
        if (int s = bjac_setup(c, sy, mixed)) return s;
    Pc<double> pc64;
    Pc<float> pc32;
    pc64.lu = (const double*)(mixed ? nullptr : lu);
    pc32.lu = (const float*)(mixed ? lu : nullptr);
    if (bjac) {
        pc64.bj = c.bjac.inv64.p; pc64.nb = c.bjac.nb; pc64.nv = c.g.N; pc64.il = il || sy0.interleaved;
        pc32.bj = c.bjac.inv32.p; pc32.nb = c.bjac.nb; pc32.nv = c.g.N; pc32.il = il || sy0.interleaved;
    }
    VFVM_HIP_TRY(hipGetLastError());
    if (!B.noalloc) VFVM_HIP_TRY(hipEventRecord(ev.e[1], c.stream));
    else {
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        host_t1 = std::chrono::steady_clock::now();
    }

    double* part_r = part + (size_t)(MAX_RESTART + 1) * w.nwav + w.nblk;
    int total_steps = 0, cycles = 0, refinements = 0;
    double rel = 0.0, bnorm = 0.0;
    int status = VFVM_ERR_LINEAR_DID_NOT_CONVERGE;
    const int max_ref = o.max_refine > 0 ? o.max_refine : o.max_outer;
    const bool graphs = !B.noalloc;
    for (;;) {
        // true FP64 residual r = b - A x, beta = ||r|| (r = b for a zero initial guess)
        if (B.x0_zero && cycles == 0)
            hipLaunchKernelGGL(k_residual_zero, dim3(nblk_r), dim3(BLK), 0, c.stream, n, (const double*)sy.b, r64, part_r, nblk_r);
        else if (int s = residual_launch(c, sy, lpr, x, r64, part_r, nblk_r)) return s;
        hipLaunchKernelGGL(k_cycle_start, dim3(1), dim3(BLK), 0, c.stream, nblk_r, part_r, w.S, w.L, w.ctrl, 0.0, 0);
        VFVM_HIP_TRY(hipMemcpyAsync(c.h_pinned + 8, w.S + w.L.misc(), 8 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        const double beta = c.h_pinned[8 + 0];
        bnorm = c.h_pinned[8 + 5];
        rel = bnorm > 0.0 ? beta / bnorm : beta;
        if (!std::isfinite(beta)) { status = VFVM_ERR_LINEAR_NON_FINITE; break; }
        if (rel <= o.tol || beta == 0.0) { status = VFVM_OK; break; }
        if (cycles >= o.max_outer || (mixed && refinements >= max_ref)) break;
        // Arnoldi target: FP64 stops on the estimate of ||b - A x||; the FP32 inner solve
        // reduces the current residual by inner_tol
        const double target_abs = mixed ? std::max(o.inner_tol, 1e-7) * beta : 0.999 * o.tol * bnorm;
        hipLaunchKernelGGL(k_cycle_start, dim3(1), dim3(BLK), 0, c.stream, nblk_r, part_r, w.S, w.L, w.ctrl, target_abs, 1);
        int steps = 0;
        if (!mixed) {
            hipLaunchKernelGGL((k_scale<double, double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, r64, B.V64, w.S, w.L.misc() + 1, w.ctrl);
            if (int s = arnoldi_cycle<double>(c, sy, w, B.V64, w64, z64, (const double*)aval, pc64, reorth, graphs, &steps)) return s;
            if (w.h_ctrl[2]) { status = VFVM_ERR_LINEAR_NON_FINITE; break; }
            // x += M^-1 (V y)
            hipLaunchKernelGGL((k_combine<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, B.V64, w.ldv, w.S, w.L, w.ctrl, w64);
            if (lu) {
                if (int s = ilu_apply<double>(c, sy, (const double*)lu, (const double*)w64, nullptr, z64, (const int*)nullptr)) return s;
                hipLaunchKernelGGL(k_xupdate, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, z64, (const double*)nullptr, x);
            } else if (bjac) {
                if (int s = bjac_apply<double>(c, pc64, (const double*)w64, nullptr, z64, (const int*)nullptr)) return s;
                hipLaunchKernelGGL(k_xupdate, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, z64, (const double*)nullptr, x);
            } else {
                hipLaunchKernelGGL(k_xupdate, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, w64, dinv_x, x);
            }
        } else {
            // inner FP32 solve of A M^-1 u = r from u = 0, V0 = r / beta in FP32
            hipLaunchKernelGGL((k_scale<double, float>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, r64, B.V32, w.S, w.L.misc() + 1, w.ctrl);
            if (int s = arnoldi_cycle<float>(c, sy, w, B.V32, B.w32, B.w32 + n, (const float*)aval, pc32, reorth, graphs, &steps)) return s;
            ++refinements;
            if (w.h_ctrl[2]) { status = VFVM_ERR_LINEAR_NON_FINITE; break; }
            // x += M^-1 (V y), V y accumulated in FP64
            hipLaunchKernelGGL((k_combine<float>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, B.V32, w.ldv, w.S, w.L, w.ctrl, z64);
            if (lu) {
                float* t32 = B.w32;
                float* z32 = B.w32 + n;
                hipLaunchKernelGGL((k_scale<double, float>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, z64, t32, w.S, w.L.misc() + 4, (const int*)nullptr);
                if (int s = ilu_apply<float>(c, sy, (const float*)lu, (const float*)t32, nullptr, z32, (const int*)nullptr)) return s;
                hipLaunchKernelGGL((k_scale<float, double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, z32, z64, w.S, w.L.misc() + 4, (const int*)nullptr);
                hipLaunchKernelGGL(k_xupdate, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, z64, (const double*)nullptr, x);
            } else if (bjac) {
                // the FP64 inverse blocks on the FP64 combination
                if (int s = bjac_apply<double>(c, pc64, (const double*)z64, nullptr, w64, (const int*)nullptr)) return s;
                hipLaunchKernelGGL(k_xupdate, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, w64, (const double*)nullptr, x);
            } else {
                hipLaunchKernelGGL(k_xupdate, dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, z64, dinv_x, x);
            }
        }
        VFVM_HIP_TRY(hipGetLastError());
        total_steps += steps;
        ++cycles;
        if (B.no_final_residual && (cycles >= o.max_outer || (mixed && refinements >= max_ref))) {
            rel = -1.0;  // not evaluated
            break;
        }
    }
    if (ilu) hipLaunchKernelGGL((k_scatter_perm<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, n, sy0.ilu->perm.p, (const double*)x, x0);
    if (il)
        hipLaunchKernelGGL((k_from_interleaved<double>), dim3(grid_for(n)), dim3(BLK), 0, c.stream, c.g.N, 4 + sy0.flow_ns, (const double*)x, x0);
    float ms_all = 0.f, ms_setup = 0.f;
    if (!B.noalloc) {
        VFVM_HIP_TRY(hipEventRecord(ev.e[2], c.stream));
        VFVM_HIP_TRY(hipEventSynchronize(ev.e[2]));
        VFVM_HIP_TRY(hipEventElapsedTime(&ms_all, ev.e[0], ev.e[2]));
        VFVM_HIP_TRY(hipEventElapsedTime(&ms_setup, ev.e[0], ev.e[1]));
    } else {
        VFVM_HIP_TRY(hipStreamSynchronize(c.stream));
        const auto host_t2 = std::chrono::steady_clock::now();
        ms_all = (float)(1e3 * std::chrono::duration<double>(host_t2 - host_t0).count());
        ms_setup = (float)(1e3 * std::chrono::duration<double>(host_t1 - host_t0).count());
    }
    if (int s = ktime_flush(c)) return s;
    st.iterations = total_steps;
    st.restarts = cycles;
    st.refinements = refinements;
    st.rel_residual = rel;
    st.seconds = ms_all * 1e-3;
    st.setup_seconds = ms_setup * 1e-3;
    st.status = status;
    if (status != VFVM_OK) {
        char buf[160];
        snprintf(buf, sizeof(buf), "GMRES did not converge: relative residual %.3e after %d steps", rel, total_steps);
        return set_error(status, buf);
    }
    return VFVM_OK;
}

}  // namespace

}  // namespace vfvm

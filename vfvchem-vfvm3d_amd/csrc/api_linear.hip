// C ABI: linear systems assembled by the caller (LinearBase, src/linear/linear_base.rs:5-10,
// with the GMRES of LinearGmres, src/linear/linear_gmres.rs:53-84).
//
// update_sparse takes the (row, col) sequence of a triplet list, merges duplicates into a
// row-sorted CSR pattern and keeps, for every CSR slot, the triplets that sum into it in
// list order (the argsort of linear_gmres.rs:59); build_matrix then sums the values of a
// triplet list with the same sequence into the slots on the device (linear_gmres.rs:71:
// duplicates summed, the first triplet of a slot first); solve_matrix runs the device GMRES
// with the caller's right-hand side and initial guess (linear_gmres.rs:76-84).  The system
// replaces the context's operator-assembled one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace vfvm;

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

// vals[s] = sum of the triplet values of slot s, in triplet-list order
__global__ void __launch_bounds__(BLK) k_trip_sum(int64_t nnz, const int64_t* __restrict__ sptr, const int64_t* __restrict__ trip,
                                                  const double* __restrict__ tv, double* __restrict__ vals) {
    const int64_t s = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (s >= nnz) return;
    const int64_t k0 = sptr[s], k1 = sptr[s + 1];
    double v = tv[trip[k0]];
    for (int64_t k = k0 + 1; k < k1; ++k) v += tv[trip[k]];
    vals[s] = v;
}

int linear_build_impl(Ctx& c, int64_t ntrip, const double* tv_dev) {
    if (!c.lin.active) return set_error(VFVM_ERR_INVALID_ARGUMENT, "build_matrix: no triplet pattern (call vfvm_linear_update_sparse)");
    if (ntrip != c.lin.ntrip)
        return set_error(VFVM_ERR_PATTERN_MISMATCH, "build_matrix: " + std::to_string(ntrip) + " triplets, the pattern was built from " +
                                                        std::to_string(c.lin.ntrip));
    if (c.nnz > 0)
        hipLaunchKernelGGL(k_trip_sum, dim3(grid_for(c.nnz)), dim3(BLK), 0, c.stream, c.nnz, (const int64_t*)c.lin.sptr.p,
                           (const int64_t*)c.lin.trip.p, tv_dev, c.vals.p);
    VFVM_HIP_TRY(hipGetLastError());
    c.has_matrix = true;
    return VFVM_OK;
}

}  // namespace

extern "C" int vfvm_linear_update_sparse(vfvm_ctx* ctx, int64_t mat_size, int64_t ntrip, const int64_t* rows, const int64_t* cols,
                                         int64_t* nnz_out) {
    CTX_GUARD(ctx);
    if (mat_size <= 0 || mat_size > INT32_MAX || ntrip < 0 || (ntrip > 0 && (!rows || !cols)))
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "update_sparse: invalid size or NULL triplet arrays");
    for (int64_t k = 0; k < ntrip; ++k)
        if (rows[k] < 0 || rows[k] >= mat_size || cols[k] < 0 || cols[k] >= mat_size)
            return set_error(VFVM_ERR_INVALID_ARGUMENT, "update_sparse: triplet " + std::to_string(k) + " lies outside the matrix");
    // argsort by (row, col), stable: a slot's triplets stay in list order
    std::vector<int64_t> ord(ntrip);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
        return rows[a] != rows[b] ? rows[a] < rows[b] : cols[a] < cols[b];
    });
    std::vector<int32_t> rp(mat_size + 1, 0), cl;
    std::vector<int64_t> sptr(1, 0);
    cl.reserve(ntrip);
    for (int64_t q = 0; q < ntrip; ++q) {
        const int64_t k = ord[q];
        const bool fresh = q == 0 || rows[k] != rows[ord[q - 1]] || cols[k] != cols[ord[q - 1]];
        if (fresh) {
            if (q > 0) sptr.push_back(q);
            cl.push_back((int32_t)cols[k]);
            rp[rows[k] + 1]++;
        }
    }
    if (ntrip > 0) sptr.push_back(ntrip);
    if ((int64_t)cl.size() > INT32_MAX) return set_error(VFVM_ERR_INVALID_ARGUMENT, "update_sparse: pattern exceeds 32-bit indices");
    for (int64_t i = 0; i < mat_size; ++i) rp[i + 1] += rp[i];
    std::vector<int32_t> dslot(mat_size, -1);
    for (int64_t i = 0; i < mat_size; ++i)
        for (int32_t s = rp[i]; s < rp[i + 1]; ++s)
            if (cl[s] == (int32_t)i) dslot[i] = s;
    // the context's system becomes this matrix
    c.has_pattern = false; c.has_matrix = false; c.ilu.ready = false; c.has_transpose = false; c.tilu.ready = false;
    c.tape.valid = false; c.flow = false; c.unk = -1;
    c.pattern_ops.clear();
    c.n = mat_size;
    c.nnz = (int64_t)cl.size();
    c.h_mrow = rp;
    c.h_mcol = cl;
    hipStream_t s = c.stream;
    VFVM_HIP_TRY(c.mrow.upload(rp.data(), rp.size(), s));
    VFVM_HIP_TRY(c.mcol.upload(cl.data(), std::max<size_t>(cl.size(), 1), s));
    VFVM_HIP_TRY(c.dslot.upload(dslot.data(), dslot.size(), s));
    VFVM_HIP_TRY(c.vals.alloc(std::max<int64_t>(c.nnz, 1)));
    VFVM_HIP_TRY(c.b.alloc(mat_size));
    VFVM_HIP_TRY(c.lin.sptr.upload(sptr.data(), sptr.size(), s));
    VFVM_HIP_TRY(c.lin.trip.upload(ord.data(), std::max<size_t>(ord.size(), 1), s));
    VFVM_HIP_TRY(c.lin.tv.alloc(std::max<int64_t>(ntrip, 1)));
    if (int st = sync_err(c)) return st;
    c.lin.active = true;
    c.lin.ntrip = ntrip;
    c.has_pattern = true;
    if (nnz_out) *nnz_out = c.nnz;
    return clear_error();
}

extern "C" int vfvm_linear_build_matrix(vfvm_ctx* ctx, int64_t ntrip, const double* vals) {
    CTX_GUARD(ctx);
    if (ntrip > 0 && !vals) return set_error(VFVM_ERR_INVALID_ARGUMENT, "build_matrix: vals is NULL");
    if (c.lin.active && ntrip == c.lin.ntrip && ntrip > 0)
        VFVM_HIP_TRY(hipMemcpyAsync(c.lin.tv.p, vals, ntrip * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (int st = linear_build_impl(c, ntrip, c.lin.tv.p)) return st;
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_linear_build_matrix_device(vfvm_ctx* ctx, int64_t ntrip, const double* vals_dev) {
    CTX_GUARD(ctx);
    if (ntrip > 0 && !vals_dev) return set_error(VFVM_ERR_INVALID_ARGUMENT, "build_matrix: vals is NULL");
    if (int st = linear_build_impl(c, ntrip, vals_dev)) return st;
    if (int st = sync_err(c)) return st;
    return clear_error();
}

extern "C" int vfvm_linear_solve_matrix(vfvm_ctx* ctx, const vfvm_solve_opts* opts, const double* b, const double* x_init, double* x,
                                        vfvm_solve_stats* stats) {
    CTX_GUARD(ctx);
    if (!opts || !b || !x) return set_error(VFVM_ERR_INVALID_ARGUMENT, "solve_matrix: NULL argument");
    if (!c.lin.active || !c.has_matrix) return set_error(VFVM_ERR_INVALID_ARGUMENT, "solve_matrix: no matrix built (vfvm_linear_build_matrix)");
    VFVM_HIP_TRY(hipMemcpyAsync(c.b.p, b, c.n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    VFVM_HIP_TRY(c.x64.alloc(c.n));
    if (x_init) VFVM_HIP_TRY(hipMemcpyAsync(c.x64.p, x_init, c.n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    else VFVM_HIP_TRY(hipMemsetAsync(c.x64.p, 0, c.n * sizeof(double), c.stream));
    vfvm_solve_stats st;
    int r = gmres_solve(c, primal_view(c), *opts, c.x64.p, st);
    if (stats) *stats = st;
    VFVM_HIP_TRY(hipMemcpyAsync(x, c.x64.p, c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (int s2 = sync_err(c)) return s2;
    if (r != VFVM_OK) return r;
    return clear_error();
}

extern "C" int vfvm_linear_solve_matrix_device(vfvm_ctx* ctx, const vfvm_solve_opts* opts, const double* b_dev, double* x_dev,
                                               vfvm_solve_stats* stats) {
    CTX_GUARD(ctx);
    if (!opts || !b_dev || !x_dev) return set_error(VFVM_ERR_INVALID_ARGUMENT, "solve_matrix: NULL argument");
    if (!c.lin.active || !c.has_matrix) return set_error(VFVM_ERR_INVALID_ARGUMENT, "solve_matrix: no matrix built (vfvm_linear_build_matrix)");
    VFVM_HIP_TRY(hipMemcpyAsync(c.b.p, b_dev, c.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    vfvm_solve_stats st;
    int r = gmres_solve(c, primal_view(c), *opts, x_dev, st);
    if (stats) *stats = st;
    if (int s2 = sync_err(c)) return s2;
    if (r != VFVM_OK) return r;
    return clear_error();
}

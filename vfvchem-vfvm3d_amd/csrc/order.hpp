// Multicolour ILU(0) ordering of a CSR pattern (host side, no HIP).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace vfvm {

// Rows sorted by (colour, spatial key): permuted row t is row perm[t] of A.  The colours are a
// greedy colouring of the symmetrised pattern graph in smallest-last order, so no two rows of
// one colour share an entry of A or A^T.  Inside a colour the rows follow the Morton (Z-order)
// code of their vertex, so that rows that are neighbours in space are neighbours in memory
// (the colour classes alone follow the vertex numbering, a lexicographic lattice order).
//
// Colour-major CSR of A' = P A P^T: prp / pcol (columns permuted, ascending), src[k] = CSR slot
// of A of entry k, pdiag[t] = entry of the diagonal.  The ILU(0) factor is kept split by
// triangle: L entries of row t are lpos[lptr[t] .. lptr[t+1]) (positions in the colour-major
// CSR, ascending), U entries upos[uptr[t] .. uptr[t+1]); each triangle's columns are lcol /
// ucol.  nnz = |L| + |U| + n.
struct IluOrder {
    std::vector<int32_t> perm, color_ptr;
    std::vector<int32_t> prp, pcol, src, pdiag;
    std::vector<int32_t> lptr, lpos, lcol, uptr, upos, ucol;
};

// n rows, CSR (rp, cl) with a diagonal entry in every row.  pos: 3 * nvert vertex coordinates
// (AoS) or null; row r belongs to vertex r % nvert when nvert divides n (a coupled system's
// unknowns are blocks of nvert rows), else the rows keep their index order inside a colour.
// Returns 0, or a non-zero code with err set.
int ilu_order_build(int64_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& cl, const double* pos,
                    int64_t nvert, IluOrder& out, std::string& err);

// 63-bit Morton code of a point of the box [lo, hi] (21 bits per axis).
uint64_t morton3(const double* p, const double* lo, const double* hi);

}  // namespace vfvm

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
    // Tiled order (empty tile_row: the plain multicolour order above).  The vertices are cut
    // into tiles (pencils along the longest axis, ILU_TILE_W mean spacings wide), the tiles
    // coloured so that no two tiles of one colour touch, and the rows of a tile ordered by
    // levels: a greedy colouring of the tile's own graph.  Rows are sorted by (tile colour,
    // tile, level, Morton code): color_ptr then bounds the tile colours, the rows of tile k are
    // [tile_row[k], tile_row[k+1]) and the tiles of colour q are [tc_ptr[q], tc_ptr[q+1]).
    // lev_row[k (ILU_MAXLEV + 1) + l] is the first row of level l of tile k (the entry
    // ILU_MAXLEV is the tile's end).  fac_rows lists the rows of every (tile colour, level) in
    // that order, group g being [fac_ptr[g], fac_ptr[g+1]).  rmax: rows of the largest tile;
    // emax_l / emax_u: L / U entries of the fullest tile.
    std::vector<int32_t> tile_row, tc_ptr, lev_row, fac_rows, fac_ptr;
    int32_t nlev = 0, rmax = 0, emax_l = 0, emax_u = 0;
};

// tiled order limits: levels per tile, tile colours, rows per tile (and the target), tile width in
// spacings
constexpr int ILU_MAXLEV = 16;
constexpr int ILU_MAXTC = 8;
constexpr int ILU_TILE_RMAX = 512;
constexpr int ILU_TILE_TARGET = 448;
constexpr int ILU_TILE_W = 2;

// n rows, CSR (rp, cl) with a diagonal entry in every row.  pos: 3 * nvert vertex coordinates
// (AoS) or null; row r belongs to vertex r % nvert when nvert divides n (a coupled system's
// unknowns are blocks of nvert rows), else the rows keep their index order inside a colour.
// tiled: the tiled order when n == nvert with coordinates and the tiles fit the limits above
// (else the multicolour order).  Returns 0, or a non-zero code with err set.
int ilu_order_build(int64_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& cl, const double* pos,
                    int64_t nvert, IluOrder& out, std::string& err, bool tiled = true);

// Roles of the one-launch sweeps of a tiled order: role records of TILE_ROLE_INTS ints {r0, r1,
// L entries begin, end, U entries begin, end, tile, kind (1 forward, 2 backward, 3 both),
// dependencies begin, end, 0 ...} and the dependency lists ((tile << 1) | 1 when only that
// tile's forward sweep is needed, tile << 1 when its backward sweep is).  Roles: the forward
// sweeps by tile colour ascending (the last tile colour's tiles do both sweeps), then the
// backward sweeps by tile colour descending.  A forward sweep depends on the adjacent tiles
// (an entry of A or A^T between them) of lower colours; a backward sweep on the adjacent tiles
// of higher colours and on its own forward sweep; every dependency is an earlier role.
// Returns false when a role has more than max_deps dependencies.
constexpr int TILE_ROLE_INTS = 16;
bool tile_roles_build(const IluOrder& o, int max_deps, std::vector<int32_t>& role, std::vector<int32_t>& dep);

// 63-bit Morton code of a point of the box [lo, hi] (21 bits per axis).
uint64_t morton3(const double* p, const double* lo, const double* hi);

}  // namespace vfvm

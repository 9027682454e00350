// C ABI: CSV and VTU output of fields (restates src/base/write_csv.rs:5-100 and
// src/base/write_vtu.rs; host only).
//
// The reference writes "{:.9}" (fixed point, 9 decimals, the exact binary value rounded
// half to even); "%.9f" is the same conversion.  Its non-finite spellings are "NaN", "inf"
// and "-inf".  digits = 0 selects "%.17g", which reads back to the same double, for parity
// dumps.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "common.hpp"
#include "mesh.hpp"

using namespace vfvm;

namespace {

// Appends v in the chosen format.  A fixed-point expansion of a large value can run to
// hundreds of characters (1e300 with 9 decimals: 311), so the text is formatted at the
// length snprintf reports, never cut to a fixed buffer.
void put_num(std::string& line, double v, int digits) {
    if (std::isnan(v)) { line += "NaN"; return; }
    if (std::isinf(v)) { line += v > 0 ? "inf" : "-inf"; return; }
    char buf[64];
    int len = digits > 0 ? std::snprintf(buf, sizeof buf, "%.*f", digits, v) : std::snprintf(buf, sizeof buf, "%.17g", v);
    if (len < 0) return;
    if ((size_t)len < sizeof buf) { line.append(buf, (size_t)len); return; }
    std::string big((size_t)len + 1, '\0');
    len = digits > 0 ? std::snprintf(&big[0], big.size(), "%.*f", digits, v) : std::snprintf(&big[0], big.size(), "%.17g", v);
    if (len > 0) line.append(big.data(), (size_t)len);
}

}  // namespace

namespace {

// vertices of a domain field (bnd < 0) or of boundary bnd's nodes, in output order
std::vector<int64_t> field_vertices(const Mesh& m, int32_t bnd) {
    std::vector<int64_t> vids;
    if (bnd < 0) {
        vids.resize(m.num_tot());
        for (int64_t v = 0; v < m.num_tot(); ++v) vids[v] = v;
    } else {
        for (int64_t k = m.bnd_ptr[bnd]; k < m.bnd_ptr[bnd + 1]; ++k) vids.push_back(m.num_pts + m.bnd_nid[k]);
    }
    return vids;
}

std::string output_path(const char* file_name, const char* file_type, int32_t ts) {
    char tsbuf[16];
    std::snprintf(tsbuf, sizeof tsbuf, "%04d", ts);
    return std::string(file_name) + "_" + tsbuf + "." + file_type;
}

}  // namespace

// VTU (write_vtu.rs): ASCII UnstructuredGrid; a domain field on the mesh's tetrahedra
// (type 10), a boundary field on one vertex cell per node (type 1); point data "value" with
// 1 or 3 components.
extern "C" int vfvm_write_vtu(const vfvm_mesh* mh, int32_t bnd, int32_t ncomp, const double* values, const char* file_name,
                              const char* file_type, int32_t ts, int32_t digits) {
    const Mesh* m = vfvm_mesh_ptr(mh);
    if (!m || !values || !file_name || !file_type || (ncomp != 1 && ncomp != 3) || ts < 0 || digits < 0 || digits > 17)
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_write_vtu: invalid argument");
    const int64_t nb = (int64_t)m->bnd_ptr.size() - 1;
    if (bnd >= nb) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_write_vtu: boundary id out of range");
    const std::vector<int64_t> vids = field_vertices(*m, bnd);
    const int64_t len = (int64_t)vids.size();
    const int64_t ncell = bnd < 0 ? (int64_t)m->tets.size() / 4 : len;
    const std::string path = output_path(file_name, file_type, ts);
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return set_error(VFVM_ERR_WRITE_VTU, "WriteVtu: Cannot write file '" + path + "'");
    std::string out;
    auto flush = [&]() -> bool {
        const bool ok = std::fputs(out.c_str(), f) >= 0;
        out.clear();
        return ok;
    };
    bool ok = true;
    out += "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n";
    out += "<Piece NumberOfPoints=\"" + std::to_string(len) + "\" NumberOfCells=\"" + std::to_string(ncell) + "\">\n";
    out += "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int64_t i = 0; i < len && ok; ++i) {
        for (int c = 0; c < 3; ++c) {
            put_num(out, m->pos[3 * vids[i] + c], digits);
            out += c < 2 ? ' ' : '\n';
        }
        if (out.size() > (1 << 16)) ok = flush();
    }
    out += "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    for (int64_t t = 0; t < ncell && ok; ++t) {
        if (bnd < 0)
            out += std::to_string(m->tets[4 * t]) + " " + std::to_string(m->tets[4 * t + 1]) + " " + std::to_string(m->tets[4 * t + 2]) +
                   " " + std::to_string(m->tets[4 * t + 3]) + "\n";
        else
            out += std::to_string(t) + "\n";
        if (out.size() > (1 << 16)) ok = flush();
    }
    out += "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (int64_t t = 0; t < ncell && ok; ++t) {
        out += std::to_string((bnd < 0 ? 4 : 1) * (t + 1)) + "\n";
        if (out.size() > (1 << 16)) ok = flush();
    }
    out += "</DataArray>\n<DataArray type=\"Int32\" Name=\"types\" format=\"ascii\">\n";
    for (int64_t t = 0; t < ncell && ok; ++t) {
        out += bnd < 0 ? "10\n" : "1\n";
        if (out.size() > (1 << 16)) ok = flush();
    }
    out += "</DataArray>\n</Cells>\n<PointData Scalars=\"value\">\n";
    out += ncomp == 1 ? "<DataArray type=\"Float64\" Name=\"value\" format=\"ascii\">\n"
                      : "<DataArray type=\"Float64\" Name=\"value\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int64_t i = 0; i < len && ok; ++i) {
        for (int c = 0; c < ncomp; ++c) {
            put_num(out, values[(int64_t)c * len + i], digits);
            out += c + 1 < ncomp ? ' ' : '\n';
        }
        if (out.size() > (1 << 16)) ok = flush();
    }
    out += "</DataArray>\n</PointData>\n<CellData>\n</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
    if (ok) ok = flush();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(VFVM_ERR_WRITE_VTU, "WriteVtu: Cannot write file '" + path + "'");
    return clear_error();
}

extern "C" int vfvm_write_csv(const vfvm_mesh* mh, int32_t bnd, int32_t ncomp, const double* values, const char* file_name,
                              const char* file_type, int32_t ts, int32_t digits) {
    const Mesh* m = vfvm_mesh_ptr(mh);
    if (!m || !values || !file_name || !file_type || (ncomp != 1 && ncomp != 3) || ts < 0 || digits < 0 || digits > 17)
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_write_csv: invalid argument");
    const int64_t nb = (int64_t)m->bnd_ptr.size() - 1;
    if (bnd >= nb) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_write_csv: boundary id out of range");
    const std::vector<int64_t> vids = field_vertices(*m, bnd);
    const int64_t len = (int64_t)vids.size();
    const std::string path = output_path(file_name, file_type, ts);
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return set_error(VFVM_ERR_WRITE_CSV, "WriteCsv: Cannot write file '" + path + "'");
    bool ok = std::fputs(ncomp == 1 ? "x,y,z,value\n" : "x,y,z,value_x,value_y,value_z\n", f) >= 0;
    std::string line;
    for (int64_t i = 0; i < len && ok; ++i) {
        const int64_t v = vids[i];
        line.clear();
        for (int c = 0; c < 3; ++c) {
            put_num(line, m->pos[3 * v + c], digits);
            line += ',';
        }
        for (int c = 0; c < ncomp; ++c) {
            put_num(line, values[(int64_t)c * len + i], digits);
            line += c + 1 < ncomp ? ',' : '\n';
        }
        ok = std::fputs(line.c_str(), f) >= 0;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(VFVM_ERR_WRITE_CSV, "WriteCsv: Cannot write file '" + path + "'");
    return clear_error();
}

// C ABI: CSV output of fields (restates src/base/write_csv.rs:5-100; host only).
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

void put_num(std::string& line, double v, int digits) {
    char buf[64];
    if (std::isnan(v)) std::snprintf(buf, sizeof buf, "NaN");
    else if (std::isinf(v)) std::snprintf(buf, sizeof buf, v > 0 ? "inf" : "-inf");
    else if (digits > 0) std::snprintf(buf, sizeof buf, "%.*f", digits, v);
    else std::snprintf(buf, sizeof buf, "%.17g", v);
    line += buf;
}

}  // namespace

extern "C" int vfvm_write_csv(const vfvm_mesh* mh, int32_t bnd, int32_t ncomp, const double* values, const char* file_name,
                              const char* file_type, int32_t ts, int32_t digits) {
    const Mesh* m = vfvm_mesh_ptr(mh);
    if (!m || !values || !file_name || !file_type || (ncomp != 1 && ncomp != 3) || ts < 0 || digits < 0 || digits > 17)
        return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_write_csv: invalid argument");
    const int64_t nb = (int64_t)m->bnd_ptr.size() - 1;
    if (bnd >= nb) return set_error(VFVM_ERR_INVALID_ARGUMENT, "vfvm_write_csv: boundary id out of range");
    // the vertices written, in order
    std::vector<int64_t> vids;
    if (bnd < 0) {
        vids.resize(m->num_tot());
        for (int64_t v = 0; v < m->num_tot(); ++v) vids[v] = v;
    } else {
        for (int64_t k = m->bnd_ptr[bnd]; k < m->bnd_ptr[bnd + 1]; ++k) vids.push_back(m->num_pts + m->bnd_nid[k]);
    }
    const int64_t len = (int64_t)vids.size();
    char tsbuf[16];
    std::snprintf(tsbuf, sizeof tsbuf, "%04d", ts);
    const std::string path = std::string(file_name) + "_" + tsbuf + "." + file_type;
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

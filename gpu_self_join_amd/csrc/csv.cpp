// csv.cpp -- text dataset import (gdsj_csv_load), host only.
//
// Same values as the reference's importNDDataset (import_dataset.cpp:49-95):
// the file is cut at every ',' (newlines stay inside the pieces); each piece
// is read with C++ formatted extraction into DTYPE, one value after another,
// until an extraction fails, which ends that piece.  Formatted extraction
// converts the decimal text straight to DTYPE (no detour through double for
// float data).  The caller reshapes the values to N x dims (N = count / dims,
// a trailing partial row dropped, import_dataset.cpp:72-84).
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <sstream>
#include <string>
#include <vector>

#include "gdsjoin.h"

namespace gdsj {
void set_error(const char *fmt, ...);
}

namespace {

template <typename T>
int load_values(const char *path, T *out, uint64_t cap, uint64_t *count) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        gdsj::set_error("cannot open %s", path);
        return GDSJ_EINVAL;
    }
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    fclose(f);
    uint64_t n = 0;
    std::istringstream ss;
    size_t pos = 0;
    while (pos <= text.size()) {
        size_t end = text.find(',', pos);
        const bool last = end == std::string::npos;
        if (last) end = text.size();
        if (end > pos) {
            ss.clear();
            ss.str(text.substr(pos, end - pos));
            T v;
            while (ss >> v) {
                if (out && n < cap) out[n] = v;
                n++;
            }
        }
        if (last) break;
        pos = end + 1;
    }
    *count = n;
    return GDSJ_OK;
}

}  // namespace

extern "C" int gdsj_csv_load(const char *path, int dtype, void *out, uint64_t cap, uint64_t *count) {
    if (!path || !count || (out == nullptr && cap)) return GDSJ_EINVAL;
    if (dtype == GDSJ_DTYPE_F32) return load_values<float>(path, (float *)out, cap, count);
    if (dtype == GDSJ_DTYPE_F64) return load_values<double>(path, (double *)out, cap, count);
    gdsj::set_error("dtype must be GDSJ_DTYPE_F32 or GDSJ_DTYPE_F64");
    return GDSJ_EINVAL;
}

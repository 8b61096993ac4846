// Walks HDF5 files through the C ABI of libusct_h5 (linked from source):
// every group listed, every dataset's shape and elements (as float64, and
// its first row through the row selection), every attribute as a number or a
// string.  Built with AddressSanitizer / UBSan by tests/test_h5.py and run on
// the golden files and on damaged copies of them: any file may fail with a
// status code, none may make the reader touch memory it does not own.
//
// usage: h5_walk file...   prints one line per file: "<file> ok <objects>"
// or "<file> err <code> <message>"; exits 0 unless the process faults.

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/usct_h5.h"

namespace {

constexpr int kMaxObjects = 4096;
constexpr int64_t kMaxElems = 1 << 24;

int walk(usct_h5_file *f, const std::string &path, int depth, int &objects) {
    if (depth > 16 || ++objects > kMaxObjects) return 0;
    int kind = 0;
    if (int rc = usct_h5_kind(f, path.c_str(), &kind)) return rc;
    int64_t need = 0;
    if (int rc = usct_h5_attr_list(f, path.c_str(), nullptr, 0, &need)) return rc;
    std::vector<char> names(size_t(need) + 1);
    if (int rc = usct_h5_attr_list(f, path.c_str(), names.data(), need, &need)) return rc;
    for (int64_t i = 0; i < need;) {
        std::string a(names.data() + i);
        i += int64_t(a.size()) + 1;
        int type = 0, rank = 0;
        int64_t n = 0;
        if (int rc = usct_h5_attr_info(f, path.c_str(), a.c_str(), &type, &n, &rank)) return rc;
        if (n < 0 || n > kMaxElems) continue;
        if (type == USCT_H5_STR) {
            std::vector<char> s(size_t(n) + 1);
            if (int rc = usct_h5_attr_read(f, path.c_str(), a.c_str(), USCT_H5_STR, s.data(), n + 1)) return rc;
        } else if (type != USCT_H5_OTHER) {
            std::vector<double> v(size_t(n) + 1);
            if (int rc = usct_h5_attr_read(f, path.c_str(), a.c_str(), USCT_H5_F64, v.data(), n)) return rc;
        }
    }
    if (kind == USCT_H5_DATASET) {
        int nd = 0, type = 0, chunked = 0;
        int64_t dims[USCT_H5_MAX_RANK];
        if (int rc = usct_h5_dataset_info(f, path.c_str(), &nd, dims, &type, &chunked)) return rc;
        if (type == USCT_H5_STR || type == USCT_H5_OTHER) return 0;
        int64_t n = 1;
        for (int i = 0; i < nd; ++i) {
            if (dims[i] < 0 || (dims[i] && n > kMaxElems / dims[i])) return 0;
            n *= dims[i];
        }
        if (n > kMaxElems) return 0;
        std::vector<double> v(size_t(n) + 1);
        if (int rc = usct_h5_read(f, path.c_str(), 0, -1, USCT_H5_F64, v.data(), n, 2)) return rc;
        if (nd > 0 && dims[0] > 0) {
            int64_t row = n / dims[0];
            std::vector<float> r(size_t(row) + 1);
            if (int rc = usct_h5_read(f, path.c_str(), dims[0] - 1, 1, USCT_H5_F32, r.data(), row, 1)) return rc;
        }
        return 0;
    }
    if (int rc = usct_h5_list(f, path.c_str(), nullptr, 0, &need)) return rc;
    std::vector<char> kids(size_t(need) + 1);
    if (int rc = usct_h5_list(f, path.c_str(), kids.data(), need, &need)) return rc;
    for (int64_t i = 0; i < need;) {
        std::string k(kids.data() + i);
        i += int64_t(k.size()) + 1;
        if (int rc = walk(f, (path == "/" ? "/" : path + "/") + k, depth + 1, objects)) return rc;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
        usct_h5_file *f = nullptr;
        int objects = 0;
        int rc = usct_h5_open(&f, argv[i]);
        if (rc == 0) rc = walk(f, "/", 0, objects);
        if (rc == 0)
            std::printf("%s ok %d\n", argv[i], objects);
        else
            std::printf("%s err %d %s\n", argv[i], rc, usct_h5_last_error());
        usct_h5_close(f);
    }
    return 0;
}

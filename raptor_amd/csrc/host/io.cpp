// Matrix files in PETSc's binary format, as the reference reads them
// (gallery/matrix_IO.cpp:24-110): int32 class id 1211216, rows, columns,
// nonzeros, then the row lengths, the column indices (int32) and the values
// (float64), all big-endian as PETSc writes them; a file whose header reads
// as little-endian is accepted as well (the reference's endianness test).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../status.hpp"
#include "setup.hpp"

using namespace ramd;

namespace {

constexpr int32_t kPetscMatCode = 1211216;

template <class T>
T swap_bytes(T v)
{
    unsigned char b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    for (size_t i = 0; i < sizeof(T) / 2; i++) std::swap(b[i], b[sizeof(T) - 1 - i]);
    std::memcpy(&v, b, sizeof(T));
    return v;
}

bool host_is_little()
{
    const uint16_t one = 1;
    unsigned char c;
    std::memcpy(&c, &one, 1);
    return c == 1;
}

struct File {
    FILE* f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

template <class T>
void read_array(FILE* f, T* out, size_t count, bool swap, const char* path)
{
    if (count && std::fread(out, sizeof(T), count, f) != count)
        fail(RAMD_E_ARG, "%s: file ends early", path);
    if (swap)
        for (size_t i = 0; i < count; i++) out[i] = swap_bytes(out[i]);
}

template <class T>
void write_array(FILE* f, const T* in, size_t count, bool swap, const char* path)
{
    std::vector<T> buf(in, in + count);
    if (swap)
        for (T& v : buf) v = swap_bytes(v);
    if (count && std::fwrite(buf.data(), sizeof(T), count, f) != count) fail(RAMD_E_ARG, "%s: write failed", path);
}

template <class T>
T* export_copy(const std::vector<T>& v)
{
    T* p = (T*)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

extern "C" {

int ramd_host_read_petsc(const char* path, int32_t* n_rows, int32_t* n_cols, int32_t** rowptr, int32_t** col,
                         double** val, int64_t* nnz)
{
    return guarded([&] {
        RAMD_CHECK(path && n_rows && n_cols && rowptr && col && val && nnz, RAMD_E_ARG,
                   "ramd_host_read_petsc: NULL argument");
        File file;
        file.f = std::fopen(path, "rb");
        RAMD_CHECK(file.f, RAMD_E_ARG, "%s: cannot open", path);
        int32_t head[4];
        read_array(file.f, head, 4, false, path);
        // PETSc writes big-endian; the class id tells which order the file uses
        bool swap = host_is_little();
        if (swap_bytes(head[0]) != kPetscMatCode && head[0] == kPetscMatCode) swap = false;
        else if (head[0] != kPetscMatCode && swap_bytes(head[0]) == kPetscMatCode) swap = true;
        else RAMD_CHECK(head[0] == kPetscMatCode, RAMD_E_ARG, "%s: not a PETSc matrix file", path);
        if (swap)
            for (int32_t& v : head) v = swap_bytes(v);
        const int32_t nr = head[1], nc = head[2], nz = head[3];
        RAMD_CHECK(nr >= 0 && nc >= 0 && nz >= 0, RAMD_E_ARG, "%s: bad header", path);
        std::vector<int32_t> lens((size_t)nr), c((size_t)nz);
        std::vector<double> v((size_t)nz);
        read_array(file.f, lens.data(), lens.size(), swap, path);
        read_array(file.f, c.data(), c.size(), swap, path);
        read_array(file.f, v.data(), v.size(), swap, path);
        std::vector<int32_t> rp((size_t)nr + 1, 0);
        for (int32_t i = 0; i < nr; i++) {
            RAMD_CHECK(lens[i] >= 0, RAMD_E_ARG, "%s: negative row length", path);
            rp[i + 1] = rp[i] + lens[i];
        }
        RAMD_CHECK(rp[nr] == nz, RAMD_E_ARG, "%s: row lengths sum to %d, header says %d", path, rp[nr], nz);
        for (int32_t j : c) RAMD_CHECK(j >= 0 && j < nc, RAMD_E_ARG, "%s: column %d out of range", path, j);
        *n_rows = nr;
        *n_cols = nc;
        *nnz = nz;
        *rowptr = export_copy(rp);
        *col = export_copy(c);
        *val = export_copy(v);
    });
}

int ramd_host_write_petsc(const char* path, int32_t n_rows, int32_t n_cols, const int32_t* rowptr,
                          const int32_t* col, const double* val)
{
    return guarded([&] {
        RAMD_CHECK(path && rowptr && n_rows >= 0 && n_cols >= 0, RAMD_E_ARG, "ramd_host_write_petsc: bad argument");
        const int64_t nnz = rowptr[n_rows];
        RAMD_CHECK(rowptr[0] == 0 && nnz <= INT32_MAX && (nnz == 0 || (col && val)), RAMD_E_ARG,
                   "ramd_host_write_petsc: bad CSR");
        File file;
        file.f = std::fopen(path, "wb");
        RAMD_CHECK(file.f, RAMD_E_ARG, "%s: cannot create", path);
        const bool swap = host_is_little();  // written big-endian
        const int32_t head[4] = {kPetscMatCode, n_rows, n_cols, (int32_t)nnz};
        write_array(file.f, head, 4, swap, path);
        std::vector<int32_t> lens((size_t)n_rows);
        for (int32_t i = 0; i < n_rows; i++) {
            RAMD_CHECK(rowptr[i + 1] >= rowptr[i], RAMD_E_ARG, "ramd_host_write_petsc: rowptr decreases at %d", i);
            lens[i] = rowptr[i + 1] - rowptr[i];
        }
        write_array(file.f, lens.data(), lens.size(), swap, path);
        write_array(file.f, col, (size_t)nnz, swap, path);
        write_array(file.f, val, (size_t)nnz, swap, path);
        RAMD_CHECK(std::fflush(file.f) == 0, RAMD_E_ARG, "%s: write failed", path);
    });
}

}  // extern "C"

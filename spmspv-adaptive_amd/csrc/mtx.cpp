// Matrix and frontier file I/O on the host (the reference's
// la/include/spmspv/mtx.hpp and readsvmdata.hpp):
//   - MatrixMarket coordinate files -> CSR with the reference's semantics
//     (ReadMTX, mtx.hpp:218-378; SortMTX :380-402; the harness's coo2csr,
//     la/test/new_test_app.cu:547-559): 1-based indices, pattern and complex
//     entries read as 1.0, integer and real entries as their value, a
//     "symmetric" file mirrored (off-diagonal entries twice, same value),
//     every other symmetry tag read as general, entries ordered by
//     (row, column), duplicates kept;
//   - the binary CSR cache ("Hola\1CSR\1" header, then values, column
//     indices, row pointers; mtx.hpp:709-819);
//   - BFS / PageRank frontier files (readNNZXFromFile,
//     readSparseXInxFromFile, extractSparseXfromFile; readsvmdata.hpp:223-348):
//     one x size per whitespace-separated integer, and one line of keys per
//     iteration with every value 1.0.
// Where the reference exits or reads out of range (missing file, index out
// of bounds, truncated data), these return SPMSPV_* codes with a message.
#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spmspv_amd.h"

namespace spmspv {
int set_error(spmspv_handle_t h, int code, const std::string& msg);
}

namespace {

using spmspv::set_error;

bool read_file(const char* path, std::string* out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out->resize(len > 0 ? (size_t)len : 0);
  const size_t got = len > 0 ? std::fread(&(*out)[0], 1, (size_t)len, f) : 0;
  std::fclose(f);
  return got == out->size();
}

// Whitespace-separated tokens over a NUL-terminated buffer.
struct Scanner {
  const char* p;
  const char* end;
  void skip_ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' || *p == '\v')) ++p;
  }
  bool token(std::string* t) {
    skip_ws();
    if (p >= end) return false;
    const char* b = p;
    while (p < end && !(*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' || *p == '\v')) ++p;
    t->assign(b, p);
    return true;
  }
  void rest_of_line() {
    while (p < end && *p != '\n') ++p;
  }
  bool integer(long long* v) {
    skip_ws();
    if (p >= end) return false;
    char* e = nullptr;
    errno = 0;
    *v = std::strtoll(p, &e, 10);
    if (e == p || errno) return false;
    p = e;
    return true;
  }
  bool real(double* v) {
    skip_ws();
    if (p >= end) return false;
    char* e = nullptr;
    *v = std::strtod(p, &e);
    if (e == p) return false;
    p = e;
    return true;
  }
};

struct Coo {
  int r, c;
  float v;
};

// The file's matrix as (row, column)-ordered COO, mirrored when symmetric.
int parse_mtx(const char* path, int* m, int* n, std::vector<Coo>* out) {
  std::string buf;
  if (!path || !read_file(path, &buf))
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, std::string("mtx: cannot read ") + (path ? path : "(null)"));
  Scanner s{buf.data(), buf.data() + buf.size()};
  std::string banner, object, format, field, symmetry;
  if (!s.token(&banner) || !s.token(&object) || !s.token(&format) || !s.token(&field) || !s.token(&symmetry))
    return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: incomplete header line");
  if (format == "array")
    return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: dense (array) files are not coordinate matrices");
  const bool real = field == "real" || field == "double";
  const bool integer = field == "integer";
  const bool pattern = field == "pattern";
  const bool complex = !real && !integer && !pattern;
  const bool sym = symmetry == "symmetric";
  s.rest_of_line();
  // comment lines, then "rows cols entries"
  std::string tok;
  long long rows = 0, cols = 0, entries = 0;
  for (;;) {
    if (!s.token(&tok)) return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: no size line");
    if (tok[0] != '%') break;
    s.rest_of_line();
  }
  rows = std::strtoll(tok.c_str(), nullptr, 10);
  if (!s.integer(&cols) || !s.integer(&entries) || rows <= 0 || cols <= 0 || entries < 0 ||
      rows > INT32_MAX || cols > INT32_MAX)
    return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: bad size line");
  out->clear();
  out->reserve((size_t)(sym ? 2 * entries : entries));
  for (long long i = 0; i < entries; ++i) {
    long long r = 0, c = 0;
    if (!s.integer(&r) || !s.integer(&c))
      return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: entry " + std::to_string(i) + " is truncated");
    if (r < 1 || r > rows || c < 1 || c > cols)
      return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: entry " + std::to_string(i) + " is out of range");
    float v = 1.0f;
    if (real) {
      double d;
      if (!s.real(&d)) return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: entry value missing");
      v = (float)d;
    } else if (integer) {
      long long d;
      if (!s.integer(&d)) return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: entry value missing");
      v = (float)(int)d;
    } else if (complex) {
      double re, im;
      if (!s.real(&re) || !s.real(&im)) return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: entry value missing");
      v = 1.0f;  // the reference keeps the pattern of complex matrices
    }
    out->push_back(Coo{(int)(r - 1), (int)(c - 1), v});
    if (sym && r != c) out->push_back(Coo{(int)(c - 1), (int)(r - 1), v});
  }
  if ((long long)out->size() > INT32_MAX)
    return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, "mtx: more than 2^31-1 stored entries");
  // (row, column) order; stable, so duplicates keep their file order
  std::stable_sort(out->begin(), out->end(),
                   [](const Coo& a, const Coo& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
  *m = (int)rows;
  *n = (int)cols;
  return SPMSPV_SUCCESS;
}

// Header of the binary CSR cache: the 9-byte magic, 3 padding bytes, then
// rows, columns and stored entries as int32 (24 bytes, as the reference's
// struct lays it out).
constexpr char kMagic[9] = {'H', 'o', 'l', 'a', 1, 'C', 'S', 'R', 1};
struct CsrHeader {
  char magic[9];
  char pad[3];
  int32_t rows, cols, nnz;
};
static_assert(sizeof(CsrHeader) == 24, "binary CSR header is 24 bytes");

int read_header(FILE* f, const char* path, CsrHeader* h) {
  if (std::fread(h, sizeof(*h), 1, f) != 1) {
    std::fclose(f);
    return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, std::string("csr: cannot read the header of ") + path);
  }
  if (std::memcmp(h->magic, kMagic, sizeof(kMagic)) != 0) {
    std::fclose(f);
    return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, std::string(path) + " is not a binary CSR file");
  }
  return SPMSPV_SUCCESS;
}

}  // namespace

extern "C" {

int spmspv_mtx_read_csr(const char* path, int* m, int* n, long long* nnz, int* row_ptr, int* col,
                        float* val) {
  if (!m || !n || !nnz) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "mtx: null size pointer");
  std::vector<Coo> coo;
  int rows = 0, cols = 0;
  int rc = parse_mtx(path, &rows, &cols, &coo);
  if (rc) return rc;
  *m = rows;
  *n = cols;
  *nnz = (long long)coo.size();
  if (!row_ptr) return SPMSPV_SUCCESS;
  if (!coo.empty() && (!col || !val)) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "mtx: null output array");
  std::fill(row_ptr, row_ptr + rows + 1, 0);
  for (const Coo& e : coo) row_ptr[e.r + 1]++;
  for (int r = 0; r < rows; ++r) row_ptr[r + 1] += row_ptr[r];
  for (size_t i = 0; i < coo.size(); ++i) {
    col[i] = coo[i].c;
    val[i] = coo[i].v;
  }
  return SPMSPV_SUCCESS;
}

int spmspv_csr_load_header(const char* path, int* m, int* n, int* nnz) {
  if (!path || !m || !n || !nnz) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "csr: null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, std::string("csr: cannot open ") + path);
  CsrHeader h;
  int rc = read_header(f, path, &h);
  if (rc) return rc;
  std::fclose(f);
  *m = h.rows;
  *n = h.cols;
  *nnz = h.nnz;
  return SPMSPV_SUCCESS;
}

int spmspv_csr_load(const char* path, int m, int n, int nnz, int* row_ptr, int* col, float* val) {
  if (!path || m < 0 || n < 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!col || !val)))
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "csr: bad arguments");
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, std::string("csr: cannot open ") + path);
  CsrHeader h;
  int rc = read_header(f, path, &h);
  if (rc) return rc;
  if (h.rows != m || h.cols != n || h.nnz != nnz) {
    std::fclose(f);
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "csr: sizes differ from the file's header");
  }
  const bool ok = std::fread(val, sizeof(float), (size_t)nnz, f) == (size_t)nnz &&
                  std::fread(col, sizeof(int), (size_t)nnz, f) == (size_t)nnz &&
                  std::fread(row_ptr, sizeof(int), (size_t)m + 1, f) == (size_t)m + 1;
  std::fclose(f);
  if (!ok) return set_error(nullptr, SPMSPV_MAT_FORMAT_ERR, std::string("csr: ") + path + " is truncated");
  return SPMSPV_SUCCESS;
}

int spmspv_csr_store(const char* path, int m, int n, int nnz, const int* row_ptr, const int* col,
                     const float* val) {
  if (!path || m < 0 || n < 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!col || !val)))
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "csr: bad arguments");
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, std::string("csr: cannot create ") + path);
  CsrHeader h{};
  std::memcpy(h.magic, kMagic, sizeof(kMagic));
  h.rows = m;
  h.cols = n;
  h.nnz = nnz;
  const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
                  std::fwrite(val, sizeof(float), (size_t)nnz, f) == (size_t)nnz &&
                  std::fwrite(col, sizeof(int), (size_t)nnz, f) == (size_t)nnz &&
                  std::fwrite(row_ptr, sizeof(int), (size_t)m + 1, f) == (size_t)m + 1;
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) return set_error(nullptr, SPMSPV_DEVICE_ERROR, std::string("csr: cannot write ") + path);
  return SPMSPV_SUCCESS;
}

int spmspv_frontier_sizes(const char* path, int* count, int* sizes) {
  if (!path || !count) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "frontier: null argument");
  std::string buf;
  if (!read_file(path, &buf))
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, std::string("frontier: cannot read ") + path);
  Scanner s{buf.data(), buf.data() + buf.size()};
  long long v;
  int c = 0;
  // integers until the first one that does not parse (the reference's
  // `while (ifs >> index)`)
  while (s.integer(&v)) {
    if (sizes) sizes[c] = (int)v;
    ++c;
  }
  *count = c;
  return SPMSPV_SUCCESS;
}

int spmspv_frontier_read(const char* path, int line, int cap, int* xnnz, int* keys, float* vals) {
  if (!path || line < 0 || cap < 0 || !xnnz || (cap > 0 && (!keys || !vals)))
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "frontier: bad arguments");
  std::string buf;
  if (!read_file(path, &buf))
    return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, std::string("frontier: cannot read ") + path);
  const char* p = buf.data();
  const char* end = p + buf.size();
  for (int l = 0; l < line && p < end; ++l) {
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
    p = nl ? nl + 1 : end;
  }
  const char* nl = p < end ? static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p))) : nullptr;
  const char* le = nl ? nl : end;
  std::string text(p, le);
  Scanner s{text.data(), text.data() + text.size()};
  std::string tok;
  int k = 0;
  while (s.token(&tok)) {
    char* e = nullptr;
    const long long key = std::strtoll(tok.c_str(), &e, 10);
    if (e == tok.c_str()) return set_error(nullptr, SPMSPV_VEC_FORMAT_ERR, "frontier: key is not an integer");
    if (k >= cap) return set_error(nullptr, SPMSPV_INVALID_ARGUMENT, "frontier: more keys than the capacity");
    keys[k] = (int)key;
    vals[k] = 1.0f;
    ++k;
  }
  *xnnz = k;
  return SPMSPV_SUCCESS;
}

}  // extern "C"

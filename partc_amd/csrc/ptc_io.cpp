// Text readers of the reference's input formats (host side).
//
//   .tns (FROSTT): one nonzero per line, d 1-based coordinates then the
//        value; the order d is fixed by the first data line.  dims[j] is the
//        largest label of mode j (what readTensor stores from pigo's
//        max_labels, mtxTnsIO.cpp:31-63, pigo.hpp:4363-4456).
//   .mtx (Matrix Market coordinate): "%%MatrixMarket matrix coordinate"
//        header, "%" comment lines, "rows cols nnz", then "i j v" 1-based
//        (readMatrix, mtxTnsIO.cpp:5-29; pigo.hpp:3149-3203): d = 2, dims =
//        the header's row and column counts, and the number of entries must
//        equal nnz.  A "pattern" matrix gets value 1.0.
//
// Coordinates become 0-based.  Values are parsed with strtod (correctly
// rounded); pigo's own float parser is not, so a decimal value can differ
// from the reference's by an ulp.  Lines starting with '#' or '%' and blank
// lines are skipped.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "partc.h"

namespace {

bool read_file(const char* path, std::string& buf) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return false;
  }
  long sz = ftell(f);
  if (sz < 0) {
    fclose(f);
    return false;
  }
  rewind(f);
  buf.resize((size_t)sz);
  size_t got = sz ? fread(&buf[0], 1, (size_t)sz, f) : 0;
  fclose(f);
  return got == (size_t)sz;
}

// Next line [p, e) of [cur, end); returns false at the end.
bool next_line(const char*& cur, const char* end, const char*& p, const char*& e) {
  while (cur < end) {
    p = cur;
    const char* nl = (const char*)memchr(cur, '\n', (size_t)(end - cur));
    e = nl ? nl : end;
    cur = nl ? nl + 1 : end;
    const char* q = p;
    while (q < e && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
    if (q == e || *q == '#' || *q == '%') continue;  // blank or comment
    p = q;
    return true;
  }
  return false;
}

// Tokens of one line: up to `max` numbers as text spans.
int split(const char* p, const char* e, const char** tok, int max) {
  int n = 0;
  while (p < e && n < max) {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    if (p >= e) break;
    tok[n++] = p;
    while (p < e && *p != ' ' && *p != '\t' && *p != '\r') ++p;
  }
  return n;
}

bool parse_label(const char* s, uint32_t* out) {
  errno = 0;
  char* end = nullptr;
  unsigned long long v = strtoull(s, &end, 10);
  if (end == s || errno || v == 0 || v > 0xFFFFFFFFull) return false;  // labels are 1-based
  *out = (uint32_t)(v - 1);
  return true;
}

bool parse_value(const char* s, double* out) {
  char* end = nullptr;
  *out = strtod(s, &end);
  return end != s;
}

int fill(ptc_host_tensor* t, uint32_t d, const std::vector<uint32_t>& h, const std::vector<double>& v,
         const uint32_t* dims) {
  t->d = d;
  t->N = v.size();
  t->hedges = (uint32_t*)malloc(sizeof(uint32_t) * (h.size() ? h.size() : 1));
  t->vals = (double*)malloc(sizeof(double) * (v.size() ? v.size() : 1));
  if (!t->hedges || !t->vals) {
    free(t->hedges);
    free(t->vals);
    t->hedges = nullptr;
    t->vals = nullptr;
    return PTC_ERR_NOMEM;
  }
  if (!h.empty()) memcpy(t->hedges, h.data(), sizeof(uint32_t) * h.size());
  if (!v.empty()) memcpy(t->vals, v.data(), sizeof(double) * v.size());
  for (uint32_t j = 0; j < PTC_MAX_MODES; ++j) t->dims[j] = j < d ? dims[j] : 0;
  return PTC_OK;
}

}  // namespace

extern "C" int ptc_read_tns(const char* path, ptc_host_tensor* out) {
  if (!path || !out) return PTC_ERR_ARG;
  memset(out, 0, sizeof(*out));
  std::string buf;
  if (!read_file(path, buf)) return PTC_ERR_IO;
  const char *cur = buf.data(), *end = buf.data() + buf.size(), *p, *e;
  const char* tok[PTC_MAX_MODES + 2];
  uint32_t d = 0, dims[PTC_MAX_MODES] = {};
  std::vector<uint32_t> h;
  std::vector<double> v;
  while (next_line(cur, end, p, e)) {
    const int nt = split(p, e, tok, PTC_MAX_MODES + 2);
    if (d == 0) {
      if (nt < 2 || nt - 1 > PTC_MAX_MODES) return PTC_ERR_IO;  // order from the first data line
      d = (uint32_t)(nt - 1);
    }
    if ((uint32_t)nt != d + 1) return PTC_ERR_IO;
    for (uint32_t j = 0; j < d; ++j) {
      uint32_t x;
      if (!parse_label(tok[j], &x)) return PTC_ERR_IO;
      h.push_back(x);
      if (x + 1 > dims[j]) dims[j] = x + 1;
    }
    double val;
    if (!parse_value(tok[d], &val)) return PTC_ERR_IO;
    v.push_back(val);
  }
  if (d == 0) return PTC_ERR_IO;
  return fill(out, d, h, v, dims);
}

extern "C" int ptc_read_mtx(const char* path, ptc_host_tensor* out) {
  if (!path || !out) return PTC_ERR_ARG;
  memset(out, 0, sizeof(*out));
  std::string buf;
  if (!read_file(path, buf)) return PTC_ERR_IO;
  static const char kHeader[] = "%%MatrixMarket matrix coordinate";
  if (buf.compare(0, sizeof(kHeader) - 1, kHeader) != 0) return PTC_ERR_UNSUPPORTED;  // pigo.hpp:3154-3155
  const char* nl = (const char*)memchr(buf.data(), '\n', buf.size());
  const std::string header(buf.data(), nl ? (size_t)(nl - buf.data()) : buf.size());
  const bool pattern = header.find("pattern") != std::string::npos;
  const char *cur = buf.data(), *end = buf.data() + buf.size(), *p, *e;
  const char* tok[4];
  if (!next_line(cur, end, p, e) || split(p, e, tok, 4) != 3) return PTC_ERR_IO;
  char* q = nullptr;
  const unsigned long long nrows = strtoull(tok[0], &q, 10), ncols = strtoull(tok[1], &q, 10),
                           nnz = strtoull(tok[2], &q, 10);
  if (nrows > 0xFFFFFFFFull || ncols > 0xFFFFFFFFull) return PTC_ERR_RANGE;
  uint32_t dims[PTC_MAX_MODES] = {(uint32_t)nrows, (uint32_t)ncols};
  std::vector<uint32_t> h;
  std::vector<double> v;
  h.reserve(2 * nnz);
  v.reserve(nnz);
  while (next_line(cur, end, p, e)) {
    const int nt = split(p, e, tok, 4);
    if (nt != (pattern ? 2 : 3)) return PTC_ERR_IO;
    uint32_t i, j;
    if (!parse_label(tok[0], &i) || !parse_label(tok[1], &j)) return PTC_ERR_IO;
    if (i >= nrows || j >= ncols) return PTC_ERR_RANGE;  // more labels than the header (pigo.hpp:3171-3182)
    double val = 1.0;
    if (!pattern && !parse_value(tok[2], &val)) return PTC_ERR_IO;
    h.push_back(i);
    h.push_back(j);
    v.push_back(val);
  }
  if (v.size() != nnz) return PTC_ERR_IO;  // header contradicts the entries (pigo.hpp:3194-3198)
  return fill(out, 2, h, v, dims);
}

extern "C" int ptc_write_tns(const char* path, uint32_t d, uint64_t N, const uint32_t* hedges, const double* vals) {
  if (!path || d == 0 || d > PTC_MAX_MODES || (N && (!hedges || !vals))) return PTC_ERR_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return PTC_ERR_IO;
  for (uint64_t i = 0; i < N; ++i) {
    for (uint32_t j = 0; j < d; ++j) fprintf(f, "%u ", hedges[i * d + j] + 1);
    fprintf(f, "%.17g\n", vals[i]);
  }
  return fclose(f) == 0 ? PTC_OK : PTC_ERR_IO;
}

extern "C" void ptc_host_tensor_free(ptc_host_tensor* t) {
  if (!t) return;
  free(t->hedges);
  free(t->vals);
  t->hedges = nullptr;
  t->vals = nullptr;
}

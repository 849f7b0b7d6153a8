// io.cpp -- file formats and start data of the reference's command-line
// programs, host only (no device code):
//
//   sgtsne_read_X_text          readXfromMTX          CU/utils.cu:216-256
//   sgtsne_read_P_mtx           buildPFromMTX         CU/utils.cu:335-409 (= HY/utils.cpp:104-170)
//   sgtsne_write_embedding_text extractEmbeddingTextT CU/utils.cu:121-138 (SoA)
//                               extractEmbeddingText  HY/utils.cpp:180-196 (AoS)
//   sgtsne_write_embedding_bin  extractEmbedding      HY/utils.cpp:199-206
//   sgtsne_write_time_info      ExtractTimeInfo       CU/utils.cu:72-102
//   sgtsne_simulation_data      loadSimulationData    CU/utils.cu:310-332
//   sgtsne_random_start         generateRandomGaussianCoord CU/utils.cu:185-192
//
// Numbers are written the way the reference's std::ofstream writes them: the
// stream's default float format, which is printf's "%.6g" of the value as a
// double, each followed by one space, one line per point.
#include "../../include/sgtsne.h"
#include "common.hpp"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

// A whitespace-separated token reader over a whole file (the reference uses
// operator>>, which skips any whitespace between numbers).
struct Tokens {
  std::string buf;
  size_t pos = 0;

  bool load(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char tmp[1 << 16];
    size_t got;
    while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.append(tmp, got);
    std::fclose(f);
    return true;
  }
  static bool space(char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f'; }
  void skip_space() {
    while (pos < buf.size() && space(buf[pos])) ++pos;
  }
  // lines that start with '%' before the header (Matrix Market comments)
  void skip_comment_lines() {
    while (pos < buf.size() && buf[pos] == '%') {
      while (pos < buf.size() && buf[pos] != '\n') ++pos;
      if (pos < buf.size()) ++pos;
    }
  }
  const char *next(size_t &len) {
    skip_space();
    const size_t s = pos;
    while (pos < buf.size() && !space(buf[pos])) ++pos;
    len = pos - s;
    return len ? buf.data() + s : nullptr;
  }
  bool read_long(long long &v) {
    size_t len;
    const char *t = next(len);
    if (!t) return false;
    std::string s(t, len);
    char *end = nullptr;
    errno = 0;
    v = std::strtoll(s.c_str(), &end, 10);
    return errno == 0 && end == s.c_str() + s.size();
  }
  bool read_float(float &v) {
    size_t len;
    const char *t = next(len);
    if (!t) return false;
    std::string s(t, len);
    char *end = nullptr;
    v = std::strtof(s.c_str(), &end);  // correctly rounded, as operator>>(float&)
    return end == s.c_str() + s.size();
  }
  bool read_double(double &v) {
    size_t len;
    const char *t = next(len);
    if (!t) return false;
    std::string s(t, len);
    char *end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end == s.c_str() + s.size();
  }
};

// "%.6g " of each value: std::ostream's default float output followed by the
// reference's separator.
inline void put_value(std::string &out, double v) {
  char tmp[64];
  const int k = std::snprintf(tmp, sizeof(tmp), "%.6g ", v);
  out.append(tmp, (size_t)k);
}

int write_all(const char *path, const std::string &s, const char *mode) {
  FILE *f = std::fopen(path, mode);
  SGT_CHECK(f != nullptr, SGTSNE_ERR_ARG, "cannot open %s for writing", path);
  const size_t put = std::fwrite(s.data(), 1, s.size(), f);
  const int closed = std::fclose(f);
  SGT_CHECK(put == s.size() && closed == 0, SGTSNE_ERR_ARG, "short write to %s", path);
  return SGTSNE_OK;
}

}  // namespace

extern "C" {

void sgtsne_free(void *p) { std::free(p); }

int sgtsne_read_X_text(const char *path, int *n, int *dim, float **X) {
  SGT_CHECK(path && n && dim && X, SGTSNE_ERR_ARG, "NULL argument");
  *X = nullptr;
  Tokens tk;
  SGT_CHECK(tk.load(path), SGTSNE_ERR_ARG, "cannot open %s", path);
  long long fn = 0, fd = 0;
  SGT_CHECK(tk.read_long(fn) && tk.read_long(fd), SGTSNE_ERR_ARG, "%s: no 'n d' header", path);
  SGT_CHECK(fn > 0 && fd > 0 && fn <= INT32_MAX && fd <= INT32_MAX, SGTSNE_ERR_ARG, "%s: bad header %lld %lld",
            path, fn, fd);
  // a positive *n on entry reads only the first *n points (the reference's
  // params.n, CU/utils.cu:233-235)
  const long long want = *n > 0 ? *n : fn;
  SGT_CHECK(want <= fn, SGTSNE_ERR_ARG, "%s holds %lld points, %lld requested", path, fn, want);
  const size_t count = (size_t)want * (size_t)fd;
  float *out = static_cast<float *>(std::malloc(sizeof(float) * (count ? count : 1)));
  SGT_CHECK(out != nullptr, SGTSNE_ERR_NOMEM, "out of host memory");
  for (size_t t = 0; t < count; ++t) {
    if (!tk.read_float(out[t])) {
      std::free(out);
      sgt::set_error("%s: value %zu of %zu is missing or malformed", path, t, count);
      return SGTSNE_ERR_ARG;
    }
  }
  *n = (int)want;
  *dim = (int)fd;
  *X = out;
  return SGTSNE_OK;
}

int sgtsne_read_P_mtx(const char *path, int *m, int *n, int64_t *nnz, int32_t **colptr, int32_t **row,
                      double **val) {
  SGT_CHECK(path && m && n && nnz && colptr && row && val, SGTSNE_ERR_ARG, "NULL argument");
  *colptr = nullptr;
  *row = nullptr;
  *val = nullptr;
  Tokens tk;
  SGT_CHECK(tk.load(path), SGTSNE_ERR_ARG, "cannot open %s", path);
  tk.skip_comment_lines();
  long long fm = 0, fn = 0, fz = 0;
  SGT_CHECK(tk.read_long(fm) && tk.read_long(fn) && tk.read_long(fz), SGTSNE_ERR_ARG, "%s: no 'm n nnz' line",
            path);
  SGT_CHECK(fm >= 0 && fn >= 0 && fz >= 0 && fm <= INT32_MAX && fn <= INT32_MAX && fz <= INT32_MAX,
            SGTSNE_ERR_ARG, "%s: bad size line", path);
  std::vector<int32_t> ri((size_t)fz), ci((size_t)fz);
  std::vector<double> vv((size_t)fz);
  for (long long l = 0; l < fz; ++l) {
    long long r = 0, c = 0;
    double v = 0.0;
    SGT_CHECK(tk.read_long(r) && tk.read_long(c) && tk.read_double(v), SGTSNE_ERR_ARG,
              "%s: entry %lld of %lld is missing or malformed", path, l, fz);
    SGT_CHECK(r >= 1 && r <= fm && c >= 1 && c <= fn, SGTSNE_ERR_ARG, "%s: entry %lld (%lld, %lld) out of range",
              path, l, r, c);
    ri[(size_t)l] = (int32_t)(r - 1);  // 1-based in the file
    ci[(size_t)l] = (int32_t)(c - 1);
    vv[(size_t)l] = v;
  }
  // COO -> CSC by a counting sort on the column, stable within a column
  int32_t *cp = static_cast<int32_t *>(std::calloc((size_t)fn + 1, sizeof(int32_t)));
  int32_t *ro = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * (size_t)(fz ? fz : 1)));
  double *va = static_cast<double *>(std::malloc(sizeof(double) * (size_t)(fz ? fz : 1)));
  if (!cp || !ro || !va) {
    std::free(cp);
    std::free(ro);
    std::free(va);
    sgt::set_error("out of host memory");
    return SGTSNE_ERR_NOMEM;
  }
  for (long long l = 0; l < fz; ++l) cp[ci[(size_t)l] + 1]++;
  for (long long j = 0; j < fn; ++j) cp[j + 1] += cp[j];
  std::vector<int32_t> fill(cp, cp + fn);
  for (long long l = 0; l < fz; ++l) {
    const int32_t dst = fill[ci[(size_t)l]]++;
    ro[dst] = ri[(size_t)l];
    va[dst] = vv[(size_t)l];
  }
  *m = (int)fm;
  *n = (int)fn;
  *nnz = fz;
  *colptr = cp;
  *row = ro;
  *val = va;
  return SGTSNE_OK;
}

int sgtsne_write_embedding_text(const char *path, const float *y, int n, int d, int layout) {
  SGT_CHECK(path && (y || n == 0) && n >= 0 && d >= 1, SGTSNE_ERR_ARG, "bad arguments");
  SGT_CHECK(layout == SGTSNE_LAYOUT_SOA || layout == SGTSNE_LAYOUT_AOS, SGTSNE_ERR_ARG, "bad layout");
  std::string out;
  out.reserve((size_t)n * d * 12);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < d; ++j)
      put_value(out, (double)(layout == SGTSNE_LAYOUT_SOA ? y[i + (size_t)j * n] : y[(size_t)i * d + j]));
    out.push_back('\n');
  }
  return write_all(path, out, "wb");
}

int sgtsne_write_embedding_text_f64(const char *path, const double *y, int n, int d, int layout) {
  SGT_CHECK(path && (y || n == 0) && n >= 0 && d >= 1, SGTSNE_ERR_ARG, "bad arguments");
  SGT_CHECK(layout == SGTSNE_LAYOUT_SOA || layout == SGTSNE_LAYOUT_AOS, SGTSNE_ERR_ARG, "bad layout");
  std::string out;
  out.reserve((size_t)n * d * 12);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < d; ++j)
      put_value(out, layout == SGTSNE_LAYOUT_SOA ? y[i + (size_t)j * n] : y[(size_t)i * d + j]);
    out.push_back('\n');
  }
  return write_all(path, out, "wb");
}

int sgtsne_write_embedding_bin(const char *path, const double *y, int n, int d) {
  SGT_CHECK(path && (y || n == 0) && n >= 0 && d >= 1, SGTSNE_ERR_ARG, "bad arguments");
  std::string out(reinterpret_cast<const char *>(y), sizeof(double) * (size_t)n * d);
  return write_all(path, out, "wb");
}

int sgtsne_write_time_info(const char *path, const double *time_info, int n, int d, int iters) {
  SGT_CHECK(path && time_info && iters >= 0, SGTSNE_ERR_ARG, "bad arguments");
  // The reference appends only to a file that already exists (it probes
  // with an input stream first) and otherwise leaves an empty file behind.
  FILE *probe = std::fopen(path, "rb");
  if (!probe) return write_all(path, std::string(), "ab");
  std::fclose(probe);
  std::string out = "N d gds matrix total: \n";
  out += std::to_string(n) + " " + std::to_string(d) + " ";
  char tmp[64];
  for (int k = 0; k < 3; ++k) {
    std::snprintf(tmp, sizeof(tmp), k < 2 ? "%.6g " : "%.6g\n", time_info[(size_t)7 * iters + k]);
    out += tmp;
  }
  out += "Fattr S2G G2G G2S Phi2F Nuconv Pre";
  for (int i = 0; i < iters; ++i) {
    for (int j = 0; j < 7; ++j) put_value(out, time_info[(size_t)i * 7 + j]);
    out.push_back('\n');
  }
  return write_all(path, out, "ab");
}

int sgtsne_simulation_data(int n, int dim, int n_clusters, float *X) {
  SGT_CHECK(n >= 0 && dim >= 1 && n_clusters >= 1 && (X || n == 0), SGTSNE_ERR_ARG, "bad arguments");
  // cluster j holds points [j n / c, (j+1) n / c); coordinate a of its points
  // is drawn from N(20, 2) when a >= j, else from N(-20, 2); one default-
  // seeded std::default_random_engine feeds both distributions in order
  std::default_random_engine gen;
  std::normal_distribution<float> lo(-20.0f, 2.0f), hi(20.0f, 2.0f);
  for (int j = 0; j < n_clusters; ++j)
    for (int i = (int)((int64_t)j * n / n_clusters); i < (int)((int64_t)(j + 1) * n / n_clusters); ++i)
      for (int a = 0; a < dim; ++a) X[(size_t)i * dim + a] = a >= j ? hi(gen) : lo(gen);
  return SGTSNE_OK;
}

int sgtsne_random_start(int n, int d, int layout, float *y) {
  SGT_CHECK(n >= 0 && d >= 1 && (y || n == 0), SGTSNE_ERR_ARG, "bad arguments");
  SGT_CHECK(layout == SGTSNE_LAYOUT_SOA || layout == SGTSNE_LAYOUT_AOS, SGTSNE_ERR_ARG, "bad layout");
  // the sequence fills [d][N] in order (the reference's layout); an AoS
  // request receives the same points transposed
  std::default_random_engine gen;
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (size_t t = 0; t < (size_t)n * d; ++t) {
    const float v = (float)(0.0001 * (double)nd(gen));
    const size_t i = t % (size_t)n, j = t / (size_t)n;
    if (layout == SGTSNE_LAYOUT_SOA)
      y[t] = v;
    else
      y[i * d + j] = v;
  }
  return SGTSNE_OK;
}

}  // extern "C"

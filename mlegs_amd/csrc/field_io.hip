// Field files: the formats of msave / mload on type(scalar)
// (src/submodules/mlegs_scalar_io.f90:6-101 and :103-248), host code.
//
// Global files hold the whole array e(n1, n2, n3) (n1 = nr + 3 rows, the
// reference's layout); per-process files ("<fn>_<rank>", is_global = 0) hold
// one process's block with its global size, block size and block start.  One
// process holds the whole field here, so its block is the global array and
// its start is (0, 0, 0) (mlegs_scalar_init.f90:47, 57-59); the file is
// "<fn>_0".
//
// Binary (Fortran unformatted stream: raw bytes, no record markers):
//   global:  int32 n1 n2 n3 | complex128 e(n1,n2,n3) column-major |
//            real64 ln | int32 nrchop_offset npchop_offset nzchop_offset |
//            char[3] space
//   block:   int32 glb(3) loc(3) st(3) | same body and trailer
// ASCII:
//   global:  one record "(3(1X,I10))" with n1 n2 n3; block: three such
//            records (glb, loc, st)
//   then per slice k, per row i: one record of 2 n2 fields "1PE24.15E3"
//            (re, im of columns 1..n2), slices separated by one blank record
//   then the trailer records "ln off1 off2 off3" and "space", list-directed.
// The readers follow mload: the data records are read as fixed 24-character
// fields, the trailer as whitespace-separated values.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mlx_internal.h"

namespace {

thread_local std::string g_file_err;

struct FileError {
  std::string msg;
};

[[noreturn]] void fail(const std::string &m) { throw FileError{m}; }

// Fortran 1PE24.15E3 of a double: one digit before the point, 15 after,
// a three-digit exponent, right-justified in 24 characters
// (formatted_num_str_len = 24, mlegs_envir.f90:21; mlegs_scalar_io.f90:15-17)
void put_e24(std::string &out, double v) {
  char b[64];
  if (std::isnan(v)) {
    std::snprintf(b, sizeof b, "%24s", "NaN");
  } else if (std::isinf(v)) {
    std::snprintf(b, sizeof b, "%24s", v < 0 ? "-Infinity" : "Infinity");
  } else {
    char m[48];
    std::snprintf(m, sizeof m, "%.15E", v);  // d.dddddddddddddddE+xx[x]
    char *e = std::strchr(m, 'E');
    const int ex = std::atoi(e + 1);  // |ex| <= 324 for a double: three digits
    *e = 0;
    std::snprintf(b, sizeof b, "%19sE%c%03d", m, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
  }
  out.append(b, 24);
}

// a Fortran real input field: blanks ignored, an all-blank field is zero,
// D or d as exponent letter, an exponent without its letter (1.0+100)
double get_real(const char *p, size_t w) {
  std::string f;
  for (size_t i = 0; i < w && p[i]; ++i)
    if (p[i] != ' ' && p[i] != '\t' && p[i] != '\r') f.push_back(p[i] == 'D' || p[i] == 'd' ? 'E' : p[i]);
  if (f.empty()) return 0.0;
  if (f == "NaN" || f == "nan" || f == "NAN") return std::nan("");
  size_t ep = f.find_first_of("Ee");
  if (ep == std::string::npos) {
    const size_t s = f.find_first_of("+-", 1);
    if (s != std::string::npos) f.insert(s, "E");
  }
  char *end = nullptr;
  errno = 0;
  const double v = std::strtod(f.c_str(), &end);
  if (end == f.c_str() || *end != 0) fail("mload: bad number field '" + std::string(p, w) + "'");
  if (errno == ERANGE && std::isinf(v)) fail("mload: number field out of range '" + std::string(p, w) + "'");
  return v;
}

std::string rank_name(const std::string &fn, bool global) { return global ? fn : fn + "_0"; }

struct Trailer {
  double ln = 0.0;
  int32_t off[3] = {0, 0, 0};
  char space[4] = {'P', 'P', 'P', 0};
};

void write_binary(const std::string &path, const double *e, const int64_t n[3], const Trailer &t, bool global) {
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) fail("msave: cannot open " + path);
  auto w = [&](const void *p, size_t sz) {
    if (std::fwrite(p, 1, sz, fp) != sz) {
      std::fclose(fp);
      fail("msave: write failed on " + path);
    }
  };
  const int32_t sz[3] = {(int32_t)n[0], (int32_t)n[1], (int32_t)n[2]};
  w(sz, sizeof sz);
  if (!global) {
    const int32_t st[3] = {0, 0, 0};
    w(sz, sizeof sz);
    w(st, sizeof st);
  }
  w(e, sizeof(double) * 2 * (size_t)(n[0] * n[1] * n[2]));
  w(&t.ln, sizeof(double));
  w(t.off, sizeof t.off);
  w(t.space, 3);
  if (std::fclose(fp) != 0) fail("msave: write failed on " + path);
}

void write_ascii(const std::string &path, const double *e, const int64_t n[3], const Trailer &t, bool global) {
  FILE *fp = std::fopen(path.c_str(), "w");
  if (!fp) fail("msave: cannot open " + path);
  std::string rec;
  auto flush = [&]() {
    if (std::fwrite(rec.data(), 1, rec.size(), fp) != rec.size()) {
      std::fclose(fp);
      fail("msave: write failed on " + path);
    }
    rec.clear();
  };
  char b[64];
  // global: the size record is closed by a list-directed write of '' (a
  // blank); per-process: three plain records
  std::snprintf(b, sizeof b, " %10d %10d %10d", (int)n[0], (int)n[1], (int)n[2]);
  rec += b;
  if (global) {
    rec += " \n";
  } else {
    rec += "\n";
    rec += b;
    rec += "\n";
    std::snprintf(b, sizeof b, " %10d %10d %10d\n", 0, 0, 0);
    rec += b;
  }
  const size_t n1 = n[0], n2 = n[1], n3 = n[2];
  for (size_t k = 0; k < n3; ++k) {
    for (size_t i = 0; i < n1; ++i) {
      for (size_t j = 0; j < n2; ++j) {
        const double *z = e + 2 * (i + n1 * (j + n2 * k));
        put_e24(rec, z[0]);
        put_e24(rec, z[1]);
      }
      rec += " \n";  // the record is closed by write(fo, *) ''
      if (rec.size() > (1u << 20)) flush();
    }
    if (k + 1 < n3) rec += " \n";  // one blank record between slices
  }
  std::snprintf(b, sizeof b, " %25.17E %11d %11d %11d\n", t.ln, t.off[0], t.off[1], t.off[2]);
  rec += b;
  rec += " ";
  rec.append(t.space, 3);
  rec += "\n";
  flush();
  if (std::fclose(fp) != 0) fail("msave: write failed on " + path);
}

void size_check(const int32_t got[3], const int64_t want[3]) {
  if (got[0] != want[0] || got[1] != want[1] || got[2] != want[2]) {
    char b[256];
    std::snprintf(b, sizeof b,
                  "mloadc: size inconsistency between data and array: axis 1 (data) %d (array) %lld, "
                  "axis 2 (data) %d (array) %lld, axis 3 (data) %d (array) %lld",
                  got[0], (long long)want[0], got[1], (long long)want[1], got[2], (long long)want[2]);
    fail(b);
  }
}

void read_binary(const std::string &path, double *e, const int64_t n[3], Trailer &t, bool global) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) fail("mload_scalar: cannot open " + path);
  auto r = [&](void *p, size_t sz) {
    if (std::fread(p, 1, sz, fp) != sz) {
      std::fclose(fp);
      fail("mload: unexpected end of file " + path);
    }
  };
  int32_t sz[3];
  r(sz, sizeof sz);
  if (!global) {
    int32_t loc[3], st[3];
    r(loc, sizeof loc);
    r(st, sizeof st);
    try {
      size_check(loc, n);
    } catch (...) {
      std::fclose(fp);
      throw;
    }
  } else {
    try {
      size_check(sz, n);
    } catch (...) {
      std::fclose(fp);
      throw;
    }
  }
  r(e, sizeof(double) * 2 * (size_t)(n[0] * n[1] * n[2]));
  r(&t.ln, sizeof(double));
  r(t.off, sizeof t.off);
  r(t.space, 3);
  t.space[3] = 0;
  std::fclose(fp);
}

struct LineReader {
  FILE *fp;
  std::string path;
  std::string line;
  bool next() {
    line.clear();
    int ch;
    bool any = false;
    while ((ch = std::fgetc(fp)) != EOF) {
      any = true;
      if (ch == '\n') break;
      line.push_back((char)ch);
    }
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return any;
  }
  void need() {
    if (!next()) fail("mload: unexpected end of file " + path);
  }
};

// "(3(1X,I10))": three integers in columns 2-11, 13-22, 24-33
void read_sizes(LineReader &lr, int32_t out[3]) {
  lr.need();
  for (int q = 0; q < 3; ++q) {
    const size_t p0 = 1 + 11 * q;
    std::string f = lr.line.size() > p0 ? lr.line.substr(p0, 10) : std::string();
    std::string d;
    for (char ch : f)
      if (ch != ' ') d.push_back(ch);
    char *end = nullptr;
    const long v = d.empty() ? 0 : std::strtol(d.c_str(), &end, 10);
    if (!d.empty() && *end != 0) fail("mload: bad size record in " + lr.path);
    out[q] = (int32_t)v;
  }
}

void read_ascii(const std::string &path, double *e, const int64_t n[3], Trailer &t, bool global) {
  FILE *fp = std::fopen(path.c_str(), "r");
  if (!fp) fail("mload_scalar: cannot open " + path);
  LineReader lr{fp, path, {}};
  try {
    int32_t sz[3];
    read_sizes(lr, sz);
    if (!global) {
      int32_t loc[3], st[3];
      read_sizes(lr, loc);
      read_sizes(lr, st);
      size_check(loc, n);
    } else {
      size_check(sz, n);
    }
    const size_t n1 = n[0], n2 = n[1], n3 = n[2];
    for (size_t k = 0; k < n3; ++k) {
      for (size_t i = 0; i < n1; ++i) {
        lr.need();
        // a record shorter than 2 n2 fields reads as blanks (zeros), as a
        // Fortran formatted read pads it
        const std::string &L = lr.line;
        for (size_t j = 0; j < 2 * n2; ++j) {
          const size_t p0 = 24 * j;
          const double v = p0 < L.size() ? get_real(L.c_str() + p0, std::min<size_t>(24, L.size() - p0)) : 0.0;
          e[2 * (i + n1 * ((j / 2) + n2 * k)) + (j & 1)] = v;
        }
      }
      if (k + 1 < n3) lr.need();  // read(fo, *): skips the blank record
    }
    // list-directed trailer: ln and the three offsets, then the space
    std::vector<std::string> tok;
    while (tok.size() < 4) {
      lr.need();
      const char *s = lr.line.c_str();
      while (*s) {
        while (*s == ' ' || *s == '\t' || *s == ',') ++s;
        if (!*s) break;
        const char *b = s;
        while (*s && *s != ' ' && *s != '\t' && *s != ',') ++s;
        tok.emplace_back(b, s - b);
      }
    }
    t.ln = get_real(tok[0].c_str(), tok[0].size());
    for (int q = 0; q < 3; ++q) {
      char *end = nullptr;
      const long v = std::strtol(tok[1 + q].c_str(), &end, 10);
      if (*end != 0) fail("mload: bad offset in the trailer of " + path);
      t.off[q] = (int32_t)v;
    }
    std::string sp;
    while (sp.empty()) {
      lr.need();
      for (char ch : lr.line)
        if (ch != ' ' && ch != '\t' && ch != '\'' && ch != '"') sp.push_back(ch);
    }
    if (sp.size() < 3) fail("mload: bad space record in " + path);
    std::memcpy(t.space, sp.data(), 3);
    t.space[3] = 0;
  } catch (...) {
    std::fclose(fp);
    throw;
  }
  std::fclose(fp);
}

void write_file(const char *path, const double *host, const int64_t n[3], const Trailer &t, int is_binary,
                int is_global) {
  if (!path || !host) fail("msave: null argument");
  if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0 || n[0] > INT32_MAX || n[1] > INT32_MAX || n[2] > INT32_MAX)
    fail("msave: bad array size");
  const std::string p = rank_name(path, is_global != 0);
  if (is_binary) write_binary(p, host, n, t, is_global != 0);
  else write_ascii(p, host, n, t, is_global != 0);
}

void read_file(const char *path, double *host, const int64_t n[3], Trailer &t, int is_binary, int is_global) {
  if (!path || !host) fail("mload: null argument");
  if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) fail("mload: bad array size");
  const std::string p = rank_name(path, is_global != 0);
  if (is_binary) read_binary(p, host, n, t, is_global != 0);
  else read_ascii(p, host, n, t, is_global != 0);
}

template <class F>
int file_guarded(F &&fn) {
  try {
    fn();
    g_file_err.clear();
    return MLX_OK;
  } catch (const FileError &e) {
    g_file_err = e.msg;
    return MLX_E_FILE;
  } catch (const std::exception &e) {
    g_file_err = e.what();
    return MLX_E_FILE;
  }
}

}  // namespace

extern "C" {

int mlx_file_write(const char *path, const double *host, int64_t n1, int64_t n2, int64_t n3, const char *space,
                   double ln, const int32_t *offsets, int32_t is_binary, int32_t is_global) {
  return file_guarded([&] {
    Trailer t;
    t.ln = ln;
    for (int q = 0; q < 3; ++q) t.off[q] = offsets ? offsets[q] : 0;
    if (!space || std::strlen(space) < 3) fail("msave: space must have three characters");
    std::memcpy(t.space, space, 3);
    const int64_t n[3] = {n1, n2, n3};
    write_file(path, host, n, t, is_binary, is_global);
  });
}

int mlx_file_read(const char *path, double *host, int64_t n1, int64_t n2, int64_t n3, char *space4, double *ln,
                  int32_t *offsets, int32_t is_binary, int32_t is_global) {
  return file_guarded([&] {
    Trailer t;
    const int64_t n[3] = {n1, n2, n3};
    read_file(path, host, n, t, is_binary, is_global);
    if (space4) std::memcpy(space4, t.space, 4);
    if (ln) *ln = t.ln;
    if (offsets)
      for (int q = 0; q < 3; ++q) offsets[q] = t.off[q];
  });
}

const char *mlx_file_last_error(void) { return g_file_err.c_str(); }

int mlx_field_save(mlx_ctx *c, const mlx_field *f, const char *path, int32_t is_binary, int32_t is_global) {
  if (!c || !f) return MLX_E_ARG;
  const int64_t n1 = c->kit.nrdim_ref, n2 = c->kit.npdim, n3 = c->kit.nzdim;
  if (n1 <= 0) {
    c->err = "msave: kit not initialised";
    return MLX_E_ARG;
  }
  std::vector<double> h((size_t)(2 * n1 * n2 * n3));
  char sp[4];
  double ln = 0.0;
  int32_t off[3];
  int rc = mlx_field_download(c, f, h.data(), n1, n2, n3, sp, &ln, off);
  if (rc) return rc;
  rc = mlx_file_write(path, h.data(), n1, n2, n3, sp, ln, off, is_binary, is_global);
  if (rc) c->err = g_file_err;
  return rc;
}

int mlx_field_load(mlx_ctx *c, mlx_field *f, const char *path, int32_t is_binary, int32_t is_global) {
  if (!c || !f) return MLX_E_ARG;
  const int64_t n1 = c->kit.nrdim_ref, n2 = c->kit.npdim, n3 = c->kit.nzdim;
  if (n1 <= 0) {
    c->err = "mload: kit not initialised";
    return MLX_E_ARG;
  }
  std::vector<double> h((size_t)(2 * n1 * n2 * n3));
  char sp[4];
  double ln = 0.0;
  int32_t off[3];
  int rc = mlx_file_read(path, h.data(), n1, n2, n3, sp, &ln, off, is_binary, is_global);
  if (rc) {
    c->err = g_file_err;
    return rc;
  }
  if (off[1] != 0 || off[2] != 0) {
    c->err = "mload: only radial chopping offsets are supported on this path";
    return MLX_E_ARG;
  }
  return mlx_field_upload(c, f, h.data(), n1, n2, n3, sp, ln, off);
}

}  // extern "C"

// Host utilities behind psdmpkg.h: argument store and getpar*, par-file
// reader, list/tuple parsers, allocators, window taper, mute curve, path
// split and velocity-file coordinates.  Written for this build; behaviour
// follows the reference entry point cited on each function.
#include "psdmpkg.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>

#include <string>
#include <vector>

int xargc;
char **xargv;
int inited;

// src/psdmpkg.cc:7-22: keep a private copy of argv.
void initargs_inner(int argc, char *argv[], const char *sdoc[]) {
  (void)sdoc;
  xargc = argc;
  xargv = (char **)malloc(sizeof(char *) * (argc + 1));
  for (int i = 0; i < argc; ++i) xargv[i] = strdup(argv[i]);
  xargv[argc] = NULL;
  inited = 1;
}

namespace {

void *checked(void *p, const char *what, size_t a, size_t b) {
  if (p == NULL) {
    fprintf(stderr, "%s error when call %s(%zu,%zu)\n", what, what, a, b);
    exit(0);
  }
  return p;
}

template <typename T>
T **rows(size_t n1, size_t n2, const char *what) {
  T **p = (T **)checked(malloc(sizeof(T *) * n2), what, n1, n2);
  p[0] = (T *)checked(malloc(sizeof(T) * n1 * n2), what, n1, n2);
  for (size_t i = 1; i < n2; ++i) p[i] = p[0] + i * n1;
  return p;
}

// getpar lookup: first stored argument of the form "key=..."
// (src/psdmpkg.cc:233-336); exit(0) before initargs.
const char *find_arg(const char *key) {
  if (!inited) {
    fprintf(stderr, "argument is not initialized \n");
    exit(0);
  }
  const size_t n = strlen(key);
  for (int i = 0; i < xargc; ++i)
    if (strncmp(key, xargv[i], n) == 0 && xargv[i][n] == '=') return xargv[i] + n + 1;
  return NULL;
}

}  // namespace

// src/psdmpkg.cc:24-107 (and the 3-D/4-D/double forms the header declares).
char *malloc1c(size_t n1) { return (char *)checked(malloc(n1), "malloc1c", n1, 1); }
int *malloc1i(size_t n1) { return (int *)checked(malloc(sizeof(int) * n1), "malloc1i", n1, 1); }
float *malloc1f(size_t n1) { return (float *)checked(malloc(sizeof(float) * n1), "malloc1f", n1, 1); }
double *malloc1d(size_t n1) { return (double *)checked(malloc(sizeof(double) * n1), "malloc1d", n1, 1); }
char **malloc2c(size_t n1, size_t n2) { return rows<char>(n1, n2, "malloc2c"); }
int **malloc2i(size_t n1, size_t n2) { return rows<int>(n1, n2, "malloc2i"); }
float **malloc2f(size_t n1, size_t n2) { return rows<float>(n1, n2, "malloc2f"); }
double **malloc2d(size_t n1, size_t n2) { return rows<double>(n1, n2, "malloc2d"); }

int ***malloc3i(size_t n1, size_t n2, size_t n3) {
  int ***p = (int ***)checked(malloc(sizeof(int **) * n3), "malloc3i", n1, n2);
  int **r = rows<int>(n1, n2 * n3, "malloc3i");
  for (size_t k = 0; k < n3; ++k) p[k] = r + k * n2;
  return p;
}

float ***malloc3f(size_t n1, size_t n2, size_t n3) {
  float ***p = (float ***)checked(malloc(sizeof(float **) * n3), "malloc3f", n1, n2);
  float **r = rows<float>(n1, n2 * n3, "malloc3f");
  for (size_t k = 0; k < n3; ++k) p[k] = r + k * n2;
  return p;
}

float ****malloc4f(size_t n1, size_t n2, size_t n3, size_t n4) {
  float ****p = (float ****)checked(malloc(sizeof(float ***) * n4), "malloc4f", n1, n2);
  float ***q = (float ***)checked(malloc(sizeof(float **) * n3 * n4), "malloc4f", n3, n4);
  float **r = rows<float>(n1, n2 * n3 * n4, "malloc4f");
  for (size_t k = 0; k < n3 * n4; ++k) q[k] = r + k * n2;
  for (size_t l = 0; l < n4; ++l) p[l] = q + l * n3;
  return p;
}

void free1c(char *p) { free(p); }
void free1i(int *p) { free(p); }
void free1f(float *p) { free(p); }
void free1d(double *p) { free(p); }
template <typename T>
static void free_rows(T **p) {
  if (p == NULL) return;
  free(p[0]);
  free(p);
}
void free2c(char **p) { free_rows(p); }
void free2i(int **p) { free_rows(p); }
void free2f(float **p) { free_rows(p); }
void free2d(double **p) { free_rows(p); }
void free3i(int ***p) {
  if (p == NULL) return;
  if (p[0]) free_rows(p[0]);
  free(p);
}
void free3f(float ***p) {
  if (p == NULL) return;
  if (p[0]) free_rows(p[0]);
  free(p);
}
void free4f(float ****p) {
  if (p == NULL) return;
  if (p[0]) {
    if (p[0][0]) free_rows(p[0][0]);
    free(p[0]);
  }
  free(p);
}

// List parser, src/psdmpkg.cc:108-204: integers separated by ',' (next
// value), '-' (inclusive range from the previous value) and ':' (a:s:b,
// stepped range replacing a).  Descending '-' ranges and zero steps are
// rejected (NULL, *n = 0); the reference's behaviour there is undefined.
int *parseint(const char *str, int *n) {
  *n = 0;
  std::vector<int> val;
  std::vector<char> sep;
  val.push_back((int)strtol(str, NULL, 10));
  for (const char *c = str; *c; ++c)
    if (*c == ',' || *c == '-' || *c == ':') {
      sep.push_back(*c);
      val.push_back((int)strtol(c + 1, NULL, 10));
    }
  std::vector<int> out;
  out.push_back(val[0]);
  for (size_t i = 0; i < sep.size(); ++i) {
    if (sep[i] == ',') {
      out.push_back(val[i + 1]);
    } else if (sep[i] == '-') {
      if (val[i + 1] < val[i]) return NULL;
      for (int v = val[i] + 1; v <= val[i + 1]; ++v) out.push_back(v);
    } else {
      if (i + 2 >= val.size() || val[i + 1] == 0) return NULL;
      const int a = val[i], s = val[i + 1], b = val[i + 2];
      out.pop_back();
      if (s > 0)
        for (int v = a; v <= b; v += s) out.push_back(v);
      else
        for (int v = a; v >= b; v += s) out.push_back(v);
      ++i;
    }
  }
  int *p = (int *)malloc(sizeof(int) * (out.size() ? out.size() : 1));
  memcpy(p, out.data(), sizeof(int) * out.size());
  *n = (int)out.size();
  return p;
}

// src/psdmpkg.cc:206.
int fexist(const char *path) { return access(path, R_OK | W_OK) == 0; }

// src/psdmpkg.cc:208-231: copy [it1, it2] with zero fill outside [0, ns),
// then raised-cosine tapers over the first and last min(200, len) samples.
int fetch_and_taper(const float *data, int ns, int it1, int it2, float *out) {
  const int len = it2 - it1 + 1;
  for (int i = 0; i < len; ++i) {
    const int it = it1 + i;
    out[i] = (it < 0 || it >= ns) ? 0.0f : data[it];
  }
  for (int i = 0; i < 200 && i < len; ++i) {
    const float c = 0.5 - 0.5 * cosf(M_PI * i / 200.0);
    out[i] *= c;
    out[len - 1 - i] *= c;
  }
  return 1;
}

int getparint(const char *key, int *val) {
  const char *v = find_arg(key);
  if (!v) return 0;
  *val = (int)strtol(v, NULL, 10);
  return 1;
}

int getparfloat(const char *key, float *val) {
  const char *v = find_arg(key);
  if (!v) return 0;
  *val = (float)strtod(v, NULL);
  return 1;
}

int getpardouble(const char *key, double *val) {
  const char *v = find_arg(key);
  if (!v) return 0;
  *val = strtod(v, NULL);
  return 1;
}

int getparstr(const char *key, char **val) {
  const char *v = find_arg(key);
  if (!v) return 0;
  *val = const_cast<char *>(v);
  return 1;
}

int getparstr(const char *key, const char **val) {
  const char *v = find_arg(key);
  if (!v) return 0;
  *val = v;
  return 1;
}

namespace {

// Logical records of a par file (src/psdmpkg.cc:413-489): physical lines
// joined across a trailing backslash, cut at the first '#', trimmed of
// surrounding white space; comment-only and blank records are dropped.
std::vector<std::string> par_records(FILE *fp) {
  std::vector<std::string> recs;
  std::string cur;
  char buf[1024];
  while (fgets(buf, sizeof(buf), fp)) {
    cur += buf;
    if (cur.empty() || cur.back() != '\n') continue;
    cur.pop_back();
    if (!cur.empty() && cur.back() == '\\') {
      cur.pop_back();
      continue;
    }
    std::string r = cur;
    cur.clear();
    size_t b = 0;
    while (b < r.size() && isspace((unsigned char)r[b])) ++b;
    r = r.substr(b);
    if (!r.empty() && r[0] == '#') continue;
    const size_t h = r.find('#');
    if (h != std::string::npos) r.resize(h);
    while (!r.empty() && isspace((unsigned char)r.back())) r.pop_back();
    if (!r.empty()) recs.push_back(r);
  }
  if (!cur.empty()) {
    // a last line without newline is not a complete record in the reference
  }
  return recs;
}

}  // namespace

// src/psdmpkg.cc:338-574.  fmt is "k1=%d,k2=%f,..." (conversions d f g s S);
// each pattern is matched with sscanf against every record; values are
// stored into the variadic pointers in the order matches occur.  Returns the
// number stored, or -1 if any pattern matched other than exactly once.
int readparfile(const char *filename, const char *fmt, ...) {
  int ncomma = 0, nequal = 0, npercent = 0;
  for (const char *c = fmt; *c; ++c) {
    ncomma += *c == ',';
    nequal += *c == '=';
    npercent += *c == '%';
  }
  if (ncomma + 1 != nequal || nequal != npercent) {
    fprintf(stderr, "Error: In FUNCTION: %s fmt[%s] [, = %%] do not Match\n", __FUNCTION__, fmt);
    exit(0);
  }
  FILE *fp = fopen(filename, "r");
  if (fp == NULL) {
    fprintf(stderr, "Error: Parameter File [%s] Open Error\n", filename);
    exit(0);
  }
  const std::vector<std::string> recs = par_records(fp);
  fclose(fp);
  std::vector<std::string> pats;
  {
    std::string f = fmt;
    size_t s = 0;
    for (;;) {
      const size_t e = f.find(',', s);
      pats.push_back(f.substr(s, e == std::string::npos ? std::string::npos : e - s));
      if (e == std::string::npos) break;
      s = e + 1;
    }
  }
  va_list ap;
  va_start(ap, fmt);
  int nret = 0;
  bool bad = false;
  for (const std::string &pat : pats) {
    int hits = 0;
    const char conv = pat.empty() ? 0 : pat.back();
    for (const std::string &r : recs) {
      bool hit = false;
      if (conv == 'd') {
        int v;
        if (sscanf(r.c_str(), pat.c_str(), &v) == 1) {
          hit = true;
          if (nret >= nequal) { va_end(ap); return -1; }
          *va_arg(ap, int *) = v;
        }
      } else if (conv == 'f' || conv == 'g') {
        float v;
        if (sscanf(r.c_str(), pat.c_str(), &v) == 1) {
          hit = true;
          if (nret >= nequal) { va_end(ap); return -1; }
          *va_arg(ap, float *) = v;
        }
      } else if (conv == 's') {
        char ss[1024];
        if (sscanf(r.c_str(), pat.c_str(), ss) == 1) {
          hit = true;
          if (nret >= nequal) { va_end(ap); return -1; }
          char *dst = va_arg(ap, char *);
          int j = 0;
          while (ss[j] && j < 1023) dst[j] = ss[j], ++j;
          dst[j] = 0;
        }
      } else if (conv == 'S') {
        const size_t k = pat.size() - 2;
        if (r.compare(0, k, pat, 0, k) == 0) {
          hit = true;
          if (nret >= nequal) { va_end(ap); return -1; }
          *va_arg(ap, char **) = strdup(r.c_str() + k);
        }
      }
      if (hit) {
        ++hits;
        ++nret;
      }
    }
    if (hits != 1) bad = true;
  }
  va_end(ap);
  return bad ? -1 : nret;
}

// src/psdmpkg.cc:576-599: make the (time, offset) pairs increasing.
void time_offset_mute_santity(float *tm, int n) {
  for (int i = 0; i < n - 1; ++i) {
    if (tm[2 * i] > tm[2 * i + 2]) tm[2 * i + 2] = 2.0 * tm[2 * i] - tm[2 * i + 2];
    if (tm[2 * i + 1] > tm[2 * i + 3]) tm[2 * i + 3] = 2.0 * tm[2 * i + 1] - tm[2 * i + 3];
  }
  for (int i = 0; i < n - 1; ++i) {
    if (tm[2 * i] >= tm[2 * i + 2]) tm[2 * i + 2] = tm[2 * i] + 0.01;
    if (tm[2 * i + 1] >= tm[2 * i + 3]) tm[2 * i + 3] = tm[2 * i + 1] + 10.0;
  }
}

// src/psdmpkg.cc:601-608.
char *strrev1(char *p) {
  if (!p || !*p) return p;
  char *a = p, *b = p + strlen(p) - 1;
  for (; a < b; ++a, --b) {
    const char t = *a;
    *a = *b;
    *b = t;
  }
  return a;
}

size_t fsize(const char *path) {
  struct stat st;
  if (stat(path, &st) != 0) return 0;
  return (size_t)st.st_size;
}

// src/psdmpkg.cc:616-631.
char *strrtrim(char *src, int c) {
  if (!src) return src;
  for (int i = (int)strlen(src) - 1; i >= 0 && src[i] == c; --i) src[i] = 0;
  return src;
}

// src/psdmpkg.cc:633-684: ".../<projname>/<iproc>/<task>" -> the pieces;
// projpath is the path up to (excluding) the "/<iproc>/<task>" tail.
void parsepath(char *path, char *projpath, char *projname, int *iproc, char *task) {
  const int len = (int)strlen(path);
  int i = len - 1;
  int j = 0;
  for (; i >= 0 && path[i] != '/'; --i) task[j++] = path[i];
  task[j] = 0;
  strrev1(task);
  --i;
  char buf[16];
  j = 0;
  for (; i >= 0 && path[i] != '/'; --i)
    if (j < 15) buf[j++] = path[i];
  strcpy(projpath, path);
  projpath[i < 0 ? 0 : i] = 0;
  buf[j] = 0;
  strrev1(buf);
  --i;
  sscanf(buf, "%d", iproc);
  j = 0;
  for (; i >= 0 && path[i] != '/'; --i) projname[j++] = path[i];
  projname[j] = 0;
  strrev1(projname);
}

namespace {
inline bool num_char(char c) { return (c >= '0' && c <= '9') || c == '.' || c == 'E' || c == 'e'; }
inline bool sep_char(char c) { return c == ' ' || c == '\t' || c == ',' || c == ';'; }
}  // namespace

// src/psdmpkg.cc:686-862: "(a,b),(c,d),..." -> ntuple x nelem floats; every
// tuple must hold as many numbers as the first (else NULL and zeros).
// Number tokens are runs of digits, '.', 'e', 'E' (no sign), as there.
float *parsetuple(const char *str, int *ntuple, int *nelem) {
  *ntuple = 0;
  *nelem = 0;
  int nl = 0, nr = 0;
  for (const char *c = str; *c; ++c) nl += *c == '(', nr += *c == ')';
  if (nl != nr || nl == 0) {
    fprintf(stderr, "Warn: String [%s]'s brackets do not match,or has no element\n", str);
    return NULL;
  }
  std::vector<std::vector<float>> tup;
  const char *c = str;
  while ((c = strchr(c, '(')) != NULL) {
    const char *e = strchr(c, ')');
    if (!e) break;
    std::vector<float> v;
    for (const char *q = c + 1; q < e;) {
      if (num_char(*q)) {
        char buf[128];
        int j = 0;
        while (q < e && num_char(*q)) {
          if (j < 127) buf[j++] = *q;
          ++q;
        }
        buf[j] = 0;
        float f;
        if (sscanf(buf, "%f", &f) != 1) {
          fprintf(stderr, "Error: Elements Parse Error\n");
          return NULL;
        }
        v.push_back(f);
      } else {
        ++q;
      }
    }
    tup.push_back(v);
    c = e + 1;
  }
  for (const auto &t : tup)
    if (t.size() != tup[0].size()) {
      fprintf(stderr, "Error: String [%s]'s Elements Count do not match\n", str);
      return NULL;
    }
  const size_t ne = tup[0].size();
  const size_t cnt = tup.size() * ne;
  float *p = (float *)malloc(sizeof(float) * (cnt ? cnt : 1));
  for (size_t i = 0; i < tup.size(); ++i)
    for (size_t k = 0; k < ne; ++k) p[i * ne + k] = tup[i][k];
  *ntuple = (int)tup.size();
  *nelem = (int)ne;
  return p;
}

// src/psdmpkg.cc:1013-1025.
extern "C" float getscale(short scalco) {
  if (scalco == 0) return 1.0f;
  if (scalco > 0) return (float)scalco;
  return -1.0f / scalco;
}

// src/psdmpkg.cc:979-1008: coordinates of the nxi velocity traces of `line`
// starting at CDP fci (line clamped to the last velocity line).
extern "C" void get_vel_coor(const char *velfilepath, int line, coor_t *vel_coor, short *vel_scalel,
                             int fci, int nxi, int fcv, int nxv, int flv, int nyv) {
  FILE *fp = fopen(velfilepath, "rb");
  if (fp == NULL) {
    fprintf(stderr, "Velocity File open Error\n");
    exit(1);
  }
  segy *tr = (segy *)malloc(sizeof(segy));
  fgettr(fp, tr);
  *vel_scalel = tr->scalco;
  if (line >= flv + nyv) line = flv + nyv - 1;
  fseek(fp, (long)(sizeof(float) * tr->ns + 240L) * ((long)(line - flv) * nxv + fci - fcv), SEEK_SET);
  for (int i = 0; i < nxi; ++i) {
    fgettr(fp, tr);
    vel_coor[i].sx = tr->sx;
    vel_coor[i].sy = tr->sy;
  }
  free(tr);
  fclose(fp);
}

// src/psdmpkg.cc:1030-1070: linear interpolation of the mute time at an
// offset from (time, offset) pairs, extrapolating the end segments.
extern "C" float time_offset_mute(float *tm, int n, float o) {
  if (n < 2) return tm[0];
  float o1, o2, t1, t2;
  if (o < tm[1]) {
    o1 = tm[1], o2 = tm[3], t1 = tm[0], t2 = tm[2];
  } else if (o >= tm[2 * (n - 1) + 1]) {
    o1 = tm[2 * (n - 2) + 1], o2 = tm[2 * (n - 1) + 1], t1 = tm[2 * (n - 2)], t2 = tm[2 * (n - 1)];
  } else {
    int k = 0;
    for (int i = 0; i < n - 1; ++i)
      if (tm[2 * i + 1] <= o && o < tm[2 * i + 3]) k = i;
    o1 = tm[2 * k + 1], o2 = tm[2 * k + 3], t1 = tm[2 * k], t2 = tm[2 * k + 2];
  }
  return t1 + (t2 - t1) / (o2 - o1) * (o - o1);
}

// ---- C-linkage forms --------------------------------------------------
extern "C" {
int qpsdm_pkg_fetch_and_taper(const float *d, int ns, int a, int b, float *o) {
  return fetch_and_taper(d, ns, a, b, o);
}
int *qpsdm_pkg_parseint(const char *s, int *n) { return parseint(s, n); }
float *qpsdm_pkg_parsetuple(const char *s, int *a, int *b) { return parsetuple(s, a, b); }
int qpsdm_pkg_readparfile_i(const char *f, const char *fmt, int *v) { return readparfile(f, fmt, v); }
int qpsdm_pkg_readparfile_f(const char *f, const char *fmt, float *v) { return readparfile(f, fmt, v); }
int qpsdm_pkg_readparfile_s(const char *f, const char *fmt, char *v) { return readparfile(f, fmt, v); }
int qpsdm_pkg_readparfile_ii(const char *f, const char *fmt, int *a, int *b) {
  return readparfile(f, fmt, a, b);
}
void qpsdm_pkg_initargs(int argc, char **argv) { initargs_inner(argc, argv, NULL); }
int qpsdm_pkg_getparint(const char *k, int *v) { return getparint(k, v); }
int qpsdm_pkg_getparfloat(const char *k, float *v) { return getparfloat(k, v); }
int qpsdm_pkg_getpardouble(const char *k, double *v) { return getpardouble(k, v); }
int qpsdm_pkg_getparstr(const char *k, char **v) { return getparstr(k, v); }
void qpsdm_pkg_time_offset_mute_santity(float *tm, int n) { time_offset_mute_santity(tm, n); }
void qpsdm_pkg_parsepath(char *p, char *a, char *b, int *i, char *t) { parsepath(p, a, b, i, t); }
char *qpsdm_pkg_strrtrim(char *s, int c) { return strrtrim(s, c); }
void qpsdm_pkg_free(void *p) { free(p); }
}

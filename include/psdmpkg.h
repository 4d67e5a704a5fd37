/*
 * psdmpkg.h -- host utility entry points of the imaging path.
 *
 * Same names, signatures, ownership and error behaviour as the reference's
 * src/psdmpkg.h:42-212 (C++ linkage, except getscale, get_vel_coor and
 * time_offset_mute which are extern "C" as there).  Errors print to stderr
 * and exit(0) or exit(1) exactly where the reference does; readparfile
 * returns the number of values assigned, or -1 when a key is missing or
 * repeated.  getparstr returns a pointer into the stored argument copy.
 * parse* return malloc'ed arrays owned by the caller; malloc2* return a
 * row-pointer array whose storage is contiguous at p[0].  Not thread-safe
 * (the argument store is global), like the reference.
 */
#ifndef PSDMPKG_H_
#define PSDMPKG_H_

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "segy.h"

#ifdef __cplusplus
extern int xargc;
extern char **xargv;
extern int inited;

#define alloc1char(n1) malloc1c(n1)
#define alloc2char(n1, n2) malloc2c(n1, n2)
#define alloc1float(n1) malloc1f(n1)
#define alloc1int(n1) malloc1i(n1)
#define alloc2float(n1, n2) malloc2f(n1, n2)
#define alloc2int(n1, n2) malloc2i(n1, n2)
#define getparstring(key, val) getparstr(key, val)
#define free1char(p) free1c(p)
#define free2char(p) free2c(p)
#define free1int(p) free1i(p)
#define free2int(p) free2i(p)
#define free3int(p) free3i(p)
#define free1float(p) free1f(p)
#define free2float(p) free2f(p)
#define free3float(p) free3f(p)
#define free4float(p) free4f(p)
#define free1double(p) free1d(p)
#define free2double(p) free2d(p)

char *malloc1c(size_t n1);
char **malloc2c(size_t n1, size_t n2);
int *malloc1i(size_t n1);
int **malloc2i(size_t n1, size_t n2);
int ***malloc3i(size_t n1, size_t n2, size_t n3);
float *malloc1f(size_t n1);
float **malloc2f(size_t n1, size_t n2);
float ***malloc3f(size_t n1, size_t n2, size_t n3);
float ****malloc4f(size_t n1, size_t n2, size_t n3, size_t n4);
double *malloc1d(size_t n1);
double **malloc2d(size_t n1, size_t n2);

#define parseline(str, n) parseint(str, n)
int *parseint(const char *str, int *n);

void free1c(char *p);
void free2c(char **p);
void free1i(int *p);
void free2i(int **p);
void free3i(int ***p);
void free1f(float *p);
void free2f(float **p);
void free3f(float ***p);
void free4f(float ****p);
void free1d(double *p);
void free2d(double **p);

void initargs_inner(int argc, char *argv[], const char *sdoc[]);

/* With no arguments and a terminal on stdin, print the self-doc and exit. */
#define initargs(argc, argv, sdoc)                                     \
  do {                                                                 \
    if ((argc) == 1 && isatty(fileno(stdin))) {                        \
      for (int i_ = 0; (sdoc)[i_] != NULL; ++i_) fputs((sdoc)[i_], stderr); \
      exit(0);                                                         \
    }                                                                  \
    initargs_inner(argc, argv, sdoc);                                  \
  } while (0)

#define egetparstring(name, value)                            \
  do {                                                        \
    if (!getparstring(name, value)) {                         \
      fprintf(stderr, "Error: %s must be specified\n", name); \
      exit(0);                                                \
    }                                                         \
  } while (0)
#define egetparint(name, value)                               \
  do {                                                        \
    if (!getparint(name, value)) {                            \
      fprintf(stderr, "Error: %s must be specified\n", name); \
      exit(0);                                                \
    }                                                         \
  } while (0)
#define egetparfloat(name, value)                             \
  do {                                                        \
    if (!getparfloat(name, value)) {                          \
      fprintf(stderr, "Error: %s must be specified\n", name); \
      exit(0);                                                \
    }                                                         \
  } while (0)
#define egetpardouble(name, value)                            \
  do {                                                        \
    if (!getpardouble(name, value)) {                         \
      fprintf(stderr, "Error: %s must be specified\n", name); \
      exit(0);                                                \
    }                                                         \
  } while (0)

size_t fsize(const char *path);
int getparint(const char *key, int *val);
int getparfloat(const char *key, float *val);
int getpardouble(const char *key, double *val);
int getparstr(const char *key, char **val);
int getparstr(const char *key, const char **val);
char *strrev1(char *p);

#define exist(_p) fexist(_p)
int fexist(const char *path);

int fetch_and_taper(const float *data, int ns, int it1, int it2, float *data1);
int readparfile(const char *filename, const char *fmt, ...);
void time_offset_mute_santity(float *tomute, int ntomute);
char *strrtrim(char *src, int c);
void parsepath(char *path, char *projpath, char *projname, int *iproc, char *task);
float *parsetuple(const char *str, int *ntuple, int *nelem);

typedef enum { POLY6THCUT = 8, TANER4THCUT = 6, TANER2TH = 4, DSR = 2, OTHERS = 2 } TRAVELTIMEMETHOD;
typedef enum { AMPLITUDE = 1, AMPLITUDE_PHASE = 2 } CORRECTION;
#endif /* __cplusplus */

typedef struct {
  int sx;
  int sy;
} coor_t;

#ifdef __cplusplus
extern "C" {
#endif

float getscale(short scalco);
void get_vel_coor(const char *velfilepath, int line, coor_t *vel_coor, short *vel_scalel, int fci,
                  int nxi, int fcv, int nxv, int flv, int nyv);
float time_offset_mute(float *tomute, int ntomute, float offset);

/* C-linkage forms of the C++ entry points for FFI callers. */
int qpsdm_pkg_fetch_and_taper(const float *data, int ns, int it1, int it2, float *data1);
int *qpsdm_pkg_parseint(const char *str, int *n);
float *qpsdm_pkg_parsetuple(const char *str, int *ntuple, int *nelem);
int qpsdm_pkg_readparfile_i(const char *file, const char *fmt, int *v);
int qpsdm_pkg_readparfile_f(const char *file, const char *fmt, float *v);
int qpsdm_pkg_readparfile_s(const char *file, const char *fmt, char *v);
int qpsdm_pkg_readparfile_ii(const char *file, const char *fmt, int *a, int *b);
void qpsdm_pkg_initargs(int argc, char **argv);
int qpsdm_pkg_getparint(const char *key, int *v);
int qpsdm_pkg_getparfloat(const char *key, float *v);
int qpsdm_pkg_getpardouble(const char *key, double *v);
int qpsdm_pkg_getparstr(const char *key, char **v);
void qpsdm_pkg_time_offset_mute_santity(float *tm, int n);
void qpsdm_pkg_parsepath(char *path, char *proj, char *name, int *iproc, char *task);
char *qpsdm_pkg_strrtrim(char *s, int c);
void qpsdm_pkg_free(void *p);

#ifdef __cplusplus
}
#endif
#endif

/* dlopen_check.c -- the binding an Infernal build uses: it defines the
 * reference's own scanners and model functions (FastCYKScan, FreeCM,
 * cm_Configure, cm_tophits_*, ...), so it loads libinfernal_gpu.so with
 * dlopen(RTLD_LOCAL) and takes the GPU library's entry points with dlsym.
 * This program stands in for such a host: it exports functions with the same
 * names (linked -rdynamic, the worst case for symbol interposition) that
 * must never run.  Passing means: dlsym returns the library's functions, and
 * the library's calls among its own functions stay inside it.
 *
 *   dlopen_check <libinfernal_gpu.so> <model.cm>
 */
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "infernal_gpu.h"

/* the host's own definitions (same names as the library's exports) */
#define HOST_STUB(ret, name, args) ret name args { fprintf(stderr, "host's own %s was called\n", #name); abort(); }
HOST_STUB(int, FastCYKScan, (CM_t *a, char *b, CM_SCAN_MX *c, int d, ESL_DSQ *e, int64_t f, int64_t g, float h, CM_TOPHITS *i,
                             int j, float k, int64_t *l, int64_t *m, float **n, float *o))
HOST_STUB(void, FreeCM, (CM_t *cm))
HOST_STUB(int, cm_Configure, (CM_t *cm, char *errbuf, int W))
HOST_STUB(int, cm_scan_mx_Create, (CM_t *cm, char *errbuf, int a, int b, CM_SCAN_MX **c))
HOST_STUB(void, cm_scan_mx_Destroy, (CM_t *cm, CM_SCAN_MX *smx))
HOST_STUB(CM_TOPHITS *, cm_tophits_Create, (void))
HOST_STUB(int, cm_tophits_CreateNextHit, (CM_TOPHITS *th, CM_HIT **h))
HOST_STUB(int, cm_tophits_CloneHitMostly, (CM_TOPHITS *s, int h, CM_TOPHITS *d))
HOST_STUB(void, FLogsumInit, (void))

typedef int (*read_fn)(const char *, CM_t **, char *);
typedef int (*configure_fn)(CM_t *, char *, int);
typedef void (*free_fn)(CM_t *);
typedef CM_TOPHITS *(*th_create_fn)(void);
typedef void (*th_destroy_fn)(CM_TOPHITS *);
typedef int (*digitize_fn)(const char *, int64_t, ESL_DSQ *);
typedef int (*cyk_fn)(CM_t *, char *, CM_SCAN_MX *, int, ESL_DSQ *, int64_t, int64_t, float, CM_TOPHITS *, int, float,
                      int64_t *, int64_t *, float **, float *);
typedef int (*trcyk_fn)(CM_t *, char *, CM_TR_SCAN_MX *, int, int, ESL_DSQ *, int64_t, int64_t, float, CM_TOPHITS *, int,
                        float, int64_t *, int64_t *, float **, char *, float *);

static unsigned bits(float x) { unsigned u; memcpy(&u, &x, 4); return u; }

#define SYM(var, type, name) type var = (type) dlsym(h, name); if (!var) { fprintf(stderr, "dlsym %s: %s\n", name, dlerror()); return 2; }

int main(int argc, char **argv)
{
  if (argc != 3) { fprintf(stderr, "usage: dlopen_check <libinfernal_gpu.so> <model.cm>\n"); return 2; }
  void *h = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
  if (!h) { fprintf(stderr, "dlopen: %s\n", dlerror()); return 2; }
  SYM(cm_read, read_fn, "infx_cm_Read")
  SYM(configure, configure_fn, "cm_Configure")
  SYM(free_cm, free_fn, "FreeCM")
  SYM(th_create, th_create_fn, "cm_tophits_Create")
  SYM(th_destroy, th_destroy_fn, "cm_tophits_Destroy")
  SYM(digitize, digitize_fn, "infx_abc_Digitize")
  SYM(cyk, cyk_fn, "FastCYKScan")
  SYM(trcyk, trcyk_fn, "RefTrCYKScan")
  if ((void *) cyk == (void *) FastCYKScan || (void *) free_cm == (void *) FreeCM) { fprintf(stderr, "dlsym gave the host's own\n"); return 1; }

  char errbuf[eslERRBUFSIZE];
  CM_t *cm = NULL;
  if (cm_read(argv[2], &cm, errbuf) != eslOK) { fprintf(stderr, "read: %s\n", errbuf); return 1; }
  cm->config_opts = CM_CONFIG_LOCAL | CM_CONFIG_SCANMX | CM_CONFIG_TRUNC | CM_CONFIG_TRSCANMX;
  if (configure(cm, errbuf, -1) != eslOK) { fprintf(stderr, "configure: %s\n", errbuf); return 1; }
  const char *seq = "GCCGGGAUAGCUCAGUUGGUAGAGCAGAGGACUGAAAAUCCUCGUGUCACCGGUUCGAUUCCGGUCCCGGCACCA";
  int64_t L = (int64_t) strlen(seq);
  ESL_DSQ dsq[128];
  if (digitize(seq, L, dsq) != eslOK) return 1;
  CM_TOPHITS *th = th_create();
  int64_t ei = 0, ej = 0; float sc1 = 0.0f, sc2 = 0.0f; char mode = 0;
  int s1 = cyk(cm, errbuf, cm->smx, SMX_QDB1_TIGHT, dsq, 1, L, 0.0f, th, 0, 0.0f, &ei, &ej, NULL, &sc1);
  printf("FastCYKScan status %d sc %08x\n", s1, bits(sc1));
  int s2 = trcyk(cm, errbuf, cm->trsmx, SMX_QDB1_TIGHT, PLI_PASS_5P_AND_3P_FORCE, dsq, 1, L, 0.0f, th, 0, 0.0f, &ei, &ej, NULL, &mode, &sc2);
  printf("RefTrCYKScan status %d sc %08x mode %d\n", s2, bits(sc2), (int) mode);
  printf("%llu hits\n", (unsigned long long) th->N);
  if (!((s1 == eslOK && s2 == eslOK) || (s1 == eslENODEVICE && s2 == eslENODEVICE))) return 1;
  th_destroy(th);
  free_cm(cm);
  dlclose(h);
  printf("dlopen ok\n");
  return 0;
}

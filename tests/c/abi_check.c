/* abi_check.c -- a C caller of libinfernal_gpu, compiled with gcc against
 * include/infernal_gpu.h only (no C++ and no HIP headers), the way Infernal's
 * cm_pipeline.c would call the scanners.
 *
 *   abi_check scan <model.cm> <seq.txt>
 *       read and configure the model (local, truncation on, both scan
 *       matrices), digitise the sequence (one line of IUPAC letters) and call
 *       the five scanners through their declared prototypes; print one line
 *       per scanner: name, status, best score (decimal and IEEE bits), best
 *       mode, envelope and hit count, then each hit.  Exit 0 when every scanner
 *       returned eslOK, or every one returned eslENODEVICE (no GPU here).
 *   abi_check teardown <model.cm>
 *       attach truncated-scanner resources to the model's truncated matrix as
 *       a GPU scan leaves them (infx_debug_attach_trdev), then FreeCM; under
 *       AddressSanitizer this checks that the matrix releases what it owns.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "infernal_gpu.h"

static CM_t *load(const char *path, int trunc)
{
  char errbuf[eslERRBUFSIZE];
  CM_t *cm = NULL;
  if (infx_cm_Read(path, &cm, errbuf) != eslOK) { fprintf(stderr, "read %s: %s\n", path, errbuf); exit(2); }
  cm->config_opts = CM_CONFIG_LOCAL | CM_CONFIG_SCANMX | (trunc ? (CM_CONFIG_TRUNC | CM_CONFIG_TRSCANMX) : 0);
  if (cm_Configure(cm, errbuf, -1) != eslOK) { fprintf(stderr, "configure: %s\n", errbuf); exit(2); }
  return cm;
}

static char *slurp(const char *path, int64_t *ret_L)
{
  FILE *fp = fopen(path, "rb");
  if (!fp) { fprintf(stderr, "open %s\n", path); exit(2); }
  size_t cap = 1 << 16, n = 0;
  char *s = malloc(cap);
  int c;
  while ((c = fgetc(fp)) != EOF) {
    if (c == '\n' || c == '\r' || c == ' ') continue;
    if (n + 1 >= cap) s = realloc(s, cap *= 2);
    s[n++] = (char) c;
  }
  s[n] = 0;
  fclose(fp);
  *ret_L = (int64_t) n;
  return s;
}

static unsigned bits(float f) { unsigned u; memcpy(&u, &f, 4); return u; }

static int report(const char *name, int status, const char *errbuf, float sc, char mode, int64_t envi, int64_t envj,
                  CM_TOPHITS *th)
{
  printf("%s status %d sc %.6f bits %08x mode %d env %lld %lld nhits %llu\n", name, status, sc, bits(sc), (int) mode,
         (long long) envi, (long long) envj, (unsigned long long) th->N);
  for (uint64_t h = 0; h < th->N; h++) {
    const CM_HIT *x = &th->unsrt[h];
    printf("  hit %lld %lld %d %d %08x\n", (long long) x->start, (long long) x->stop, x->root, (int) x->mode, bits(x->score));
  }
  if (status != eslOK) printf("  errbuf: %s\n", errbuf);
  return status;
}

static int run_scans(const char *cmpath, const char *seqpath)
{
  char errbuf[eslERRBUFSIZE];
  CM_t *cm = load(cmpath, 1);
  int64_t L;
  char *seq = slurp(seqpath, &L);
  ESL_DSQ *dsq = malloc((size_t) L + 2);
  if (infx_abc_Digitize(seq, L, dsq) != eslOK) { fprintf(stderr, "digitize failed\n"); return 2; }
  int st[5], nok = 0, nodev = 0;
  for (int s = 0; s < 5; s++) {
    CM_TOPHITS *th = cm_tophits_Create();
    int64_t envi = 0, envj = 0;
    float sc = 0.f, *vsc = NULL;
    char mode = TRMODE_J;
    const char *name = "";
    errbuf[0] = 0;
    if (s == 1 || s == 2 || s == 4) cm->search_opts |= CM_SEARCH_INSIDE;
    else                            cm->search_opts &= ~CM_SEARCH_INSIDE;
    switch (s) {
    case 0: name = "FastCYKScan";
      st[s] = FastCYKScan(cm, errbuf, cm->smx, SMX_QDB1_TIGHT, dsq, 1, L, 0.0f, th, 0, 2.0f, &envi, &envj, &vsc, &sc); break;
    case 1: name = "FastFInsideScan";
      st[s] = FastFInsideScan(cm, errbuf, cm->smx, SMX_QDB2_LOOSE, dsq, 1, L, 0.0f, th, 0, 2.0f, &envi, &envj, &vsc, &sc); break;
    case 2: name = "FastIInsideScan";
      st[s] = FastIInsideScan(cm, errbuf, cm->smx, SMX_QDB2_LOOSE, dsq, 1, L, 0.0f, th, 0, 2.0f, &envi, &envj, &vsc, &sc); break;
    case 3: name = "RefTrCYKScan";
      st[s] = RefTrCYKScan(cm, errbuf, cm->trsmx, SMX_QDB1_TIGHT, PLI_PASS_5P_AND_3P_FORCE, dsq, 1, L, 0.0f, th, 0, 2.0f,
                           &envi, &envj, &vsc, &mode, &sc); break;
    case 4: name = "RefITrInsideScan";
      st[s] = RefITrInsideScan(cm, errbuf, cm->trsmx, SMX_QDB2_LOOSE, PLI_PASS_5P_ONLY_FORCE, dsq, 1, L, 0.0f, th, 0, 2.0f,
                               &envi, &envj, &vsc, &mode, &sc); break;
    }
    report(name, st[s], errbuf, sc, mode, envi, envj, th);
    free(vsc);                 /* the caller owns *ret_vsc */
    cm_tophits_Destroy(th);
    nok += st[s] == eslOK;
    nodev += st[s] == eslENODEVICE;
  }
  free(dsq); free(seq);
  FreeCM(cm);
  return (nok == 5 || nodev == 5) ? 0 : 1;
}

static int run_teardown(const char *cmpath)
{
  CM_t *cm = load(cmpath, 1);
  if (!cm->trsmx) { fprintf(stderr, "no truncated scan matrix\n"); return 2; }
  int status = infx_debug_attach_trdev(cm, cm->trsmx, 3);
  if (status != eslOK) { fprintf(stderr, "attach failed: %d\n", status); return 2; }
  FreeCM(cm);
  printf("teardown ok\n");
  return 0;
}

int main(int argc, char **argv)
{
  if (argc == 4 && strcmp(argv[1], "scan") == 0) return run_scans(argv[2], argv[3]);
  if (argc == 3 && strcmp(argv[1], "teardown") == 0) return run_teardown(argv[2]);
  fprintf(stderr, "usage: abi_check scan <model.cm> <seq.txt> | abi_check teardown <model.cm>\n");
  return 2;
}

/* cm_model.cpp -- covariance model input and configuration (host side).
 *
 * Produces the score tables the scanners read: tsc/itsc, oesc/ioesc,
 * lmesc/rmesc, beginsc/endsc, el_selfsc, QDB bands.  Each function cites the
 * reference routine whose results it reproduces.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include <string>
#include <vector>
#include <mutex>

#include "cm_internal.h"

/* ------------------------------------------------------------------------ */
/* errors                                                                    */

extern "C" int infx_fail(int status, char *errbuf, const char *fmt, ...)
{
  if (errbuf) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(errbuf, eslERRBUFSIZE, fmt, ap);
    va_end(ap);
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* RNA alphabet: "ACGU-RYMKSWHBVDN*~" (Easel's eslRNA layout).              */

const char infx_sym[INFX_KP + 1] = "ACGU-RYMKSWHBVDN*~";
/*                                       A C G U */
const char infx_degen[INFX_KP][INFX_K] = {
  {1,0,0,0}, {0,1,0,0}, {0,0,1,0}, {0,0,0,1},   /* A C G U */
  {0,0,0,0},                                    /* - gap */
  {1,0,1,0},                                    /* R = AG */
  {0,1,0,1},                                    /* Y = CU */
  {1,1,0,0},                                    /* M = AC */
  {0,0,1,1},                                    /* K = GU */
  {0,1,1,0},                                    /* S = CG */
  {1,0,0,1},                                    /* W = AU */
  {1,1,0,1},                                    /* H = ACU */
  {0,1,1,1},                                    /* B = CGU */
  {1,1,1,0},                                    /* V = ACG */
  {1,0,1,1},                                    /* D = AGU */
  {1,1,1,1},                                    /* N = ACGU */
  {0,0,0,0},                                    /* * */
  {0,0,0,0}                                     /* ~ */
};
const int infx_ndegen[INFX_KP] = {1,1,1,1, 0, 2,2,2,2,2,2, 3,3,3,3, 4, 0,0};

extern "C" int infx_abc_Digitize(const char *seq, int64_t L, ESL_DSQ *dsq)
{
  /* built once, thread-safely (a function-local static) */
  struct Map { int m[256]; };
  static const Map map = [] {
    Map t;
    for (int c = 0; c < 256; c++) t.m[c] = -1;
    for (int x = 0; x < INFX_KP; x++) {
      t.m[(unsigned char) infx_sym[x]] = x;
      t.m[tolower((unsigned char) infx_sym[x])] = x;
    }
    t.m['T'] = t.m['t'] = 3;  /* T read as U */
    t.m['X'] = t.m['x'] = 15; /* X read as N */
    t.m['I'] = t.m['i'] = 0;  /* inosine read as A */
    t.m['_'] = t.m['.'] = 4;  /* gap synonyms */
    return t;
  }();
  dsq[0] = 255;
  for (int64_t i = 0; i < L; i++) {
    int x = map.m[(unsigned char) seq[i]];
    if (x < 0) return eslEINVAL;
    dsq[i + 1] = (ESL_DSQ) x;
  }
  dsq[L + 1] = 255;
  return eslOK;
}

/* Easel esl_abc_FCount: a degenerate residue spreads its weight evenly over
 * its canonical residues. */
static void abc_FCount(float *ct, int x, float wt)
{
  if (x < INFX_K) { ct[x] += wt; return; }
  for (int y = 0; y < INFX_K; y++)
    if (infx_degen[x][y]) ct[y] += wt / (float) infx_ndegen[x];
}

extern "C" void infx_abc_DCount(double *ct, ESL_DSQ x, double wt)
{
  if (x < INFX_K) { ct[x] += wt; return; }
  if (x >= INFX_KP) return;
  for (int y = 0; y < INFX_K; y++)
    if (infx_degen[x][y]) ct[y] += wt / (double) infx_ndegen[x];
}

/* Easel esl_abc_FAvgScore: mean score over the residues a degenerate code
 * stands for. */
static float abc_FAvgScore(int x, const float *sc)
{
  float result = 0.;
  for (int i = 0; i < INFX_K; i++)
    if (infx_degen[x][i]) result += sc[i];
  result /= (float) infx_ndegen[x];
  return result;
}

/* Easel esl_abc_FExpectScVec / IExpectScVec: expected score of degenerate
 * codes K+1..Kp-3 under background p. */
static void abc_FExpectScVec(float *sc, const float *p)
{
  for (int x = INFX_K + 1; x <= INFX_KP - 3; x++) {
    float result = 0., denom = 0.;
    for (int i = 0; i < INFX_K; i++)
      if (infx_degen[x][i]) { result += sc[i] * p[i]; denom += p[i]; }
    sc[x] = result / denom;
  }
}
static void abc_IExpectScVec(int *sc, const float *p)
{
  for (int x = INFX_K + 1; x <= INFX_KP - 3; x++) {
    float result = 0., denom = 0.;
    for (int i = 0; i < INFX_K; i++)
      if (infx_degen[x][i]) { result += (float) sc[i] * p[i]; denom += p[i]; }
    result /= denom;
    sc[x] = (result < 0) ? (int) (result - 0.5) : (int) (result + 0.5);
  }
}

/* ------------------------------------------------------------------------ */
/* Easel vector helpers.  esl_vec_FSum uses Kahan compensated summation.      */

extern "C" float infx_vec_FSum(const float *vec, int n)
{
  float sum = 0., c = 0., y, t;
  for (int i = 0; i < n; i++) {
    y   = vec[i] - c;
    t   = sum + y;
    c   = (t - sum) - y;
    sum = t;
  }
  return sum;
}

extern "C" void infx_vec_FNorm(float *vec, int n)
{
  float sum = infx_vec_FSum(vec, n);
  if (sum != 0.0) for (int x = 0; x < n; x++) vec[x] /= sum;
  else            for (int x = 0; x < n; x++) vec[x] = 1. / (float) n;
}

/* ------------------------------------------------------------------------ */
/* score conversions (src/infernal.h:153-154, src/cm.c:4185-4235)             */

static inline double sreLOG2(double x) { return (x > 0) ? log(x) * 1.44269504 : IMPOSSIBLE; }
static inline double sreEXP2(double x) { return exp(x * 0.69314718); }

static int Prob2Score(float p, float null)
{
  if (p == 0.0) return -INFTY;
  return (int) floor(0.5 + INTSCALE * sreLOG2(p / null));
}

extern "C" float Scorify(int sc)
{
  if (sc == -INFTY) return IMPOSSIBLE;
  return ((float) sc / INTSCALE);
}

/* ------------------------------------------------------------------------ */
/* state helpers (src/cm.c:1177-1270)                                         */

extern "C" int StateDelta(int t)
{
  switch (t) { case MP_st: return 2; case ML_st: case MR_st: case IL_st: case IR_st: return 1; default: return 0; }
}
extern "C" int StateLeftDelta(int t)
{
  switch (t) { case MP_st: case ML_st: case IL_st: return 1; default: return 0; }
}
extern "C" int StateRightDelta(int t)
{
  switch (t) { case MP_st: case MR_st: case IR_st: return 1; default: return 0; }
}
extern "C" int Emitmode(int t)
{
  switch (t) {
  case IL_st: case ML_st: return EMITLEFT;
  case IR_st: case MR_st: return EMITRIGHT;
  case MP_st: return EMITPAIR;
  default: return EMITNONE;
  }
}
/* src/cm.c:1775: a detached insert is the state right before an END_E. */
extern "C" int StateIsDetached(const CM_t *cm, int v)
{
  return (v + 1 < cm->M && cm->stid[v + 1] == END_E) ? 1 : 0;
}

/* src/cm.c:1506 DeriveUniqueStateCode */
static int unique_state_code(int nd, int st)
{
  switch (nd) {
  case BIF_nd:  return st == B_st ? BIF_B : -1;
  case MATP_nd:
    switch (st) { case D_st: return MATP_D; case MP_st: return MATP_MP; case ML_st: return MATP_ML;
                  case MR_st: return MATP_MR; case IL_st: return MATP_IL; case IR_st: return MATP_IR; }
    return -1;
  case MATL_nd: return st == D_st ? MATL_D : st == ML_st ? MATL_ML : st == IL_st ? MATL_IL : -1;
  case MATR_nd: return st == D_st ? MATR_D : st == MR_st ? MATR_MR : st == IR_st ? MATR_IR : -1;
  case BEGL_nd: return st == S_st ? BEGL_S : -1;
  case BEGR_nd: return st == S_st ? BEGR_S : st == IL_st ? BEGR_IL : -1;
  case ROOT_nd: return st == S_st ? ROOT_S : st == IL_st ? ROOT_IL : st == IR_st ? ROOT_IR : -1;
  case END_nd:  return st == E_st ? END_E : -1;
  }
  return -1;
}

static int state_code(const char *s)
{
  static const char *names[] = {"D", "MP", "ML", "MR", "IL", "IR", "S", "E", "B", "EL"};
  for (int i = 0; i < 10; i++) if (strcmp(s, names[i]) == 0) return i;
  return -1;
}
static int node_code(const char *s)
{
  static const char *names[] = {"BIF", "MATP", "MATL", "MATR", "BEGL", "BEGR", "ROOT", "END"};
  for (int i = 0; i < 8; i++) if (strcmp(s, names[i]) == 0) return i;
  return -1;
}

/* ------------------------------------------------------------------------ */
/* allocation                                                                 */

template <typename T> static T **alloc_2d(int rows, int cols)
{
  T **a = (T **) malloc(sizeof(T *) * (rows > 0 ? rows : 1));
  T  *b = (T *)  calloc((size_t) (rows > 0 ? rows : 1) * (cols > 0 ? cols : 1), sizeof(T));
  for (int i = 0; i < rows; i++) a[i] = b + (size_t) i * cols;
  if (rows == 0) a[0] = b;
  return a;
}
template <typename T> static void free_2d(T **a) { if (a) { free(a[0]); free(a); } }

static CM_t *create_cm_body(int M, int nodes)
{
  CM_t *cm = (CM_t *) calloc(1, sizeof(CM_t));
  cm->M = M; cm->nodes = nodes;
  cm->null    = (float *) calloc(INFX_K, sizeof(float));
  cm->sttype  = (char *) calloc(M + 1, 1);
  cm->stid    = (char *) calloc(M + 1, 1);
  cm->ndidx   = (int *) calloc(M, sizeof(int));
  cm->cfirst  = (int *) calloc(M, sizeof(int));
  cm->cnum    = (int *) calloc(M, sizeof(int));
  cm->plast   = (int *) calloc(M, sizeof(int));
  cm->pnum    = (int *) calloc(M, sizeof(int));
  cm->nodemap = (int *) calloc(nodes, sizeof(int));
  cm->ndtype  = (char *) calloc(nodes + 1, 1);
  cm->t       = alloc_2d<float>(M, MAXCONNECT);
  cm->e       = alloc_2d<float>(M, INFX_K * INFX_K);
  cm->begin   = (float *) calloc(M, sizeof(float));
  cm->end     = (float *) calloc(M, sizeof(float));
  cm->tsc     = alloc_2d<float>(M, MAXCONNECT);
  cm->esc     = alloc_2d<float>(M, INFX_K * INFX_K);
  cm->lmesc   = alloc_2d<float>(M, INFX_KP);
  cm->rmesc   = alloc_2d<float>(M, INFX_KP);
  cm->beginsc = (float *) calloc(M, sizeof(float));
  cm->endsc   = (float *) calloc(M, sizeof(float));
  cm->itsc    = alloc_2d<int>(M, MAXCONNECT);
  cm->iesc    = alloc_2d<int>(M, INFX_K * INFX_K);
  cm->ilmesc  = alloc_2d<int>(M, INFX_KP);
  cm->irmesc  = alloc_2d<int>(M, INFX_KP);
  cm->ibeginsc = (int *) calloc(M, sizeof(int));
  cm->iendsc   = (int *) calloc(M, sizeof(int));
  cm->qdbinfo  = (CM_QDBINFO *) calloc(1, sizeof(CM_QDBINFO));
  cm->qdbinfo->dmin1 = (int *) calloc(M, sizeof(int));
  cm->qdbinfo->dmax1 = (int *) calloc(M, sizeof(int));
  cm->qdbinfo->dmin2 = (int *) calloc(M, sizeof(int));
  cm->qdbinfo->dmax2 = (int *) calloc(M, sizeof(int));
  cm->pbegin = 0.05f; cm->pend = 0.05f;
  cm->null3_omega = (float) (1. / 65536.);
  cm->beta_W = 1e-7;
  return cm;
}

static void free_emit_scores(CM_t *cm)
{
  if (cm->oesc)  { free(cm->oesc[0] ? cm->oesc[0] : NULL); }
  free(cm->oesc); cm->oesc = NULL;
  if (cm->ioesc) { free(cm->ioesc[0] ? cm->ioesc[0] : NULL); }
  free(cm->ioesc); cm->ioesc = NULL;
}

static void free_trp(CM_TR_PENALTIES *trp)
{
  if (!trp) return;
  free_2d(trp->g_ptyAA); free_2d(trp->l_ptyAA); free_2d(trp->ig_ptyAA); free_2d(trp->il_ptyAA);
  free(trp);
}

extern "C" void FreeCM(CM_t *cm)
{
  if (!cm) return;
  if (cm->smx)   cm_scan_mx_Destroy(cm, cm->smx);
  if (cm->trsmx) cm_tr_scan_mx_Destroy(cm, cm->trsmx);
  free(cm->name); free(cm->null); free(cm->sttype); free(cm->stid); free(cm->ndidx);
  free(cm->cfirst); free(cm->cnum); free(cm->plast); free(cm->pnum); free(cm->nodemap); free(cm->ndtype);
  free_2d(cm->t); free_2d(cm->e); free(cm->begin); free(cm->end);
  free_2d(cm->tsc); free_2d(cm->esc); free_2d(cm->lmesc); free_2d(cm->rmesc);
  free(cm->beginsc); free(cm->endsc);
  free_2d(cm->itsc); free_2d(cm->iesc); free_2d(cm->ilmesc); free_2d(cm->irmesc);
  free(cm->ibeginsc); free(cm->iendsc);
  free_emit_scores(cm);
  free(cm->root_trans);
  if (cm->qdbinfo) {
    free(cm->qdbinfo->dmin1); free(cm->qdbinfo->dmax1); free(cm->qdbinfo->dmin2); free(cm->qdbinfo->dmax2);
    free(cm->qdbinfo);
  }
  if (cm->emap) { free(cm->emap->lpos); free(cm->emap->rpos); free(cm->emap->epos); free(cm->emap); }
  free_trp(cm->trp);
  free(cm);
}

/* ------------------------------------------------------------------------ */
/* emit map (src/display.c:1110 CreateEmitMap): consensus coordinates of the
 * left/right emission of each node, by a preorder walk of the node tree.    */

static int create_emit_map(CM_t *cm)
{
  CMEmitMap_t *map = (CMEmitMap_t *) calloc(1, sizeof(CMEmitMap_t));
  map->lpos = (int *) malloc(sizeof(int) * cm->nodes);
  map->rpos = (int *) malloc(sizeof(int) * cm->nodes);
  map->epos = (int *) malloc(sizeof(int) * cm->nodes);
  for (int nd = 0; nd < cm->nodes; nd++) map->lpos[nd] = map->rpos[nd] = map->epos[nd] = -1;

  /* explicit stack of (node, side) */
  std::vector<std::pair<int,int>> stack;
  int cpos = 0;
  stack.push_back({0, 0});
  while (!stack.empty()) {
    int nd = stack.back().first, on_right = stack.back().second;
    stack.pop_back();
    if (nd < 0 || nd >= cm->nodes) { cm->emap = map; return eslEINVAL; }
    if (on_right) {
      map->rpos[nd] = cpos + 1;
      if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATR_nd) cpos++;
    } else {
      if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd) cpos++;
      map->lpos[nd] = cpos;
      if (cm->ndtype[nd] == BIF_nd) {
        int v = cm->nodemap[nd];
        stack.push_back({nd, 1});
        stack.push_back({cm->ndidx[cm->cnum[v]], 0});    /* right child visited after the left */
        stack.push_back({cm->ndidx[cm->cfirst[v]], 0});
      } else {
        stack.push_back({nd, 1});
        if (cm->ndtype[nd] != END_nd) stack.push_back({nd + 1, 0});
      }
    }
  }
  for (int nd = cm->nodes - 1; nd >= 0; nd--) {
    if (cm->ndtype[nd] == END_nd) cpos = map->lpos[nd];
    else if (cm->ndtype[nd] == BIF_nd) cpos = map->epos[cm->ndidx[cm->cnum[cm->nodemap[nd]]]];
    map->epos[nd] = cpos;
  }
  map->clen = map->rpos[0] - 1;
  cm->emap = map;
  for (int nd = 0; nd < cm->nodes; nd++)
    if (map->lpos[nd] == -1 || map->rpos[nd] == -1 || map->epos[nd] == -1) return eslEINVAL;
  return eslOK;
}

/* ------------------------------------------------------------------------ */
/* INFERNAL1/a reader (src/cm_file.c:1539-2118)                               */

/* src/cm_file.c:3353 ascii2prob */
static float ascii2prob(const char *s, float null)
{
  return (*s == '*') ? 0. : exp(atof(s) / 1.44269504) * null;
}

static void tokenize(const std::string &line, std::vector<std::string> &tok)
{
  tok.clear();
  size_t i = 0, n = line.size();
  while (i < n) {
    while (i < n && isspace((unsigned char) line[i])) i++;
    if (i >= n) break;
    size_t j = i;
    while (j < n && !isspace((unsigned char) line[j])) j++;
    tok.push_back(line.substr(i, j - i));
    i = j;
  }
}

extern "C" int infx_cm_ReadText(const char *text, CM_t **ret_cm, char *errbuf)
{
  *ret_cm = NULL;
  std::vector<std::string> lines;
  {
    const char *p = text;
    while (*p) {
      const char *q = strchr(p, '\n');
      if (!q) q = p + strlen(p);
      lines.emplace_back(p, q - p);
      p = (*q) ? q + 1 : q;
    }
  }
  size_t li = 0;
  std::vector<std::string> tok;
  while (li < lines.size()) { tokenize(lines[li], tok); if (!tok.empty()) break; li++; }
  if (li >= lines.size()) return infx_fail(eslEOF, errbuf, "empty CM file");
  if (tok[0] != "INFERNAL1/a") return infx_fail(eslEFORMAT, errbuf, "Didn't find INFERNAL1/a tag: bad format or not an INFERNAL save file?");
  li++;

  int M = 0, nodes = 0, clen = 0, W = 0, have_el = 0, have_null = 0;
  std::string name = "";
  float el_selfsc = 0, pbegin = 0.05f, pend = 0.05f, n3omega = (float) (1. / 65536.);
  double beta_W = 1e-7, beta1 = 1e-7, beta2 = 1e-15;
  float nullv[INFX_K];
  for (; li < lines.size(); li++) {
    tokenize(lines[li], tok);
    if (tok.empty()) continue;
    const std::string &tag = tok[0];
    if (tag == "CM") { li++; break; }
    if (tok.size() < 2) continue;
    if      (tag == "NAME")     name = tok[1];
    else if (tag == "STATES")   M = atoi(tok[1].c_str());
    else if (tag == "NODES")    nodes = atoi(tok[1].c_str());
    else if (tag == "CLEN")     clen = atoi(tok[1].c_str());
    else if (tag == "W")        W = atoi(tok[1].c_str());
    else if (tag == "ALPH")     { if (tok[1] != "RNA" && tok[1] != "DNA") return infx_fail(eslEFORMAT, errbuf, "only RNA/DNA models are supported, not %s", tok[1].c_str()); }
    else if (tag == "PBEGIN")   pbegin = atof(tok[1].c_str());
    else if (tag == "PEND")     pend = atof(tok[1].c_str());
    else if (tag == "WBETA")    beta_W = atof(tok[1].c_str());
    else if (tag == "QDBBETA1") beta1 = atof(tok[1].c_str());
    else if (tag == "QDBBETA2") beta2 = atof(tok[1].c_str());
    else if (tag == "N3OMEGA")  n3omega = atof(tok[1].c_str());
    else if (tag == "ELSELF")   { el_selfsc = atof(tok[1].c_str()); have_el = 1; }
    else if (tag == "NULL") {
      if (tok.size() < 1 + INFX_K) return infx_fail(eslEFORMAT, errbuf, "Too few fields on NULL line");
      for (int x = 0; x < INFX_K; x++) nullv[x] = ascii2prob(tok[1 + x].c_str(), (1. / (float) INFX_K));
      have_null = 1;
    }
  }
  if (M < 1)     return infx_fail(eslEFORMAT, errbuf, "Failed to read STATES line in header section");
  if (nodes < 1) return infx_fail(eslEFORMAT, errbuf, "Failed to read NODES line in header section");
  if (clen < 1)  return infx_fail(eslEFORMAT, errbuf, "Failed to read CLEN line in header section");
  if (W < 1)     return infx_fail(eslEFORMAT, errbuf, "Failed to read W line in header section");
  if (!have_el)  return infx_fail(eslEFORMAT, errbuf, "Failed to read ELSELF line in header section");
  if (!have_null) return infx_fail(eslEFORMAT, errbuf, "Failed to read NULL line in header section");

  CM_t *cm = create_cm_body(M, nodes);
  cm->name = strdup(name.c_str());
  cm->W = W; cm->clen = clen; cm->el_selfsc = el_selfsc;
  cm->pbegin = pbegin; cm->pend = pend; cm->beta_W = beta_W; cm->null3_omega = n3omega;
  cm->qdbinfo->beta1 = beta1; cm->qdbinfo->beta2 = beta2;
  for (int x = 0; x < INFX_K; x++) cm->null[x] = nullv[x];

  int nd = -1, v = 0, status = eslOK;
  int counted_clen = 0;
  for (; li < lines.size() && v < M; li++) {
    tokenize(lines[li], tok);
    if (tok.empty()) continue;
    if (tok[0] == "[") {
      if (tok.size() < 3) { status = infx_fail(eslEFORMAT, errbuf, "Too few fields on node line"); goto ERROR; }
      int x = node_code(tok[1].c_str());
      if (x < 0) { status = infx_fail(eslEFORMAT, errbuf, "Invalid node type %s", tok[1].c_str()); goto ERROR; }
      nd = atoi(tok[2].c_str());
      if (nd < 0 || nd >= nodes) { status = infx_fail(eslEFORMAT, errbuf, "Invalid node index %s", tok[2].c_str()); goto ERROR; }
      cm->ndtype[nd] = (char) x;
      if (x == MATP_nd) counted_clen += 2; else if (x == MATL_nd || x == MATR_nd) counted_clen++;
      cm->nodemap[nd] = v;
      continue;
    }
    /* state line: <type> <v> <plast> <pnum> <cfirst> <cnum> <dmin2> <dmin1> <dmax1> <dmax2> <t...> <e...> */
    if (tok.size() < 10) { status = infx_fail(eslEFORMAT, errbuf, "Too few fields on state %d line", v); goto ERROR; }
    if (nd < 0) { status = infx_fail(eslEFORMAT, errbuf, "state line before any node line"); goto ERROR; }
    {
      int st = state_code(tok[0].c_str());
      if (st < 0) { status = infx_fail(eslEFORMAT, errbuf, "Invalid state type %s", tok[0].c_str()); goto ERROR; }
      if (atoi(tok[1].c_str()) != v) { status = infx_fail(eslEFORMAT, errbuf, "Invalid state index on state line: should be %d, not %s", v, tok[1].c_str()); goto ERROR; }
      cm->sttype[v] = (char) st;
      cm->plast[v]  = atoi(tok[2].c_str());
      cm->pnum[v]   = atoi(tok[3].c_str());
      cm->cfirst[v] = atoi(tok[4].c_str());
      cm->cnum[v]   = atoi(tok[5].c_str());
      cm->qdbinfo->dmin2[v] = atoi(tok[6].c_str());
      cm->qdbinfo->dmin1[v] = atoi(tok[7].c_str());
      cm->qdbinfo->dmax1[v] = atoi(tok[8].c_str());
      cm->qdbinfo->dmax2[v] = atoi(tok[9].c_str());
      size_t k = 10;
      if (st != B_st) {
        if (cm->cnum[v] > MAXCONNECT || cm->cnum[v] < 0) { status = infx_fail(eslEFORMAT, errbuf, "bad cnum %d for state %d", cm->cnum[v], v); goto ERROR; }
        for (int x = 0; x < cm->cnum[v]; x++) {
          if (k >= tok.size()) { status = infx_fail(eslEFORMAT, errbuf, "Too few fields on state %d line", v); goto ERROR; }
          cm->t[v][x] = ascii2prob(tok[k++].c_str(), 1.);
        }
      }
      if (st == ML_st || st == MR_st || st == IL_st || st == IR_st) {
        for (int x = 0; x < INFX_K; x++) {
          if (k >= tok.size()) { status = infx_fail(eslEFORMAT, errbuf, "Too few fields on state %d line", v); goto ERROR; }
          cm->e[v][x] = ascii2prob(tok[k++].c_str(), cm->null[x]);
        }
      } else if (st == MP_st) {
        for (int x = 0; x < INFX_K; x++)
          for (int y = 0; y < INFX_K; y++) {
            if (k >= tok.size()) { status = infx_fail(eslEFORMAT, errbuf, "Too few fields on state %d line", v); goto ERROR; }
            cm->e[v][x * INFX_K + y] = ascii2prob(tok[k++].c_str(), cm->null[x] * cm->null[y]);
          }
      }
      cm->ndidx[v] = nd;
      cm->stid[v]  = (char) unique_state_code(cm->ndtype[nd], st);
      if (cm->stid[v] < 0) { status = infx_fail(eslEFORMAT, errbuf, "state %d: type %s not allowed in node type %d", v, tok[0].c_str(), cm->ndtype[nd]); goto ERROR; }
    }
    v++;
  }
  if (v < M) { status = infx_fail(eslEFORMAT, errbuf, "Premature end of data in main model: no state %d line", v); goto ERROR; }
  for (; li < lines.size(); li++) { tokenize(lines[li], tok); if (!tok.empty()) break; }
  if (li >= lines.size() || tok[0] != "//") { status = infx_fail(eslEFORMAT, errbuf, "Expected closing //"); goto ERROR; }
  cm->clen = counted_clen;

  /* src/cm.c:344 CMRenormalize */
  infx_vec_FNorm(cm->null, INFX_K);
  for (v = 0; v < M; v++) {
    if (cm->cnum[v] > 0 && cm->sttype[v] != B_st) infx_vec_FNorm(cm->t[v], cm->cnum[v]);
    if (cm->sttype[v] == ML_st || cm->sttype[v] == MR_st || cm->sttype[v] == IL_st || cm->sttype[v] == IR_st)
      infx_vec_FNorm(cm->e[v], INFX_K);
    if (cm->sttype[v] == MP_st) infx_vec_FNorm(cm->e[v], INFX_K * INFX_K);
  }
  /* structural sanity: children inside the model, child counts the scanners handle */
  for (v = 0; v < M; v++) {
    int st = cm->sttype[v];
    if (st == E_st) continue;
    if (st == B_st) {
      if (cm->cfirst[v] <= v || cm->cfirst[v] >= M || cm->cnum[v] <= v || cm->cnum[v] >= M ||
          cm->stid[cm->cfirst[v]] != BEGL_S || cm->stid[cm->cnum[v]] != BEGR_S) {
        status = infx_fail(eslEFORMAT, errbuf, "B state %d has bad children", v); goto ERROR;
      }
      continue;
    }
    if (cm->cfirst[v] < v || cm->cfirst[v] + cm->cnum[v] > M) { status = infx_fail(eslEFORMAT, errbuf, "state %d has children outside the model", v); goto ERROR; }
    if (cm->cfirst[v] == v && st != IL_st && st != IR_st) { status = infx_fail(eslEFORMAT, errbuf, "state %d: only inserts may loop", v); goto ERROR; }
    if (v > 0 && cm->stid[v] != BEGL_S && (cm->cnum[v] < 2 || cm->cnum[v] > 6)) {
      status = infx_fail(eslEFORMAT, errbuf, "state %d has %d children; the scanners handle 2..6", v, cm->cnum[v]); goto ERROR;
    }
    for (int y = cm->cfirst[v]; y < cm->cfirst[v] + cm->cnum[v]; y++)
      if (cm->stid[y] == BEGL_S) { status = infx_fail(eslEFORMAT, errbuf, "state %d has a BEGL_S child", v); goto ERROR; }
  }
  if ((status = create_emit_map(cm)) != eslOK) { infx_fail(status, errbuf, "After reading complete model, failed to create an emit map"); goto ERROR; }
  *ret_cm = cm;
  return eslOK;

ERROR:
  FreeCM(cm);
  return status;
}

extern "C" int infx_cm_Read(const char *path, CM_t **ret_cm, char *errbuf)
{
  *ret_cm = NULL;
  FILE *fp = fopen(path, "rb");
  if (!fp) return infx_fail(eslENOTFOUND, errbuf, "CM file %s not found", path);
  std::string text;
  char buf[65536];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, n);
  fclose(fp);
  /* keep only the first model: stop at the first line that is exactly "//" */
  size_t pos = 0;
  while (pos < text.size()) {
    size_t e = text.find('\n', pos);
    if (e == std::string::npos) e = text.size();
    size_t a = pos; while (a < e && isspace((unsigned char) text[a])) a++;
    size_t b = e;   while (b > a && isspace((unsigned char) text[b - 1])) b--;
    if (b - a == 2 && text[a] == '/' && text[a + 1] == '/') { text.resize(e); break; }
    pos = e + 1;
  }
  return infx_cm_ReadText(text.c_str(), ret_cm, errbuf);
}

/* ------------------------------------------------------------------------ */
/* localization (src/cm_modelconfig.c:404-548)                                */

static void calculate_local_begin_probs(CM_t *cm, float p_internal_start, float *begin)
{
  int nstarts = 0;
  for (int nd = 2; nd < cm->nodes; nd++)
    if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd || cm->ndtype[nd] == MATR_nd || cm->ndtype[nd] == BIF_nd)
      nstarts++;
  for (int v = 0; v < cm->M; v++) begin[v] = 0.;
  begin[cm->nodemap[1]] = 1. - p_internal_start;
  float p = p_internal_start / (float) nstarts;
  for (int nd = 2; nd < cm->nodes; nd++)
    if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd || cm->ndtype[nd] == MATR_nd || cm->ndtype[nd] == BIF_nd)
      begin[cm->nodemap[nd]] = p;
}

static int is_exit_node(const CM_t *cm, int nd)
{
  return (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd || cm->ndtype[nd] == MATR_nd ||
          cm->ndtype[nd] == BEGL_nd || cm->ndtype[nd] == BEGR_nd) &&
         (nd + 1 < cm->nodes) && cm->ndtype[nd + 1] != END_nd;
}

static void cm_localize(CM_t *cm, float p_internal_start, float p_internal_exit)
{
  calculate_local_begin_probs(cm, p_internal_start, cm->begin);
  if (cm->root_trans == NULL) {
    cm->root_trans = (float *) malloc(sizeof(float) * (cm->cnum[0] > 0 ? cm->cnum[0] : 1));
    for (int x = 0; x < cm->cnum[0]; x++) cm->root_trans[x] = cm->t[0][x];
  }
  for (int x = 0; x < cm->cnum[0]; x++) cm->t[0][x] = 0.;
  cm->flags |= CMH_LOCAL_BEGIN;

  int nexits = 0;
  for (int nd = 1; nd < cm->nodes; nd++) if (is_exit_node(cm, nd)) nexits++;
  for (int v = 0; v < cm->M; v++) cm->end[v] = 0.;
  for (int nd = 1; nd < cm->nodes; nd++) {
    if (!is_exit_node(cm, nd)) continue;
    int v = cm->nodemap[nd];
    cm->end[v] = p_internal_exit / (float) nexits;
    float denom = infx_vec_FSum(cm->t[v], cm->cnum[v]);
    denom += cm->end[v];
    float scale = 1. / denom;   /* esl_vec_FScale takes a float */
    for (int x = 0; x < cm->cnum[v]; x++) cm->t[v][x] *= scale;
  }
  cm->flags |= CMH_LOCAL_END;
  cm->flags &= ~CMH_BITS;
}

/* ------------------------------------------------------------------------ */
/* log odds scores (src/cm.c:784-900, :3604-3950)                             */

static float **calc_optimized_emit_scores(CM_t *cm)
{
  const int K = INFX_K, Kp = INFX_KP;
  int nsing = 0, npair = 0;
  for (int v = 0; v < cm->M; v++) {
    int t = cm->sttype[v];
    if (t == IL_st || t == ML_st || t == IR_st || t == MR_st) nsing++;
    else if (t == MP_st) npair++;
  }
  float left[INFX_KP][INFX_K], right[INFX_KP][INFX_K];
  for (int a = K + 1; a < Kp - 1; a++) {
    for (int x = 0; x < K; x++) left[a][x] = right[a][x] = 0.;
    abc_FCount(left[a], a, 1.);
    abc_FCount(right[a], a, 1.);
  }
  float **esc = (float **) calloc(cm->M, sizeof(float *));
  float *block = (float *) malloc(sizeof(float) * ((size_t) Kp * nsing + (size_t) Kp * Kp * npair + 1));
  size_t cur = 0;
  for (int v = 0; v < cm->M; v++) {
    int t = cm->sttype[v];
    if (t == IL_st || t == ML_st || t == IR_st || t == MR_st) {
      float *e = esc[v] = block + cur; cur += Kp;
      for (int a = 0; a < K; a++) e[a] = cm->esc[v][a];
      e[K] = IMPOSSIBLE;
      for (int a = K + 1; a < Kp - 1; a++) e[a] = (infx_ndegen[a] > 0) ? abc_FAvgScore(a, cm->esc[v]) : (float) IMPOSSIBLE;
      e[Kp - 1] = IMPOSSIBLE;
    } else if (t == MP_st) {
      float *e = esc[v] = block + cur; cur += (size_t) Kp * Kp;
      for (int i = 0; i < Kp * Kp; i++) e[i] = IMPOSSIBLE;
      for (int a = 0; a < K; a++) for (int b = 0; b < K; b++) e[a * Kp + b] = cm->esc[v][a * K + b];
      for (int a = K + 1; a < Kp - 1; a++) for (int b = 0; b < K; b++) {
        float sc = 0.;
        for (int l = 0; l < K; l++) sc += cm->esc[v][l * K + b] * left[a][l];
        e[a * Kp + b] = sc;
      }
      for (int a = 0; a < K; a++) for (int b = K + 1; b < Kp - 1; b++) {
        float sc = 0.;
        for (int r = 0; r < K; r++) sc += cm->esc[v][a * K + r] * right[b][r];
        e[a * Kp + b] = sc;
      }
      for (int a = K + 1; a < Kp - 1; a++) for (int b = K + 1; b < Kp - 1; b++) {
        float sc = 0.;
        for (int l = 0; l < K; l++) for (int r = 0; r < K; r++) sc += cm->esc[v][l * K + r] * left[a][l] * right[b][r];
        e[a * Kp + b] = sc;
      }
    }
  }
  if (cur == 0) { free(block); block = NULL; }
  /* esc[] rows point into block; remember block at a stable place: row of first emitter */
  float **out = (float **) calloc(cm->M + 1, sizeof(float *));
  for (int v = 0; v < cm->M; v++) out[v + 1] = esc[v];
  out[0] = block;
  free(esc);
  return out;   /* out[0] = block owner, out[v+1] = row of state v */
}

static int **copy_int_emit_scores(CM_t *cm, float **oesc_rows /* indexed v */)
{
  const int K = INFX_K, Kp = INFX_KP;
  size_t total = 0;
  for (int v = 0; v < cm->M; v++) {
    int t = cm->sttype[v];
    if (t == IL_st || t == ML_st || t == IR_st || t == MR_st) total += Kp;
    else if (t == MP_st) total += (size_t) Kp * Kp;
  }
  int **out = (int **) calloc(cm->M + 1, sizeof(int *));
  int *block = (int *) malloc(sizeof(int) * (total + 1));
  out[0] = block;
  size_t cur = 0;
  for (int v = 0; v < cm->M; v++) {
    int t = cm->sttype[v];
    const float *o = oesc_rows[v];
    if (t == IL_st || t == ML_st || t == IR_st || t == MR_st) {
      int *e = out[v + 1] = block + cur; cur += Kp;
      for (int a = 0; a < K; a++) e[a] = cm->iesc[v][a];
      e[K] = -INFTY;
      for (int a = K + 1; a < Kp - 1; a++) e[a] = (int) floor(0.5 + INTSCALE * o[a]);
      e[Kp - 1] = -INFTY;
    } else if (t == MP_st) {
      int *e = out[v + 1] = block + cur; cur += (size_t) Kp * Kp;
      for (int i = 0; i < Kp * Kp; i++) e[i] = -INFTY;
      for (int a = 0; a < K; a++) for (int b = 0; b < K; b++) e[a * Kp + b] = cm->iesc[v][a * K + b];
      for (int a = K + 1; a < Kp - 1; a++) for (int b = 0; b < K; b++) e[a * Kp + b] = (int) floor(0.5 + INTSCALE * o[a * Kp + b]);
      for (int a = 0; a < K; a++) for (int b = K + 1; b < Kp - 1; b++) e[a * Kp + b] = (int) floor(0.5 + INTSCALE * o[a * Kp + b]);
      for (int a = K + 1; a < Kp - 1; a++) for (int b = K + 1; b < Kp - 1; b++) e[a * Kp + b] = (int) floor(0.5 + INTSCALE * o[a * Kp + b]);
    }
  }
  return out;
}

static int cm_logoddsify(CM_t *cm)
{
  const int K = INFX_K, Kp = INFX_KP;
  for (int v = 0; v < cm->M; v++) for (int x = 0; x < Kp; x++) { cm->lmesc[v][x] = 0.; cm->rmesc[v][x] = 0.; }
  for (int v = 0; v < cm->M; v++) {
    cm->lmesc[v][K] = cm->rmesc[v][K] = IMPOSSIBLE;
    cm->lmesc[v][Kp - 2] = cm->rmesc[v][Kp - 2] = IMPOSSIBLE;
    cm->lmesc[v][Kp - 1] = cm->rmesc[v][Kp - 1] = IMPOSSIBLE;
    cm->ilmesc[v][K] = cm->irmesc[v][K] = -INFTY;
    cm->ilmesc[v][Kp - 2] = cm->irmesc[v][Kp - 2] = -INFTY;
    cm->ilmesc[v][Kp - 1] = cm->irmesc[v][Kp - 1] = -INFTY;
    int st = cm->sttype[v];
    if (st != B_st && st != E_st)
      for (int x = 0; x < cm->cnum[v]; x++) {
        cm->tsc[v][x]  = sreLOG2(cm->t[v][x]);
        cm->itsc[v][x] = Prob2Score(cm->t[v][x], 1.0);
      }
    if (st == MP_st) {
      for (int x = 0; x < K; x++)
        for (int y = 0; y < K; y++) {
          cm->esc[v][x * K + y]  = sreLOG2(cm->e[v][x * K + y] / (cm->null[x] * cm->null[y]));
          cm->iesc[v][x * K + y] = Prob2Score(cm->e[v][x * K + y], (cm->null[x] * cm->null[y]));
          cm->lmesc[v][x] += cm->e[v][x * K + y];
          cm->rmesc[v][y] += cm->e[v][x * K + y];
        }
      for (int x = 0; x < K; x++) {
        cm->ilmesc[v][x] = Prob2Score(cm->lmesc[v][x], cm->null[x]);
        cm->irmesc[v][x] = Prob2Score(cm->rmesc[v][x], cm->null[x]);
        cm->lmesc[v][x]  = sreLOG2(cm->lmesc[v][x] / cm->null[x]);
        cm->rmesc[v][x]  = sreLOG2(cm->rmesc[v][x] / cm->null[x]);
      }
      abc_FExpectScVec(cm->lmesc[v], cm->null);
      abc_FExpectScVec(cm->rmesc[v], cm->null);
      abc_IExpectScVec(cm->ilmesc[v], cm->null);
      abc_IExpectScVec(cm->irmesc[v], cm->null);
    } else if (st == ML_st || st == MR_st || st == IL_st || st == IR_st) {
      for (int x = 0; x < K; x++) {
        cm->esc[v][x]  = sreLOG2(cm->e[v][x] / cm->null[x]);
        cm->iesc[v][x] = Prob2Score(cm->e[v][x], cm->null[x]);
      }
      if (st == ML_st || st == IL_st) {
        for (int x = 0; x < K; x++) { cm->lmesc[v][x] = cm->esc[v][x]; cm->ilmesc[v][x] = cm->iesc[v][x]; cm->rmesc[v][x] = 0.; cm->irmesc[v][x] = 0; }
        abc_FExpectScVec(cm->lmesc[v], cm->null);
        abc_IExpectScVec(cm->ilmesc[v], cm->null);
        for (int x = K + 1; x < Kp - 2; x++) { cm->rmesc[v][x] = 0.; cm->irmesc[v][x] = 0; }
      } else {
        for (int x = 0; x < K; x++) { cm->lmesc[v][x] = 0.; cm->ilmesc[v][x] = 0; cm->rmesc[v][x] = cm->esc[v][x]; cm->irmesc[v][x] = cm->iesc[v][x]; }
        for (int x = K + 1; x < Kp - 2; x++) { cm->lmesc[v][x] = 0.; cm->ilmesc[v][x] = 0; }
        abc_FExpectScVec(cm->rmesc[v], cm->null);
        abc_IExpectScVec(cm->irmesc[v], cm->null);
      }
    }
    cm->beginsc[v]  = sreLOG2(cm->begin[v]);
    cm->ibeginsc[v] = Prob2Score(cm->begin[v], 1.0);
    cm->endsc[v]    = sreLOG2(cm->end[v]);
    cm->iendsc[v]   = Prob2Score(cm->end[v], 1.0);
  }
  cm->iel_selfsc = Prob2Score(sreEXP2(cm->el_selfsc), 1.0);

  free_emit_scores(cm);
  float **frows = calc_optimized_emit_scores(cm);
  /* store with the reference's indexing: oesc[v] (NULL for non-emitters); the
   * allocation block is recovered from the first emitter row at free time */
  cm->oesc = (float **) calloc(cm->M, sizeof(float *));
  for (int v = 0; v < cm->M; v++) cm->oesc[v] = frows[v + 1];
  int **irows = copy_int_emit_scores(cm, cm->oesc);
  cm->ioesc = (int **) calloc(cm->M, sizeof(int *));
  for (int v = 0; v < cm->M; v++) cm->ioesc[v] = irows[v + 1];
  /* v = 0 (ROOT_S) never emits, so row 0 is free to own the block */
  cm->oesc[0]  = frows[0];
  cm->ioesc[0] = irows[0];
  free(frows); free(irows);
  cm->flags |= CMH_BITS;
  return eslOK;
}

/* ------------------------------------------------------------------------ */
/* init scores (src/cm.c:4100-4183)                                           */

extern "C" float **FCalcInitDPScores(const CM_t *cm)
{
  std::vector<float> el(cm->W + 1);
  for (int d = 0; d <= cm->W; d++) el[d] = cm->el_selfsc * d;
  float **a = alloc_2d<float>(cm->M, cm->W + 1);
  for (int v = 0; v < cm->M; v++) {
    if (NOT_IMPOSSIBLE(cm->endsc[v])) for (int d = 0; d <= cm->W; d++) a[v][d] = el[d] + cm->endsc[v];
    else                              for (int d = 0; d <= cm->W; d++) a[v][d] = IMPOSSIBLE;
  }
  return a;
}

extern "C" int **ICalcInitDPScores(const CM_t *cm)
{
  std::vector<int> el(cm->W + 1);
  for (int d = 0; d <= cm->W; d++) el[d] = cm->iel_selfsc * d;
  int **a = alloc_2d<int>(cm->M, cm->W + 1);
  for (int v = 0; v < cm->M; v++) {
    if (cm->iendsc[v] != -INFTY) for (int d = 0; d <= cm->W; d++) a[v][d] = el[d] + cm->iendsc[v];
    else                         for (int d = 0; d <= cm->W; d++) a[v][d] = -INFTY;
  }
  return a;
}

/* ------------------------------------------------------------------------ */
/* log-sum tables (src/logsum.c:72-169)                                       */

static int   ilogsum_lookup[LOGSUM_TBL];
static float flogsum_lookup[LOGSUM_TBL];
static std::once_flag ilogsum_once, flogsum_once;   /* filled once, also when threads race */

extern "C" void init_ilogsum(void)
{
  std::call_once(ilogsum_once, [] {
    for (int i = 0; i < LOGSUM_TBL; i++)
      ilogsum_lookup[i] = rint(INTSCALE * (sreLOG2(1. + sreEXP2((double) -i / INTSCALE))));
  });
}
extern "C" int ILogsum(int s1, int s2)
{
  const int max = ESL_MAX(-INFTY, ESL_MAX(s1, s2));
  const int min = ESL_MIN(s1, s2);
  return (min <= -INFTY || (max - min) >= LOGSUM_TBL) ? max : max + ilogsum_lookup[max - min];
}
extern "C" void FLogsumInit(void)
{
  std::call_once(flogsum_once, [] {
    for (int i = 0; i < LOGSUM_TBL; i++)
      flogsum_lookup[i] = sreLOG2(1. + sreEXP2((double) -i / INTSCALE));
  });
}
extern "C" float FLogsum(float s1, float s2)
{
  const float max = ESL_MAX(s1, s2);
  const float min = ESL_MIN(s1, s2);
  return (min == -INFINITY || (max - min) >= 23.f) ? max : max + flogsum_lookup[(int) ((max - min) * INTSCALE)];
}
extern "C" const float *infx_flogsum_table(void) { FLogsumInit(); return flogsum_lookup; }
extern "C" const int   *infx_ilogsum_table(void) { init_ilogsum(); return ilogsum_lookup; }

/* ------------------------------------------------------------------------ */
/* NULL3 (src/cm_parsetree.c:2352-2380)                                       */

extern "C" void ScoreCorrectionNull3(const float *null0, float *comp, int len, float omega, float *ret_sc)
{
  float score = 0.;
  for (int a = 0; a < INFX_K; a++) score += sreLOG2(comp[a] / null0[a]) * comp[a] * len;
  score += sreLOG2(omega);
  /* LogSum2(0., score) (src/logsum.c:151) */
  FLogsumInit();
  const float s1 = 0.f, s2 = score;
  const float max = ESL_MAX(s1, s2), min = ESL_MIN(s1, s2);
  score = (min == -INFINITY || (max - min) >= 23.f) ? max : max + flogsum_lookup[(int) ((max - min) * INTSCALE)];
  *ret_sc = score;
}

/* ------------------------------------------------------------------------ */
/* configuration (src/cm_modelconfig.c:139-373, subset)                       */

extern "C" int infx_tr_penalties_create(CM_t *cm, char *errbuf);   /* cm_trunc.cpp */

extern "C" int cm_Configure(CM_t *cm, char *errbuf, int W_from_cmdline)
{
  int status;
  if (cm == NULL) return infx_fail(eslEINVAL, errbuf, "cm_Configure: NULL cm");
  if (cm->flags & CM_IS_CONFIGURED) return infx_fail(eslEINCOMPAT, errbuf, "CM is already configured");
  if (W_from_cmdline != -1) {
    if (W_from_cmdline < 1) return infx_fail(eslEINVAL, errbuf, "bad W %d", W_from_cmdline);
    cm->W = W_from_cmdline;
  }
  FLogsumInit();
  init_ilogsum();
  if (cm->config_opts & CM_CONFIG_SCANMX) {
    if ((status = cm_scan_mx_Create(cm, errbuf, 1, 1, &cm->smx)) != eslOK) return status;
  }
  if (cm->config_opts & CM_CONFIG_TRSCANMX) {
    if ((status = cm_tr_scan_mx_Create(cm, errbuf, 1, 1, &cm->trsmx)) != eslOK) return status;
  }
  /* truncation penalties are computed from the global model, before localization */
  if (cm->config_opts & CM_CONFIG_TRUNC) {
    if ((status = infx_tr_penalties_create(cm, errbuf)) != eslOK) return status;
  }
  if (cm->config_opts & CM_CONFIG_LOCAL) cm_localize(cm, cm->pbegin, cm->pend);
  if ((cm->el_selfsc * cm->W) < IMPOSSIBLE) {
    cm->el_selfsc  = (IMPOSSIBLE / (cm->W + 1));
    cm->iel_selfsc = -INFTY;
  }
  if ((status = cm_logoddsify(cm)) != eslOK) return infx_fail(status, errbuf, "problem logoddsifying CM");
  cm->flags |= CM_IS_CONFIGURED;
  return eslOK;
}

/* ------------------------------------------------------------------------ */
/* work counts                                                               */

/* Cells per residue at steady state (jp >= W) and algorithmic ops per
 * residue: non-B states 2*cnum (+1 if emitting) per cell, B states
 * 2*(kmax-kmin+1) per cell, E states excluded.  The root state v=0 counts
 * its own children only. */
extern "C" int infx_CountSearchWork(CM_t *cm, int qdbidx, int truncated, double *ret_cells, double *ret_ops, double *ret_red)
{
  if (!cm || qdbidx < 0 || qdbidx > 2) return eslEINVAL;
  const int W = cm->W;
  const int *dmin = qdbidx == 1 ? cm->qdbinfo->dmin1 : qdbidx == 2 ? cm->qdbinfo->dmin2 : NULL;
  const int *dmax = qdbidx == 1 ? cm->qdbinfo->dmax1 : qdbidx == 2 ? cm->qdbinfo->dmax2 : NULL;
  double cells = 0, ops = 0, red = 0;
  for (int v = 0; v < cm->M; v++) {
    int st = cm->sttype[v];
    if (st == E_st) continue;
    int dn, dx;
    if (truncated) { dn = 1; dx = dmax ? ESL_MIN(dmax[v], W) : W; }
    else {
      dn = (st == MP_st) ? 2 : 1;
      if (dmin) dn = ESL_MAX(dmin[v], dn);
      dn = ESL_MIN(dn, W);
      dx = dmax ? ESL_MIN(dmax[v], W) : W;
    }
    if (dn > dx) continue;
    cells += dx - dn + 1;
    if (st == B_st) {
      int w = cm->cfirst[v], y = cm->cnum[v];
      for (int d = dn; d <= dx; d++) {
        int kmin, kmax;
        if (dmin && !truncated) {
          int dn_y = ESL_MIN(dmin[y], W), dx_y = ESL_MIN(dmax[y], W);
          int dn_w = ESL_MIN(dmin[w], W), dx_w = ESL_MIN(dmax[w], W);
          kmin = ESL_MAX(0, ESL_MAX(dn_y, d - dx_w));
          kmax = ESL_MIN(dx_y, d - dn_w);
        } else if (dmax && truncated) {
          int dx_y = ESL_MIN(dmax[y], W), dx_w = ESL_MIN(dmax[w], W);
          kmin = ESL_MAX(0, d - dx_w);
          kmax = ESL_MIN(dx_y, d);
        } else { kmin = 0; kmax = d; }
        if (kmax >= kmin) { ops += 2.0 * (kmax - kmin + 1); red += kmax - kmin + 1; }
      }
    } else {
      ops += (double) (dx - dn + 1) * (2 * cm->cnum[v] + (Emitmode(st) != EMITNONE ? 1 : 0));
      red += (double) (dx - dn + 1) * cm->cnum[v];
    }
  }
  if (ret_red) *ret_red = red;
  if (ret_cells) *ret_cells = cells;
  if (ret_ops)   *ret_ops = ops;
  return eslOK;
}

/* cm_CountSearchDPCalcs (src/cm_dpsearch.c:3099-3205): millions of DP
 * calculations per residue that a standard scan of L residues does, per
 * sub-model (ret_vcalcs[v], the caller frees it) and for the whole model
 * (ret_calcs = vcalcs[0]).  The reference's definition of a calculation:
 *   non-B state v >= 1, per j:  1 + (cnum+1) (dx-dn+1), plus (dx-dn+1) if v emits
 *   B state, per d with kmax >= kmin:  1 + (kmax-kmin+1)
 *   root, per j:  (cnum+1) (dx-dn+1), plus each finite local begin's dx-dn+1
 * with dn = max(dmin, 1 or 2 for MP) and dx = min(j, dmax, W) (no bands:
 * dmin/dmax NULL).  correct_for_first_W counts only j = W+1..L.  The sums are
 * float, each term added as (float) n / 1e6 in the reference's j, v, d order,
 * so the result rounds as the reference's does.  Each state then adds the
 * sub-model below it (B: both children; others: state v+1), and every entry
 * is divided by the residues counted. */
extern "C" int cm_CountSearchDPCalcs(CM_t *cm, char *errbuf, int L, int *dmin, int *dmax, int W, int correct_for_first_W,
                                     float **ret_vcalcs, float *ret_calcs)
{
  if (W > L && correct_for_first_W)
    return infx_fail(eslFAIL, errbuf, "gross misuse of cm_CountSearchDPCalcs(), W: %d > L: %d and correct_for_first_W is TRUE.\n", W, L);
  const bool banded = (dmin != NULL && dmax != NULL);
  if (W > L) W = L;
  const int M = cm->M;
  float *vc = (float *) calloc((size_t) M, sizeof(float));
  if (!vc) return infx_fail(eslEMEM, errbuf, "cm_CountSearchDPCalcs(): memory error.");
  auto add = [](float &acc, long n) { acc += ((float) n) / 1000000.; };   /* float += double, rounded to float */
  /* the d band of state v at end position j (cells d = dn..dx) */
  auto band = [&](int v, int j, bool mp_floor, int *dn, int *dx) {
    const int lo = (mp_floor && cm->sttype[v] == MP_st) ? 2 : 1;
    *dn = banded ? ESL_MAX(dmin[v], lo) : lo;
    *dx = ESL_MIN(j, W);
    if (banded) *dx = ESL_MIN(*dx, dmax[v]);
  };
  const int Leff = correct_for_first_W ? (L - W) : L;
  for (int j = correct_for_first_W ? W + 1 : 1; j <= L; j++) {
    int dn, dx;
    for (int v = M - 1; v > 0; v--) {
      const int st = cm->sttype[v];
      if (st == E_st) continue;
      band(v, j, true, &dn, &dx);
      if (st == B_st) {
        const int w = cm->cfirst[v], y = cm->cnum[v];
        for (int d = dn; d <= dx; d++) {
          const int kmin = banded ? ESL_MAX(ESL_MAX(dmin[y], d - dmax[w]), 0) : 0;
          const int kmax = banded ? ESL_MIN(dmax[y], d - dmin[w]) : d;
          if (kmax >= kmin) add(vc[v], 1 + (kmax - kmin + 1));
        }
      } else if (dx >= dn) {
        add(vc[v], 1 + (long) (cm->cnum[v] + 1) * (dx - dn + 1));
        if (StateDelta(st) > 0) add(vc[v], dx - dn + 1);
      }
    }
    band(0, j, false, &dn, &dx);
    if (dx >= dn) add(vc[0], (long) (cm->cnum[0] + 1) * (dx - dn + 1));
    if (cm->flags & CMH_LOCAL_BEGIN)
      for (int y = 1; y < M; y++) {
        /* no bands: dn = 1 for every begin state, MP included */
        band(y, j, banded, &dn, &dx);
        if (dx >= dn && NOT_IMPOSSIBLE(cm->beginsc[y])) add(vc[0], dx - dn + 1);
      }
  }
  for (int v = M - 1; v >= 0; v--) {
    if (cm->sttype[v] == B_st)      vc[v] += vc[cm->cnum[v]] + vc[cm->cfirst[v]];
    else if (cm->sttype[v] != E_st) vc[v] += vc[v + 1];
  }
  for (int v = M - 1; v >= 0; v--) vc[v] /= Leff;
  if (ret_calcs) *ret_calcs = vc[0];
  if (ret_vcalcs) *ret_vcalcs = vc; else free(vc);
  return eslOK;
}

extern "C" int infx_CountSearchCells(CM_t *cm, int qdbidx, int truncated, double *ret_cells, double *ret_ops)
{
  return infx_CountSearchWork(cm, qdbidx, truncated, ret_cells, ret_ops, NULL);
}

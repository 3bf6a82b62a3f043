/* cm_hits.cpp -- hit lists and hit resolution (host side).
 *
 * The GPU scanners hand back a compact candidate set {(j, d, sc, root, mode)}
 * that contains every cell the reference's per-j hit reporters could act on.
 * This file replays those reporters over the candidates in the reference's
 * order, so the resulting hit list is the one the reference builds:
 *   ReportHitsGreedily     src/cm_mx.c:7547-7628
 *   UpdateGammaHitMx       src/cm_mx.c:7440-7530, TBackGammaHitMx :7640-7669
 *   overlap removal        src/cm_dpsearch.c:651-661, src/cm_tophits.c:220-245,
 *                          :400-420, :1097-1275, CloneHitMostly :877-915
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <vector>

#include "cm_internal.h"

/* ------------------------------------------------------------------------ */
/* CM_TOPHITS                                                                */

extern "C" CM_TOPHITS *cm_tophits_Create(void)
{
  CM_TOPHITS *th = (CM_TOPHITS *) calloc(1, sizeof(CM_TOPHITS));
  th->Nalloc = 256;
  th->unsrt = (CM_HIT *) calloc(th->Nalloc, sizeof(CM_HIT));
  th->hit   = (CM_HIT **) calloc(th->Nalloc, sizeof(CM_HIT *));
  return th;
}

extern "C" void cm_tophits_Destroy(CM_TOPHITS *th)
{
  if (!th) return;
  free(th->unsrt); free(th->hit); free(th);
}

extern "C" int cm_tophits_Grow(CM_TOPHITS *th)
{
  if (th->N < th->Nalloc) return eslOK;
  uint64_t n = th->Nalloc * 2;
  CM_HIT *u = (CM_HIT *) realloc(th->unsrt, n * sizeof(CM_HIT));
  CM_HIT **h = (CM_HIT **) realloc(th->hit, n * sizeof(CM_HIT *));
  if (!u || !h) return eslEMEM;
  th->unsrt = u; th->hit = h; th->Nalloc = n;
  th->is_sorted_for_overlap_removal = 0;
  return eslOK;
}

/* src/cm_tophits.c:135 */
extern "C" int cm_tophits_CreateNextHit(CM_TOPHITS *th, CM_HIT **ret_hit)
{
  int status;
  if ((status = cm_tophits_Grow(th)) != eslOK) { *ret_hit = NULL; return status; }
  CM_HIT *hit = &th->unsrt[th->N];
  th->N++;
  if (th->N >= 2) th->is_sorted_for_overlap_removal = 0;
  memset(hit, 0, sizeof(CM_HIT));
  hit->start = 1; hit->stop = 1; hit->in_rc = 0; hit->score = 0.0;
  hit->cm_idx = -1; hit->seq_idx = -1; hit->pass_idx = -1;
  hit->hit_idx = (int64_t) th->N - 1;
  hit->srcL = -1; hit->hmmonly = 0; hit->flags = CM_HIT_FLAGS_DEFAULT;
  *ret_hit = hit;
  return eslOK;
}

/* src/cm_tophits.c:220 hit_sorter_for_overlap_removal, as a strict weak
 * order.  Ties on every key keep creation order: the sort is stable, as
 * glibc's merge-sort qsort is. */
static bool hit_before(const CM_HIT *h1, const CM_HIT *h2)
{
  if (h1->cm_idx  != h2->cm_idx)  return h1->cm_idx  < h2->cm_idx;
  if (h1->seq_idx != h2->seq_idx) return h1->seq_idx < h2->seq_idx;
  if (h1->in_rc   != h2->in_rc)   return h1->in_rc   < h2->in_rc;
  if (h1->score   != h2->score)   return h1->score   > h2->score;
  if (h1->start   != h2->start)   return h1->start   < h2->start;
  return false;
}

/* src/cm_tophits.c:400 */
extern "C" int cm_tophits_SortForOverlapRemoval(CM_TOPHITS *th)
{
  if (th->is_sorted_for_overlap_removal) return eslOK;
  for (uint64_t i = 0; i < th->N; i++) th->hit[i] = th->unsrt + i;
  if (th->N > 1) std::stable_sort(th->hit, th->hit + th->N, hit_before);
  th->is_sorted_for_overlap_removal = 1;
  return eslOK;
}

/* src/cm_tophits.c:1203 with remove_or_mark_overlaps_one_seq_memeff
 * (:1097-1150): within each (cm, seq, strand) run of the sorted list, a hit
 * is removed if it overlaps an earlier hit that was kept.  The set of kept
 * intervals is held in an ordered map, so the pass is O(N log N); the kept
 * set is the same as the reference's O(N^2) scan. */
extern "C" int cm_tophits_RemoveOrMarkOverlaps(CM_TOPHITS *th, int do_clans_only, char *errbuf)
{
  if (!th->is_sorted_for_overlap_removal) return infx_fail(eslEINVAL, errbuf, "cm_tophits_RemoveOrMarkOverlaps() list is not sorted appropriately");
  if (do_clans_only) return infx_fail(eslEINVAL, errbuf, "clan-restricted overlap removal is not supported");
  if (th->N < 2) return eslOK;
  uint64_t i = 0;
  while (i < th->N) {
    uint64_t j = i + 1;
    while (j < th->N && th->hit[j]->cm_idx == th->hit[i]->cm_idx && th->hit[j]->seq_idx == th->hit[i]->seq_idx &&
           th->hit[j]->in_rc == th->hit[i]->in_rc) j++;
    for (uint64_t h = i; h < j; h++) {
      int64_t srcL = th->hit[i]->srcL;
      if (th->hit[h]->srcL != srcL) return infx_fail(eslEINVAL, errbuf, "removing overlapping hits, srcL inconsistent");
      if (th->hit[h]->start < 1 || th->hit[h]->start > srcL || th->hit[h]->stop < 1 || th->hit[h]->stop > srcL)
        return infx_fail(eslERANGE, errbuf, "removing overlapping hits, hit position outside 1..srcL");
    }
    if (j > i + 1) {
      std::map<int64_t, int64_t> kept;   /* lo -> hi of kept hits, disjoint */
      for (uint64_t h = i; h < j; h++) {
        CM_HIT *x = th->hit[h];
        if (x->flags & CM_HIT_IS_REMOVED_DUPLICATE) continue;
        int64_t lo = x->in_rc ? x->stop : x->start;
        int64_t hi = x->in_rc ? x->start : x->stop;
        bool overlap = false;
        auto it = kept.upper_bound(hi);           /* first kept interval starting after hi */
        if (it != kept.begin()) { --it; if (it->second >= lo) overlap = true; }
        if (overlap) {
          x->flags |= CM_HIT_IS_REMOVED_DUPLICATE;
          x->flags &= ~CM_HIT_IS_REPORTED;
          x->flags &= ~CM_HIT_IS_INCLUDED;
        } else {
          kept[lo] = hi;
        }
      }
    }
    i = j;
  }
  return eslOK;
}

/* src/cm_tophits.c:877 */
extern "C" int cm_tophits_CloneHitMostly(CM_TOPHITS *src_th, int h, CM_TOPHITS *dest_th)
{
  CM_HIT *hit = NULL;
  int status;
  if ((status = cm_tophits_CreateNextHit(dest_th, &hit)) != eslOK) return status;
  const CM_HIT *s = src_th->hit[h];
  int64_t keep_idx = hit->hit_idx;
  *hit = *s;
  hit->hit_idx = keep_idx;
  return eslOK;
}

/* ------------------------------------------------------------------------ */
/* truncation pass helpers                                                   */

/* src/cm_dpsearch_trunc.c:3566 */
extern "C" int cm_TrFillFromPassIdx(int pass_idx, int *ret_fill_L, int *ret_fill_R, int *ret_fill_T)
{
  int L = 0, R = 0, T = 0, bad = 0;
  switch (pass_idx) {
  case PLI_PASS_5P_AND_3P_FORCE: L = R = T = 1; break;
  case PLI_PASS_5P_AND_3P_ANY:   L = R = T = 1; break;
  case PLI_PASS_5P_ONLY_FORCE:   R = 1; break;
  case PLI_PASS_3P_ONLY_FORCE:   L = 1; break;
  default: bad = 1; break;
  }
  if (ret_fill_L) *ret_fill_L = L;
  if (ret_fill_R) *ret_fill_R = R;
  if (ret_fill_T) *ret_fill_T = T;
  return bad ? eslEINVAL : eslOK;
}

/* src/cm_trunc.c:464 */
extern "C" int cm_tr_penalties_IdxForPass(int pass_idx)
{
  switch (pass_idx) {
  case PLI_PASS_5P_ONLY_FORCE:   return TRPENALTY_5P_ONLY;
  case PLI_PASS_3P_ONLY_FORCE:   return TRPENALTY_3P_ONLY;
  case PLI_PASS_5P_AND_3P_FORCE: return TRPENALTY_5P_AND_3P;
  case PLI_PASS_5P_AND_3P_ANY:   return TRPENALTY_5P_AND_3P;
  }
  return -1;
}

/* src/cm_pipeline.c:2396 */
extern "C" int cm_pli_PassAllowsTruncation(int pass_idx)
{
  switch (pass_idx) {
  case PLI_PASS_5P_ONLY_FORCE:
  case PLI_PASS_3P_ONLY_FORCE:
  case PLI_PASS_5P_AND_3P_FORCE:
  case PLI_PASS_5P_AND_3P_ANY: return 1;
  }
  return 0;
}

/* src/cm_tophits.c:3048 */
extern "C" int cm_hit_AllowTruncation(CM_t *cm, int pass_idx, int64_t start, int64_t stop, int64_t i0, int64_t j0, char mode, int b)
{
  int in_local_mode = (cm->flags & CMH_LOCAL_BEGIN) ? 1 : 0;
  int nd   = cm->ndidx[b];
  int lpos = (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd) ? cm->emap->lpos[nd] : cm->emap->lpos[nd] + 1;
  int rpos = (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATR_nd) ? cm->emap->rpos[nd] : cm->emap->rpos[nd] - 1;
  if (pass_idx == PLI_PASS_STD_ANY || pass_idx == PLI_PASS_5P_AND_3P_ANY || pass_idx == PLI_PASS_HMM_ONLY_ANY) return 1;
  if (start == i0 && stop == j0) return 1;
  if (in_local_mode) {
    switch (mode) {
    case TRMODE_J: return 1;
    case TRMODE_L: return (stop == j0) ? 1 : 0;
    case TRMODE_R: return (start == i0) ? 1 : 0;
    default:       return 0;
    }
  }
  switch (mode) {
  case TRMODE_J: return (lpos == 1 && rpos == cm->clen) ? 1 : 0;
  case TRMODE_L: return (stop == j0 && lpos == 1) ? 1 : 0;
  case TRMODE_R: return (start == i0 && rpos == cm->clen) ? 1 : 0;
  default:       return 0;
  }
}

/* ------------------------------------------------------------------------ */
/* replay of the per-j hit reporters over the candidate set                 */

static bool cand_before(const INFX_CAND &a, const INFX_CAND &b)
{
  if (a.j != b.j) return a.j < b.j;
  return a.d < b.d;
}

/* The permutation that orders key[0..n-1] ascending, ties in index order (a
 * stable LSD radix sort over 11-bit digits that carries the keys with the
 * indices, so every pass reads sequentially; digits equal in every key are
 * skipped).  O(n) per digit: the comparison sorts it replaces took most of
 * the replay for the tens of millions of candidates a long target gives. */
static void radix_order(std::vector<uint64_t> &key, std::vector<uint32_t> &perm)
{
  const size_t n = key.size();
  perm.resize(n);
  for (size_t i = 0; i < n; i++) perm[i] = (uint32_t) i;
  uint64_t all_or = 0, all_and = ~0ull;
  for (size_t i = 0; i < n; i++) { all_or |= key[i]; all_and &= key[i]; }
  std::vector<uint64_t> tkey(n);
  std::vector<uint32_t> tperm(n);
  std::vector<size_t> cnt(2049);
  for (int sh = 0; sh < 64; sh += 11) {
    if ((((all_or ^ all_and) >> sh) & 0x7ffull) == 0) continue;   /* this digit is the same in every key */
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; i++) cnt[((key[i] >> sh) & 0x7ff) + 1]++;
    for (size_t b = 1; b <= 2048; b++) cnt[b] += cnt[b - 1];
    for (size_t i = 0; i < n; i++) {
      const size_t o = cnt[(key[i] >> sh) & 0x7ff]++;
      tkey[o] = key[i];
      tperm[o] = perm[i];
    }
    key.swap(tkey);
    perm.swap(tperm);
  }
}

/* Candidates in (j, d) order: radix-ordered by j (every j of a call lies in
 * one 2^32 range above its smallest), then each j's run checked for
 * ascending d and sorted where it is not (the device writes one row's cells
 * in ascending d, so the runs come out sorted).  Small sets use std::sort.
 * The (j, d) pairs of one call are distinct, so the order is the one a
 * comparison sort gives. */
static void sort_cands(INFX_CAND *cand, int64_t n)
{
  if (n < ((int64_t) 1 << 16) || n > (int64_t) UINT32_MAX) { std::sort(cand, cand + n, cand_before); return; }
  int32_t jmin = INT32_MAX;
  for (int64_t k = 0; k < n; k++) jmin = std::min(jmin, cand[k].j);
  std::vector<uint64_t> key((size_t) n);
  for (int64_t k = 0; k < n; k++) key[k] = (uint64_t) ((int64_t) cand[k].j - (int64_t) jmin);
  std::vector<uint32_t> perm;
  radix_order(key, perm);
  std::vector<INFX_CAND> out((size_t) n);
  for (int64_t k = 0; k < n; k++) out[k] = cand[perm[k]];
  for (int64_t c = 0; c < n; ) {
    int64_t e = c + 1;
    while (e < n && out[e].j == out[c].j) e++;
    if (!std::is_sorted(out.begin() + c, out.begin() + e, cand_before)) std::sort(out.begin() + c, out.begin() + e, cand_before);
    c = e;
  }
  memcpy(cand, out.data(), sizeof(INFX_CAND) * (size_t) n);
}

/* order-preserving unsigned key of a non-NaN float (-0 and +0 equal) */
static uint32_t float_key(float x)
{
  uint32_t u;
  memcpy(&u, &x, 4);
  if (u == 0x80000000u) u = 0;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* Sliding act vectors: act[jp % (W+1)] = residue counts over i0..i0+jp-1,
 * accumulated in the reference's order (src/cm_dpsearch.c:223-230). */
struct ActRing {
  int W;
  std::vector<double> a;   /* (W+1) x K */
  int64_t jp_done = 0;
  ActRing(int W_) : W(W_), a((size_t) (W_ + 1) * INFX_K, 0.0) {}
  void advance_to(int64_t jp, const ESL_DSQ *dsq, int64_t i0) {
    while (jp_done < jp) {
      int64_t n = jp_done + 1;
      double *dst = &a[(size_t) (n % (W + 1)) * INFX_K];
      const double *src = &a[(size_t) ((n - 1) % (W + 1)) * INFX_K];
      for (int x = 0; x < INFX_K; x++) dst[x] = src[x];
      infx_abc_DCount(dst, dsq[i0 + n - 1], 1.);
      jp_done = n;
    }
  }
  const double *row(int64_t jp) const { return &a[(size_t) (jp % (W + 1)) * INFX_K]; }
};

static float null3_correction(const CM_t *cm, const ActRing &act, int64_t ip, int64_t jp, int d)
{
  float comp[INFX_K];
  const double *hi = act.row(jp), *lo = act.row(ip - 1);
  for (int a = 0; a < INFX_K; a++) comp[a] = hi[a] - lo[a];
  infx_vec_FNorm(comp, INFX_K);
  float corr = 0.;
  ScoreCorrectionNull3(cm->null, comp, d, cm->null3_omega, &corr);
  return corr;
}

extern "C" int infx_resolve_hits(CM_t *cm, char *errbuf, int pass_idx, int truncated, INFX_CAND *cand, int64_t ncand,
                                 ESL_DSQ *dsq, int64_t i0, int64_t j0, int W, float cutoff, int do_null3,
                                 int greedy, CM_TOPHITS *hitlist)
{
  if (hitlist == NULL) return eslOK;
  /* every candidate lies in the call's range (a cell the scan computed); a
   * set that breaks this is refused rather than indexed out of bounds */
  for (int64_t k = 0; k < ncand; k++)
    if (cand[k].j < i0 || cand[k].j > j0 || cand[k].d < 1 || (int64_t) cand[k].j - cand[k].d + 1 < i0)
      return infx_fail(eslEINVAL, errbuf, "hit replay: candidate (j %d, d %d) outside %lld..%lld", cand[k].j, cand[k].d,
                       (long long) i0, (long long) j0);
  sort_cands(cand, ncand);
  ActRing act(W);
  int status;

  if (greedy) {
    /* ReportHitsGreedily per j, then the overlap removal of the scan call
     * (src/cm_dpsearch.c:651-661) on compact records: sorted by score (desc),
     * start (asc) and creation order -- the order of the stable sort in
     * cm_tophits_SortForOverlapRemoval, since every hit of one call has the
     * same cm, sequence and strand -- and kept unless they overlap a kept hit
     * (a bitmap of kept positions).  Only the survivors become CM_HITs, built
     * and cloned exactly as cm_tophits_CloneHitMostly would. */
    /* positions as offsets from i0 (a call spans fewer than 2^31 residues) */
    struct CHit { int32_t lo, hi; float score; int32_t root; int32_t mode; uint32_t idx; };
    if (j0 - i0 >= (int64_t) INT32_MAX || ncand > (int64_t) UINT32_MAX)
      return infx_fail(eslERANGE, errbuf, "hit replay: %lld positions, %lld candidates", (long long) (j0 - i0 + 1), (long long) ncand);
    std::vector<CHit> hs;
    hs.reserve((size_t) ncand);
    int64_t c = 0;
    while (c < ncand) {
      int64_t j = cand[c].j, e = c;
      while (e < ncand && cand[e].j == j) e++;
      float max_sc_reported = IMPOSSIBLE;
      for (int64_t k = c; k < e; k++) {
        int d = cand[k].d;
        int64_t i = j - d + 1;
        float hit_sc = cand[k].sc;
        if (hit_sc > max_sc_reported && hit_sc >= cutoff && NOT_IMPOSSIBLE(hit_sc)) {
          int do_report = 1;
          if (truncated) do_report = cm_hit_AllowTruncation(cm, pass_idx, i, j, i0, j0, (char) cand[k].mode, cand[k].r);
          if (do_report && do_null3) {
            int64_t ip = i - i0 + 1, jp = j - i0 + 1;
            act.advance_to(jp, dsq, i0);
            hit_sc -= null3_correction(cm, act, ip, jp, d);
            do_report = (hit_sc > max_sc_reported && hit_sc >= cutoff) ? 1 : 0;
          }
          if (do_report) {
            hs.push_back(CHit{(int32_t) (i - i0), (int32_t) (j - i0), hit_sc, cand[k].r, cand[k].mode, (uint32_t) hs.size()});
            max_sc_reported = hit_sc;
          }
        }
      }
      c = e;
    }
    /* the order of the hits: score descending, start ascending, creation */
    std::vector<uint32_t> order;
    if (hs.size() < ((size_t) 1 << 16)) {
      order.resize(hs.size());
      for (size_t k = 0; k < hs.size(); k++) order[k] = (uint32_t) k;
      std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        const CHit &a = hs[x], &b = hs[y];
        if (a.score != b.score) return a.score > b.score;
        if (a.lo != b.lo) return a.lo < b.lo;
        return a.idx < b.idx;
      });
    } else {
      /* the same order by one 64-bit key: score descending, then start
       * ascending; hs is in creation order, and the radix sort is stable */
      std::vector<uint64_t> key(hs.size());
      for (size_t k = 0; k < hs.size(); k++)
        key[k] = ((uint64_t) ~float_key(hs[k].score) << 32) | (uint64_t) (uint32_t) hs[k].lo;
      radix_order(key, order);
    }
    const int64_t L = j0 - i0 + 1;
    std::vector<uint64_t> used((size_t) (L + 64) / 64 + 1, 0);
    auto any_used = [&](int64_t lo, int64_t hi) {        /* positions lo..hi, 0-based in the call's range */
      for (int64_t x = lo; x <= hi; ) {
        if ((x & 63) == 0 && x + 63 <= hi) { if (used[x >> 6]) return true; x += 64; }
        else { if (used[x >> 6] >> (x & 63) & 1ull) return true; x++; }
      }
      return false;
    };
    CM_TOPHITS *tmp = cm_tophits_Create();
    for (uint32_t q : order) {
      const CHit &h = hs[q];
      const int64_t lo = h.lo, hi = h.hi;
      if (any_used(lo, hi)) continue;
      for (int64_t x = lo; x <= hi; x++) used[x >> 6] |= 1ull << (x & 63);
      CM_HIT *hit;
      if ((status = cm_tophits_CreateNextHit(tmp, &hit)) != eslOK) { cm_tophits_Destroy(tmp); return infx_fail(status, errbuf, "out of memory"); }
      hit->start = i0 + h.lo; hit->stop = i0 + h.hi; hit->root = h.root;
      hit->mode = (char) h.mode; hit->score = h.score; hit->hmmonly = 0; hit->srcL = j0;
    }
    for (uint64_t h = 0; h < tmp->N; h++) tmp->hit[h] = tmp->unsrt + h;   /* already in sorted order */
    tmp->is_sorted_for_overlap_removal = 1;
    for (uint64_t h = 0; h < tmp->N; h++)
      if ((status = cm_tophits_CloneHitMostly(tmp, (int) h, hitlist)) != eslOK) { cm_tophits_Destroy(tmp); return infx_fail(status, errbuf, "problem copying hit to hitlist, out of memory?"); }
    cm_tophits_Destroy(tmp);
    return eslOK;
  }

  /* gamma semi-HMM (non-greedy) */
  int64_t L = j0 - i0 + 1;
  std::vector<float> mx(L + 1), savesc(L + 1);
  std::vector<int64_t> gback(L + 1);
  std::vector<int> saver(L + 1), savemode(L + 1);
  mx[0] = 0; gback[0] = -1;
  int64_t c = 0;
  for (int64_t j = i0; j <= j0; j++) {
    int64_t jp = j - i0 + 1;
    mx[jp] = mx[jp - 1] + 0;
    gback[jp] = -1; savesc[jp] = IMPOSSIBLE; saver[jp] = -1; savemode[jp] = -1;
    while (c < ncand && cand[c].j < j) c++;
    for (; c < ncand && cand[c].j == j; c++) {
      int d = cand[c].d;
      int64_t i = j - d + 1, ip = jp - d + 1;
      float hit_sc = cand[c].sc;
      float cumulative_sc = mx[ip - 1] + hit_sc;
      if (cumulative_sc > mx[jp] && NOT_IMPOSSIBLE(hit_sc)) {
        int do_report = 1;
        if (truncated) do_report = cm_hit_AllowTruncation(cm, pass_idx, i, j, i0, j0, (char) cand[c].mode, cand[c].r);
        if (do_report && do_null3) {
          act.advance_to(jp, dsq, i0);
          float corr = null3_correction(cm, act, ip, jp, d);
          hit_sc -= corr;
          cumulative_sc -= corr;
          do_report = (cumulative_sc > mx[jp]) ? 1 : 0;
        }
        if (do_report) {
          mx[jp] = cumulative_sc; gback[jp] = i; savesc[jp] = hit_sc;
          saver[jp] = cand[c].r; savemode[jp] = truncated ? cand[c].mode : TRMODE_J;
        }
      }
    }
  }
  int64_t j = j0;
  while (j >= i0) {
    int64_t jp = j - i0 + 1;
    if (gback[jp] == -1) j--;
    else {
      if (savesc[jp] >= cutoff) {
        CM_HIT *hit;
        if ((status = cm_tophits_CreateNextHit(hitlist, &hit)) != eslOK) return infx_fail(status, errbuf, "out of memory");
        hit->start = gback[jp]; hit->stop = j; hit->root = saver[jp];
        hit->mode = (char) savemode[jp]; hit->score = savesc[jp]; hit->hmmonly = 0;
      }
      j = gback[jp] - 1;
    }
  }
  return eslOK;
}

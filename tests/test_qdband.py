"""QDB band calculation (infx_CalculateQueryDependentBands, a restatement of
CalculateQueryDependentBands / BandCalculationEngine, src/cm_qdband.c:45-430)
pinned against the bands and W that the reference's cmbuild wrote into the
shipped model files (dmin2/dmin1/dmax1/dmax2 columns, W header line).

cmsearch uses the file's bands as they are (src/cm_modelconfig.c:192-205
recomputes them only when the betas change); cmbuild computed them from its
unrounded probabilities.  The files print the transition scores log2(t) to
three decimals, so every transition read back is within a factor
2^(+-0.0005) of cmbuild's.  A parse emitting n residues takes at most
n + nodes transitions (one state per node, plus one per inserted residue), so
every tail mass sum_{i>=n} gamma_v(i) or sum_{i<=n} gamma_v(i) moves by at
most a factor 2^(0.0005 (n + nodes)); a band edge can only move where the
tail mass at the edge lies that close to beta.  The test requires:

* W equal to the file's;
* every band column equal to the file's, or off by one at a near-tie: the
  engine's own density (ret_gamma) puts the tail mass at the contested
  length within that factor of beta.  Over the seven models 117 of 14,368
  columns are such ties, the largest 1.5% from beta (the factor allows
  2-26%, growing with the model and the length).  An engine mistake that
  moves edges (a wrong beta, an off-by-one in a tail sum) moves columns
  whose tail mass is far from beta and fails here.

tests/test_cmbuild.py checks the same engine on models rebuilt from the
tutorial alignments, whose probabilities are not rounded to three decimals:
there every column is exact except ties within 1e-3 of beta.
"""
import ctypes as C
import os

import numpy as np
import pytest

import infernal_amd as ia

CMDIR = os.path.join(os.path.dirname(__file__), "golden", "cm")
MODELS = ["tRNA.c.cm", "Vault.c.cm", "5S_rRNA.c.cm", "Cobalamin.c.cm", "Plant_SRP.c.cm", "rnaseP-eubact.c.cm", "tRNA5.c.cm"]
BANDS = ("dmin1", "dmax1", "dmin2", "dmax2")


def qdb_bands(cm):
    q = cm.c.qdbinfo.contents
    return [np.array([getattr(q, nm)[v] for v in range(cm.M)]) for nm in BANDS]


def engine_bands(cm, beta_W):
    """Bands, W, Z and the densities gamma[M][Z+1] from the engine, computed
    in place over cm's qdbinfo."""
    lib = ia.lib()
    f = lib.infx_CalculateQueryDependentBands
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    g = lib.infx_BandCalculationEngine
    g.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_int), C.c_void_p]
    W, Z, W2 = C.c_int(), C.c_int(), C.c_int()
    eb = C.create_string_buffer(128)
    q = C.cast(cm.c.qdbinfo, C.c_void_p)
    assert f(C.cast(cm.ptr, C.c_void_p), eb, q, beta_W, C.byref(W), C.byref(Z)) == 0, eb.value
    gamma = np.zeros((cm.M, Z.value + 1))
    assert g(C.cast(cm.ptr, C.c_void_p), Z.value, q, beta_W, C.byref(W2), gamma.ctypes.data) == 0
    assert W2.value == W.value
    return qdb_bands(cm), W.value, Z.value, gamma


def contested(got, want, gamma, beta1, beta2):
    """Per differing column: (band, state, file value, engine value, contested
    length n, tail mass at n / beta).  A left edge dmin is the first n with
    sum_{i<=n} gamma > beta, a right edge dmax the last n with
    sum_{i>=n} gamma > beta, so two edges one apart disagree exactly on
    whether the tail mass at min (left) or max (right) of the two exceeds
    beta."""
    out = []
    for k, nm in enumerate(BANDS):
        beta = beta1 if nm.endswith("1") else beta2
        for v in np.nonzero(got[k] != want[k])[0]:
            f, o = int(want[k][v]), int(got[k][v])
            if nm.startswith("dmin"):
                n = min(f, o)
                mass = gamma[v, :n + 1].sum()
            else:
                n = max(f, o)
                mass = gamma[v, n:].sum()
            out.append((nm, int(v), f, o, n, mass / beta))
    return out


def test_density_closed_forms():
    """ret_gamma against laws that follow from the model's structure alone:
    * an IL state whose children are itself (self-loop t0) and an END's E
      state (t1) emits a geometric number of residues:
      gamma(n) = t1 t0^(n-1) for n >= 1 (divided by its sum when that
      exceeds 1, as the engine renormalises);
    * a bifurcation's length is the sum of two independent lengths, so its
      mean and variance are the sums of its BEGL and BEGR children's;
    * every density sums to 1 within the float precision of the rows."""
    for name in ("tRNA.c.cm", "Cobalamin.c.cm"):
        cm = ia.CM.read(os.path.join(CMDIR, name))
        _, _, Z, gamma = engine_bands(cm, cm.c.beta_W)
        c = cm.c
        st = cm.sttype()
        n = np.arange(Z + 1, dtype=np.float64)
        geo = bif = 0
        for v in range(1, cm.M):
            if st[v] == ia.E_st:
                continue
            assert abs(gamma[v].sum() - 1.0) < 1e-4, (name, v)
            if st[v] == ia.IL_st and c.cnum[v] == 2 and st[c.cfirst[v] + 1] == ia.E_st:
                t0, t1 = float(np.float32(c.t[v][0])), float(np.float32(c.t[v][1]))
                want = np.zeros(Z + 1)
                want[1:] = t1 * t0 ** (n[1:] - 1)
                if want.sum() > 1.0:
                    want /= want.sum()
                np.testing.assert_allclose(gamma[v], want, rtol=1e-9, atol=1e-300)
                geo += 1
            if st[v] == ia.B_st:
                def moments(g):
                    m = (n * g).sum() / g.sum()
                    return m, ((n - m) ** 2 * g).sum() / g.sum()
                mb, vb = moments(gamma[v])
                ml, vl = moments(gamma[c.cfirst[v]])
                mr, vr = moments(gamma[c.cnum[v]])
                assert abs(mb - (ml + mr)) <= 1e-6 * mb, (name, v)
                assert abs(vb - (vl + vr)) <= 1e-6 * vb, (name, v)
                bif += 1
        assert geo >= 1 and bif >= 1, (name, geo, bif)


@pytest.mark.parametrize("name", MODELS)
def test_bands_match_shipped_files(name):
    cm = ia.CM.read(os.path.join(CMDIR, name))          # unconfigured: the file's probabilities
    want = qdb_bands(cm)
    W_file = cm.W
    got, W, Z, gamma = engine_bands(cm, cm.c.beta_W)
    q = cm.c.qdbinfo.contents
    keep = cm.sttype() != ia.E_st
    for g, w, nm in zip(got, want, BANDS):
        diff = np.abs(g[keep] - w[keep])
        assert diff.max() <= 1, (name, nm, np.nonzero(diff > 1)[0][:10])
    assert W == W_file, (W, W_file)
    ties = contested(got, want, gamma, q.beta1, q.beta2)
    for nm, v, f, o, n, ratio in ties:
        reach = 2.0 ** (0.0005 * (n + cm.nodes)) - 1.0
        assert abs(ratio - 1.0) <= reach, (name, nm, v, f, o, ratio, reach)
    worst = max((abs(t[5] - 1.0) for t in ties), default=0.0)
    print(name, "W", W, "Z", Z, "columns off by one: %d of %d, tail mass at most %.2g from beta" %
          (len(ties), 4 * int(keep.sum()), worst))


@pytest.mark.parametrize("name", MODELS)
def test_bands_independent_of_Z(name):
    """The reference's bandcyk-truncation-test (src/bandcyk-truncation-test.c,
    test 1): bands and W computed with two lengths Z are identical, since
    the engine refuses a Z whose truncated tail could move them (eslERANGE,
    checked here on a Z far too short)."""
    cm = ia.CM.read(os.path.join(CMDIR, name))
    bands1, W1, Z1, _ = engine_bands(cm, cm.c.beta_W)
    lib = ia.lib()
    g = lib.infx_BandCalculationEngine
    g.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_int), C.c_void_p]
    q = C.cast(cm.c.qdbinfo, C.c_void_p)
    W2 = C.c_int()
    assert g(C.cast(cm.ptr, C.c_void_p), 3 * Z1, q, cm.c.beta_W, C.byref(W2), None) == 0
    bands2 = qdb_bands(cm)
    assert W2.value == W1
    keep = cm.sttype() != ia.E_st
    for a, b in zip(bands1, bands2):
        assert np.array_equal(a[keep], b[keep])
    st = g(C.cast(cm.ptr, C.c_void_p), max(4, cm.clen // 4), q, cm.c.beta_W, C.byref(W2), None)
    assert st == 16          # eslERANGE: Z too short

"""Calibration on the GPU scanners vs the tRNA5 tutorial model's shipped
ECMLC / ECMLI lines (cmcalibrate's defaults: 1.6 Mb from the generative HMM,
QDB2, NULL3, every hit).  The random stream is not Easel's, so the bar is
statistical: hit count within 3%, lambda within 15% (the local tail has
1,200 points, a standard error near 3%), mu_orig within 0.5 bit and
mu_extrap within 0.6 bit."""
import os

import pytest

from infernal_amd import calibrate as cal

pytestmark = pytest.mark.gpu

CM = os.path.join(os.path.dirname(__file__), "golden", "cm", "tRNA5.c.cm")


def test_calibrate_local_tails_match_shipped():
    print("case calibrate tRNA5 LC LI 1.6 Mb", flush=True)
    got, scores = cal.calibrate(CM, L_mb=1.6, seed=181, modes=("LC", "LI"))
    ship = cal.read_exp_infos(CM)
    for mode in ("LC", "LI"):
        g, s = got[mode], ship[mode]
        assert abs(g.nrandhits / s.nrandhits - 1) < 0.03, (mode, g.nrandhits, s.nrandhits)
        assert abs(g.lambda_ / s.lambda_ - 1) < 0.15, (mode, g.lambda_, s.lambda_)
        assert abs(g.mu_orig - s.mu_orig) < 0.5, (mode, g.mu_orig, s.mu_orig)
        assert abs(g.mu_extrap - s.mu_extrap) < 0.6, (mode, g.mu_extrap, s.mu_extrap)

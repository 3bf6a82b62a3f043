"""Calibrate the synthetic benchmark models (E-value tails, as cmcalibrate
does) with the GPU scanners and write the calibrated models to
gpurun_out/calibrated/; copy them over tests/golden/cm/ afterwards.  Runs on
a machine with the GPU, from the repository root:

    python tests/golden/calibrate_synthetic_cms.py [names...]

The 1.6 Mb of generated sequence is scanned as one target per mode."""
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from infernal_amd import calibrate as cal  # noqa: E402

CMDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cm")


def main(names):
    out = os.path.join("gpurun_out", "calibrated")
    os.makedirs(out, exist_ok=True)
    for name in names:
        t = time.time()
        path = os.path.join(CMDIR, name + ".cm")
        infos, scores = cal.calibrate(path, joined=True)
        with open(path) as fh:
            text = cal.with_exp_lines(fh.read(), infos)
        with open(os.path.join(out, name + ".cm"), "w") as fh:
            fh.write(text)
        print(name, "%.1f s" % (time.time() - t), {m: (round(e.lambda_, 4), round(e.mu_extrap, 3), e.nrandhits)
                                                  for m, e in infos.items()}, flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["synth_cloverleaf76", "synth_multiloop350"])

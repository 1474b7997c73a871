"""CPU: the FP64 oracle adjoint (tests/adjoint_ref.py) of a small 2D SRD
channel agrees with central finite differences of the oracle's own
observable taken with common random numbers (the same Philox draws).  The
step h is small enough that no particle changes cell between the two runs,
which is the regime the pathwise gradient describes."""
import pytest

import adjoint_ref as A


@pytest.mark.parametrize("which", ["g", "angle"])
def test_cpu_adjoint_matches_finite_differences(which):
    Lx, Ly, ppc, K, seed = 12, 8, 10, 30, 41
    g, ang = 5e-4, 2.0
    band = (3, 5)
    Q, V = A.initial(Lx, Ly, ppc, 7)
    cfg = A.channel_cfg(Lx, Ly, ppc, g, ang)
    J, steps = A.forward(cfg, Q, V, K, seed, band)
    dg, da = A.backward(cfg, steps, K)
    h = 1e-9
    if which == "g":
        Jp, _ = A.forward(A.channel_cfg(Lx, Ly, ppc, g + h, ang), Q, V, K, seed, band, tape=False)
        Jm, _ = A.forward(A.channel_cfg(Lx, Ly, ppc, g - h, ang), Q, V, K, seed, band, tape=False)
        fd, ad = (Jp - Jm) / (2 * h), dg[0]
    else:
        Jp, _ = A.forward(A.channel_cfg(Lx, Ly, ppc, g, ang + h), Q, V, K, seed, band, tape=False)
        Jm, _ = A.forward(A.channel_cfg(Lx, Ly, ppc, g, ang - h), Q, V, K, seed, band, tape=False)
        fd, ad = (Jp - Jm) / (2 * h), da
    print(which, J, fd, ad)
    assert ad == pytest.approx(fd, rel=1e-5)

"""Golden rate vectors for the REACLIB 7-parameter formula (reaclib.F90:431-582), evaluated with
mpmath at 50 significant digits and rounded once to double: for each reaction
v = max(exp(a0 + a1/T9 + a2 T9^-1/3 + a3 T9^1/3 + a4 T9 + a5 T9^5/3 + a6 ln T9) * (rho if two-body), 1e-99)
with T9 clamped to >= 0.01.  Inputs are the first 400 reactions of the 270-species synthetic
network plus hand-picked coefficient rows (pure decay, overflow-free extremes).
Run:  python tests/golden/make_rates_golden.py
"""
import importlib.util
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(os.path.dirname(HERE)), "nuppn_modular2-ininet3_amd")
spec = importlib.util.spec_from_file_location("nuppn_network", os.path.join(PKG, "network.py"))
network = importlib.util.module_from_spec(spec)
sys.modules["nuppn_network"] = network
spec.loader.exec_module(network)

mp.mp.dps = 50


def rate(a, rt, t9, rho):
    t = max(mp.mpf(t9), mp.mpf("0.01"))
    e = (a[0] + a[1] / t + a[2] * t ** (mp.mpf(-1) / 3) + a[3] * t ** (mp.mpf(1) / 3) + a[4] * t
         + a[5] * t ** (mp.mpf(5) / 3) + a[6] * mp.log(t))
    r = mp.exp(e)
    if rt == network.RTYPE_CAPTURE:
        r = r * mp.mpf(rho)
    return max(float(r), 1e-99)


def main():
    net = network.build_network("c0_270", 30, -2, 6, None, network.SEED_BASE)
    nr = 400
    coef = [list(map(mp.mpf, net.coef[:, l])) for l in range(nr)]
    rtype = list(net.rtype[:nr])
    extra = [([20.0, -0.1, -1.0, 0.3, -0.2, 0.1, -1.4], 0),
             ([-10.0, -40.0, -30.0, 0.0, 0.0, 0.0, 0.0], 0),
             ([float(np.log(3.0e-5)), 0, 0, 0, 0, 0, 0], 2),
             ([24.9, -0.0, -0.0, 1.5, 1.5, 1.5, 0.0], 1)]
    for a, rt in extra:
        coef.append(list(map(mp.mpf, a)))
        rtype.append(rt)
    points = [(0.005, 1e3), (0.05, 10.0), (0.2, 1e2), (0.5, 1e4), (1.0, 3e5), (3.0, 1e8), (4.0, 1e9)]
    v = np.zeros((len(points), len(coef)))
    for i, (t9, rho) in enumerate(points):
        for l, a in enumerate(coef):
            v[i, l] = rate(a, rtype[l], t9, rho)
    np.savez(os.path.join(HERE, "rates_mpmath.npz"),
             coef=np.array([[float(x) for x in a] for a in coef]).T.copy(),
             rtype=np.array(rtype, dtype=np.int32), t9=np.array([p[0] for p in points]),
             rho=np.array([p[1] for p in points]), v=v)
    print("wrote rates_mpmath.npz", v.shape)


if __name__ == "__main__":
    main()

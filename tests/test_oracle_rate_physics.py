"""Pins of the oracle's rate physics (oracle/nuppn_oracle.c: partition_functions, reverse_gfac,
reverse_pairs, detailed_balance, screening, orc_rates_full) by what the physics and mathematics
fix, each chosen so that a plausible slip fails it:

* Gfac: a REACLIB reverse rate with partition functions on, divided by the same rate without,
  is G(ejectile) G(residual) / G(target) with the tables interpolated linearly at the clipped
  temperature (numpy's interp; a swapped participant, a wrong interval or no clipping fails);
  no other rate changes;
* detailed balance: reverse_init pairs every photodisintegration with the capture of mirrored
  topology; the reverse-to-forward ratio is the Saha constant written with quantum
  concentrations, n_Q(m) = (2 pi m k T / h^2)^(3/2): v_rev / v_fwd = G_a n_Q(m_a) G_b n_Q(m_b) /
  (G_c n_Q(m_c)) exp(-Q / kT) / (rho N_A) (a wrong exponent, mass ratio or sign of Q fails);
  for an identical pair the reference's Q leaves out the second reactant's binding energy;
* screening: no factor for a reaction with an uncharged reactant; in the weak-screening limit
  the factor tends to exp(sqrt(3) Gamma^(3/2)) (Debye-Hueckel, which Chugunov's fit is built to
  reproduce) with Gamma = Z1 Z2 e^2 / (a kT) and a the ion-sphere radius of the plasma's ions
  (a wrong charge product or density fails); an identical pair screens Z with itself; below
  0.05 T_p the factor no longer depends on the temperature (the fade's floor);
* the T9 cache built with partition functions holds the reverse rates with their factors.
"""
import math

import numpy as np
import pytest

AMU = 1.6605402e-24
KB = 1.3806485279e-16
QE = 4.80320425e-10
NA = 6.022140857e23
H_ERG = 4.135667662e-21 * 1.602177e-12 * 1.0e6
MEV2ERG = 1.602177e-12 * 1.0e6


@pytest.fixture(scope="module")
def c1(pkg, oracle):
    w = pkg.network.config(1, nz=4)
    rp = pkg.network.rate_physics(w.net)
    return w, rp


def _oracle(oracle, w, rp=None, **over):
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    if rp is not None:
        import dataclasses
        o.set_rate_physics(dataclasses.replace(rp, **over))
    return o


def _interp(rp, s, t9):
    return np.interp(t9, rp.pf_t9, rp.pf_tab[s])


@pytest.mark.parametrize("t9", [0.05, 0.1, 0.35, 1.0, 2.2, 10.0, 12.0])
def test_gfac_is_the_partition_function_ratio(oracle, c1, t9):
    w, rp = c1
    bare = _oracle(oracle, w).rates(t9, 1e4)
    o = _oracle(oracle, w, rp, screen=0, detailed_balance=0)
    v = o.rates(t9, 1e4)
    photo = rp.rev_ch > 0
    np.testing.assert_array_equal(v[~photo], bare[~photo])
    k = w.net.k
    idx = np.where(photo & (bare > 1e-80) & (v > 1e-80))[0]
    assert idx.size > 100
    tgt, ej, res = rp.rev_part[0, idx], rp.rev_part[1, idx], rp.rev_part[2, idx]
    g = np.array([_interp(rp, s, t9) for s in range(w.net.nsp)])
    np.testing.assert_allclose(v[idx] / bare[idx], g[ej] * g[res] / g[tgt], rtol=1e-13)
    # participants as the network defines the reaction: target k1, products k5 / k7
    np.testing.assert_array_equal(tgt, k[0, idx])
    np.testing.assert_array_equal(res, k[6, idx])


def test_reverse_pairs(oracle, c1):
    w, rp = c1
    o = _oracle(oracle, w, rp)
    irev = o.reverse_pairs()
    k, rt = w.net.k, w.net.rtype
    cap = np.where(rt == 0)[0]
    assert np.all(irev[rt != 0] == -1)
    assert np.all(irev[cap] >= 0)
    j = irev[cap]
    assert np.all(rt[j] == 1)
    # mirrored topology: the photodisintegration's target is the capture's product, and it
    # gives back the capture's target and projectile
    np.testing.assert_array_equal(k[0, j], k[6, cap])
    np.testing.assert_array_equal(k[6, j], k[0, cap])
    assert len(set(j.tolist())) == cap.size


@pytest.mark.parametrize("t9,rho", [(0.3, 1e3), (1.5, 1e6), (3.0, 1e2)])
def test_detailed_balance_is_the_saha_constant(oracle, c1, t9, rho):
    w, rp = c1
    o = _oracle(oracle, w, rp, screen=0)
    v = o.rates_full(t9, rho, w.x0[0])
    irev = o.reverse_pairs()
    k = w.net.k
    kT = KB * t9 * 1e9
    g = np.array([_interp(rp, s, t9) for s in range(w.net.nsp)]) * rp.g_spin

    def nq(m):
        return (2.0 * math.pi * m * kT / H_ERG ** 2) ** 1.5

    checked = 0
    for i in np.where(irev >= 0)[0]:
        a, b, c = k[0, i], k[2, i], k[6, i]
        if k[1, i] == 2:
            # an identical pair (alpha + alpha): the reference takes G and m of the second
            # reactant from the first but its binding energy from k3, the gamma after folding
            # (reverse.F90:236,253-256,302), so Q misses one reactant; checked below
            continue
        # (Q in the reference's order: it cancels to a small difference of large binding
        # energies, so another order moves exp(-Q / kT) by more than the comparison allows)
        q = -rp.q_mev[a] - rp.q_mev[b] + rp.q_mev[c]
        x = -q * MEV2ERG / kT
        if x > 190 or v[i] <= 1e-90:
            continue
        K = g[a] * nq(rp.mass_g[a]) * g[b] * nq(rp.mass_g[b]) / (g[c] * nq(rp.mass_g[c])) \
            * math.exp(x) / (rho * NA)
        if not 1e-90 < K < 1e90 or v[i] * K < 1e-90:
            continue
        assert v[irev[i]] == pytest.approx(v[i] * K, rel=1e-12)
        checked += 1
    assert checked > 500
    for i in np.where((irev >= 0) & (k[1] == 2))[0]:
        a, c = k[0, i], k[6, i]
        x = min(-(-rp.q_mev[a] - rp.q_mev[k[2, i]] + rp.q_mev[c]) * MEV2ERG / kT, 200.0)
        K = g[a] * nq(rp.mass_g[a]) * g[a] * nq(rp.mass_g[a]) / (g[c] * nq(rp.mass_g[c])) \
            * math.exp(x) / (rho * NA)
        K = min(max(K, 1e-99), 1e99)
        assert v[irev[i]] == pytest.approx(max(v[i] * K, 1e-99), rel=1e-12)


def _plasma_gamma(w, x, t9, rho, z1z2):
    # ion-sphere radius of the ions: a = (3 / (4 pi n_ion))^(1/3), n_ion = rho / (abar amu)
    a = w.net.aw[1:]
    y = x[1:] / a
    abar = np.sum(a * y) / np.sum(y)
    n_ion = rho / (abar * AMU)
    rad = (3.0 / (4.0 * math.pi * n_ion)) ** (1.0 / 3.0)
    return z1z2 * QE ** 2 / (rad * KB * t9 * 1e9)


def test_screening_factors(oracle, c1):
    w, rp = c1
    x = w.x0[0]
    bare = _oracle(oracle, w)
    # screening alone: no partition functions (an empty grid), no detailed balance
    scr = _oracle(oracle, w, rp, pf_t9=rp.pf_t9[:0], pf_tab=rp.pf_tab[:, :0],
                  detailed_balance=0)
    k, zw = w.net.k, w.net.zw
    r2 = np.where(k[1] == 2, k[0], k[2])
    z1z2 = zw[k[0]] * zw[r2]
    # weak screening: hot and thin
    t9, rho = 5.0, 1e-2
    f = scr.rates_full(t9, rho, x) / bare.rates(t9, rho)
    charged = (z1z2 > 0) & (bare.rates(t9, rho) > 1e-90)
    np.testing.assert_array_equal(f[z1z2 == 0], 1.0)
    gam = _plasma_gamma(w, x, t9, rho, z1z2[charged])
    assert gam.max() < 2e-3
    np.testing.assert_allclose(np.log(f[charged]), math.sqrt(3.0) * gam ** 1.5, rtol=0.03)
    # the alpha + alpha pair: Z = 2 with itself
    aa = np.where((k[1] == 2) & (zw[k[0]] == 2))[0]
    assert aa.size >= 1 and np.all(z1z2[aa] == 4.0)
    # below 0.05 T_p the temperature no longer enters (the fade's floor tk = 0.1 T_p)
    rho = 1e11   # T_p ~ 2 GK: T9 0.01 and 0.02 are both below 0.05 T_p
    fa = scr.rates_full(0.01, rho, x) / bare.rates(0.01, rho)
    fb = scr.rates_full(0.02, rho, x) / bare.rates(0.02, rho)
    ok = charged & (fa > 1.0) & (bare.rates(0.01, rho) > 1e-80) & (bare.rates(0.02, rho) > 1e-80)
    assert ok.sum() > 50
    np.testing.assert_allclose(fa[ok], fb[ok], rtol=1e-12)


def test_cache_holds_reverse_factors(oracle, c1):
    w, rp = c1
    o = _oracle(oracle, w, rp, screen=0, detailed_balance=0)
    o.set_rate_cache(True)
    tab = o.rate_cache()
    o2 = _oracle(oracle, w, rp, screen=0, detailed_balance=0)
    for i in (0, 77, 2500, 5000):
        t9 = 0.01 + i * (10.0 - 0.01) / 5000
        np.testing.assert_array_equal(tab[i], o2.rates(t9, 1.0))

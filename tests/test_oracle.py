"""CPU pins of the oracle (oracle/mpf_oracle.c).

Each oracle function is checked against something it cannot be tautologically
equal to:
  * golden vectors written by tests/golden/make_golden.py from the reference's
    own kernels executed in IEEE double (bit-exact with use_libm=1);
  * the a(theta)/b(theta) known-answer tables printed by the reference
    notebook (GPU-multi-phase-field-model-solification_2d.ipynb, cells 16-17);
  * mathematical limits: the isotropic limit of the anisotropy terms, the
    derivative of a(|u.g|) by finite differences, conservation and bounds.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _g(name):
    return np.load(os.path.join(GOLD, name))


def _scal(arr):
    return dict(zip(O.PARAM_KEYS, arr))


def _tabs3():
    gi = _g("inputs.npz")
    return tuple(gi[k].astype(np.float64) for k in ("wij", "aij", "mij", "n111"))


@pytest.mark.parametrize("tag,aniso,torque,n", [("full", 1, 1, 3), ("notorque", 1, 0, 2), ("iso", 0, 0, 2)])
def test_update3d_bit_exact_vs_reference(tag, aniso, torque, n):
    g = _g("step3d.npz")
    scal = _scal(g["scalars"])
    for k in range(1, n + 1):
        ph, _, done, _ = O.run(g["phi0"], g["temp0"], k, _tabs3(), scal, kmax=50, aniso=aniso, torque=torque,
                               use_libm=True)
        assert done == k
        assert np.array_equal(ph, g[f"{tag}_phi{k}"]), f"{tag} step {k}"
    # the GPU-identical transcendentals stay within 1e-12 of libm's
    ph, _, _, _ = O.run(g["phi0"], g["temp0"], n, _tabs3(), scal, kmax=50, aniso=aniso, torque=torque)
    assert np.abs(ph - g[f"{tag}_phi{n}"]).max() <= 1e-12


def test_apt_bit_exact_vs_reference():
    g = _g("step3d.npz")
    mf, nf, st = O.apt(g["phi0"], 50)
    assert np.array_equal(nf, g["full_nf1"])
    assert np.array_equal(st, g["full_st1"])
    mask = np.arange(50)[:, None, None, None] < nf[None]
    assert np.array_equal(np.where(mask, mf, -1), np.where(mask, g["full_mf1"], -1))
    mf2, nf2, st2 = O.apt(g["full_phi1"], 50)
    assert np.array_equal(nf2, g["full_nf2"]) and np.array_equal(st2, g["full_st2"])


def test_directional_cooling_column_vs_reference():
    g = _g("step3d.npz")
    scal = _scal(g["scalars"])
    from mpf import inputs
    from cases import ref_params, tables_for
    d, tabs = tables_for(ref_params(), g["dir_quats"])
    ph, T, done, _ = O.run(g["dir_phi0"], g["dir_temp0"], 3, tabs, scal, kmax=50,
                           cooling_rate=float(g["dir_cool"]), use_libm=True)
    assert np.array_equal(ph, g["dir_phi3"])
    assert np.array_equal(T, g["dir_temp3"])
    assert inputs is not None


def test_update2d_bit_exact_vs_reference():
    g = _g("step2d.npz")
    from mpf import inputs
    n111 = inputs.compute_rotated_n111(g["quats"]).astype(np.float64)
    tabs = (g["wij"].astype(np.float64), g["aij"].astype(np.float64), g["mij"].astype(np.float64), n111)
    scal = _scal(g["scalars"])
    p1, _, _, _ = O.run(g["phi0"], g["temp0"], 1, tabs, scal, use_libm=True)
    assert np.array_equal(p1, g["phi1"])
    p3, _, _, _ = O.run(g["phi0"], g["temp0"], 3, tabs, scal, use_libm=True)
    assert np.array_equal(p3, g["phi3"])


def test_per_function_probes_vs_reference():
    g = _g("probes.npz")
    st = _g("step3d.npz")
    scal = _scal(st["scalars"])
    phi0 = st["phi0"]
    n111 = _tabs3()[3]
    for row in g["probes"]:
        l, m, k, s = (int(v) for v in row[:4])
        pr = O.probe3(phi0, n111, scal, l, m, k, s, use_libm=True)
        assert pr["aniso_term1"] == row[4]
        assert pr["torque"] == row[5]
        assert pr["best_cos"] == row[6]
        assert pr["b_sl"] == row[7]
        np.testing.assert_array_equal(pr["A12"], row[8:12])
        np.testing.assert_array_equal(pr["A13"], row[12:15])
    for c, a in zip(g["cos_grid"], g["a_vals"]):
        assert O.calc_a(c, scal, use_libm=True) == a
    sb = dict(scal, ksi=0.3, theta_c_rad=10.0 * math.pi / 180.0)
    for t, b in zip(np.linspace(0.0, 12.0, 49), g["b_vals"]):
        assert O.calc_b(math.cos(math.radians(t)), sb, use_libm=True) == b


# Known-answer tables printed by the reference notebook (cell 17 and cell 16).
A_KAT = {0.0: 0.85314282, 10.0: 0.89069032, 20.0: 0.93233553, 35.0: 0.97768503, 45.0: 0.99457414,
         54.7: 1.00007059, 60.0: 0.99847758, 70.0: 0.98674219, 90.0: 0.94007131}
B_KAT = {2.0: 0.52648053, 5.0: 0.83311591, 8.0: 0.97636397, 10.0: 1.0, 12.0: 1.0, 20.0: 1.0,
         9.5: 0.99855831, 9.75: 0.99964002, 9.95: 0.99998561}


@pytest.mark.parametrize("use_libm", [True, False])
def test_notebook_known_answers(use_libm):
    s = dict(zip(O.PARAM_KEYS, [0.0] * 13))
    s.update(a0=math.radians(54.7), delta_a=0.36, mu_a=0.6156, p_round=0.05, ksi=0.3,
             theta_c_rad=math.radians(10.0))
    for th, want in A_KAT.items():
        assert O.calc_a(math.cos(math.radians(th)), s, use_libm) == pytest.approx(want, abs=5e-9)
    for th, want in B_KAT.items():
        assert O.calc_b(math.cos(math.radians(th)), s, use_libm) == pytest.approx(want, abs=5e-9)
    # cell 18 evaluates the kernel's calc_b (eps = 1e-6): b(0) = ksi, b(5 deg) = 0.8331
    assert O.calc_b(1.0, s, use_libm) == pytest.approx(0.3, abs=5e-5)


def test_isotropic_limit_of_anisotropy_terms():
    """delta_a = 0 makes a(theta) == mu_a: aniso_term1 reduces to the flux form
    of (eps0 mu)^2 lap(phi_L) and the torque vanishes."""
    st = _g("step3d.npz")
    scal = dict(_scal(st["scalars"]), delta_a=0.0)
    phi = st["phi0"]
    n111 = _tabs3()[3]
    dx = scal["dx"]
    e2 = (scal["eps0_sl"] * scal["mu_a"]) ** 2
    L = phi[0]
    for (l, m, k) in [(3, 2, 4), (5, 4, 6), (11, 7, 3)]:
        pr = O.probe3(phi, n111, scal, l, m, k, 1)
        nx, ny, nz = L.shape
        lap = (L[(l + 1) % nx, m, k] + L[(l - 1) % nx, m, k] + L[l, (m + 1) % ny, k] + L[l, (m - 1) % ny, k]
               + L[l, m, min(k + 1, nz - 1)] + L[l, m, max(k - 1, 0)] - 6 * L[l, m, k]) / dx ** 2
        assert pr["aniso_term1"] == pytest.approx(e2 * lap, rel=1e-12, abs=1e-6)
        assert pr["torque"] == 0.0


def test_A12_is_the_gradient_derivative_of_a():
    """da/dphi_p (A12) equals the central-difference derivative of
    a(|u . g/|g||) with respect to the gradient components, away from facet
    switches."""
    st = _g("step3d.npz")
    scal = _scal(st["scalars"])
    n111 = _tabs3()[3]
    phi = st["phi0"].copy()
    l, m, k, s = 3, 2, 4, 2
    pr = O.probe3(phi, n111, scal, l, m, k, s)
    dx = scal["dx"]
    L = phi[0]
    nx, ny, nz = L.shape
    g = np.array([(L[(l + 1) % nx, m, k] - L[(l - 1) % nx, m, k]) / (2 * dx),
                  (L[l, (m + 1) % ny, k] - L[l, (m - 1) % ny, k]) / (2 * dx),
                  (L[l, m, k + 1] - L[l, m, k - 1]) / (2 * dx)])
    normals = n111[s]

    def a_of(gv):
        u = gv / np.linalg.norm(gv)
        c = np.max(np.abs(normals @ u))
        return O.calc_a(c, scal)

    for p in range(3):
        h = 1e-6 * np.linalg.norm(g)
        e = np.zeros(3)
        e[p] = h
        fd = (a_of(g + e) - a_of(g - e)) / (2 * h)
        assert pr["A12"][1 + p] == pytest.approx(fd, rel=1e-5, abs=1e-12)


def test_conservation_and_bounds_over_a_run():
    g = _g("step3d.npz")
    scal = _scal(g["scalars"])
    ph, _, _, st = O.run(g["phi0"], g["temp0"], 8, _tabs3(), scal, kmax=50)
    assert np.abs(ph.sum(0) - 1.0).max() < 1e-12
    assert ph.min() >= 0.0 and ph.max() <= 1.0
    # inactive phases are exactly zero: a phase can only be non-zero where the
    # APT of the previous state listed it
    mf, nf, _ = O.apt(ph, 50)
    assert st[1] <= 50


def test_apt_probes_vs_reference():
    g = _g("apt.npz")
    mf, nf, st = O.apt(g["kmax_phi"].astype(np.float64), 50, depth=52)
    assert list(st) == [1, 52] and list(g["kmax_st"]) == [1, 52]
    assert np.array_equal(nf, g["kmax_nf"])
    assert np.array_equal(mf[:50, 1, 1, 1], g["kmax_mf"][:50, 1, 1, 1])
    # the reference scan reports the first overflowing cell at (0, 1, 1)
    first = next((l, m, k) for l in range(4) for m in range(4) for k in range(4)
                 if _count(g["kmax_phi"], l, m, k) > 50)
    assert first == (0, 1, 1)
    mf, nf, st = O.apt(g["bc_phi"].astype(np.float64), 50)
    assert np.array_equal(nf, g["bc_nf"])
    mask = np.arange(50)[:, None, None, None] < nf[None]
    assert np.array_equal(np.where(mask, mf, -1), np.where(mask, g["bc_mf"], -1))

    def has(cell, gid):
        l, m, k = cell
        return gid in mf[:nf[l, m, k], l, m, k].tolist()
    assert has((5, 2, 2), 1)       # x periodic wrap
    assert has((3, 4, 2), 2)       # y periodic wrap
    assert has((3, 2, 1), 3)       # z mirror neighbour
    assert not has((3, 2, 4), 3)   # no wrap in z


def _count(phi, l, m, k):
    N, nx, ny, nz = phi.shape
    n = 0
    for i in range(N):
        v = phi[i, l, m, k]
        nb = [phi[i, (l + 1) % nx, m, k], phi[i, (l - 1) % nx, m, k], phi[i, l, (m + 1) % ny, k],
              phi[i, l, (m - 1) % ny, k], phi[i, l, m, min(k + 1, nz - 1)], phi[i, l, m, max(k - 1, 0)]]
        if v > 0 or (v == 0 and any(x > 0 for x in nb)):
            n += 1
    return n


def test_halt_on_kmax_keeps_state():
    g = _g("apt.npz")
    phi = g["kmax_phi"].astype(np.float64)
    N = phi.shape[0]
    from cases import ref_params, tables_for
    tabs = tables_for(ref_params(), np.tile([0, 0, 0, 1.0], (N, 1)))[1]
    scal = _scal(_g("step3d.npz")["scalars"])
    out, T, done, st = O.run(phi, np.full((4, 4, 4), 1680.0), 3, tabs, scal, halt_on_kmax=True,
                             cooling_rate=0.5)
    assert done == 0 and list(st) == [1, 52]
    assert np.array_equal(out, phi)
    assert np.all(T == 1679.5)   # the temperature step ran before the APT

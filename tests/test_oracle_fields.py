"""The oracle's planar near fields (the reference's rcwa::solve_with_fields + field_at + field_map, built from its
headers into oracle/_ref) pinned by physics the fields must obey, so that the GPU sampler's parity against them
means something:

* Poynting flux: the x-averaged z-flux sum_j Re(Ex_j Hy_j* - Ey_j Hx_j*) (orders are orthogonal over a period)
  is the same at every depth of a lossless stack -- in both outer media, inside grating layers and films --
  and falls monotonically through absorbing layers; the incidence-side flux times sum(DEt) / (1 - sum(DEr))
  is the exit-side flux
* tangential continuity: Ey, Hx (TE) and Hy, Ex (TM) take the same harmonic values just above and below every
  interface
* the grid is the carrier sum of the harmonics: field_map at x_k = k period / nx equals
  sum_j exp(-i kxm_j x_k) h_j (rcwa_fields.hpp:326-355), with kxm_j = k0 (nI sin(theta) - m_j lambda / period)
  (rcwa1d.hpp:410-415)
* a planted mistake (a+ and a- exchanged in a layer) breaks the flux invariant
No GPU is used.
"""
import numpy as np
import pytest

LOSSLESS = dict(nI=1.2, nII=1.4, period=0.6)


def lossless_stack():
    from stablercwa_amd.types import Layer, Stack
    return Stack(1.2, [Layer.grating(.2, [0., .21, .6], [1.7, 1.]), Layer.uniform(.09, 1.46),
                       Layer.grating(.13, [0., .12, .6], [1.1, 1.9])], 1.4, .6)


def lossy_stack():
    """The reference's own field-test stack (tests/test_rcwa_public_fields.cpp:28-33)."""
    from stablercwa_amd.types import Layer, Stack
    return Stack(1.2, [Layer.grating(.2, [0., .21, .6], [1.7, 1.]), Layer.uniform(.09, 1.46 + .03j),
                       Layer.grating(.13, [0., .12, .6], [1.1, 1.9 + .07j])], 1.4, .6)


ZS = np.array([-0.3, -1e-9, 0.0, 0.05, 0.2 - 1e-9, 0.2 + 1e-9, 0.25, 0.29 - 1e-9, 0.29 + 1e-9, 0.35, 0.42 - 1e-9,
               0.42 + 1e-9, 0.5, 2.0])


def flux(h):
    return np.sum((h[0] * np.conj(h[4]) - h[1] * np.conj(h[3])).real, axis=0)


@pytest.mark.parametrize("pol", [0, 1])
def test_flux_is_conserved_through_a_lossless_stack(ref_lib, pol):
    from oracle import ref
    o = ref.planar_fields(ref_lib, lossless_stack().flat(), .83, 27., 3, pol, ZS)
    assert o["available"]
    S = flux(o["harm"])
    assert np.max(np.abs(S / S[0] - 1.0)) <= 1e-10, S / S[0]
    # incidence side: S_inc (1 - R); exit side: S_inc T
    R, T = o["DEr"].sum(), o["DEt"].sum()
    assert abs(S[-1] - S[0] * T / (1.0 - R)) <= 1e-10 * abs(S[0])


@pytest.mark.parametrize("pol", [0, 1])
def test_flux_falls_through_absorbing_layers(ref_lib, pol):
    from oracle import ref
    zs = np.linspace(-0.1, 0.5, 61)
    o = ref.planar_fields(ref_lib, lossy_stack().flat(), .83, 27., 3, pol, zs)
    S = flux(o["harm"])
    assert np.all(np.diff(S) <= 1e-12 * abs(S[0]))
    assert S[-1] < S[0]
    R, T = o["DEr"].sum(), o["DEt"].sum()
    assert abs(S[-1] - S[0] * T / (1.0 - R)) <= 1e-10 * abs(S[0])


@pytest.mark.parametrize("pol", [0, 1])
def test_tangential_fields_are_continuous_across_interfaces(ref_lib, pol):
    from oracle import ref
    h = ref.planar_fields(ref_lib, lossy_stack().flat(), .83, 27., 3, pol,
                          [s + e for s in (0.0, 0.2, 0.29, 0.42) for e in (-1e-10, 1e-10)])["harm"]
    comps = (1, 3) if pol == 0 else (4, 0)
    for c in comps:
        above, below = h[c][:, 0::2], h[c][:, 1::2]
        assert np.max(np.abs(above - below)) <= 1e-7 * max(1.0, np.max(np.abs(above)))


@pytest.mark.parametrize("pol", [0, 1])
def test_grid_is_the_carrier_sum_of_the_harmonics(ref_lib, pol):
    from oracle import ref
    M, lam, theta, nx, y = 3, .83, 27., 11, .11
    o = ref.planar_fields(ref_lib, lossy_stack().flat(), lam, theta, M, pol, ZS, nx=nx, y=y)
    k0 = 2 * np.pi / lam
    kxm = k0 * (1.2 * np.sin(np.radians(theta)) - np.arange(-M, M + 1) * lam / .6)
    x = np.arange(nx) / nx * .6
    carrier = np.exp(-1j * np.outer(x, kxm))
    for c in range(6):
        want = (carrier @ o["harm"][c]).T
        assert np.max(np.abs(o["grid"][c] - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


def test_planted_amplitude_swap_breaks_the_flux_invariant(ref_lib):
    """Rebuild the fields inside the uniform film from its two one-sided amplitudes: the correct anchoring
    (a+ decays from the top, a- from the bottom) keeps the flux, exchanging the anchors does not."""
    from oracle import ref
    st = lossless_stack()
    lam, M, pol = .83, 3, 0
    d, z0, n = .09, .2, 1.46
    o = ref.planar_fields(ref_lib, st.flat(), lam, 27., M, pol, [z0 + 1e-12, z0 + d - 1e-12])
    k0 = 2 * np.pi / lam
    kx = 1.2 * np.sin(np.radians(27.)) - np.arange(-M, M + 1) * lam / .6
    q = np.sqrt(kx.astype(complex) ** 2 - n ** 2)
    q = np.where(q.real < 0, -q, q)
    q = np.where((q.real < 1e-10) & (q.imag > 0), -q, q)
    Ey = o["harm"][1]
    Hx_over_i = o["harm"][3] / 1j
    # S = a+ e^{-k0 z q} + a- e^{-k0 (d - z) q}, U = -a+ e^{..} + a- e^{..}, Hx = i q U: solve a+-, a- at the top
    U = Hx_over_i[:, 0] / q
    ap = (Ey[:, 0] - U) / 2 - 0j
    am_top = (Ey[:, 0] + U) / 2
    am = am_top / np.exp(-k0 * d * q)
    good = ap * np.exp(-k0 * d * q) + am
    swapped = am * np.exp(-k0 * d * q) + ap
    assert np.max(np.abs(good - Ey[:, 1])) <= 1e-9 * np.max(np.abs(Ey))
    assert np.max(np.abs(swapped - Ey[:, 1])) > 1e-3 * np.max(np.abs(Ey))


# ---- conical near fields (rcwa1d_conical_fields.hpp) -------------------------------------------------------
CONICAL_ZS = np.array([-0.3, -1e-9, 0.0, 0.1, 0.2 - 1e-10, 0.2 + 1e-10, 0.25, 0.29 - 1e-10, 0.29 + 1e-10, 0.42, 0.6])


def test_conical_flux_is_conserved_for_any_jones_vector(ref_lib):
    from oracle import ref
    for jones in [(1, 0), (0, 1), (0.3 + 0.2j, -0.7 + 0.5j)]:
        o = ref.conical_fields(ref_lib, lossless_stack().flat(), .83, 27., 31., 3, jones, CONICAL_ZS)
        assert o["available"]
        S = flux(o["harm"])
        assert np.max(np.abs(S / S[0] - 1.0)) <= 1e-10, (jones, S / S[0])


def test_conical_fields_are_linear_in_the_jones_vector(ref_lib):
    """The capture holds the s and p responses; the fields of (Es, Ep) / |(Es, Ep)| are that combination."""
    from oracle import ref
    st = lossy_stack().flat()
    hs = ref.conical_fields(ref_lib, st, .83, 27., 31., 3, (1, 0), CONICAL_ZS)["harm"]
    hp = ref.conical_fields(ref_lib, st, .83, 27., 31., 3, (0, 1), CONICAL_ZS)["harm"]
    es, ep = 0.3 + 0.2j, -0.7 + 0.5j
    nrm = np.hypot(abs(es), abs(ep))
    h = ref.conical_fields(ref_lib, st, .83, 27., 31., 3, (es, ep), CONICAL_ZS)["harm"]
    want = (es * hs + ep * hp) / nrm
    assert np.max(np.abs(h - want)) <= 1e-12 * np.max(np.abs(want))


def test_conical_tangential_fields_are_continuous(ref_lib):
    from oracle import ref
    zs = [s + e for s in (0.0, 0.2, 0.29, 0.42) for e in (-1e-10, 1e-10)]
    h = ref.conical_fields(ref_lib, lossy_stack().flat(), .83, 27., 31., 3, (0.6, 0.8j), zs)["harm"]
    for c in (0, 1, 3, 4):            # Ex, Ey, Hx, Hy
        above, below = h[c][:, 0::2], h[c][:, 1::2]
        assert np.max(np.abs(above - below)) <= 1e-7 * max(1.0, np.max(np.abs(above)))


def test_conical_grid_carries_the_ky_phase(ref_lib):
    """field_map at height y: sum_j exp(-i (kxm_j x + ky y)) h_j with ky = k0 nI sin(theta) sin(phi)."""
    from oracle import ref
    lam, th, ph, M, nx, y = .83, 27., 31., 3, 9, .37
    o = ref.conical_fields(ref_lib, lossy_stack().flat(), lam, th, ph, M, (1, 0), CONICAL_ZS, nx=nx, y=y)
    k0 = 2 * np.pi / lam
    s = 1.2 * np.sin(np.radians(th))
    kxm = k0 * (s * np.cos(np.radians(ph)) - np.arange(-M, M + 1) * lam / .6)
    ky = k0 * s * np.sin(np.radians(ph))
    x = np.arange(nx) / nx * .6
    carrier = np.exp(-1j * (np.outer(x, kxm) + ky * y))
    for c in range(6):
        want = (carrier @ o["harm"][c]).T
        assert np.max(np.abs(o["grid"][c] - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


def test_conical_schur_layer_fields_are_continuous_and_linear(ref_lib):
    """The reference's degenerate (Schur-basis) conical cases: tangential Ex, Ey, Hx, Hy continuous at every
    interface and the fields linear in the Jones vector, so the Schur-basis propagation is pinned too."""
    from conftest import load_golden
    from oracle import ref
    from stablercwa_amd.types import Stack
    for c in [c for c in load_golden("conical_ref.json")["cases"] if c["n_schur"] > 0][:4]:
        st = Stack.from_flat(np.array(c["stack"]))
        tops = np.cumsum([0.0] + [l.d for l in st.layers])
        zs = [t + e for t in tops for e in (-1e-10, 1e-10)]
        args = (ref_lib, c["stack"], c["wavelength"], c["theta"], c["phi"], c["M"])
        hs = ref.conical_fields(*args, (1, 0), zs)
        assert hs["available"] and hs["n_schur"] > 0
        hp = ref.conical_fields(*args, (0, 1), zs)["harm"]
        h = ref.conical_fields(*args, (0.6, 0.8j), zs)["harm"]
        assert np.max(np.abs(h - (0.6 * hs["harm"] + 0.8j * hp))) <= 1e-12 * np.max(np.abs(h))
        for comp in (0, 1, 3, 4):
            above, below = h[comp][:, 0::2], h[comp][:, 1::2]
            assert np.max(np.abs(above - below)) <= 1e-6 * max(1.0, np.max(np.abs(above))), (c["name"], comp)

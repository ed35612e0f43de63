"""GPU: planar near fields (SRCWA_PLAN_FIELDS; srcwa_plan1d_field_at / _field_map) against the reference's own
rcwa::solve_with_fields + field_at + field_map compiled from its headers (oracle/_ref), whose fields are pinned
by flux conservation and tangential continuity in tests/test_oracle_fields.py.

* harmonic fields (6 components x N orders x depths) and the x-z grid within 1e-9 of the reference (relative to
  max(1, |field|)) for TE and TM: the reference's field-test stack (two gratings around an absorbing film,
  tests/test_rcwa_public_fields.cpp:28-33) at its depths (in both outer media, inside each layer, on and just
  beside every interface, repeated depths), a grating on a 20-film mirror at M = 20 and M = 60 (through the
  whole 3.44-deep stack), a pure film stack, an FFT-sampled grating (TM: the [[eps]] solve of Ez), and normal
  incidence on a symmetric grating (degenerate modes)
* TE leaves Ex, Ez, Hy and TM leaves Ey, Hx, Hz exactly zero (every component of the output is written)
* depths keep their order and repeats; one problem of a multi-problem plan samples the same as alone
* flux: the GPU fields carry the same z-flux at every depth of a lossless stack
* the failure contract: a solve that fails (exit-side grazing) leaves the context unavailable; non-finite depths,
  nx <= 0 and a non-finite y are invalid arguments (rcwa_fields.hpp:253-292)
"""
import numpy as np
import pytest

from test_oracle_fields import ZS, flux, lossless_stack, lossy_stack

pytestmark = pytest.mark.gpu

TOL = 1e-9


@pytest.fixture(scope="module")
def fctx():
    from stablercwa_amd import _lib
    return _lib.Context(0, plan_flags=_lib.PLAN_FIELDS)


def mirror_stack():
    from stablercwa_amd.types import Layer, Stack
    films = [Layer.uniform(1.053 / (4 * n), n) for _ in range(10) for n in (2.1, 1.46)]
    return Stack(1.0, [Layer.grating(0.38, [0.0, 0.3 * 0.574, 0.574], [1.46, 1.0])] + films, 1.46, 0.574)


def films_stack():
    from stablercwa_amd.types import Layer, Stack
    return Stack(1.0, [Layer.uniform(0.12, 2.1), Layer.uniform(0.2, 1.46 + 0.01j), Layer.uniform(0.07, 2.1)], 1.5, 0.6)


def fft_stack():
    from stablercwa_amd.types import FourierMethod, Layer, Stack
    g = Layer.grating(0.25, [0.0, 0.2, 0.45, 0.6], [1.9, 1.0, 1.5 + 0.02j])
    g.fourier_method = FourierMethod.FFT
    return Stack(1.1, [g, Layer.uniform(0.1, 1.7)], 1.45, 0.6)


def _plan(ctx, stack, problems):
    from stablercwa_amd import rcwa
    b = rcwa.PlanarBatch([stack], [(0,) + p for p in problems], ctx)
    b.execute()
    b.sync()
    return b


def _close(a, b, what):
    a = np.asarray(a); b = np.asarray(b)
    err = np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))) if b.size else 0.0
    assert err <= TOL, (what, err)


CASES = [
    ("reference field stack", lossy_stack, .83, 27., 3, ZS),
    ("grating on a mirror, M=20", mirror_stack, 1.053, 66.36, 20, np.linspace(-0.2, 1.6, 23)),
    ("pure films", films_stack, .9, 35., 4, np.array([-0.1, 0.0, 0.06, 0.12, 0.25, 0.32, 0.39, 0.5])),
    ("FFT grating", fft_stack, .8, 12., 6, np.array([-0.05, 0.1, 0.25, 0.3, 0.35, 0.6])),
    ("grating on a mirror, M=60", mirror_stack, 1.053, 66.36, 60, np.linspace(-0.3, 3.6, 40)),
    # normal incidence on a symmetric grating: degenerate eigenvalues, so the eigenvector bases of the GPU and
    # the reference differ inside each degenerate pair; the fields are physical and must still agree
    ("normal incidence, degenerate modes", lossless_stack, .83, 0.0, 5,
     np.array([-0.2, 0.0, 0.1, 0.2, 0.2 + 5e-13, 0.25, 0.29, 0.3, 0.42, 0.42 + 5e-13, 0.7])),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("pol", [0, 1])
def test_fields_match_reference(fctx, ref_lib, case, pol):
    from oracle import ref
    name, make, lam, theta, M, zs = case
    st = make()
    b = _plan(fctx, st, [(lam, theta, M, pol)])
    assert b.field_available(0)
    o = ref.planar_fields(ref_lib, st.flat(), lam, theta, M, pol, zs, nx=11, y=.11)
    assert o["available"]
    _close(b.result(0).DEr, o["DEr"], name + " DEr")
    h = b.field_at(0, zs)
    _close(h, o["harm"], name + " harmonics")
    _close(b.field_map(0, zs, 11, .11), o["grid"], name + " grid")
    # the polarization's other three components are exactly zero, as the reference's are
    assert not np.any(h[[0, 2, 4] if pol == 0 else [1, 3, 5]])


def test_order_repeats_and_batch_independence(fctx):
    st = lossy_stack()
    zs = np.array([.2, -.12, .2, 100., .083, .42, -100., .083])
    alone = _plan(fctx, st, [(.83, 27., 3, 1)])
    many = _plan(fctx, st, [(.9, 10., 5, 0), (.83, 27., 3, 1), (.7, 40., 2, 1)])
    h = alone.field_at(0, zs)
    assert np.array_equal(h, many.field_at(1, zs))
    assert np.array_equal(h[:, :, 0], h[:, :, 2]) and np.array_equal(h[:, :, 4], h[:, :, 7])
    one = np.stack([alone.field_at(0, z)[:, :, 0] for z in zs], axis=2)
    assert np.array_equal(h, one)
    assert alone.field_at(0, []).shape == (6, 7, 0)
    assert alone.field_map(0, [], 5).shape == (6, 0, 5)


@pytest.mark.parametrize("pol", [0, 1])
def test_gpu_fields_conserve_flux(fctx, pol):
    b = _plan(fctx, lossless_stack(), [(.83, 27., 3, pol)])
    S = flux(b.field_at(0, ZS))
    assert np.max(np.abs(S / S[0] - 1.0)) <= 1e-10


def test_failed_solve_is_unavailable_and_arguments_are_checked(fctx):
    from stablercwa_amd import _lib
    from stablercwa_amd.types import Layer, Stack
    p, f = 0.5747, 0.30
    grazing = Stack(1.0, [Layer.grating(0.38, [-f * p / 2, f * p / 2, p - f * p / 2], [1.45, 1.0])], 2.0, p)
    b = _plan(fctx, grazing, [(p, 0.0, 8, 0)])
    assert not b.field_available(0)
    with pytest.raises(_lib.FieldUnavailable):
        b.field_at(0, [0.1])
    with pytest.raises(_lib.FieldUnavailable):
        b.field_at(0, [])
    ok = _plan(fctx, lossy_stack(), [(.83, 27., 3, 0)])
    for bad in ([np.inf], [0.0, np.nan]):
        with pytest.raises(ValueError):
            ok.field_at(0, bad)
    with pytest.raises(ValueError):
        ok.field_map(0, ZS, 0)
    with pytest.raises(ValueError):
        ok.field_map(0, ZS, -1)
    with pytest.raises(ValueError):
        ok.field_map(0, ZS, 3, np.inf)


def test_plans_without_the_flag_keep_no_fields(ctx):
    b = _plan(ctx, lossy_stack(), [(.83, 27., 3, 0)])
    with pytest.raises(ValueError):
        b.field_at(0, [0.1])


def test_cpp_solve_with_fields_matches_reference(ref_lib):
    """The C++ API with the reference's signatures (include/stablercwa_amd/fields.hpp) through the C++ driver."""
    from oracle import ref
    from test_gpu_cpp_api import _drive, _fields, _rec
    st = lossy_stack()
    zs = ZS
    lines = [f"fields {pol} 3 0.83 27.0 {len(zs)} " + " ".join(repr(float(z)) for z in zs) + f" 11 0.11 {_rec(st.flat())}"
             for pol in (0, 1)]
    for pol, block in zip((0, 1), _drive(lines)):
        f = _fields(block)
        assert f["head"][0] == "ok" and f["head"][3] == "1", block[0]
        o = ref.planar_fields(ref_lib, st.flat(), .83, 27., 3, pol, zs, nx=11, y=.11)
        harm = f["harm"].view(np.complex128).reshape(6, len(zs), 7).transpose(0, 2, 1)
        grid = f["grid"].view(np.complex128).reshape(6, len(zs), 11)
        _close(harm, o["harm"], "C++ harmonics")
        _close(grid, o["grid"], "C++ grid")
        assert abs(f["z_bottom"][0] - 0.42) <= 1e-15
    # a failed solve: the context is unavailable and field_at throws std::logic_error
    from stablercwa_amd.types import Layer, Stack
    p, fl = 0.5747, 0.30
    grazing = Stack(1.0, [Layer.grating(0.38, [-fl * p / 2, fl * p / 2, p - fl * p / 2], [1.45, 1.0])], 2.0, p)
    block, = _drive([f"fields 0 8 {p!r} 0.0 1 0.1 5 0.0 {_rec(grazing.flat())}"])
    assert block[0].split()[3] == "0"
    assert any(ln.startswith("logic_error rcwa: planar field context is unavailable") for ln in block)


def test_python_solve_with_fields(ref_lib):
    from oracle import ref
    from stablercwa_amd import _lib, rcwa
    st = lossy_stack()
    r, f = rcwa.solve_with_fields(st, .83, 27., 3, rcwa.Polarization.TM)
    assert f.available()
    o = ref.planar_fields(ref_lib, st.flat(), .83, 27., 3, 1, ZS, nx=9, y=0.0)
    _close(r.DEt, o["DEt"], "DEt")
    _close(f.field_at(ZS), o["harm"], "harmonics")
    _close(f.field_map(ZS, 9), o["grid"], "grid")
    from stablercwa_amd.types import Layer, Stack
    p, fl = 0.5747, 0.30
    grazing = Stack(1.0, [Layer.grating(0.38, [-fl * p / 2, fl * p / 2, p - fl * p / 2], [1.45, 1.0])], 2.0, p)
    r2, f2 = rcwa.solve_with_fields(grazing, p, 0.0, 8)
    assert not f2.available() and np.all(np.isnan(r2.DEr))
    with pytest.raises(_lib.FieldUnavailable):
        f2.field_at([0.0])


def test_field_plans_must_fit_one_chunk_and_be_planar_or_conical(fctx):
    """The samplers read the modes where the solve left them: a field plan is refused when its workspace would
    be split into chunks; crossed-grating plans capture no fields in this build."""
    from stablercwa_amd import rcwa
    st = mirror_stack()
    probs = [(0, 1.0 + 0.01 * i, 66.36, 20, 0) for i in range(4)]
    with pytest.raises(ValueError, match="one workspace chunk"):
        rcwa.PlanarBatch([st], probs, fctx, max_workspace_bytes=1 << 20)
    with pytest.raises(ValueError, match="one workspace chunk"):
        rcwa.ConicalBatch([st], [(0, 1.0 + 0.01 * i, 40., 30., 20) for i in range(4)], fctx, max_workspace_bytes=1 << 20)
    from stablercwa_amd.types import Layer2D, Stack2D
    with pytest.raises(NotImplementedError):
        rcwa.CrossedBatch([Stack2D(nI=1.0, nII=1.45, layers=[Layer2D.uniform(0.1, 1.5)], Lx=1.0, Ly=1.0)],
                          [(0, 1.3, 10., 20., 1, 1, (1.0, 0.0))], fctx)


def test_fields_on_a_pooled_context(ref_lib):
    """A context with a device pool: the field plan and the samplers' buffers come from the pool, and the pool
    gets every block back once the plan is destroyed."""
    import ctypes
    from oracle import ref
    from stablercwa_amd import _lib, rcwa
    lib = _lib.load()
    pool = ctypes.c_void_p()
    _lib.check(lib.srcwa_pool_create(0, 256 << 20, ctypes.byref(pool)))
    ctx = _lib.Context(0, plan_flags=_lib.PLAN_FIELDS)
    try:
        _lib.check(lib.srcwa_context_set_pool(ctx.handle, pool))
        free0 = lib.srcwa_pool_free_bytes(pool)
        st = lossy_stack()
        b = _plan(ctx, st, [(.83, 27., 3, 1)])
        o = ref.planar_fields(ref_lib, st.flat(), .83, 27., 3, 1, ZS, nx=7)
        for _ in range(3):
            _close(b.field_at(0, ZS), o["harm"], "pooled harmonics")
            _close(b.field_map(0, ZS, 7), o["grid"], "pooled grid")
        b.close()
        assert lib.srcwa_pool_free_bytes(pool) == free0
    finally:
        ctx.close()
        lib.srcwa_pool_destroy(pool)


@pytest.mark.parametrize("extra", ["diagnostics", "point_ops"])
def test_fields_with_other_plan_flags(ref_lib, extra):
    """Near fields combine with the other plan flags: diagnostics (per-task operators, kept layer matrices) and
    per-point operators give the same fields as the reference."""
    from oracle import ref
    from stablercwa_amd import _lib
    flag = _lib.PLAN_DIAGNOSTICS if extra == "diagnostics" else _lib.PLAN_POINT_OPS
    ctx = _lib.Context(0, plan_flags=_lib.PLAN_FIELDS | flag)
    st = lossy_stack()
    for pol in (0, 1):
        b = _plan(ctx, st, [(.83, 27., 3, pol), (.9, 10., 3, pol)])
        for i, (lam, th) in enumerate([(.83, 27.), (.9, 10.)]):
            o = ref.planar_fields(ref_lib, st.flat(), lam, th, 3, pol, ZS)
            _close(b.field_at(i, ZS), o["harm"], f"{extra} pol {pol} point {i}")
        b.close()
    ctx.close()


def test_fields_edge_stacks(fctx, ref_lib):
    """No layers at all (z = 0 belongs to the exit side), a zero-thickness layer (its interface is owned by the
    layer above, rcwa_fields.hpp:253-279), and many depths in one group."""
    from oracle import ref
    from stablercwa_amd.types import Layer, Stack
    bare = Stack(1.3, [], 1.6, 0.7)
    zero = Stack(1.0, [Layer.grating(0.2, [0.0, 0.3, 0.6], [1.8, 1.2]), Layer.uniform(0.0, 2.2),
                       Layer.grating(0.0, [0.0, 0.1, 0.6], [1.4, 2.0]), Layer.uniform(0.15, 1.5)], 1.45, 0.6)
    cases = [(bare, np.array([-0.5, -1e-12, 0.0, 1e-12, 0.4])),
             (zero, np.array([-0.1, 0.0, 0.1, 0.2, 0.2 + 1e-13, 0.25, 0.35, 0.5])),
             (zero, np.linspace(0.2, 0.35, 3000))]
    for st, zs in cases:
        for pol in (0, 1):
            b = _plan(fctx, st, [(.9, 20., 4, pol)])
            o = ref.planar_fields(ref_lib, st.flat(), .9, 20., 4, pol, zs, nx=5, y=.3)
            assert b.field_available(0) == o["available"] == True
            _close(b.field_at(0, zs), o["harm"], f"edge harmonics L={len(st.layers)} nz={len(zs)}")
            _close(b.field_map(0, zs, 5, .3), o["grid"], "edge grid")
            b.close()


# ---- conical near fields (srcwa_plan_conical_field_at / _field_map) ----------------------------------------
CONICAL_CASES = [
    ("reference field stack", lossy_stack, .83, 27., 31., 3, ZS),
    ("grating on a mirror, M=20", mirror_stack, 1.053, 40., 30., 20, np.linspace(-0.2, 3.6, 25)),
    ("pure films", films_stack, .9, 35., 20., 4, np.array([-0.1, 0.0, 0.06, 0.12, 0.25, 0.32, 0.39, 0.5])),
    ("FFT grating", fft_stack, .8, 12., 50., 6, np.array([-0.05, 0.1, 0.25, 0.3, 0.35, 0.6])),
]
JONES = [(1.0, 0.0), (0.0, 1.0), (0.3 + 0.2j, -0.7 + 0.5j)]


def _conical_plan(ctx, stack, problems):
    from stablercwa_amd import rcwa
    b = rcwa.ConicalBatch([stack], [(0,) + p for p in problems], ctx)
    b.execute()
    b.sync()
    return b


@pytest.mark.parametrize("case", CONICAL_CASES, ids=[c[0] for c in CONICAL_CASES])
def test_conical_fields_match_reference(fctx, ref_lib, case):
    from oracle import ref
    name, make, lam, theta, phi, M, zs = case
    st = make()
    b = _conical_plan(fctx, st, [(lam, theta, phi, M)])
    assert b.field_available(0)
    for jones in JONES:
        o = ref.conical_fields(ref_lib, st.flat(), lam, theta, phi, M, jones, zs, nx=9, y=.37)
        assert o["available"] and o["n_schur"] == 0
        _close(b.field_at(0, zs, jones), o["harm"], f"conical {name} {jones} harmonics")
        _close(b.field_map(0, zs, 9, .37, jones), o["grid"], f"conical {name} {jones} grid")
    b.close()


def test_conical_fields_zero_thickness_layers_and_batches(fctx, ref_lib):
    """Zero-thickness layers hold no fields (skipped by the depth groups); one problem of a batch samples the
    same as alone; the GPU fields conserve flux through a lossless stack."""
    from oracle import ref
    from stablercwa_amd.types import Layer, Stack
    st = Stack(1.0, [Layer.grating(0.2, [0.0, 0.3, 0.6], [1.8, 1.2]), Layer.uniform(0.0, 2.2),
                     Layer.grating(0.0, [0.0, 0.1, 0.6], [1.4, 2.0]), Layer.uniform(0.15, 1.5)], 1.45, 0.6)
    zs = np.array([-0.1, 0.0, 0.1, 0.2, 0.2 + 1e-13, 0.25, 0.35, 0.5])
    alone = _conical_plan(fctx, st, [(.9, 20., 35., 4)])
    many = _conical_plan(fctx, st, [(.8, 10., 5., 3), (.9, 20., 35., 4)])
    o = ref.conical_fields(ref_lib, st.flat(), .9, 20., 35., 4, (0.6, 0.8j), zs)
    h = alone.field_at(0, zs, (0.6, 0.8j))
    _close(h, o["harm"], "conical zero-thickness layers")
    assert np.array_equal(h, many.field_at(1, zs, (0.6, 0.8j)))
    b = _conical_plan(fctx, lossless_stack(), [(.83, 27., 31., 3)])
    S = flux(b.field_at(0, ZS, (0.3 + 0.2j, -0.7 + 0.5j)))
    assert np.max(np.abs(S / S[0] - 1.0)) <= 1e-10


def test_conical_fields_through_schur_layers(fctx, ref_lib):
    """Degenerate layers propagate in the Schur basis: the field plan keeps their generator G and samples
    exp(-k0 z G) per depth; the reference's cases with Schur layers (tests/golden/conical_ref.json) agree."""
    from conftest import load_golden
    from oracle import ref
    from stablercwa_amd.types import Stack
    cases = [c for c in load_golden("conical_ref.json")["cases"] if c["n_schur"] > 0][:4]
    assert cases
    for c in cases:
        st = Stack.from_flat(np.array(c["stack"]))
        total = sum(l.d for l in st.layers)
        zs = np.concatenate([[-0.2, 0.0], np.linspace(0.0, total, 9)[1:], [total + 0.3]])
        b = _conical_plan(fctx, st, [(c["wavelength"], c["theta"], c["phi"], c["M"])])
        assert b.results()["status"][0] == 0 and b.field_available(0), c["name"]
        for jones in JONES[1:]:
            o = ref.conical_fields(ref_lib, c["stack"], c["wavelength"], c["theta"], c["phi"], c["M"], jones, zs,
                                   nx=5, y=.2)
            assert o["available"] and o["n_schur"] > 0, c["name"]
            _close(b.field_at(0, zs, jones), o["harm"], f"schur {c['name']} harmonics")
            _close(b.field_map(0, zs, 5, .2, jones), o["grid"], f"schur {c['name']} grid")
        b.close()


def test_conical_field_contract(fctx):
    """Jones vectors that are zero or not finite and non-finite depths are invalid arguments; the planar entry
    points refuse a conical plan; a failed conical solve leaves the context unavailable."""
    from stablercwa_amd import _lib
    from stablercwa_amd.types import Layer, Stack
    ok = _conical_plan(fctx, lossy_stack(), [(.83, 27., 31., 3)])
    for bad in [(0.0, 0.0), (np.inf, 1.0), (1.0, complex(0, np.nan))]:
        with pytest.raises(ValueError):
            ok.field_at(0, [0.1], bad)
    with pytest.raises(ValueError):
        ok.field_at(0, [np.nan])
    with pytest.raises(ValueError):
        ok.field_map(0, ZS, 0)
    out = np.zeros(12 * 7)
    z = np.array([0.1])
    with pytest.raises(ValueError):
        _lib.check(ok.lib.srcwa_plan1d_field_at(ok.handle, 0, 1, _lib.ptr(z), _lib.ptr(out)))
    # an eigensolve stopped after one sweep: eig_info > 0, NaN results, no fields
    st = Stack(1.0, [Layer.grating(0.38, [0.0, 0.1722, 0.574], [1.46, 1.0])], 1.46, 0.574)
    b = rcwa_conical_batch(fctx, st, (1.0, 30.0, 20.0, 6))
    b.set_max_iter(1)
    b.execute()
    b.sync()
    assert b.results()["eig_info"][0] > 0 and not b.field_available(0)
    with pytest.raises(_lib.FieldUnavailable):
        b.field_at(0, [0.1])


def rcwa_conical_batch(ctx, stack, point):
    from stablercwa_amd import rcwa
    return rcwa.ConicalBatch([stack], [(0,) + tuple(point)], ctx)


def test_cpp_solve_conical_with_fields_matches_reference(ref_lib):
    from oracle import ref
    from test_gpu_cpp_api import _drive, _fields, _rec
    st = lossy_stack()
    es, ep = 0.3 + 0.2j, -0.7 + 0.5j
    line = (f"conical_fields 3 0.83 27.0 31.0 {es.real!r} {es.imag!r} {ep.real!r} {ep.imag!r} {len(ZS)} "
            + " ".join(repr(float(z)) for z in ZS) + f" 9 0.37 {_rec(st.flat())}")
    block, = _drive([line])
    f = _fields(block)
    assert f["head"][0] == "ok" and f["head"][3] == "1", block[0]
    o = ref.conical_fields(ref_lib, st.flat(), .83, 27., 31., 3, (es, ep), ZS, nx=9, y=.37)
    _close(f["harm"].view(np.complex128).reshape(6, len(ZS), 7).transpose(0, 2, 1), o["harm"], "C++ conical harmonics")
    _close(f["grid"].view(np.complex128).reshape(6, len(ZS), 9), o["grid"], "C++ conical grid")


def test_cpp_conical_metadata_marks_schur_layers(ref_lib):
    """FieldLayerInfo::propagation is Schur for the layers the solve propagated in the Schur basis, as many as
    the reference's n_schur."""
    from conftest import load_golden
    from oracle import ref
    from test_gpu_cpp_api import _drive, _rec
    c = [c for c in load_golden("conical_ref.json")["cases"] if c["n_schur"] > 0][0]
    line = (f"conical_fields {c['M']} {c['wavelength']!r} {c['theta']!r} {c['phi']!r} 1.0 0.0 0.0 0.0 1 0.01 3 0.0 "
            f"{_rec(c['stack'])}")
    block, = _drive([line])
    assert block[0].split()[3] == "1", block[0]
    marks = [int(x) for ln in block if ln.startswith("schur_layers") for x in ln.split()[1:]]
    o = ref.conical_fields(ref_lib, c["stack"], c["wavelength"], c["theta"], c["phi"], c["M"], (1, 0), [0.01])
    assert sum(marks) == o["n_schur"] > 0, (marks, o["n_schur"])


@pytest.mark.parametrize("pol", [0, 1])
def test_fields_through_a_grazing_film_order(fctx, ref_lib, pol):
    """Normal incidence with lambda / period = 1.5 exactly and a film of index 1.5: orders +-1 have q = 0 exactly in
    the film, which then propagates in the linear finite-zero basis (rcwa1d_fields.hpp:124-131); the metadata
    counts those two modes."""
    import ctypes
    from oracle import ref
    from stablercwa_amd import _lib
    from stablercwa_amd.types import Layer, Stack
    st = Stack(1.0, [Layer.grating(0.2, [0.0, 0.2, 0.5], [1.9, 1.0]), Layer.uniform(0.1, 1.5),
                     Layer.uniform(0.05, 2.0)], 1.45, 0.5)
    zs = np.array([-0.1, 0.1, 0.2, 0.23, 0.27, 0.3, 0.33, 0.4])
    b = _plan(fctx, st, [(0.75, 0.0, 3, pol)])
    o = ref.planar_fields(ref_lib, st.flat(), 0.75, 0.0, 3, pol, zs, nx=5)
    assert b.field_available(0) and o["available"]
    _close(b.field_at(0, zs), o["harm"], "finite-zero film harmonics")
    _close(b.field_map(0, zs, 5), o["grid"], "finite-zero film grid")
    zeros = np.zeros(3, dtype=np.int32)
    _lib.check(b.lib.srcwa_plan1d_field_layers(b.handle, 0, _lib.ptr(zeros)))
    assert zeros.tolist() == [0, 2, 0]


def test_python_solve_conical_with_fields(ref_lib):
    from oracle import ref
    from stablercwa_amd import rcwa
    st = lossy_stack()
    r, f = rcwa.solve_conical_with_fields(st, .83, 27., 31., 3)
    assert f.available() and len(r.DEr) == 7
    o = ref.conical_fields(ref_lib, st.flat(), .83, 27., 31., 3, (0.6, 0.8j), ZS, nx=9, y=.2)
    _close(f.field_at(ZS, (0.6, 0.8j)), o["harm"], "python conical harmonics")
    _close(f.field_map(ZS, 9, .2, (0.6, 0.8j)), o["grid"], "python conical grid")

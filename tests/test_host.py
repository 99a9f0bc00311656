"""CPU tests of the boundary and the host logic (no kernel launches).

- libmivegas.so loads, has the ABI version and struct layouts of the
  ctypes binding, and exports every function include/mivegas.h declares;
  argument errors come back as status codes before any device call;
- the integrator classes validate their arguments like the reference
  (test_misc.py:85-97), compute the VEGAS+ stratification, write and read
  the grid JSON checkpoint (vflow.py:271-347), and check integrand shapes
  at compile time (monte_carlo.py:602-633);
- configuration flags (test_config.py) and the mask helper (test_utils.py).
"""
import ctypes
import importlib
import json
import math
import os
import re
import tempfile

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="session")
def lib():
    from vegasflow_amd import _abi, _build

    if not os.path.exists(_abi.lib_path()):
        _build.build()
    return _abi.load()


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mivegas.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mv_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name
    from vegasflow_amd import _abi

    assert sorted(_abi.SIGNATURES) == names


def test_abi_version_and_struct_layouts(lib):
    from vegasflow_amd import _abi

    assert lib.mv_abi_version() == _abi.ABI_VERSION
    for i, st in enumerate(_abi.STRUCTS):
        assert lib.mv_struct_size(i) == ctypes.sizeof(st), st.__name__
    assert lib.mv_struct_size(99) == -1
    assert lib.mv_partial_stride(8, 1, 1) == 2 + 400
    assert lib.mv_partial_stride(8, 3, 0) == 6
    assert lib.mv_producer_blocks(1) == 1
    assert lib.mv_producer_blocks(10**9) == _abi.MAX_BLOCKS
    assert lib.mv_producer_rows(10**9) == _abi.ROWS


def test_argument_errors_are_reported_without_device_calls(lib):
    from vegasflow_amd import _abi, kernels

    rc = lib.mv_vegas_fused(None, None, None, None, None)
    assert rc == _abi.ERR_ARG
    assert "step is NULL" in _abi.last_error()
    step = kernels.make_step(17, 100)
    fn = _abi.Integrand()
    rc = lib.mv_vegas_fused(ctypes.byref(step), ctypes.byref(fn), None, None, None)
    assert rc == _abi.ERR_UNSUPPORTED and "exceeds 16" in _abi.last_error()
    step = kernels.make_step(4, 100, precision="f64")
    fn.family = 42
    rc = lib.mv_vegas_fused(ctypes.byref(step), ctypes.byref(fn), None, None, None)
    assert rc == _abi.ERR_UNSUPPORTED and "family" in _abi.last_error()
    fz = _abi.Finalize()
    fz.n_dim, fz.n_out, fz.n_events = 4, 1, 100.0
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: the call is rejected first
    rc = lib.mv_vegas_finalize(ctypes.byref(fz), fake, 202, _abi.ROWS + 1, None, fake, None, None)
    assert rc == _abi.ERR_ARG and "n_rows" in _abi.last_error()
    with pytest.raises(ValueError):
        _abi.check(_abi.ERR_ARG)
    with pytest.raises(NotImplementedError):
        _abi.check(_abi.ERR_UNSUPPORTED)
    with pytest.raises(_abi.MivegasError):
        _abi.check(_abi.ERR_LAUNCH)


# ------------------------------------------------------------ host logic
def test_integration_limits_checks():
    from vegasflow_amd import PlainFlow

    with pytest.raises(ValueError):
        PlainFlow(1, 10, xmin=[10], xmax=[1])
    with pytest.raises(ValueError):
        PlainFlow(1, 10, xmin=[10])
    with pytest.raises(ValueError):
        PlainFlow(1, 10, xmax=[10])
    with pytest.raises(ValueError):
        PlainFlow(2, 10, xmin=[0], xmax=[1])
    with pytest.raises(ValueError):
        PlainFlow(2, 10, xmin=[0, 1], xmax=[1])


def test_vplus_stratification_matches_table():
    from vegasflow_amd import VegasFlowPlus

    table = json.load(open(os.path.join(ROOT, "tests", "golden", "vplus_strata.json")))["cases"]
    for row in table:
        if row["n_cubes"] > 10000:
            continue
        inst = VegasFlowPlus(row["n_dim"], row["n_events"], adaptive=row["adaptive"], verbose=False)
        assert inst._n_strat == row["n_strat"]
        assert inst._n_cubes == row["n_cubes"]
        assert inst.min_neval_hcube == row["n_ev"]
        assert inst.n_events == row["total"]
        assert tuple(inst._hypercubes.shape) == (row["n_cubes"], row["n_dim"])


def test_vplus_events_limit_and_adaptive_rules():
    from vegasflow_amd import VegasFlowPlus

    inst = VegasFlowPlus(3, 1000, events_limit=10)
    assert inst._events_limit == 1000
    inst = VegasFlowPlus(14, 100000, adaptive=True)
    assert inst._adaptive is False
    inst = VegasFlowPlus(4, 1_000_000, adaptive=True)
    assert inst._events_bound() >= inst._n_cubes * inst.min_neval_hcube + 500_000


def test_strata_division_by_fma_corrections_is_correctly_rounded():
    """The fused VEGAS+ kernel divides t = (h + u) * 50 by n_strat with
    y = 1/n_strat and two FMA residual corrections (div_by_strat in
    vplus.hpp). FMA is emulated exactly here (a rational sum rounded once),
    and the quotient must equal the IEEE division the reference performs
    (vflowplus.py:72) on draws, the technical cut, and points within a few
    ulp of bin edges, where a wrong last bit would change the bin."""
    from fractions import Fraction

    def fma(a, b, c):
        return float(Fraction(a) * Fraction(b) + Fraction(c))

    def div5(t, n, y):
        q = t * y
        r = fma(-q, n, t)
        q = fma(r, y, q)
        r = fma(-q, n, t)
        return fma(r, y, q)

    rng = np.random.default_rng(72)
    checked = 0
    for n in list(range(1, 33)) + [47, 64, 99, 100, 127]:
        fn, y = float(n), 1.0 / n
        for k in range(120):
            h = int(rng.integers(n))
            if k % 4 == 0:
                u = float(rng.integers(1 << 53)) * 2.0**-53
            elif k % 4 == 1:
                u = float(rng.integers(1 << 32)) * 2.0**-32
            elif k % 4 == 2:
                u = 1e-8 + (1.0 - 2e-8) * float(rng.random())
            else:  # t within a few ulp of an edge m * n / 50
                m = int(rng.integers(50 * n))
                u = m / 50.0 - h + int(rng.integers(-1000, 1001)) * 2.0**-52
                if not 0.0 <= u < 1.0:
                    u = 0.5
            t = (h + u) * 50.0
            q = div5(t, fn, y)
            assert q == t / fn, (n, h, u)
            assert int(q) == int(t / fn)
            checked += 1
    assert checked == 37 * 120


def test_grid_json_roundtrip_and_errors():
    """test_algs.py:102-148 on host data."""
    from vegasflow_amd import VegasFlow

    inst = VegasFlow(2, 1000)
    grid_shape = tuple(inst.divisions.shape)
    assert grid_shape == (2, 51)
    assert inst.divisions.is_contiguous()
    np.testing.assert_array_equal(inst.divisions.cpu().numpy()[0], np.linspace(0, 1, 51))
    tmp = tempfile.mktemp()
    inst.save_grid(tmp)
    doc = json.load(open(tmp))
    assert doc["dimensions"] == 2 and doc["BINS"] == 51 and doc["ALPHA"] == 1.5
    assert doc["integrand"] == ""
    tmp_grid = np.random.default_rng(0).random(grid_shape)
    with open(tmp, "w") as f:
        json.dump({"grid": tmp_grid.tolist()}, f)
    inst.load_grid(file_name=tmp)
    np.testing.assert_equal(inst.divisions.cpu().numpy(), tmp_grid)
    tmp_grid = np.random.default_rng(1).random(grid_shape)
    inst.load_grid(numpy_grid=tmp_grid)
    np.testing.assert_equal(inst.divisions.cpu().numpy(), tmp_grid)
    for bad in ({"grid": tmp_grid.tolist(), "BINS": 0},
                {"grid": tmp_grid.tolist(), "BINS": 51, "dimensions": -4}):
        with open(tmp, "w") as f:
            json.dump(bad, f)
        with pytest.raises(ValueError):
            inst.load_grid(file_name=tmp)
    with pytest.raises(ValueError):
        inst.load_grid(file_name=tmp, numpy_grid=tmp_grid)
    with pytest.raises(ValueError):
        inst.load_grid()
    os.remove(tmp)


def test_compile_checks_integrand_shapes_on_host():
    from vegasflow_amd import PlainFlow, VegasFlow, VegasFlowPlus, Gaussian

    def wrong(x):
        return torch.sum(x)

    def wrong_vec(x):
        return x.T

    def vec(x, weight=None):
        return torch.exp(-torch.square((x - 1.0) ** 2)) / 0.845

    for f in (wrong, wrong_vec):
        with pytest.raises(ValueError):
            PlainFlow(2, 100).compile(f)
    with pytest.raises(NotImplementedError):
        VegasFlowPlus(2, 100).compile(vec)
    with pytest.raises(ValueError):
        VegasFlow(3, 100, main_dimension=5).compile(vec)
    inst = VegasFlow(3, 100).compile(vec)
    assert inst._vectorial and inst._fused is None
    inst = VegasFlow(4, 100).compile(Gaussian())
    assert inst._fused is not None and inst._fused.family == 1
    assert VegasFlow(17, 100).compile(Gaussian())._fused is None


def test_signature_discovery():
    from vegasflow_amd.monte_carlo import _argnames
    from vegasflow_amd import Gaussian

    def f0(x, weight=None):
        return x

    def f1(xarr, n_dim=None, weight=None):
        return xarr

    assert _argnames(f0) == ["weight"]
    assert _argnames(f1) == ["n_dim", "weight"]
    assert _argnames(Gaussian()) == ["weight"]


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU failure")
def test_run_without_gpu_fails_loudly():
    from vegasflow_amd import VegasFlow, Gaussian

    inst = VegasFlow(2, 100).compile(Gaussian())
    with pytest.raises(RuntimeError):
        inst.run_integration(1)


def test_builtin_exact_integrals():
    from vegasflow_amd import BreitWignerPower, Gaussian, GaussianMix, Product

    assert abs(Gaussian(0.1).exact(4) - 0.99999999999385025) < 1e-15
    assert abs(Gaussian(0.1).exact(8) - 0.99999999998770051) < 1e-15
    assert Product().exact(8) == 2.0 ** -8
    assert abs(GaussianMix().exact(4) - 0.9999969422972157) < 1e-13
    bw = BreitWignerPower(0.5, 0.02, -0.7)
    assert abs(bw.exact(6) - 3.2908976726601815) < 1e-12


def test_builtin_structs():
    from vegasflow_amd import Gaussian, GaussianMix

    s = Gaussian(0.2, [0.1, 0.2, 0.3]).to_struct(3)
    assert s.comp[0][1] == 5.0 and abs(s.comp[0][2] - (1 / (0.2 * math.sqrt(math.pi))) ** 3) < 1e-14
    assert list(s.centre[0])[:3] == [0.1, 0.2, 0.3]
    with pytest.raises(ValueError):
        Gaussian(0.2, [0.1, 0.2]).to_struct(3)
    with pytest.raises(ValueError):
        GaussianMix((1,) * 5, (0.1,) * 5, (0.5,) * 5)


# --------------------------------------------------------------- config
def test_float_and_int_env():
    import vegasflow_amd.configflow as cf

    try:
        os.environ["VEGASFLOW_FLOAT"] = "32"
        importlib.reload(cf)
        assert cf.DTYPE == torch.float32 and cf.precision_of(cf.DTYPE) == "mixed"
        os.environ["VEGASFLOW_INT"] = "64"
        importlib.reload(cf)
        assert cf.DTYPEINT == torch.int64
    finally:
        os.environ["VEGASFLOW_FLOAT"] = "64"
        os.environ["VEGASFLOW_INT"] = "32"
        importlib.reload(cf)
    assert cf.DTYPE == torch.float64 and cf.DTYPEINT == torch.int32
    assert cf.float_me(4.0).dtype == cf.DTYPE
    assert cf.int_me(4).dtype == cf.DTYPEINT


def test_generate_condition_function():
    from vegasflow_amd.utils import generate_condition_function

    rng = np.random.default_rng(0)
    masks, vals = 4, 15
    np_masks = rng.integers(2, size=(masks, vals)).astype(bool)
    t_masks = [torch.tensor(m) for m in np_masks]
    f_and = generate_condition_function(masks, "and")
    f_or = generate_condition_function(masks, "or")
    a, ia = f_and(*t_masks)
    o, io = f_or(*t_masks)
    np.testing.assert_array_equal(a.numpy(), np.all(np_masks, axis=0))
    np.testing.assert_array_equal(o.numpy(), np.any(np_masks, axis=0))
    np.testing.assert_array_equal(ia.numpy(), np.array(np.all(np_masks, axis=0).nonzero()).T)
    f_comb = generate_condition_function(3, ["and", "or"])
    c, _ = f_comb(*t_masks[:3])
    np.testing.assert_array_equal(c.numpy(), np_masks[0] & np_masks[1] | np_masks[2])
    for args in ((1, "and"), (5, "bad"), (5, ["or", "and"]), (3, ["or", "bad"])):
        with pytest.raises(ValueError):
            generate_condition_function(*args)


def test_chi2_per_dof_closed_form():
    from vegasflow_amd.monte_carlo import chi2_per_dof

    res = [(1.0, 0.1), (1.2, 0.2), (0.9, 0.1)]
    w = np.array([100.0, 25.0, 100.0])
    r = np.array([1.0, 1.2, 0.9])
    avg = np.sum(w * r) / np.sum(w)
    want = np.sum(w * (r - avg) ** 2) / 2
    assert abs(chi2_per_dof(res, avg) - want) < 1e-14
    assert np.isnan(chi2_per_dof(res[:1], 1.0))
    vec = [(np.array([1.0, 2.0]), np.array([0.1, 0.1])), (np.array([1.1, 2.0]), np.array([0.1, 0.2]))]
    out = chi2_per_dof(vec, np.array([1.05, 2.0]))
    np.testing.assert_allclose(out, [0.5, 0.0])


def test_issue_cost_probe_library_loads():
    """libmvprobe.so (the issue-cost measurement bench.py reports) is built
    beside the main library and names one result per bench label; argument
    errors return before any device call."""
    import bench
    from vegasflow_amd import _build

    path = _build.probe_path()
    if not os.path.exists(path):
        pytest.skip("probe library not built (run __graft_entry__.build())")
    probe = ctypes.CDLL(path)
    assert probe.mvp_kinds() == len(bench.ISSUE_KINDS)
    d = ctypes.c_double()
    assert probe.mvp_issue_cost(-1, 4096, ctypes.byref(d), ctypes.byref(d), ctypes.byref(d)) == -1
    assert probe.mvp_issue_cost(probe.mvp_kinds(), 4096, ctypes.byref(d), ctypes.byref(d), ctypes.byref(d)) == -1
    assert probe.mvp_lds_kinds() == len(bench.LDS_KINDS)
    assert probe.mvp_lds_cost(-1, 4096, ctypes.byref(d), ctypes.byref(d), ctypes.byref(d)) == -1
    assert probe.mvp_lds_cost(probe.mvp_lds_kinds(), 4096, ctypes.byref(d), ctypes.byref(d), ctypes.byref(d)) == -1

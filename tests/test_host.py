"""Host-side checks that need no GPU: the C ABI library exports every entry
point declared in include/binder_c.h, the product's Philox4x32-10 matches
the Random123 known-answer vectors, bench.py refuses multi-GPU runs, the
synthetic inputs are deterministic, and the binning helpers follow
BINDER/utils.py:12-82."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "binder_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(binder_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    lib_path = os.path.join(ROOT, "binder_amd", "libbinder_hip.so")
    if not os.path.exists(lib_path):
        pytest.skip("libbinder_hip.so not built")
    lib = ctypes.CDLL(lib_path)
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from binder_amd import _core
    assert set(_core.exported_symbols()) <= set(names)
    assert lib.binder_backend and _core.backend == "hip"


def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32_10."""
    from binder_amd import _core
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, exp in kat:
        assert tuple(int(v) for v in _core.philox4x32_10(ctr, key)) == exp


def test_bench_refuses_multiple_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_synthetic_pair_is_deterministic():
    from binder_amd import synth
    a = synth.make_pair((20, 18, 16), 4, 2.0, seed=9, semi=(3, 7))
    b = synth.make_pair((20, 18, 16), 4, 2.0, seed=9, semi=(3, 7))
    assert np.array_equal(a.binned_fixed, b.binned_fixed)
    assert np.array_equal(a.binned_moving, b.binned_moving)
    d = a.displacement
    assert np.isclose(np.sqrt(np.mean(np.sum(d ** 2, -1))), 2.0)
    assert a.binned_fixed.min() >= 0 and a.binned_fixed.max() <= 31


def test_bin_signal_matches_definition():
    from binder_amd.utils import bin_signal, bin_signal_robust
    x = np.linspace(0.0, 1.0, 101)
    # (signal - min) / ((max - min) / (bins - 1)), rounded (BINDER/utils.py:12-16)
    assert np.array_equal(bin_signal(x, 11), np.round(x / 0.1).astype(np.int32))
    assert bin_signal(x, 11).min() == 0 and bin_signal(x, 11).max() == 10
    rng = np.random.default_rng(0)
    s = np.concatenate([rng.uniform(0, 1, 10000), [-50.0, 50.0]])
    b, edges = bin_signal_robust(s, 32)
    assert b.dtype == np.int32 and b.min() == 0 and b.max() == 31
    assert b[-2] == 0 and b[-1] == 31  # outliers clipped to the end bins
    assert len(edges) == 33
    hist, centers = np.histogram(s, 1000)
    cdf = np.cumsum(hist / hist.sum())
    lo = centers[np.where(cdf <= 0.01)[0][-1]]
    hi = centers[np.where(cdf <= 0.99)[0][-1]]
    inside = (s >= lo) & (s <= hi)
    assert np.array_equal(b[inside], np.floor((s[inside] - lo) / ((hi - lo) / 31)).astype(np.int32))


def test_dct_eigenvalues_closed_form():
    from binder_amd.utils import initialize_fourier_domain_smoothing
    D = initialize_fourier_domain_smoothing([4, 5, 6], True, 2, 3, voxel_size=np.array([1., 2., 0.5]))
    k = np.meshgrid(np.arange(4), np.arange(5), np.arange(6), indexing="ij")
    exp = ((2 - 2 * np.cos(np.pi * k[0] / 4)) / 1 + (2 - 2 * np.cos(np.pi * k[1] / 5)) / 4 +
           (2 - 2 * np.cos(np.pi * k[2] / 6)) / 0.25) ** 2
    assert np.allclose(D, exp)
    assert D[0, 0, 0] == 0.0


@pytest.mark.parametrize("shape,nd,vs,order", [
    ((160, 192, 160), 3, (1.0, 1.0, 1.0), 2),
    ((20, 24, 20), 3, (8.0, 8.0, 8.0), 2),
    ((40, 33, 7), 3, (1.2, 0.8, 2.0), 3),
    ((192, 160), 2, (1.0, 1.0, 1.0), 2),
    ((17, 9), 2, (2.0, 3.0, 1.0), 1),
    ((13,), 1, (1.5, 1.0, 1.0), 2)])
def test_dct_prior_eigenvalues_binding(shape, nd, vs, order):
    """The library's host loop gives exactly the float32 rounding of the
    NumPy restatement of BINDER/utils.py:331-349 (host code, no GPU)."""
    from binder_amd import _core
    from binder_amd.utils import initialize_fourier_domain_smoothing
    got = _core.dct_prior_eigenvalues(shape, order, nd, voxel_size=np.array(vs))
    exp = np.asarray(initialize_fourier_domain_smoothing(list(shape), True, order, nd,
                                                         voxel_size=np.array(vs)), np.float32)
    assert got.shape == exp.shape and got.dtype == np.float32
    assert np.array_equal(got, exp)


def test_spline_binding(orc):
    """Spline(order, is2D).spline / log_spline of the HIP library
    (BINDER/tests/cpu/test_check_spline.py:5-76 properties, host code, no
    GPU): partition of unity, exp(log) = weights, no NaN at the range edges,
    and the oracle's weights (pinned to the reference build)."""
    from binder_amd import _core
    for dim in (2, 3):
        for order in (1, 2, 3):
            sp = _core.Spline(order, 1 if dim == 2 else 0)
            ds = [0.4, 0.3] if dim == 2 else [0.2, 0.4, 0.1]
            ws = [np.array(sp.spline(d))[:order + 1] for d in ds]
            ls = [np.array(sp.log_spline(d))[:order + 1] for d in ds]
            prod = ws[0][:, None, None] * ws[1][None, :, None]
            if dim == 3:
                prod = prod * ws[2][None, None, :]
            assert np.isclose(prod.sum(), 1.0)
            for w, lw in zip(ws, ls):
                assert np.allclose(w, np.exp(lw))
            for d in ds:
                assert np.allclose(np.array(sp.spline(d))[:order + 1],
                                   orc.spline(order, d)[:order + 1], rtol=0, atol=1e-15)
            edge = [0.0, 1.0] if dim == 2 else [0.0, 1.0, -0.3]
            ws = [np.array(sp.spline(d))[:order + 1] for d in edge]
            prod = ws[0][:, None, None] * ws[1][None, :, None]
            if dim == 3:
                prod = prod * ws[2][None, None, :]
            assert np.isclose(prod.sum(), 1.0)
            for d in edge:
                assert not np.any(np.isnan(np.array(sp.log_spline(d))[:order + 1]))


def test_host_histogram_and_binning_match_numpy(core):
    """The library's multithreaded host histogram and floor binning (used by
    bin_signal_robust, BINDER/utils.py:20-82) give numpy's counts, edges and
    bins exactly: smooth and noisy images, values on bin edges, a constant
    image, a tiny range and negative values."""
    from binder_amd import utils
    rng = np.random.default_rng(5)
    cases = [rng.random((40, 33, 27)),
             np.round(rng.normal(0.3, 0.2, (50, 41, 9)), 3),
             np.full((17, 5, 3), 0.25),
             1e-9 * rng.random((64, 64, 2)) + 7.0,
             rng.normal(-3.0, 2.0, (200_000,))]
    edges_case = np.linspace(0.0, 1.0, 1001)
    cases.append(np.concatenate([edges_case, edges_case[::7], rng.random(70_000)]))
    for x in cases:
        c, e = core.histogram_uniform(x, 1000)
        cn, en = np.histogram(x, 1000)
        assert np.array_equal(e, en) and np.array_equal(c, cn)
        for L in (32, 7):
            got, ge = utils.bin_signal_robust(x, L)
            # the same computation with numpy's histogram and binning
            hist, centers = np.histogram(np.asarray(x, np.float64), 1000)
            cdf = np.cumsum(hist / np.sum(hist))

            def edge(p):
                idx = np.where(cdf <= p)[0]
                add = 0.01
                while len(idx) == 0:
                    idx = np.where(cdf <= p + add)[0]
                    add += 0.01
                return centers[idx[-1]]
            lo, hi = edge(0.01), edge(0.99)
            width = (hi - lo) / (L - 1)
            with np.errstate(divide="ignore", invalid="ignore"):
                b = np.floor((x - lo) / width).astype(np.int32)
            b[x < lo] = 0
            b[x > hi] = L - 1
            assert np.array_equal(got, b), (x.shape, L)
    with pytest.raises(ValueError):
        core.histogram_uniform(np.array([1.0, np.nan]), 10)

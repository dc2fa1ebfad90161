"""Host-side API: input utilities, result types, device dispatch (CPU only).

Known answers follow the reference's unit tests
(tests/test_periodfind.py:64-306): significance closed forms 4.0 and 6.0,
argmax (1,2) -> 6.0, argmin (1,1) -> 0.5, top-3 9/7/5, best_params.
"""

import warnings

import numpy as np
import pytest

import periodfind_amd
from periodfind_amd import Periodogram, Statistics
from periodfind_amd._utils import (ensure_float32, magnitude_mode, prepare_magnitudes,
                                   validate_errs, validate_inputs)


class TestUtils:
    def test_magnitude_mode_mirrors_prepare_magnitudes(self):
        # The drop-in classes prepare on the device; what the host still does
        # (mode choice, warning, numpy's errors on empty curves) must match.
        one = [np.ones(3, np.float32)]
        assert magnitude_mode(one, False, False) is None
        assert magnitude_mode(one, False, True) == "normalize"
        assert magnitude_mode(one, True, False) == "center"
        with pytest.warns(RuntimeWarning, match="conflicting settings"):
            assert magnitude_mode(one, True, True) == "center"
        empty = [np.ones(3, np.float32), np.zeros(0, np.float32)]
        with pytest.raises(ValueError) as want:
            prepare_magnitudes(empty, False, True)
        with pytest.raises(ValueError) as got:
            magnitude_mode(empty, False, True)
        assert str(got.value) == str(want.value)
        with pytest.warns(RuntimeWarning, match="Mean of empty slice"):
            magnitude_mode(empty, True, False)

    def test_normalize_range_and_formula(self):
        m = [np.array([1.0, 2.0, 3.0], dtype=np.float32)]
        r = prepare_magnitudes(m, center=False, normalize=True)[0]
        assert r.dtype == np.float32
        np.testing.assert_array_equal(
            r, ((m[0] - 1.0) / 2.0) * np.float32(0.999) + np.float32(5e-4))
        assert r.min() >= 0.0 and r.max() <= 1.0

    def test_center(self):
        r = prepare_magnitudes([np.array([1.0, 2.0, 3.0], np.float32)], True, False)[0]
        assert abs(np.mean(r)) < 1e-6

    def test_passthrough(self):
        m = [np.array([1.0, 2.0, 3.0], np.float32)]
        assert prepare_magnitudes(m, False, False)[0] is m[0]

    def test_conflict_warns_and_centers(self):
        m = [np.array([1.0, 2.0, 3.0], np.float32)]
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            r = prepare_magnitudes(m, center=True, normalize=True)[0]
        assert any(issubclass(x.category, RuntimeWarning) for x in w)
        assert abs(np.mean(r)) < 1e-6

    def test_validate_inputs(self):
        validate_inputs([np.zeros(2, np.float32)], [np.zeros(2, np.float32)])
        with pytest.raises(ValueError, match="same number of light curves"):
            validate_inputs([np.zeros(2), np.zeros(2)], [np.zeros(2)])
        with pytest.raises(ValueError, match="different lengths"):
            validate_inputs([np.zeros(3)], [np.zeros(2)])

    def test_ensure_float32(self):
        ensure_float32([np.zeros(2, np.float32)], "t")
        with pytest.raises(TypeError, match="float64.*expected float32"):
            ensure_float32([np.zeros(2, np.float64)], "t")

    def test_validate_errs(self):
        t = [np.zeros(3, np.float32)]
        assert validate_errs(None, t) is None
        with pytest.raises(ValueError, match="same number of arrays"):
            validate_errs([np.ones(3, np.float32)] * 2, t)
        with pytest.raises(ValueError, match="different lengths"):
            validate_errs([np.ones(2, np.float32)], t)
        with pytest.raises(TypeError, match="expected float32"):
            validate_errs([np.ones(3, np.float64)], t)


class TestStatistics:
    def test_significance_closed_forms(self):
        assert Statistics([1.0], 3.0, 1.0, 0.5, 1.0, 0.4).significance == pytest.approx(4.0)
        s = Statistics([1.0], 3.0, 1.0, 0.5, 1.2, 0.3, significance_type="madmedian")
        assert s.significance == pytest.approx(6.0)
        with pytest.raises(NotImplementedError):
            _ = Statistics([1.0], 3.0, 1.0, 0.5, 1.0, 0.4, significance_type="bogus").significance

    def test_repr(self):
        r = repr(Statistics([1.0, 0.0], 0.5, 1.0, 0.2, 1.0, 0.15))
        assert "Statistics" in r and "0.5" in r and "stdmean" in r

    def test_from_data_argmax_argmin(self):
        data = np.array([[1.0, 2.0, 3.0], [4.0, 0.5, 6.0]], np.float32)
        params = [np.array([10.0, 20.0], np.float32), np.array([0.0, 0.1, 0.2], np.float32)]
        s = Statistics.statistics_from_data(data, params, use_max=True)
        assert s.value == pytest.approx(6.0) and s.params == pytest.approx([20.0, 0.2])
        s = Statistics.statistics_from_data(data, params, use_max=False)
        assert s.value == pytest.approx(0.5) and s.params == pytest.approx([20.0, 0.1])

    def test_from_data_top3_sorted(self):
        data = np.array([5.0, 1.0, 9.0, 3.0, 7.0], np.float32)
        params = [np.array([0.1, 0.2, 0.3, 0.4, 0.5], np.float32)]
        r = Statistics.statistics_from_data(data, params, use_max=True, n=3)
        assert [x.value for x in r] == pytest.approx([9.0, 7.0, 5.0])

    def test_from_data_precomputed_and_computed(self):
        data = np.array([1.0, 2.0, 3.0, 4.0, 5.0], np.float32)
        params = [np.arange(5, dtype=np.float32)]
        s = Statistics.statistics_from_data(data, params, True, mean=100.0, std=50.0,
                                            median=99.0, mad=45.0)
        assert (s.mean, s.std, s.median, s.mad) == (100.0, 50.0, 99.0, 45.0)
        s = Statistics.statistics_from_data(data, params, True)
        assert s.mean == np.mean(data) and s.std == np.std(data)
        assert s.median == np.median(data)
        assert s.mad == np.median(np.abs(data - np.median(data)))


class TestPeriodogram:
    def test_best_params(self):
        data = np.array([[1.0, 5.0], [3.0, 2.0]], np.float32)
        params = [np.array([0.1, 0.2], np.float32), np.array([0.0, 0.01], np.float32)]
        b = Periodogram(data, params, use_max=True).best_params()
        assert b.value == pytest.approx(5.0) and b.params == pytest.approx([0.1, 0.01])
        b = Periodogram(data, params, use_max=False).best_params()
        assert b.value == pytest.approx(1.0) and b.params == pytest.approx([0.1, 0.0])

    def test_best_params_multiple_and_populates(self):
        data = np.array([1.0, 5.0, 3.0, 9.0, 2.0], np.float32)
        p = Periodogram(data, [np.arange(5, dtype=np.float32)], use_max=True)
        assert p.mean is None
        r = p.best_params(n=2)
        assert [x.value for x in r] == pytest.approx([9.0, 5.0])
        assert None not in (p.mean, p.std, p.median, p.mad)
        assert "Periodogram" in repr(p) and "True" in repr(p)

    def test_madmedian(self):
        data = np.array([1.0, 2.0, 3.0, 4.0, 100.0], np.float32)
        b = Periodogram(data, [np.arange(5, dtype=np.float32)], True).best_params(
            significance_type="madmedian")
        assert b.significance_type == "madmedian" and b.significance > 0


class TestDeviceAPI:
    def setup_method(self):
        periodfind_amd._default_device = None

    def teardown_method(self):
        periodfind_amd._default_device = None

    def test_set_get(self):
        periodfind_amd.set_device("GPU")
        assert periodfind_amd.get_device() == "gpu"
        periodfind_amd.set_device("cpu")
        assert periodfind_amd.get_device() == "cpu"

    def test_unknown_device(self):
        with pytest.raises(ValueError, match="Unknown device"):
            periodfind_amd.set_device("tpu")
        with pytest.raises(ValueError, match="Unknown device"):
            periodfind_amd.ConditionalEntropy(device="tpu")

    def test_cpu_device_not_served(self):
        with pytest.raises(NotImplementedError, match="GPU"):
            periodfind_amd.AOV(device="cpu")

    def test_auto_detect_matches_runtime(self):
        from periodfind_amd import _pfhip

        expect = "gpu" if _pfhip.device_count() > 0 else "cpu"
        assert periodfind_amd.get_device() == expect

    def test_factory_returns_gpu_class(self):
        from periodfind_amd import gpu

        ce = periodfind_amd.ConditionalEntropy(n_phase=12, n_mag=8, device="gpu")
        assert isinstance(ce, gpu.ConditionalEntropy)
        assert (ce.n_phase, ce.n_mag) == (12, 8)
        bls = periodfind_amd.BoxLeastSquares(n_bins=40, qmin=0.02, qmax=0.3, device="gpu")
        assert bls._params() == dict(n_bins=40, qmin=0.02, qmax=0.3)


class TestGpuClassValidation:
    """Validation runs before any device call, so it is testable without a GPU."""

    def _data(self):
        t = [np.linspace(0, 10, 20, dtype=np.float32)]
        return t, [np.ones(20, np.float32)]

    def test_float64_rejected(self):
        from periodfind_amd import gpu

        t, m = self._data()
        with pytest.raises(TypeError, match="expected float32"):
            gpu.LombScargle().calc([x.astype(np.float64) for x in t], m, np.ones(3), np.zeros(1))

    def test_mismatch_rejected(self):
        from periodfind_amd import gpu

        t, m = self._data()
        with pytest.raises(ValueError, match="same number of light curves"):
            gpu.AOV().calc(t + t, m, np.ones(3, np.float32), np.zeros(1, np.float32))
        with pytest.raises(ValueError, match="different lengths"):
            gpu.FPW().calc(t, m, np.ones(3, np.float32), np.zeros(1, np.float32),
                           errs=[np.ones(5, np.float32)])


def test_autograd_refuses_cpu_tensors():
    """The differentiable LS has no CPU path: CPU tensors raise before any
    library call."""
    torch = pytest.importorskip("torch")
    from periodfind_amd import autograd as AG

    t = torch.linspace(0, 10, 20, dtype=torch.float64)
    with pytest.raises(ValueError, match="GPU tensors"):
        AG.ls_power(t, torch.sin(t), torch.tensor([1.0, 2.0], dtype=torch.float64))

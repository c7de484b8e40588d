"""Budget policy arithmetic on the CPU: realized coverage, the calibration
search, the Hoeffding constant and label algebra, against closed forms."""
import math

import numpy as np
import pytest

from ptwalker_amd.budget import _smallest, hoeffding_c2, labels_from_constants, realized_coverage


def test_realized_coverage_closed_form():
    c = np.array([1.0, 2.0])
    shape = np.array([[0.9, 1.1], [1.0, 1.0]])
    eps = np.array([1e-3, 1e-3])
    # W = 1e6 gives errors 0.9e-3 and 1.1e-3 for source 0; W = 4e6 gives exactly 1e-3 for source 1
    assert realized_coverage([1e6, 4e6], c, shape, eps) == 0.75
    assert realized_coverage([1.21e6, 4e6], c, shape, eps) == 1.0
    assert realized_coverage([1.0, 1.0], c, shape, eps) == 0.0


def test_smallest_scale_search():
    c = np.array([1.0, 2.0])
    shape = np.array([[0.9, 1.1], [1.0, 1.0]])
    eps = np.array([1e-3, 1e-3])
    k = _smallest(lambda k: realized_coverage(k * np.array([1.0, 4.0]) / 1e-6, c, shape, eps), 1.0)
    assert k == pytest.approx(1.21, rel=1e-9)  # (1.1 c_0)^2 / eps^2 = 1.21e6
    # half coverage needs only source 1's exact budget
    k = _smallest(lambda k: realized_coverage(k * np.array([1.0, 4.0]) / 1e-6, c, shape, eps), 0.75)
    assert k == pytest.approx(1.0, rel=1e-9)


def test_hoeffding_constant():
    # 2 (m ln 2 + ln(1/delta)) with m = k + 1 cells
    assert hoeffding_c2(100, 0.05) == pytest.approx(2 * (101 * math.log(2) + math.log(20)))
    assert hoeffding_c2(100, 0.05) / 1e-3 ** 2 == pytest.approx(146.0e6, rel=2e-3)


def test_labels_are_log_budgets():
    c = np.array([0.5, 1.0, 2.0])
    y = labels_from_constants(c, 1e-3)
    assert np.allclose(np.exp(y), (c / 1e-3) ** 2)
    # expected L1 at that budget equals eps
    assert np.allclose(c / np.sqrt(np.exp(y)), 1e-3)

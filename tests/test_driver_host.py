"""Host-side rules of the driver (no GPU): B-solve and Ap-solve degrees,
the frequency interval conversion, stream split and KPM degree."""
import math

import numpy as np
import pytest

from normalmodes_amd import driver


def test_reference_bsolve_degrees():
    # src/mod_matvec.f90:88-92: 45 for P2, 25 for P1
    assert driver.chebiter_degree(2, 0.3, 4.5) == 45
    assert driver.chebiter_degree(1, 0.3, 4.5) == 25
    assert driver.chebiter_degree(2, 0.3, 4.5, policy=17) == 17
    with pytest.raises(ValueError):
        driver.chebiter_degree(2, 0.3, 4.5, policy="fastest")


@pytest.mark.parametrize("kappa", [4.0, 15.0, 16.3, 60.0])
def test_auto_bsolve_degree_is_the_smallest_meeting_the_chebyshev_bound(kappa):
    """The Chebyshev semi-iteration on [lmin, lmax] reduces the error by
    1/T_c(theta/delta) (theta, delta the centre and half-width).  The auto
    degree is the smallest c whose 2 rho^c bound is below eps; the exact
    factor, evaluated with numpy's Chebyshev series, is then below eps too,
    and one degree less misses the 2 rho^c bound."""
    lmin, lmax, eps = 0.25, 0.25 * kappa, 1e-13
    c = driver.chebiter_degree(2, lmin, lmax, policy="auto", eps=eps)
    s = (lmax + lmin) / (lmax - lmin)
    t_c = np.polynomial.chebyshev.chebval(s, [0] * c + [1])
    assert 1.0 / t_c <= eps
    rho = (math.sqrt(kappa) - 1) / (math.sqrt(kappa) + 1)
    assert 2 * rho ** (c - 1) > eps
    # the bound the policy uses is the textbook one: 1/T_c <= 2 rho^c
    assert 1.0 / t_c <= 2 * rho ** c * (1 + 1e-12)


def test_ap_degree_is_100_for_p2():
    # src/mod_matvec.f90:167-171: max(min(int(k), 15) * 3, 100) never exceeds 100 at P2
    for k in (1.5, 15.0, 33.4, 1e6):
        assert driver.ap_chebiter_degree(2, 1.0, k) == 100
    assert driver.ap_chebiter_degree(1, 1.0, 50.0) == 25


def test_ap_degree_auto_meets_its_bound():
    """apdeg="auto": the smallest c with 2 rho^c <= 1e-13 for the Ap~ bounds."""
    for lo, hi in ((0.28, 4.35), (0.1, 9.0), (1.0, 1.5)):
        c = driver.ap_chebiter_degree(2, lo, hi, "auto")
        rho = (math.sqrt(hi / lo) - 1) / (math.sqrt(hi / lo) + 1)
        assert 2 * rho ** c <= 1e-13 < 2 * rho ** (c - 1)
    assert driver.ap_chebiter_degree(2, 0.28, 4.35, 77) == 77
    with pytest.raises(ValueError):
        driver.ap_chebiter_degree(2, 0.28, 4.35, "fast")


def test_interval_conversion_keeps_the_reference_single_precision():
    """lowfreq/upfreq and PI are default REALs in the reference
    (src/mod_para.f90:51-52, src/mod_pevsl.f90:43): the interval differs from
    the exact double conversion by float32 rounding, not more."""
    for f in (0.1, 0.2, 1.0, 2.0, 9.7):
        lam = driver.freq_to_lambda(f)
        exact = (2 * math.pi * f) ** 2 * 1e-6
        rel = abs(lam / exact - 1)
        assert rel < 3e-7
        assert abs(driver.lambda_to_freq(lam) / f - 1) < 2e-7
    assert abs(driver.freq_to_lambda(0.1) / ((2 * math.pi * 0.1) ** 2 * 1e-6) - 1) > 1e-9
    assert driver.lambda_to_freq(-1e-12) == 0.0


def test_split_parts():
    assert [driver.split_parts(b) for b in (1, 2, 4, 8, 16)] == [0, 0, 0, 2, 2]
    assert driver.split_parts(8, 4) == 4 and driver.split_parts(8, 0) == 0


def test_kpm_degree_grows_with_the_slices_and_stays_bounded():
    xintv = (0.5, 12.0, 0.1, 1.0e4)
    d = [driver.kpm_degree(xintv, s) for s in (1, 2, 4, 8, 16)]
    assert d == sorted(d) and 50 <= d[0] and d[-1] <= 3000
    assert driver.kpm_degree((0.5, 0.5 + 1e-9, 0.1, 1.0e4), 4) == 3000


def test_slice_waves_partition_and_fit():
    """Waves are consecutive, cover every slice once, and each fits the
    budget unless it is a single slice too large by itself."""
    n = 8_600_000
    rng = np.random.default_rng(4)
    counts = rng.integers(20, 251, size=16)              # a C4-like spread of slice counts
    per = 16.0 * n
    for budget in (50e9, 200e9, 1e15):
        waves = driver.plan_slice_waves(counts, n, budget)
        assert waves[0][0] == 0 and waves[-1][1] == counts.size
        assert all(w[1] == v[0] for w, v in zip(waves, waves[1:]))
        for i0, i1 in waves:
            need = sum((3.0 * c + 40) * per for c in counts[i0:i1])
            assert need <= budget or i1 - i0 == 1
        if budget == 1e15:
            assert waves == [(0, 16)]
    # at 200 GB (a 288 GB device holding the operator and work besides) the
    # 16 slices need several waves: the whole basis would not fit at once
    total = sum((3.0 * c + 40) * per for c in counts)
    assert total > 288e9
    assert len(driver.plan_slice_waves(counts, n, 200e9)) >= 2
    # width cap
    assert all(i1 - i0 <= 4 for i0, i1 in driver.plan_slice_waves(np.ones(10), 100, 1e12, max_width=4))


def test_refine_inner_degree_meets_its_per_pass_bound():
    """The inner degree of a refinement pass: the smallest c with
    2 rho^c <= the per-pass target (the FP32 iterates' own floor is ~1e-6)."""
    for lo, hi, target in ((0.25, 4.35, 1e-6), (0.3, 3.0, 1e-5), (1.0, 2.0, 1e-7)):
        c = driver.refine_inner_degree(lo, hi, target)
        rho = (math.sqrt(hi / lo) - 1) / (math.sqrt(hi / lo) + 1)
        assert 2 * rho ** c <= target < 2 * rho ** (c - 1)
    # two passes of the default reach about the square of the per-pass error
    assert driver.refine_inner_degree(0.25, 4.35) < driver.chebiter_degree(2, 0.25, 4.35, "auto")

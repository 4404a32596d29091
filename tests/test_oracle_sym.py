"""The symmetric-Aladin matrix update of the oracle against the reference build and closed forms.

reg_aladin_sym averages the forward matrix with the inverted backward one (and vice versa) by
NiftyReg::Mat44Avg2 = expm((logm A + logm B) / 2), which the reference evaluates with Eigen's
matrix functions.  The oracle computes the same principal logarithm and exponential in double by
its own method (inverse scaling and squaring with a Gauss-Legendre Pade form; Taylor scaling and
squaring), so after the float roundings the reference applies the two agree to a float ulp.
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build absent")


def _ulps(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b), dtype=np.float32)).astype(np.float64)
    return np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(sp, 1e-45)


def _rotz(t, shift=(0, 0, 0)):
    m = np.eye(4, dtype=np.float32)
    m[:2, :2] = [[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]]
    m[:3, 3] = shift
    return m


def _affines(n, seed=0):
    return [phantoms.random_affine(seed + i).astype(np.float32) for i in range(n)]


@needs_ref
def test_nifti_inverse_bit_exact():
    for A in _affines(20):
        assert np.array_equal(O.orc_nifti_inverse(A), O.ref_mat44_inverse(A))


@needs_ref
def test_avg2_matches_reference():
    worst = 0.0
    for i, A in enumerate(_affines(30)):
        B = phantoms.random_affine(100 + i).astype(np.float32)
        got, want = O.orc_avg2(A, B), O.ref_mat44_avg2(A, B)
        # zero entries of the exact result carry rounding noise of order 1e-16 in either build
        close_to_zero = np.abs(want) < 1e-6
        assert np.all(np.abs(got - want)[close_to_zero] < 1e-6)
        worst = max(worst, _ulps(got, want)[~close_to_zero].max())
    assert worst <= 2, worst


@needs_ref
def test_symmetric_update_matches_reference_steps():
    for i, F in enumerate(_affines(10, 7)):
        Bw = (np.linalg.inv(F.astype(np.float64)) @ phantoms.random_affine(50 + i, scale=(0.98, 1.02))).astype(np.float32)
        Bw[3] = [0, 0, 0, 1]
        nf, nb = O.orc_symmetric_update(F, Bw)
        wf = O.ref_mat44_avg2(F, O.ref_mat44_inverse(Bw))
        wb = O.ref_mat44_avg2(O.ref_mat44_inverse(F), Bw)
        wf[3], wb[3] = [0, 0, 0, 1], [0, 0, 0, 1]
        assert np.abs(nf - wf).max() < 1e-5 and np.abs(nb - wb).max() < 1e-5


def test_avg2_closed_forms():
    # the average of two rotations about one axis is the rotation by the mean angle
    got = O.orc_avg2(_rotz(0.1), _rotz(0.3))
    assert np.abs(got - _rotz(0.2)).max() < 2e-7
    # pure translations average arithmetically
    a, b = np.eye(4, dtype=np.float32), np.eye(4, dtype=np.float32)
    a[:3, 3], b[:3, 3] = [2, -4, 6], [4, 0, -2]
    assert np.abs(O.orc_avg2(a, b)[:3, 3] - [3, -2, 2]).max() < 1e-6
    # a matrix averaged with itself is itself; with its inverse, the identity
    for A in _affines(10, 3):
        assert np.abs(O.orc_avg2(A, A) - A).max() < 1e-5
        assert np.abs(O.orc_avg2(A, O.orc_nifti_inverse(A)) - np.eye(4)).max() < 1e-5
    # scalings average geometrically
    s1, s2 = np.diag([1.2, 0.8, 1.0, 1.0]).astype(np.float32), np.diag([0.9, 1.1, 1.25, 1.0]).astype(np.float32)
    want = np.diag([np.sqrt(1.08), np.sqrt(0.88), np.sqrt(1.25), 1.0])
    assert np.abs(O.orc_avg2(s1, s2) - want).max() < 1e-6


def test_symmetric_update_of_consistent_pair_is_fixed_point():
    for A in _affines(10, 11):
        nf, nb = O.orc_symmetric_update(A, O.orc_nifti_inverse(A))
        assert np.abs(nf - A).max() < 1e-5
        assert np.abs(nb @ nf - np.eye(4)).max() < 1e-5

"""Eisenstat-Walker forcing term of the Newton driver's linear solves (host logic only).

eta_0 = ksp_rtol; eta_k = gamma (|g_k| / |g_{k-1}|)^alpha with gamma = 0.9, alpha = 2,
raised to gamma eta_{k-1}^alpha when that exceeds 0.1, capped by ksp_rtol and floored at
1e-8 (Eisenstat & Walker 1996, choice 2); the fixed policy always returns ksp_rtol.
"""

import pytest

from fenics_nonlinear_energies_amd.drivers import LinearSolveContext


class _Vec:
    def __init__(self, n):
        self.n = n

    def norm(self):
        return self.n


class _Asm:
    class ksp:  # noqa: N801 - attribute bag standing in for the assembler's KSP
        _amg = None
        _inv_diag = None


def _ctx(forcing, rtol=0.1):
    return LinearSolveContext(_Asm(), ksp_rtol=rtol, ksp_max_it=100, ksp_forcing=forcing)


def test_fixed_policy_returns_ksp_rtol():
    c = _ctx(None)
    assert [c._rtol_for(_Vec(g)) for g in (10.0, 1.0, 1e-3)] == [0.1, 0.1, 0.1]


def test_eisenstat_walker_sequence():
    c = _ctx("ew")
    assert c._rtol_for(_Vec(1.0)) == 0.1  # first solve: ksp_rtol
    # fast decrease: 0.9 * (0.01)^2 = 9e-5, safeguard 0.9 * 0.1^2 = 9e-3 <= 0.1 -> not applied
    assert c._rtol_for(_Vec(0.01)) == pytest.approx(9e-5)
    # slow decrease: 0.9 * 0.9^2 = 0.729 -> capped by ksp_rtol
    assert c._rtol_for(_Vec(0.009)) == pytest.approx(0.1)
    # after eta = 0.1 the safeguard 0.9 * 0.1^2 = 9e-3 is below 0.1, so the plain term holds
    assert c._rtol_for(_Vec(0.009 * 1e-3)) == pytest.approx(max(0.9 * 1e-6, 1e-8))
    c.reset()
    assert c._rtol_for(_Vec(5.0)) == 0.1  # reset starts over


def test_eisenstat_walker_safeguard_and_floor():
    c = _ctx("ew", rtol=0.9)
    c._rtol_for(_Vec(1.0))  # eta_0 = 0.9
    # plain term 0.9 * (1e-6)^2 is tiny, safeguard 0.9 * 0.9^2 = 0.729 > 0.1 keeps it up
    assert c._rtol_for(_Vec(1e-6)) == pytest.approx(0.729)
    c2 = _ctx("ew", rtol=0.05)
    c2._rtol_for(_Vec(1.0))
    assert c2._rtol_for(_Vec(1e-9)) == pytest.approx(1e-8)  # floor


def test_unknown_forcing_rejected():
    with pytest.raises(ValueError):
        _ctx("bogus")

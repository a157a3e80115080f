"""Oracle element energies with hand-derived gradients and Hessians (numpy, FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Vectorised over elements.
"""

from __future__ import annotations

import numpy as np


# ------------------------------------------------------------------------------------------
# A1 p-Laplace: plaplace/jax_petsc/reordered_element_assembler.py:16-19 (integrand) and
# :38-63 (energy * vol); serial J plaplace/jax/jax_energy.py:7-15.
# ------------------------------------------------------------------------------------------
def plaplace(v, dvx, dvy, vol, p, want=("e", "g", "h")):
    """v, dvx, dvy: (N, 3); vol: (N,).  Returns dict with e (N,), g (N,3), h (N,3,3)."""
    v = np.asarray(v, np.float64)
    fx = np.sum(v * dvx, axis=1)
    fy = np.sum(v * dvy, axis=1)
    s = fx * fx + fy * fy
    out = {}
    if "e" in want:
        out["e"] = (1.0 / p) * s ** (p / 2.0) * vol
    t = fx[:, None] * dvx + fy[:, None] * dvy  # g . grad(phi_a)
    with np.errstate(divide="ignore", invalid="ignore"):
        c1 = np.where(s > 0.0, s ** (p / 2.0 - 1.0), 1.0 if p == 2.0 else 0.0)
        c2 = np.where(s > 0.0, (p - 2.0) * s ** (p / 2.0 - 2.0), 0.0)
    if "g" in want:
        out["g"] = (vol * c1)[:, None] * t
    if "h" in want:
        K = dvx[:, :, None] * dvx[:, None, :] + dvy[:, :, None] * dvy[:, None, :]
        out["h"] = vol[:, None, None] * (c1[:, None, None] * K + c2[:, None, None] * t[:, :, None] * t[:, None, :])
    return out


# ------------------------------------------------------------------------------------------
# A2 Ginzburg-Landau: ginzburg_landau/jax_petsc/reordered_element_assembler.py:18-24;
# serial ginzburg_landau/jax/jax_energy.py:7-16.  ip: (3, nq), w: (nq,)
# ------------------------------------------------------------------------------------------
def ginzburg_landau(v, dvx, dvy, vol, ip, w, eps, want=("e", "g", "h")):
    v = np.asarray(v, np.float64)
    fx = np.sum(v * dvx, axis=1)
    fy = np.sum(v * dvy, axis=1)
    s = fx * fx + fy * fy
    uq = v @ ip  # (N, nq)
    out = {}
    if "e" in want:
        out["e"] = (0.5 * eps * s + 0.25 * (((uq * uq) - 1.0) ** 2) @ w) * vol
    if "g" in want:
        t = fx[:, None] * dvx + fy[:, None] * dvy
        pot = ((uq * uq - 1.0) * uq * w[None, :]) @ ip.T  # (N, 3)
        out["g"] = vol[:, None] * (eps * t + pot)
    if "h" in want:
        K = dvx[:, :, None] * dvx[:, None, :] + dvy[:, :, None] * dvy[:, None, :]
        c = (3.0 * uq * uq - 1.0) * w[None, :]  # (N, nq)
        M = np.einsum("nq,aq,bq->nab", c, ip, ip)
        out["h"] = vol[:, None, None] * (eps * K + M)
    return out


# ------------------------------------------------------------------------------------------
# A3 compressible Neo-Hookean: hyperelasticity/jax_petsc/parallel_hessian_dof.py:17-52
# (signed det; use_abs_det=True gives the serial |det| form of jax/jax_energy.py:26).
# v: (N, 12) interleaved deformed positions [x0,y0,z0,x1,...]; dphi*: (N, 4)
# ------------------------------------------------------------------------------------------
def neohookean(v, dphix, dphiy, dphiz, vol, C1, D1, use_abs_det=False, want=("e", "g", "h")):
    v = np.asarray(v, np.float64)
    N = v.shape[0]
    y = v.reshape(N, 4, 3)  # node a, component i
    D = np.stack((dphix, dphiy, dphiz), axis=2)  # (N, 4, 3): d phi_a / d X_j
    F = np.einsum("nai,naj->nij", y, D)
    det = np.linalg.det(F)
    J = np.abs(det) if use_abs_det else det
    I1 = np.sum(F * F, axis=(1, 2))
    out = {}
    if "e" in want:
        with np.errstate(invalid="ignore", divide="ignore"):
            W = C1 * (I1 - 3.0 - 2.0 * np.log(J)) + D1 * (J - 1.0) ** 2
        out["e"] = W * vol
    if "g" in want or "h" in want:
        G = np.transpose(np.linalg.inv(F), (0, 2, 1))  # F^{-T}
        q = np.einsum("nij,naj->nai", G, D)  # q_a = G dphi_a
    if "g" in want:
        P = 2.0 * C1 * F + (-2.0 * C1 + 2.0 * D1 * (J - 1.0) * J)[:, None, None] * G
        g = vol[:, None, None] * np.einsum("nij,naj->nai", P, D)
        out["g"] = g.reshape(N, 12)
    if "h" in want:
        dd = np.einsum("naj,nbj->nab", D, D)
        alpha = 2.0 * C1 - 2.0 * D1 * J * (J - 1.0)
        beta = 2.0 * D1 * J * (2.0 * J - 1.0)
        eye = np.eye(3)
        # H[(a,i),(b,k)] = vol [2C1 d_ik (Da.Db) + alpha q_b,i q_a,k + beta q_a,i q_b,k]
        t0 = 2.0 * C1 * np.einsum("nab,ik->naibk", dd, eye)
        t1 = alpha[:, None, None, None, None] * np.einsum("nbi,nak->naibk", q, q)
        t2 = beta[:, None, None, None, None] * np.einsum("nai,nbk->naibk", q, q)
        H = t0 + t1 + t2
        out["h"] = (vol[:, None, None, None, None] * H).reshape(N, 12, 12)
    return out


# ------------------------------------------------------------------------------------------
# A4 Mohr-Coulomb plane strain: slope_stability/jax/jax_energy.py:11-30 (tensors),
# :45-108 (density with return mapping), :121-142 (element energy = sum_q w_q psi_q).
# ------------------------------------------------------------------------------------------
def elastic_tensors(E, nu):
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    C = np.array([[lam + 2.0 * mu, lam, 0.0], [lam, lam + 2.0 * mu, 0.0], [0.0, 0.0, mu]])
    S = np.array(
        [
            [(1.0 - nu * nu) / E, -(nu * (1.0 + nu)) / E, 0.0],
            [-(nu * (1.0 + nu)) / E, (1.0 - nu * nu) / E, 0.0],
            [0.0, 0.0, 1.0 / mu],
        ]
    )
    return C, S


def mc_stress_tangent(eps_e, E, nu, phi_deg, cohesion, reg=1e-12, elastic_only=False):
    """Per quadrature point: psi (M,), sigma (M,3), D = d sigma / d eps (M,3,3).

    eps_e: (M, 3) elastic trial strain (engineering shear).  The tangent of the line
    return is derived by hand from sigma = R(theta_trial) diag(s1p, s2p); the apex return
    gives a constant stress (zero tangent).
    """
    C, S = elastic_tensors(E, nu)
    eps_e = np.asarray(eps_e, np.float64)
    M = eps_e.shape[0]
    sig_tr = eps_e @ C.T
    if elastic_only:
        psi = np.sum(sig_tr * eps_e, axis=1) - 0.5 * np.einsum("mi,ij,mj->m", sig_tr, S, sig_tr)
        return psi, sig_tr, np.broadcast_to(C, (M, 3, 3)).copy()
    sxx, syy, txy = sig_tr[:, 0], sig_tr[:, 1], sig_tr[:, 2]
    m = 0.5 * (sxx + syy)
    d = 0.5 * (sxx - syy)
    r = np.sqrt(d * d + txy * txy + reg * reg)
    s1 = m + r
    s2 = m - r
    phi = np.deg2rad(phi_deg)
    alpha = 1.0 + np.sin(phi)
    beta = 1.0 - np.sin(phi)
    kappa = 2.0 * cohesion * np.cos(phi)
    f_tr = alpha * s1 - beta * s2 - kappa
    A11, A12, A22 = S[0, 0], S[0, 1], S[1, 1]
    n1, n2 = alpha, -beta
    det = A11 * A22 - A12 * A12
    z1 = (A22 * n1 - A12 * n2) / det
    z2 = (-A12 * n1 + A11 * n2) / det
    nz = n1 * z1 + n2 * z2
    lam_mc = (n1 * s1 + n2 * s2 - kappa) / nz
    s1_line = s1 - lam_mc * z1
    s2_line = s2 - lam_mc * z2
    safe_den = alpha - beta if abs(alpha - beta) > 1e-14 else 1.0
    s_apex = kappa / safe_den
    line = s1_line >= s2_line
    s1p = np.where(line, s1_line, s_apex)
    s2p = np.where(line, s2_line, s_apex)
    mp = 0.5 * (s1p + s2p)
    rp = 0.5 * (s1p - s2p)
    two_theta = np.arctan2(2.0 * txy, sxx - syy)
    c2, sn2 = np.cos(two_theta), np.sin(two_theta)
    sig_pl = np.stack((mp + rp * c2, mp - rp * c2, rp * sn2), axis=1)
    elastic = f_tr <= 0.0
    sig = np.where(elastic[:, None], sig_tr, sig_pl)
    psi = np.sum(sig * eps_e, axis=1) - 0.5 * np.einsum("mi,ij,mj->m", sig, S, sig)

    # tangent of the line return: (s1p, s2p) = P (s1, s2) + const, P = I - z n^T / (n.z)
    P11 = 1.0 - z1 * n1 / nz
    P12 = -z1 * n2 / nz
    P21 = -z2 * n1 / nz
    P22 = 1.0 - z2 * n2 / nz
    # mp, rp as linear functions of (m, r): s1 = m + r, s2 = m - r
    am = 0.5 * (P11 + P12 + P21 + P22)
    ar = 0.5 * (P11 - P12 + P21 - P22)
    bm = 0.5 * (P11 + P12 - P21 - P22)
    br = 0.5 * (P11 - P12 - P21 + P22)
    r0 = np.sqrt(d * d + txy * txy)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv_r = np.where(r > 0.0, 1.0 / r, 0.0)
        inv_r0 = np.where(r0 > 0.0, 1.0 / r0, 0.0)
    # derivatives w.r.t. (sxx, syy, txy): dm = (1/2, 1/2, 0), dd = (1/2, -1/2, 0), dtau = (0,0,1)
    dm = np.array([0.5, 0.5, 0.0])
    ddv = np.array([0.5, -0.5, 0.0])
    dtau = np.array([0.0, 0.0, 1.0])
    dr = (d * inv_r)[:, None] * ddv[None, :] + (txy * inv_r)[:, None] * dtau[None, :]
    dphi = (sn2 * inv_r0)[:, None] * ddv[None, :] - (c2 * inv_r0)[:, None] * dtau[None, :]
    dmp = am * dm[None, :] + ar * dr
    drp = bm * dm[None, :] + br * dr
    Dsig = np.empty((M, 3, 3))
    Dsig[:, 0, :] = dmp + c2[:, None] * drp + (rp * sn2)[:, None] * dphi
    Dsig[:, 1, :] = dmp - c2[:, None] * drp - (rp * sn2)[:, None] * dphi
    Dsig[:, 2, :] = sn2[:, None] * drp - (rp * c2)[:, None] * dphi
    D_line = Dsig @ C  # d sigma / d eps_e
    D = np.where(elastic[:, None, None], C[None, :, :], np.where(line[:, None, None], D_line, 0.0))
    return psi, sig, D


def mohr_coulomb(u, B, qw, eps_p_old, E, nu, phi_deg, cohesion, reg=1e-12, elastic_only=False, want=("e", "g", "h")):
    """u: (N, 30); B: (N, nq, 3, 30); qw: (N, nq); eps_p_old: (N, nq, 3)."""
    N, nq = qw.shape
    strain = np.einsum("nqij,nj->nqi", B, u)
    eps_e = (strain - eps_p_old).reshape(N * nq, 3)
    psi, sig, D = mc_stress_tangent(eps_e, E, nu, phi_deg, cohesion, reg, elastic_only)
    out = {}
    if "e" in want:
        out["e"] = np.sum(qw * psi.reshape(N, nq), axis=1)
    if "g" in want:
        out["g"] = np.einsum("nq,nqij,nqi->nj", qw, B, sig.reshape(N, nq, 3))
    if "h" in want:
        Dq = D.reshape(N, nq, 3, 3) * qw[:, :, None, None]
        out["h"] = np.einsum("nqia,nqij,nqjb->nab", B, Dq, B)
    return out


def davis_b_reduction(c0, phi_deg, psi_deg, lam):
    """slope_stability/support/reduction.py:8-29 -> (reduced cohesion, reduced phi in degrees)."""
    if lam <= 0.0:
        raise ValueError("lambda must be positive")
    phi = np.radians(phi_deg)
    psi = np.radians(psi_deg)
    c01 = c0 / lam
    phi1 = np.arctan(np.tan(phi) / lam)
    psi1 = np.arctan(np.tan(psi) / lam)
    beta = np.cos(phi1) * np.cos(psi1) / (1.0 - np.sin(phi1) * np.sin(psi1))
    return float(beta * c01), float(np.degrees(np.arctan(beta * np.tan(phi1))))

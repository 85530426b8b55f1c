"""Shared helpers for the parity tests: seeded perturbed states and error norms."""
import numpy as np


def perturbed(setup, amp=1e-2, seed=0):
    """the set-up's state plus seeded noise on every evolved field, scaled by
    the local full density / total energy so density and pressure stay
    positive on stratified backgrounds"""
    rng = np.random.default_rng(seed)
    D = setup.ndim
    w = setup.w.copy()
    bg = setup.bg
    rho = w[0] + bg[1]
    etot = np.abs(w[1 + D] + bg[0])
    vs = np.sqrt(etot / rho)
    N = lambda: rng.standard_normal(setup.ncell)
    w[0] = w[0] + amp * rho * N()
    for d in range(D):
        w[1 + d] = w[1 + d] + amp * rho * vs * N()
        w[2 + D + d] = w[2 + D + d] + amp * np.sqrt(etot) * N()
    w[1 + D] = w[1 + D] + amp * etot * N()
    return w


def interior_mask(setup):
    n = setup.n
    m = np.zeros((n[2], n[1], n[0]), dtype=bool)
    if setup.ndim == 3:
        m[2:-2, 2:-2, 2:-2] = True
    else:
        m[:, 2:-2, 2:-2] = True
    return m.reshape(-1)


def rel_err(a, b, mask=None):
    """per-field max |a-b| / max |b| (L-infinity, normalised by the field's amplitude)"""
    if mask is not None:
        a = a[:, mask]
        b = b[:, mask]
    out = []
    for f in range(a.shape[0]):
        s = np.abs(b[f]).max()
        d = np.abs(a[f] - b[f]).max()
        out.append(0.0 if s == 0 and d == 0 else d / (s if s > 0 else 1.0))
    return np.array(out)


def rel_l2(a, b, mask=None):
    if mask is not None:
        a = a[:, mask]
        b = b[:, mask]
    out = []
    for f in range(a.shape[0]):
        s = np.sqrt(np.mean(b[f] ** 2))
        d = np.sqrt(np.mean((a[f] - b[f]) ** 2))
        out.append(0.0 if s == 0 and d == 0 else d / (s if s > 0 else 1.0))
    return np.array(out)

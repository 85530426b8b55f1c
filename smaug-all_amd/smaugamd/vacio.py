"""VAC/SAC ascii configuration files (the reference's .ini format,
dev/smaug/src/readwrite.c:982-1540; e.g. configs/zero1_BW_asc.ini).

line 1 name; line 2 ``it t ndim neqpar nw``; line 3 cells per dim (with
ghosts); line 4 eqpar (gamma eta g1..gD 0 0); line 5 variable names; then
one row per cell, i fastest: positions x_1..x_D then the NVAR fields
rho m_1..m_D e b_1..b_D e_b rho_b b_b1..b_bD.
"""
from __future__ import annotations

import numpy as np

from .configs import Setup


def write_ascii(path: str, s: Setup, it: int = 0, t: float = 0.0, name: str = "smaug"):
    D = s.ndim
    n = s.n[:D]
    eq = [s.gamma, 0.0] + list(s.g[:D]) + [0.0, 0.0]
    names = ("x y z"[: 2 * D - 1] + " h " + " ".join(f"m{d+1}" for d in range(D)) + " e " +
             " ".join(f"b{d+1}" for d in range(D)) + " eb rhob " +
             " ".join(f"bg{d+1}" for d in range(D)) + "   gamma eta   " +
             " ".join(f"grav{d+1}" for d in range(D)))
    coords = []
    shape = (s.n[2], s.n[1], s.n[0]) if D == 3 else (s.n[1], s.n[0])
    if D == 3:
        k, j, i = np.meshgrid(np.arange(s.n[2]), np.arange(s.n[1]), np.arange(s.n[0]), indexing="ij")
        idx = [i, j, k]
    else:
        j, i = np.meshgrid(np.arange(s.n[1]), np.arange(s.n[0]), indexing="ij")
        idx = [i, j]
    for d in range(D):
        coords.append(s.pos[d][idx[d]].reshape(-1))
    data = np.vstack(coords + [s.w, s.bg]).T
    with open(path, "w") as f:
        f.write(name + "\n")
        f.write(f"{it:7d} {t:23.16E} {D:2d} {len(eq):2d} {s.nve + s.nbg:2d}\n")
        f.write(" ".join(f"{v:4d}" for v in n) + "\n")
        f.write("".join(f" {v:23.16E}" for v in eq) + "\n")
        f.write(names + "\n")
        np.savetxt(f, data, fmt="%24.16E")


def read_ascii(path: str):
    """returns (header dict, positions (D, ncell), fields (NVAR, ncell))"""
    with open(path) as f:
        name = f.readline().rstrip("\n")
        it, t, ndim, neqpar, nw = f.readline().split()
        dims = [int(v) for v in f.readline().split()]
        eqpar = [float(v) for v in f.readline().split()]
        names = f.readline().split()
        data = np.loadtxt(f)
    D = int(ndim)
    hdr = {"name": name, "it": int(it), "t": float(t), "ndim": D, "neqpar": int(neqpar),
           "nw": int(nw), "dims": dims, "eqpar": eqpar, "names": names}
    data = np.atleast_2d(data)
    return hdr, data[:, :D].T.copy(), data[:, D:].T.copy()

"""The oracle restatement of the C driver's onestep loop with moddton
(dev/smaug/src/iosmaug.cpp:855-898 and the step that follows), used as the
reference by the moddton tests:

  - dt = courant / maxcourant when that is above 1e-8, where maxcourant is
    the maximum of cfast_d/dx_d along the LAST direction only: the reference's
    cucomputemaxcourant resets p->maxcourant on every call
    (computedervfields.cu:2154), so the driver's loop over directions keeps
    the last one;
  - from the second step on, dt is capped by the hyperviscous limit
    0.25 min dx^2 / numax (sm_dtvisc), numax being the largest coefficient
    the previous step's hyperdiffusion formed (p->maxviscoef);
  - with -tend the step is shortened to land on tend (SAC tmaxexact);
  - every term of the forward-Euler step takes that dt, except the energy
    field's own hyperdiffusion, which the reference passes the initial dt
    (iosmaug.cpp:1041), and the user source's clock p->qt, which advances
    by the initial dt before the step (iosmaug.cpp:918).
"""
import copy

import numpy as np

from oracle.oracle import Oracle


def onestep_moddton(s, w0, dt0, courant=0.3, nmax=10, tend=0.0, t0=0.0, dts_given=None):
    """returns (state, steps run, elapsed time, the dt of every step, the
    Courant dt and the hyperviscous cap of every step).  With dts_given the
    steps take those dt instead (a replay of another run's sequence); the
    limits are still evaluated on the replayed states."""
    o = Oracle(s)
    se = copy.copy(s)
    se.chyp = s.chyp.copy()
    se.chyp[1 + s.ndim] = 0.0  # the energy field's own hyperdiffusion off
    oe = Oracle(se)
    w = o.boundary(w0)
    dt, elapsed, qt = dt0, 0.0, t0
    numax_prev = 0.0
    dts, cour, visc = [], [], []
    n = 0
    if dts_given is not None:
        nmax = len(dts_given)
    for n in range(1, nmax + 1):
        mc = o.maxcourant_dim(w, s.ndim - 1)
        if courant / mc > 1e-8:
            dt = courant / mc
        cour.append(dt)
        dv = o.dtvisc(numax_prev) if n > 1 else 0.0
        visc.append(dv)
        if dv > 1e-8 and dt > dv:
            dt = dv
        if tend > 0 and elapsed + dt > tend:
            dt = tend - elapsed
        if dts_given is not None:
            dt = float(dts_given[n - 1])
        qt += dt0
        cm = o.cmax(w)
        r = o.rhs(w, t=qt, cmax=cm)
        if s.hyperdif:
            re = oe.rhs(w, t=qt, cmax=cm)
            numax_prev = o.numax(w, cmax=cm)
            w = w + dt * r + (dt0 - dt) * (r - re)
        else:
            w = w + dt * r
        w = o.boundary(w)
        elapsed += dt
        dts.append(dt)
        if tend > 0 and elapsed >= tend and dts_given is None:
            break
    return w, n, elapsed, np.array(dts), np.array(cour), np.array(visc)

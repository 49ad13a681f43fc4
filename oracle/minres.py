"""MINRES restatement (oracle, FP64).

The reference solves each Newton correction with ``scipy.sparse.linalg.minres``
(``small_strain_fft.py:281-299``: x0 = 0, ``rtol = linear_solver_tolerance``, no
preconditioner, no shift, default ``maxiter = 5 n``).  SciPy is a third-party
dependency (unpinned in ``pyproject.toml:19-33``; 1.15.3 here).  This restates the
Paige-Saunders recurrence of SciPy's ``_isolve/minres.py`` (the loop and stopping
tests, lines 203-321 of that file) for exactly that configuration, and is checked
iterate-for-iterate against SciPy in ``tests/test_oracle_minres.py``.
"""

import numpy as np

EPS = np.finfo(np.float64).eps


def minres(matvec, b, rtol=1e-5, maxiter=None, callback=None):
    """Return (x, itn, istop) for A x = b with A symmetric, x0 = 0."""
    b = np.asarray(b, float)
    n = b.size
    if maxiter is None:
        maxiter = 5 * n
    x = np.zeros(n)
    r1 = b.copy()
    y = r1
    beta1 = float(np.dot(r1, y))
    if beta1 == 0.0:
        return x, 0, 0
    if np.linalg.norm(b) == 0.0:
        return b.copy(), 0, 0
    beta1 = np.sqrt(beta1)

    oldb = 0.0
    beta = beta1
    dbar = 0.0
    epsln = 0.0
    phibar = beta1
    rhs1 = beta1
    rhs2 = 0.0
    tnorm2 = 0.0
    gmax = 0.0
    gmin = np.finfo(np.float64).max
    cs = -1.0
    sn = 0.0
    w = np.zeros(n)
    w2 = np.zeros(n)
    r2 = r1
    istop = 0
    itn = 0
    while itn < maxiter:
        itn += 1
        v = (1.0 / beta) * y
        y = matvec(v)
        if itn >= 2:
            y = y - (beta / oldb) * r1
        alfa = float(np.dot(v, y))
        y = y - (alfa / beta) * r2
        r1 = r2
        r2 = y
        oldb = beta
        beta = float(np.dot(r2, y))
        if beta < 0:
            raise ValueError("non-symmetric matrix")
        beta = np.sqrt(beta)
        tnorm2 += alfa**2 + oldb**2 + beta**2
        if itn == 1 and beta / beta1 <= 10 * EPS:
            istop = -1

        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        root = np.linalg.norm([gbar, dbar])
        gamma = max(np.linalg.norm([gbar, beta]), EPS)
        cs = gbar / gamma
        sn = beta / gamma
        phi = cs * phibar
        phibar = sn * phibar

        denom = 1.0 / gamma
        w1 = w2
        w2 = w
        w = (v - oldeps * w1 - delta * w2) * denom
        x = x + phi * w

        gmax = max(gmax, gamma)
        gmin = min(gmin, gamma)
        z = rhs1 / gamma
        rhs1 = rhs2 - delta * z
        rhs2 = -epsln * z

        anorm = np.sqrt(tnorm2)
        ynorm = np.linalg.norm(x)
        epsx = anorm * ynorm * EPS
        rnorm = phibar
        test1 = np.inf if (ynorm == 0 or anorm == 0) else rnorm / (anorm * ynorm)
        test2 = np.inf if anorm == 0 else root / anorm
        acond = gmax / gmin
        if istop == 0:
            if 1 + test2 <= 1:
                istop = 2
            if 1 + test1 <= 1:
                istop = 1
            if itn >= maxiter:
                istop = 6
            if acond >= 0.1 / EPS:
                istop = 4
            if epsx >= beta1:
                istop = 3
            if test2 <= rtol:
                istop = 2
            if test1 <= rtol:
                istop = 1
        if callback is not None:
            callback(x)
        if istop != 0:
            break
    return x, itn, istop

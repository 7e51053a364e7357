"""COCG, restarted GMRES and the IterativeMethod wrapper (oracle; test infrastructure only).

`apply` and `precondition` are callables on complex128 vectors (the reference's convertToHandle,
src/+rooftopMoM/+solver/+utils/convertToHandle.m:10-16, accepts a matrix, a handle or an object with apply).
"""

import numpy as np


def as_handle(operator):
    """Matrix, callable or object with .apply -> callable. +solver/+utils/convertToHandle.m:10-16."""
    if operator is None:
        return None
    if callable(operator) and not hasattr(operator, 'apply'):
        return operator
    if isinstance(operator, np.ndarray):
        return lambda x: operator @ x
    return operator.apply


def cocg(apply, b, x0, tol, max_iter, precondition=None):
    """COCG (complex-symmetric CG, unconjugated transposes), optional preconditioner.

    Follows src/+rooftopMoM/+solver/ConjugateGradient.m:32-59 and PreconditionedConjugateGradient.m:11-23. The monitored
    norm is the conjugated 2-norm of the recursive residual over norm(b); a non-finite norm stops the loop.
    Returns (x, converged, residual_norms). Pinned: backslash on the synthetic symmetric systems at 1e-9, iteration
    records, breakdown on diag([1, -1]) (ConjugateGradientTest.m:5-72).
    """
    b = np.asarray(b, dtype=np.complex128)
    bnorm = np.linalg.norm(b)
    if x0 is None:
        x = np.zeros_like(b)
        r = b.copy()
    else:
        x = np.array(x0, dtype=np.complex128)
        r = b - apply(x)
    norms = [np.linalg.norm(r) / bnorm]
    converged = norms[0] < tol
    it = 0
    d = None
    rho = None
    while not converged and it < max_iter:
        it += 1
        z = r if precondition is None else precondition(r)
        rho_next = r @ z
        d = z.copy() if it == 1 else z + (rho_next / rho) * d
        rho = rho_next
        ad = apply(d)
        with np.errstate(divide='ignore', invalid='ignore', over='ignore'):
            alpha = rho / (d @ ad)
            x = x + alpha * d
            r = r - alpha * ad
            norms.append(np.linalg.norm(r) / bnorm)
        converged = norms[-1] < tol
        if not np.isfinite(norms[-1]):
            break
    return x, bool(converged), np.array(norms)


def givens(a, b):
    """Complex Givens [conj(c), conj(s); -s, c] mapping [a; b] to [norm; 0].
    @GeneralizedMinimumResidual/private/generateRotation_.m:11-18."""
    radius = np.linalg.norm([a, b])
    if radius > 0:
        return a / radius, b / radius
    return 1.0 + 0j, 0.0 + 0j


def cgs2(v, basis, count):
    """Classical Gram-Schmidt run twice against the first `count` columns.
    @GeneralizedMinimumResidual/private/orthogonalize_.m:15-23."""
    q = basis[:, :count]
    c = q.conj().T @ v
    v = v - q @ c
    corr = q.conj().T @ v
    v = v - q @ corr
    return v, c + corr


def gmres(apply, b, x0, tol, max_iter, restart, precondition=None):
    """Restarted GMRES, left preconditioned, CGS2, complex Givens; restarts from the true preconditioned residual.

    Follows src/+rooftopMoM/+solver/@GeneralizedMinimumResidual/iterate_.m:1-85. Monitors |g_{j+1}| / norm(P b).
    Returns (x, converged, residual_norms). Pinned: backslash on the synthetic non-symmetric systems at 1e-9, the
    iteration limit, restart capping and the preconditioned residual record (GeneralizedMinimumResidualTest.m:5-65).
    """
    P = (lambda v: v) if precondition is None else precondition
    b = np.asarray(b, dtype=np.complex128)
    n = b.size
    restart = n if restart is None or not np.isfinite(restart) else int(restart)
    m = int(min(restart, n, max_iter))
    reference = np.linalg.norm(P(b))
    if x0 is None:
        x = np.zeros(n, dtype=np.complex128)
        r = P(b)
    else:
        x = np.array(x0, dtype=np.complex128)
        r = P(b - apply(x))
    rnorm = np.linalg.norm(r)
    norms = [rnorm / reference]
    converged = norms[0] < tol
    it = 0
    basis = np.zeros((n, m + 1), dtype=np.complex128)
    H = np.zeros((m + 1, m), dtype=np.complex128)
    cs = np.zeros(m, dtype=np.complex128)
    sn = np.zeros(m, dtype=np.complex128)
    while not converged and it < max_iter:
        basis[:, 0] = r / rnorm
        g = np.zeros(m + 1, dtype=np.complex128)
        g[0] = rnorm
        step = 0
        while not converged and step < m and it < max_iter:
            step += 1
            it += 1
            v, coeffs = cgs2(P(apply(basis[:, step - 1])), basis, step)
            H[:step, step - 1] = coeffs
            H[step, step - 1] = np.linalg.norm(v)
            for k in range(step - 1):
                h0, h1 = H[k, step - 1], H[k + 1, step - 1]
                H[k, step - 1] = np.conj(cs[k]) * h0 + np.conj(sn[k]) * h1
                H[k + 1, step - 1] = -sn[k] * h0 + cs[k] * h1
            cs[step - 1], sn[step - 1] = givens(H[step - 1, step - 1], H[step, step - 1])
            sub = H[step, step - 1]
            H[step - 1, step - 1] = np.conj(cs[step - 1]) * H[step - 1, step - 1] + np.conj(sn[step - 1]) * sub
            H[step, step - 1] = 0
            g[step] = -sn[step - 1] * g[step - 1]
            g[step - 1] = np.conj(cs[step - 1]) * g[step - 1]
            norms.append(abs(g[step]) / reference)
            converged = norms[-1] < tol
            if not converged and sub != 0:
                basis[:, step] = v / sub
        y = np.linalg.solve(np.triu(H[:step, :step]), g[:step])
        x = x + basis[:, :step] @ y
        if not converged and it < max_iter:
            r = P(b - apply(x))
            rnorm = np.linalg.norm(r)
    return x, bool(converged), np.array(norms)


def solve(method, operator, b, x0=None, tol=1e-6, max_iter=1000, restart=50, preconditioner=None):
    """IterativeMethod.solve: zero right-hand side -> zero solution; otherwise iterate and report the true residual.

    Follows src/+rooftopMoM/+solver/IterativeMethod.m:27-70 and Solver.m:42-120 (method 'cg' -> COCG, 'gmres' -> GMRES).
    Returns a dict with solution, converged, iterationCount, residualNorms, residualNorm.
    Pinned: true residual at 1e-12, zero RHS, exact initial guess, non-convergence (IterativeMethodTest.m:17-91).
    """
    apply = as_handle(operator)
    precondition = as_handle(preconditioner)
    b = np.asarray(b, dtype=np.complex128)
    if np.any(b != 0):
        if method == 'cg':
            x, conv, norms = cocg(apply, b, x0, tol, max_iter, precondition)
        elif method == 'gmres':
            x, conv, norms = gmres(apply, b, x0, tol, max_iter, restart, precondition)
        else:
            raise ValueError('unknown method %r' % method)
        true = np.linalg.norm(b - apply(x)) / np.linalg.norm(b)
    else:
        x, conv, norms, true = np.zeros_like(b), True, np.array([0.0]), 0.0
    return {'solution': x, 'converged': conv, 'iterationCount': norms.size - 1, 'residualNorms': norms,
            'residualNorm': true}

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


def block_orth(W, tau):
    """A basis of the columns of W (n x s): the Cholesky factor L of the Gram matrix W^H W, taken column by column,
    keeps column j when what is left of it after the earlier kept columns exceeds tau^2 of its own squared norm (the
    rank-revealing drop). Returns (Q, C, T, kept): Q = W C with C = L^-H on the kept columns (zero elsewhere), and the
    upper triangular T, T[i, j] = conj(L[j, i]) for kept i and j >= i (zero in the rows of dropped i), so that W = Q T
    up to the dropped parts. Pinned: Q^H Q = I on the kept columns, W = Q T and W in the span of Q, on full-rank and
    exactly rank-deficient blocks."""
    s = W.shape[1]
    G = W.conj().T @ W
    L = np.zeros((s, s), dtype=np.complex128)
    kept = np.zeros(s, dtype=bool)
    for j in range(s):
        d = G[j, j].real - np.sum(np.abs(L[j, :j]) ** 2)
        if np.isfinite(d) and d > tau * tau * G[j, j].real:
            kept[j] = True
            L[j, j] = np.sqrt(d)
            for i in range(j + 1, s):
                L[i, j] = (G[i, j] - np.sum(L[i, :j] * L[j, :j].conj())) / L[j, j]
    C = np.zeros((s, s), dtype=np.complex128)
    k = np.flatnonzero(kept)
    if k.size:
        C[np.ix_(k, k)] = np.linalg.inv(L[np.ix_(k, k)].conj().T)
    T = np.triu(L.conj().T)
    T[~kept, :] = 0
    return W @ C, C, T, kept


BLOCK_DROP = 1e-6   # the rank-revealing threshold of the residual and direction blocks


def block_cocg(apply, B, tol, max_iter, precondition=None, tau=BLOCK_DROP):
    """Block COCG on the columns of B (n x s): block conjugate gradients with the unconjugated products of COCG
    (ConjugateGradient.m:32-59 per column). The residual block is carried as an orthonormal basis and coordinates,
    R = Q S (the residual orthonormalization of Dubrulle's retooled block CG), and the direction block P is an
    orthonormal basis of Z + P beta (the breakdown-free block CG of Ji and Li, 2017):
        A P, D = P^T A P, a = D^-1 P^T Q, X += P a S, W = Q - A P a = Q' T (block_orth), S' = T S,
        Z = M Q', beta = -D^-1 (A P)^T Z, P' = orth(Z + P beta).
    These are the iterates of block COCG (a block Krylov method depends on its blocks only through their spans) with
    every coefficient matrix formed from orthonormal blocks, so residuals that converge at different rates, or nearly
    dependent right-hand sides, drop directions instead of making a coefficient matrix singular. The monitored norm of
    column j is norm(W S_j) / norm(b_j); column j stops (solution and record frozen) below tol or at max_iter. A
    singular D (on the kept directions) ends every running column with a NaN norm, as COCG's breakdown
    (ConjugateGradient.m:56-59); an empty direction block ends them unconverged.
    No reference counterpart (the reference solves one right-hand side at a time); pinned by: one column is COCG,
    finite termination within ceil(n / s) iterations on a well-conditioned system, backslash, and block_orth.
    Returns (X, converged per column, iterations per column, monitored norms per column).
    """
    B = np.asarray(B, dtype=np.complex128)
    n, s = B.shape
    apply_block = lambda V: np.stack([apply(V[:, j]) for j in range(V.shape[1])], axis=1)
    prec_block = (lambda V: V.copy()) if precondition is None else (
        lambda V: np.stack([precondition(V[:, j]) for j in range(V.shape[1])], axis=1))
    X = np.zeros_like(B)
    R = B.copy()
    bnorm = np.linalg.norm(B, axis=0)
    norms = [[np.linalg.norm(R[:, j]) / bnorm[j] if bnorm[j] > 0 else 0.0] for j in range(s)]
    conv = np.array([bnorm[j] == 0 or norms[j][0] < tol for j in range(s)])
    R[:, conv] = 0
    running = ~conv
    iters = np.zeros(s, dtype=int)
    Q, _, S, _ = block_orth(R, tau)
    P, _, _, kept = block_orth(prec_block(Q), tau)
    it = 0
    while running.any() and it < max_iter:
        if not kept.any():
            running[:] = False
            break
        it += 1
        AP = apply_block(P)
        D = P.T @ AP
        k = np.flatnonzero(kept)
        Dk = D[np.ix_(k, k)]
        if not np.all(np.isfinite(Dk)) or np.linalg.matrix_rank(Dk, tol=0) < k.size or \
                np.linalg.cond(Dk) > 1 / np.finfo(float).eps:
            for j in np.flatnonzero(running):
                norms[j].append(np.nan)
                iters[j] = it
            running[:] = False
            break
        a = np.zeros((s, s), dtype=np.complex128)
        a[k] = np.linalg.solve(Dk, (P.T @ Q)[k])
        X[:, running] += (P @ (a @ S))[:, running]
        W = Q - AP @ a
        for j in np.flatnonzero(running):
            norms[j].append(np.linalg.norm(W @ S[:, j]) / bnorm[j])
            iters[j] = it
            if norms[j][-1] < tol:
                conv[j] = True
                running[j] = False
            elif not np.isfinite(norms[j][-1]) or it >= max_iter:
                running[j] = False
        Q, _, T, _ = block_orth(W, tau)
        S = T @ S
        if not running.any():
            break
        Z = prec_block(Q)
        beta = np.zeros((s, s), dtype=np.complex128)
        beta[k] = -np.linalg.solve(Dk, (AP.T @ Z)[k])
        P, _, _, kept = block_orth(Z + P @ beta, tau)
    return X, conv, iters, [np.array(v) for v in norms]


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

"""OrthonormalizationMultigrid restated in numpy (test infrastructure only).

Follows L3-solution/BoSSS.Solution.AdvancedSolvers/OrthonormalizationMultigrid.cs:
* Solve (:797-1080): per iteration the pre-smoother's correction, the coarse-grid
  correction (m_omega times), NoOfPostSmootherSweeps post-smoother corrections; the
  termination criterion is checked at the start of an iteration and after every
  correction; a cancellation in a post-smoother correction skips the next
  pre-smoother (:1047-1055);
* AddSol (:219-340): A c normalised, orthogonalised against the history, severe
  cancellation (|A c| <= 1e-5) restarts from the rescaled remainder (a random
  vector if nothing is left, FillXwithRandom :224-240), at most 10 re-cycles;
* MinimizeResidual (:149-215): alpha = <A X_i, r_0>, x += alpha X_i,
  r -= alpha A X_i, the residual recomputed (RecalcResiPeriod = 1), a residual
  increase beyond max(0.2 max(|r|, |r_old|), 1e-7) raises;
* termination as IterativeSolverConfig.DefaultTermination
  (AS/IterativeSolverConfig.cs:48-58).
The device solver orthogonalises by classical Gram-Schmidt in two passes and
recomputes the residual as b - A x; so does this restatement (the reference uses
one modified Gram-Schmidt pass and Res0 + A (Sol0 - x): equal in exact arithmetic).
The random direction is the same counter hash as the device's (splitmix64
finaliser), so a restart is reproducible.
"""
import numpy as np


class OrthoArithmeticError(ArithmeticError):
    pass


def random_direction(n, k):
    """x_i in [0, 1) from splitmix64 of (seed_k + i * golden), as the device fill."""
    M = (1 << 64) - 1
    seed = (0x5DEECE66D + 7919 * k) & M
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def keep_running(it, r0, r, tol, maxit, minit=0, eps=2.2204460492503131e-16 * 1000):
    """(keep going, success) as the reference's DefaultTermination."""
    if it <= minit:
        return True, False
    if r < r0 * tol + tol:
        return False, True
    if it > maxit:
        return False, False
    if r0 < eps or r < eps:
        return False, True
    return True, False


class OrthoMG:
    """apply(x) -> A x; pre(r), coarse(r), post(r), learned(r) -> corrections (any may be None)."""

    def __init__(self, apply, pre=None, coarse=None, post=None, post_sweeps=1, coarse_reps=1, learned=None,
                 tol=1e-10, maxit=100, max_hist=128, trace=False):
        self.apply, self.pre, self.coarse, self.post = apply, pre, coarse, post
        self.post_sweeps, self.coarse_reps, self.learned = post_sweeps, coarse_reps, learned
        self.tol, self.maxit, self.max_hist = tol, maxit, max_hist
        self.trace = trace
        self.n_random = self.n_cancel = self.n_restart = self.n_added = 0

    def _random(self, n):
        self.n_random += 1
        x = random_direction(n, self.n_random)
        q = self.apply(x)
        nq = np.linalg.norm(q)
        if not (nq > 0 and np.isfinite(nq)):
            raise OrthoArithmeticError("numerical breakdown with a random vector")
        return x, q, nq

    def _add_sol(self, c):
        c = np.array(c, dtype=np.float64)
        q = self.apply(c)
        nq = np.linalg.norm(q)
        if not np.isfinite(nq):
            raise OrthoArithmeticError("numerical breakdown")
        if nq == 0:
            c, q, nq = self._random(len(c))
        c, q = c / nq, q / nq
        cancel = False
        self.used_more = False
        for jj in range(11):
            self.used_more = jj > 0
            if self.X:
                Xh, Qh = np.array(self.X), np.array(self.Q)
                for _ in range(2):
                    beta = Qh @ q
                    c = c - beta @ Xh
                    q = q - beta @ Qh
            nn = np.linalg.norm(q)
            if nn <= 1e-5:
                cancel = True
                self.n_cancel += 1
                if jj == 10:
                    raise OrthoArithmeticError("numerical breakdown")
                xn = np.linalg.norm(c)
                if xn < 1e-200 or jj == 9:
                    c, q, nm = self._random(len(c))
                else:
                    c = c / xn
                    q = self.apply(c)
                    nm = np.linalg.norm(q)
                    if nm == 0:
                        c, q, nm = self._random(len(c))
                c, q = c / nm, q / nm
            else:
                c, q = c / nn, q / nn
                break
        if len(self.X) == self.max_hist:
            self.n_restart += 1
            self.sol0, self.res0 = self.x.copy(), self.r.copy()
            self.X, self.Q = [], []
        self.X.append(c)
        self.Q.append(q)
        self.n_added += 1
        return cancel

    def _minimize(self, b):
        old = np.linalg.norm(self.r)
        alpha = self.Q[-1] @ self.res0
        self.x = self.x + alpha * self.X[-1]
        self.r = b - self.apply(self.x)
        rn = np.linalg.norm(self.r)
        tol = max(max(rn, old) * 0.2, 1e-7)
        if not rn <= old + tol:
            raise OrthoArithmeticError(f"residual increase: {old} -> {rn}")
        self.alphas.append(alpha)
        self.resnorms.append(rn)
        return rn

    def _correct(self, c, b, kind):
        if self.trace:
            self.raw.append((kind, np.array(c, dtype=np.float64)))
        cancel = self._add_sol(c)
        return self._minimize(b), cancel

    def solve(self, b, x0=None):
        b = np.asarray(b, np.float64)
        self.x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
        self.sol0 = self.x.copy()
        self.res0 = b - self.apply(self.x)
        self.r = self.res0.copy()
        self.X, self.Q, self.alphas, self.resnorms, self.raw = [], [], [], [], []
        r0 = rn = np.linalg.norm(self.r)
        skip_pre, ok, completed = False, False, 0
        it = 1
        while True:
            go, ok = keep_running(it, r0, rn, self.tol, self.maxit)
            if not go:
                break
            completed = it
            if self.pre is not None and not skip_pre:
                rn, _ = self._correct(self.pre(self.r), b, "pre")
                go, ok = keep_running(it, r0, rn, self.tol, self.maxit)
                if not go:
                    break
            skip_pre = False
            if self.coarse is not None:
                for _ in range(self.coarse_reps):
                    rn, _ = self._correct(self.coarse(self.r), b, "coarse")
            go, ok = keep_running(it, r0, rn, self.tol, self.maxit)
            if not go:
                break
            stop = False
            if self.post is not None:
                for _ in range(self.post_sweeps):
                    rn, cancel = self._correct(self.post(self.r), b, "post")
                    go, ok = keep_running(it, r0, rn, self.tol, self.maxit)
                    if not go:
                        stop = True
                        break
                    if cancel:
                        skip_pre = True
                        break
            if stop:
                break
            if self.learned is not None:
                rn, _ = self._correct(self.learned(self.r), b, "learned")
            it += 1
        self.iterations, self.converged, self.res0_norm, self.res_norm = completed, ok, r0, rn
        return self.x

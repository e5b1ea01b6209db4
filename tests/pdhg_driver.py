"""The reference's PDHG host loop (runPdhg, src/pdhg/pdhg.cpp:483-687)
restated over any object with the Backend op set (src/pdhg/pdhg.cpp:51-70):
the oracle's serial CpuBackend (oracle_lib.OracleOps) or the device backend
(pdhg_amd.Backend).  It is how a caller that keeps the reference's own host
loop drives the op-level C-ABI, and it lets the tests compare the two backends
op by op and over whole solves.  Test infrastructure only."""
import math

import numpy as np

X, XP, X0, ATY, ATYP, C, LO, UP, Y, YP, Y0, AX, AXP, AX0, RL, RU = range(16)
INF_BOUND = 1e20  # src/util/common.h:15


def scaled_csr_pair(lp, pr):
    """CSR of the scaled A (from the prepared set-up) and of its transpose,
    the latter in the CSC order of A (csrOfTranspose, src/pdhg/pdhg.cpp:35-43)."""
    import scipy.sparse as sp

    cols = np.repeat(np.arange(lp.n), np.diff(lp.start))
    tval = lp.value * (pr["R"][lp.index] * pr["C"][cols])  # applyScaling, src/lp/scaling.cpp:28-32
    A = sp.csr_matrix((pr["val"], pr["idx"], pr["ptr"]), shape=(lp.m, lp.n))
    AT = sp.csr_matrix((tval, lp.index.copy(), lp.start.copy()), shape=(lp.n, lp.m))
    return A, AT


def case_data(lp, pr, costs):
    """Scaled k-interleaved case vectors, norms of the originals and initial
    primal weights (src/pdhg/pdhg.cpp:461-545)."""
    k = costs.shape[0]
    fin = lambda v: np.abs(v) < INF_BOUND  # noqa: E731
    R, Cs, beta, gamma = pr["R"], pr["C"], pr["beta"], pr["gamma"]
    c = (costs * Cs[None, :] / gamma).T.copy()
    lo = np.where(fin(lp.l), lp.l / Cs / beta, -np.inf)[:, None].repeat(k, 1)
    up = np.where(fin(lp.u), lp.u / Cs / beta, np.inf)[:, None].repeat(k, 1)
    rl = np.where(fin(lp.L), lp.L * R / beta, -np.inf)[:, None].repeat(k, 1)
    ru = np.where(fin(lp.U), lp.U * R / beta, np.inf)[:, None].repeat(k, 1)
    x = np.minimum(np.maximum(0.0, lo), up)
    omega = np.ones(k)
    cnorm, rnorm = np.zeros(k), np.zeros(k)
    bn0 = 0.0  # serial sums in the reference's order, so the norms match bit for bit
    for lv, uv in zip(lp.L.tolist(), lp.U.tolist()):
        if abs(lv) < INF_BOUND:
            bn0 += lv * lv
        if abs(uv) < INF_BOUND and uv != lv:
            bn0 += uv * uv
    for b in range(k):
        cn = 0.0
        for v in c[:, b].tolist():
            cn += v * v
        bn = 0.0
        for lv, uv in zip(rl[:, b].tolist(), ru[:, b].tolist()):
            if abs(lv) < INF_BOUND:
                bn += lv * lv
            if abs(uv) < INF_BOUND and uv != lv:
                bn += uv * uv
        if cn > 1e-20 and bn > 1e-20:
            omega[b] = math.sqrt(cn) / math.sqrt(bn)
        c0 = 0.0
        for v in costs[b].tolist():
            c0 += v * v
        cnorm[b] = math.sqrt(c0)
        rnorm[b] = math.sqrt(bn0)
    return dict(c=c, lo=lo, up=up, rl=rl, ru=ru, x=x, omega=omega, cnorm=cnorm, rnorm=rnorm)


def run_pdhg(be, lp, pr, costs, tol=1e-6, max_iter=1_000_000, check_every=64, rho=1.0):
    """runPdhg's loop (src/pdhg/pdhg.cpp:506-687) over backend `be`; returns a
    result dict per case (status 1 optimal / 6 iteration limit, iterations,
    objective, relative_kkt, restarts, primal_weight, x, y)."""
    k = costs.shape[0]
    n, m = lp.n, lp.m
    d = case_data(lp, pr, costs)
    for v, a in ((C, d["c"]), (LO, d["lo"]), (UP, d["up"]), (RL, d["rl"]), (RU, d["ru"]), (X, d["x"]),
                 (X0, d["x"])):
        be.set(v, a)
    be.set(Y, np.zeros((m, k)))
    be.set(Y0, np.zeros((m, k)))
    be.spmv(X, AX)
    be.set(AX0, be.get(AX))
    omega = d["omega"].copy()
    eta = 0.998 / pr["normA"]
    be.set_steps(eta / omega, eta * omega)
    res = [None] * k
    done = np.zeros(k, bool)
    fp_r, fp_prev = np.full(k, np.inf), np.full(k, np.inf)
    since = np.zeros(k, np.int64)
    restarts = np.zeros(k, int)
    scale_obj = pr["beta"] * pr["gamma"]

    def extract(b, status, it, kkt, pobj, dobj):
        xs, ys = be.get(XP).reshape(n, k), be.get(YP).reshape(m, k)
        res[b] = dict(status=status, iterations=it, relative_kkt=kkt, objective=pobj + lp.offset,
                      dual_objective=dobj + lp.offset, restarts=int(restarts[b]), primal_weight=float(omega[b]),
                      x=xs[:, b] * pr["C"] * pr["beta"], y=ys[:, b] * pr["R"] * pr["gamma"])

    def measures(b, r3, c3):
        pobj = c3[3 * b + 1] * scale_obj
        dobj = (r3[3 * b + 1] + c3[3 * b + 2]) * scale_obj
        relP = math.sqrt(r3[3 * b]) / (1.0 + d["rnorm"][b])
        relD = math.sqrt(c3[3 * b]) / (1.0 + d["cnorm"][b])
        gap = abs(pobj - dobj) / (1.0 + abs(pobj) + abs(dobj))
        return pobj, dobj, relP, relD, gap

    it = 0
    remaining = k
    while it < max_iter:
        be.spmvT(Y, ATY)
        be.primal_step()
        be.spmv(XP, AXP)
        be.dual_step()
        mask = None
        if (it + 1) % check_every == 0 or it + 1 == max_iter:
            be.spmvT(YP, ATYP)
            r3, c3 = be.kkt()
            dx2, dy2 = be.diff2(X, XP), be.diff2(Y, YP)
            mask = np.zeros(k, np.int32)
            for b in range(k):
                if done[b]:
                    continue
                pobj, dobj, relP, relD, gap = measures(b, r3, c3)
                if relP <= tol and relD <= tol and gap <= tol:
                    extract(b, 1, it, max(relP, relD, gap), pobj, dobj)
                    done[b] = True
                    remaining -= 1
                    continue
                since[b] += check_every
                fp = math.sqrt(omega[b] * dx2[b] + dy2[b] / omega[b])
                if fp_r[b] == np.inf:
                    fp_r[b] = fp
                if fp <= 0.2 * fp_r[b] or (fp <= 0.8 * fp_r[b] and fp > fp_prev[b]) or \
                        since[b] >= int(0.36 * (it + 1)):
                    mask[b] = 1
                fp_prev[b] = fp
            if mask.any():
                ax2, ay2 = be.diff2(XP, X0), be.diff2(YP, Y0)
                for b in np.nonzero(mask)[0]:
                    dxn, dyn = math.sqrt(ax2[b]), math.sqrt(ay2[b])
                    if dxn > 1e-10 and dyn > 1e-10:
                        omega[b] = math.exp(0.5 * math.log(dyn / dxn) + 0.5 * math.log(omega[b]))
                    fp_r[b] = math.sqrt(omega[b] * dx2[b] + dy2[b] / omega[b])
                    fp_prev[b] = np.inf
                    since[b] = 0
                    restarts[b] += 1
                be.set_steps(eta / omega, eta * omega)
            if remaining == 0:
                break
        be.halpern(rho)
        if mask is not None and mask.any():
            be.restart(mask)
        it += 1
    if remaining > 0:
        be.spmvT(YP, ATYP)
        r3, c3 = be.kkt()
        for b in range(k):
            if not done[b]:
                pobj, dobj, relP, relD, gap = measures(b, r3, c3)
                extract(b, 6, it, max(relP, relD, gap), pobj, dobj)
    return res

"""CPU model of the fused forward kernel's indexing (tests only).

Executes, in numpy float64, the algorithm of `fused_forward_kernel`
(jaxley-pytorch_amd/csrc/jxhip_fused.hip) on the exact descriptor arrays that
`jaxhip.fused.ForwardLaunch` builds (programs, topology, per-node tables, program
layout). It lets the CPU test suite check the host-side tables against the oracle
without a GPU; GPU tests then check the kernel itself.
"""
import numpy as np
import torch


def _np(t):
    return t.detach().cpu().numpy()


def _xoe(x, y):
    z = min(x / y, 20.0)
    if abs(z) < 1e-4:
        return y * (1 - z / 2 + z * z / 12)
    return x / np.expm1(z)


def _se(x):
    return np.exp(min(x, 20.0))


def _gate(x, dt, a, b):
    tau = 1 / (a + b)
    xinf = a * tau
    e = _se(-dt / tau)
    return x * e + xinf * (1 - e)


def run_forward_model(run):
    """Fill `run.rec`, `run.v` and the gates as the kernel would (in place, CPU copy)."""
    from jaxhip import _lib

    progs = _np(run.progs)
    topo = _np(run.topos)[0]
    n, H = int(topo[0]), int(topo[1])
    order = _np(run.st.order)
    lptr = _np(run.st.level_ptr)
    ni = _np(run.st.node_i)
    nf = _np(run.node_f).astype(np.float64)
    cidx = _np(run.st.child_idx)
    rec_node = _np(run.st.rec_node)
    n_rec = run.st.n_rec
    a = run.args
    S = run.S
    dt = a.dt
    params = {}
    from jaxhip.fused import ALL_PARAMS

    for j, name in enumerate(ALL_PARAMS):
        mode, t = run.params.get(name, (0, None))
        params[j] = (mode, None if t is None else _np(t).astype(np.float64))
    v = _np(run.v).astype(np.float64).reshape(-1)
    gm = _np(run.gm).astype(np.float64).reshape(-1)
    gh = _np(run.gh).astype(np.float64).reshape(-1)
    gn = _np(run.gn).astype(np.float64).reshape(-1)
    rec = np.zeros(run.rec.numel())
    stim = None if run.stim is None else _np(run.stim).astype(np.float64)
    stim_scale = None if run.stim_scale is None else _np(run.stim_scale).astype(np.float64)
    stim_col = _np(run.st.stim_col)

    def P(j, pg, k, s):
        mode, arr = params[j]
        if mode == 0:
            return arr[k]
        if mode == 1:
            return arr[pg[5] + s]
        return arr[pg[4] + k * S + s]

    for pg in progs:
        nsim, t0, t1, soff, simoff, roff = pg[1], pg[2], pg[3], pg[4], pg[5], pg[6]
        xs = v[soff:soff + n * S].copy()
        ds = np.zeros(n * S)
        for i in range(n_rec * S):
            r, s = divmod(i, S)
            rec[roff + i] = xs[rec_node[r] * S + s]
        for step in range(t1 - t0):
            t = t0 + step
            for i in range(n * S):
                k, s = divmod(i, S)
                flags = ni[k, 3]
                dbase = nf[k, 2]
                if not (flags & 1) or s >= nsim:
                    ds[i], xs[i] = dbase, 0.0
                    continue
                vv = xs[i]
                lin = cst = 0.0
                if flags & 2:
                    am = 0.1 * _xoe(-(vv + 40), 10)
                    bm = 4.0 * _se(-(vv + 65) / 18)
                    ah = 0.07 * _se(-(vv + 65) / 20)
                    bh = 1.0 / (_se(-(vv + 35) / 10) + 1)
                    an = 0.01 * _xoe(-(vv + 55), 10)
                    bn = 0.125 * _se(-(vv + 65) / 80)
                    gi = soff + i
                    gm[gi] = _gate(gm[gi], dt, am, bm)
                    gh[gi] = _gate(gh[gi], dt, ah, bh)
                    gn[gi] = _gate(gn[gi], dt, an, bn)
                    gNa = P(0, pg, k, s) * gm[gi] ** 3 * gh[gi]
                    gK = P(1, pg, k, s) * gn[gi] ** 4
                    gL = P(2, pg, k, s)
                    lin += 1000 * (gNa + gK + gL)
                    cst += 1000 * (gNa * P(3, pg, k, s) + gK * P(4, pg, k, s) + gL * P(5, pg, k, s))
                if flags & 4:
                    gl = P(6, pg, k, s)
                    lin += 1000 * gl
                    cst += 1000 * gl * P(7, pg, k, s)
                rhs = vv + dt * cst * nf[k, 3]
                for e in range(ni[k, 4], ni[k, 5]):
                    col = stim_col[e]
                    sb = stim.shape[2]
                    bidx = simoff + s if sb > 1 else 0
                    rhs += stim[t, col, bidx] * stim_scale[e]
                ds[i] = dbase + dt * lin * nf[k, 3]
                xs[i] = rhs
            for h in range(1, H + 1):
                for pos in range(lptr[h], lptr[h + 1]):
                    k = order[pos]
                    for s in range(S):
                        dk, xk = ds[k * S + s], xs[k * S + s]
                        for j in range(ni[k, 1], ni[k, 2]):
                            c = cidx[j]
                            f = nf[c, 1] / ds[c * S + s]
                            dk -= nf[c, 0] * f
                            xk -= xs[c * S + s] * f
                        ds[k * S + s], xs[k * S + s] = dk, xk
            for h in range(H, -1, -1):
                for pos in range(lptr[h], lptr[h + 1]):
                    k = order[pos]
                    p = ni[k, 0]
                    for s in range(S):
                        xk = xs[k * S + s]
                        if p >= 0:
                            xk -= nf[k, 0] * xs[p * S + s]
                        xs[k * S + s] = xk / ds[k * S + s]
            for i in range(n_rec * S):
                r, s = divmod(i, S)
                rec[roff + ((step + 1) * n_rec + r) * S + s] = xs[rec_node[r] * S + s]
        v[soff:soff + n * S] = xs
    shape = run.rec.shape
    run.rec.copy_(torch.as_tensor(rec.reshape(shape), dtype=run.rec.dtype))
    run.v.copy_(torch.as_tensor(v.reshape(run.v.shape), dtype=run.v.dtype))

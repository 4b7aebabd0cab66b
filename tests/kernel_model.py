"""CPU model of the fused forward kernel's algorithm and indexing (tests only).

Executes, in numpy float64, what `fused_forward_kernel`
(jaxley-pytorch_amd/csrc/jxhip_fused.hip) computes, on the exact descriptor arrays that
`jaxhip.fused.ForwardLaunch` builds: programs, kernel-order node tables, segments,
program layout, the increment form A dv = b - A v. It lets the CPU test suite check the
host-side tables against the oracle without a GPU; GPU tests then check the kernel.
"""
import numpy as np
import torch


def _np(t):
    return t.detach().cpu().numpy()


def _xoe(x, y):
    z = np.minimum(x / y, 20.0)
    small = np.abs(z) < 1e-4
    zs = np.where(small, 1.0, z)
    return np.where(small, y * (1 - z / 2 + z * z / 12), x / np.expm1(zs))


def _se(x):
    return np.exp(np.minimum(x, 20.0))


def _gate(x, dt, a, b):
    tau = 1 / (a + b)
    e = _se(-dt / tau)
    return x * e + a * tau * (1 - e)


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def _synapse_phase(syn, prm, scale, V, Sst, dt, diff):
    """Slope and current of every (target, simulation) and the new synapse states, the
    synapses visited in the kernel's jagged-diagonal order (KineticSynapse and
    SigmoidRateSynapse)."""
    ptr, pre, tgtn = syn["ptr"], syn["pre"], syn["tgt"]
    n_tgt = len(tgtn)
    TL = np.zeros((n_tgt, V.shape[1]))
    TC = np.zeros((n_tgt, V.shape[1]))
    for tgt in range(n_tgt):
        node = tgtn[tgt]
        j = 0
        while tgt < ptr[j + 1] - ptr[j]:
            q = ptr[j] + tgt
            vpre = V[pre[q]]
            if syn["model"] == 0:
                e_syn, alpha, beta, v_th, delta = (prm[i][q] for i in range(1, 6))
                a = alpha * _sigmoid((vpre - v_th) / delta)
                rate = a + beta
                sinf = a / rate
                Sst[q] = sinf + (Sst[q] - sinf) * np.exp(-dt * rate)
                g = prm[0][q] * Sst[q] * scale[q]
                TL[tgt] += g
                TC[tgt] += g * (V[node] - e_syn)
            elif syn["model"] == 1:
                xo, slope = prm[1][q], prm[2][q]
                sa = _sigmoid((vpre - xo) * slope)
                sb = _sigmoid(((vpre + diff) - xo) * slope)
                TL[tgt] += -prm[0][q] * (sb - sa) * scale[q] / diff
                TC[tgt] += -prm[0][q] * sa * scale[q]
            else:
                raise NotImplementedError("the CPU model covers Kinetic and SigmoidRate synapses")
            j += 1
    return TL, TC


def run_forward_model(run):
    """Fill `run.rec`, `run.v` and the gates as the kernel would (in place)."""
    from jaxhip.fused import ALL_PARAMS

    st = run.st
    progs = _np(run.progs)
    n, Lmax = int(st.topo_row[0]), int(st.topo_row[1])
    seg = _np(st.seg)
    lptr = _np(st.seg_level_ptr)
    sch = _np(st.seg_child)
    ni = _np(st.node_i)
    nf = _np(run.node_f).astype(np.float64)
    cidx = _np(st.child_idx)
    rec_node = _np(st.rec_node)
    n_rec = st.n_rec
    S = run.S
    dt = run.args.dt
    node_p = _np(run.node_p).astype(np.float64)
    bmode = run.bmode if run.bparam is not None else 0
    bparam = None if run.bparam is None else _np(run.bparam).astype(np.float64).reshape(-1, 4)
    bcol = {0: 0, 1: 1, 2: 2, 6: 3}
    v = _np(run.v).astype(np.float64).reshape(-1)
    gm = _np(run.gm).astype(np.float64).reshape(-1)
    gh = _np(run.gh).astype(np.float64).reshape(-1)
    gn = _np(run.gn).astype(np.float64).reshape(-1)
    rec = np.zeros(run.rec.numel())
    stim = None if run.stim is None else _np(run.stim).astype(np.float64)
    scale = None if run.stim_scale is None else _np(run.stim_scale).astype(np.float64)
    scol = _np(st.stim_col)
    ss = np.arange(S)
    syn = None
    if st.syn is not None:
        syn = {"model": st.syn["model"], "ptr": _np(st.syn["ptr"]), "pre": _np(st.syn["pre"]),
               "tgt": _np(st.syn["tgt"])}
        sprm = [_np(x).astype(np.float64) for x in run.syn_params]
        sscale = _np(run.syn_scale).astype(np.float64)
        n_syn = st.syn["n_syn"]
        s_all = None if run.syn_s is None else _np(run.syn_s).astype(np.float64).reshape(-1)

    def P(j, pg, k):
        if bmode and j in bcol:
            if bmode == 1:
                return bparam[pg[5] + ss, bcol[j]]
            return bparam[pg[4] + k * S + ss, bcol[j]]
        return np.full(S, node_p[k, j])

    for pg in progs:
        nsim, t0, t1, soff, simoff, roff = pg[1], pg[2], pg[3], pg[4], pg[5], pg[6]
        act = ss < nsim
        V = v[soff:soff + n * S].reshape(n, S).copy()
        Gm = gm[soff:soff + n * S].reshape(n, S).copy()
        Gh = gh[soff:soff + n * S].reshape(n, S).copy()
        Gn = gn[soff:soff + n * S].reshape(n, S).copy()
        for r in range(n_rec):
            rec[roff + r * S:roff + (r + 1) * S] = V[rec_node[r]]
        if syn is not None:
            synoff = pg[7]
            Sst = (np.zeros((n_syn, S)) if s_all is None
                   else s_all[synoff:synoff + n_syn * S].reshape(n_syn, S).copy())
        for step in range(t1 - t0):
            t = t0 + step
            D = np.zeros((n, S))
            X = np.zeros((n, S))
            if syn is not None:
                TL, TC = _synapse_phase(syn, sprm, sscale, V, Sst, dt, run.args.diff)
            for k in range(n):
                flags = ni[k, 3]
                p = ni[k, 0]
                ax = np.zeros(S)
                if p >= 0:
                    ax += nf[k, 0] * (V[k] - V[p])
                for j in range(ni[k, 1], ni[k, 2]):
                    c = cidx[j]
                    ax += nf[c, 1] * (V[k] - V[c])
                if not (flags & 1):
                    D[k], X[k] = nf[k, 2], ax
                    continue
                g = np.zeros(S)
                cur = np.zeros(S)
                vv = V[k]
                if flags & 2:
                    am = 0.1 * _xoe(-(vv + 40), 10)
                    bm = 4.0 * _se(-(vv + 65) / 18)
                    ah = 0.07 * _se(-(vv + 65) / 20)
                    bh = 1.0 / (_se(-(vv + 35) / 10) + 1)
                    an = 0.01 * _xoe(-(vv + 55), 10)
                    bn = 0.125 * _se(-(vv + 65) / 80)
                    Gm[k] = _gate(Gm[k], dt, am, bm)
                    Gh[k] = _gate(Gh[k], dt, ah, bh)
                    Gn[k] = _gate(Gn[k], dt, an, bn)
                    gNa = P(0, pg, k) * Gm[k] ** 3 * Gh[k]
                    gK = P(1, pg, k) * Gn[k] ** 4
                    gL = P(2, pg, k)
                    g += gNa + gK + gL
                    cur += gNa * (vv - P(3, pg, k)) + gK * (vv - P(4, pg, k)) + gL * (vv - P(5, pg, k))
                if flags & 4:
                    gl = P(6, pg, k)
                    g += gl
                    cur += gl * (vv - P(7, pg, k))
                lin, net = 1000 * g, -1000 * cur
                if syn is not None and ni[k, 6] >= 0:
                    lin, net = lin + TL[ni[k, 6]], net - TC[ni[k, 6]]
                r = ax + dt * nf[k, 3] * net
                for e in range(ni[k, 4], ni[k, 5]):
                    sb = stim.shape[2]
                    b = simoff + ss if sb > 1 else np.zeros(S, np.int64)
                    r = r + stim[t, scol[e], b] * scale[e]
                D[k] = np.where(act, nf[k, 2] + dt * nf[k, 3] * lin, nf[k, 2] + 1)
                X[k] = np.where(act, r, 0.0)
            # Elimination by segment level, chains bottom-up; 1/d kept in D.
            for lv in range(Lmax + 1):
                for q in range(lptr[lv], lptr[lv + 1]):
                    beg, end, cb, ce = seg[q]
                    k = end - 1
                    dk, xk = D[k].copy(), X[k].copy()
                    for c in sch[cb:ce]:
                        f = nf[c, 1] * D[c]
                        dk -= nf[c, 0] * f
                        xk -= X[c] * f
                    inv = 1 / dk
                    D[k], X[k] = inv, xk
                    for k in range(end - 2, beg - 1, -1):
                        f = nf[k + 1, 1] * inv
                        d2 = D[k] - nf[k + 1, 0] * f
                        xk = X[k] - xk * f
                        inv = 1 / d2
                        D[k], X[k] = inv, xk
            for lv in range(Lmax, -1, -1):
                for q in range(lptr[lv], lptr[lv + 1]):
                    beg, end = seg[q, 0], seg[q, 1]
                    p = ni[beg, 0]
                    xp = X[p] if p >= 0 else np.zeros(S)
                    for k in range(beg, end):
                        xk = X[k].copy()
                        if k > beg or p >= 0:
                            xk -= nf[k, 0] * xp
                        xk *= D[k]
                        X[k] = xk
                        xp = xk
            V = V + X
            for r in range(n_rec):
                o = roff + ((step + 1) * n_rec + r) * S
                rec[o:o + S] = V[rec_node[r]]
        v[soff:soff + n * S] = V.reshape(-1)
        gm[soff:soff + n * S] = Gm.reshape(-1)
        gh[soff:soff + n * S] = Gh.reshape(-1)
        gn[soff:soff + n * S] = Gn.reshape(-1)
    run.rec.copy_(torch.as_tensor(rec.reshape(run.rec.shape), dtype=run.rec.dtype))
    run.v.copy_(torch.as_tensor(v.reshape(run.v.shape), dtype=run.v.dtype))
    for buf, arr in ((run.gm, gm), (run.gh, gh), (run.gn, gn)):
        buf.copy_(torch.as_tensor(arr.reshape(buf.shape), dtype=buf.dtype))

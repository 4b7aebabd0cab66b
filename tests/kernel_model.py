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


def _synapse_phase(syn, prm, scale, Vg, V, Sst, dt, diff):
    """Slope and current of every (target, simulation) of one part and its synapses' new
    states, the synapses visited in the kernel's jagged-diagonal order (KineticSynapse
    and SigmoidRateSynapse); presynaptic voltages from the group's voltages Vg at the
    step's start, postsynaptic from the part's V."""
    ptr, pre, tgtn, base = syn["ptr"], syn["pre"], syn["tgt"], syn["base"]
    n_tgt = len(tgtn)
    TL = np.zeros((n_tgt, V.shape[1]))
    TC = np.zeros((n_tgt, V.shape[1]))
    for tgt in range(n_tgt):
        node = tgtn[tgt]
        j = 0
        while tgt < ptr[j + 1] - ptr[j]:
            q = ptr[j] + tgt                       # part-relative synapse
            g_q = base + q                         # in the synapse arrays
            vpre = Vg[pre[g_q]]
            if syn["model"] == 0:
                e_syn, alpha, beta, v_th, delta = (prm[i][g_q] for i in range(1, 6))
                a = alpha * _sigmoid((vpre - v_th) / delta)
                rate = a + beta
                sinf = a / rate
                Sst[q] = sinf + (Sst[q] - sinf) * np.exp(-dt * rate)
                g = prm[0][g_q] * Sst[q] * scale[g_q]
                TL[tgt] += g
                TC[tgt] += g * (V[node] - e_syn)
            elif syn["model"] == 1:
                xo, slope = prm[1][g_q], prm[2][g_q]
                sa = _sigmoid((vpre - xo) * slope)
                sb = _sigmoid(((vpre + diff) - xo) * slope)
                TL[tgt] += -prm[0][g_q] * (sb - sa) * scale[g_q] / diff
                TC[tgt] += -prm[0][g_q] * sa * scale[g_q]
            else:
                raise NotImplementedError("the CPU model covers Kinetic and SigmoidRate synapses")
            j += 1
    return TL, TC


def run_forward_model(run):
    """Fill `run.rec`, `run.v` and the gates as the kernel would (in place). The programs
    of a group (the parts of a split network) advance in lockstep: every part's synapses
    read the group's voltages at the start of the step."""
    st = run.st
    progs = _np(run.progs)
    topo = np.atleast_2d(st.topo_row)
    seg_all = _np(st.seg)
    lptr_all = _np(st.seg_level_ptr)
    sch_all = _np(st.seg_child)
    ni_all = _np(st.node_i)
    nf_all = _np(run.node_f).astype(np.float64)
    cidx_all = _np(st.child_idx)
    rec_node_all = _np(st.rec_node)
    S = run.S
    dt = run.args.dt
    node_p_all = _np(run.node_p).astype(np.float64)
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
    N = st.n
    parts = st.parts
    syn_g = None
    if st.syn is not None:
        sprm = [_np(x).astype(np.float64) for x in run.syn_params]
        sscale = _np(run.syn_scale).astype(np.float64)
        syn_pre = _np(st.syn["pre"])
        syn_ptr = _np(st.syn["ptr"])
        syn_tgt = _np(st.syn["tgt"])
        s_all = None if run.syn_s is None else _np(run.syn_s).astype(np.float64).reshape(-1)
    P = len(progs) // parts
    for gi in range(P):
        group = progs[gi * parts:(gi + 1) * parts]
        gbase = group[0][4] - topo[group[0][0], 2] * S    # the group's block
        Vg = v[gbase:gbase + N * S].reshape(N, S).copy()
        st_p = []
        for pg in group:
            tp = topo[pg[0]]
            n_p, nb = tp[0], tp[2]
            d = {"pg": pg, "tp": tp, "n": n_p, "nb": nb,
                 "Gm": gm[pg[4]:pg[4] + n_p * S].reshape(n_p, S).copy(),
                 "Gh": gh[pg[4]:pg[4] + n_p * S].reshape(n_p, S).copy(),
                 "Gn": gn[pg[4]:pg[4] + n_p * S].reshape(n_p, S).copy()}
            if st.syn is not None:
                d["syn"] = {"model": st.syn["model"], "ptr": syn_ptr[tp[9]:], "pre": syn_pre,
                            "tgt": syn_tgt[tp[8]:tp[8] + tp[7]], "base": tp[10]}
                ns = tp[11]
                d["Sst"] = (np.zeros((ns, S)) if s_all is None
                            else s_all[pg[7]:pg[7] + ns * S].reshape(ns, S).copy())
            st_p.append(d)
        T_steps = group[0][3] - group[0][2]
        for d in st_p:
            _record(d, Vg, rec, rec_node_all, S, 0)
        for step in range(T_steps):
            Vn = Vg.copy()
            for d in st_p:
                _program_step(d, Vg, Vn, step, ni_all, nf_all, cidx_all, seg_all, lptr_all, sch_all,
                              node_p_all, bmode, bparam, bcol, stim, scale, scol, ss, S, dt, run,
                              sprm if st.syn is not None else None, sscale if st.syn is not None else None)
            Vg = Vn
            for d in st_p:
                _record(d, Vg, rec, rec_node_all, S, step + 1)
        v[gbase:gbase + N * S] = Vg.reshape(-1)
        for d in st_p:
            pg, n_p = d["pg"], d["n"]
            gm[pg[4]:pg[4] + n_p * S] = d["Gm"].reshape(-1)
            gh[pg[4]:pg[4] + n_p * S] = d["Gh"].reshape(-1)
            gn[pg[4]:pg[4] + n_p * S] = d["Gn"].reshape(-1)
    run.rec.copy_(torch.as_tensor(rec.reshape(run.rec.shape), dtype=run.rec.dtype))
    run.v.copy_(torch.as_tensor(v.reshape(run.v.shape), dtype=run.v.dtype))
    for buf, arr in ((run.gm, gm), (run.gh, gh), (run.gn, gn)):
        buf.copy_(torch.as_tensor(arr.reshape(buf.shape), dtype=buf.dtype))


def _record(d, Vg, rec, rec_node_all, S, slot):
    pg, tp = d["pg"], d["tp"]
    n_rec, rb, nb = tp[5], tp[6], tp[2]
    for r in range(n_rec):
        o = pg[6] + (slot * n_rec + r) * S
        rec[o:o + S] = Vg[nb + rec_node_all[rb + r]]


def _program_step(d, Vg, Vn, step, ni_all, nf_all, cidx_all, seg_all, lptr_all, sch_all, node_p_all,
                  bmode, bparam, bcol, stim, scale, scol, ss, S, dt, run, sprm, sscale):
    """One step of one program: its part's voltages from the group's Vg into Vn."""
    pg, tp, n, nb = d["pg"], d["tp"], d["n"], d["nb"]
    Lmax, lb, cb, sb, scb = tp[1], tp[3], tp[4], tp[12], tp[13]
    ni = ni_all[nb:nb + n]
    nf = nf_all[nb:nb + n]
    node_p = node_p_all[nb:nb + n]
    cidx = cidx_all[cb:]
    lptr = lptr_all[lb:lb + Lmax + 2]
    seg = seg_all[sb:]
    sch = sch_all[scb:]
    nsim, t0, simoff = pg[1], pg[2], pg[5]
    act = ss < nsim
    V = Vg[nb:nb + n]
    Gm, Gh, Gn = d["Gm"], d["Gh"], d["Gn"]

    def P(j, k):
        if bmode and j in bcol:
            if bmode == 1:
                return bparam[simoff + ss, bcol[j]]
            return bparam[pg[4] + k * S + ss, bcol[j]]
        return np.full(S, node_p[k, j])

    t = t0 + step
    D = np.zeros((n, S))
    X = np.zeros((n, S))
    if "syn" in d:
        TL, TC = _synapse_phase(d["syn"], sprm, sscale, Vg, V, d["Sst"], dt, run.args.diff)
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
            gNa = P(0, k) * Gm[k] ** 3 * Gh[k]
            gK = P(1, k) * Gn[k] ** 4
            gL = P(2, k)
            g += gNa + gK + gL
            cur += gNa * (vv - P(3, k)) + gK * (vv - P(4, k)) + gL * (vv - P(5, k))
        if flags & 4:
            gl = P(6, k)
            g += gl
            cur += gl * (vv - P(7, k))
        lin, net = 1000 * g, -1000 * cur
        if "syn" in d and ni[k, 6] >= 0:
            lin, net = lin + TL[ni[k, 6]], net - TC[ni[k, 6]]
        r = ax + dt * nf[k, 3] * net
        for e in range(ni[k, 4], ni[k, 5]):
            sbn = stim.shape[2]
            b = simoff + ss if sbn > 1 else np.zeros(S, np.int64)
            r = r + stim[t, scol[e], b] * scale[e]
        D[k] = np.where(act, nf[k, 2] + dt * nf[k, 3] * lin, nf[k, 2] + 1)
        X[k] = np.where(act, r, 0.0)
    # Elimination by segment level, chains bottom-up; 1/d kept in D.
    for lv in range(Lmax + 1):
        for q in range(lptr[lv], lptr[lv + 1]):
            beg, end, cb_, ce = seg[q]
            k = end - 1
            dk, xk = D[k].copy(), X[k].copy()
            for c in sch[cb_:ce]:
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
    Vn[nb:nb + n] = V + X

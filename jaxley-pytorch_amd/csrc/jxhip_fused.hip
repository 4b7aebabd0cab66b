// Fused whole-simulation kernels for gfx950.
//
// Every workgroup ("program") owns S simulations of one topology (a cell, a network of
// cells, or a pack of trees) for a range of time steps and runs, per step,
//   synapses -> channels (HH, Leak) + stimulus -> implicit Euler system -> tree solve
//   -> voltage update -> record,
// i.e. `_step_with_plan` (modules/base.py:3214-3400) with the DHS solve
// (solver_voltage.py:294-359), inside the time loop of `integrate` (integrate.py:624-662).
// Simulations are independent: no workgroup waits for another (no grid-wide sync).
//
// Numerics. The backward-Euler system A v(t+1) = b is solved for the increment
//   A dv = r,  r = b - A v(t),  v(t+1) = v(t) + dv,
// which is the same solution; r is formed from the ionic currents and the axial
// currents (differences of neighbouring voltages), so the solve and the gates can run
// in float32 while the voltage state is accumulated in float64: the increment is small,
// its float32 rounding is far below a float32 ulp of the voltage. In float64 mode every
// stage is float64.
//
// Layout. Nodes are renumbered by the host into "kernel order": segments (maximal
// single-child chains, top to bottom) grouped by segment level (leaf segments first).
// The triangularization runs segment level by segment level, one thread per (segment,
// simulation) walking its chain bottom-up in registers; the back-substitution walks
// top-down. LDS holds the diagonal and rhs/increment, [node][S] (simulations fastest).
// In the forward kernel every thread owns the same (node, simulation) items for the
// whole run, so gates and voltages of its items stay in VGPRs; a float64 copy of the
// voltages in HBM gives neighbours' voltages to the axial currents.
#include "jxhip_common.h"
#include "../../include/jxhip.h"

namespace jxhip {

// Column layout of the tables (see jxhip.h).
enum : int { kNiParent = 0, kNiChildBeg, kNiChildEnd, kNiFlags, kNiStimBeg, kNiStimEnd, kNiTgt };
enum : int { kNfLo = 0, kNfUp, kNfDbase, kNfInvCm };
enum : int32_t { kComp = 1, kHH = 2, kLeak = 4 };

template <typename G, typename L>
struct Fused {
  const int64_t* progs;
  const int32_t* topos;
  const int32_t* seg;        // (n_seg_total, 4): beg, end, child_beg, child_end
  const int32_t* seg_lptr;   // per topology: L+2 segment offsets per segment level
  const int32_t* seg_child;  // child segment tops (kernel node index)
  const int32_t* node_i;
  const L* node_f;
  const int32_t* child_idx;
  const int32_t* stim_col;
  const G* stim_scale;
  const int32_t* rec_node;
  const G* node_p;     // (n_nodes_total, JXHIP_NPARAM) shared channel parameters
  const G* bparam;     // batched conductances (gNa, gK, gLeak of HH, gLeak of Leak)
  int32_t bmode;       // 0 none, 1 per simulation, 2 per (node, simulation)
  double* v;
  G* gm;
  G* gh;
  G* gn;
  const G* stim;
  int64_t stim_cols;
  int64_t stim_b;
  G* rec;
  int32_t syn_model;
  const int32_t* syn_pre;
  const int32_t* syn_ptr;
  const int32_t* tgt_node;
  const G* syn_param[JXHIP_NSYNPARAM];
  const G* syn_scale;
  G* syn_s;
  G dt;
  G diff;
  int32_t S;
  int32_t log2S;
  double* ckpt_v;
  G* ckpt_m;
  G* ckpt_h;
  G* ckpt_n;
  G* ckpt_s;
  int64_t ckpt_every;
  int64_t ckpt_stride;
  int64_t ckpt_syn_stride;
};

template <typename G, typename L>
Fused<G, L> unpack(const jxhip_fused_args& a) {
  Fused<G, L> f;
  f.progs = a.progs;
  f.topos = a.topos;
  f.seg = a.seg;
  f.seg_lptr = a.seg_level_ptr;
  f.seg_child = a.seg_child;
  f.node_i = a.node_i;
  f.node_f = static_cast<const L*>(a.node_f);
  f.child_idx = a.child_idx;
  f.stim_col = a.stim_col;
  f.stim_scale = static_cast<const G*>(a.stim_scale);
  f.rec_node = a.rec_node;
  f.node_p = static_cast<const G*>(a.node_p);
  f.bparam = static_cast<const G*>(a.bparam);
  f.bmode = a.bparam_mode;
  f.v = static_cast<double*>(a.v);
  f.gm = static_cast<G*>(a.gate_m);
  f.gh = static_cast<G*>(a.gate_h);
  f.gn = static_cast<G*>(a.gate_n);
  f.stim = static_cast<const G*>(a.stim);
  f.stim_cols = a.stim_cols;
  f.stim_b = a.stim_b;
  f.rec = static_cast<G*>(a.rec);
  f.syn_model = a.syn_model;
  f.syn_pre = a.syn_pre;
  f.syn_ptr = a.syn_ptr;
  f.tgt_node = a.tgt_node;
  for (int j = 0; j < JXHIP_NSYNPARAM; ++j) f.syn_param[j] = static_cast<const G*>(a.syn_param[j]);
  f.syn_scale = static_cast<const G*>(a.syn_scale);
  f.syn_s = static_cast<G*>(a.syn_s);
  f.dt = static_cast<G>(a.dt);
  f.diff = static_cast<G>(a.diff);
  f.S = a.S;
  int l = 0;
  while ((1 << l) < a.S) ++l;
  f.log2S = l;
  f.ckpt_v = static_cast<double*>(a.ckpt_v);
  f.ckpt_m = static_cast<G*>(a.ckpt_m);
  f.ckpt_h = static_cast<G*>(a.ckpt_h);
  f.ckpt_n = static_cast<G*>(a.ckpt_n);
  f.ckpt_s = static_cast<G*>(a.ckpt_s);
  f.ckpt_every = a.ckpt_every;
  f.ckpt_stride = a.ckpt_stride;
  f.ckpt_syn_stride = a.ckpt_syn_stride;
  return f;
}

struct Prog {
  int32_t topo, nsim;
  int64_t t0, t1, soff, simoff, roff, synoff;
};
struct Topo {
  int32_t n, L, nb, lb, cb, n_rec, rb, n_tgt, tgb, spb, synb, n_syn, sb, scb;
};

__device__ __forceinline__ Prog load_prog(const int64_t* progs, int64_t p) {
  const int64_t* r = progs + p * JXHIP_PROG_COLS;
  Prog g;
  g.topo = static_cast<int32_t>(r[0]);
  g.nsim = static_cast<int32_t>(r[1]);
  g.t0 = r[2];
  g.t1 = r[3];
  g.soff = r[4];
  g.simoff = r[5];
  g.roff = r[6];
  g.synoff = r[7];
  return g;
}

__device__ __forceinline__ Topo load_topo(const int32_t* topos, int32_t t) {
  const int32_t* r = topos + static_cast<int64_t>(t) * JXHIP_TOPO_COLS;
  Topo o;
  o.n = r[0];
  o.L = r[1];
  o.nb = r[2];
  o.lb = r[3];
  o.cb = r[4];
  o.n_rec = r[5];
  o.rb = r[6];
  o.n_tgt = r[7];
  o.tgb = r[8];
  o.spb = r[9];
  o.synb = r[10];
  o.n_syn = r[11];
  o.sb = r[12];
  o.scb = r[13];
  return o;
}

// Reciprocal: hardware rcp in float32 (the operands are increments and diagonals of a
// diagonally dominant system; 1 ulp is far below the tolerance), IEEE in float64.
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

// Channel parameter J of node k, simulation s: the batched conductances (J = 0, 1, 2:
// gNa, gK, gLeak of HH; J = 6: gLeak of Leak) come from `bparam` when batched, every
// other value from the per-node table.
template <int J, typename G, typename L>
__device__ __forceinline__ G prm(const Fused<G, L>& a, const Prog& pg, const Topo& tp, int k, int s) {
  constexpr int col = J == 0 ? 0 : J == 1 ? 1 : J == 2 ? 2 : J == 6 ? 3 : -1;
  if constexpr (col >= 0) {
    if (a.bmode == 1) return a.bparam[((pg.simoff + s) << 2) + col];
    if (a.bmode == 2)
      return a.bparam[((pg.soff + (static_cast<int64_t>(k) << a.log2S) + s) << 2) + col];
  }
  return a.node_p[static_cast<int64_t>(tp.nb + k) * JXHIP_NPARAM + J];
}

// Synapses of one target and simulation (modules/network.py:382-478,
// kernels/synapse_step.py:121-182): `lin` = summed slope of i / area * 1e5 (the
// reference's finite difference with both voltages perturbed by `diff`), `cur` = summed
// i / area * 1e5 at the unperturbed voltages.
template <typename G, typename L>
__device__ __forceinline__ void synapse_target(const Fused<G, L>& a, const Prog& pg,
                                               const Topo& tp, int tgt, int s, G& lin, G& cur) {
  G sl = G(0), sc = G(0);
  const G dt = a.dt;
  const int32_t* sptr = a.syn_ptr + tp.spb;
  const int S = a.S;
  const double* v = a.v + pg.soff;
  for (int syn = sptr[tgt]; syn < sptr[tgt + 1]; ++syn) {
    const int64_t gsyn = tp.synb + syn;
    const int pre = a.syn_pre[gsyn];
    const G vpre = static_cast<G>(v[pre * S + s]);
    const G scale = a.syn_scale[gsyn];
    const G gs = a.syn_param[0][gsyn];
    if (a.syn_model == 0) {
      const int node = a.tgt_node[tp.tgb + tgt];
      const G e_syn = a.syn_param[1][gsyn];
      const G alpha = a.syn_param[2][gsyn];
      const G beta = a.syn_param[3][gsyn];
      const G v_th = a.syn_param[4][gsyn];
      const G delta = a.syn_param[5][gsyn];
      G* sp = a.syn_s + pg.synoff + static_cast<int64_t>(syn) * S + s;
      const G st = *sp;
      const G tr = sigmoid((vpre - v_th) / delta);
      const G aa = alpha * tr;
      const G rate = aa + beta;
      const G sinf = aa / rate;
      const G snew = sinf + (st - sinf) * dexp(-dt * rate);
      *sp = snew;
      // i = gS s (v_post - e_syn) is linear in v_post: the slope is gS s.
      const G g = gs * snew * scale;
      sl += g;
      sc += g * static_cast<G>(v[node * S + s] - static_cast<double>(e_syn));
    } else {
      const G xo = a.syn_param[1][gsyn];
      const G slope = a.syn_param[2][gsyn];
      // sig(b) - sig(a) = (1 - sig(a)) sig(b) (-expm1(-(b - a))) without cancellation.
      const G za = (vpre - xo) * slope;
      const G zb = ((vpre + a.diff) - xo) * slope;
      const G sa = sigmoid(za);
      const G sb = sigmoid(zb);
      const G dsig = (G(1) - sa) * sb * (-dexpm1(-(zb - za)));
      sl += -gs * dsig * scale / a.diff;
      sc += -gs * sa * scale;
    }
  }
  lin = sl;
  cur = sc;
}

// LDS of a program.
template <typename L>
struct Lds {
  L* xs;   // rhs, then the increment (solution)
  L* ds;   // diagonal, then its reciprocal after elimination
  L* tl;   // synapse slope per (target, simulation)
  L* tc;   // synapse current per (target, simulation)
  L* cl;   // lower / upper coefficient per node
};

// Triangularization leaves first, one thread per (segment, simulation) walking its
// chain bottom-up; back-substitution top-down. kT: solve with A^T (lower <-> upper).
template <typename G, typename L, bool kT>
__device__ __forceinline__ void tree_solve(const Fused<G, L>& a, const Topo& tp, L* xs, L* ds,
                                           const L* cl) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int S = a.S, lS = a.log2S;
  const int32_t* sg = a.seg + static_cast<int64_t>(tp.sb) * 4;
  const int32_t* lptr = a.seg_lptr + tp.lb;
  const int32_t* sch = a.seg_child + tp.scb;
  const int32_t* ni = a.node_i + static_cast<int64_t>(tp.nb) * JXHIP_NODE_ICOLS;
  // cl: [lo, up] per node in LDS.
  constexpr int kLo = kT ? 1 : 0;
  constexpr int kUp = kT ? 0 : 1;
  for (int lv = 0; lv <= tp.L; ++lv) {
    const int lo = lptr[lv], w = (lptr[lv + 1] - lo) << lS;
    for (int i = tid; i < w; i += nthr) {
      const int q = lo + (i >> lS), s = i & (S - 1);
      const int beg = sg[4 * q], end = sg[4 * q + 1];
      int k = end - 1;
      L dk = ds[(k << lS) + s], xk = xs[(k << lS) + s];
      for (int j = sg[4 * q + 2]; j < sg[4 * q + 3]; ++j) {
        const int c = sch[j];
        // ds[c] holds 1 / d[c] of the already eliminated child segment top.
        const L f = cl[2 * c + kUp] * ds[(c << lS) + s];
        dk -= cl[2 * c + kLo] * f;
        xk -= xs[(c << lS) + s] * f;
      }
      L inv = recip(dk);
      ds[(k << lS) + s] = inv;
      xs[(k << lS) + s] = xk;
      for (k = end - 2; k >= beg; --k) {
        const L f = cl[2 * (k + 1) + kUp] * inv;
        const L d2 = ds[(k << lS) + s] - cl[2 * (k + 1) + kLo] * f;
        xk = xs[(k << lS) + s] - xk * f;
        inv = recip(d2);
        ds[(k << lS) + s] = inv;
        xs[(k << lS) + s] = xk;
      }
    }
    __syncthreads();
  }
  for (int lv = tp.L; lv >= 0; --lv) {
    const int lo = lptr[lv], w = (lptr[lv + 1] - lo) << lS;
    for (int i = tid; i < w; i += nthr) {
      const int q = lo + (i >> lS), s = i & (S - 1);
      const int beg = sg[4 * q], end = sg[4 * q + 1];
      const int p = ni[beg * JXHIP_NODE_ICOLS + kNiParent];
      L xp = p >= 0 ? xs[(p << lS) + s] : L(0);
      for (int k = beg; k < end; ++k) {
        L xk = xs[(k << lS) + s];
        if (k > beg || p >= 0) xk -= cl[2 * k + kLo] * xp;
        xk *= ds[(k << lS) + s];
        xs[(k << lS) + s] = xk;
        xp = xk;
      }
    }
    __syncthreads();
  }
}

// Copies the lower / upper coefficients of the topology into LDS (once per launch).
template <typename G, typename L>
__device__ __forceinline__ void load_coefficients(const Fused<G, L>& a, const Topo& tp, L* cl) {
  const L* nf = a.node_f + static_cast<int64_t>(tp.nb) * JXHIP_NODE_FCOLS;
  for (int k = threadIdx.x; k < tp.n; k += blockDim.x) {
    cl[2 * k] = nf[k * JXHIP_NODE_FCOLS + kNfLo];
    cl[2 * k + 1] = nf[k * JXHIP_NODE_FCOLS + kNfUp];
  }
}

// System of one (node, simulation) item: diagonal d and residual r = b - A v.
// Gates are read from / written to the references (registers or memory).
template <typename G, typename L>
__device__ __forceinline__ void build_item(const Fused<G, L>& a, const Prog& pg, const Topo& tp,
                                           const Lds<L>& Ld, const double* vg, int k, int s,
                                           double v, int64_t t, G& m, G& h, G& n) {
  const int lS = a.log2S;
  const int i = (k << lS) + s;
  const int32_t* nik = a.node_i + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS;
  const L* nfk = a.node_f + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_FCOLS;
  const int32_t flags = nik[kNiFlags];
  // Axial part of the residual: lo_k (v_k - v_parent) + sum_c up_c (v_k - v_c).
  L ax = L(0);
  const int p = nik[kNiParent];
  if (p >= 0) ax += nfk[kNfLo] * static_cast<L>(v - vg[(p << lS) + s]);
  const int32_t* cidx = a.child_idx + tp.cb;
  for (int j = nik[kNiChildBeg]; j < nik[kNiChildEnd]; ++j) {
    const int c = cidx[j];
    ax += a.node_f[static_cast<int64_t>(tp.nb + c) * JXHIP_NODE_FCOLS + kNfUp] *
          static_cast<L>(v - vg[(c << lS) + s]);
  }
  if (!(flags & kComp)) {
    Ld.ds[i] = nfk[kNfDbase];
    Ld.xs[i] = ax;
    return;
  }
  const G dt = a.dt;
  const G inv_cm = static_cast<G>(nfk[kNfInvCm]);
  const G vv = static_cast<G>(v);
  G g = G(0);    // total conductance (mS/cm^2 -> x1000 below)
  G cur = G(0);  // ionic current sum g_i (v - E_i), uA/cm^2 / 1000
  if (flags & kHH) {
    G am, bm, ah, bh, an, bn;
    hh_rates(vv, am, bm, ah, bh, an, bn);
    m = gate_step(m, dt, am, bm);
    h = gate_step(h, dt, ah, bh);
    n = gate_step(n, dt, an, bn);
    const G gNa = prm<0>(a, pg, tp, k, s) * (m * m * m) * h;
    const G n2 = n * n;
    const G gK = prm<1>(a, pg, tp, k, s) * (n2 * n2);
    const G gL = prm<2>(a, pg, tp, k, s);
    g += gNa + gK + gL;
    cur += gNa * static_cast<G>(v - static_cast<double>(prm<3>(a, pg, tp, k, s))) +
           gK * static_cast<G>(v - static_cast<double>(prm<4>(a, pg, tp, k, s))) +
           gL * static_cast<G>(v - static_cast<double>(prm<5>(a, pg, tp, k, s)));
  }
  if (flags & kLeak) {
    const G gl = prm<6>(a, pg, tp, k, s);
    g += gl;
    cur += gl * static_cast<G>(v - static_cast<double>(prm<7>(a, pg, tp, k, s)));
  }
  G lin = G(1000) * g;
  G net = -G(1000) * cur;  // (cst - lin v) of the channels
  const int tg = nik[kNiTgt];
  if (tg >= 0) {
    lin += static_cast<G>(Ld.tl[(tg << lS) + s]);
    net -= static_cast<G>(Ld.tc[(tg << lS) + s]);
  }
  L r = ax + static_cast<L>(dt * inv_cm * net);
  for (int e = nik[kNiStimBeg]; e < nik[kNiStimEnd]; ++e) {
    const int64_t col = a.stim_col[e];
    const int64_t bidx = a.stim_b > 1 ? (pg.simoff + s) : 0;
    r += static_cast<L>(a.stim[(t * a.stim_cols + col) * a.stim_b + bidx] * a.stim_scale[e]);
  }
  Ld.ds[i] = nfk[kNfDbase] + static_cast<L>(dt * inv_cm * lin);
  Ld.xs[i] = r;
}

// Items of simulations beyond `nsim` (padding): a well-conditioned row, zero rhs.
template <typename G, typename L>
__device__ __forceinline__ void idle_item(const Fused<G, L>& a, const Topo& tp, const Lds<L>& Ld,
                                          int k, int i) {
  const int32_t flags = a.node_i[static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS + kNiFlags];
  Ld.ds[i] = a.node_f[static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_FCOLS + kNfDbase] +
             ((flags & kComp) ? L(1) : L(0));
  Ld.xs[i] = L(0);
}

template <typename G, typename L>
__device__ __forceinline__ void synapse_phase(const Fused<G, L>& a, const Prog& pg, const Topo& tp,
                                              const Lds<L>& Ld) {
  if (a.syn_model < 0) return;
  const int lS = a.log2S, S = a.S;
  for (int i = threadIdx.x; i < (tp.n_tgt << lS); i += blockDim.x) {
    const int tgt = i >> lS, s = i & (S - 1);
    G lin = G(0), cur = G(0);
    if (s < pg.nsim) synapse_target(a, pg, tp, tgt, s, lin, cur);
    Ld.tl[i] = static_cast<L>(lin);
    Ld.tc[i] = static_cast<L>(cur);
  }
  __syncthreads();
}

// Forward. The first IPT items of a thread, i = tid + j * blockDim (j < IPT), keep their
// gates in VGPRs for the whole launch; further items stream their gates from HBM.
// Voltages live in HBM (float64, program layout); LDS holds the system and the
// coefficients of the tree.
template <typename G, typename L, int IPT>
__global__ void __launch_bounds__(1024) fused_forward_kernel(Fused<G, L> a) {
  extern __shared__ unsigned char smem_raw[];
  const Prog pg = load_prog(a.progs, blockIdx.x);
  const Topo tp = load_topo(a.topos, pg.topo);
  const int lS = a.log2S, S = a.S;
  const int nS = tp.n << lS;
  Lds<L> Ld;
  Ld.xs = reinterpret_cast<L*>(smem_raw);
  Ld.ds = Ld.xs + nS;
  Ld.tl = Ld.ds + nS;
  Ld.tc = Ld.tl + (tp.n_tgt << lS);
  Ld.cl = Ld.tc + (tp.n_tgt << lS);
  load_coefficients(a, tp, Ld.cl);
  const int tid = threadIdx.x, nthr = blockDim.x;
  double* vg = a.v + pg.soff;
  G* gm = a.gm + pg.soff;
  G* gh = a.gh + pg.soff;
  G* gn = a.gn + pg.soff;
  constexpr int kR = IPT > 0 ? IPT : 1;
  G mr[kR], hr[kR], nr[kR];
  if constexpr (IPT > 0) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * nthr;
      if (i < nS) {
        mr[j] = gm[i];
        hr[j] = gh[i];
        nr[j] = gn[i];
      }
    }
  }
  for (int i = tid; i < (tp.n_rec << lS); i += nthr) {
    const int r = i >> lS, s = i & (S - 1);
    a.rec[pg.roff + i] = static_cast<G>(vg[(a.rec_node[tp.rb + r] << lS) + s]);
  }
  const int64_t nsteps = pg.t1 - pg.t0;
  for (int64_t step = 0; step < nsteps; ++step) {
    if (a.ckpt_v != nullptr && step % a.ckpt_every == 0) {
      const int64_t off = pg.soff + (step / a.ckpt_every) * a.ckpt_stride;
      for (int i = tid; i < nS; i += nthr) a.ckpt_v[off + i] = vg[i];
      if constexpr (IPT > 0) {
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          const int i = tid + j * nthr;
          if (i < nS) {
            a.ckpt_m[off + i] = mr[j];
            a.ckpt_h[off + i] = hr[j];
            a.ckpt_n[off + i] = nr[j];
          }
        }
      }
      for (int i = tid + IPT * nthr; i < nS; i += nthr) {
        a.ckpt_m[off + i] = gm[i];
        a.ckpt_h[off + i] = gh[i];
        a.ckpt_n[off + i] = gn[i];
      }
      if (a.syn_model == 0 && a.ckpt_s != nullptr) {
        const int64_t nsyn = static_cast<int64_t>(tp.n_syn) << lS;
        const int64_t so = pg.synoff + (step / a.ckpt_every) * a.ckpt_syn_stride;
        for (int64_t i = tid; i < nsyn; i += nthr) a.ckpt_s[so + i] = a.syn_s[pg.synoff + i];
      }
    }
    synapse_phase(a, pg, tp, Ld);
    const int64_t t = pg.t0 + step;
    if constexpr (IPT > 0) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        int i = tid + j * nthr;
        // Opaque to the optimizer: the item's addresses are recomputed every step instead
        // of being hoisted out of the time loop for all IPT items (register pressure).
        asm volatile("" : "+v"(i));
        if (i < nS) {
          const int k = i >> lS, s = i & (S - 1);
          if (s < pg.nsim) {
            build_item(a, pg, tp, Ld, vg, k, s, vg[i], t, mr[j], hr[j], nr[j]);
          } else {
            idle_item(a, tp, Ld, k, i);
          }
        }
      }
    }
    // Items beyond the register-resident ones: gates streamed from HBM.
    for (int i = tid + IPT * nthr; i < nS; i += nthr) {
      const int k = i >> lS, s = i & (S - 1);
      if (s < pg.nsim) {
        G m = gm[i], h = gh[i], n = gn[i];
        build_item(a, pg, tp, Ld, vg, k, s, vg[i], t, m, h, n);
        gm[i] = m;
        gh[i] = h;
        gn[i] = n;
      } else {
        idle_item(a, tp, Ld, k, i);
      }
    }
    __syncthreads();
    tree_solve<G, L, false>(a, tp, Ld.xs, Ld.ds, Ld.cl);
    for (int i = tid; i < nS; i += nthr) vg[i] += static_cast<double>(Ld.xs[i]);
    __syncthreads();
    for (int i = tid; i < (tp.n_rec << lS); i += nthr) {
      const int r = i >> lS, s = i & (S - 1);
      a.rec[pg.roff + ((step + 1) * tp.n_rec + r) * S + s] =
          static_cast<G>(vg[(a.rec_node[tp.rb + r] << lS) + s]);
    }
  }
  if constexpr (IPT > 0) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * nthr;
      if (i < nS) {
        gm[i] = mr[j];
        gh[i] = hr[j];
        gn[i] = nr[j];
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------------

template <typename G>
struct Adj {
  const G* g_rec;
  G* g_param[JXHIP_NPARAM];
  G* g_v0;
  double* wv;
  G* wm;
  G* wh;
  G* wn;
  G* lm;
  G* lh;
  G* ln;
  int64_t work_stride;
};

// d/dx of x_over_expm1(x, y) as autograd differentiates solver_gate.py:14-30.
template <typename T>
__device__ __forceinline__ T d_x_over_expm1(T x, T y) {
  const T z = x / y;
  const T zc = z < T(20) ? z : T(20);
  const T pass = z <= T(20) ? T(1) : T(0);
  if (fabs(zc) < T(1e-4)) return pass * (T(-0.5) + zc / T(6));
  const T em1 = dexpm1(zc);
  return T(1) / em1 - x * (em1 + T(1)) / (em1 * em1) * pass / y;
}

template <typename T>
__device__ __forceinline__ T d_save_exp(T u) {
  return u <= T(20) ? dexp(u) : T(0);
}

// dx'/dv of the exponential-Euler gate step through alpha(v), beta(v); e = dx'/dx.
template <typename T>
__device__ __forceinline__ T d_gate_dv(T x, T dt, T al, T be, T dal, T dbe, T& e) {
  const T tau = ddiv(T(1), al + be);
  const T xinf = al * tau;
  e = save_exp(-dt * (al + be));
  const T de = -dt * e;
  const T dxa = be * tau * tau * (T(1) - e) + (x - xinf) * de;
  const T dxb = -al * tau * tau * (T(1) - e) + (x - xinf) * de;
  return dxa * dal + dxb * dbe;
}

template <typename G, typename L>
__global__ void __launch_bounds__(512) fused_adjoint_kernel(Fused<G, L> a, Adj<G> g) {
  extern __shared__ unsigned char smem_raw[];
  const Prog pg = load_prog(a.progs, blockIdx.x);
  const Topo tp = load_topo(a.topos, pg.topo);
  const int lS = a.log2S, S = a.S;
  const int nS = tp.n << lS;
  Lds<L> Ld;
  Ld.xs = reinterpret_cast<L*>(smem_raw);
  Ld.ds = Ld.xs + nS;
  L* lv = Ld.ds + nS;  // adjoint of the voltages
  Ld.tl = lv;          // unused (no synapses in the adjoint)
  Ld.tc = lv;
  Ld.cl = lv + nS;
  load_coefficients(a, tp, Ld.cl);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const G dt = a.dt;
  const int64_t nsteps = pg.t1 - pg.t0;
  const int64_t K = a.ckpt_every;
  const int64_t C = (nsteps + K - 1) / K;
  const int64_t wbase = static_cast<int64_t>(blockIdx.x) * g.work_stride;
  G* lm = g.lm + pg.soff;
  G* lh = g.lh + pg.soff;
  G* ln = g.ln + pg.soff;
  for (int i = tid; i < nS; i += nthr) {
    lv[i] = L(0);
    lm[i] = G(0);
    lh[i] = G(0);
    ln[i] = G(0);
  }
  __syncthreads();
  double* wv = g.wv + wbase;
  G* wm = g.wm + wbase;
  G* wh = g.wh + wbase;
  G* wn = g.wn + wbase;
  for (int64_t c = C - 1; c >= 0; --c) {
    const int64_t ts = c * K;
    const int64_t te = ts + K < nsteps ? ts + K : nsteps;
    // Recompute the segment from its checkpoint into work[0 .. te - ts]; the voltage
    // copy of step j is work_v[j] (program layout, as the forward's HBM copy).
    const int64_t coff = pg.soff + c * a.ckpt_stride;
    for (int i = tid; i < nS; i += nthr) {
      wv[i] = a.ckpt_v[coff + i];
      wm[i] = a.ckpt_m[coff + i];
      wh[i] = a.ckpt_h[coff + i];
      wn[i] = a.ckpt_n[coff + i];
    }
    __syncthreads();
    for (int64_t j = 0; j < te - ts; ++j) {
      for (int i = tid; i < nS; i += nthr) {
        const int k = i >> lS, s = i & (S - 1);
        G m = wm[j * nS + i], h = wh[j * nS + i], n = wn[j * nS + i];
        if (s < pg.nsim) {
          build_item(a, pg, tp, Ld, wv + j * nS, k, s, wv[j * nS + i], pg.t0 + ts + j, m, h, n);
        } else {
          idle_item(a, tp, Ld, k, i);
        }
        wm[(j + 1) * nS + i] = m;
        wh[(j + 1) * nS + i] = h;
        wn[(j + 1) * nS + i] = n;
      }
      __syncthreads();
      tree_solve<G, L, false>(a, tp, Ld.xs, Ld.ds, Ld.cl);
      for (int i = tid; i < nS; i += nthr) wv[(j + 1) * nS + i] = wv[j * nS + i] + static_cast<double>(Ld.xs[i]);
      __syncthreads();
    }
    for (int64_t j = te - ts - 1; j >= 0; --j) {
      const int64_t t1 = ts + j + 1;
      for (int s = tid; s < S; s += nthr) {
        for (int r = 0; r < tp.n_rec; ++r) {
          const int node = a.rec_node[tp.rb + r];
          lv[(node << lS) + s] += static_cast<L>(g.g_rec[pg.roff + (t1 * tp.n_rec + r) * S + s]);
        }
      }
      __syncthreads();
      const G* m1 = wm + (j + 1) * nS;
      const G* h1 = wh + (j + 1) * nS;
      const G* n1 = wn + (j + 1) * nS;
      for (int i = tid; i < nS; i += nthr) {
        const int k = i >> lS, s = i & (S - 1);
        const int32_t* nik = a.node_i + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS;
        const L* nfk = a.node_f + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_FCOLS;
        const int32_t flags = nik[kNiFlags];
        G lin = G(0);
        if ((flags & kComp) && s < pg.nsim) {
          if (flags & kHH) {
            const G mm = m1[i], hh = h1[i], nn = n1[i];
            const G n2 = nn * nn;
            lin += G(1000) * (prm<0>(a, pg, tp, k, s) * (mm * mm * mm) * hh +
                              prm<1>(a, pg, tp, k, s) * (n2 * n2) + prm<2>(a, pg, tp, k, s));
          }
          if (flags & kLeak) lin += G(1000) * prm<6>(a, pg, tp, k, s);
          Ld.ds[i] = nfk[kNfDbase] + static_cast<L>(dt * static_cast<G>(nfk[kNfInvCm]) * lin);
        } else {
          Ld.ds[i] = nfk[kNfDbase] + ((flags & kComp) ? L(1) : L(0));
        }
        Ld.xs[i] = lv[i];
      }
      __syncthreads();
      tree_solve<G, L, true>(a, tp, Ld.xs, Ld.ds, Ld.cl);  // xs = mu = A^-T lambda_v(t+1)
      const double* v0 = wv + j * nS;
      const double* v1 = wv + (j + 1) * nS;
      const G* m0 = wm + j * nS;
      const G* h0 = wh + j * nS;
      const G* n0 = wn + j * nS;
      for (int i = tid; i < nS; i += nthr) {
        const int k = i >> lS, s = i & (S - 1);
        const int32_t* nik = a.node_i + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS;
        const L* nfk = a.node_f + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_FCOLS;
        const int32_t flags = nik[kNiFlags];
        if (!(flags & kComp) || s >= pg.nsim) {
          lv[i] = L(0);
          continue;
        }
        const G inv_cm = static_cast<G>(nfk[kNfInvCm]);
        const G mu = static_cast<G>(Ld.xs[i]);
        const G gl = -mu * static_cast<G>(v1[i]) * dt * inv_cm;  // dL/dlin
        const G gc = mu * dt * inv_cm;                           // dL/dcst
        G lvn = mu;
        if (flags & kHH) {
          const G gNa = prm<0>(a, pg, tp, k, s);
          const G gK = prm<1>(a, pg, tp, k, s);
          const G eNa = prm<3>(a, pg, tp, k, s);
          const G eK = prm<4>(a, pg, tp, k, s);
          const G eL = prm<5>(a, pg, tp, k, s);
          const G mm = m1[i], hh = h1[i], nn = n1[i];
          const G m3 = mm * mm * mm, n3 = nn * nn * nn, n4 = n3 * nn;
          const G wNa = G(1000) * (gl + gc * eNa);
          const G wK = G(1000) * (gl + gc * eK);
          const G wL = G(1000) * (gl + gc * eL);
          const int64_t pi = pg.soff + i;
          if (g.g_param[0]) g.g_param[0][pi] += wNa * m3 * hh;
          if (g.g_param[1]) g.g_param[1][pi] += wK * n4;
          if (g.g_param[2]) g.g_param[2][pi] += wL;
          if (g.g_param[3]) g.g_param[3][pi] += G(1000) * gc * gNa * m3 * hh;
          if (g.g_param[4]) g.g_param[4][pi] += G(1000) * gc * gK * n4;
          if (g.g_param[5]) g.g_param[5][pi] += G(1000) * gc * prm<2>(a, pg, tp, k, s);
          const G lam_m1 = lm[i] + wNa * gNa * G(3) * mm * mm * hh;
          const G lam_h1 = lh[i] + wNa * gNa * m3;
          const G lam_n1 = ln[i] + wK * gK * G(4) * n3;
          const G v = static_cast<G>(v0[i]);
          const G xm = -(v + G(40)), xn = -(v + G(55));
          const G um = -(v + G(65)) / G(18), uh = -(v + G(65)) / G(20);
          const G ub = -(v + G(35)) / G(10), un = -(v + G(65)) / G(80);
          const G am = G(0.1) * x_over_expm1(xm, G(10));
          const G dam = -G(0.1) * d_x_over_expm1(xm, G(10));
          const G bm = G(4) * save_exp(um);
          const G dbm = -G(4) / G(18) * d_save_exp(um);
          const G ah = G(0.07) * save_exp(uh);
          const G dah = -G(0.07) / G(20) * d_save_exp(uh);
          const G Eb = save_exp(ub);
          const G bh = G(1) / (Eb + G(1));
          const G dbh = d_save_exp(ub) / (G(10) * (Eb + G(1)) * (Eb + G(1)));
          const G an = G(0.01) * x_over_expm1(xn, G(10));
          const G dan = -G(0.01) * d_x_over_expm1(xn, G(10));
          const G bn = G(0.125) * save_exp(un);
          const G dbn = -G(0.125) / G(80) * d_save_exp(un);
          G em, eh, en;
          const G dmv = d_gate_dv(m0[i], dt, am, bm, dam, dbm, em);
          const G dhv = d_gate_dv(h0[i], dt, ah, bh, dah, dbh, eh);
          const G dnv = d_gate_dv(n0[i], dt, an, bn, dan, dbn, en);
          lvn += lam_m1 * dmv + lam_h1 * dhv + lam_n1 * dnv;
          lm[i] = lam_m1 * em;
          lh[i] = lam_h1 * eh;
          ln[i] = lam_n1 * en;
        }
        if (flags & kLeak) {
          const G eLk = prm<7>(a, pg, tp, k, s);
          const int64_t pi = pg.soff + i;
          if (g.g_param[6]) g.g_param[6][pi] += G(1000) * (gl + gc * eLk);
          if (g.g_param[7]) g.g_param[7][pi] += G(1000) * gc * prm<6>(a, pg, tp, k, s);
        }
        lv[i] = static_cast<L>(lvn);
      }
      __syncthreads();
    }
  }
  for (int s = tid; s < S; s += nthr) {
    for (int r = 0; r < tp.n_rec; ++r) {
      const int node = a.rec_node[tp.rb + r];
      lv[(node << lS) + s] += static_cast<L>(g.g_rec[pg.roff + r * S + s]);
    }
  }
  __syncthreads();
  for (int i = tid; i < nS; i += nthr) g.g_v0[pg.soff + i] = static_cast<G>(lv[i]);
}

template <typename K>
hipError_t set_lds(K kern, size_t lds) {
  if (lds <= 64 * 1024) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
}

inline int check_args(const jxhip_fused_args& a) {
  if (a.n_progs > 0x7fffffffLL || a.threads < 64 || a.threads > 1024 || a.threads % 64)
    return hipErrorInvalidValue;
  if (a.S < 1 || a.S > 64 || (a.S & (a.S - 1))) return hipErrorInvalidValue;
  if (a.lds_bytes < 0 || a.lds_bytes > 160 * 1024) return hipErrorInvalidValue;
  return hipSuccess;
}

template <typename G, typename L>
int launch_forward(const jxhip_fused_args* args, hipStream_t stream) {
  if (args == nullptr) return hipErrorInvalidValue;
  const jxhip_fused_args& a = *args;
  if (a.n_progs <= 0) return hipSuccess;
  if (int e = check_args(a)) return e;
  const Fused<G, L> f = unpack<G, L>(a);
  const dim3 grid(static_cast<unsigned>(a.n_progs)), block(a.threads);
  const size_t lds = static_cast<size_t>(a.lds_bytes);
#define JXHIP_FWD_CASE(IPT)                                         \
  case IPT: {                                                       \
    auto kern = fused_forward_kernel<G, L, IPT>;                    \
    if (hipError_t e = set_lds(kern, lds)) return e;                \
    hipLaunchKernelGGL(kern, grid, block, lds, stream, f);          \
    break;                                                          \
  }
  switch (a.items_per_thread) {
    JXHIP_FWD_CASE(0)
    JXHIP_FWD_CASE(1)
    JXHIP_FWD_CASE(2)
    JXHIP_FWD_CASE(4)
    JXHIP_FWD_CASE(8)
    JXHIP_FWD_CASE(12)
    JXHIP_FWD_CASE(16)
    default:
      return hipErrorInvalidValue;
  }
#undef JXHIP_FWD_CASE
  return hipGetLastError();
}

template <typename G, typename L>
int launch_adjoint(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                   hipStream_t stream) {
  if (args == nullptr || adj == nullptr) return hipErrorInvalidValue;
  const jxhip_fused_args& a = *args;
  if (a.n_progs <= 0) return hipSuccess;
  if (int e = check_args(a)) return e;
  if (a.syn_model >= 0 || a.ckpt_v == nullptr || a.ckpt_every <= 0) return hipErrorInvalidValue;
  if (a.threads > 512) return hipErrorInvalidValue;  // fused_adjoint_kernel's launch bound
  const Fused<G, L> f = unpack<G, L>(a);
  Adj<G> g;
  g.g_rec = static_cast<const G*>(adj->g_rec);
  for (int j = 0; j < JXHIP_NPARAM; ++j) g.g_param[j] = static_cast<G*>(adj->g_param[j]);
  g.g_v0 = static_cast<G*>(adj->g_v0);
  g.wv = static_cast<double*>(adj->work_v);
  g.wm = static_cast<G*>(adj->work_m);
  g.wh = static_cast<G*>(adj->work_h);
  g.wn = static_cast<G*>(adj->work_n);
  g.lm = static_cast<G*>(adj->lam_m);
  g.lh = static_cast<G*>(adj->lam_h);
  g.ln = static_cast<G*>(adj->lam_n);
  g.work_stride = adj->work_stride;
  const dim3 grid(static_cast<unsigned>(a.n_progs)), block(a.threads);
  const size_t lds = static_cast<size_t>(a.lds_bytes);
  auto kern = fused_adjoint_kernel<G, L>;
  if (hipError_t e = set_lds(kern, lds)) return e;
  hipLaunchKernelGGL(kern, grid, block, lds, stream, f, g);
  return hipGetLastError();
}

}  // namespace jxhip

extern "C" {

int jxhip_fused_forward_f32(const jxhip_fused_args* args, void* stream) {
  return jxhip::launch_forward<float, float>(args, static_cast<hipStream_t>(stream));
}

int jxhip_fused_forward_f64(const jxhip_fused_args* args, void* stream) {
  return jxhip::launch_forward<double, double>(args, static_cast<hipStream_t>(stream));
}

int jxhip_fused_forward_mixed(const jxhip_fused_args* args, void* stream) {
  return jxhip::launch_forward<double, float>(args, static_cast<hipStream_t>(stream));
}

int jxhip_fused_adjoint_mixed(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                              void* stream) {
  return jxhip::launch_adjoint<double, float>(args, adj, static_cast<hipStream_t>(stream));
}

int jxhip_fused_adjoint_f32(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                            void* stream) {
  return jxhip::launch_adjoint<float, float>(args, adj, static_cast<hipStream_t>(stream));
}

int jxhip_fused_adjoint_f64(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                            void* stream) {
  return jxhip::launch_adjoint<double, double>(args, adj, static_cast<hipStream_t>(stream));
}

}  // extern "C"

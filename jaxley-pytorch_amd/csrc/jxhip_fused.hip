// Fused whole-simulation kernel for gfx950: every workgroup ("program") owns S
// simulations of one topology for a range of time steps and runs, per step,
//   synapses (regular layout)  -> channels (HH, Leak) + stimulus -> implicit Euler
//   system -> leaves-first elimination by height -> back-substitution -> record,
// i.e. `_step_with_plan` (modules/base.py:3214-3400) with the DHS solve
// (solver_voltage.py:294-359), inside the time loop of `integrate`
// (integrate.py:624-662). Voltages, the diagonal and the right-hand side live in LDS
// for the whole launch ([node][S], simulations fastest); gates and synapse states are
// streamed from HBM in the same [node][S] program layout (coalesced). Simulations are
// independent, so no workgroup ever waits for another: there is no grid-wide sync.
#include "jxhip_common.h"
#include "../../include/jxhip.h"

namespace jxhip {

template <typename T>
struct Fused {
  const int64_t* progs;
  const int32_t* topos;
  const int32_t* order;
  const int32_t* level_ptr;
  const int32_t* node_i;
  const T* node_f;
  const int32_t* child_idx;
  const int32_t* stim_col;
  const T* stim_scale;
  const int32_t* rec_node;
  const T* param[JXHIP_NPARAM];
  int32_t param_mode[JXHIP_NPARAM];
  T* v;
  T* gm;
  T* gh;
  T* gn;
  const T* stim;
  int64_t stim_cols;
  int64_t stim_b;
  T* rec;
  int32_t syn_model;
  const int32_t* syn_pre;
  const int32_t* syn_ptr;
  const int32_t* tgt_node;
  const T* syn_param[JXHIP_NSYNPARAM];
  const T* syn_scale;
  T* syn_s;
  T dt;
  T diff;
  T* ckpt_v;
  T* ckpt_m;
  T* ckpt_h;
  T* ckpt_n;
  T* ckpt_s;
  int64_t ckpt_every;
  int64_t ckpt_stride;
  int64_t ckpt_syn_stride;
};

template <typename T>
Fused<T> unpack(const jxhip_fused_args& a) {
  Fused<T> f;
  f.progs = a.progs;
  f.topos = a.topos;
  f.order = a.order;
  f.level_ptr = a.level_ptr;
  f.node_i = a.node_i;
  f.node_f = static_cast<const T*>(a.node_f);
  f.child_idx = a.child_idx;
  f.stim_col = a.stim_col;
  f.stim_scale = static_cast<const T*>(a.stim_scale);
  f.rec_node = a.rec_node;
  for (int j = 0; j < JXHIP_NPARAM; ++j) {
    f.param[j] = static_cast<const T*>(a.param[j]);
    f.param_mode[j] = a.param_mode[j];
  }
  f.v = static_cast<T*>(a.v);
  f.gm = static_cast<T*>(a.gate_m);
  f.gh = static_cast<T*>(a.gate_h);
  f.gn = static_cast<T*>(a.gate_n);
  f.stim = static_cast<const T*>(a.stim);
  f.stim_cols = a.stim_cols;
  f.stim_b = a.stim_b;
  f.rec = static_cast<T*>(a.rec);
  f.syn_model = a.syn_model;
  f.syn_pre = a.syn_pre;
  f.syn_ptr = a.syn_ptr;
  f.tgt_node = a.tgt_node;
  for (int j = 0; j < JXHIP_NSYNPARAM; ++j) f.syn_param[j] = static_cast<const T*>(a.syn_param[j]);
  f.syn_scale = static_cast<const T*>(a.syn_scale);
  f.syn_s = static_cast<T*>(a.syn_s);
  f.dt = static_cast<T>(a.dt);
  f.diff = static_cast<T>(a.diff);
  f.ckpt_v = static_cast<T*>(a.ckpt_v);
  f.ckpt_m = static_cast<T*>(a.ckpt_m);
  f.ckpt_h = static_cast<T*>(a.ckpt_h);
  f.ckpt_n = static_cast<T*>(a.ckpt_n);
  f.ckpt_s = static_cast<T*>(a.ckpt_s);
  f.ckpt_every = a.ckpt_every;
  f.ckpt_stride = a.ckpt_stride;
  f.ckpt_syn_stride = a.ckpt_syn_stride;
  return f;
}

// Program / topology descriptors (column meaning in jxhip.h).
struct Prog {
  int32_t topo, nsim;
  int64_t t0, t1, soff, simoff, roff, synoff;
};
struct Topo {
  int32_t n, H, nb, lb, cb, n_rec, rb, n_tgt, tgb, spb, synb, n_syn;
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
  o.H = r[1];
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
  return o;
}

enum : int32_t { kComp = 1, kHH = 2, kLeak = 4 };

template <typename T, int S>
__device__ __forceinline__ T param_at(const Fused<T>& a, int j, const Prog& pg, const Topo& tp,
                                      int k, int s) {
  const int32_t mode = a.param_mode[j];
  if (mode == 0) return a.param[j][tp.nb + k];
  if (mode == 1) return a.param[j][pg.simoff + s];
  return a.param[j][pg.soff + static_cast<int64_t>(k) * S + s];
}

// Synapse step for one (target, simulation): the regular-layout synapse step
// (modules/network.py:382-478, kernels/synapse_step.py:121-182). Returns the summed
// slope (lin) and minus the summed offset (cst) in uA/cm^2 (per mV).
template <typename T, int S>
__device__ __forceinline__ void synapse_target(const Fused<T>& a, const Prog& pg, const Topo& tp,
                                               const T* xs, int tgt, int s, T vp, T& lin,
                                               T& cst) {
  T sl = T(0), sc = T(0);
  const T dt = a.dt;
  // Synapses of a target: [syn_ptr[t], syn_ptr[t+1]) (sorted by target; regular layouts
  // have syn_ptr[t] = t k).
  const int32_t* sptr = a.syn_ptr + tp.spb;
  for (int syn = sptr[tgt]; syn < sptr[tgt + 1]; ++syn) {
    const int64_t gsyn = tp.synb + syn;
    const int pre = a.syn_pre[gsyn];
    const T vpre = xs[pre * S + s];
    const T scale = a.syn_scale[gsyn];
    const T gs = a.syn_param[0][gsyn];
    T vt, ct;
    if (a.syn_model == 0) {
      const T e_syn = a.syn_param[1][gsyn];
      const T alpha = a.syn_param[2][gsyn];
      const T beta = a.syn_param[3][gsyn];
      const T v_th = a.syn_param[4][gsyn];
      const T delta = a.syn_param[5][gsyn];
      T* sp = a.syn_s + pg.synoff + static_cast<int64_t>(syn) * S + s;
      const T st = *sp;
      const T tr = sigmoid((vpre - v_th) / delta);
      const T aa = alpha * tr;
      const T rate = aa + beta;
      const T sinf = aa / rate;
      const T snew = sinf + (st - sinf) * dexp(-dt * rate);
      *sp = snew;
      // i = gS s (v_post - e_syn) is linear in v_post: slope gS s, offset -gS s e_syn.
      vt = gs * snew * scale;
      ct = -vt * e_syn;
    } else {
      const T xo = a.syn_param[1][gsyn];
      const T slope = a.syn_param[2][gsyn];
      // The reference perturbs v_pre and v_post by `diff` (network.py:452-457); the
      // slope is the difference of sigmoid((v_pre - x_off) slope) at v_pre and
      // v_pre + diff, computed without cancellation:
      //   sig(b) - sig(a) = (1 - sig(a)) sig(b) (-expm1(-(b - a))).
      const T za = (vpre - xo) * slope;
      const T zb = ((vpre + a.diff) - xo) * slope;
      const T sa = sigmoid(za);
      const T sb = sigmoid(zb);
      const T dsig = (T(1) - sa) * sb * (-dexpm1(-(zb - za)));
      const T d0 = -gs * sa * scale;
      vt = -gs * dsig * scale / a.diff;
      ct = d0 - vt * vp;
    }
    sl += vt;
    sc += ct;
  }
  lin = sl;
  cst = -sc;
}

// LDS of a program: xs (voltages / rhs / solution) and ds (diagonal), [node][S]; then
// the synapse slope and offset per (target, simulation).
template <typename T, int S>
struct Lds {
  T* xs;
  T* ds;
  T* tl;
  T* tc;
};

// Gate pointers of one step: read at `*_in`, written at `*_out` (program blocks).
template <typename T>
struct Gates {
  const T* m_in;
  const T* h_in;
  const T* n_in;
  T* m_out;
  T* h_out;
  T* n_out;
};

// Triangularization, leaves first (a node gathers its already eliminated children), and
// back-substitution, roots first. `kTransposed` swaps lower and upper (A^T).
template <typename T, int S, bool kTransposed>
__device__ __forceinline__ void tree_solve(const Fused<T>& a, const Topo& tp, T* xs, T* ds) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int32_t* ni = a.node_i + static_cast<int64_t>(tp.nb) * JXHIP_NODE_ICOLS;
  const T* nf = a.node_f + static_cast<int64_t>(tp.nb) * JXHIP_NODE_FCOLS;
  const int32_t* order = a.order + tp.nb;
  const int32_t* lptr = a.level_ptr + tp.lb;
  const int32_t* cidx = a.child_idx + tp.cb;
  constexpr int kLo = kTransposed ? 1 : 0;
  constexpr int kUp = kTransposed ? 0 : 1;
  for (int h = 1; h <= tp.H; ++h) {
    const int lo = lptr[h], w = (lptr[h + 1] - lo) * S;
    for (int i = tid; i < w; i += nthr) {
      const int k = order[lo + i / S], s = i % S;
      const int32_t* nik = ni + k * JXHIP_NODE_ICOLS;
      T dk = ds[k * S + s], xk = xs[k * S + s];
      for (int j = nik[1]; j < nik[2]; ++j) {
        const int c = cidx[j];
        const T f = nf[c * JXHIP_NODE_FCOLS + kUp] / ds[c * S + s];
        dk -= nf[c * JXHIP_NODE_FCOLS + kLo] * f;
        xk -= xs[c * S + s] * f;
      }
      ds[k * S + s] = dk;
      xs[k * S + s] = xk;
    }
    __syncthreads();
  }
  for (int h = tp.H; h >= 0; --h) {
    const int lo = lptr[h], w = (lptr[h + 1] - lo) * S;
    for (int i = tid; i < w; i += nthr) {
      const int k = order[lo + i / S], s = i % S;
      const int p = ni[k * JXHIP_NODE_ICOLS + 0];
      T xk = xs[k * S + s];
      if (p >= 0) xk -= nf[k * JXHIP_NODE_FCOLS + kLo] * xs[p * S + s];
      xs[k * S + s] = xk / ds[k * S + s];
    }
    __syncthreads();
  }
}

// One time step t of `_step_with_plan`: synapses, channels + stimulus, the implicit
// Euler system and its solve. On entry xs holds v(t); on exit v(t+1).
template <typename T, int S>
__device__ __forceinline__ void step_body(const Fused<T>& a, const Prog& pg, const Topo& tp,
                                          const Lds<T, S>& L, int64_t t, const Gates<T>& G) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nS = tp.n * S;
  const T dt = a.dt;
  const int32_t* ni = a.node_i + static_cast<int64_t>(tp.nb) * JXHIP_NODE_ICOLS;
  const T* nf = a.node_f + static_cast<int64_t>(tp.nb) * JXHIP_NODE_FCOLS;
  T* xs = L.xs;
  T* ds = L.ds;
  // Synapses (they read the voltages of the previous step).
  if (a.syn_model >= 0) {
    for (int i = tid; i < tp.n_tgt * S; i += nthr) {
      const int tgt = i / S, s = i % S;
      const int node = a.tgt_node[tp.tgb + tgt];
      T lin = T(0), cst = T(0);
      if (s < pg.nsim) synapse_target<T, S>(a, pg, tp, xs, tgt, s, xs[node * S + s], lin, cst);
      L.tl[i] = lin;
      L.tc[i] = cst;
    }
    __syncthreads();
  }
  // Channels, stimulus and the implicit Euler system (solver_voltage.py:250-291).
  for (int i = tid; i < nS; i += nthr) {
    const int k = i / S, s = i % S;
    const int32_t* nik = ni + k * JXHIP_NODE_ICOLS;
    const int32_t flags = nik[3];
    const T dbase = nf[k * JXHIP_NODE_FCOLS + 2];
    if (!(flags & kComp) || s >= pg.nsim) {
      ds[i] = dbase;
      xs[i] = T(0);
      continue;
    }
    const T inv_cm = nf[k * JXHIP_NODE_FCOLS + 3];
    const T v = xs[i];
    T lin = T(0), cst = T(0);
    if (flags & kHH) {
      T am, bm, ah, bh, an, bn;
      hh_rates(v, am, bm, ah, bh, an, bn);
      const T mm = gate_step(G.m_in[i], dt, am, bm);
      const T hh = gate_step(G.h_in[i], dt, ah, bh);
      const T nn = gate_step(G.n_in[i], dt, an, bn);
      G.m_out[i] = mm;
      G.h_out[i] = hh;
      G.n_out[i] = nn;
      const T gNa = param_at<T, S>(a, 0, pg, tp, k, s) * (mm * mm * mm) * hh;
      const T nn2 = nn * nn;
      const T gK = param_at<T, S>(a, 1, pg, tp, k, s) * (nn2 * nn2);
      const T gL = param_at<T, S>(a, 2, pg, tp, k, s);
      // I = gNa (v - eNa) + gK (v - eK) + gL (v - eL) is linear in v: its slope is the
      // total conductance and the constant term is gNa eNa + gK eK + gL eL (x1000:
      // mA/cm^2 -> uA/cm^2, modules/base.py:3500-3502).
      lin += T(1000) * (gNa + gK + gL);
      cst += T(1000) * (gNa * param_at<T, S>(a, 3, pg, tp, k, s) +
                        gK * param_at<T, S>(a, 4, pg, tp, k, s) +
                        gL * param_at<T, S>(a, 5, pg, tp, k, s));
    } else if (G.m_out != G.m_in) {
      G.m_out[i] = G.m_in[i];
      G.h_out[i] = G.h_in[i];
      G.n_out[i] = G.n_in[i];
    }
    if (flags & kLeak) {
      const T gl = param_at<T, S>(a, 6, pg, tp, k, s);
      lin += T(1000) * gl;
      cst += T(1000) * gl * param_at<T, S>(a, 7, pg, tp, k, s);
    }
    const int tg = nik[6];
    if (tg >= 0) {
      lin += L.tl[tg * S + s];
      cst += L.tc[tg * S + s];
    }
    T rhs = v + dt * cst * inv_cm;
    for (int e = nik[4]; e < nik[5]; ++e) {
      const int64_t col = a.stim_col[e];
      const int64_t bidx = a.stim_b > 1 ? (pg.simoff + s) : 0;
      rhs += a.stim[(t * a.stim_cols + col) * a.stim_b + bidx] * a.stim_scale[e];
    }
    ds[i] = dbase + dt * lin * inv_cm;
    xs[i] = rhs;
  }
  __syncthreads();
  tree_solve<T, S, false>(a, tp, xs, ds);
}

template <typename T, int S>
__global__ void __launch_bounds__(1024) fused_forward_kernel(Fused<T> a) {
  extern __shared__ unsigned char smem_raw[];
  const Prog pg = load_prog(a.progs, blockIdx.x);
  const Topo tp = load_topo(a.topos, pg.topo);
  const int nS = tp.n * S;
  Lds<T, S> L;
  L.xs = reinterpret_cast<T*>(smem_raw);
  L.ds = L.xs + nS;
  L.tl = L.ds + nS;
  L.tc = L.tl + tp.n_tgt * S;
  const int tid = threadIdx.x, nthr = blockDim.x;
  T* xs = L.xs;
  for (int i = tid; i < nS; i += nthr) xs[i] = a.v[pg.soff + i];
  __syncthreads();
  for (int i = tid; i < tp.n_rec * S; i += nthr) {
    const int r = i / S, s = i % S;
    a.rec[pg.roff + i] = xs[a.rec_node[tp.rb + r] * S + s];
  }
  Gates<T> G;
  G.m_in = G.m_out = a.gm + pg.soff;
  G.h_in = G.h_out = a.gh + pg.soff;
  G.n_in = G.n_out = a.gn + pg.soff;
  const int64_t nsteps = pg.t1 - pg.t0;
  for (int64_t step = 0; step < nsteps; ++step) {
    if (a.ckpt_v != nullptr && step % a.ckpt_every == 0) {
      const int64_t off = pg.soff + (step / a.ckpt_every) * a.ckpt_stride;
      for (int i = tid; i < nS; i += nthr) {
        a.ckpt_v[off + i] = xs[i];
        a.ckpt_m[off + i] = G.m_in[i];
        a.ckpt_h[off + i] = G.h_in[i];
        a.ckpt_n[off + i] = G.n_in[i];
      }
      if (a.syn_model == 0 && a.ckpt_s != nullptr) {
        const int64_t nsyn = static_cast<int64_t>(tp.n_syn) * S;
        const int64_t so = pg.synoff + (step / a.ckpt_every) * a.ckpt_syn_stride;
        for (int64_t i = tid; i < nsyn; i += nthr) a.ckpt_s[so + i] = a.syn_s[pg.synoff + i];
      }
    }
    step_body<T, S>(a, pg, tp, L, pg.t0 + step, G);
    for (int i = tid; i < tp.n_rec * S; i += nthr) {
      const int r = i / S, s = i % S;
      a.rec[pg.roff + ((step + 1) * tp.n_rec + r) * S + s] = xs[a.rec_node[tp.rb + r] * S + s];
    }
    __syncthreads();
  }
  for (int i = tid; i < nS; i += nthr) a.v[pg.soff + i] = xs[i];
}

// ---------------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------------

template <typename T>
struct Adj {
  const T* g_rec;
  T* g_param[JXHIP_NPARAM];
  T* g_v0;
  T* wv;
  T* wm;
  T* wh;
  T* wn;
  T* lm;
  T* lh;
  T* ln;
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

// Derivative of save_exp(u) = exp(min(u, 20)) with respect to u.
template <typename T>
__device__ __forceinline__ T d_save_exp(T u) {
  return u <= T(20) ? dexp(u) : T(0);
}

// dx'/dv of the exponential-Euler gate step through alpha(v), beta(v).
template <typename T>
__device__ __forceinline__ T d_gate_dv(T x, T dt, T al, T be, T dal, T dbe, T& e) {
  const T tau = T(1) / (al + be);
  const T xinf = al * tau;
  e = save_exp(-dt / tau);
  const T de = -dt * e;  // d e / d(alpha + beta)
  const T dxa = be * tau * tau * (T(1) - e) + (x - xinf) * de;
  const T dxb = -al * tau * tau * (T(1) - e) + (x - xinf) * de;
  return dxa * dal + dxb * dbe;
}

template <typename T, int S>
__global__ void __launch_bounds__(1024) fused_adjoint_kernel(Fused<T> a, Adj<T> g) {
  extern __shared__ unsigned char smem_raw[];
  const Prog pg = load_prog(a.progs, blockIdx.x);
  const Topo tp = load_topo(a.topos, pg.topo);
  const int nS = tp.n * S;
  Lds<T, S> L;
  L.xs = reinterpret_cast<T*>(smem_raw);
  L.ds = L.xs + nS;
  L.tl = L.ds + nS;
  L.tc = L.tl;
  T* lv = L.ds + nS;  // adjoint of the voltages (no synapses in the adjoint)
  T* xs = L.xs;
  T* ds = L.ds;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const T dt = a.dt;
  const int32_t* ni = a.node_i + static_cast<int64_t>(tp.nb) * JXHIP_NODE_ICOLS;
  const T* nf = a.node_f + static_cast<int64_t>(tp.nb) * JXHIP_NODE_FCOLS;
  const int64_t nsteps = pg.t1 - pg.t0;
  const int64_t K = a.ckpt_every;
  const int64_t C = (nsteps + K - 1) / K;
  const int64_t wbase = static_cast<int64_t>(blockIdx.x) * g.work_stride;
  T* lm = g.lm + pg.soff;
  T* lh = g.lh + pg.soff;
  T* ln = g.ln + pg.soff;
  for (int i = tid; i < nS; i += nthr) {
    lv[i] = T(0);
    lm[i] = T(0);
    lh[i] = T(0);
    ln[i] = T(0);
  }
  __syncthreads();
  for (int64_t c = C - 1; c >= 0; --c) {
    const int64_t ts = c * K;
    const int64_t te = ts + K < nsteps ? ts + K : nsteps;
    // Recompute the segment from its checkpoint; trajectory into work[0 .. te - ts].
    const int64_t coff = pg.soff + c * a.ckpt_stride;
    T* wv = g.wv + wbase;
    T* wm = g.wm + wbase;
    T* wh = g.wh + wbase;
    T* wn = g.wn + wbase;
    for (int i = tid; i < nS; i += nthr) {
      xs[i] = a.ckpt_v[coff + i];
      wv[i] = xs[i];
      wm[i] = a.ckpt_m[coff + i];
      wh[i] = a.ckpt_h[coff + i];
      wn[i] = a.ckpt_n[coff + i];
    }
    __syncthreads();
    for (int64_t j = 0; j < te - ts; ++j) {
      Gates<T> G;
      G.m_in = wm + j * nS;
      G.h_in = wh + j * nS;
      G.n_in = wn + j * nS;
      G.m_out = wm + (j + 1) * nS;
      G.h_out = wh + (j + 1) * nS;
      G.n_out = wn + (j + 1) * nS;
      step_body<T, S>(a, pg, tp, L, pg.t0 + ts + j, G);
      for (int i = tid; i < nS; i += nthr) wv[(j + 1) * nS + i] = xs[i];
      __syncthreads();
    }
    // Adjoint of the steps of the segment, last first.
    for (int64_t j = te - ts - 1; j >= 0; --j) {
      const int64_t t1 = ts + j + 1;  // local time of v(t+1) (recording index)
      // Gradient of the recording of v(t+1); one thread per simulation adds all
      // recordings (a node may be recorded more than once).
      for (int s = tid; s < S; s += nthr) {
        for (int r = 0; r < tp.n_rec; ++r) {
          const int node = a.rec_node[tp.rb + r];
          lv[node * S + s] += g.g_rec[pg.roff + (t1 * tp.n_rec + r) * S + s];
        }
      }
      __syncthreads();
      // Diagonal of step t (gates at t+1) and the right-hand side lambda_v(t+1).
      const T* m1 = wm + (j + 1) * nS;
      const T* h1 = wh + (j + 1) * nS;
      const T* n1 = wn + (j + 1) * nS;
      for (int i = tid; i < nS; i += nthr) {
        const int k = i / S, s = i % S;
        const int32_t flags = ni[k * JXHIP_NODE_ICOLS + 3];
        const T dbase = nf[k * JXHIP_NODE_FCOLS + 2];
        T lin = T(0);
        if ((flags & kComp) && s < pg.nsim) {
          if (flags & kHH) {
            const T mm = m1[i], hh = h1[i], nn = n1[i];
            const T nn2 = nn * nn;
            lin += T(1000) * (param_at<T, S>(a, 0, pg, tp, k, s) * (mm * mm * mm) * hh +
                              param_at<T, S>(a, 1, pg, tp, k, s) * (nn2 * nn2) +
                              param_at<T, S>(a, 2, pg, tp, k, s));
          }
          if (flags & kLeak) lin += T(1000) * param_at<T, S>(a, 6, pg, tp, k, s);
          ds[i] = dbase + dt * lin * nf[k * JXHIP_NODE_FCOLS + 3];
        } else {
          ds[i] = dbase;
        }
        xs[i] = lv[i];
      }
      __syncthreads();
      tree_solve<T, S, true>(a, tp, xs, ds);  // xs = mu = A^-T lambda_v(t+1)
      // Adjoint of the system build, the currents and the gates.
      const T* v0 = wv + j * nS;
      const T* v1 = wv + (j + 1) * nS;
      const T* m0 = wm + j * nS;
      const T* h0 = wh + j * nS;
      const T* n0 = wn + j * nS;
      for (int i = tid; i < nS; i += nthr) {
        const int k = i / S, s = i % S;
        const int32_t flags = ni[k * JXHIP_NODE_ICOLS + 3];
        if (!(flags & kComp) || s >= pg.nsim) {
          lv[i] = T(0);
          continue;
        }
        const T inv_cm = nf[k * JXHIP_NODE_FCOLS + 3];
        const T mu = xs[i];
        const T gl = -mu * v1[i] * dt * inv_cm;  // dL/dlin
        const T gc = mu * dt * inv_cm;           // dL/dcst
        T lvn = mu;                              // rhs = v(t) + ...
        if (flags & kHH) {
          const T gNa = param_at<T, S>(a, 0, pg, tp, k, s);
          const T gK = param_at<T, S>(a, 1, pg, tp, k, s);
          const T eNa = param_at<T, S>(a, 3, pg, tp, k, s);
          const T eK = param_at<T, S>(a, 4, pg, tp, k, s);
          const T eL = param_at<T, S>(a, 5, pg, tp, k, s);
          const T mm = m1[i], hh = h1[i], nn = n1[i];
          const T m3 = mm * mm * mm, n3 = nn * nn * nn, n4 = n3 * nn;
          const T wNa = T(1000) * (gl + gc * eNa);
          const T wK = T(1000) * (gl + gc * eK);
          const T wL = T(1000) * (gl + gc * eL);
          const int64_t pi = pg.soff + i;
          if (g.g_param[0]) g.g_param[0][pi] += wNa * m3 * hh;
          if (g.g_param[1]) g.g_param[1][pi] += wK * n4;
          if (g.g_param[2]) g.g_param[2][pi] += wL;
          if (g.g_param[3]) g.g_param[3][pi] += T(1000) * gc * gNa * m3 * hh;
          if (g.g_param[4]) g.g_param[4][pi] += T(1000) * gc * gK * n4;
          if (g.g_param[5]) g.g_param[5][pi] += T(1000) * gc * param_at<T, S>(a, 2, pg, tp, k, s);
          const T lam_m1 = lm[i] + wNa * gNa * T(3) * mm * mm * hh;
          const T lam_h1 = lh[i] + wNa * gNa * m3;
          const T lam_n1 = ln[i] + wK * gK * T(4) * n3;
          // Gate step at v(t) (channels/hh.py:87-108 and their derivatives).
          const T v = v0[i];
          const T xm = -(v + T(40)), xn = -(v + T(55));
          const T um = -(v + T(65)) / T(18), uh = -(v + T(65)) / T(20);
          const T ub = -(v + T(35)) / T(10), un = -(v + T(65)) / T(80);
          const T am = T(0.1) * x_over_expm1(xm, T(10));
          const T dam = -T(0.1) * d_x_over_expm1(xm, T(10));
          const T bm = T(4) * save_exp(um);
          const T dbm = -T(4) / T(18) * d_save_exp(um);
          const T ah = T(0.07) * save_exp(uh);
          const T dah = -T(0.07) / T(20) * d_save_exp(uh);
          const T Eb = save_exp(ub);
          const T bh = T(1) / (Eb + T(1));
          const T dbh = d_save_exp(ub) / (T(10) * (Eb + T(1)) * (Eb + T(1)));
          const T an = T(0.01) * x_over_expm1(xn, T(10));
          const T dan = -T(0.01) * d_x_over_expm1(xn, T(10));
          const T bn = T(0.125) * save_exp(un);
          const T dbn = -T(0.125) / T(80) * d_save_exp(un);
          T em, eh, en;
          const T dmv = d_gate_dv(m0[i], dt, am, bm, dam, dbm, em);
          const T dhv = d_gate_dv(h0[i], dt, ah, bh, dah, dbh, eh);
          const T dnv = d_gate_dv(n0[i], dt, an, bn, dan, dbn, en);
          lvn += lam_m1 * dmv + lam_h1 * dhv + lam_n1 * dnv;
          lm[i] = lam_m1 * em;
          lh[i] = lam_h1 * eh;
          ln[i] = lam_n1 * en;
        }
        if (flags & kLeak) {
          const T eL = param_at<T, S>(a, 7, pg, tp, k, s);
          const int64_t pi = pg.soff + i;
          if (g.g_param[6]) g.g_param[6][pi] += T(1000) * (gl + gc * eL);
          if (g.g_param[7]) g.g_param[7][pi] += T(1000) * gc * param_at<T, S>(a, 6, pg, tp, k, s);
        }
        lv[i] = lvn;
      }
      __syncthreads();
    }
  }
  // The initial recording (t = 0) and the gradient of the initial voltages.
  for (int s = tid; s < S; s += nthr) {
    for (int r = 0; r < tp.n_rec; ++r) {
      const int node = a.rec_node[tp.rb + r];
      lv[node * S + s] += g.g_rec[pg.roff + r * S + s];
    }
  }
  __syncthreads();
  for (int i = tid; i < nS; i += nthr) g.g_v0[pg.soff + i] = lv[i];
}

template <typename T>
int launch_forward(const jxhip_fused_args* args, hipStream_t stream) {
  if (args == nullptr) return hipErrorInvalidValue;
  const jxhip_fused_args& a = *args;
  if (a.n_progs <= 0) return hipSuccess;
  if (a.n_progs > 0x7fffffffLL || a.threads < 64 || a.threads > 1024 || a.threads % 64)
    return hipErrorInvalidValue;
  if (a.lds_bytes < 0 || a.lds_bytes > 160 * 1024) return hipErrorInvalidValue;
  const Fused<T> f = unpack<T>(a);
  const dim3 grid(static_cast<unsigned>(a.n_progs)), block(a.threads);
  const size_t lds = static_cast<size_t>(a.lds_bytes);
#define JXHIP_FWD_CASE(SV)                                                             \
  case SV: {                                                                           \
    auto kern = fused_forward_kernel<T, SV>;                                           \
    if (lds > 64 * 1024) {                                                             \
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),          \
                                         hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                         static_cast<int>(lds));                       \
      if (e != hipSuccess) return e;                                                   \
    }                                                                                  \
    hipLaunchKernelGGL(kern, grid, block, lds, stream, f);                             \
    break;                                                                             \
  }
  switch (a.S) {
    JXHIP_FWD_CASE(1)
    JXHIP_FWD_CASE(2)
    JXHIP_FWD_CASE(4)
    JXHIP_FWD_CASE(8)
    JXHIP_FWD_CASE(16)
    JXHIP_FWD_CASE(32)
    JXHIP_FWD_CASE(64)
    default:
      return hipErrorInvalidValue;
  }
#undef JXHIP_FWD_CASE
  return hipGetLastError();
}

template <typename T>
int launch_adjoint(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                   hipStream_t stream) {
  if (args == nullptr || adj == nullptr) return hipErrorInvalidValue;
  const jxhip_fused_args& a = *args;
  if (a.n_progs <= 0) return hipSuccess;
  if (a.n_progs > 0x7fffffffLL || a.threads < 64 || a.threads > 1024 || a.threads % 64)
    return hipErrorInvalidValue;
  if (a.syn_model >= 0 || a.ckpt_v == nullptr || a.ckpt_every <= 0) return hipErrorInvalidValue;
  if (a.lds_bytes < 0 || a.lds_bytes > 160 * 1024) return hipErrorInvalidValue;
  const Fused<T> f = unpack<T>(a);
  Adj<T> g;
  g.g_rec = static_cast<const T*>(adj->g_rec);
  for (int j = 0; j < JXHIP_NPARAM; ++j) g.g_param[j] = static_cast<T*>(adj->g_param[j]);
  g.g_v0 = static_cast<T*>(adj->g_v0);
  g.wv = static_cast<T*>(adj->work_v);
  g.wm = static_cast<T*>(adj->work_m);
  g.wh = static_cast<T*>(adj->work_h);
  g.wn = static_cast<T*>(adj->work_n);
  g.lm = static_cast<T*>(adj->lam_m);
  g.lh = static_cast<T*>(adj->lam_h);
  g.ln = static_cast<T*>(adj->lam_n);
  g.work_stride = adj->work_stride;
  const dim3 grid(static_cast<unsigned>(a.n_progs)), block(a.threads);
  const size_t lds = static_cast<size_t>(a.lds_bytes);
#define JXHIP_ADJ_CASE(SV)                                                             \
  case SV: {                                                                           \
    auto kern = fused_adjoint_kernel<T, SV>;                                           \
    if (lds > 64 * 1024) {                                                             \
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),          \
                                         hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                         static_cast<int>(lds));                       \
      if (e != hipSuccess) return e;                                                   \
    }                                                                                  \
    hipLaunchKernelGGL(kern, grid, block, lds, stream, f, g);                          \
    break;                                                                             \
  }
  switch (a.S) {
    JXHIP_ADJ_CASE(1)
    JXHIP_ADJ_CASE(2)
    JXHIP_ADJ_CASE(4)
    JXHIP_ADJ_CASE(8)
    JXHIP_ADJ_CASE(16)
    JXHIP_ADJ_CASE(32)
    JXHIP_ADJ_CASE(64)
    default:
      return hipErrorInvalidValue;
  }
#undef JXHIP_ADJ_CASE
  return hipGetLastError();
}

}  // namespace jxhip

extern "C" {

int jxhip_fused_adjoint_f32(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                            void* stream) {
  return jxhip::launch_adjoint<float>(args, adj, static_cast<hipStream_t>(stream));
}

int jxhip_fused_adjoint_f64(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                            void* stream) {
  return jxhip::launch_adjoint<double>(args, adj, static_cast<hipStream_t>(stream));
}

int jxhip_fused_forward_f32(const jxhip_fused_args* args, void* stream) {
  return jxhip::launch_forward<float>(args, static_cast<hipStream_t>(stream));
}

int jxhip_fused_forward_f64(const jxhip_fused_args* args, void* stream) {
  return jxhip::launch_forward<double>(args, static_cast<hipStream_t>(stream));
}

}  // extern "C"

// Fused whole-simulation kernels for gfx950 (templates; the entry points for each
// element type are in jxhip_fused_{f32,f64,mixed}.hip).
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
#pragma once
#include <hip/hip_cooperative_groups.h>

#include <type_traits>

#include "jxhip_common.h"
#include "../../include/jxhip.h"

namespace jxhip {

// Column layout of the tables (see jxhip.h).
enum : int { kNiParent = 0, kNiChildBeg, kNiChildEnd, kNiFlags, kNiStimBeg, kNiStimEnd, kNiTgt };
enum : int { kNfLo = 0, kNfUp, kNfDbase, kNfInvCm };
enum : int32_t { kComp = 1, kHH = 2, kLeak = 4, kExport = 8, kPNa = 16, kPK = 32, kPKm = 64 };
constexpr int32_t kPosp = kPNa | kPK | kPKm;

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
  G* ax;               // streaming kernel: axial part of every item's residual
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
  const int32_t* syn_kind;
  G dt;
  G dtv;     // step of the voltage solve (dt, or dt / 2 for Crank-Nicolson)
  G rscale;  // right-hand-side factor (1, or 2 for Crank-Nicolson)
  int32_t expl;  // forward Euler: no solve (branchpoint rows only)
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
  const G* node_p2;  // (n_nodes_total, JXHIP_NPARAM2) Pospischil parameters, or null
  G* gx;             // Pospischil gates m, h (Na), n (K), p (Km) at q * gx_stride + item
  int64_t gx_stride;
  // Split networks: programs per group of S simulations, nodes of a group, and with coop
  // the grid-wide barrier per step and the exported voltages by step parity.
  int32_t parts;
  int32_t coop;
  int64_t group_nodes;
  double* vx;
  int64_t vx_stride;
  G* traj_v;           // voltage trajectory of the checkpointed forward (gate type), or null
  int64_t traj_steps;
  const L* node_sym;   // (symmetric coupling coefficient, row scale) per node, or null
  unsigned int* gsync;      // lockstep programs: one counter per group (this launch's), or null
  unsigned int* gsync_err;  // set when a group barrier gave up waiting
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
  f.ax = static_cast<G*>(a.work_ax);
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
  f.node_p2 = static_cast<const G*>(a.node_p2);
  f.gx = static_cast<G*>(a.gate_ext);
  f.gx_stride = a.gate_ext_stride;
  f.syn_s = static_cast<G*>(a.syn_s);
  f.syn_kind = a.syn_model == JXHIP_SYN_MIXED ? a.syn_kind : nullptr;
  f.dt = static_cast<G>(a.dt);
  f.dtv = static_cast<G>(a.dt_v > 0 ? a.dt_v : a.dt);
  f.rscale = static_cast<G>(a.rhs_scale > 0 ? a.rhs_scale : 1.0);
  f.expl = a.explicit_step;
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
  f.parts = a.parts > 1 ? a.parts : 1;
  f.coop = a.coop;
  f.group_nodes = a.group_nodes;
  f.vx = static_cast<double*>(a.vx);
  f.vx_stride = a.vx_stride;
  f.traj_v = static_cast<G*>(a.traj_v);
  f.traj_steps = a.traj_steps;
  f.node_sym = static_cast<const L*>(a.node_sym);
  f.gsync = a.group_sync;
  f.gsync_err = a.group_sync != nullptr ? a.group_sync + a.n_progs / f.parts : nullptr;
  return f;
}

struct Prog {
  int32_t topo, nsim;
  int64_t t0, t1, soff, simoff, roff, synoff;
};

// Barrier of every workgroup of a cooperative launch (the programs of split networks):
// the writes of all workgroups before it are visible to all after it.
__device__ __forceinline__ void grid_barrier() { cooperative_groups::this_grid().sync(); }

// Barrier of the lockstep programs of one group (the parts of a split network run for
// one group of simulations): the writes of all of them before it are visible to all
// after it. With a.gsync the group's programs count their arrivals on their own counter
// (no wait on the other groups, no contention with them); without it the whole grid
// synchronizes. `nbar` counts the barriers this program has passed (uniform). The wait
// gives up after about 2^26 polls and sets the error word (the host reports it).
// kCoherentOnly: the only data crossing the barrier was written with store_coherent and
// is read with load_coherent (device-scope accesses that each L2 passes through to the
// device's coherence point); the barrier then waits for every thread's own stores to
// complete instead of the device-scope fences, which write back and invalidate the whole
// L2 of the XCD and made the forward's step of a split network 15-16 us slower.
template <bool kCoherentOnly = false, typename F>
__device__ __forceinline__ void group_barrier(const F& a, unsigned int& nbar) {
  if (a.gsync == nullptr) {
    grid_barrier();
    return;
  }
  ++nbar;
  if (kCoherentOnly) __builtin_amdgcn_s_waitcnt(0);  // this thread's stores acknowledged
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int* ctr = a.gsync + blockIdx.x / a.parts;
    const unsigned int target = nbar * static_cast<unsigned int>(a.parts);
    if (!kCoherentOnly) __threadfence();
    atomicAdd(ctr, 1u);
    unsigned int polls = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++polls == (1u << 26)) {
        atomicExch(a.gsync_err, 1u);
        break;
      }
    }
    if (!kCoherentOnly) __threadfence();
  }
  __syncthreads();
}

// Device-scope relaxed store and load of one exported value (see group_barrier).
template <typename V>
__device__ __forceinline__ void store_coherent(V* p, V x) {
  __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename V>
__device__ __forceinline__ V load_coherent(const V* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
struct Topo {
  int32_t n, L, nb, lb, cb, n_rec, rb, n_tgt, tgb, spb, synb, n_syn, sb, scb, ppb;
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
  o.ppb = r[14];
  return o;
}

// Reciprocal of a pivot (diagonals of a diagonally dominant system, > 0): hardware rcp
// in float32, hardware estimate + two Newton steps in float64.
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return drcp(x); }

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

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
// `v`: the program's voltages (program layout); `s_in` / `s_out`: the program's synapse
// states before / after the step ([syn][S]; the same array in the forward). The target's
// synapses j0, j0 + jstep, ... (jstep <= 8: the pointer list's 8 repeated end entries
// keep every read of it in range).
// kCoh: vpre_g is a split network's exported copy, read with load_coherent.
template <typename G, typename L, typename V, bool kCoh = false>
__device__ __forceinline__ void synapse_target(const Fused<G, L>& a, const Prog& pg,
                                               const Topo& tp, int tgt, int s, G& lin, G& cur,
                                               const V* vpre_g, const V* v,
                                               const G* s_in, G* s_out, int j0 = 0,
                                               int jstep = 1) {
  G sl = G(0), sc = G(0);
  const G dt = a.dt;
  const int32_t* jdp = a.syn_ptr + tp.spb;
  const int S = a.S;
  for (int j = j0;; j += jstep) {
    // The target's j-th synapse (jagged-diagonal order: consecutive targets, consecutive
    // synapses).
    const int beg = jdp[j];
    if (tgt >= jdp[j + 1] - beg) break;
    const int syn = beg + tgt;
    const int64_t gsyn = tp.synb + syn;
    const int pre = a.syn_pre[gsyn];
    const int64_t ip = static_cast<int64_t>(pre) * S + s;
    const G vpre = static_cast<G>(kCoh ? load_coherent(vpre_g + ip) : vpre_g[ip]);
    const G scale = a.syn_scale[gsyn];
    const G gs = a.syn_param[0][gsyn];
    const int model = a.syn_kind != nullptr ? a.syn_kind[gsyn] : a.syn_model;
    if (model == 0) {
      const int node = a.tgt_node[tp.tgb + tgt];
      const G e_syn = a.syn_param[1][gsyn];
      const G alpha = a.syn_param[2][gsyn];
      const G beta = a.syn_param[3][gsyn];
      const G v_th = a.syn_param[4][gsyn];
      const G delta = a.syn_param[5][gsyn];
      const int64_t si = static_cast<int64_t>(syn) * S + s;
      const G st = s_in[si];
      const G tr = sigmoid((vpre - v_th) / delta);
      const G aa = alpha * tr;
      const G rate = aa + beta;
      const G sinf = aa / rate;
      const G snew = sinf + (st - sinf) * dexp(-dt * rate);
      s_out[si] = snew;
      // i = gS s (v_post - e_syn) is linear in v_post: the slope is gS s.
      const G g = gs * snew * scale;
      sl += g;
      sc += g * static_cast<G>(v[node * S + s] - static_cast<double>(e_syn));
    } else if (model == 2) {
      // IonotropicSynapse (synapses/ionotropic.py:44-80): u = (v_pre - v_th) / delta,
      // s_inf = 1 / (1 + save_exp(-u)), rate = k_minus (1 + exp(clamp(u, -20, 50))),
      // s' = s e + s_inf (1 - e) with e = save_exp(-rate dt); i = gS s' (v_post - e_syn),
      // linear in v_post (slope gS s').
      const int node = a.tgt_node[tp.tgb + tgt];
      const G e_syn = a.syn_param[1][gsyn];
      const G k_minus = a.syn_param[2][gsyn];
      const G v_th = a.syn_param[3][gsyn];
      const G delta = a.syn_param[4][gsyn];
      const int64_t si = static_cast<int64_t>(syn) * S + s;
      const G st = s_in[si];
      const G u = (vpre - v_th) / delta;
      const G sinf = G(1) / (G(1) + save_exp(-u));
      const G uc = u < G(-20) ? G(-20) : (u > G(50) ? G(50) : u);
      const G rate = k_minus * (G(1) + dexp(uc));
      const G e = save_exp(-rate * dt);
      const G snew = st * e + sinf * (G(1) - e);
      s_out[si] = snew;
      const G g = gs * snew * scale;
      sl += g;
      sc += g * static_cast<G>(v[node * S + s] - static_cast<double>(e_syn));
    } else if (model == 3) {
      // TestSynapse (synapses/test.py:25-59): s_bar = 1 / (1 + save_exp((v_th - v_pre) /
      // delta)), tau = (1 - s_bar) / k_minus, c' = c e + s_bar (1 - e) with e =
      // save_exp(-dt / tau); v_th = -35, delta = 10, k_minus = 1 / 40; i = gC c' v_post.
      const int node = a.tgt_node[tp.tgb + tgt];
      const int64_t si = static_cast<int64_t>(syn) * S + s;
      const G st = s_in[si];
      const G sbar = G(1) / (G(1) + save_exp((G(-35) - vpre) / G(10)));
      const G tau = (G(1) - sbar) / G(1.0 / 40.0);
      const G e = save_exp((G(-1) / tau) * dt);
      const G snew = st * e + sbar * (G(1) - e);
      s_out[si] = snew;
      const G g = gs * snew * scale;
      sl += g;
      sc += g * static_cast<G>(v[node * S + s]);
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

// Solve schedule of a topology staged in LDS: segment level pointers (L + 2), segments
// (beg, end, child_beg, child_end), the parent of every segment top, the child segment
// tops.
struct Sched {
  const int* lptr;
  const int4* sq;
  const int* sp;
  const int* sch;
  int lw;  // levels lw .. L hold at most 64 (segment, simulation) pairs each
};

// First level from which every level holds at most 64 (segment, simulation) pairs.
template <typename G, typename L>
__device__ __forceinline__ int wave_levels(const Fused<G, L>& a, const Topo& tp) {
  int lw = tp.L + 1;
  while (lw > 0 && ((a.seg_lptr[tp.lb + lw] - a.seg_lptr[tp.lb + lw - 1]) << a.log2S) <= 64) --lw;
  return lw;
}


// Bytes of a staged schedule (a multiple of 16).
__host__ __device__ __forceinline__ int sched_bytes(int nseg, int L, int ncs) {
  return (16 * nseg + 4 * (L + 2 + nseg + ncs) + 15) & ~15;
}

// LDS of a program.
template <typename L>
struct Lds {
  L* xs;   // rhs, then the increment (solution)
  L* ds;   // diagonal, then its reciprocal after elimination
  L* tl;   // synapse slope per (target, simulation)
  L* tc;   // synapse current per (target, simulation)
  L* cl;   // lower / upper coefficient per node
  L* cs;   // symmetric coupling coefficient per node (solve of the row-scaled system)
  Sched sc;
};

struct SchedSize {
  int nseg, ncs, bytes;
};

template <typename G, typename L>
__device__ __forceinline__ SchedSize schedule_size(const Fused<G, L>& a, const Topo& tp) {
  SchedSize z;
  z.nseg = a.seg_lptr[tp.lb + tp.L + 1];
  z.ncs = z.nseg > 0 ? a.seg[(static_cast<int64_t>(tp.sb) + z.nseg - 1) * 4 + 3] : 0;
  z.bytes = sched_bytes(z.nseg, tp.L, z.ncs);
  return z;
}

// Copies the schedule of the topology into LDS at `base` (16-byte aligned).
template <typename G, typename L>
__device__ __forceinline__ Sched stage_schedule(const Fused<G, L>& a, const Topo& tp,
                                                const SchedSize& z, unsigned char* base) {
  int4* sq = reinterpret_cast<int4*>(base);
  int* lp = reinterpret_cast<int*>(sq + z.nseg);
  int* sp = lp + tp.L + 2;
  int* sch = sp + z.nseg;
  const int4* gsq = reinterpret_cast<const int4*>(a.seg + static_cast<int64_t>(tp.sb) * 4);
  const int32_t* ni = a.node_i + static_cast<int64_t>(tp.nb) * JXHIP_NODE_ICOLS;
  for (int q = threadIdx.x; q < z.nseg; q += blockDim.x) {
    const int4 r = gsq[q];
    sq[q] = r;
    sp[q] = ni[r.x * JXHIP_NODE_ICOLS + kNiParent];
  }
  for (int l = threadIdx.x; l < tp.L + 2; l += blockDim.x) lp[l] = a.seg_lptr[tp.lb + l];
  for (int j = threadIdx.x; j < z.ncs; j += blockDim.x) sch[j] = a.seg_child[tp.scb + j];
  Sched sc;
  sc.lptr = lp;
  sc.sq = sq;
  sc.sp = sp;
  sc.sch = sch;
  sc.lw = wave_levels(a, tp);
  return sc;
}

// Lower / upper coefficient pair of node k in one load; with kT (solve with A^T) the two
// swap roles. kCS = 0: packed pairs (LDS copy); kCS = 1: the node table rows (lo, up,
// dbase, inv_cm) in HBM / L2; kCS = 2: one coefficient per node, lo = up (the system
// with every row scaled so that it is symmetric).
template <typename L, bool kT, int kCS = 0>
__device__ __forceinline__ void coef(const L* cl, int k, L& lo, L& up) {
  if constexpr (kCS == 2) {
    lo = up = cl[k];
  } else {
    using V = typename Vec2<L>::type;
    const V c = reinterpret_cast<const V*>(cl)[k << kCS];
    lo = kT ? c.y : c.x;
    up = kT ? c.x : c.y;
  }
}

// Eliminates segment q of simulation s, its child segments being eliminated: merges the
// child segment tops into the bottom node, then walks the chain bottom-up. ds keeps
// 1 / (eliminated diagonal), xs the eliminated rhs. The walk is latency-bound (each node
// depends on the previous one): the LDS operands of kB nodes (kC children) are loaded
// together before their dependent updates.
template <typename L, bool kT, int kCS>
__device__ __forceinline__ void elim_segment(const Sched& sc, int q, int s, int lS, L* xs, L* ds,
                                             const L* cl) {
  constexpr int kB = 4;
  const int4 sg = sc.sq[q];
  const int beg = sg.x, end = sg.y, cb = sg.z, nch = sg.w - sg.z;
  const int id = ((end - 1) << lS) + s;
  L dk = ds[id], xk = xs[id];
  constexpr int kC = 2;
  for (int j0 = 0; j0 < nch; j0 += kC) {
    int c4[kC];
    L lo4[kC], up4[kC], dc4[kC], xc4[kC];
#pragma unroll
    for (int u = 0; u < kC; ++u) c4[u] = sc.sch[cb + (j0 + u < nch ? j0 + u : nch - 1)];
#pragma unroll
    for (int u = 0; u < kC; ++u) {
      coef<L, kT, kCS>(cl, c4[u], lo4[u], up4[u]);
      dc4[u] = ds[(c4[u] << lS) + s];
      xc4[u] = xs[(c4[u] << lS) + s];
    }
#pragma unroll
    for (int u = 0; u < kC; ++u) {
      if (j0 + u < nch) {
        // ds[c] holds 1 / d[c] of the already eliminated child segment top.
        const L f = up4[u] * dc4[u];
        dk -= lo4[u] * f;
        xk -= xc4[u] * f;
      }
    }
  }
  L inv = recip(dk);
  ds[id] = inv;
  xs[id] = xk;
  // Node k uses the coefficients of k + 1 and its own diagonal and rhs: full batches of
  // kB nodes without guards, then the remaining nodes one by one.
  int k = end - 2;
  for (; k - (kB - 1) >= beg; k -= kB) {
    L lo8[kB], up8[kB], d8[kB], x8[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      coef<L, kT, kCS>(cl, k - u + 1, lo8[u], up8[u]);
      d8[u] = ds[((k - u) << lS) + s];
      x8[u] = xs[((k - u) << lS) + s];
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const L f = up8[u] * inv;
      const L d2 = d8[u] - lo8[u] * f;
      xk = x8[u] - xk * f;
      inv = recip(d2);
      const int idu = ((k - u) << lS) + s;
      ds[idu] = inv;
      xs[idu] = xk;
    }
  }
  for (; k >= beg; --k) {
    L lo1, up1;
    coef<L, kT, kCS>(cl, k + 1, lo1, up1);
    const int id1 = (k << lS) + s;
    const L f = up1 * inv;
    const L d2 = ds[id1] - lo1 * f;
    xk = xs[id1] - xk * f;
    inv = recip(d2);
    ds[id1] = inv;
    xs[id1] = xk;
  }
}

// Back-substitution of segment q of simulation s, its parent segment being solved.
template <typename L, bool kT, int kCS>
__device__ __forceinline__ void back_segment(const Sched& sc, int q, int s, int lS, L* xs,
                                             const L* ds, const L* cl) {
  constexpr int kB = 4;
  const int4 sg = sc.sq[q];
  const int beg = sg.x, end = sg.y;
  const int p = sc.sp[q];
  // Roots have zero coupling coefficients, so xp = 0 there adds nothing.
  L xp = p >= 0 ? xs[(p << lS) + s] : L(0);
  int k = beg;
  for (; k + (kB - 1) < end; k += kB) {
    L lo8[kB], d8[kB], x8[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      L up_unused;
      coef<L, kT, kCS>(cl, k + u, lo8[u], up_unused);
      d8[u] = ds[((k + u) << lS) + s];
      x8[u] = xs[((k + u) << lS) + s];
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      xp = (x8[u] - lo8[u] * xp) * d8[u];
      xs[((k + u) << lS) + s] = xp;
    }
  }
  for (; k < end; ++k) {
    L lo1, up_unused;
    coef<L, kT, kCS>(cl, k, lo1, up_unused);
    const int id1 = (k << lS) + s;
    xp = (xs[id1] - lo1 * xp) * ds[id1];
    xs[id1] = xp;
  }
}

// Orders the LDS accesses of one wavefront between two levels it processes alone.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Tree solve of all S simulations of a program, segment level by segment level: leaves
// first for the triangularization, root first for the back-substitution, one thread per
// (segment, simulation). The upper levels with at most 64 pairs each (sc.lw .. L) are
// processed by the first wavefront alone, without workgroup barriers. kT: solve with A^T
// (lower <-> upper).
template <typename G, typename L, bool kT, int kCS = 0>
__device__ __forceinline__ void tree_solve(const Fused<G, L>& a, const Topo& tp, const Sched& sc,
                                           L* xs, L* ds, const L* cl) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int S = a.S, lS = a.log2S;
  int lo = sc.lptr[0];
  for (int lv = 0; lv < sc.lw; ++lv) {
    const int hi = sc.lptr[lv + 1], w = (hi - lo) << lS;
    for (int i = tid; i < w; i += nthr) elim_segment<L, kT, kCS>(sc, lo + (i >> lS), i & (S - 1), lS, xs, ds, cl);
    lo = hi;
    __syncthreads();
  }
  if (sc.lw <= tp.L) {
    if (tid < 64) {
      for (int lv = sc.lw; lv <= tp.L; ++lv) {
        const int hi = sc.lptr[lv + 1], w = (hi - lo) << lS;
        if (tid < w) elim_segment<L, kT, kCS>(sc, lo + (tid >> lS), tid & (S - 1), lS, xs, ds, cl);
        lo = hi;
        wave_sync();
      }
      for (int lv = tp.L; lv >= sc.lw; --lv) {
        const int l0 = sc.lptr[lv], w = (sc.lptr[lv + 1] - l0) << lS;
        if (tid < w) back_segment<L, kT, kCS>(sc, l0 + (tid >> lS), tid & (S - 1), lS, xs, ds, cl);
        wave_sync();
      }
    }
    __syncthreads();
  }
  for (int lv = sc.lw - 1; lv >= 0; --lv) {
    const int l0 = sc.lptr[lv], w = (sc.lptr[lv + 1] - l0) << lS;
    for (int i = tid; i < w; i += nthr) back_segment<L, kT, kCS>(sc, l0 + (i >> lS), i & (S - 1), lS, xs, ds, cl);
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

// Axial part of the residual of item (k, s), lo_k (x_k - x_parent) + sum_c up_c (x_k - x_c),
// in float64 from values x in program layout (voltages, or increments of voltages); the
// coefficients as in `coef` (kCS).
template <int kCS = 0, typename G, typename L, typename X>
__device__ __forceinline__ double axial(const Fused<G, L>& a, const Topo& tp, const int32_t* nik,
                                        int k, const L* cl, const X* x, int s, double xk) {
  const int lS = a.log2S;
  double ax = 0.0;
  const int p = nik[kNiParent];
  if (p >= 0) {
    ax += static_cast<double>(cl[k << (kCS + 1)]) * (xk - static_cast<double>(x[(p << lS) + s]));
  }
  const int32_t* cidx = a.child_idx + tp.cb;
  for (int j = nik[kNiChildBeg]; j < nik[kNiChildEnd]; ++j) {
    const int c = cidx[j];
    ax += static_cast<double>(cl[(c << (kCS + 1)) + 1]) * (xk - static_cast<double>(x[(c << lS) + s]));
  }
  return ax;
}

// Forward Euler (step_voltage_explicit, solver_voltage.py:412-502) replaces the tree
// solve: a compartment's increment is its residual (dt times the vector field, with the
// branchpoints at the conductance-weighted mean of their neighbours), and a branchpoint
// solves its own row for the neighbours' increments, dv_k = (r_k - lo_k dv_p - sum_c
// up_c dv_c) / d_k, which puts it at the weighted mean of their new voltages. A
// branchpoint's neighbours are compartments, so the items do not depend on each other.
template <int kCS = 0, typename G, typename L>
__device__ void branchpoint_pass(const Fused<G, L>& a, const Topo& tp, const int32_t* ni,
                                 const Lds<L>& Ld) {
  const int lS = a.log2S, S = a.S;
  const int nS = tp.n << lS;
  const int32_t* cidx = a.child_idx + tp.cb;
  for (int i = threadIdx.x; i < nS; i += blockDim.x) {
    const int k = i >> lS, s = i & (S - 1);
    const int32_t* nik = ni + k * JXHIP_NODE_ICOLS;
    if (nik[kNiFlags] & kComp) continue;
    double num = static_cast<double>(Ld.xs[i]);
    const int p = nik[kNiParent];
    if (p >= 0) num -= static_cast<double>(Ld.cl[k << (kCS + 1)]) * static_cast<double>(Ld.xs[(p << lS) + s]);
    for (int j = nik[kNiChildBeg]; j < nik[kNiChildEnd]; ++j) {
      const int c = cidx[j];
      num -= static_cast<double>(Ld.cl[(c << (kCS + 1)) + 1]) * static_cast<double>(Ld.xs[(c << lS) + s]);
    }
    Ld.xs[i] = static_cast<L>(num / static_cast<double>(Ld.ds[i]));
  }
  __syncthreads();
}

// Forward Euler: every branchpoint voltage set to the conductance-weighted mean of its
// neighbours (the value the reference computes from the compartments at every step),
// before the axial residuals are formed at the start of a launch.
template <int kCS = 0, typename G, typename L>
__device__ void branchpoint_means(const Fused<G, L>& a, const Topo& tp, const int32_t* ni,
                                  const L* cl, double* vg) {
  const int lS = a.log2S, S = a.S;
  const int nS = tp.n << lS;
  const int32_t* cidx = a.child_idx + tp.cb;
  for (int i = threadIdx.x; i < nS; i += blockDim.x) {
    const int k = i >> lS, s = i & (S - 1);
    const int32_t* nik = ni + k * JXHIP_NODE_ICOLS;
    if (nik[kNiFlags] & kComp) continue;
    double num = 0.0, den = 0.0;
    const int p = nik[kNiParent];
    if (p >= 0) {
      const double w = static_cast<double>(cl[k << (kCS + 1)]);
      num += w * vg[(p << lS) + s];
      den += w;
    }
    for (int j = nik[kNiChildBeg]; j < nik[kNiChildEnd]; ++j) {
      const int c = cidx[j];
      const double w = static_cast<double>(cl[(c << (kCS + 1)) + 1]);
      num += w * vg[(c << lS) + s];
      den += w;
    }
    if (den != 0.0) vg[i] = num / den;
  }
  __syncthreads();
}

// Pospischil et al. (2008) Na, K and Km channels (channels/pospischil.py:74-250) of one
// item: parameters from the node's row of node_p2 (Na gNa, eNa, vt, K gK, eK, Km gKm,
// taumax), gates read from and written to gx (m, h of Na, n of K, p of Km; q * gx_stride
// + item), each advanced at the old voltage by exponential Euler; adds the conductances
// to g and g (v - E) to cur. Compartments without these channels never enter.
template <typename G, typename L>
__device__ __forceinline__ void pospischil_step(const Fused<G, L>& a, int32_t flags, int64_t row,
                                                int64_t item, double v, G& g, G& cur) {
  const G* pr = a.node_p2 + row * JXHIP_NPARAM2;
  const G dt = a.dt;
  const G vv = static_cast<G>(v);
  G* gx = a.gx + item;
  const int64_t st = a.gx_stride;
  const G u = vv - pr[2];  // v - vt
  if (flags & kPNa) {
    const G am = G(0.32) * x_over_expm1(G(-0.25) * (u - G(13)), G(1)) / G(0.25);
    const G bm = G(0.28) * x_over_expm1(G(0.2) * (u - G(40)), G(1)) / G(0.2);
    const G ah = G(0.128) * save_exp(-(u - G(17)) / G(18));
    const G bh = G(4) / (save_exp(-(u - G(40)) / G(5)) + G(1));
    const G m = gate_step(gx[0], dt, am, bm);
    const G h = gate_step(gx[st], dt, ah, bh);
    gx[0] = m;
    gx[st] = h;
    const G gna = pr[0] * (m * m * m) * h;
    g += gna;
    cur += gna * static_cast<G>(v - static_cast<double>(pr[1]));
  }
  if (flags & kPK) {
    const G an = G(0.032) * x_over_expm1(G(-0.2) * (u - G(15)), G(1)) / G(0.2);
    const G bn = G(0.5) * save_exp(-(u - G(10)) / G(40));
    const G n = gate_step(gx[2 * st], dt, an, bn);
    gx[2 * st] = n;
    const G n2 = n * n;
    const G gk = pr[3] * (n2 * n2);
    g += gk;
    cur += gk * static_cast<G>(v - static_cast<double>(pr[4]));
  }
  if (flags & kPKm) {
    const G vp = vv + G(35);
    const G pinf = G(1) / (G(1) + save_exp(G(-0.1) * vp));
    const G tau = pr[6] / (G(3.3) * save_exp(G(0.05) * vp) + save_exp(G(-0.05) * vp));
    const G e = save_exp(-G(1) / tau * dt);
    const G p = gx[3 * st] * e + pinf * (G(1) - e);
    gx[3 * st] = p;
    const G gkm = pr[5] * p;
    g += gkm;
    cur += gkm * static_cast<G>(v - static_cast<double>(pr[4]));
  }
}

// System of one (node, simulation) item from its voltage v and the axial part ax of its
// residual: diagonal d and residual r = b - A v into LDS. The gates are read from /
// written to the references (registers or memory).
template <typename G, typename L>
__device__ __forceinline__ void build_core(const Fused<G, L>& a, const Prog& pg, const Topo& tp,
                                           const Lds<L>& Ld, const int32_t* nik, int k, int s,
                                           int i, double v, double ax, int64_t t, G& m, G& h,
                                           G& n) {
  const int lS = a.log2S;
  const L* nfk = a.node_f + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_FCOLS;
  const int32_t flags = nik[kNiFlags];
  if (!(flags & kComp)) {
    Ld.ds[i] = nfk[kNfDbase];
    Ld.xs[i] = static_cast<L>(a.rscale * static_cast<G>(ax));
    return;
  }
  const G dt = a.dt;
  const G inv_cm = static_cast<G>(nfk[kNfInvCm]);
  const G vv = static_cast<G>(v);
  G g = G(0);    // total conductance (mS/cm^2 -> x1000 below)
  G cur = G(0);  // ionic current sum g_i (v - E_i), uA/cm^2 / 1000
  if (flags & kHH) {
    G am, bm, ah, bh, an, bn;
    hh_rates_step(vv, am, bm, ah, bh, an, bn);
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
  if (flags & kPosp) pospischil_step(a, flags, static_cast<int64_t>(tp.nb) + k, pg.soff + i, v, g, cur);
  G lin = G(1000) * g;
  G net = -G(1000) * cur;  // (cst - lin v) of the channels
  const int tg = nik[kNiTgt];
  if (tg >= 0) {
    lin += static_cast<G>(Ld.tl[(tg << lS) + s]);
    net -= static_cast<G>(Ld.tc[(tg << lS) + s]);
  }
  G r = static_cast<G>(ax) + a.dtv * inv_cm * net;
  for (int e = nik[kNiStimBeg]; e < nik[kNiStimEnd]; ++e) {
    const int64_t col = a.stim_col[e];
    const int64_t bidx = a.stim_b > 1 ? (pg.simoff + s) : 0;
    r += a.stim[(t * a.stim_cols + col) * a.stim_b + bidx] * a.stim_scale[e];
  }
  Ld.ds[i] = nfk[kNfDbase] + static_cast<L>(a.dtv * inv_cm * lin);
  Ld.xs[i] = static_cast<L>(a.rscale * r);
}

// build_core with the axial residual from the float64 voltages vg (program layout).
template <int kCS = 0, typename G, typename L>
__device__ __forceinline__ void build_item(const Fused<G, L>& a, const Prog& pg, const Topo& tp,
                                           const Lds<L>& Ld, const double* vg, int k, int s,
                                           double v, int64_t t, G& m, G& h, G& n) {
  const int32_t* nik = a.node_i + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS;
  const double ax = axial<kCS>(a, tp, nik, k, Ld.cl, vg, s, v);
  build_core(a, pg, tp, Ld, nik, k, s, (k << a.log2S) + s, v, ax, t, m, h, n);
}

// Inputs of one (node, simulation) item of the streaming build, in two parts: its state
// in HBM (voltage, gates, axial part of the residual), loaded one item ahead of its use
// so that the HBM latency overlaps the arithmetic of the previous item, and its node's
// tables (flags, stimulus range, synapse target, diagonal base, 1 / cm, channel
// parameters; small, cache-resident), loaded at the start of the item's own turn and
// before the next item's state, so that waiting for them does not wait for that state
// (vector memory counters complete in order). The arithmetic itself then loads nothing
// but LDS (and the stimulus of stimulated nodes).
template <typename G>
struct ItemState {
  double v;
  G m, h, n, ax;
};

template <typename G, typename L>
struct ItemConst {
  int32_t flags, sbeg, send, tgt;
  L dbase, inv_cm;
  L w;  // row scale of the symmetric system (kSym), else 1
  G p[JXHIP_NPARAM];
};

template <typename G, typename L>
__device__ __forceinline__ ItemState<G> load_state(const Fused<G, L>& a, const Prog& pg, int i) {
  const int64_t gi = pg.soff + i;
  ItemState<G> st;
  st.v = a.v[gi];
  st.m = a.gm[gi];
  st.h = a.gh[gi];
  st.n = a.gn[gi];
  st.ax = a.ax[gi];
  return st;
}

template <bool kSym = false, typename G, typename L>
__device__ __forceinline__ ItemConst<G, L> load_const(const Fused<G, L>& a, const Prog& pg,
                                                      const Topo& tp, int i) {
  using VG = typename Vec2<G>::type;
  using VL = typename Vec2<L>::type;
  const int k = i >> a.log2S, s = i & (a.S - 1);
  const int64_t row = static_cast<int64_t>(tp.nb) + k;
  ItemConst<G, L> c;
  const int4* ni4 = reinterpret_cast<const int4*>(a.node_i + row * JXHIP_NODE_ICOLS);
  const int4 r0 = ni4[0], r1 = ni4[1];
  c.flags = r0.w;
  c.sbeg = r1.x;
  c.send = r1.y;
  c.tgt = r1.z;
  const VL f = reinterpret_cast<const VL*>(a.node_f + row * JXHIP_NODE_FCOLS)[1];
  c.dbase = f.x;
  c.inv_cm = f.y;
  c.w = kSym ? a.node_sym[row * 2 + 1] : L(1);
  const G* prow = a.node_p + row * JXHIP_NPARAM;
  const VG* np = reinterpret_cast<const VG*>(prow);
#pragma unroll
  for (int q = 0; q < JXHIP_NPARAM / 2; ++q) {
    const VG w = np[q];
    c.p[2 * q] = w.x;
    c.p[2 * q + 1] = w.y;
  }
  // Batched conductances (one load from a uniformly selected row, no branch): per
  // simulation (mode 1) or per item (mode 2); mode 0 reads the node row again and keeps
  // the node's values.
  const G* brow = a.bmode == 1   ? a.bparam + ((pg.simoff + s) << 2)
                  : a.bmode == 2 ? a.bparam + ((pg.soff + i) << 2)
                                 : prow;
  const VG b0 = reinterpret_cast<const VG*>(brow)[0], b1 = reinterpret_cast<const VG*>(brow)[1];
  const bool bat = a.bmode != 0;
  c.p[0] = bat ? b0.x : c.p[0];
  c.p[1] = bat ? b0.y : c.p[1];
  c.p[2] = bat ? b1.x : c.p[2];
  c.p[6] = bat ? b1.y : c.p[6];
  return c;
}

// build_core from loaded inputs: diagonal and residual r = b - A v of item i into LDS,
// the updated gates into `it`. kSym: the row scaled by c.w (the symmetric system).
template <bool kSym = false, typename G, typename L>
__device__ __forceinline__ void build_loaded(const Fused<G, L>& a, const Prog& pg,
                                             const Topo& tp, const Lds<L>& Ld, ItemState<G>& it,
                                             const ItemConst<G, L>& c, int i, int s,
                                             int64_t t) {
  if (!(c.flags & kComp)) {
    Ld.ds[i] = kSym ? c.w * c.dbase : c.dbase;
    Ld.xs[i] = kSym ? c.w * static_cast<L>(a.rscale * it.ax) : static_cast<L>(a.rscale * it.ax);
    return;
  }
  const int lS = a.log2S;
  const G dt = a.dt;
  const G inv_cm = static_cast<G>(c.inv_cm);
  const double v = it.v;
  const G vv = static_cast<G>(v);
  G g = G(0);
  G cur = G(0);
  if (c.flags & kHH) {
    G am, bm, ah, bh, an, bn;
    hh_rates_step(vv, am, bm, ah, bh, an, bn);
    it.m = gate_step(it.m, dt, am, bm);
    it.h = gate_step(it.h, dt, ah, bh);
    it.n = gate_step(it.n, dt, an, bn);
    const G gNa = c.p[0] * (it.m * it.m * it.m) * it.h;
    const G n2 = it.n * it.n;
    const G gK = c.p[1] * (n2 * n2);
    const G gL = c.p[2];
    g += gNa + gK + gL;
    cur += gNa * static_cast<G>(v - static_cast<double>(c.p[3])) +
           gK * static_cast<G>(v - static_cast<double>(c.p[4])) +
           gL * static_cast<G>(v - static_cast<double>(c.p[5]));
  }
  if (c.flags & kLeak) {
    const G gl = c.p[6];
    g += gl;
    cur += gl * static_cast<G>(v - static_cast<double>(c.p[7]));
  }
  if (c.flags & kPosp)
    pospischil_step(a, c.flags, static_cast<int64_t>(tp.nb) + (i >> a.log2S), pg.soff + i, v, g, cur);
  G lin = G(1000) * g;
  G net = -G(1000) * cur;
  if (c.tgt >= 0) {
    lin += static_cast<G>(Ld.tl[(c.tgt << lS) + s]);
    net -= static_cast<G>(Ld.tc[(c.tgt << lS) + s]);
  }
  G r = it.ax + a.dtv * inv_cm * net;
  for (int e = c.sbeg; e < c.send; ++e) {
    const int64_t col = a.stim_col[e];
    const int64_t bidx = a.stim_b > 1 ? (pg.simoff + s) : 0;
    r += a.stim[(t * a.stim_cols + col) * a.stim_b + bidx] * a.stim_scale[e];
  }
  const L d = c.dbase + static_cast<L>(a.dtv * inv_cm * lin);
  const L x = static_cast<L>(a.rscale * r);
  Ld.ds[i] = kSym ? c.w * d : d;
  Ld.xs[i] = kSym ? c.w * x : x;
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

// Synapse step of every (target, simulation) of a program; the presynaptic voltages are
// read from the group block vpre_g (the program's own block of a.v, or the exported
// voltages of all parts of a split network).
// log2 of the threads per (target, simulation) pair of a synapse pass over `pairs` pairs:
// up to 8 while the pairs still take one pass of the workgroup.
__device__ __forceinline__ int syn_lanes_log2(int pairs) {
  int l = 0;
  while (l < 3 && (pairs << (l + 1)) <= static_cast<int>(blockDim.x)) ++l;
  return l;
}

// Sum of x over the 2^l adjacent lanes of a pair, in a fixed order (every lane of the
// group ends with the same bits: each addition pairs the same two operands).
template <typename T>
__device__ __forceinline__ T pair_sum(T x, int l) {
  for (int o = 1; o < (1 << l); o <<= 1) x += __shfl_xor(x, o);
  return x;
}

template <bool kCoh, typename G, typename L>
__device__ __forceinline__ void synapse_phase(const Fused<G, L>& a, const Prog& pg, const Topo& tp,
                                              const Lds<L>& Ld, const double* vpre_g) {
  if (a.syn_model < 0) return;
  const int lS = a.log2S, S = a.S;
  const int nts = tp.n_tgt << lS;
  // Threads per (target, simulation): thread `sub` of a pair walks the target's synapses
  // sub, sub + nsub, ... (a target's synapses are walked one after the other, each
  // waiting for its loads).
  const int lsub = syn_lanes_log2(nts);
  const int nsub = 1 << lsub;
  for (int i = threadIdx.x; i < (nts << lsub); i += blockDim.x) {
    const int pr = i >> lsub, sub = i & (nsub - 1);
    const int tgt = pr >> lS, s = pr & (S - 1);
    G lin = G(0), cur = G(0);
    if (s < pg.nsim) {
      G* st = a.syn_s + pg.synoff;
      synapse_target<G, L, double, kCoh>(a, pg, tp, tgt, s, lin, cur, vpre_g, a.v + pg.soff, st, st, sub,
                                        nsub);
    }
    lin = pair_sum(lin, lsub);
    cur = pair_sum(cur, lsub);
    if (sub == 0) {
      Ld.tl[pr] = static_cast<L>(lin);
      Ld.tc[pr] = static_cast<L>(cur);
    }
  }
  __syncthreads();
}

// LDS of a forward program: the staged schedule, the lower / upper coefficient per node
// (kCM = 0; kCM = 1: read from the node table in HBM / L2; kCM = 2: one symmetric
// coefficient per node for the solve, the pairs of the residual updates from the node
// table), rhs / increment and diagonal per (node, simulation), synapse terms per
// (target, simulation). Stages the schedule and the coefficients (the caller
// synchronizes).
template <typename G, typename L, int kCM = 0>
__device__ __forceinline__ Lds<L> forward_lds(const Fused<G, L>& a, unsigned char* smem, int nS,
                                              const Topo& tp, int lS) {
  Lds<L> Ld;
  const SchedSize z = schedule_size(a, tp);
  Ld.sc = stage_schedule(a, tp, z, smem);
  if constexpr (kCM != 0) {
    Ld.cl = const_cast<L*>(a.node_f + static_cast<int64_t>(tp.nb) * JXHIP_NODE_FCOLS);
    Ld.cs = Ld.cl;
    Ld.xs = reinterpret_cast<L*>(smem + z.bytes);
    if constexpr (kCM == 2) {
      // Rounded up to 16 bytes, as the pairs.
      Ld.cs = reinterpret_cast<L*>(smem + z.bytes);
      Ld.xs = Ld.cs + ((tp.n * sizeof(L) + 15) & ~15) / sizeof(L);
      const L* sym = a.node_sym + static_cast<int64_t>(tp.nb) * 2;
      for (int k = threadIdx.x; k < tp.n; k += blockDim.x) Ld.cs[k] = sym[2 * k];
    }
  } else {
    Ld.cl = reinterpret_cast<L*>(smem + z.bytes);
    Ld.cs = Ld.cl;
    Ld.xs = Ld.cl + 2 * tp.n;
    load_coefficients(a, tp, Ld.cl);
  }
  Ld.ds = Ld.xs + nS;
  Ld.tl = Ld.ds + nS;
  Ld.tc = Ld.tl + (tp.n_tgt << lS);
  return Ld;
}

template <typename G, typename L>
__device__ __forceinline__ void record_step(const Fused<G, L>& a, const Prog& pg, const Topo& tp,
                                            const double* vg, int64_t slot) {
  const int lS = a.log2S, S = a.S;
  for (int i = threadIdx.x; i < (tp.n_rec << lS); i += blockDim.x) {
    const int r = i >> lS, s = i & (S - 1);
    a.rec[pg.roff + (slot * tp.n_rec + r) * S + s] =
        static_cast<G>(vg[(a.rec_node[tp.rb + r] << lS) + s]);
  }
}

template <typename G, typename L>
__device__ __forceinline__ void checkpoint_syn(const Fused<G, L>& a, const Prog& pg, const Topo& tp,
                                               int64_t c) {
  if (a.syn_model != 1 && a.ckpt_s != nullptr) {
    const int64_t nsyn = static_cast<int64_t>(tp.n_syn) << a.log2S;
    const int64_t so = pg.synoff + c * a.ckpt_syn_stride;
    for (int64_t i = threadIdx.x; i < nsyn; i += blockDim.x) a.ckpt_s[so + i] = a.syn_s[pg.synoff + i];
  }
}

// Node data of the axial change of one item after the solve: its node row (parent, child
// range, flags), own lower coefficient and the upper coefficient of node k + 1 (its child
// when it has exactly one).
struct UpdConst {
  int4 r0;
  double lo, upc;
};

template <int kCS, typename L>
__device__ __forceinline__ UpdConst load_update_const(const Topo& tp, const int32_t* ni, const L* cl,
                                                      int k) {
  const int kc = k + 1 < tp.n ? k + 1 : k;
  UpdConst u;
  u.r0 = reinterpret_cast<const int4*>(ni + k * JXHIP_NODE_ICOLS)[0];
  u.lo = static_cast<double>(cl[k << (kCS + 1)]);
  u.upc = static_cast<double>(cl[(kc << (kCS + 1)) + 1]);
  return u;
}

// Change of the axial part of the residual of item (k, s) for the increments x (LDS), as
// `axial`: the parent from the node row, the single child of a chain node at k + 1 (kernel
// order places a one-child node's child right after it), the child list otherwise.
template <int kCS, typename G, typename L>
__device__ __forceinline__ double apply_update(const Fused<G, L>& a, const Topo& tp,
                                               const UpdConst& u, int k, const L* cl, const L* x,
                                               int s, double xk) {
  const int lS = a.log2S;
  double ax = 0.0;
  if (u.r0.x >= 0) ax += u.lo * (xk - static_cast<double>(x[(u.r0.x << lS) + s]));
  if (u.r0.z - u.r0.y == 1) {
    ax += u.upc * (xk - static_cast<double>(x[((k + 1) << lS) + s]));
  } else {
    const int32_t* cidx = a.child_idx + tp.cb;
    for (int j = u.r0.y; j < u.r0.z; ++j) {
      const int c = cidx[j];
      ax += static_cast<double>(cl[(c << (kCS + 1)) + 1]) * (xk - static_cast<double>(x[(c << lS) + s]));
    }
  }
  return ax;
}

// The kernel's by-value argument block (the first and only explicit argument, at offset
// 0 of the kernarg segment) read through a pointer the compiler cannot prove constant:
// inside a loop its fields are re-read with scalar loads (scalar cache hits) where they
// are used instead of being held in SGPRs for the whole kernel, which with the float64
// item code spills SGPRs into VGPR lanes (a v_readlane per use).
template <typename T>
__device__ __forceinline__ const T& kernarg() {
  typedef const __attribute__((address_space(4))) T* KP;
  KP p = (KP)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return *(const T*)p;
}

// Streaming forward: every step reads the state of the program (voltages, gates and the
// axial part of each item's residual) from HBM and writes it back, for programs whose
// items do not fit the registers of their threads. The axial part is computed from the
// voltages once per launch and then updated from the increments of every solve (the
// neighbours' increments are in LDS), so the build of an item reads only its own state.
// After a solve, one pass over the items leaves each item's axial change in its pivot
// slot (free after the solve) and brings exported voltages (recordings, synapse ends) up
// to date in HBM; the next step's build applies the increment and the axial change from
// the item's own LDS slots when it loads the item's state (and takes the checkpoints).
// The last step applies every increment in that pass.
// kSyn: the topologies carry synapses; kCkpt: checkpoints are written for the adjoint;
// kCM, where the tree's coupling coefficients are: 0, a copy of the (lower, upper) pairs
// in LDS; 1, the node table (HBM / L2), which leaves the LDS to the system of more
// simulations; 2, the node table for the residual updates, and for the solve one
// coefficient per node in LDS, half the pairs' bytes: every row of the system is scaled
// by its node's w (w_child / w_parent = upper / lower of the edge), which makes the
// matrix symmetric without changing the solution.
template <typename G, typename L, bool kSyn, bool kCkpt, int kCM>
__global__ void __launch_bounds__(1024) fused_stream_kernel(Fused<G, L> a) {
  extern __shared__ unsigned char smem_raw[];
  const Prog pg = load_prog(a.progs, blockIdx.x);
  const Topo tp = load_topo(a.topos, pg.topo);
  const int lS = a.log2S, S = a.S;
  const int nS = tp.n << lS;
  const Lds<L> Ld = forward_lds<G, L, kCM>(a, smem_raw, nS, tp, lS);
  constexpr int kCS = kCM ? 1 : 0;
  constexpr bool kSym = kCM == 2;
  __syncthreads();
  const int tid = threadIdx.x, nthr = blockDim.x;
  double* vg = a.v + pg.soff;
  G* gm = a.gm + pg.soff;
  G* gh = a.gh + pg.soff;
  G* gn = a.gn + pg.soff;
  G* axg = a.ax + pg.soff;
  const int32_t* ni = a.node_i + static_cast<int64_t>(tp.nb) * JXHIP_NODE_ICOLS;
  // Every thread owns the items tid + j * nthr throughout, all of simulation s_own (the
  // thread count is a multiple of S); no barrier is needed between the writes and reads
  // of an item's own state or LDS slots.
  const int s_own = tid & (S - 1);
  if (a.expl) branchpoint_means<kCS>(a, tp, ni, Ld.cl, vg);
  for (int i = tid; i < nS; i += nthr) {
    const int k = i >> lS, s = i & (S - 1);
    axg[i] = static_cast<G>(axial<kCS>(a, tp, ni + k * JXHIP_NODE_ICOLS, k, Ld.cl, vg, s, vg[i]));
  }
  record_step(a, pg, tp, vg, 0);
  // The group's block of the state layout (the programs of a split network share it).
  const int64_t gbase = pg.soff - (static_cast<int64_t>(tp.nb) << lS);
  unsigned int nbar = 0;  // group barriers passed
  if (kSyn && a.coop) {
    // Split network: every part's voltages at the start into the exported copy of step 0.
    for (int i = tid; i < nS; i += nthr) store_coherent(a.vx + pg.soff + i, vg[i]);
    __syncthreads();
    group_barrier<true>(a, nbar);
  }
  const int64_t nsteps = pg.t1 - pg.t0;
  // The voltage trajectory: per group of simulations [T + 1][group nodes][S], the
  // program's nodes at its node base.
  const int64_t trow = static_cast<int64_t>(a.group_nodes > 0 ? a.group_nodes : tp.n) << lS;
  const int64_t tbase = (pg.soff - (static_cast<int64_t>(tp.nb) << lS)) * a.traj_steps +
                        (static_cast<int64_t>(tp.nb) << lS);
  for (int64_t step = 0; step < nsteps; ++step) {
    const bool pend = step > 0;  // the last solve's increments wait in LDS
    const bool ck = kCkpt && (step % a.ckpt_every == 0);
    const int64_t coff = ck ? pg.soff + (step / a.ckpt_every) * a.ckpt_stride : 0;
    if constexpr (kSyn && kCkpt) {
      if (ck) checkpoint_syn(a, pg, tp, step / a.ckpt_every);
    }
    if constexpr (kSyn) {
      if (a.coop)
        synapse_phase<true>(a, pg, tp, Ld, a.vx + (step & 1) * a.vx_stride + gbase);
      else
        synapse_phase<false>(a, pg, tp, Ld, a.v + gbase);
    }
    const int64_t t = pg.t0 + step;
    if (tid < nS) {
      ItemState<G> cur = load_state(a, pg, tid);
      for (int i = tid; i < nS; i += nthr) {
        const Fused<G, L>& a = kernarg<Fused<G, L>>();  // re-read per item (SGPR pressure)
        const int inx = i + nthr < nS ? i + nthr : i;
        const ItemConst<G, L> c = load_const<kSym>(a, pg, tp, i);
        __builtin_amdgcn_sched_barrier(0);
        const ItemState<G> nxt = load_state(a, pg, inx);
        __builtin_amdgcn_sched_barrier(0);
        if (s_own < pg.nsim) {
          if (pend) {
            if (!(c.flags & kExport)) cur.v += static_cast<double>(Ld.xs[i]);
            cur.ax += static_cast<G>(Ld.ds[i]);
          }
          if constexpr (kCkpt) {
            if (ck) {
              a.ckpt_v[coff + i] = cur.v;
              a.ckpt_m[coff + i] = cur.m;
              a.ckpt_h[coff + i] = cur.h;
              a.ckpt_n[coff + i] = cur.n;
            }
            if (a.traj_v != nullptr) a.traj_v[tbase + step * trow + i] = static_cast<G>(cur.v);
          }
          vg[i] = cur.v;
          axg[i] = cur.ax;
          build_loaded<kSym>(a, pg, tp, Ld, cur, c, i, s_own, t);
          gm[i] = cur.m;
          gh[i] = cur.h;
          gn[i] = cur.n;
        } else {
          const L d = c.dbase + ((c.flags & kComp) ? L(1) : L(0));
          Ld.ds[i] = kSym ? c.w * d : d;
          Ld.xs[i] = L(0);
        }
        cur = nxt;
      }
    }
    __syncthreads();
    if (a.expl) {
      branchpoint_pass<kCS>(a, tp, ni, Ld);
    } else {
      tree_solve<G, L, false, kSym ? 2 : kCS>(a, tp, Ld.sc, Ld.xs, Ld.ds, Ld.cs);
    }
    const bool last = step + 1 == nsteps;
    if (s_own < pg.nsim && tid < nS) {
      UpdConst c = load_update_const<kCS>(tp, ni, Ld.cl, tid >> lS);
      for (int i = tid; i < nS; i += nthr) {
        const int k = i >> lS;
        const int inx = i + nthr < nS ? i + nthr : i;
        const UpdConst cn = load_update_const<kCS>(tp, ni, Ld.cl, inx >> lS);  // next item's
        const double dv = static_cast<double>(Ld.xs[i]);
        if (last || (c.r0.w & kExport)) {
          vg[i] += dv;
          if (kSyn && a.coop) store_coherent(a.vx + ((step + 1) & 1) * a.vx_stride + pg.soff + i, vg[i]);
          if (kCkpt && last && a.traj_v != nullptr) a.traj_v[tbase + nsteps * trow + i] = static_cast<G>(vg[i]);
        }
        if (!last) Ld.ds[i] = static_cast<L>(apply_update<kCS>(a, tp, c, k, Ld.cl, Ld.xs, s_own, dv));
        c = cn;
      }
    }
    __syncthreads();
    // Split network: every part's exported voltages of the step written before any part
    // starts the next step's synapses.
    if (kSyn && a.coop) group_barrier<true>(a, nbar);
    record_step(a, pg, tp, vg, step + 1);
  }
}

// Register-resident forward. Every thread owns the items i = tid + j * blockDim (j < IPT)
// for the whole launch and keeps their voltage, gates and the axial part of their
// residual in VGPRs: after the solve the axial part is updated from the increments of the
// item and of its neighbours (in LDS), so no voltage is read from HBM inside the time
// loop. Only nodes flagged kExport (recorded nodes, synapse pre- and postsynaptic nodes)
// copy their voltage to HBM every step. The host guarantees n * S <= IPT * blockDim for
// every program of the launch (a program that does not fit is left untouched).
// The item loops are not unrolled (one copy of the item code keeps the kernel inside the
// instruction cache): the per-item arrays are indexed with the uniform loop counter,
// which the compiler lowers to VGPR indexing (the build promotes them to vectors).
template <typename G, typename L, int IPT, int MAXT, bool kSyn, bool kCkpt>
__global__ void __launch_bounds__(MAXT) fused_regs_kernel(Fused<G, L> a) {
  extern __shared__ unsigned char smem_raw[];
  const Prog pg = load_prog(a.progs, blockIdx.x);
  const Topo tp = load_topo(a.topos, pg.topo);
  const int lS = a.log2S, S = a.S;
  const int nS = tp.n << lS;
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (nS > IPT * nthr) return;
  const Lds<L> Ld = forward_lds(a, smem_raw, nS, tp, lS);
  __syncthreads();
  double* vg = a.v + pg.soff;
  G* gm = a.gm + pg.soff;
  G* gh = a.gh + pg.soff;
  G* gn = a.gn + pg.soff;
  const int32_t* ni = a.node_i + static_cast<int64_t>(tp.nb) * JXHIP_NODE_ICOLS;
  // The axial part of the residual accumulates in the gate type (float64 except in the
  // all-float32 mode).
  double vr[IPT];
  G ar[IPT], mr[IPT], hr[IPT], nr[IPT];
  if (a.expl) branchpoint_means(a, tp, ni, Ld.cl, vg);
#pragma unroll 1
  for (int j = 0; j < IPT; ++j) {
    const int i = tid + j * nthr;
    vr[j] = 0.0;
    ar[j] = mr[j] = hr[j] = nr[j] = G(0);
    if (i < nS) {
      const int k = i >> lS, s = i & (S - 1);
      vr[j] = vg[i];
      mr[j] = gm[i];
      hr[j] = gh[i];
      nr[j] = gn[i];
      ar[j] = static_cast<G>(axial(a, tp, ni + k * JXHIP_NODE_ICOLS, k, Ld.cl, vg, s, vr[j]));
    }
  }
  record_step(a, pg, tp, vg, 0);
  const int64_t nsteps = pg.t1 - pg.t0;
  // The voltage trajectory (layout as in the streaming kernel).
  const int64_t trow = static_cast<int64_t>(a.group_nodes > 0 ? a.group_nodes : tp.n) << lS;
  const int64_t tbase = (pg.soff - (static_cast<int64_t>(tp.nb) << lS)) * a.traj_steps +
                        (static_cast<int64_t>(tp.nb) << lS);
  for (int64_t step = 0; step < nsteps; ++step) {
    if constexpr (kCkpt) {
      if (step % a.ckpt_every == 0) {
        const int64_t c = step / a.ckpt_every;
        const int64_t off = pg.soff + c * a.ckpt_stride;
#pragma unroll 1
        for (int j = 0; j < IPT; ++j) {
          const int i = tid + j * nthr;
          if (i < nS) {
            a.ckpt_v[off + i] = vr[j];
            a.ckpt_m[off + i] = mr[j];
            a.ckpt_h[off + i] = hr[j];
            a.ckpt_n[off + i] = nr[j];
          }
        }
        if constexpr (kSyn) checkpoint_syn(a, pg, tp, c);
      }
      if (a.traj_v != nullptr) {
#pragma unroll 1
        for (int j = 0; j < IPT; ++j) {
          const int i = tid + j * nthr;
          if (i < nS) a.traj_v[tbase + step * trow + i] = static_cast<G>(vr[j]);
        }
      }
    }
    if constexpr (kSyn) synapse_phase<false>(a, pg, tp, Ld, a.v + pg.soff - (static_cast<int64_t>(tp.nb) << lS));
    const int64_t t = pg.t0 + step;
#pragma unroll 1
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * nthr;
      if (i < nS) {
        const int k = i >> lS, s = i & (S - 1);
        if (s < pg.nsim) {
          build_core(a, pg, tp, Ld, ni + k * JXHIP_NODE_ICOLS, k, s, i, vr[j], ar[j], t, mr[j],
                     hr[j], nr[j]);
        } else {
          idle_item(a, tp, Ld, k, i);
        }
      }
    }
    __syncthreads();
    if (a.expl) {
      branchpoint_pass(a, tp, ni, Ld);
    } else {
      tree_solve<G, L, false>(a, tp, Ld.sc, Ld.xs, Ld.ds, Ld.cl);
    }
#pragma unroll 1
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * nthr;
      if (i < nS) {
        const int k = i >> lS, s = i & (S - 1);
        const int32_t* nik = ni + k * JXHIP_NODE_ICOLS;
        const double dv = static_cast<double>(Ld.xs[i]);
        vr[j] += dv;
        ar[j] += static_cast<G>(axial(a, tp, nik, k, Ld.cl, Ld.xs, s, dv));
        if (nik[kNiFlags] & kExport) vg[i] = vr[j];
      }
    }
    __syncthreads();
    record_step(a, pg, tp, vg, step + 1);
  }
#pragma unroll 1
  for (int j = 0; j < IPT; ++j) {
    const int i = tid + j * nthr;
    if (i < nS) {
      vg[i] = vr[j];
      gm[i] = mr[j];
      gh[i] = hr[j];
      gn[i] = nr[j];
      if (kCkpt && a.traj_v != nullptr) a.traj_v[tbase + nsteps * trow + i] = static_cast<G>(vr[j]);
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
  G* lv;  // adjoint of the voltages when not kept in VGPRs (IPT == 0)
  int64_t work_stride;
  const int32_t* rec_first;  // per node: first recording, or INT32_MAX
  const int32_t* rec_next;   // per recording: the node's next one, or INT32_MAX
  // Synapses: by presynaptic node (CSR), recomputed states, state adjoints, presynaptic
  // shares, gradient of gS.
  const int32_t* pre_ptr;
  const int32_t* pre_pos;  // per synapse: its slot in the presynaptic grouping
  G* ws;
  int64_t work_syn_stride;
  G* lam_s;
  G* wq;
  G* g_gs;
  G* wl;  // per program [K][n_tgt][S]: the synaptic slope of each recomputed step
  int64_t work_lin_stride;
  int64_t wq_stride;  // elements between the two step-parity buffers of wq (0: one buffer)
};

// Derivatives of a KineticSynapse step (synapses/kinetic.py:37-53) at the presynaptic
// voltage vpre: the new state, E = ds'/ds and ds'/dvpre.
template <typename G>
__device__ __forceinline__ void kinetic_step_d(G st, G vpre, G alpha, G beta, G v_th, G delta, G dt,
                                               G& snew, G& E, G& dsdv) {
  const G tr = sigmoid((vpre - v_th) / delta);
  const G aa = alpha * tr;
  const G rate = aa + beta;
  const G rinv = G(1) / rate;
  const G sinf = aa * rinv;
  E = dexp(-dt * rate);
  snew = sinf + (st - sinf) * E;
  const G daa = aa * (G(1) - tr) / delta;     // d(alpha T) / dvpre
  const G dsinf = beta * rinv * rinv * daa;    // d s_inf / dvpre
  dsdv = dsinf * (G(1) - E) + (st - sinf) * (-dt * E * daa);
}

// The exponential-Euler gate step (the arithmetic of gate_step) with its derivatives:
// returns x'; dv = dx'/dv through alpha(v), beta(v); e = dx'/dx.
template <typename T>
__device__ __forceinline__ T gate_step_d(T x, T dt, T al, T be, T dal, T dbe, T& dv, T& e) {
  const T tau = ddiv(T(1), al + be);
  const T xinf = al * tau;
  e = save_exp(-dt * (al + be));
  const T de = -dt * e;
  const T dxa = be * tau * tau * (T(1) - e) + (x - xinf) * de;
  const T dxb = -al * tau * tau * (T(1) - e) + (x - xinf) * de;
  dv = dxa * dal + dxb * dbe;
  return fma(x - xinf, e, xinf);
}

// Reverse pass. Every thread owns the items i = tid + j * blockDim for the whole launch
// and keeps their voltage adjoints in VGPRs (IPT > 0 items per thread) or in HBM
// (IPT == 0: g.lv, program layout, for programs too large for the registers); LDS holds
// the staged schedule,
// the tree's coupling coefficients (or, kGC, they are read from the node table in L2),
// the system of the transposed solve and, per node, its first recording. Per K-step
// segment (last first): recompute the segment from its checkpoint into the work arrays,
// then run the adjoint of each step backwards.
// Launch bound of the reverse pass: 512 threads with the adjoints in VGPRs; with them in
// HBM (IPT 0) as many threads as the registers allow without spills, 768 with float64
// gates (170 VGPRs), 1024 with float32 (128).
template <typename G>
constexpr int adjoint_max_threads(int ipt) {
  return ipt > 0 ? 512 : (sizeof(G) == 8 ? 768 : 1024);
}

template <typename G, typename L, int IPT, bool kGC, bool kSyn, bool kTraj>
__global__ void __launch_bounds__(adjoint_max_threads<G>(IPT)) fused_adjoint_kernel(Fused<G, L> a, Adj<G> g) {
  extern __shared__ unsigned char smem_raw[];
  const Prog pg = load_prog(a.progs, blockIdx.x);
  const Topo tp = load_topo(a.topos, pg.topo);
  const int lS = a.log2S, S = a.S;
  const int nS = tp.n << lS;
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (IPT > 0 && nS > IPT * nthr) return;  // the host sizes IPT for every program
  constexpr int kJ = IPT > 0 ? IPT : 0x7fffffff;  // item loop bound (items end at nS)
  constexpr int kCS = kGC ? 1 : 0;
  const SchedSize z = schedule_size(a, tp);
  Lds<L> Ld;
  Ld.sc = stage_schedule(a, tp, z, smem_raw);
  unsigned char* lp = smem_raw + z.bytes;
  if constexpr (kGC) {
    Ld.cl = const_cast<L*>(a.node_f + static_cast<int64_t>(tp.nb) * JXHIP_NODE_FCOLS);
  } else {
    Ld.cl = reinterpret_cast<L*>(lp);
    lp += 2 * static_cast<size_t>(tp.n) * sizeof(L);
    load_coefficients(a, tp, Ld.cl);
  }
  Ld.xs = reinterpret_cast<L*>(lp);
  Ld.ds = Ld.xs + nS;
  // Synapses: per (target, simulation) the slope and current of the recomputed forward;
  // in the reverse, the target's dL/dlin and dL/dcst, then its postsynaptic share.
  Ld.tl = Ld.ds + nS;
  Ld.tc = Ld.tl + (tp.n_tgt << lS);
  constexpr int kNone = 0x7fffffff;
  const int ntS = tp.n_tgt << lS;
  const int64_t nsynS = static_cast<int64_t>(tp.n_syn) << lS;
  G* ws = kSyn ? g.ws + static_cast<int64_t>(blockIdx.x) * g.work_syn_stride : nullptr;
  G* lam_s = kSyn ? g.lam_s + pg.synoff : nullptr;
  // Presynaptic shares: per group of simulations, slots of the group's presynaptic
  // grouping (a split network's parts write each other's), by step parity.
  G* wq = kSyn ? g.wq + pg.synoff - (static_cast<int64_t>(tp.synb) << lS) : nullptr;
  G* wl = kSyn ? g.wl + static_cast<int64_t>(blockIdx.x) * g.work_lin_stride : nullptr;
  if constexpr (kSyn) {
    if (a.syn_model == 0) {
      for (int64_t q = tid; q < nsynS; q += nthr) lam_s[q] = G(0);
    }
  }
  const G dt = a.dt;
  const int64_t nsteps = pg.t1 - pg.t0;
  const int64_t K = a.ckpt_every;
  unsigned int nbar = 0;  // group barriers passed
  const int64_t C = (nsteps + K - 1) / K;
  // Recomputed trajectory: per group of simulations [K + 1][group nodes][S] (the parts of
  // a split network side by side, so presynaptic voltages of other parts are in reach);
  // the program's own nodes at its node base.
  const int64_t rowS = (a.group_nodes > 0 ? a.group_nodes : static_cast<int64_t>(tp.n)) << lS;
  const int64_t gwork = static_cast<int64_t>(blockIdx.x / a.parts) * g.work_stride;
  const int64_t wbase = gwork + (static_cast<int64_t>(tp.nb) << lS);
  G* lm = g.lm + pg.soff;
  G* lh = g.lh + pg.soff;
  G* ln = g.ln + pg.soff;
  L lvr[IPT > 0 ? IPT : 1];  // adjoint of the voltages of the thread's items (IPT > 0)
  G* lvg = g.lv + pg.soff;    // or in HBM (IPT == 0)
  // Gate adjoints: in VGPRs with the voltage adjoints (IPT > 0), else in HBM (lm, lh, ln).
  G lmr[IPT > 0 ? IPT : 1], lhr[IPT > 0 ? IPT : 1], lnr[IPT > 0 ? IPT : 1];
  // Conductances batched per simulation (bparam mode 1): every item of a thread belongs
  // to simulation tid mod S, so the thread sums the gradients of gNa, gK, gLeak (HH) and
  // gLeak (Leak) over its items and all steps in VGPRs; the sums are reduced across the
  // threads of a simulation in a fixed order at the end (no atomics).
  const bool simgrad = a.bmode == 1;
  G acc[4] = {G(0), G(0), G(0), G(0)};
  auto grad = [&](int j, int q, int64_t pi, G x) {
    if (g.g_param[j] == nullptr) return;
    if (q >= 0 && simgrad) {
      acc[q] += x;
    } else {
      g.g_param[j][pi] += x;
    }
  };
  auto lv_get = [&](int jj, int i) -> L {
    if constexpr (IPT > 0) {
      return lvr[jj];
    } else {
      return static_cast<L>(lvg[i]);
    }
  };
  auto lv_set = [&](int jj, int i, L x) {
    if constexpr (IPT > 0) {
      lvr[jj] = x;
    } else {
      lvg[i] = static_cast<G>(x);
    }
  };
#pragma unroll 1
  for (int j = 0; j < kJ; ++j) {
    const int i = tid + j * nthr;
    if (i >= nS) break;
    lv_set(j, i, L(0));
    if constexpr (IPT > 0) {
      lmr[j] = lhr[j] = lnr[j] = G(0);
    } else {
      lm[i] = G(0);
      lh[i] = G(0);
      ln[i] = G(0);
    }
  }
  // Adds the gradient of the recordings at slot t1 to the voltage adjoint of item i.
  auto add_rec = [&](int i, int64_t t1, L& lv) {
    const int k = i >> lS, s = i & (S - 1);
    for (int r = g.rec_first[tp.nb + k]; r != kNone; r = g.rec_next[tp.rb + r]) {
      lv += static_cast<L>(g.g_rec[pg.roff + (t1 * tp.n_rec + r) * S + s]);
    }
  };
  __syncthreads();
  double* wv = g.wv + wbase;
  const double* wvg = g.wv + gwork;  // the group's block (presynaptic voltages)
  G* wm = g.wm + wbase;
  G* wh = g.wh + wbase;
  G* wn = g.wn + wbase;
  // With the forward's voltage trajectory the segment's voltages are read from it and
  // only the gates (and synapse states) are recomputed: recurrences local to an item (a
  // synapse), no tree solve and no grid-wide barrier. Its layout is the work arrays':
  // per group of simulations [T + 1][group nodes][S], the program's nodes at its base.
  // kTraj: the trajectory is in the gate type (the voltages the gates, synapses and
  // linearization computed with in the forward).
  constexpr bool traj = kTraj;
  const G* tvg = kTraj ? a.traj_v + (pg.soff - (static_cast<int64_t>(tp.nb) << lS)) * a.traj_steps : nullptr;
  const G* tv = kTraj ? tvg + (static_cast<int64_t>(tp.nb) << lS) : nullptr;
  // The voltages the reverse steps read: the trajectory or the recomputed float64 copy.
  using TV = std::conditional_t<kTraj, G, double>;
  for (int64_t c = C - 1; c >= 0; --c) {
    const int64_t ts = c * K;
    const int64_t te = ts + K < nsteps ? ts + K : nsteps;
    const int64_t coff = pg.soff + c * a.ckpt_stride;
    if (traj) {
      // Gates of every step of the segment from the checkpoint's, with the forward's
      // arithmetic (rates at the step's starting voltage, exponential Euler).
#pragma unroll 1
      for (int i = tid; i < nS; i += nthr) {
        const int k = i >> lS, s = i & (S - 1);
        G m = a.ckpt_m[coff + i], h = a.ckpt_h[coff + i], n = a.ckpt_n[coff + i];
        wm[i] = m;
        wh[i] = h;
        wn[i] = n;
        const int32_t flags = a.node_i[static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS + kNiFlags];
        if (!((flags & kComp) && (flags & kHH)) || s >= pg.nsim) continue;
        // The voltages of the next four steps in flight (the trajectory is read-only; the
        // row index stops at the segment's end, te <= the trajectory's last row).
        auto row = [&](int64_t t) { return (t < te ? t : te) * rowS + i; };
        G b0 = tv[row(ts)];
        G b1 = tv[row(ts + 1)];
        G b2 = tv[row(ts + 2)];
        G b3 = tv[row(ts + 3)];
#pragma unroll 1
        for (int64_t j = 0; j < te - ts; ++j) {
          const G vv = b0;
          b0 = b1;
          b1 = b2;
          b2 = b3;
          b3 = tv[row(ts + j + 4)];
          G am, bm, ah, bh, an, bn;
          hh_rates_step(vv, am, bm, ah, bh, an, bn);
          m = gate_step(m, dt, am, bm);
          h = gate_step(h, dt, ah, bh);
          n = gate_step(n, dt, an, bn);
          wm[(j + 1) * rowS + i] = m;
          wh[(j + 1) * rowS + i] = h;
          wn[(j + 1) * rowS + i] = n;
        }
      }
      __syncthreads();
    }
    // Recompute the segment from its checkpoint into work[0 .. te - ts]; the voltage
    // copy of step j is work_v[j] (program layout, as the forward's HBM copy).
    for (int i = tid; i < (traj ? 0 : nS); i += nthr) {
      wv[i] = a.ckpt_v[coff + i];
      wm[i] = a.ckpt_m[coff + i];
      wh[i] = a.ckpt_h[coff + i];
      wn[i] = a.ckpt_n[coff + i];
    }
    if constexpr (kSyn) {
      if (a.syn_model == 0) {
        const G* cs = a.ckpt_s + pg.synoff + c * a.ckpt_syn_stride;
        for (int64_t q = tid; q < nsynS; q += nthr) ws[q] = cs[q];
      }
    }
    __syncthreads();
    // Split network: every part's checkpoint loaded before any part reads presynaptic
    // voltages from it (the same after every recomputed step); with the trajectory, the
    // previous segment's presynaptic shares summed by every part before this segment's
    // reverse steps write the step-parity buffers again.
    if (kSyn && a.coop) group_barrier(a, nbar);
    if constexpr (kSyn) {
      if (traj) {
        // Synapse states and synaptic slopes of every step of the segment, per (target,
        // simulation) over the steps, the voltages from the trajectory; the synapses of a
        // pair split over its lanes as in the forward.
        const int lsub = syn_lanes_log2(ntS);
        for (int ii = tid; ii < (ntS << lsub); ii += nthr) {
          const int i2 = ii >> lsub, sub = ii & ((1 << lsub) - 1);
          const int tgt = i2 >> lS, s = i2 & (S - 1);
#pragma unroll 1
          for (int64_t j = 0; j < te - ts; ++j) {
            G lin = G(0), cur = G(0);
            if (s < pg.nsim)
              synapse_target(a, pg, tp, tgt, s, lin, cur, tvg + (ts + j) * rowS, tv + (ts + j) * rowS,
                             ws + j * nsynS, ws + (j + 1) * nsynS, sub, 1 << lsub);
            lin = pair_sum(lin, lsub);
            if (sub == 0) wl[j * ntS + i2] = lin;
          }
        }
        __syncthreads();
      }
    }
    for (int64_t j = 0; j < (traj ? 0 : te - ts); ++j) {
      if constexpr (kSyn) {
        // The synapse step of the recomputed forward (states into work_s[j + 1]).
        const int lsub = syn_lanes_log2(ntS);
        for (int ii = tid; ii < (ntS << lsub); ii += nthr) {
          const int i2 = ii >> lsub, sub = ii & ((1 << lsub) - 1);
          const int tgt = i2 >> lS, s = i2 & (S - 1);
          G lin = G(0), cur = G(0);
          if (s < pg.nsim)
            synapse_target(a, pg, tp, tgt, s, lin, cur, wvg + j * rowS, wv + j * rowS, ws + j * nsynS,
                           ws + (j + 1) * nsynS, sub, 1 << lsub);
          lin = pair_sum(lin, lsub);
          cur = pair_sum(cur, lsub);
          if (sub != 0) continue;
          Ld.tl[i2] = static_cast<L>(lin);
          Ld.tc[i2] = static_cast<L>(cur);
          wl[j * ntS + i2] = lin;
        }
        __syncthreads();
      }
      // The item's state loaded one item ahead (its loads in flight during the build).
      double vc = 0.0;
      G mc = G(0), hc = G(0), nc = G(0);
      if (tid < nS) {
        vc = wv[j * rowS + tid];
        mc = wm[j * rowS + tid];
        hc = wh[j * rowS + tid];
        nc = wn[j * rowS + tid];
      }
      for (int i = tid; i < nS; i += nthr) {
        const int k = i >> lS, s = i & (S - 1);
        const int inx = i + nthr < nS ? i + nthr : i;
        __builtin_amdgcn_sched_barrier(0);
        const double vx = wv[j * rowS + inx];
        const G mx = wm[j * rowS + inx], hx = wh[j * rowS + inx], nx = wn[j * rowS + inx];
        __builtin_amdgcn_sched_barrier(0);
        G m = mc, h = hc, n = nc;
        if (s < pg.nsim) {
          build_item<kCS>(a, pg, tp, Ld, wv + j * rowS, k, s, vc, pg.t0 + ts + j, m, h, n);
        } else {
          idle_item(a, tp, Ld, k, i);
        }
        vc = vx;
        mc = mx;
        hc = hx;
        nc = nx;
        wm[(j + 1) * rowS + i] = m;
        wh[(j + 1) * rowS + i] = h;
        wn[(j + 1) * rowS + i] = n;
      }
      __syncthreads();
      tree_solve<G, L, false, kCS>(a, tp, Ld.sc, Ld.xs, Ld.ds, Ld.cl);
      for (int i = tid; i < nS; i += nthr) wv[(j + 1) * rowS + i] = wv[j * rowS + i] + static_cast<double>(Ld.xs[i]);
      __syncthreads();
      if (kSyn && a.coop) group_barrier(a, nbar);
    }
    // The transposed system of a step: its diagonal at the gates after the step, its
    // right-hand side the voltage adjoint at the step's end (recordings included). The
    // segment's last step takes it from the recomputed trajectory; every other step's
    // system is written by the previous (later) step's item pass.
    auto system = [&](int i, int64_t t1, G mm, G hh, G nn, L lv) {
      const int k = i >> lS, s = i & (S - 1);
      const int32_t* nik = a.node_i + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS;
      const L* nfk = a.node_f + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_FCOLS;
      const int32_t flags = nik[kNiFlags];
      G lin = G(0);
      if ((flags & kComp) && s < pg.nsim) {
        if (flags & kHH) {
          const G n2 = nn * nn;
          lin += G(1000) * (prm<0>(a, pg, tp, k, s) * (mm * mm * mm) * hh +
                            prm<1>(a, pg, tp, k, s) * (n2 * n2) + prm<2>(a, pg, tp, k, s));
        }
        if (flags & kLeak) lin += G(1000) * prm<6>(a, pg, tp, k, s);
        if constexpr (kSyn) {
          // The synaptic slope of the step ending at t1 (a diagonal term of its system).
          const int tg = nik[kNiTgt];
          if (tg >= 0) lin += wl[(t1 - ts - 1) * ntS + (tg << lS) + s];
        }
        Ld.ds[i] = nfk[kNfDbase] + static_cast<L>(dt * static_cast<G>(nfk[kNfInvCm]) * lin);
      } else {
        Ld.ds[i] = nfk[kNfDbase] + ((flags & kComp) ? L(1) : L(0));
      }
      add_rec(i, t1, lv);
      Ld.xs[i] = lv;
    };
    {
      const int64_t jl = te - ts;
#pragma unroll 1
      for (int jj = 0; jj < kJ; ++jj) {
        const int i = tid + jj * nthr;
        if (i >= nS) break;
        system(i, te, wm[jl * rowS + i], wh[jl * rowS + i], wn[jl * rowS + i], lv_get(jj, i));
      }
      __syncthreads();
    }
    for (int64_t j = te - ts - 1; j >= 0; --j) {
      tree_solve<G, L, true, kCS>(a, tp, Ld.sc, Ld.xs, Ld.ds, Ld.cl);  // xs = mu = A^-T lambda_v(t+1)
      const TV* v0;
      const TV* v1;
      if constexpr (kTraj) {
        v0 = tv + (ts + j) * rowS;
        v1 = tv + (ts + j + 1) * rowS;
      } else {
        v0 = wv + j * rowS;
        v1 = wv + (j + 1) * rowS;
      }
      const G* m0 = wm + j * rowS;
      const G* h0 = wh + j * rowS;
      const G* n0 = wn + j * rowS;
      const G* m1 = wm + (j + 1) * rowS;
      const G* h1 = wh + (j + 1) * rowS;
      const G* n1 = wn + (j + 1) * rowS;
      // The trajectory of an item around the step, loaded one item ahead (the loads of
      // the next item are in flight while the current one computes). float64 gates: the
      // gates after the step are recomputed from those before it (the forward's gate
      // step; fewer loads, measured faster); float32: loaded (measured faster).
      constexpr bool kLoadG1 = sizeof(G) == 4;
      struct RevItem {
        TV v0, v1;
        G m0, h0, n0, m1, h1, n1;
        G lm, lh, ln;  // gate adjoints (IPT == 0: in HBM, loaded with the item)
      };
      auto load_rev = [&](int i) {
        RevItem r;
        r.v0 = v0[i];
        r.v1 = v1[i];
        r.m0 = m0[i];
        r.h0 = h0[i];
        r.n0 = n0[i];
        if constexpr (kLoadG1) {
          r.m1 = m1[i];
          r.h1 = h1[i];
          r.n1 = n1[i];
        } else {
          r.m1 = r.h1 = r.n1 = G(0);
        }
        if constexpr (IPT == 0) {
          r.lm = lm[i];
          r.lh = lh[i];
          r.ln = ln[i];
        } else {
          r.lm = r.lh = r.ln = G(0);
        }
        return r;
      };
      RevItem cur = load_rev(tid < nS ? tid : 0);
#pragma unroll 1
      for (int jj = 0; jj < kJ; ++jj) {
        const int i = tid + jj * nthr;
        if (i >= nS) break;
        const int inx = i + nthr < nS ? i + nthr : i;
        __builtin_amdgcn_sched_barrier(0);
        const RevItem nxt = load_rev(inx);
        __builtin_amdgcn_sched_barrier(0);
        const int k = i >> lS, s = i & (S - 1);
        const int32_t* nik = a.node_i + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS;
        const L* nfk = a.node_f + static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_FCOLS;
        const int32_t flags = nik[kNiFlags];
        G lvn = G(0);
        if ((flags & kComp) && s < pg.nsim) {
          const G inv_cm = static_cast<G>(nfk[kNfInvCm]);
          const G mu = static_cast<G>(Ld.xs[i]);
          const G gl = -mu * static_cast<G>(cur.v1) * dt * inv_cm;  // dL/dlin
          const G gc = mu * dt * inv_cm;                           // dL/dcst
          lvn = mu;
          if (flags & kHH) {
            const G gNa = prm<0>(a, pg, tp, k, s);
            const G gK = prm<1>(a, pg, tp, k, s);
            const G eNa = prm<3>(a, pg, tp, k, s);
            const G eK = prm<4>(a, pg, tp, k, s);
            const G eL = prm<5>(a, pg, tp, k, s);
            // The gates' step and its derivatives through the rates at the step's start
            // (evaluated first when they give the gates after the step).
            G em, eh, en, dmv, dhv, dnv, mm, hh, nn;
            auto gate_derivatives = [&]() {
              G am, bm, ah, bh, an, bn, dam, dbm, dah, dbh, dan, dbn;
              hh_rates_dv_step(static_cast<G>(cur.v0), am, bm, ah, bh, an, bn, dam, dbm, dah, dbh, dan, dbn);
              const G m1c = gate_step_d(cur.m0, dt, am, bm, dam, dbm, dmv, em);
              const G h1c = gate_step_d(cur.h0, dt, ah, bh, dah, dbh, dhv, eh);
              const G n1c = gate_step_d(cur.n0, dt, an, bn, dan, dbn, dnv, en);
              if constexpr (!kLoadG1) {
                mm = m1c;
                hh = h1c;
                nn = n1c;
              }
            };
            if constexpr (kLoadG1) {
              mm = cur.m1;
              hh = cur.h1;
              nn = cur.n1;
            } else {
              gate_derivatives();
            }
            const G m3 = mm * mm * mm, n3 = nn * nn * nn, n4 = n3 * nn;
            const G wNa = G(1000) * (gl + gc * eNa);
            const G wK = G(1000) * (gl + gc * eK);
            const G wL = G(1000) * (gl + gc * eL);
            const int64_t pi = pg.soff + i;
            grad(0, 0, pi, wNa * m3 * hh);
            grad(1, 1, pi, wK * n4);
            grad(2, 2, pi, wL);
            grad(3, -1, pi, G(1000) * gc * gNa * m3 * hh);
            grad(4, -1, pi, G(1000) * gc * gK * n4);
            grad(5, -1, pi, G(1000) * gc * prm<2>(a, pg, tp, k, s));
            G lam_m0, lam_h0, lam_n0;
            if constexpr (IPT > 0) {
              lam_m0 = lmr[jj];
              lam_h0 = lhr[jj];
              lam_n0 = lnr[jj];
            } else {
              lam_m0 = cur.lm;
              lam_h0 = cur.lh;
              lam_n0 = cur.ln;
            }
            const G lam_m1 = lam_m0 + wNa * gNa * G(3) * mm * mm * hh;
            const G lam_h1 = lam_h0 + wNa * gNa * m3;
            const G lam_n1 = lam_n0 + wK * gK * G(4) * n3;
            if constexpr (kLoadG1) gate_derivatives();
            lvn += lam_m1 * dmv + lam_h1 * dhv + lam_n1 * dnv;
            if constexpr (IPT > 0) {
              lmr[jj] = lam_m1 * em;
              lhr[jj] = lam_h1 * eh;
              lnr[jj] = lam_n1 * en;
            } else {
              lm[i] = lam_m1 * em;
              lh[i] = lam_h1 * eh;
              ln[i] = lam_n1 * en;
            }
          }
          if (flags & kLeak) {
            const G eLk = prm<7>(a, pg, tp, k, s);
            const int64_t pi = pg.soff + i;
            grad(6, 3, pi, G(1000) * (gl + gc * eLk));
            grad(7, -1, pi, G(1000) * gc * prm<6>(a, pg, tp, k, s));
          }
          if constexpr (kSyn) {
            const int tg = nik[kNiTgt];
            if (tg >= 0) {
              Ld.tl[(tg << lS) + s] = static_cast<L>(gl);
              Ld.tc[(tg << lS) + s] = static_cast<L>(gc);
            }
          }
        }
        if constexpr (kSyn) {
          lv_set(jj, i, static_cast<L>(lvn));  // completed by the synapse shares below
        } else if (j > 0) {
          // The item's own slots of the earlier step's system (the solve above read every
          // slot before its closing barrier); at the segment's first step the adjoint
          // waits for the next segment in VGPRs / HBM.
          system(i, ts + j, cur.m0, cur.h0, cur.n0, static_cast<L>(lvn));
        } else {
          lv_set(jj, i, static_cast<L>(lvn));
        }
        cur = nxt;
      }
      __syncthreads();
      if constexpr (kSyn) {
        // Synapse adjoint per (target, simulation) (kernels/synapse_step.py:184-330 of
        // the reference for its semantics): dL/dslope = gl + gc v_post and dL/dcurrent =
        // -gc at the target; each synapse's state adjoint, gS gradient and share of its
        // presynaptic voltage's adjoint (work_q); the sigmoid-rate synapses' share of the
        // postsynaptic voltage's adjoint (the slope multiplies v_post in the right-hand
        // side) into tl.
        const int32_t* jdp = a.syn_ptr + tp.spb;
        const TV* v0;
        const TV* vg0;
        if constexpr (kTraj) {
          v0 = tv + (ts + j) * rowS;
          vg0 = tvg + (ts + j) * rowS;
        } else {
          v0 = wv + j * rowS;
          vg0 = wvg + j * rowS;
        }
        const G* s0 = ws + j * nsynS;
        G* wqp = wq + (j & 1) * g.wq_stride;
        // The synapses of a pair split over its lanes (each synapse's results are its own;
        // the postsynaptic shares are summed over the lanes).
        const int lsub = syn_lanes_log2(ntS);
        for (int ii = tid; ii < (ntS << lsub); ii += nthr) {
          const int i2 = ii >> lsub, sub = ii & ((1 << lsub) - 1);
          const int tgt = i2 >> lS, s = i2 & (S - 1);
          G post = G(0);
          if (s < pg.nsim) {
            const G gl = static_cast<G>(Ld.tl[i2]), gc = static_cast<G>(Ld.tc[i2]);
            const int qn = a.tgt_node[tp.tgb + tgt];
            for (int q = sub;; q += 1 << lsub) {
              const int beg = jdp[q];
              if (tgt >= jdp[q + 1] - beg) break;
              const int syn = beg + tgt;
              const int64_t gsyn = tp.synb + syn;
              const int64_t si = static_cast<int64_t>(syn) * S + s;
              const G vpre = static_cast<G>(vg0[static_cast<int64_t>(a.syn_pre[gsyn]) * S + s]);
              const G scale = a.syn_scale[gsyn];
              const G gs = a.syn_param[0][gsyn];
              if (a.syn_model == 0) {
                G snew, E, dsdv;
                kinetic_step_d(s0[si], vpre, a.syn_param[2][gsyn], a.syn_param[3][gsyn],
                               a.syn_param[4][gsyn], a.syn_param[5][gsyn], dt, snew, E, dsdv);
                const G dg = gl + gc * a.syn_param[1][gsyn];  // dL/d(gS s' scale)
                const G lsp = lam_s[si] + dg * gs * scale;     // dL/ds'
                if (g.g_gs) g.g_gs[pg.synoff + si] += dg * snew * scale;
                lam_s[si] = lsp * E;
                store_coherent(wqp + static_cast<int64_t>(g.pre_pos[gsyn]) * S + s, lsp * dsdv);
              } else {
                const G xo = a.syn_param[1][gsyn];
                const G slope = a.syn_param[2][gsyn];
                const G za = (vpre - xo) * slope;
                const G zb = ((vpre + a.diff) - xo) * slope;
                const G sa = sigmoid(za);
                const G sb = sigmoid(zb);
                const G dsig = (G(1) - sa) * sb * (-dexpm1(-(zb - za)));
                const G vt = -gs * dsig * scale / a.diff;
                const G dvt = gl + gc * static_cast<G>(v0[qn * S + s]);  // dL/dslope
                const G dd0 = -gc;                                       // dL/dcurrent
                post += gc * vt;
                const G dvt_dv = -gs * scale / a.diff * slope * (sb * (G(1) - sb) - sa * (G(1) - sa));
                const G dd0_dv = -gs * scale * slope * sa * (G(1) - sa);
                store_coherent(wqp + static_cast<int64_t>(g.pre_pos[gsyn]) * S + s, dvt * dvt_dv + dd0 * dd0_dv);
                if (g.g_gs) g.g_gs[pg.synoff + si] += dvt * (-dsig * scale / a.diff) + dd0 * (-sa * scale);
              }
            }
          }
          post = pair_sum(post, lsub);
          // Every lane of the pair read tl / tc above; they are replaced after the sum.
          if (sub == 0) Ld.tl[i2] = static_cast<L>(post);
        }
        __syncthreads();
        // Split network: the presynaptic shares of every part written before any part
        // sums its nodes' slots (the shares are the only data crossing this barrier; the
        // recomputed voltages crossed the barriers of the recomputation).
        if (a.coop) group_barrier<true>(a, nbar);
        // The voltage adjoints completed by the synapse shares: the target's own, and the
        // presynaptic shares of every synapse leaving the node (by presynaptic node).
        const int32_t* pptr = g.pre_ptr + tp.ppb;
#pragma unroll 1
        for (int jj = 0; jj < kJ; ++jj) {
          const int i = tid + jj * nthr;
          if (i >= nS) break;
          const int k = i >> lS, s = i & (S - 1);
          L lv = lv_get(jj, i);
          if (s < pg.nsim) {
            G add = G(0);
            const int tg = a.node_i[static_cast<int64_t>(tp.nb + k) * JXHIP_NODE_ICOLS + kNiTgt];
            if (tg >= 0) add += static_cast<G>(Ld.tl[(tg << lS) + s]);
            for (int e = pptr[k]; e < pptr[k + 1]; ++e) add += load_coherent(wqp + static_cast<int64_t>(e) * S + s);
            lv = static_cast<L>(static_cast<G>(lv) + add);
          }
          if (j > 0) {
            system(i, ts + j, m0[i], h0[i], n0[i], lv);
          } else {
            lv_set(jj, i, lv);
          }
        }
        __syncthreads();
      }
    }
  }
#pragma unroll 1
  for (int jj = 0; jj < kJ; ++jj) {
    const int i = tid + jj * nthr;
    if (i >= nS) break;
    L lv = lv_get(jj, i);
    add_rec(i, 0, lv);
    g.g_v0[pg.soff + i] = static_cast<G>(lv);
  }
  if (simgrad) {
    // Per-simulation sums: thread partials into LDS (the system's space is free now),
    // then thread s < S adds the partials of simulation s in thread order and stores
    // them at its node-0 item (the caller sums each accumulator over the nodes).
    G* red = reinterpret_cast<G*>(smem_raw);
    constexpr int kQ[4] = {0, 1, 2, 6};
    for (int q = 0; q < 4; ++q) {
      __syncthreads();
      red[tid] = acc[q];
      __syncthreads();
      if (tid < S && g.g_param[kQ[q]] != nullptr) {
        G sum = G(0);
        for (int t = tid; t < nthr; t += S) sum += red[t];
        g.g_param[kQ[q]][pg.soff + tid] += sum;
      }
    }
  }
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
  if (a.syn_model < -1 || a.syn_model > JXHIP_SYN_MIXED) return hipErrorInvalidValue;
  if (a.syn_model == JXHIP_SYN_MIXED && a.syn_kind == nullptr) return hipErrorInvalidValue;
  if (a.dt_v < 0 || a.rhs_scale < 0 || a.explicit_step < 0 || a.explicit_step > 1) return hipErrorInvalidValue;
  if (a.node_p2 != nullptr && (a.gate_ext == nullptr || a.gate_ext_stride <= 0)) return hipErrorInvalidValue;
  if (a.parts < 0 || a.coop < 0 || a.coop > 1 || a.group_nodes < 0) return hipErrorInvalidValue;
  if (a.coef_global < 0 || a.coef_global > 2) return hipErrorInvalidValue;
  // Lockstep programs: whole groups of parts.
  if (a.coop && (a.parts < 2 || a.n_progs % a.parts)) return hipErrorInvalidValue;
  return hipSuccess;
}

// One launch of the programs; with a.coop a cooperative launch (every workgroup resident,
// so the grid-wide barriers complete), in consecutive launches of as many whole groups as
// fit the device at once when the programs do not.
template <typename K, typename F, typename... Rest>
int launch_kernel(K kern, const jxhip_fused_args& a, hipStream_t stream, F f, Rest... rest) {
  const dim3 block(a.threads);
  const size_t lds = static_cast<size_t>(a.lds_bytes);
  if (hipError_t e = set_lds(kern, lds)) return e;
  if (!a.coop) {
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(a.n_progs)), block, lds, stream, f, rest...);
    return hipGetLastError();
  }
  int per_cu = 0, dev = 0, cus = 0;
  if (hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, a.threads, lds)) return e;
  if (hipError_t e = hipGetDevice(&dev)) return e;
  if (hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) return e;
  int64_t cap = static_cast<int64_t>(per_cu) * cus;
  cap -= cap % a.parts;
  if (cap <= 0) return hipErrorCooperativeLaunchTooLarge;
  for (int64_t p0 = 0; p0 < a.n_progs; p0 += cap) {
    const int64_t cnt = a.n_progs - p0 < cap ? a.n_progs - p0 : cap;
    F fc = f;
    fc.progs = f.progs + p0 * JXHIP_PROG_COLS;
    if (f.gsync != nullptr) fc.gsync = f.gsync + p0 / a.parts;
    void* kargs[] = {&fc, &rest...};
    if (hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern),
                                                  dim3(static_cast<unsigned>(cnt)), block, kargs,
                                                  static_cast<unsigned>(lds), stream))
      return e;
  }
  return hipSuccess;
}

// Register-resident item counts with a kernel, by state element size: float gates
// (<= 12 items at 1024 threads, <= 16 at 512), float64 gates (<= 4 at 1024, <= 12 at 512).
template <typename G, typename L, bool kSyn, bool kCkpt>
int launch_forward_feat(const jxhip_fused_args& a, const Fused<G, L>& f, hipStream_t st) {
  if (a.items_per_thread == 0) {
    if (a.work_ax == nullptr) return hipErrorInvalidValue;
    if (kSyn && a.coop && (a.vx == nullptr || a.vx_stride <= 0)) return hipErrorInvalidValue;
    if (a.coef_global == 2) {
      // The symmetric system: its coefficients and row scales; not with forward Euler
      // (its branchpoint pass reads the unscaled rows).
      if (a.node_sym == nullptr || a.explicit_step) return hipErrorInvalidValue;
      return launch_kernel(fused_stream_kernel<G, L, kSyn, kCkpt, 2>, a, st, f);
    }
    if (a.coef_global) return launch_kernel(fused_stream_kernel<G, L, kSyn, kCkpt, 1>, a, st, f);
    return launch_kernel(fused_stream_kernel<G, L, kSyn, kCkpt, 0>, a, st, f);
  }
  if (a.coef_global || a.coop) return hipErrorInvalidValue;  // lockstep programs: streaming only
  if (a.max_items > static_cast<int64_t>(a.items_per_thread) * a.threads) return hipErrorInvalidValue;
  constexpr bool kF = sizeof(G) == 4;
#define JXHIP_REGS(IPT, MAXT)                                                            \
  case IPT:                                                                              \
    if (a.threads > MAXT) return hipErrorInvalidValue;                                   \
    return launch_kernel(fused_regs_kernel<G, L, IPT, MAXT, kSyn, kCkpt>, a, st, f);
  if constexpr (kF) {
    switch (a.items_per_thread) {
      JXHIP_REGS(1, 1024)
      JXHIP_REGS(2, 1024)
      JXHIP_REGS(3, 1024)
      JXHIP_REGS(4, 1024)
      JXHIP_REGS(6, 1024)
      JXHIP_REGS(8, 1024)
      JXHIP_REGS(10, 1024)
      JXHIP_REGS(12, 1024)
      JXHIP_REGS(16, 512)
      default:
        return hipErrorInvalidValue;
    }
  } else {
    switch (a.items_per_thread) {
      JXHIP_REGS(1, 1024)
      JXHIP_REGS(2, 1024)
      JXHIP_REGS(3, 1024)
      JXHIP_REGS(4, 1024)
      JXHIP_REGS(6, 512)
      JXHIP_REGS(8, 512)
      JXHIP_REGS(12, 512)
      default:
        return hipErrorInvalidValue;
    }
  }
#undef JXHIP_REGS
}

template <typename G, typename L>
int launch_forward(const jxhip_fused_args* args, hipStream_t stream) {
  if (args == nullptr) return hipErrorInvalidValue;
  const jxhip_fused_args& a = *args;
  if (a.n_progs <= 0) return hipSuccess;
  if (int e = check_args(a)) return e;
  const Fused<G, L> f = unpack<G, L>(a);
  const bool syn = a.syn_model >= 0, ckpt = a.ckpt_v != nullptr;
  if (ckpt && a.ckpt_every <= 0) return hipErrorInvalidValue;
  // A voltage trajectory is kept by the checkpointed kernels only.
  if (a.traj_v != nullptr && (!ckpt || a.traj_steps < 1)) return hipErrorInvalidValue;
  if (syn && ckpt) return launch_forward_feat<G, L, true, true>(a, f, stream);
  if (syn) return launch_forward_feat<G, L, true, false>(a, f, stream);
  if (ckpt) return launch_forward_feat<G, L, false, true>(a, f, stream);
  return launch_forward_feat<G, L, false, false>(a, f, stream);
}

template <typename G, typename L>
int launch_adjoint(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                   hipStream_t stream) {
  if (args == nullptr || adj == nullptr) return hipErrorInvalidValue;
  const jxhip_fused_args& a = *args;
  if (a.n_progs <= 0) return hipSuccess;
  if (int e = check_args(a)) return e;
  if (a.syn_model > 1 || a.ckpt_v == nullptr || a.ckpt_every <= 0) return hipErrorInvalidValue;
  const bool syn = a.syn_model >= 0;
  if (adj->rec_first == nullptr || adj->rec_next == nullptr) return hipErrorInvalidValue;
  if (syn && (adj->pre_ptr == nullptr || adj->pre_pos == nullptr || adj->work_q == nullptr ||
              adj->work_lin == nullptr || adj->work_lin_stride <= 0))
    return hipErrorInvalidValue;
  if (syn && a.syn_model == 0 && (adj->work_s == nullptr || adj->lam_s == nullptr || a.ckpt_s == nullptr))
    return hipErrorInvalidValue;
  if (a.node_p2 != nullptr) return hipErrorInvalidValue;  // HH / Leak cells only
  if (a.threads > adjoint_max_threads<G>(a.items_per_thread)) return hipErrorInvalidValue;
  // The reverse pass differentiates the backward-Euler step only.
  if ((a.dt_v > 0 && a.dt_v != a.dt) || (a.rhs_scale > 0 && a.rhs_scale != 1.0) || a.explicit_step)
    return hipErrorInvalidValue;
  if (a.items_per_thread < 0) return hipErrorInvalidValue;
  if (a.items_per_thread > 0 && a.max_items > static_cast<int64_t>(a.items_per_thread) * a.threads)
    return hipErrorInvalidValue;
  if (a.items_per_thread == 0 && adj->lam_v == nullptr) return hipErrorInvalidValue;
  if (a.traj_v != nullptr && a.traj_steps < 1) return hipErrorInvalidValue;
  if (a.coef_global > 1) return hipErrorInvalidValue;  // the symmetric system: forward only
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
  g.lv = static_cast<G*>(adj->lam_v);
  g.work_stride = adj->work_stride;
  g.rec_first = adj->rec_first;
  g.rec_next = adj->rec_next;
  g.pre_ptr = adj->pre_ptr;
  g.pre_pos = adj->pre_pos;
  g.ws = static_cast<G*>(adj->work_s);
  g.work_syn_stride = adj->work_syn_stride;
  g.lam_s = static_cast<G*>(adj->lam_s);
  g.wq = static_cast<G*>(adj->work_q);
  g.g_gs = static_cast<G*>(adj->g_syn_gs);
  g.wl = static_cast<G*>(adj->work_lin);
  g.work_lin_stride = adj->work_lin_stride;
  g.wq_stride = adj->work_q_stride;
#define JXHIP_ADJ_T(IPT, TR)                                                                        \
  if (syn)                                                                                          \
    return a.coef_global ? launch_kernel(fused_adjoint_kernel<G, L, IPT, true, true, TR>, a, stream, f, g) \
                         : launch_kernel(fused_adjoint_kernel<G, L, IPT, false, true, TR>, a, stream, f, g); \
  return a.coef_global ? launch_kernel(fused_adjoint_kernel<G, L, IPT, true, false, TR>, a, stream, f, g) \
                       : launch_kernel(fused_adjoint_kernel<G, L, IPT, false, false, TR>, a, stream, f, g);
#define JXHIP_ADJ(IPT)                                                                              \
  case IPT:                                                                                         \
    if (a.traj_v != nullptr) {                                                                      \
      JXHIP_ADJ_T(IPT, true)                                                                        \
    }                                                                                               \
    JXHIP_ADJ_T(IPT, false)
  switch (a.items_per_thread) {
    JXHIP_ADJ(0)
    JXHIP_ADJ(1)
    JXHIP_ADJ(2)
    JXHIP_ADJ(4)
    JXHIP_ADJ(8)
    JXHIP_ADJ(12)
    JXHIP_ADJ(16)
    default:
      return hipErrorInvalidValue;
  }
#undef JXHIP_ADJ
#undef JXHIP_ADJ_T
}

}  // namespace jxhip

/* jxhip.h — C ABI of libjxhip.so, the MI355X (gfx950) kernels of the per-time-step
 * hot path of jaxley-pytorch's `jx.integrate` (tree solve + HH gates + synapse step).
 *
 * Conventions
 *  - Every entry point returns a hipError_t value as int (0 = success) and launches on
 *    the given stream (`void* stream` is a hipStream_t; NULL = default stream). No
 *    entry point synchronizes, allocates, or copies host<->device, so launches can be
 *    captured in a HIP graph.
 *  - All pointers are device pointers unless stated otherwise. Ownership stays with the
 *    caller (PyTorch tensors in the Python binding).
 *  - `f32` / `f64` suffix: element type of the floating point arrays.
 *  - Applicability errors are reported as hipErrorInvalidValue (the Python layer raises).
 *
 * Reference interfaces replaced (paths relative to the jaxley-pytorch checkout):
 *  - jxhip_tree_solve_levels_*: custom op `jaxley_pytorch::tree_solve_levels`
 *    (jaxley_pytorch/kernels/tree_solve.py:589-630, kernel :290-375), launched by
 *    `LevelTreeSolveKernel.solve/.fused` (:433-482).
 *  - jxhip_tree_solve_regs_*: custom op `jaxley_pytorch::tree_solve`
 *    (jaxley_pytorch/kernels/tree_solve.py:518-587, kernel :228-287), launched by
 *    `TreeSolveKernel.solve/.fused` (:378-430).
 *  - jxhip_tree_solve_grads_*: the backward of both ops after the transposed solve
 *    (tree_solve.py:575-587, 619-630; solver_voltage.py:170-187).
 *  - jxhip_hh_step_*: the eager HH update + linearization of `_step_channels`
 *    (modules/base.py:3402-3542, channels/hh.py:37-67), which has no kernel boundary in
 *    the reference.
 *  - jxhip_synapse_step_* / jxhip_synapse_step_backward_*: custom ops
 *    `jaxley_pytorch::synapse_step` / `::synapse_step_backward`
 *    (jaxley_pytorch/kernels/synapse_step.py:336-421, kernels :121-330).
 *  - jxhip_fused_forward_* / jxhip_fused_adjoint_*: the whole time loop of
 *    `integrate` (integrate.py:624-662, `nested_checkpoint_scan` utils/torch_utils.py:141-198)
 *    around `_step_with_plan` (modules/base.py:3214-3400), forward and hand-written
 *    reverse pass with K-step checkpoints.
 */
#ifndef JXHIP_H
#define JXHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Library version (major * 10000 + minor * 100 + patch). Host-only, no GPU call. */
int jxhip_version(void);

/* ---------------------------------------------------------------------------------
 * Standalone tree solve (drop-in for the reference's two custom ops:
 * jxhip_tree_solve_levels_* replaces `jaxley_pytorch::tree_solve_levels`
 * (jaxley_pytorch/kernels/tree_solve.py:589-630, kernel :290-375), jxhip_tree_solve_regs_*
 * replaces `jaxley_pytorch::tree_solve` (tree_solve.py:518-587, kernel :228-287); both
 * are reached from `dhs_solve` (solver_voltage.py:237-240)).
 *
 * Solves A x = s for every row of the (n_rows, n) row-major, contiguous inputs with
 *   A[i,i] = d[i],  A[p,c] = u[c],  A[c,p] = l[c]   (p = parent of c).
 * The forest is in the reference's solve order (every tree contiguous, root first,
 * parents before children). Schedule arrays (int32, built on the host):
 *   block_info (n_blocks, 4): start, length, max height H, level_ptr offset
 *   level_ptr  per block H+2 entries: positions in `order` of each height 0..H
 *   order      (n): block-relative node of each position (sorted by height, stable)
 *   parent     (n): block-relative parent, -1 for roots
 *   child_rng  (n, 2): [begin, end) into child_idx (block-relative children)
 * ------------------------------------------------------------------------------- */
int jxhip_tree_solve_levels_f32(const float* d, const float* l, const float* u, const float* s,
                                float* x, float* work, int64_t n_rows, int64_t n,
                                const int32_t* block_info, int32_t n_blocks, int32_t max_block,
                                const int32_t* level_ptr, const int32_t* order,
                                const int32_t* parent, const int32_t* child_rng,
                                const int32_t* child_idx, void* stream);
int jxhip_tree_solve_levels_f64(const double* d, const double* l, const double* u, const double* s,
                                double* x, double* work, int64_t n_rows, int64_t n,
                                const int32_t* block_info, int32_t n_blocks, int32_t max_block,
                                const int32_t* level_ptr, const int32_t* order,
                                const int32_t* parent, const int32_t* child_rng,
                                const int32_t* child_idx, void* stream);

/* Wave-resident variant: one wavefront per (block, row), block <= 64 * npl nodes held
 * in VGPRs (npl in {1,2,4,8,16,32}); same schedule arrays. `work` is unused (NULL). */
int jxhip_tree_solve_regs_f32(const float* d, const float* l, const float* u, const float* s,
                              float* x, int64_t n_rows, int64_t n,
                              const int32_t* block_info, int32_t n_blocks, int32_t npl,
                              const int32_t* level_ptr, const int32_t* order,
                              const int32_t* parent, const int32_t* child_rng,
                              const int32_t* child_idx, void* stream);
int jxhip_tree_solve_regs_f64(const double* d, const double* l, const double* u, const double* s,
                              double* x, int64_t n_rows, int64_t n,
                              const int32_t* block_info, int32_t n_blocks, int32_t npl,
                              const int32_t* level_ptr, const int32_t* order,
                              const int32_t* parent, const int32_t* child_rng,
                              const int32_t* child_idx, void* stream);

/* Gradients after the transposed solve lam = A^-T g (solver_voltage.py:170-187):
 *   g_d = -lam x,  g_u[c] = -lam[p] x[c],  g_l[c] = -lam[c] x[p]  (0 at roots).
 * parent_glob (n): forest-global parent index, -1 for roots. Any of g_d/g_l/g_u may be
 * NULL. */
int jxhip_tree_solve_grads_f32(const float* lam, const float* x, const int32_t* parent_glob,
                               float* g_d, float* g_l, float* g_u, int64_t n_rows, int64_t n,
                               void* stream);
int jxhip_tree_solve_grads_f64(const double* lam, const double* x, const int32_t* parent_glob,
                               double* g_d, double* g_l, double* g_u, int64_t n_rows, int64_t n,
                               void* stream);

/* ---------------------------------------------------------------------------------
 * HH channel step on compartments `inds` of (n_rows, n_nodes) voltages
 * (modules/base.py:3414-3542 for one HH channel):
 *   gates m,h,n (n_rows, n_ch) updated at the old v (exponential Euler), then the
 *   current I (mA/cm^2) at v with the new gates, its slope `lin` = dI/dv and
 *   `cst` = -(I - lin v); outputs are (n_rows, n_ch).
 * params: pointer to 6 device pointers (gNa, gK, gLeak, eNa, eK, eLeak), each with
 * row stride p_row_stride[j] (0 = shared over rows) and unit column stride.
 * ------------------------------------------------------------------------------- */
int jxhip_hh_step_f32(const float* v, int64_t n_nodes, const int32_t* inds, int64_t n_ch,
                      const float* m, const float* h, const float* n,
                      const float* const* params, const int64_t* p_row_stride,
                      float dt, float* m_out, float* h_out, float* n_out, float* i_out,
                      float* lin_out, float* cst_out, int64_t n_rows, void* stream);
int jxhip_hh_step_f64(const double* v, int64_t n_nodes, const int32_t* inds, int64_t n_ch,
                      const double* m, const double* h, const double* n,
                      const double* const* params, const int64_t* p_row_stride,
                      double dt, double* m_out, double* h_out, double* n_out, double* i_out,
                      double* lin_out, double* cst_out, int64_t n_rows, void* stream);

/* ---------------------------------------------------------------------------------
 * Synapse step of one synapse type in a regular layout (k synapses per target, sorted
 * by target; syn_utils.py:13-27). model 0 = KineticSynapse (params gS, e_syn, alpha,
 * beta, v_th, delta), 1 = SigmoidRateSynapse (gS, x_offset, slope).
 *   v (B, n_nodes); s, s_new, i_syn (B, E); params: 6 device pointers of (E,);
 *   area (E,) postsynaptic area; lin, cst (B, T): the finite-difference slope and
 *   offset (network.py:382-478 / synapse_step.py:121-182) summed over the k synapses.
 * targets (T) node of each target (ignored when targets_are_nodes).
 * ------------------------------------------------------------------------------- */
int jxhip_synapse_step_f32(const float* v, const float* s, const int32_t* pre,
                           const int32_t* targets, const float* const* params,
                           const float* area, float dt, float diff, float* s_new, float* i_syn,
                           float* lin, float* cst, int64_t B, int64_t n_nodes, int64_t E,
                           int64_t k, int32_t model, int32_t targets_are_nodes, void* stream);
int jxhip_synapse_step_f64(const double* v, const double* s, const int32_t* pre,
                           const int32_t* targets, const double* const* params,
                           const double* area, double dt, double diff, double* s_new,
                           double* i_syn, double* lin, double* cst, int64_t B, int64_t n_nodes,
                           int64_t E, int64_t k, int32_t model, int32_t targets_are_nodes,
                           void* stream);

/* Backward of jxhip_synapse_step_* (synapse_step.py:184-330): given upstream gradients
 * g_s_new (B,E) or NULL, g_i (B,E) or NULL, g_lin, g_cst (B,T), writes
 *   g_v (B, n_nodes) (overwritten: presynaptic + postsynaptic contributions),
 *   g_s (B, E) (kinetic only), g_params (n_params, E) summed over the batch
 *   (rows in parameter order; written for every parameter).
 * pre_nodes (n_pre) / inv (n_pre, width): transposed pre->synapse index (-1 padded)
 * used for a deterministic presynaptic sum; work: 2 * E * B elements of scratch. */
int jxhip_synapse_step_backward_f32(const float* v, const float* s, const int32_t* pre,
                                    const int32_t* targets, const float* const* params,
                                    const float* area, float dt, float diff,
                                    const float* g_s_new, const float* g_i, const float* g_lin,
                                    const float* g_cst, float* g_v, float* g_s, float* g_params,
                                    float* work, const int32_t* pre_nodes, const int32_t* inv,
                                    int64_t n_pre, int64_t inv_width, int64_t B, int64_t n_nodes,
                                    int64_t E, int64_t k, int32_t model, int32_t targets_are_nodes,
                                    void* stream);
int jxhip_synapse_step_backward_f64(const double* v, const double* s, const int32_t* pre,
                                    const int32_t* targets, const double* const* params,
                                    const double* area, double dt, double diff,
                                    const double* g_s_new, const double* g_i,
                                    const double* g_lin, const double* g_cst, double* g_v,
                                    double* g_s, double* g_params, double* work,
                                    const int32_t* pre_nodes, const int32_t* inv, int64_t n_pre,
                                    int64_t inv_width, int64_t B, int64_t n_nodes, int64_t E,
                                    int64_t k, int32_t model, int32_t targets_are_nodes,
                                    void* stream);

/* ---------------------------------------------------------------------------------
 * Fused whole-simulation kernels. One workgroup ("program") owns S simulations of one
 * topology (a cell, a network of cells, or a pack of trees) for a range of time
 * steps, with voltages and the elimination in LDS; nothing crosses workgroups, so no
 * grid-wide synchronization is needed.
 *
 * Programs: int64 (n_progs, JXHIP_PROG_COLS):
 *   topo, nsim (active simulations <= S), t_begin, t_end (global step range: stimulus
 *   time index), state_off (element offset of the program's [node][S] blocks),
 *   sim_off (global index of its first simulation), rec_off, syn_off ([syn][S] blocks)
 * Topologies: int32 (n_topo, JXHIP_TOPO_COLS):
 *   n (nodes), L (max segment level), node_base, seg_level_base, child_base, n_rec,
 *   rec_base, n_tgt, tgt_base, syn_ptr_base, syn_base, n_syn, seg_base, seg_child_base,
 *   pre_ptr_base (the reverse pass's synapses by presynaptic node), 0
 * Nodes are in "kernel order": segments (maximal single-child chains, top first) grouped
 * by segment level, leaf segments first. A segment's level is 0 if its bottom node has
 * no children, else 1 + the largest level of its child segments.
 * Per-node int32 (JXHIP_NODE_ICOLS) at node_base + k: parent (-1 root), child_beg,
 *   child_end (into child_idx at child_base), flags (1 comp, 2 HH, 4 Leak, 8 export:
 *   voltage copied to HBM every step; set for recorded nodes and synapse pre- and
 *   postsynaptic nodes; 16, 32, 64: Pospischil Na, K, Km, channels/pospischil.py:74-250),
 *   stim_beg,
 *   stim_end (stimulus entries), target index (-1 none), 0
 * Per-node solve type (JXHIP_NODE_FCOLS): lo = -dt g(parent->k), up = -dt g(k->parent),
 *   dbase = [comp] + sum of dt g over edges into k, inv_cm (0 at branchpoints)
 * Segments int32 (n_seg, 4) at seg_base: first node, end node, child_beg, child_end (into
 *   seg_child at seg_child_base: the top nodes of the child segments); seg_level_ptr:
 *   L + 2 offsets (segments of level l are [ptr[l], ptr[l+1])).
 * Voltages are float64 in both modes (the solve runs on increments, see jxhip_fused.hip);
 * every other floating array has the entry point's element type.
 * ------------------------------------------------------------------------------- */
#define JXHIP_PROG_COLS 8
#define JXHIP_TOPO_COLS 16
#define JXHIP_NODE_ICOLS 8
#define JXHIP_NODE_FCOLS 4
#define JXHIP_SYN_MIXED 4   /* syn_model: the model of every synapse in syn_kind */
#define JXHIP_NPARAM 8      /* HH gNa gK gLeak eNa eK eLeak, Leak gLeak eLeak */
#define JXHIP_NPARAM2 7     /* Pospischil: Na gNa, eNa, vt, K gK, eK, Km gKm, taumax */
#define JXHIP_NSYNPARAM 6

typedef struct jxhip_fused_args {
  const int64_t* progs;
  int64_t n_progs;
  const int32_t* topos;
  const int32_t* seg;       /* (n_seg_total, 4) */
  const int32_t* seg_level_ptr;
  const int32_t* seg_child;
  const int32_t* node_i;    /* (n_nodes_total, JXHIP_NODE_ICOLS) */
  const void* node_f;       /* (n_nodes_total, JXHIP_NODE_FCOLS), element type */
  const int32_t* child_idx; /* topology-relative children at child_base */
  const int32_t* stim_col;  /* stimulus column of each stimulus entry */
  const void* stim_scale;   /* dt * 1e5 / (area cm) per stimulus entry */
  const int32_t* rec_node;  /* recorded nodes at rec_base */
  /* channel parameters: node_p (n_nodes_total, JXHIP_NPARAM) per node; the
     conductances gNa, gK, gLeak (HH) and gLeak (Leak) may instead be batched in
     bparam, 4 values per entry: bparam_mode 1 = per simulation ([sim_off + s][4]),
     2 = per (node, simulation) ([state_off + k S + s][4]); 0 = not batched */
  const void* node_p;
  const void* bparam;
  int32_t bparam_mode;
  /* state in program layout [node][S] at state_off; v is float64 */
  void* v;
  void* gate_m;
  void* gate_h;
  void* gate_n;
  /* stimulus: (T_total, stim_cols, stim_b), stim_b = total simulations (indexed by
     sim_off + s) or 1 (shared by all simulations) */
  const void* stim;
  int64_t stim_cols;
  int64_t stim_b;
  /* recordings: per program [t][rec][S] at rec_off, t = 0 .. t_end - t_begin */
  void* rec;
  /* synapses in jagged-diagonal order: targets numbered by non-increasing in-degree; the
     j-th synapse of target t is syn_ptr[j] + t while t < syn_ptr[j+1] - syn_ptr[j] (the
     pointer list ends with 8 entries equal to the topology's synapse count: the forward
     walks a target's synapses with up to 8 threads, each reading up to 7 entries past
     the target's last one) */
  int32_t syn_model;        /* -1 none, 0 kinetic, 1 sigmoid rate, 2 ionotropic, 3 test
                               synapse, JXHIP_SYN_MIXED: per synapse in syn_kind */
  const int32_t* syn_pre;   /* presynaptic node (of the group's nodes: the topology's own
                               unless the network is split), at syn_base */
  const int32_t* syn_ptr;   /* max in-degree + 9 per topology at syn_ptr_base */
  const int32_t* tgt_node;  /* topology-relative target node, at tgt_base */
  const void* syn_param[JXHIP_NSYNPARAM]; /* per synapse at syn_base */
  const void* syn_scale;    /* 1e5 / area(post) per synapse at syn_base */
  void* syn_s;              /* synapse states, per program [syn][S] at syn_off */
  int64_t sims_total;
  double dt;
  double diff;              /* finite difference of the synapse slope (1e-3) */
  int32_t S;                /* simulations per program: 1, 2, 4, ..., 64 */
  int32_t threads;          /* threads per workgroup: multiple of 64, <= 1024 */
  int32_t items_per_thread; /* forward: 0 = streaming (state read from and written to
                               HBM every step); > 0 = every thread keeps that many
                               (node, simulation) items in VGPRs for the whole launch:
                               float32 gates 1, 2, 3, 4, 6, 8, 10, 12 (threads <=
                               1024) or 16 (threads <= 512); float64 gates 1, 2, 3, 4
                               (<= 1024) or 6, 8, 12 (<= 512) */
  int64_t lds_bytes;        /* dynamic LDS per workgroup (<= 160 KiB) */
  /* checkpoints (for the adjoint): per program, snapshot c of the state before step
     c * ckpt_every at ckpt_* + c * ckpt_stride + state_off (syn: ckpt_syn_stride,
     syn_off); NULL = none */
  void* ckpt_v;             /* float64 */
  void* ckpt_m;
  void* ckpt_h;
  void* ckpt_n;
  void* ckpt_s;
  int64_t ckpt_every;
  int64_t ckpt_stride;
  int64_t ckpt_syn_stride;
  int64_t max_items;        /* largest n * S over the programs of the launch; with
                               items_per_thread > 0 it must be <= items_per_thread *
                               threads (else the call fails with an invalid argument) */
  int32_t coef_global;      /* forward, streaming only: 1 = the tree's coupling
                               coefficients are read from node_f (HBM / L2) and lds_bytes
                               excludes their LDS copy; 0 = copied to LDS; 2 = the solve
                               runs on the row-scaled symmetric system of node_sym, whose
                               one coefficient per node is copied to LDS (lds_bytes
                               includes n * sizeof(solve type) rounded up to 16 in place
                               of the pairs), and the residual updates read node_f from
                               HBM / L2 (not with explicit_step; the reverse pass
                               rejects it) */
  void* work_ax;            /* forward, streaming only (required there): scratch of the
                               axial part of every item's residual, program layout
                               [node][S] at state_off, gate type; computed from v at the
                               start of the launch and updated from the increments after
                               every solve. Unused by the register-resident kernel. */
  /* The voltage step (base.py:3327-3366 of jaxley-pytorch): backward Euler solves with
     dt; Crank-Nicolson solves with dt / 2 and takes v' = 2 x - v, which in the increment
     form is a right-hand side scaled by 2. Gates and synapse states always advance by
     dt. node_f's coupling coefficients and stim_scale are built with dt_v. 0 = the
     backward-Euler values (dt_v = dt, rhs_scale = 1); the reverse pass accepts only
     those. */
  double dt_v;
  double rhs_scale;
  /* 1 = forward Euler (step_voltage_explicit, solver_voltage.py:412-502): no tree solve;
     a compartment advances by dt times the vector field and every branchpoint sits at
     the conductance-weighted mean of its neighbours (set so at the start of the launch
     and after every step). 0 = the implicit step above. Forward only. */
  int32_t explicit_step;
  /* With syn_model == JXHIP_SYN_MIXED (several synapse types in one network,
     network.py:350-478 steps every type): the model (0-3) of every synapse at syn_base.
     syn_param[j] then holds each synapse's j-th parameter in its own model's order, and
     every synapse has a state slot in syn_s (unused by the sigmoid-rate model). */
  const int32_t* syn_kind;
  /* Pospischil channels (node flags 16 Na, 32 K, 64 Km; Pospischil et al. 2008 as in
     channels/pospischil.py:74-250 of jaxley-pytorch): node_p2 (n_nodes_total,
     JXHIP_NPARAM2) their parameters per node, unbatched; gate_ext their gates in the
     program layout of gate_m, Na m at gate_ext, Na h at + gate_ext_stride, K n at
     + 2 stride, Km p at + 3 stride (elements). null node_p2: no such channels. Forward
     only (the reverse pass rejects them). */
  const void* node_p2;
  void* gate_ext;
  int64_t gate_ext_stride;
  /* Networks split into parts (groups of whole cells, one topology each, node_base the
     part's first node): a group of S simulations is run by `parts` consecutive programs,
     one per part, whose [node][S] state blocks lie back to back (group_nodes nodes in
     all, group block at state_off - node_base * S); syn_pre indexes the group's nodes.
     parts 0 or 1: one program per group. With coop = 1 (synapses between parts) the
     programs of a group advance in lockstep: the launch is cooperative (in consecutive
     launches of whole groups when not all programs fit the device at once), every step
     ends with a grid-wide barrier, and the presynaptic voltages are read from vx, the
     exported voltages in the state layout by step parity ([2][vx_stride], float64).
     Streaming kernel only (items_per_thread 0). */
  int32_t parts;
  int32_t coop;
  int64_t group_nodes;
  void* vx;
  int64_t vx_stride;
  /* Voltage trajectory, optional, in the element type of the gates (the float64
     voltages rounded to it: the values the gates, synapses and linearization compute
     with): the checkpointed forward (ckpt_v set) also stores every item's voltage at
     every step t = 0 ..
     traj_steps - 1 (traj_steps = t_end - t_begin + 1), per group of simulations
     [traj_steps][group nodes][S] at traj_v + (state_off - node_base * S) * traj_steps
     (the program's nodes at node_base * S in each row; group nodes = group_nodes, or n
     without parts); the reverse pass given the same trajectory then recomputes only the
     gates and synapse states between checkpoints (no tree solve) and reads the voltages
     from it. NULL = checkpoints only. */
  void* traj_v;
  int64_t traj_steps;
  /* coef_global = 2 only (else ignored): per kernel node, rows as node_f, the pair
     (s_k, w_k) in the solve type. w_k > 0 scales row k of the system (the tree's root
     rows 1, then w_child = w_parent * upper / lower of the child's edge), and s_k =
     w_k * lower_k = w_parent * upper_k is the coupling of node k and its parent in the
     scaled, symmetric system (0 at roots). The solution is that of the unscaled system;
     the rounding differs from it. */
  const void* node_sym;
  /* coop = 1 only (else ignored): n_progs / parts + 1 zeroed 32-bit words. The programs
     of one group (the parts of a split network for one group of simulations) then
     synchronize through their group's counter instead of a grid-wide barrier; the last
     word is set to 1 if a barrier gave up waiting (after about 2^26 polls; the results
     are then invalid). Zero the words before every launch (the forward and the reverse
     pass count their barriers from 0). NULL: grid-wide barriers. */
  unsigned int* group_sync;
} jxhip_fused_args;

int jxhip_fused_forward_f32(const jxhip_fused_args* args, void* stream);
int jxhip_fused_forward_f64(const jxhip_fused_args* args, void* stream);
/* Mixed precision: float64 gates, rates, parameters, stimulus and recordings (every
 * "element type" array is float64) with the tree solve on float32 increments (node_f
 * and LDS in float32). Meets float64-level trace accuracy at float32 solve cost. */
int jxhip_fused_forward_mixed(const jxhip_fused_args* args, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc). The same
 * forward as jxhip_fused_forward_f64, but the streaming kernel's scratch (work_ax, which
 * must be NULL here) is a 64 MiB array inside the library, one per device, held until the
 * process ends: state_elems (the number of elements of the state arrays, i.e. the
 * largest state_off + n * S over the programs) must be <= 8388608, else the call fails
 * with hipErrorInvalidValue. Calls nothing that allocates. The first call on a device
 * looks the array up (hipGetSymbolAddress) with that device current; make it outside a
 * graph capture. Calls that overlap on one device share the array and race: order them
 * on one stream. */
int jxhip_fused_forward_noalloc_f64(const jxhip_fused_args* args, int64_t state_elems, void* stream);

/* Reverse pass of jxhip_fused_forward_* for HH / Leak cells and networks of them with
 * one synapse type, KineticSynapse (syn_model 0) or SigmoidRateSynapse (1) (no stimulus
 * gradient): for every K-step segment, from the last to the first, recomputes the
 * segment from its checkpoint (storing the trajectory in work_*), then runs the adjoint
 * of each step backwards: transposed tree solve, adjoint of the linearized currents and
 * of the exponential-Euler gates, and of the synapse step (state update, current and its
 * finite-difference slope; each synapse's share of its presynaptic voltage's adjoint
 * is stored at its slot of the presynaptic grouping, pre_pos, and summed per node over
 * the node's contiguous slots, without atomics).
 *   g_rec:   gradient of the recordings, per program [t][rec][S] at rec_off
 *   g_param: per (node, simulation) contributions to the parameter gradients, program
 *            layout [node][S] at state_off (the caller reduces them per parameter mode);
 *            NULL entries are skipped
 *   g_v0:    gradient of the initial voltages, program layout
 *   work_*:  per group of simulations [K + 1][group node][S] at (group index) * work_stride
 *            (a group's programs: the parts of a split network, else one program;
 *            the index is relative to the launch)
 *   lam_*:   adjoint gates, program layout (scratch)
 * Launch of the reverse pass (fields of fwd): threads <= 512; items_per_thread = the voltage
 * adjoints every thread keeps in VGPRs, 1, 2, 4, 8, 12 or 16 with n * S <= items_per_thread
 * * threads, or 0 to keep them in lam_v; coef_global = 1 reads the tree's coupling
 * coefficients from node_f (L2) instead of an LDS copy; lds_bytes = staged schedule +
 * (2 * n + 2 * n_tgt) * S * sizeof(solve type) (+ 2 * n * sizeof(solve type) with LDS
 * coefficients) <= 160 KiB. The forward that wrote the checkpoints must have used the
 * same S and written the synapse states' checkpoints (ckpt_s, KineticSynapse). */
typedef struct jxhip_adjoint_args {
  const void* g_rec;
  void* g_param[JXHIP_NPARAM];
  void* g_v0;
  void* work_v;             /* float64 */
  void* work_m;
  void* work_h;
  void* work_n;
  void* lam_m;
  void* lam_h;
  void* lam_n;
  int64_t work_stride;
  int64_t n_steps;          /* total steps of the forward */
  void* lam_v;              /* adjoint of the voltages, program layout (scratch); required
                               when fwd->items_per_thread == 0, unused otherwise */
  const int32_t* rec_first; /* per node (n_nodes_total): the first recording of the node
                               (topology-relative index), INT32_MAX if none */
  const int32_t* rec_next;  /* per recording (at rec_base): the node's next recording,
                               INT32_MAX after its last */
  /* Synapses (fwd->syn_model 0 or 1): */
  const int32_t* pre_ptr;   /* per topology at pre_ptr_base: n + 1 offsets, the slots of
                               the synapses leaving each node */
  const int32_t* pre_pos;   /* per synapse at syn_base: its slot (topology-relative) */
  void* work_s;             /* recomputed synapse states, per program [K + 1][syn][S] at
                               (program index) * work_syn_stride */
  int64_t work_syn_stride;
  void* lam_s;              /* adjoint of the synapse states, per program [syn][S] at syn_off
                               (scratch) */
  void* work_q;             /* per program [slot][S] at syn_off (scratch): each synapse's
                               share of the presynaptic voltage's adjoint, at its slot */
  void* g_syn_gs;           /* gradient of gS per (synapse, simulation), [syn][S] at syn_off,
                               added to; NULL = not wanted */
  void* work_lin;           /* recomputed synaptic slope of every (target, simulation), per
                               program [K][n_tgt][S] at (program index) * work_lin_stride:
                               the transposed system's diagonal */
  int64_t work_lin_stride;
  int64_t work_q_stride;    /* elements between work_q's two step-parity buffers (split
                               networks: parts write each other's slots); 0 = one buffer */
} jxhip_adjoint_args;

int jxhip_fused_adjoint_f32(const jxhip_fused_args* fwd, const jxhip_adjoint_args* adj,
                            void* stream);
int jxhip_fused_adjoint_f64(const jxhip_fused_args* fwd, const jxhip_adjoint_args* adj,
                            void* stream);
int jxhip_fused_adjoint_mixed(const jxhip_fused_args* fwd, const jxhip_adjoint_args* adj,
                              void* stream);

/* Last HIP error string for an error code (host-only helper). */
const char* jxhip_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif /* JXHIP_H */

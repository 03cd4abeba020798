// cbm_rule.h -- the kernel skeleton around a user agentRule body (struct
// UserRule, compiled with hipRTC).  It restates the generated rule loop of
// src/AgentStructure/functionRule.jl:121-160 with @addAgent (:9-79) and
// @removeAgent (:89-111) in two passes over the agents: the decision pass
// writes each agent's additions and removal by logical slot; after their
// exclusive prefixes, the apply pass evaluates the body again on the same
// inputs and writes its effects, new agents at N + (additions before this
// slot), the reference's serial append order.  The body sees only its own
// agent, so both passes see the same values and make the same decisions.
#ifndef CBM_RULE_H
#define CBM_RULE_H

#include "cbm_stage.h"

namespace cbm {

constexpr int RULE_MAX_ADD = 2;           // @addAgent calls per agent and step
constexpr uint32_t RULE_PURPOSE = 0x600u; // Philox purpose word of the k-th draw: RULE_PURPOSE + k

// Rule error bits in RuleArgs::err.
constexpr int32_t RULE_ERR_CAPACITY = 1;  // more agents than n_max
constexpr int32_t RULE_ERR_ADDS = 2;      // more than RULE_MAX_ADD additions by one agent
constexpr int32_t RULE_ERR_NBVIEW = 4;    // a neighbor loop without the neighbor view (no loop run)
constexpr int32_t RULE_ERR_MEDIUM = 8;    // medium access in a model without a medium (reads 0, adds nothing)


// The medium node of a position: cellInMesh (src/metrics.jl:18-20) per axis,
// clamped to the grid (x beyond the box takes the edge node).
__device__ __forceinline__ int64_t medium_node(const RuleArgs& a, const double* x) {
  int64_t q[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c >= a.dims) break;
    int64_t v;
    if (x[c] > a.med_hi[c]) v = a.med_n[c] - 1;
    else if (x[c] < a.med_lo[c]) v = 0;
    else v = floor_div(x[c] - a.med_lo[c], a.med_dx[c]);
    q[c] = v > a.med_n[c] - 1 ? a.med_n[c] - 1 : v;
  }
  return q[0] + (int64_t)a.med_n[0] * (q[1] + (int64_t)a.med_n[1] * q[2]);
}

// The values of one agent a rule body reads and writes.
struct RuleRec {
  double x[3];
  double vol, growth, vbirth;
  double u[USER_PARAMS];
  int64_t lf;
};

// One neighbor as a rule body's neighbor loop sees it (values at the start
// of the step).
struct RuleNeighbor {
  int64_t row, id;
  double x[3];                 // its position
  double dx[3];                // the agent's position minus it (minimum image in a periodic box)
  double d;                    // euclidean distance (src/metrics.jl:34-36)
  const RuleArgs* args;
  __device__ double vol() const { return args->nb_vol[row]; }
  __device__ double u(int k) const { return args->nb_u[k][row]; }
};

// What a UserRule body receives: the agent's values at the start of the step
// (old), its new values (nw, initialised to old) and the new agents it adds.
// New agents are passed by value and add() merges the record into every
// slot with value selects (never a store to a computed address), so the
// whole object stays in registers also when add() sits in a loop.
struct RuleAgent {
  int64_t id, step;
  int32_t sim, dims;
  double dt;
  uint64_t seed;
  RuleRec old, nw;
  RuleRec dn[RULE_MAX_ADD];
  int32_t n_add, dead, over;
  uint32_t n_draw;
  const RuleArgs* args;
  int64_t row;
  double madd;                 // this agent's medium additions, in call order

  // a new agent's values before the body sets any: the parent's old values
  // (@addAgent copies every parameter it is not given, functionRule.jl:46-55)
  __device__ RuleRec child() const { return old; }
  // @addAgent
  __device__ void add(const RuleRec& d) {
#pragma unroll
    for (int k = 0; k < RULE_MAX_ADD; ++k) pick(dn[k], d, n_add == k);
    if (n_add < RULE_MAX_ADD) n_add++; else over = 1;
  }
  __device__ static void pick(RuleRec& t, const RuleRec& s, bool c) {
#pragma unroll
    for (int i = 0; i < 3; ++i) t.x[i] = c ? s.x[i] : t.x[i];
    t.vol = c ? s.vol : t.vol;
    t.growth = c ? s.growth : t.growth;
    t.vbirth = c ? s.vbirth : t.vbirth;
#pragma unroll
    for (int i = 0; i < USER_PARAMS; ++i) t.u[i] = c ? s.u[i] : t.u[i];
    t.lf = c ? s.lf : t.lf;
  }
  // @removeAgent
  __device__ void remove() { dead = 1; }

  // the medium at the agent's node, as the step started (a medium variable
  // in agent code, vectorizeMediumInAgents: auxiliar.jl:353-365)
  __device__ double medium() const {
    if (!args->med_c) { atomicOr(args->err, RULE_ERR_MEDIUM); return 0.0; }
    return args->med_c[medium_node(*args, old.x)];
  }
  // `p += v` on the agent's node (the docs' secretion idiom, Usage_Agent.md
  // :390-395); lands in the field before the step's medium update
  __device__ void medium_add(double v) { madd = madd + v; }

  // CBMDistributions (src/random.jl:1-23) on this agent's k-th draw of the step
  __device__ double rand() {
    const u4 q = draw(seed, step, id, sim, RULE_PURPOSE + n_draw++);
    return u53(q.x, q.y);
  }
  __device__ double uniform(double l0, double l1) { return (l1 - l0) * rand() + l0; }
  __device__ double normal(double mu, double sigma) {
    double z0, z1;
    normal2(draw(seed, step, id, sim, RULE_PURPOSE + n_draw++), z0, z1);
    return sigma * z0 + mu;
  }
  __device__ double exponential(double theta) { return -log(1.0 - rand()) * theta; }

  // @loopOverNeighbors (src/AgentStructure/macros.jl:97-126): f(RuleNeighbor)
  // for every candidate of the neighbor structure, in its walk order
  template <class F>
  __device__ void forNeighbors(F&& f) const {
    // the host sets up the view when the body's text names forNeighbors
    if (!args->nb_vol) { atomicOr(args->err, RULE_ERR_NBVIEW); return; }
    const SimDev& sd = args->g.sims[sim];
    const int64_t N = args->g.st[sim].N;
    auto visit = [&](int64_t j) {
      RuleNeighbor n;
      n.row = j;
      n.id = args->id[j];
      n.args = args;
      double s2 = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        n.x[c] = 0.0;
        n.dx[c] = 0.0;
        if (c < dims) {
          n.x[c] = args->nb_x[c][j];
          double e = old.x[c] - n.x[c];
          if (args->periodic) e = min_image(e, sd.L[c]);
          n.dx[c] = e;
          s2 = c == 0 ? e * e : s2 + e * e;
        }
      }
      n.d = sqrt(s2);
      f(n);
    };
    switch (args->nb_kind * 4 + dims) {
      case NB_FULL * 4 + 1: for_candidates<1, NB_FULL>(*args, sd, N, row, visit); break;
      case NB_FULL * 4 + 2: for_candidates<2, NB_FULL>(*args, sd, N, row, visit); break;
      case NB_FULL * 4 + 3: for_candidates<3, NB_FULL>(*args, sd, N, row, visit); break;
      case NB_CELL * 4 + 1: for_candidates<1, NB_CELL>(*args, sd, N, row, visit); break;
      case NB_CELL * 4 + 2: for_candidates<2, NB_CELL>(*args, sd, N, row, visit); break;
      case NB_CELL * 4 + 3: for_candidates<3, NB_CELL>(*args, sd, N, row, visit); break;
      case NB_LIST * 4 + 1: for_candidates<1, NB_LIST>(*args, sd, N, row, visit); break;
      case NB_LIST * 4 + 2: for_candidates<2, NB_LIST>(*args, sd, N, row, visit); break;
      default: for_candidates<3, NB_LIST>(*args, sd, N, row, visit); break;
    }
  }
};

__device__ __forceinline__ void rule_load(const RuleArgs& a, int64_t r, RuleRec& v) {
#pragma unroll
  for (int c = 0; c < 3; ++c) v.x[c] = c < a.dims ? a.pos[c][r] : 0.0;
  v.vol = a.vol[r];
  v.growth = a.gr[r];
  v.vbirth = a.vb[r];
#pragma unroll
  for (int k = 0; k < USER_PARAMS; ++k) v.u[k] = a.u[k][r];
  v.lf = a.lf[r];
}

template <class Rule>
__device__ __forceinline__ void rule_eval(const RuleArgs& a, int sim, int64_t r, int64_t step, RuleAgent& ag) {
  ag.id = a.id[r];
  ag.step = step;
  ag.sim = sim;
  ag.dims = a.dims;
  ag.dt = a.dt;
  ag.seed = a.seed;
  ag.n_add = 0; ag.dead = 0; ag.over = 0; ag.n_draw = 0;
  ag.args = &a;
  ag.row = r;
  ag.madd = 0.0;
  rule_load(a, r, ag.old);
  ag.nw = ag.old;
  const Rule rule(a.rp, a.g.sims[sim].p);
  // inlined whatever the body's size, so that ag never needs an address
  [[clang::always_inline]] rule(ag);
}

template <class Rule>
__device__ void user_rule_decide(const RuleArgs& a) {
  const RowIdx ri = row_index(a.g);
  if (ri.local >= a.g.st[ri.sim].N) return;
  RuleAgent ag;
  rule_eval<Rule>(a, ri.sim, ri.row, *a.step_ptr, ag);
  const int32_t s = a.slot[ri.row];
  a.n_add_slot[s] = ag.n_add;
  a.dead_slot[s] = ag.dead;
  if (ag.over) atomicOr(a.err, RULE_ERR_ADDS);
}

template <class Rule>
__device__ void user_rule_apply(const RuleArgs& a) {
  const RowIdx ri = row_index(a.g);
  SimState& st = a.g.st[ri.sim];
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = st.N;
  const int64_t nmax = sd.n_max;
  const int64_t A = a.off_add[nmax];
  if (ri.local == 0) {
    st.n_add = A;
    st.n_remove = a.off_dead[nmax];
    st.n_tot = (N + A <= nmax) ? N + A : nmax;
    if (N + A > nmax) atomicOr(a.err, RULE_ERR_CAPACITY);   // reference D12 drops the agent silently
  }
  if (ri.local >= N) return;
  const int64_t r = ri.row;
  RuleAgent ag;
  rule_eval<Rule>(a, ri.sim, r, *a.step_ptr, ag);
  const int D = a.dims;
  if (ag.madd != 0.0) {
    if (a.med_add)
      atomicAdd(a.med_add + medium_node(a, ag.old.x), (unsigned long long)llrint(ag.madd * MEDIUM_ADD_SCALE));
    else atomicOr(a.err, RULE_ERR_MEDIUM);
  }
  a.dead[r] = (uint8_t)ag.dead;
  // positions and vol are DE state (written in place, read by the stages);
  // growth and vbirth are write-new, committed by update!
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if (c < D) a.pos[c][r] = ag.nw.x[c];
  a.vol[r] = ag.nw.vol;
  a.gr_new[r] = ag.nw.growth;
  a.vb_new[r] = ag.nw.vbirth;
#pragma unroll
  for (int k = 0; k < USER_PARAMS; ++k) a.u[k][r] = ag.nw.u[k];
  a.lf[r] = ag.nw.lf;
  if (ag.n_add == 0) return;
  const int64_t o = a.off_add[a.slot[r]];
#pragma unroll
  for (int k = 0; k < RULE_MAX_ADD; ++k) {
    if (k >= ag.n_add) return;
    const int64_t nw = N + o + k;
    if (nw >= nmax) return;
    const int64_t dst = sd.row_off + nw;
    const RuleRec& d = ag.dn[k];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (c >= D) break;
      const double x = a.periodic ? wrap_box(d.x[c], sd.lo[c], sd.L[c]) : d.x[c];
      a.pos[c][dst] = x;
      a.pos_alt[c][dst] = x;
      if (a.lagged) { a.F[c][dst] = a.F[c][r]; a.Fnew[c][dst] = a.F[c][r]; }
    }
    a.vol[dst] = d.vol; a.vol_alt[dst] = d.vol;
    ((double*)a.gr)[dst] = d.growth; a.gr_new[dst] = d.growth;
    ((double*)a.vb)[dst] = d.vbirth; a.vb_new[dst] = d.vbirth;
#pragma unroll
    for (int q = 0; q < USER_PARAMS; ++q) { a.u[q][dst] = d.u[q]; a.u_alt[q][dst] = d.u[q]; }
    a.lf[dst] = d.lf;
    a.id[dst] = st.id_max + 1 + o + k;
    a.slot[dst] = (int32_t)nw;
    a.dead[dst] = 0;
  }
}

}  // namespace cbm

#endif  // CBM_RULE_H

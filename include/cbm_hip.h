/*
 * cbm_hip.h -- C ABI of libcbm_hip.so, the MI355X (gfx950) per-step agent
 * engine behind the CellBasedModels.jl `ABM` / `Community` / `evolve!` API.
 *
 * Everything here is plain C: integers, doubles and pointers.  No HIP or
 * torch type crosses the boundary.  A host binding (the Julia `ccall`
 * stubs, or the Python ctypes mirror in cellbasedmodels.jl_amd/cbm) calls
 * these entry points; the library owns every device buffer, its HIP stream
 * and its compiled code objects.
 *
 * Conventions
 *  - Return value 0 (CBM_OK) on success, a negative CBM_ERR_* on failure; the
 *    message is available from cbm_last_error(ctx).  The reference raises a
 *    Julia `error()` in the same situations
 *    (src/CommunityStructure/communityStructure.jl:122-124,718-722).
 *  - Agent data crosses the boundary in the reference's *logical slot order*
 *    (slot s = Julia index s+1 of `com.x`, `com.id`, ...).  Internally the
 *    library keeps agents sorted by grid cell; that permutation never leaks.
 *  - Calls on one context are serialised on its stream and are asynchronous
 *    except cbm_download*, cbm_get_*, cbm_sync and cbm_adjoint, which block.
 *  - FP64 state, int64 ids, int32 indices inside the library.
 */
#ifndef CBM_HIP_H
#define CBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum cbm_status {
  CBM_OK = 0,
  CBM_ERR_CONFIG = -1,        /* invalid model / community description      */
  CBM_ERR_CAPACITY = -2,      /* more agents than n_max (reference D12 drops them silently) */
  CBM_ERR_LIST_OVERFLOW = -3, /* a Verlet list needed more than max_neighbors entries */
  CBM_ERR_COMPILE = -4,       /* hipRTC rejected a user interaction body    */
  CBM_ERR_HIP = -5,           /* HIP runtime error                          */
  CBM_ERR_ARG = -6,           /* bad argument (field id, size, sim index)   */
  CBM_ERR_NONFINITE = -7      /* NaN/Inf in an uploaded agent field or medium */
};

/* ---------------------------------------------------------- integrators */
/* src/integrators.jl:465 (Euler), :516 (Heun), :583 (RungeKutta4), :683 (EM),
 * :744 (EulerHeun: the CPU branch, whose corrector draws fresh noise) */
enum cbm_integrator { CBM_EULER = 0, CBM_HEUN = 1, CBM_RK4 = 2, CBM_EM = 3, CBM_EULER_HEUN = 4 };

/* ------------------------------------------------------ neighbor engines */
/* src/neighbors.jl:27 (Full), :196 (VerletTime), :310 (VerletDisplacement),
 * :587 (CellLinked), :997 (CLVD) */
enum cbm_neighbors {
  CBM_FULL = 0,
  CBM_CELL_LINKED = 1,
  CBM_VERLET_DISPLACEMENT = 2,
  CBM_VERLET_TIME = 3,
  CBM_CLVD = 4
};

/* ------------------------------------------------------ interaction laws */
/* Built-in agentODE interaction bodies (the soft-sphere idioms of
 * src/models/softSpheres.jl and docs/src/Aggregation.md:62-123).  Each is a
 * pair force summed over neighbors: F_ij = (s(d)/d) * (x_i - x_j) for LINEAR
 * and REP_ADH, and s(d) * ((x_i - x_j)/d) per axis for AGGREGATION (the
 * reference's dirx = (x - x[i2])/d).
 *   CBM_LAW_LINEAR       p = {k, c}:            s = k (c - d)           for d < c
 *   CBM_LAW_REP_ADH      p = {k, r_eq, eps, f}: c = 2 r_eq, dc = f r_eq
 *                          s = k (c - d)                              d < c
 *                          s = -eps 4 (d - c)(dc - d)/((dc - c)(dc - c))  c <= d < dc
 *   CBM_LAW_AGGREGATION  p = {rRep, fRep, rAtr, fAtr, walls}:
 *                          s = fRep (rRep - d)  d < rRep;  s = -fAtr (rAtr - d)  d < rAtr
 *                          plus the box-wall push fRep within rRep/2 of a wall
 *   CBM_LAW_CUSTOM       user body compiled with hipRTC (cbm_compile)
 *   CBM_LAW_SOFT_SPHERES p = {b, mu, f0[1][1], f0[1][2], f0[2][1], f0[2][2]}:
 *     the inertial soft-sphere model of src/models/softSpheres.jl:1-81.  The
 *     DE variables are the positions and the velocities (CBM_F_VX..VZ); each
 *     agent has a radius (CBM_F_RADIUS), a mass (CBM_F_MASS) and a cellFate
 *     in {1, 2} (CBM_F_FATE, int64).  With r_ij = r_i + r_j, for 0 < d < mu r_ij
 *       F_i += (f0[fate_i][fate_j] (r_ij/d - 1)(mu r_ij/d - 1)(x_i - x_j))/d
 *     and dv/dt = (-b v)/m + F/m, dx/dt = v.  No rules or adjoint.
 *   CBM_LAW_RODS p = {kn, gamma_n, gamma_t, mu_cc, mu_cw, beta, beta_w}:
 *     the 2-D rod model of Volfson et al. (src/models/rods.jl:21-124, the
 *     rod geometry of src/metrics.jl:138-246).  The DE variables are x, y,
 *     the velocities (CBM_F_VX, VY), the angle theta (CBM_F_THETA) and its
 *     rate omega (CBM_F_OMEGA); each rod has a diameter (CBM_F_DIAMETER), a
 *     length (CBM_F_LENGTH) and a mass (CBM_F_MASS).  Over the neighbors in
 *     walk order the closest virtual spheres of the two rods give the
 *     contact force (Fx, Fy) and torque W of rodForces, summed into fx, fy,
 *     W and pressure (+= |F|), written every stage (CBM_F_FX, FY,
 *     CBM_F_TORQUE, CBM_F_PRESSURE); then dt(v) = -beta v + f/m, dt(theta) =
 *     omega, dt(omega) = W/(m (d + l)^2/12 + m d^2) - beta_w omega.  2-D,
 *     no periodic box, no rules, Euler, Heun or RK4.                      */
enum cbm_law {
  CBM_LAW_NONE = 0,
  CBM_LAW_LINEAR = 1,
  CBM_LAW_REP_ADH = 2,
  CBM_LAW_AGGREGATION = 3,
  CBM_LAW_CUSTOM = 4,
  CBM_LAW_SOFT_SPHERES = 5,
  CBM_LAW_RODS = 6
};

/* ---------------------------------------------------------- medium solvers */
/* CBM_MEDIUM_FTCS   explicit c' = c + dt f(c) with zero-flux mirror ghosts.
 * CBM_MEDIUM_DGADI  the reference's Douglas-Gunn ADI (src/integrators.jl:
 *   821-935 with the tridiagonal sweeps of :141-428): the reaction part of
 *   the body (its Laplacian terms zero) as c += dt f, then one implicit sweep
 *   per axis (x, then y, then z) solving (1 + 2 D r) c_i - D r (c_{i-1} +
 *   c_{i+1}) = c_i^old on the interior nodes by the Thomas algorithm, with
 *   r = dt / (dims dx^2) of the last axis for every sweep and boundary nodes
 *   held at their initial values (the field at creation or at the last
 *   cbm_upload_medium).  Needs >= 4 nodes per axis. */
enum cbm_medium_alg {
  CBM_MEDIUM_FTCS = 0,
  CBM_MEDIUM_DGADI = 1
};

/* Border of the medium on one side of one axis (cbm_model_desc.medium_border).
 * MIRROR: the border nodes integrate like the others, with a ghost node
 *   mirroring them (zero flux).
 * DIRICHLET: the border nodes take medium_border_value each step (`p = v`).
 * COPY: the border nodes take the value of their inward neighbor along that
 *   axis at the step start (`p = p[2, i2_]`, the reference's reflective
 *   idiom).
 * A node on borders of several sides takes the first of them in the order
 * x-, x+, y-, y+, z-, z+ (the order of an if/elseif chain over the sides);
 * inside nodes read the border nodes' step-start values. */
enum cbm_medium_border { CBM_BORDER_MIRROR = 0, CBM_BORDER_DIRICHLET = 1, CBM_BORDER_COPY = 2 };

/* ------------------------------------------------------------ agent rules */
/* Built-in agentRule bodies (src/AgentStructure/functionRule.jl:9-111).
 *   CBM_RULE_GROWTH_DIVISION rule_p = {g_mean, g_spread, split, ratio}:
 *     vol grows by dvol/dt = growth*vol (integrated with the agent DE);
 *     when vol >= ratio*vbirth the agent divides: parent and one new agent
 *     are displaced -/+ split along a uniform random axis, each gets vol/2,
 *     vbirth = vol/2 and a fresh growth ~ U[g_mean(1-g_spread), g_mean(1+g_spread)].
 *   CBM_RULE_PARITY  the reference add/remove known-answer rule
 *     (test/testCommunity.jl:241-316): x += 1; odd id -> @removeAgent();
 *     even id -> @addAgent(lf = id).
 *   CBM_RULE_DEATH   rule_p = {rate}: each agent is removed with probability
 *     rate*dt per step (Philox uniform keyed by seed, step, id).
 *   CBM_RULE_CUSTOM  a user agentRule body compiled by cbm_compile_model
 *     (required before stepping).  The agents carry vol, growth, vbirth
 *     (dvol/dt = growth*vol, as for CBM_RULE_GROWTH_DIVISION), lf and the
  *     user parameters CBM_F_U0..U7 for the body to read and write.
 *   CBM_RULE_RODS_GROWTH  rule_p = {growth, sigma_lTarget, lMax, alpha}:
 *     rods2dGrowth (src/models/rods.jl:123-164) on CBM_LAW_RODS.  The length
 *     l becomes a DE variable, dt(l) = growth/(1 + alpha*pressure) with the
 *     stage's pressure; a rod with l > lTarget (CBM_F_LTARGET) is replaced by
 *     two daughters at +-(l + d)/4 along its axis with length (l - d)/2,
 *     omega +- ww, ww uniform in [-0.1, 0.1), and lTarget uniform in
 *     [lMax - sigma, lMax + sigma) each (Philox keyed seed, step, id, sim,
 *     0x700/0x701); every other field is the parent's.                    */
enum cbm_rule {
  CBM_RULE_NONE = 0,
  CBM_RULE_GROWTH_DIVISION = 1,
  CBM_RULE_PARITY = 2,
  CBM_RULE_DEATH = 3,
  CBM_RULE_CUSTOM = 4,
  CBM_RULE_RODS_GROWTH = 5
};

/* ------------------------------------------------------ per-agent fields */
/* Field ids for cbm_upload / cbm_download.  Doubles unless marked int64. */
enum cbm_field {
  CBM_F_X = 0, CBM_F_Y = 1, CBM_F_Z = 2,       /* positions (DE variables)    */
  CBM_F_ID = 3,                                /* int64, 1-based like `id`    */
  CBM_F_FX = 4, CBM_F_FY = 5, CBM_F_FZ = 6,    /* committed force accumulator (lagged mode) */
  CBM_F_VOL = 7, CBM_F_GROWTH = 8, CBM_F_VBIRTH = 9,
  CBM_F_LF = 10,                               /* int64 payload (`lf` of the add/remove KAT) */
  CBM_F_VX = 11, CBM_F_VY = 12, CBM_F_VZ = 13,  /* velocities (CBM_LAW_SOFT_SPHERES)        */
  CBM_F_RADIUS = 14, CBM_F_MASS = 15,
  CBM_F_FATE = 16,                             /* int64 cellFate, 1-based                   */
  CBM_F_U0 = 17, CBM_F_U1 = 18, CBM_F_U2 = 19, CBM_F_U3 = 20,   /* user agent parameters (CBM_RULE_CUSTOM) */
  CBM_F_THETA = 21, CBM_F_OMEGA = 22,          /* rods: angle and its rate (DE variables)  */
  CBM_F_DIAMETER = 23, CBM_F_LENGTH = 24,      /* rods: per-rod geometry                    */
  CBM_F_TORQUE = 25, CBM_F_PRESSURE = 26,      /* rods: W and pressure of the last stage    */
  CBM_F_U4 = 27, CBM_F_U5 = 28, CBM_F_U6 = 29, CBM_F_U7 = 30,   /* user agent parameters U4..U7 */
  CBM_F_LTARGET = 31,                          /* rods with CBM_RULE_RODS_GROWTH: division length */
  CBM_F_COUNT = 32
};

/* --------------------------------------------------- model description */
/* What `ABM(...)` + the neighbor/integrator constructors fix for a model.
 * Replaces ABM(dims; agentODE, agentSDE, agentRule, neighborsAlg, agentAlg)
 * (src/AgentStructure/agentStructure.jl:131-304). */
typedef struct cbm_model_desc {
  int32_t dims;          /* 1, 2 or 3                                        */
  int32_t integrator;    /* enum cbm_integrator                              */
  int32_t neighbors;     /* enum cbm_neighbors                               */
  int32_t law;           /* enum cbm_law                                     */
  int32_t rule;          /* enum cbm_rule                                    */
  int32_t periodic;      /* 1: periodic box, minimum-image distances         */
  int32_t lagged;        /* 1: dt(x) reads the previous step's committed force
                            accumulator (reference read-old/write-new idiom,
                            src/AgentStructure/agentStructure.jl:352-453)     */
  int32_t max_neighbors; /* Verlet list capacity (nMaxNeighbors)             */
  double dt;
  double cell_edge;      /* CellLinked / CLVD cellEdge                       */
  double skin;           /* Verlet list radius (the reference's `skin`)      */
  double rebuild_disp;   /* displacement that triggers a Verlet rebuild;
                            <= 0 selects the reference rule skin/2
                            (src/neighbors.jl:403)                          */
  double dt_recompute;   /* VerletTime dtNeighborRecompute                   */
  double drift;          /* agentODE linear term: dt(x) += -drift*x          */
  double noise;          /* agentSDE diagonal term g (EM only)               */
  double rule_p[4];      /* rule parameters (see enum cbm_rule)              */
  uint64_t seed;         /* Philox4x32-10 key for all in-kernel randomness   */
  /* medium (diffusion of one field on a cell-centred grid) */
  int32_t medium_n[3];   /* NMedium; all zero = no medium                    */
  int32_t medium_alg;    /* enum cbm_medium_alg                              */
  double medium_D;       /* diffusion coefficient                            */
  double medium_consume; /* uptake rate per agent in a node: dc/dt -= rate*c*n */
  double medium_c0;      /* initial concentration                            */
  int32_t user_de;       /* U0..U(user_de-1) are DE variables of the agentODE
                            body (UserLaw::de, dt(u) of functionDE.jl:13-116);
                            0..4, needs CBM_LAW_CUSTOM, no soft spheres     */
  int32_t walk;          /* 0: neighbor lists walked one agent per lane; a
                            CellLinked model with a built-in law (LINEAR,
                            REP_ADH, AGGREGATION) under Euler or Heun and no
                            rules runs on a skin list (no per-step cell sort;
                            each agent's entries ordered as the cell walk).
                            1: as 0, but CellLinked sorts its grid every step
                            and walks it on LDS brick tiles.  512..2048: as
                            0, with list stages flattened over each block's
                            lanes in rounds of that many list entries (a
                            round's terms sit in LDS: 2048 keeps a 3-D block
                            under 64 KiB).  The results are the same bits for
                            every value.                                    */
  /* medium borders, per axis and side in the order x-, x+, y-, y+, z-, z+
   * (enum cbm_medium_border): the nodes of a side with a Dirichlet or copy
   * border are assigned each step instead of integrated, as a mediumODE
   * body's `@mediumBorder(axis, side)` branch assigns them
   * (src/AgentStructure/macros.jl:284-311; docs/src/Usage_Agent.md:397-403)
   * while `@mediumInside()` nodes integrate dt(c).  FTCS only.           */
  int32_t medium_border[6];
  double medium_border_value[6];  /* the Dirichlet values                    */
} cbm_model_desc;

/* Per-simulation description.  One context may hold many independent
 * simulations (the batched sweep); each has its own agents, box and law
 * parameters.  Replaces Community(abm; N, simBox, ...)
 * (src/CommunityStructure/communityStructure.jl:108-187). */
typedef struct cbm_sim_desc {
  int64_t n;             /* initial agent count N                            */
  int64_t n_max;         /* capacity (N + preallocateAgents)                 */
  double box_lo[3];      /* simBox[:,1]                                      */
  double box_hi[3];      /* simBox[:,2]                                      */
  double p[8];           /* law parameters (see enum cbm_law)                */
} cbm_sim_desc;

typedef struct cbm_ctx cbm_ctx;

/* ------------------------------------------------------------ lifecycle */
/* ABM(...) + Community(...) + loadToPlatform! allocation
 * (src/CommunityStructure/communityStructure.jl:468-655). */
int cbm_create(cbm_ctx** out, const cbm_model_desc* model, int32_t n_sims,
               const cbm_sim_desc* sims);
void cbm_destroy(cbm_ctx* ctx);
const char* cbm_last_error(const cbm_ctx* ctx);
/* Library version string and device name (for records). */
const char* cbm_version(void);

/* compileABM! / Main.eval (src/AgentStructure/functionDE.jl:110): compile a
 * user interaction body with hipRTC into the stage kernel skeleton and use it
 * as the model's law (model law CBM_LAW_CUSTOM; required before stepping).
 * The body is HIP C++ defining, inside namespace cbm:
 *   struct UserLaw {
 *     static constexpr bool active = true, per_axis_division = false, walls = false;
 *     __device__ explicit UserLaw(const double* p);   // the simulation's p[8]
 *     __device__ bool operator()(double d, double& s) const;  // s(d); true if
 *   };                                                // the pair interacts
 * The force on agent i is sum_j (s/d)(x_i - x_j) (s*((x_i - x_j)/d) per axis
 * when per_axis_division).  A UserLaw may also define
 *     __device__ void self(const double* x, double* f) const;
 * the agentODE body's per-agent terms: f (zeroed, D values) is added to
 * dt(x) at the stage positions x, after the pair sum and `drift`.  With
 * user_de > 0 it defines
 *     __device__ void de(const double* x, const double* u, double* du) const;
 * dt of the user DE variables: du[k] (zeroed) = dt(U_k) at the stage
 * positions x and stage values u[k] of U_k, k < user_de, integrated by the
 * model's integrator beside the positions (one normal each under EM).
 * A general agentODE body (src/AgentStructure/functionDE.jl:13-116: neighbor
 * loops that read the neighbors' parameters and sum into any variable)
 * defines instead of operator():
 *     static constexpr int sums = K;                     // 1..8
 *     __device__ void pair(const LawAgent& i, const LawAgent& j,
 *                          const double* dx, double d, double* s) const;
 *     __device__ void ode(const LawAgent& i, const double* s,
 *                         double* dxdt, double* dudt) const;
 * pair is called for every neighbor candidate j of agent i in the reference
 * loop order (dx = x_i - x_j, minimum image in a periodic box, d = |dx|)
 * and adds its terms to the sums s (zeroed first) in the order it writes
 * them; ode then gives dt of the positions (dxdt) and of the user DE
 * variables (dudt) from the sums.  A LawAgent holds row, id, x[3] and
 * u(k), vol(): U_k and vol at the stage state for DE variables, as stored
 * otherwise.  Such a body walks its candidates row by row (no brick
 * tiles) and takes no lagged forces.  Compile errors return
 * CBM_ERR_COMPILE with the hipRTC log in cbm_last_error. */
int cbm_compile(cbm_ctx* ctx, const char* pair_body_hip);

/* Compile user bodies as cbm_compile_model would for a model described by
 * `model` (its dims, neighbors, law, rule, periodic and user_de), without a
 * context or a GPU: the check of compileABM! (src/AgentStructure/
 * functionRule.jl:121-160, functionDE.jl:13-116) before a run.  Returns
 * CBM_OK, or CBM_ERR_COMPILE with the compiler log in `log` (log_cap bytes
 * with the terminating NUL; log may be NULL). */
int cbm_compile_check(const cbm_model_desc* model, const char* law_body, const char* noise_body,
                      const char* rule_body, const char* medium_body, char* log, int64_t log_cap);

/* compileABM! for the agentODE and agentSDE bodies together
 * (src/AgentStructure/functionDE.jl:13-116 builds both f and g).  law_body
 * as for cbm_compile (required with CBM_LAW_CUSTOM, NULL otherwise: the
 * built-in law is compiled around the other body).  noise_body (EM or
 * EulerHeun models; NULL = the model's constant `noise` on the positions)
 * is HIP C++ defining, inside namespace cbm:
 *   struct UserNoise {
 *     __device__ explicit UserNoise(const double* p);   // the simulation's p[8]
 *     __device__ double operator()(int var, const double* u) const;
 *   };
 * the diagonal noise amplitude g of DE variable `var` at the agent state u
 * (positions, then velocities or volume), evaluated at each stage's state as
 * the reference's g: EM u' = u + f dt + g xi sqrt(dt)
 * (src/integrators.jl:709-727); EulerHeun averages g of both stages with a
 * fresh normal (:775-803).  One normal per DE variable, keyed by (seed, step,
 * agent id) as for the constant noise.
 * rule_body (functionRule.jl:121-160 with @addAgent/@removeAgent, :9-111;
 * required with CBM_RULE_CUSTOM, NULL otherwise) is HIP C++ defining, inside
 * namespace cbm:
 *   struct UserRule {
 *     __device__ UserRule(const double* rp, const double* p);  // rule_p[4], p[8]
 *     __device__ void operator()(RuleAgent& a) const;
 *   };
 * run once per agent and step before the DE stages.  RuleAgent (cbm_rule.h)
 * holds a.old, the agent's values at the start of the step (x[3], vol,
 * growth, vbirth, u[8], lf), and a.nw, its new values (initialised to old);
 * a.id, a.step, a.dt and a.dims; a.child() is a new agent's record
 * (the parent's old values) and a.add(d) adds it (@addAgent; at most 2 per
 * agent and step); a.remove() removes the agent (@removeAgent);
 * a.uniform(l0, l1), a.normal(mu, sigma) and a.exponential(theta) are the
 * reference's CBMDistributions (src/random.jl:1-23) on Philox(seed, step,
 * id, sim, 0x600 + k), k counting the agent's draws in the step;
 * a.forNeighbors(f) is @loopOverNeighbors (macros.jl:97-126): f(const
 * RuleNeighbor& n) for every candidate of the neighbor structure in its walk
 * order, with n.x, n.dx (minimum image), n.d, n.id, n.vol() and n.u(k) as
 * they were before the rule ran.  New agents get ids idMax + 1 + (prefix of
 * additions in slot order), the reference's serial append order.
 * a.medium() is the medium at the agent's node (cellInMesh of its position)
 * as the step started, and a.medium_add(v) is `p += v` there (a medium
 * variable in agent code, auxiliar.jl:353-365; the secretion idiom of
 * docs/src/Usage_Agent.md:390-395): the additions of all agents are summed
 * per node in 32.32 fixed point (exact, so their order does not matter)
 * and land in the field before the step's medium update.
 * medium_body (mediumODE, functionDE.jl:13-116 with the medium loop of
 * auxiliar.jl:168-172,184-210 and @∂2 of macros.jl:203-257; models with a
 * medium, NULL = the built-in diffusion with uptake) is HIP C++ defining,
 * inside namespace cbm:
 *   struct UserMedium {
 *     __device__ UserMedium(double D, double consume, const double* p);   // medium_D, medium_consume, p[8]
 *     __device__ double operator()(const MediumNode& m) const;            // dt(c) at the node
 *   };
 * MediumNode (cbm_medium.h) holds c, lap[3] (@∂2 per axis with zero-flux
 * mirror ghosts), laplacian(), the node centre x[3] and index idx[3], the
 * number of agents in the node (cellInMesh, metrics.jl:18-20), dt, the
 * neighbors' values nb[axis][0 below / 1 above], border[axis] (-1 / +1 on a
 * lower / upper border, @mediumBorder) and inside() (@mediumInside); the
 * step is c' = c + dt * dt(c) for every node, except nodes of a side with a
 * Dirichlet or copy border (cbm_model_desc.medium_border) and nodes for
 * which an optional `__device__ bool assign(const MediumNode&, double& v)
 * const` returns true: those are set to v (the `p = ...` branches).
 * Each call replaces the previous module, so pass every body of the model in
 * one call. */
int cbm_compile_model(cbm_ctx* ctx, const char* law_body, const char* noise_body,
                      const char* rule_body, const char* medium_body);

/* ------------------------------------------------------------- snapshots */
/* evolve!(saveEach) with saveRAM! (src/CommunityStructure/step.jl:124-155,
 * IO.jl:6-31) kept on the device: the gathers of each snapshot are queued on
 * the context's stream and nothing synchronises the host until a snapshot is
 * read back.
 * cbm_snapshot_config: room for `capacity` snapshots of the listed fields
 *   (8-byte fields only; every simulation; plus the medium when the model has
 *   one).  Replaces an earlier configuration and discards its snapshots.
 * cbm_snapshot_take: a snapshot of the current state (saveCurrentState).
 * cbm_step_save: nsteps steps as cbm_step, with a snapshot after step i for
 *   every i % save_each == 0 (i = 1..nsteps).  CBM_ERR_CAPACITY before any
 *   step if the snapshots would not fit.
 * cbm_snapshot_get: snapshot k's values of `field` for simulation `sim` in
 *   slot order (its N values; *n_out = N, *t_out = t at the snapshot);
 *   CBM_ERR_ARG if N > n_cap. */
int cbm_snapshot_config(cbm_ctx* ctx, const int32_t* fields, int32_t nfields, int32_t capacity);
int cbm_snapshot_take(cbm_ctx* ctx);
int cbm_step_save(cbm_ctx* ctx, int64_t nsteps, int64_t save_each);
int cbm_snapshot_count(cbm_ctx* ctx, int32_t* count);
int cbm_snapshot_get(cbm_ctx* ctx, int32_t k, int32_t field, int32_t sim, void* host, int64_t n_cap,
                     int64_t* n_out, double* t_out);
int cbm_snapshot_get_medium(cbm_ctx* ctx, int32_t k, double* host, int64_t n);

/* ------------------------------------------------------- data transfer */
/* loadToPlatform! / bringFromPlatform! (communityStructure.jl:606-711).
 * `n` values in logical slot order for simulation `sim`.  Uploading CBM_F_X
 * of simulation `sim` does not change N; N is set at creation or with
 * cbm_set_n.  Int64 fields take/return int64_t, the others double; a NaN or
 * infinite double is refused with CBM_ERR_NONFINITE (nothing uploaded). */
int cbm_upload(cbm_ctx* ctx, int32_t field, const void* host, int64_t n, int32_t sim);
int cbm_download(cbm_ctx* ctx, int32_t field, void* host, int64_t n, int32_t sim);
int cbm_set_n(cbm_ctx* ctx, int32_t sim, int64_t n);
/* Medium field download/upload (row-major x fastest), sim 0 only. */
int cbm_upload_medium(cbm_ctx* ctx, const double* host, int64_t n);
int cbm_download_medium(cbm_ctx* ctx, double* host, int64_t n);

/* -------------------------------------------------------------- stepping */
/* computeNeighbors! (src/neighbors.jl:18): (re)build the neighbor
 * structures from the current positions.  Called once after upload. */
int cbm_neighbors(cbm_ctx* ctx);
/* step! x nsteps (src/CommunityStructure/step.jl:96-107): rule -> DE stages
 * -> medium -> commit/compaction -> neighbor update, all device resident. */
int cbm_step(cbm_ctx* ctx, int64_t nsteps);
/* Record nsteps of the step loop into a HIP graph and replay it `replays`
 * times (the evolve! loop without per-step host work). */
int cbm_step_graph(cbm_ctx* ctx, int64_t nsteps, int64_t replays);
int cbm_sync(cbm_ctx* ctx);

/* ------------------------------------------------------------- queries */
int cbm_get_n(cbm_ctx* ctx, int32_t sim, int64_t* n);        /* com.N      */
int cbm_get_t(cbm_ctx* ctx, double* t, int64_t* step);      /* com.t      */
int cbm_get_id_max(cbm_ctx* ctx, int32_t sim, int64_t* id_max); /* idMax_ */
/* Number of Verlet rebuilds (or CL binnings) performed so far. */
int cbm_get_rebuilds(cbm_ctx* ctx, int64_t* rebuilds);
/* Cell-brick launch geometry (no reference counterpart; diagnostics of the
 * LDS-tiled CellLinked stage and Verlet list build): the brick edge in cells
 * per axis (bdim[3]), the agents one brick block stages in LDS (cap), the
 * number of bricks, and how many brick launches so far found more than cap
 * agents in their halo and left its rows to the row-parallel kernel (same
 * results, slower).  Zeros for Full. */
int cbm_get_bricks(cbm_ctx* ctx, int32_t* bdim, int32_t* cap, int64_t* nbricks,
                   int64_t* overflows);
/* The sub-brick table of a CellLinked context with rules (no reference
 * counterpart; diagnostics as cbm_get_bricks): a finer tiling whose bricks
 * each lie inside one main brick, launched after the main bricks to stage
 * the rows of those too dense for LDS, each sub-brick sized for 32 agents
 * per cell.  Zeros when the context has none. */
int cbm_get_sub_bricks(cbm_ctx* ctx, int32_t* bdim, int32_t* cap, int64_t* nbricks);
/* Neighbor lists of simulation `sim` in logical slots: counts[n] and
 * lists[n*cap] (row i holds the slots of agent i's list, unordered).  For
 * CellLinked/Full the "list" is the candidate set within `radius` of the
 * current structure.  Returns CBM_ERR_LIST_OVERFLOW if cap is too small. */
int cbm_get_neighbor_lists(cbm_ctx* ctx, int32_t sim, double radius,
                           int32_t* counts, int64_t* lists, int64_t cap);
/* The neighbor-count agentRule of the reference KAT
 * (test/testCommunity.jl:361-415): out[i] = #{j in structure : d_ij < radius}.
 * For Verlet models, radius <= 0 returns the list lengths instead. */
int cbm_count_neighbors(cbm_ctx* ctx, int32_t sim, double radius, int64_t* out);

/* cbm_count_neighbors without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code that must not call
 * hipMalloc/hipFree, e.g. around graph capture or in tight serving loops).
 * The counts go through one file-scope device array per GPU (8 MiB, held
 * until the process ends), so at most 1,048,576 agents per call; a larger
 * population returns CBM_ERR_ARG.  Calls that overlap on one GPU (contexts
 * used from several host threads) share that array and race. */
int cbm_count_neighbors_noalloc(cbm_ctx* ctx, int32_t sim, double radius, int64_t* out);

/* --------------------------------------------------------------- adjoint */
/* Gradient of a trajectory loss through the overdamped-Euler mechanics step
 * (no reference counterpart: the reference has no AD).  The loss is
 *   L = sum_{m} mean_i |x_i(t_m) - x*_i(m)|^2      (mean over agents and
 * save points is taken as 1/(n_saves*N) sum ...), for the REP_ADH law with
 * parameters (k, r_eq, eps) of every simulation.  target holds n_saves
 * snapshots of dims*N doubles (SoA per snapshot: x[N], y[N], z[N]) per sim,
 * concatenated over sims.  Checkpoints every `checkpoint_every` (<= 64)
 * steps are kept on device and the steps between them recomputed in the
 * backward pass.  The Verlet lists of the forward pass are kept too while
 * they fit in half of the free device memory (or in CBM_ADJ_SNAPSHOT_MB MiB
 * when that environment variable is set); otherwise the lists are checked
 * and rebuilt along the backward pass, with identical results.  target
 * NULL reuses the targets of the previous call (same n_saves and N).  Work
 * buffers and list snapshots stay in the context for the next call and are
 * released by cbm_destroy.
 * grad[3*s + {0,1,2}] = dL_s/d(k, r_eq, eps); loss[s] = L_s. */
typedef struct cbm_loss_desc {
  int64_t n_steps;
  int64_t checkpoint_every;
  int32_t n_saves;
  int32_t pad;
  const int64_t* save_steps;   /* n_saves step indices in (0, n_steps]     */
  const double* target;        /* see above; NULL: the previous call's     */
} cbm_loss_desc;
int cbm_adjoint(cbm_ctx* ctx, const cbm_loss_desc* loss_desc, double* grad,
                double* loss);

/* Model-scope parameters of simulation `sim` (the law parameters p[0..n-1]);
 * replaces assigning com.k etc. between runs of a fitting loop
 * (test/testFitting.jl:4-29). */
int cbm_set_params(cbm_ctx* ctx, int32_t sim, const double* p, int32_t n);

/* The current model-scope parameters p[0..n-1] of simulation `sim`, as a
 * model-scope body has left them (com.k etc. after a run). */
int cbm_get_params(cbm_ctx* ctx, int32_t sim, double* p, int32_t n);

/* modelRule / modelODE / modelSDE (src/AgentStructure/agentStructure.jl:43-45,
 * run by modelStepRule! and modelStepDE! after the medium step,
 * src/CommunityStructure/step.jl:96-107): compile a model-scope body with
 * hipRTC; every later step runs it once per simulation.  The model
 * variables are the simulation's parameters p[8].  The body is HIP C++
 * defining, inside namespace cbm, with every member optional:
 *   struct UserModel {
 *     __device__ void rule(ModelScope& m) const;                   // modelRule
 *     static constexpr unsigned de = MASK;                         // bit k: p[k] is a DE variable
 *     __device__ void ode(const ModelScope& m, double* dp) const;  // modelODE: dp[k] = dt(p[k])
 *     __device__ void noise(const ModelScope& m, double* g) const; // modelSDE: g of p[k]
 *   };
 * A ModelScope holds t (the step's start time), dt, N, step, sim, p[8] (the
 * values at the step's start) and pn[8] (the new values, initialised to p;
 * the rule writes them).  DE parameters step by Euler, or Euler-Maruyama
 * with noise: pn[k] = (p[k] + dp[k] dt) + g[k] z sqrt(dt), z normal from
 * Philox keyed (seed, step, 0, sim, 0x600 + k/2).  pn is committed at the
 * step's end, as update! commits write-new values: agents and the medium
 * read p during the step and pn from the next step on.  CBM_ERR_CONFIG in
 * a context running CellLinked on a skin list (the lists are sized for the
 * law's reach at creation: create it with walk = 1); CBM_ERR_COMPILE with
 * the hipRTC log in cbm_last_error. */
int cbm_compile_model_scope(cbm_ctx* ctx, const char* model_body_hip);

/* cbm_compile_model_scope's compilation alone, without a context or GPU;
 * the log as for cbm_compile_check. */
int cbm_compile_model_scope_check(const char* model_body_hip, char* log, int64_t log_cap);

/* -------------------------------------------------------------- autotune */
/* Pick block size (and, for Verlet models, list radius margin) by timing a
 * few steps; the choice is cached per (model, device) in `cache_path` (may
 * be NULL).  Writes the chosen block size to *block. */
int cbm_autotune(cbm_ctx* ctx, const char* cache_path, int32_t* block);
int cbm_set_block(cbm_ctx* ctx, int32_t block);

/* ------------------------------------------------------------ profiling */
/* Run nsteps steps with HIP events around every launch on the context
 * stream (no reference counterpart: the measurement hook of the benchmark).
 * avg_ms[k] / launches[k] for k = 0..7: the integrator stage kernels
 * (0 Euler, 1 EM, 2 Heun stage 1, 3 Heun stage 2, 4-7 RK4 stages 1-4);
 * k = 8: the neighbor update of one step (total, launches = steps);
 * k = 9: whole step; k = 10, 11: EulerHeun stages 1 and 2; k = 12 is
 * reserved (no launches: CellLinked on a skin list orders its lists inside
 * the first stage).  ncat must be >= 13. */
int cbm_profile(cbm_ctx* ctx, int64_t nsteps, double* avg_ms, int64_t* launches, int32_t ncat);
/* Average device time (ms) of the last cbm_step call's force kernels, from
 * HIP events on the context stream; *launches = force-kernel launches timed. */
int cbm_last_timing(cbm_ctx* ctx, double* ms_total, double* ms_force,
                    int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* CBM_HIP_H */

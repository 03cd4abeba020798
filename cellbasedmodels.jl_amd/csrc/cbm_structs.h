// cbm_structs.h -- device-visible structures of the step kernels, shared by
// libcbm_hip.so and the hipRTC programs built around user interaction
// bodies (so their layout is identical on both sides).
#ifndef CBM_STRUCTS_H
#define CBM_STRUCTS_H

#ifdef CBM_HIPRTC
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
using __hip_internal::uint8_t;
using __hip_internal::uint16_t;
#else
#include <stdint.h>
#endif

namespace cbm {

constexpr int MAX_COLS = 20;      // permutable per-agent columns
constexpr int USER_PARAMS = 8;    // user agent parameters U0..U7 (the first user_de of them DE variables)
constexpr int MAX_STATE = 8;      // DE variables: 3 positions + vol + 4 user, or 3 positions + 3 velocities
constexpr int BUILTIN_STATE = 6;  // without user DE variables

// Per-simulation constants (device array, one entry per simulation).
// Row space: simulation s owns rows [row_off, row_off + n_max).
// Cell space: simulation s owns global buckets [cell_off, cell_off + ncells + 2):
// its grid cells, then a bucket for removed agents, then a "gap" bucket
// sized so that a global exclusive scan of bucket counts lands every bucket
// start on its global row index.
struct SimDev {
  double lo[3], hi[3], L[3];
  double p[8];
  double gedge[3];
  int32_t gn[3];
  int32_t pad0;
  int64_t row_off, n_max;
  int64_t cell_off, ncells;
  int64_t rank_off;     // offset of this simulation's cell -> Morton-rank table
  int64_t pad1;
  // CellLinked on a skin list: the model's own cell grid (the reference's
  // padded grid of edge cellEdge, or the periodic grid of edge L/n), whose
  // cells order each agent's candidates; gedge/gn above are then the list
  // build's grid
  double cedge[3];
  int32_t cgn[3];
  int32_t pad2;
};

// Per-simulation mutable scalars (device array).
struct SimState {
  int64_t N;        // com.N (agents committed)
  int64_t n_tot;    // rows populated this step (N + NAdd_ before compaction)
  int64_t n_add;    // NAdd_
  int64_t n_remove; // NRemove_
  int64_t id_max;   // idMax_
  int64_t pad[3];
};

// Launch geometry shared by row-parallel kernels: block b serves
// simulation block_sim[b], rows (b - block_off[sim]) * blockDim + tid.
struct RowGeom {
  const int32_t* block_sim;
  const int32_t* block_off;
  const SimDev* sims;
  SimState* st;
};

// Integrator stage kinds (one kernel launch each).  See src/integrators.jl.
enum StageKind {
  ST_EULER = 0,     // u' = u + du*dt                        (:496-498)
  ST_EM = 1,        // u' = (u + f*dt) + (g*xi)*sqrt(dt)      (:724-727)
  ST_HEUN1 = 2,     // h = u + du*dt                          (:551-553)
  ST_HEUN2 = 3,     // u' = (h + (u + du*dt))/2               (:564-566)
  ST_RK4_1 = 4,     // k = k1;        h = u + k1*dt/2         (:620-624)
  ST_RK4_2 = 5,     // k = k + 2*k2;  h = u + k2*dt/2
  ST_RK4_3 = 6,     // k = k + 2*k3;  h = u + k3*dt
  ST_RK4_4 = 7,     // u' = u + ((k + k4)*dt)/6                (:659-661)
  ST_EH1 = 8,       // k = f;  h = (u + f*dt) + (g*xi1)*sqrt(dt)   EulerHeun (:787-789)
  ST_EH2 = 9        // u' = u + ((k + f)*dt + ((g + g)*xi2)*sqrt(dt))/2  (:801-803)
};

enum NeighborKind { NB_FULL = 0, NB_CELL = 1, NB_LIST = 2 };

// Rows [s, e) of one grid cell after a rebuild, indexed by linear cell.
struct alignas(8) CellRange { int32_t s, e; };

// LDS of one brick block: a header of int32 words, then D position arrays
// (and radius for soft spheres) of cap doubles, int32 fate (soft spheres),
// the halo cells' LDS offsets (H + 1) and first rows (H), the per-lane
// queues of kept candidates, the uint16 halo cell of each staged agent and
// the uint16 LDS index of each home row.
constexpr int BRICK_HDR = 16;        // int32 words: total, per-wave sums; int64 at byte 48
constexpr int BRICK_LIST = 12;       // queue entries per lane
constexpr int BRICK_MAX_BLOCK = 256;
constexpr int BRICK_WORDS = 6;       // brick table entry: sim, origin x, y, z, first bucket, end bucket
constexpr double CL_LIST_GROW = 1.05;  // radius of the stage-2 lists, in units of the law's reach

__host__ __device__ constexpr int brick_halo_cells(int D, int bx, int by, int bz) {
  return (bx + 2) * (D > 1 ? by + 2 : 1) * (D > 2 ? bz + 2 : 1);
}

__host__ __device__ inline int64_t brick_lds_bytes(int D, bool soft, int cap, int H) {
  int64_t b = 4 * BRICK_HDR;
  b += (int64_t)8 * cap * (D + (soft ? 1 : 0));
  if (soft) b += (int64_t)4 * cap;
  b += (int64_t)4 * (2 * H + 1);
  b += (int64_t)2 * BRICK_LIST * BRICK_MAX_BLOCK;
  b += (int64_t)4 * cap;      // owner and home_li, uint16 each
  return (b + 15) & ~(int64_t)15;
}

// Inputs of the kernels that tile cell bricks (the stage, and the Verlet
// list build).
struct BrickSrc {
  const SimDev* sims;
  const int32_t* block_off;       // row-kernel block of each simulation's first row
  const int32_t* brick;           // BRICK_WORDS per brick
  const int32_t* bucket_start;    // first row of each global bucket
  const CellRange* cell_range;    // rank_off + linear cell -> rows
  int32_t bdim[3];
  int32_t cap;                    // agents a brick can stage
  uint8_t* brick_skip;            // per row: left to the row-parallel kernel
  uint8_t* blk_flag;              // per row-kernel block: holds such rows
  unsigned long long* brick_over; // count of such bricks
  int32_t sub;                    // sub-brick pass: only bricks whose rows an earlier pass left over
};

// LDS of one Verlet-build brick block: the header, D position arrays of cap
// doubles, int32 slots, the halo offsets and first rows, the per-lane
// slot-sorted queues (VBRICK_Q uint16 LDS indices per lane), then the uint16
// owner and home-row arrays.
constexpr int VBRICK_Q = 16;

__host__ __device__ inline int64_t vbrick_lds_bytes(int D, int cap, int H) {
  int64_t b = 4 * BRICK_HDR;
  b += (int64_t)8 * cap * D + (int64_t)4 * cap;
  b += (int64_t)4 * (2 * H + 1);
  b += (int64_t)2 * VBRICK_Q * BRICK_MAX_BLOCK;
  b += (int64_t)4 * cap;
  return (b + 15) & ~(int64_t)15;
}

// What the skin-list order kernel reports to the host each step (host
// memory mapped into the device): the step it ran for and, as of the end of
// the previous step and since the last list build, whether a final stage
// predicted an agent to pass half the skin soon (the host then schedules a
// build) and the largest squared displacement reported (bit pattern).
struct ClvRecord {
  int64_t step;
  unsigned long long soon;
  unsigned long long e2max;
  int64_t pad;
};

struct StageArgs {
  RowGeom g;
  const double* S[MAX_STATE];   // stage state read by the force (neighbors + self)
  const double* U[MAX_STATE];   // step-start state u
  double* out[MAX_STATE];       // stage output (h, or u')
  double* kacc[MAX_STATE];      // RK4 accumulator
  double* gacc[MAX_STATE];      // EulerHeun with a user agentSDE body: g of stage 1
  const double* F[3];           // committed force accumulator (lagged mode)
  double* Fnew[3];              // force accumulator written by this stage
  const double* xref[3];        // positions at the last Verlet build
  const double* growth;         // per-agent growth rate (vol variable)
  const double* rad;            // soft spheres: radius, mass, cellFate (1-based) per row
  const double* mass;
  const int64_t* fate;
  const double* rod_d;          // rods: diameter and length per row (mass in mass)
  const double* rod_l;
  double* rod_out[4];           // rods: fx, fy, W, pressure of this stage
  int32_t* rod_err;             // rods: set when a contact's overlap d_i - r_ij is negative
  double rod_growth, rod_alpha;  // rods with CBM_RULE_RODS_GROWTH: dt(l) = growth/(1 + alpha pressure)
  const double* upar[USER_PARAMS];   // user agent parameter columns U0..U7 (null when absent)
  const int64_t* id;
  const int32_t* list;          // ELL Verlet list, entry q of row i at q*ldl + i
  const int32_t* list_n;
  int64_t ldl;
  const CellRange* cell_range;  // rank_off + linear local cell -> its rows (cell-linked)
  const int32_t* cell_row;      // row -> linear local cell (cell-linked)
  const int32_t* brick;         // cell-linked brick blocks: BRICK_WORDS per block
  const int32_t* bucket_start;  // first row of each global bucket (the rebuild's scan)
  unsigned long long* brick_over;   // count of brick blocks left to the row-parallel kernel
  uint8_t* brick_skip;          // per row: 1 = left to the row-parallel kernel (cell-linked; it clears it)
  uint8_t* blk_flag;            // per row-kernel block: holds rows left to it (cleared by it)
  // Cell-linked two-stage integrators: stage 1 writes, per agent, its
  // candidates within CL_LIST_GROW x the law's reach (walk order) so that
  // stage 2 can walk them as a list; lgate is cleared when they may miss a
  // pair (a stage-1 displacement over half the margin, a list over wcap, a
  // brick too dense for LDS).  A launch with gate != NULL runs only when
  // *gate == gate_run.
  int32_t* wlist;
  int32_t* wlist_n;
  int32_t* lgate;
  const int32_t* gate;
  int32_t* flag;                // rebuild flag written by the final stage
  double dt, sqdt, drift, noise;
  double thr2;                  // largest d2 with sqrt(d2) < rebuild displacement
  uint64_t seed;
  const int64_t* step_ptr;     // device step counter (Philox counter word)
  int32_t kind, periodic, lagged, walls, nvar, disp_check, write_fnew, pad;
  int32_t bdim[3];              // brick edge in cells per axis
  int32_t brick_cap;            // agents a brick block can stage in LDS
  // Sub-bricks (cell-linked contexts whose population can concentrate): a
  // finer brick table, each of whose bricks lies inside one brick of the
  // main table.  After the main brick launch a second launch over it (its
  // copy of these arguments has brick_sub = 1) takes the rows of the main
  // bricks too dense for LDS; what it cannot stage either is left to the
  // row-parallel kernel.  brick2 is NULL when there is no such table.
  const int32_t* brick2;
  int64_t nbricks2;
  int64_t brick_lds2;
  int32_t bdim2[3];
  int32_t brick_cap2;
  int32_t brick_sub;
  int32_t wcap;                 // entries per agent of wlist
  int32_t gate_run;
  int32_t fchunk;               // flattened list stage: pairs per LDS round
  int32_t fcap;                 // flattened list stage: the list capacity (entries per agent)
  // CellLinked on a skin list (clv != 0): packed cells (cl_pack) of the
  // step-start positions on the model's grid, read by the stages; the final
  // stage writes those of its output to ccell_out.  Stage 1 of a two-stage
  // integrator adds its agents' |h - x_ref|^2 to *emax (bit pattern, max);
  // stage 2 reads it.  Agents whose list may miss a pair take the exact
  // walk (clv_walk): clv_e1 bounds |x - x_ref| at step start, clv_skin is the
  // list margin.
  const int32_t* ccell;
  int32_t* ccell_out;
  unsigned long long* emax;
  const int32_t* cslot;         // logical slot of every row (clv_walk's order)
  double clv_e1, clv_skin;
  double clv_half, clv_half2;   // skin/2 (a little above) and its square: stage 1 reports
                                // only agents past it, and E is at least clv_half
  int32_t clv, lcap;            // lcap: entries of a full (possibly truncated) skin list
  int32_t* fix_rows;            // agents deferred to the exact walk (clv_fix_body)
  int32_t* fix_n;               // their count, per stage ([0] stage 1, [1] stage 2)
  int32_t fix_stage, nsims;
  // since the last list build (both cleared by it): efin[0], the largest
  // |x - x_ref|^2 past half the skin of a final stage (bit pattern, max;
  // stage 1 bounds the step-start displacement by it); efin[1], set when a
  // final stage predicts an agent to pass half the skin within clv_ahead
  // step displacements
  unsigned long long* efin;
  double clv_ahead;
  // clv_order: this stage (the step's first) orders the skin list (a.list,
  // a.list_n) into a.wlist/a.wlist_n itself, and clears *emax_next (the
  // next step's stage-1 maximum).  Every stage queues its deferred agents
  // in fix_n[fix_q] and clears fix_n[1 - fix_q].  book: the first stage
  // also writes rec (step_no, efin) for the host, and the final stage's
  // fix kernel advances *tick (the device step).
  unsigned long long* emax_next;
  ClvRecord* rec;
  int64_t* tick;
  int64_t step_no;
  int32_t book, clv_order, fix_q, pad3;
};

// Flattened list stage (flat_body): per block, the agents' list entries are
// numbered agent-major; every lane evaluates pairs of that sequence and an
// agent's terms are then added in list order.  LDS: the pair terms of one
// round (chunk + cap entries: a round takes whole agents), the home
// positions, the agents' pair offsets, the rounds' first pairs, the scan's
// wave sums and the owner (agent) of each pair of the round.
constexpr int FLAT_MAX_CAP = 64;       // longer lists take the per-agent walk
constexpr int FLAT_MAX_ROUNDS = (BRICK_MAX_BLOCK * FLAT_MAX_CAP) / 512 + 2;

__host__ __device__ inline int64_t flat_lds_bytes(int D, int chunk, int cap) {
  const int64_t ch = (int64_t)chunk + cap;
  int64_t b = 8 * (int64_t)D * ch;                    // pair terms
  b += 8 * (int64_t)D * BRICK_MAX_BLOCK;              // home positions
  b += 4 * (int64_t)(BRICK_MAX_BLOCK + 1);            // pair offsets of the agents
  b += 4 * (int64_t)(FLAT_MAX_ROUNDS + 1);            // first pair of every round
  b += 4 * 8;                                         // wave sums, total
  b += ch;                                            // owner of each pair (uint8)
  return (b + 15) & ~(int64_t)15;
}

// Agent rules (agentRule with @addAgent/@removeAgent) and the canonical
// compaction of update!; also the argument of the hipRTC user rule kernels.
// Arrays marked (slot) are indexed by logical slot.
// Medium additions of rule bodies are summed per node in 32.32 fixed point:
// exact and the same in any order (the reference's unsynchronized `p += v`
// of several agents in one node has no defined sum).  |sum of one agent's
// additions in a step| < 2^31.
constexpr double MEDIUM_ADD_SCALE = 4294967296.0;

// The rod columns of RuleArgs::rcol (CBM_RULE_RODS_GROWTH).
enum RodCol { RC_VX = 0, RC_VY, RC_TH, RC_OM, RC_LEN, RC_DIAM, RC_MASS, RC_LT, RC_FX, RC_FY, RC_TQ, RC_PR, RC_N };

struct RuleArgs {
  RowGeom g;
  int32_t rule, dims, lagged, growth;
  int32_t periodic, pad;        // periodic box: new agents' positions are wrapped into it
  double rp[4];
  double dt;
  uint64_t seed;
  const int64_t* step_ptr;
  double* pos[3];               // current positions (DE state, written in place)
  double* vol;                  // current volume (DE state)
  double* pos_alt[3];           // the other DE-state buffers: new agents are
  double* vol_alt;              // written to both (out-of-place stages flip them)
  const double* gr;             // growth rate (read-old)
  double* gr_new;               // growth rate (write-new)
  const double* vb;             // birth volume (read-old)
  double* vb_new;               // birth volume (write-new)
  double* F[3];                 // committed force accumulator (lagged)
  double* Fnew[3];
  int64_t* id;
  int64_t* lf;
  double* u[USER_PARAMS];       // user agent parameters (user rules, user DE variables), written in place
  double* u_alt[USER_PARAMS];   // their other buffers: new agents are written to both (user DE
                                // variables flip with the positions under Euler and EM)
  int32_t* slot;
  uint8_t* dead;                // per row
  int32_t* n_add_slot;          // (slot) daughters of this agent, consumed by the scan
  int32_t* dead_slot;           // (slot) 1 if removed, consumed by the scan
  const int32_t* off_add;       // (slot) exclusive prefix of n_add_slot, [n_max] = total
  const int32_t* off_dead;      // (slot) exclusive prefix of dead_slot, [n_max] = total
  int32_t* holes;               // ascending hole slots below the new N
  int32_t* err;                 // capacity error flag
  int32_t* rebuild_flag;        // set when agents were added or removed
  // neighbor loops of a user rule: the structure of the last neighbor update
  // and a copy of the values a neighbor exposes, taken before the rule's
  // passes (the apply pass writes the agents' own values in place)
  const int32_t* list;
  const int32_t* list_n;
  int64_t ldl;
  const int32_t* cell_row;
  const CellRange* cell_range;
  int32_t nb_kind, pad2;
  const double* nb_x[3];
  const double* nb_vol;
  const double* nb_u[USER_PARAMS];
  // rods (CBM_RULE_RODS_GROWTH): the rod columns a daughter takes from its
  // parent (both buffers), RC_* order
  double* rcol[RC_N];
  double* rcol_alt[RC_N];
  // the medium at the agent's node (cellInMesh of its step-start position,
  // src/AgentStructure/auxiliar.jl:353-365): read from med_c (the field at
  // the step start, NULL without a medium); additions are summed per node
  // in med_add as 32.32 fixed point by the apply pass
  const double* med_c;
  unsigned long long* med_add;  // two's complement
  int32_t med_n[3], med_pad;
  double med_lo[3], med_hi[3], med_dx[3];
};

// Medium: one field on a cell-centred grid over the box of simulation 0,
// stepped by explicit FTCS with zero-flux (mirror) boundaries; dt(c) is the
// built-in diffusion with agent uptake or a user mediumODE body (@∂2,
// src/AgentStructure/macros.jl:203-257).  Also the argument of the hipRTC
// user medium kernel.
struct MediumArgs {
  RowGeom g;
  const double* pos[3];         // agent positions after the DE stages
  const double* c;              // concentration (read)
  double* cn;                   // concentration (write)
  int32_t* count;               // agents per node, consumed (zeroed) by the stencil
  int32_t n[3];                 // NMedium
  int32_t dims;
  int32_t lap, pad;             // lap = 0: the body's Laplacian terms are zero (ADI reaction part)
  double dx[3];                 // (hi - lo) / n per axis
  double lo[3], hi[3];
  double D, uptake, dt;
  int64_t nodes;
  int32_t bmode[6];             // border per axis side x-, x+, y-, y+, z-, z+ (cbm_medium_border)
  double bval[6];               // Dirichlet values
};

// Arguments of the model-scope kernel (modelStepRule! and modelStepDE!,
// src/CommunityStructure/step.jl:21-30, :66-75): one thread per simulation.
struct ModelArgs {
  SimDev* sims;                 // the simulations' parameters p[8] (read, then committed)
  const SimState* st;           // N per simulation
  const int64_t* step_ptr;      // the device step counter
  double* t_buf;                // the time at the start of step s in t_buf[s & 1]
  double dt;
  uint64_t seed;
  int32_t S, pad;
};

}  // namespace cbm

#endif  // CBM_STRUCTS_H

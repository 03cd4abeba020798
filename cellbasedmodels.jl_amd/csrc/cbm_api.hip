// cbm_api.hip -- host side of libcbm_hip.so: the C ABI declared in
// include/cbm_hip.h, context/buffer management and the step orchestration
// that replaces step!/update!/computeNeighbors! of the reference
// (src/CommunityStructure/step.jl:96-107, update.jl:177-321,
// src/neighbors.jl:18).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <limits>
#include <vector>
#include <algorithm>
#include <mutex>
#include <chrono>
#include "cbm_internal.h"
#include "cbm_device.h"
#include <hip/hiprtc.h>
#include "cbm_rtc_src.inc"

// Skin-list builds scheduled by the host (clv_schedule): the host runs at
// most CLV_HOST_LEAD steps ahead of the device, and a final stage cues a
// build when CLV_AHEAD more displacements like its step's would carry an
// agent past half the skin (the lead, the step that ends with the build and
// one and a half more).
constexpr int64_t CLV_HOST_LEAD = 3;
// Builds that follow each other within CLV_DENSE_GAP steps twice in a row
// mean agents cross the skin faster than the host's lead: the stages would
// defer ever more agents to the exact walk.  (One such gap is common while
// random initial overlaps relax.)  The context then builds after
// every step for CLV_DENSE_STEPS steps before it schedules again.  The
// closest two scheduled builds can be is the lead + 2 steps: the build
// clears the cue, the next step's final stage raises it, the step after
// records it, and the host reads that record CLV_HOST_LEAD steps later.
constexpr int64_t CLV_DENSE_GAP = CLV_HOST_LEAD + 2;
constexpr int64_t CLV_DENSE_STEPS = 64;
constexpr double CLV_AHEAD = 2.5 + (double)CLV_HOST_LEAD;


using namespace cbm;

#define CBM_VERSION "cbm_hip 0.1 (gfx950, FP64)"

// ------------------------------------------------------------------ context
struct Column {
  void* buf[2] = {nullptr, nullptr};
  int cidx = 0;
  int bytes = 8;
  void* cur() const { return buf[cidx]; }
  void* alt() const { return buf[1 - cidx]; }
};

enum ColId {
  C_X = 0, C_Y, C_Z, C_VOL,            // DE state group (flip together)
  C_VX, C_VY, C_VZ,                     // velocities (soft spheres), DE state too
  C_ID, C_SLOT, C_FX, C_FY, C_FZ, C_GR, C_VB, C_LF,
  C_RAD, C_MASS, C_FATE,                // per-agent soft-sphere properties
  C_U0, C_U1, C_U2, C_U3, C_U4, C_U5, C_U6, C_U7,   // user agent parameters U0..U7
  C_TH, C_OM,                           // rods: angle and its rate, DE state too
  C_DIAM, C_LEN, C_TQ, C_PR,            // rods: diameter, length (a DE variable with growth), torque, pressure
  C_LT,                                 // rods with CBM_RULE_RODS_GROWTH: the division length lTarget
  C_CCELL,                              // CellLinked on a skin list: packed cell of the step start (int32)
  C_NCOLS
};

struct cbm_ctx {
  cbm_model_desc m;
  std::vector<cbm_sim_desc> sims;
  int S = 1;
  int dims = 1, nvar = 1;
  bool growth = false, lagged = false, verlet = false, cl = false, full = false, vtime = false, rules = false;
  bool inertial = false;                  // CBM_LAW_SOFT_SPHERES: velocities are DE variables
  bool rods = false;                      // CBM_LAW_RODS: x, y, vx, vy, theta, omega are DE variables
  // model scope (modelStepRule!/modelStepDE!): a hipRTC module of its own
  hipModule_t model_mod = nullptr;
  hipFunction_t model_fn = nullptr;
  double* model_t = nullptr;              // the step's start time by step parity (device)
  // CellLinked executed on a skin list (cl && clv): no per-step cell sort;
  // each stage orders an agent's list entries by its cell among the 3^D
  // neighbor cells (the reference loop order) from the agents' packed cells
  bool clv = false;
  double clv_reach = 0.0, clv_skin = 0.0;  // largest law reach at creation, list margin
  unsigned long long* emax = nullptr;       // max |h - x_ref|^2 of stage 1 (bit pattern)
  int32_t* wlist = nullptr;                 // the step's lists in cell-walk order (clv)
  int32_t* wlist_n = nullptr;
  int32_t* fix_rows = nullptr;              // agents deferred to the exact walk, and
  int32_t* fix_n = nullptr;                 // their counts per stage
  unsigned long long* efin = nullptr;       // displacement bound and build cue since the last build
  unsigned long long* med_add = nullptr;    // a rule body's medium additions per node (32.32 fixed point)
  // the host schedules the skin-list builds of eager steps from the record
  // the order kernel writes each step (mapped host memory); graph capture
  // keeps the device-side trigger
  ClvRecord* rec_host = nullptr;
  ClvRecord* rec_dev = nullptr;
  int64_t last_build_step = -1;             // step whose end saw the last build
  int clv_short = 0;                        // consecutive builds CLV_DENSE_GAP or fewer steps apart
  int64_t clv_dense_until = -1;             // build after every step before this one
  int64_t rec_first = 0;                    // first step whose record the device writes
  bool capturing = false;
  bool build_now = false;
  int nb_kind = NB_FULL;
  int block = 256;
  int device = 0;
  hipStream_t stream = nullptr;

  // geometry
  std::vector<int32_t> h_block_sim, h_block_off;
  int nblocks = 0;
  int32_t* d_block_sim = nullptr;
  int32_t* d_block_off = nullptr;
  std::vector<SimDev> h_sims;
  SimDev* d_sims = nullptr;
  SimState* d_st = nullptr;
  int64_t C = 0;      // rows
  int64_t NB = 0;     // buckets
  int64_t NR = 0;     // rank-table entries (sum of cells)
  std::vector<int32_t> h_crank;
  double list_radius = 0.0, list_r2 = 0.0, thr = 0.0;

  // per-agent columns
  Column col[C_NCOLS];
  bool has_col[C_NCOLS] = {};
  double* Fnew[3] = {nullptr, nullptr, nullptr};
  double* gr_new = nullptr;     // growth rule write-new buffers
  double* vb_new = nullptr;
  int32_t* n_add_slot = nullptr; // rule decisions by logical slot
  int32_t* dead_slot = nullptr;
  int32_t* off_add = nullptr;    // their exclusive prefixes
  int32_t* off_dead = nullptr;
  int32_t* holes = nullptr;
  double* h[2][MAX_STATE] = {};

  double* kacc[MAX_STATE] = {};
  double* xref[3] = {nullptr, nullptr, nullptr};
  uint8_t* dead = nullptr;

  // neighbor engine
  int32_t* gbucket = nullptr;             // rank_off + cell -> global bucket (cell_off + Morton rank)
  CellRange* cell_range = nullptr;        // rank_off + cell -> its rows after a rebuild
  int32_t* lin = nullptr;
  int32_t *key = nullptr, *count = nullptr, *cursor = nullptr, *start = nullptr, *tmp = nullptr, *cell_row = nullptr;
  unsigned long long* scan_status = nullptr;   // one status word per scan tile
  int64_t scan_tiles = 0;                      // capacity of scan_status in tiles
  unsigned int* scan_ticket = nullptr;    // [0] ticket, [1] done, [2] epoch
  int32_t* list = nullptr;
  int32_t* list_n = nullptr;
  int cap = 0;
  bool wl = false;                        // cell-linked two-stage: stage 2 walks lists written by stage 1
  int32_t wcap = 0;
  int32_t* bricks = nullptr;              // cell-linked brick blocks (BRICK_WORDS each)
  uint8_t* brick_skip = nullptr;          // per row: left to the row-parallel stage kernel
  uint8_t* blk_flag = nullptr;            // per row-kernel block: holds such rows
  int64_t nbricks = 0;
  int bdim[3] = {1, 1, 1};
  int brick_cap = 0;                      // agents staged per brick block
  size_t brick_lds = 0;                   // its LDS bytes
  int32_t* bricks2 = nullptr;             // sub-bricks for dense regions (StageArgs::brick2)
  int64_t nbricks2 = 0;
  int bdim2[3] = {1, 1, 1};
  int brick_cap2 = 0;
  size_t brick_lds2 = 0;

  // adjoint work buffers, kept across calls and grown on demand
  struct DevBuf { void* p = nullptr; size_t bytes = 0; };
  DevBuf adj_tgt, adj_tbase, adj_invn, adj_ck, adj_seg, adj_lam, adj_gpart, adj_lpart;
  int adj_tgt_saves = 0;                  // shape of the targets held in adj_tgt
  std::vector<int64_t> adj_tgt_n;
  std::vector<int32_t*> snap_l, snap_n;   // Verlet-list snapshots of the forward pass
  int32_t* flags = nullptr;               // [0..1] rebuild flags by step parity, [2] list overflow, [3] capacity
  int64_t* counters = nullptr;            // [0] rebuilds, [1] device step, [2] brick overflows
  bool structure_valid = false;

  // medium
  bool medium = false;
  double* med_c[2] = {nullptr, nullptr};
  int med_cur = 0;
  int32_t* med_count = nullptr;
  int64_t med_nodes = 0;
  // DGADI: the initial field (held on the boundary), a sweep buffer, and per
  // axis c*[i] and m[i] (n entries each) of the constant-diagonal Thomas solve
  bool adi = false;
  double* med_init = nullptr;
  double* med_tmp = nullptr;
  double* adi_tab[3] = {nullptr, nullptr, nullptr};
  double adi_sub = 0.0, adi_main = 0.0, adi_sup = 0.0;

  // staging for transfers
  void* d_staging = nullptr;
  // evolve!(saveEach) snapshots on the device: per field capacity x C values
  // at [k * C + row_off + slot], N per simulation, the medium, t on the host
  std::vector<int32_t> snap_fields;
  std::vector<void*> snap_buf;
  int64_t* snap_cnt = nullptr;
  double* snap_med = nullptr;
  int32_t snap_cap = 0, snap_count = 0;
  std::vector<double> snap_t;
  int64_t staging_bytes = 0;

  // time
  int64_t step = 0;
  double t = 0.0;
  double t_next = 0.0;

  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool profiling = false;
  std::vector<hipEvent_t> prof_ev;       // pairs
  std::vector<int> prof_cat;
  size_t prof_used = 0;
  double last_ms = 0.0;
  int64_t last_launches = 0;

  // user interaction body compiled with hipRTC (law CBM_LAW_CUSTOM)
  hipModule_t user_mod = nullptr;
  hipFunction_t user_fn[2] = {nullptr, nullptr};   // [0] deterministic stages, [1] noisy (EM, EulerHeun)
  hipFunction_t user_brick[2] = {nullptr, nullptr};   // the same, cell-linked brick tiles
  hipFunction_t user_rule[2] = {nullptr, nullptr};    // agentRule body: decide, apply
  hipFunction_t user_medium = nullptr;                // mediumODE body: the FTCS step
  double* rule_nb[4 + USER_PARAMS] = {};  // a neighbor-looping rule body's view: x, y, z, vol, u0..u7
  int user_nb = NB_FULL;                  // neighbor walk the module was built for
  bool user_noise = false;                // the module holds a user agentSDE body
  double* gacc[MAX_STATE] = {};           // EulerHeun with a user agentSDE body: g of stage 1

  std::string err;
};

static int fail(cbm_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                        \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail((ctx), CBM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
static hipError_t dalloc(T** p, int64_t n) {
  *p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e == hipSuccess) e = hipMemset(*p, 0, sizeof(T) * (size_t)n);
  return e;
}

static RowGeom geom(cbm_ctx* c) {
  RowGeom g;
  g.block_sim = c->d_block_sim;
  g.block_off = c->d_block_off;
  g.sims = c->d_sims;
  g.st = c->d_st;
  return g;
}

static void prof_begin(cbm_ctx* c, int cat) {
  if (!c->profiling || c->prof_used + 2 > c->prof_ev.size()) return;
  c->prof_cat[c->prof_used / 2] = cat;
  (void)hipEventRecord(c->prof_ev[c->prof_used], c->stream);
}
static void prof_end(cbm_ctx* c) {
  if (!c->profiling || c->prof_used + 2 > c->prof_ev.size()) return;
  (void)hipEventRecord(c->prof_ev[c->prof_used + 1], c->stream);
  c->prof_used += 2;
}

// ----------------------------------------------------------------- create
static int validate(cbm_ctx* c) {
  const cbm_model_desc& m = c->m;
  if (m.dims < 1 || m.dims > 3) return fail(c, CBM_ERR_CONFIG, "dims must be 1, 2 or 3");
  if (m.integrator < CBM_EULER || m.integrator > CBM_EULER_HEUN) return fail(c, CBM_ERR_CONFIG, "unknown integrator");
  if (m.neighbors < CBM_FULL || m.neighbors > CBM_CLVD) return fail(c, CBM_ERR_CONFIG, "unknown neighbor algorithm");
  if (m.law < CBM_LAW_NONE || m.law > CBM_LAW_RODS) return fail(c, CBM_ERR_CONFIG, "unknown interaction law");
  if (m.law == CBM_LAW_SOFT_SPHERES && m.rule != CBM_RULE_NONE)
    return fail(c, CBM_ERR_CONFIG, "the soft-sphere model takes no agent rules");
  if (m.law == CBM_LAW_RODS) {
    // rods2D (src/models/rods.jl:61-121): a 2-D agentODE without rules
    if (m.dims != 2) return fail(c, CBM_ERR_CONFIG, "the rod model is 2-D");
    if (m.periodic) return fail(c, CBM_ERR_CONFIG, "the rod model has no periodic box (its contact geometry uses raw coordinates)");
    if (m.lagged) return fail(c, CBM_ERR_CONFIG, "the rod model sums its contacts at the stage state (no lagged forces)");
    if (m.integrator != CBM_EULER && m.integrator != CBM_HEUN && m.integrator != CBM_RK4)
      return fail(c, CBM_ERR_CONFIG, "the rod model is an agentODE: Euler, Heun or RK4");
    if (m.medium_n[0] > 0) return fail(c, CBM_ERR_CONFIG, "the rod model takes no medium");
    if (m.drift != 0.0) return fail(c, CBM_ERR_CONFIG, "the rod model takes no drift term");
  }
  if (!(m.dt > 0)) return fail(c, CBM_ERR_CONFIG, "dt must be positive");
  if (m.walk != 0 && m.walk != 1 && (m.walk < 512 || m.walk > 2048))
    return fail(c, CBM_ERR_CONFIG, "walk is 0, 1 or a round of 512..2048 list entries");
  if (m.user_de < 0 || m.user_de > 4 || (m.user_de > 0 && m.law != CBM_LAW_CUSTOM))
    return fail(c, CBM_ERR_CONFIG, "user_de is 0..4 and needs a custom law (UserLaw::de)");
  if (m.neighbors == CBM_CELL_LINKED && !(m.cell_edge > 0)) return fail(c, CBM_ERR_CONFIG, "CellLinked needs cellEdge > 0");
  bool verlet = m.neighbors == CBM_VERLET_DISPLACEMENT || m.neighbors == CBM_VERLET_TIME || m.neighbors == CBM_CLVD;
  if (verlet && !(m.skin > 0)) return fail(c, CBM_ERR_CONFIG, "Verlet algorithms need skin > 0");
  if (verlet && (m.max_neighbors < 1 || m.max_neighbors > 4096)) return fail(c, CBM_ERR_CONFIG, "nMaxNeighbors must be in [1, 4096]");
  if (m.neighbors == CBM_VERLET_TIME && !(m.dt_recompute > 0)) return fail(c, CBM_ERR_CONFIG, "VerletTime needs dtNeighborRecompute > 0");
  if (m.rule < CBM_RULE_NONE || m.rule > CBM_RULE_RODS_GROWTH) return fail(c, CBM_ERR_CONFIG, "unknown agent rule");
  if ((m.rule == CBM_RULE_RODS_GROWTH) != (m.law == CBM_LAW_RODS && m.rule != CBM_RULE_NONE))
    return fail(c, CBM_ERR_CONFIG, "the rod model takes CBM_RULE_RODS_GROWTH or no rule, and that rule needs it");
  if (m.rule != CBM_RULE_NONE && c->S != 1) return fail(c, CBM_ERR_CONFIG, "agent rules (division/death) need a single simulation per context");
  if (m.medium_n[0] > 0) {
    if (c->S != 1) return fail(c, CBM_ERR_CONFIG, "a medium needs a single simulation per context");
    for (int a = 0; a < m.dims; ++a)
      if (m.medium_n[a] < 1) return fail(c, CBM_ERR_CONFIG, "NMedium must be >= 1 on every axis");
    if (!(m.medium_D >= 0.0) || !(m.medium_consume >= 0.0)) return fail(c, CBM_ERR_CONFIG, "medium D and uptake must be >= 0");
    if (m.medium_alg != CBM_MEDIUM_FTCS && m.medium_alg != CBM_MEDIUM_DGADI)
      return fail(c, CBM_ERR_CONFIG, "unknown medium solver");
    if (m.medium_alg == CBM_MEDIUM_DGADI)
      for (int a = 0; a < m.dims; ++a)
        if (m.medium_n[a] < 4) return fail(c, CBM_ERR_CONFIG, "DGADI needs NMedium >= 4 on every axis");
    for (int k = 0; k < 6; ++k) {
      const int32_t b = m.medium_border[k];
      if (b < CBM_BORDER_MIRROR || b > CBM_BORDER_COPY) return fail(c, CBM_ERR_CONFIG, "unknown medium border");
      if (b != CBM_BORDER_MIRROR && k >= 2 * m.dims) return fail(c, CBM_ERR_CONFIG, "medium border on an axis the model does not have");
      if (b != CBM_BORDER_MIRROR && m.medium_alg == CBM_MEDIUM_DGADI)
        return fail(c, CBM_ERR_CONFIG, "DGADI holds its boundary at the initial field: medium borders need FTCS");
      if (b == CBM_BORDER_DIRICHLET && !std::isfinite(m.medium_border_value[k]))
        return fail(c, CBM_ERR_CONFIG, "a Dirichlet medium border needs a finite value");
    }
  }
  for (int s = 0; s < c->S; ++s) {
    const cbm_sim_desc& sd = c->sims[s];
    if (sd.n < 0 || sd.n_max < sd.n || sd.n_max < 1) return fail(c, CBM_ERR_CONFIG, "need 0 <= N <= n_max and n_max >= 1");
    for (int a = 0; a < m.dims; ++a)
      if (!(sd.box_hi[a] > sd.box_lo[a])) return fail(c, CBM_ERR_CONFIG, "simBox must have hi > lo on every axis");
  }
  return CBM_OK;
}

// Bucket order of the cells: rank of each linear cell (x fastest) along the
// Morton (Z-order) curve, so that rows sorted by bucket are spatially compact.
static uint64_t spread_bits(uint64_t v, int dims) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (b * dims);
  return r;
}

// Largest s with sqrt(s) < r.  A correctly rounded sqrt is monotone, so the
// list test sqrt(d2) < r of the reference (src/neighbors.jl:114-126) is
// exactly d2 <= sq_threshold(r), with no square root per candidate.
static double sq_threshold(double r) {
  if (!(r > 0)) return -1.0;
  double a = r * r;
  while (a > 0 && !(std::sqrt(a) < r)) a = std::nextafter(a, 0.0);
  while (std::sqrt(std::nextafter(a, INFINITY)) < r) a = std::nextafter(a, INFINITY);
  return a;
}

// Sorted list entries a lane keeps in LDS during the list build.
static int verlet_lds_r(const cbm_ctx* c) {
  const int r = VERLET_LDS_BYTES / (8 * c->block);
  return std::max(1, std::min(r, (int)c->cap));
}

static void morton_ranks(const int32_t* gn, int dims, std::vector<int32_t>& out) {
  const int64_t nc = (int64_t)gn[0] * gn[1] * gn[2];
  std::vector<std::pair<uint64_t, int32_t>> key((size_t)nc);
  for (int64_t l = 0; l < nc; ++l) {
    uint64_t cx = (uint64_t)(l % gn[0]), cy = (uint64_t)((l / gn[0]) % gn[1]), cz = (uint64_t)(l / ((int64_t)gn[0] * gn[1]));
    uint64_t code = spread_bits(cx, dims);
    if (dims > 1) code |= spread_bits(cy, dims) << 1;
    if (dims > 2) code |= spread_bits(cz, dims) << 2;
    key[(size_t)l] = {code, (int32_t)l};
  }
  std::sort(key.begin(), key.end());
  const size_t base = out.size();
  out.resize(base + (size_t)nc);
  for (int64_t r = 0; r < nc; ++r) out[base + (size_t)key[(size_t)r].second] = (int32_t)r;
}

// Agents per halo cell a sub-brick tile is sized for.
constexpr double SUB_BRICK_DENSITY = 48.0;

// The brick table for edges bd: per simulation, its bricks in Morton order
// of their coordinates, BRICK_WORDS each (simulation, origin cell, bucket
// range of its cells).
static void brick_table(cbm_ctx* c, const int* bd, std::vector<int32_t>& tab) {
  const int D = c->dims;
  for (int s = 0; s < c->S; ++s) {
    const SimDev& sd = c->h_sims[s];
    int64_t nb[3];
    for (int a = 0; a < 3; ++a) nb[a] = (sd.gn[a] + bd[a] - 1) / bd[a];
    std::vector<std::pair<uint64_t, int64_t>> key;
    key.reserve((size_t)(nb[0] * nb[1] * nb[2]));
    for (int64_t z = 0; z < nb[2]; ++z)
      for (int64_t y = 0; y < nb[1]; ++y)
        for (int64_t x = 0; x < nb[0]; ++x) {
          uint64_t code = spread_bits((uint64_t)x, D);
          if (D > 1) code |= spread_bits((uint64_t)y, D) << 1;
          if (D > 2) code |= spread_bits((uint64_t)z, D) << 2;
          key.push_back({code, x + nb[0] * (y + nb[1] * z)});
        }
    std::sort(key.begin(), key.end());
    for (const auto& k : key) {
      const int64_t x = k.second % nb[0], y = (k.second / nb[0]) % nb[1], z = k.second / (nb[0] * nb[1]);
      // the brick's cells have consecutive Morton ranks: its buckets are
      // [cell_off + lowest rank, + number of cells)
      const int32_t o[3] = {(int32_t)(x * bd[0]), (int32_t)(y * bd[1]), (int32_t)(z * bd[2])};
      int64_t lo = std::numeric_limits<int64_t>::max(), cnt = 0;
      for (int32_t cz = o[2]; cz < std::min(o[2] + bd[2], sd.gn[2]); ++cz)
        for (int32_t cy = o[1]; cy < std::min(o[1] + bd[1], sd.gn[1]); ++cy)
          for (int32_t cx = o[0]; cx < std::min(o[0] + bd[0], sd.gn[0]); ++cx) {
            const int64_t l = cx + (int64_t)sd.gn[0] * (cy + (int64_t)sd.gn[1] * cz);
            lo = std::min<int64_t>(lo, c->h_crank[(size_t)(sd.rank_off + l)]);
            ++cnt;
          }
      tab.push_back(s);
      for (int a = 0; a < 3; ++a) tab.push_back(o[a]);
      tab.push_back((int32_t)(sd.cell_off + lo));
      tab.push_back((int32_t)(sd.cell_off + lo + cnt));
    }
  }
}

// Bricks of the cell-linked stage: 2^m cells, Morton-aligned (x takes the
// extra bits first, so a brick's cells have consecutive Morton ranks and
// its rows are one range).  m is the largest size whose expected population
// (the densest simulation's mean density; capacity when rules change the
// population) stays within ~1.25 blocks of 256 threads, lowered while the
// LDS for 1.2x the expected halo population would exceed 64 KiB (a denser
// brick is left to the row-parallel stage kernel).  Bricks
// are listed per simulation in Morton order of their coordinates.
static int setup_bricks(cbm_ctx* c) {
  const int D = c->dims;
  const bool soft = c->inertial;
  double rho = 0.0;
  for (int s = 0; s < c->S; ++s) {
    const double n = (double)std::max<int64_t>(c->rules ? c->sims[s].n_max : c->sims[s].n, 1);
    rho = std::max(rho, n / (double)c->h_sims[s].ncells);
  }
  const int mmin = D, mmax = D == 3 ? 9 : 10;
  int m = mmin;
  while (m < mmax && rho * (double)(1 << (m + 1)) <= 320.0) ++m;
  const int64_t budget = 64 * 1024;
  // LDS bytes per staged agent and in all of a tile: the Verlet list build
  // stages positions and slots, the cell-linked stage positions (and radius
  // and fate for soft spheres); both add an owner and a home-row index
  const bool build = c->verlet || c->clv;   // the bricks serve the list build
  const int64_t per = build ? 8 * D + 4 + 4 : 8 * (D + (soft ? 1 : 0)) + (soft ? 4 : 0) + 4;
  auto tile_bytes = [&](int cp, int H) {
    return build ? vbrick_lds_bytes(D, cp, H) : brick_lds_bytes(D, soft, cp, H);
  };
  int bd[3] = {1, 1, 1};
  int cap = 0;
  auto edges = [&](int mm, int* e) {
    e[0] = 1 << ((mm + D - 1) / D);
    e[1] = D > 1 ? 1 << ((mm + D - 2) / D) : 1;
    e[2] = D > 2 ? 1 << ((mm + D - 3) / D) : 1;
  };
  for (;; --m) {
    edges(m, bd);
    const int H = brick_halo_cells(D, bd[0], bd[1], bd[2]);
    const int64_t fixed = tile_bytes(0, H) + 16;
    const int64_t capmax = (budget - fixed) / per;
    const double want = 1.2 * rho * H + 32.0;
    if (want <= (double)capmax || m == mmin) {
      cap = (int)std::max<int64_t>(1, std::min<int64_t>(capmax, (int64_t)std::max(256.0, want)));
      break;
    }
  }
  for (int a = 0; a < 3; ++a) c->bdim[a] = bd[a];
  c->brick_cap = cap;
  c->brick_lds = (size_t)tile_bytes(cap, brick_halo_cells(D, bd[0], bd[1], bd[2]));
  std::vector<int32_t> tab;
  brick_table(c, bd, tab);
  c->nbricks = (int64_t)tab.size() / BRICK_WORDS;
  if (c->nbricks > ((int64_t)1 << 31) - 1) return fail(c, CBM_ERR_CONFIG, "too many cell bricks; raise cellEdge");
  HIPCHK(c, dalloc(&c->bricks, (int64_t)tab.size()));
  HIPCHK(c, hipMemcpy(c->bricks, tab.data(), sizeof(int32_t) * tab.size(), hipMemcpyHostToDevice));

  // Sub-bricks for a cell-linked stage whose population can concentrate
  // (rules add agents where they are): the largest size below m, and at most
  // 4^D cells, whose tile holds SUB_BRICK_DENSITY agents per cell of its halo.  Their edges
  // divide the main edges and both are Morton-aligned, so each lies inside
  // one main brick.
  if (!build && c->rules && m > mmin) {
    for (int m2 = std::min(m - 1, 2 * D); m2 >= mmin; --m2) {
      int b2[3];
      edges(m2, b2);
      const int H2 = brick_halo_cells(D, b2[0], b2[1], b2[2]);
      const int64_t capmax = (budget - tile_bytes(0, H2) - 16) / per;
      const double want = 1.2 * SUB_BRICK_DENSITY * H2 + 32.0;
      if (want > (double)capmax) continue;
      for (int a = 0; a < 3; ++a) c->bdim2[a] = b2[a];
      c->brick_cap2 = (int)std::min<int64_t>(capmax, (int64_t)std::max(256.0, want));
      c->brick_lds2 = (size_t)tile_bytes(c->brick_cap2, H2);
      std::vector<int32_t> tab2;
      brick_table(c, b2, tab2);
      c->nbricks2 = (int64_t)tab2.size() / BRICK_WORDS;
      if (c->nbricks2 > ((int64_t)1 << 31) - 1) return fail(c, CBM_ERR_CONFIG, "too many cell bricks; raise cellEdge");
      HIPCHK(c, dalloc(&c->bricks2, (int64_t)tab2.size()));
      HIPCHK(c, hipMemcpy(c->bricks2, tab2.data(), sizeof(int32_t) * tab2.size(), hipMemcpyHostToDevice));
      break;
    }
  }
  return CBM_OK;
}

// Reach of a built-in law from a simulation's parameters: no pair at
// d >= reach interacts (the PairLaw::reach of the kernels).
static double law_reach(int law, const double* p) {
  switch (law) {
    case CBM_LAW_LINEAR: return p[1];
    case CBM_LAW_REP_ADH: return p[3] * p[1];
    case CBM_LAW_AGGREGATION: return std::max(p[0], p[2]);
    default: return std::numeric_limits<double>::infinity();
  }
}

// CellLinked on a skin list: for a built-in law with a reach, Euler or Heun,
// no rules, and grids the packed cells can hold.  The list radius is the
// largest reach plus a skin of reach/8, rebuilt when an agent has moved
// half the skin; the list grid has cells of at least the list radius.
static void choose_clv(cbm_ctx* c) {
  const cbm_model_desc& m = c->m;
  c->clv = c->cl && m.walk != 1 && !c->rules && !c->inertial &&
           (m.law == CBM_LAW_LINEAR || m.law == CBM_LAW_REP_ADH || m.law == CBM_LAW_AGGREGATION) &&
           (m.integrator == CBM_EULER || m.integrator == CBM_HEUN);
  if (!c->clv) return;
  // the order kernel packs a row into 23 bits and addresses the lists by
  // 32-bit byte offsets: at most 2^23 rows
  int64_t rows = 0;
  for (int s = 0; s < c->S; ++s) rows += c->sims[s].n_max;
  if (rows > ((int64_t)1 << 23)) { c->clv = false; return; }
  double reach = 0.0;
  for (int s = 0; s < c->S; ++s) reach = std::max(reach, law_reach(m.law, c->sims[s].p));
  const double rl = (reach + 0.125 * reach) * (1.0 + 1e-9);
  bool ok = reach > 0.0 && std::isfinite(reach);
  for (int s = 0; s < c->S && ok; ++s)
    for (int a = 0; a < m.dims && ok; ++a) {
      const double L = c->sims[s].box_hi[a] - c->sims[s].box_lo[a];
      const double n = m.periodic ? std::floor(L / m.cell_edge) : std::ceil(L / m.cell_edge + 2.0);
      ok = n <= 1023.0 && (!m.periodic || std::floor(L / rl) >= 3.0);
    }
  if (!ok) { c->clv = false; return; }
  c->clv_reach = reach;
  c->clv_skin = 0.125 * reach;
  c->list_radius = rl;
  c->list_r2 = sq_threshold(rl);
  c->thr = 0.5 * c->clv_skin;
  c->nb_kind = NB_LIST;
  // list capacity from the densest simulation's expected list length
  double lam = 0.0;
  for (int s = 0; s < c->S; ++s) {
    double vol = 1.0;
    for (int a = 0; a < m.dims; ++a) vol *= c->sims[s].box_hi[a] - c->sims[s].box_lo[a];
    const double ball = m.dims == 1 ? 2.0 * rl : (m.dims == 2 ? M_PI * rl * rl : 4.0 / 3.0 * M_PI * rl * rl * rl);
    lam = std::max(lam, (double)c->sims[s].n_max / vol * ball);
  }
  const int want = (int)std::ceil(lam + 6.0 * std::sqrt(lam) + 8.0);
  c->cap = std::min(FLAT_MAX_CAP, std::max(16, (want + 7) / 8 * 8));
}

static int setup_geometry(cbm_ctx* c) {
  const cbm_model_desc& m = c->m;
  c->h_sims.resize(c->S);
  int64_t row = 0, cell = 0;
  double edge = m.cell_edge;
  if (c->verlet) edge = std::max(m.cell_edge, m.skin);
  if (c->clv) edge = c->list_radius;
  for (int s = 0; s < c->S; ++s) {
    const cbm_sim_desc& in = c->sims[s];
    SimDev& sd = c->h_sims[s];
    std::memset(&sd, 0, sizeof(sd));
    for (int a = 0; a < 3; ++a) {
      sd.lo[a] = in.box_lo[a]; sd.hi[a] = in.box_hi[a];
      sd.L[a] = in.box_hi[a] - in.box_lo[a];
      sd.gn[a] = 1; sd.gedge[a] = 1.0;
    }
    for (int q = 0; q < 8; ++q) sd.p[q] = in.p[q];
    for (int a = 0; a < 3; ++a) { sd.cgn[a] = 1; sd.cedge[a] = 1.0; }
    if (c->clv)
      for (int a = 0; a < m.dims; ++a) {   // the model's own grid (src/neighbors.jl:604), for the cell order
        const double L = sd.L[a];
        if (m.periodic) {
          const int64_t n = (int64_t)std::floor(L / m.cell_edge);
          sd.cgn[a] = (int32_t)n;
          sd.cedge[a] = L / (double)n;
        } else {
          sd.cgn[a] = (int32_t)std::ceil(L / m.cell_edge + 2.0);
          sd.cedge[a] = m.cell_edge;
        }
      }
    if (!c->full) {
      for (int a = 0; a < m.dims; ++a) {
        double L = sd.L[a];
        if (m.periodic) {
          int64_t n = (int64_t)std::floor(L / edge);
          if (n < 3) return fail(c, CBM_ERR_CONFIG, "periodic cell grid needs >= 3 cells per axis (box too small for the cell edge)");
          sd.gn[a] = (int32_t)n;
          sd.gedge[a] = L / (double)n;
        } else {
          double n = std::ceil(L / edge + 2.0);   // src/neighbors.jl:604
          if (n > 2.0e9) return fail(c, CBM_ERR_CONFIG, "cell grid too large");
          sd.gn[a] = (int32_t)n;
          sd.gedge[a] = edge;
        }
      }
    }
    int64_t nc = (int64_t)sd.gn[0] * sd.gn[1] * sd.gn[2];
    if (nc > (int64_t)1 << 30) return fail(c, CBM_ERR_CONFIG, "cell grid too large; raise cellEdge");
    sd.row_off = row;
    sd.n_max = in.n_max;
    sd.cell_off = cell;
    sd.ncells = nc;
    sd.rank_off = (int64_t)c->h_crank.size();
    morton_ranks(sd.gn, m.dims, c->h_crank);
    row += in.n_max;
    cell += nc + 2;
  }
  if (row > ((int64_t)1 << 31) - 1) return fail(c, CBM_ERR_CONFIG, "total capacity exceeds 2^31 agents");
  if (cell > ((int64_t)1 << 31) - 1) return fail(c, CBM_ERR_CONFIG, "total cell count exceeds 2^31");
  c->C = row;
  c->NB = cell;
  c->NR = (int64_t)c->h_crank.size();
  return CBM_OK;
}

static int setup_blocks(cbm_ctx* c) {
  c->h_block_sim.clear();
  c->h_block_off.assign(c->S, 0);
  for (int s = 0; s < c->S; ++s) {
    c->h_block_off[s] = (int32_t)c->h_block_sim.size();
    int64_t nb = (c->sims[s].n_max + c->block - 1) / c->block;
    for (int64_t b = 0; b < nb; ++b) c->h_block_sim.push_back(s);
  }
  c->nblocks = (int)c->h_block_sim.size();
  if (c->d_block_sim) { (void)hipFree(c->d_block_sim); c->d_block_sim = nullptr; }
  if (c->d_block_off) { (void)hipFree(c->d_block_off); c->d_block_off = nullptr; }
  HIPCHK(c, dalloc(&c->d_block_sim, c->nblocks));
  HIPCHK(c, dalloc(&c->d_block_off, c->S));
  HIPCHK(c, hipMemcpy(c->d_block_sim, c->h_block_sim.data(), sizeof(int32_t) * c->nblocks, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_block_off, c->h_block_off.data(), sizeof(int32_t) * c->S, hipMemcpyHostToDevice));
  return CBM_OK;
}

static int statecol_k(const cbm_ctx* c, int v);

static int alloc_column(cbm_ctx* c, int k, int bytes) {
  c->has_col[k] = true;
  c->col[k].bytes = bytes;
  for (int b = 0; b < 2; ++b) {
    if (bytes == 8) { double* p; HIPCHK(c, dalloc(&p, c->C)); c->col[k].buf[b] = p; }
    else { int32_t* p; HIPCHK(c, dalloc(&p, c->C)); c->col[k].buf[b] = p; }
  }
  return CBM_OK;
}

static void destroy_ctx(cbm_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  for (int k = 0; k < C_NCOLS; ++k) { F(c->col[k].buf[0]); F(c->col[k].buf[1]); }
  for (int a = 0; a < 3; ++a) { F(c->Fnew[a]); F(c->xref[a]); }
  for (int v = 0; v < MAX_STATE; ++v) { F(c->h[0][v]); F(c->h[1][v]); F(c->kacc[v]); F(c->gacc[v]); }
  F(c->gr_new); F(c->vb_new); F(c->dead);
  F(c->n_add_slot); F(c->dead_slot); F(c->off_add); F(c->off_dead); F(c->holes);
  F(c->med_c[0]); F(c->med_c[1]); F(c->med_count);
  for (int k = 0; k < 4 + USER_PARAMS; ++k) F(c->rule_nb[k]);
  F(c->med_init); F(c->med_tmp); F(c->adi_tab[0]); F(c->adi_tab[1]); F(c->adi_tab[2]);
  F(c->key); F(c->lin); F(c->gbucket); F(c->cell_range); F(c->count); F(c->cursor); F(c->start); F(c->tmp); F(c->cell_row);
  F(c->emax); F(c->wlist); F(c->wlist_n); F(c->fix_rows); F(c->fix_n);
  F(c->scan_status); F(c->scan_ticket); F(c->list); F(c->list_n); F(c->bricks); F(c->bricks2); F(c->brick_skip); F(c->blk_flag); F(c->flags); F(c->counters);
  for (cbm_ctx::DevBuf* b : {&c->adj_tgt, &c->adj_tbase, &c->adj_invn, &c->adj_ck, &c->adj_seg, &c->adj_lam,
                             &c->adj_gpart, &c->adj_lpart}) F(b->p);
  for (int32_t* p : c->snap_l) F(p);
  for (int32_t* p : c->snap_n) F(p);
  for (void* p : c->snap_buf) F(p);
  F(c->snap_cnt); F(c->snap_med);
  F(c->d_staging); F(c->d_block_sim); F(c->d_block_off); F(c->d_sims); F(c->d_st);
  for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  F(c->efin);
  F(c->med_add);
  if (c->rec_host) (void)hipHostFree(c->rec_host);
  if (c->user_mod) (void)hipModuleUnload(c->user_mod);
  if (c->model_mod) (void)hipModuleUnload(c->model_mod);
  if (c->model_t) (void)hipFree(c->model_t);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static int init_ctx(cbm_ctx* c) {
  const cbm_model_desc& m = c->m;
  int rc = validate(c);
  if (rc) return rc;
  c->dims = m.dims;
  c->growth = m.rule == CBM_RULE_GROWTH_DIVISION || m.rule == CBM_RULE_CUSTOM;
  c->rules = m.rule != CBM_RULE_NONE;
  c->inertial = m.law == CBM_LAW_SOFT_SPHERES;
  c->rods = m.law == CBM_LAW_RODS;
  c->nvar = c->rods ? (m.rule == CBM_RULE_RODS_GROWTH ? 7 : 6)
                    : (c->inertial ? 2 * c->dims : c->dims + (c->growth ? 1 : 0) + c->m.user_de);
  c->lagged = m.lagged != 0;
  c->full = m.neighbors == CBM_FULL;
  c->cl = m.neighbors == CBM_CELL_LINKED;
  c->verlet = !c->full && !c->cl;
  c->vtime = m.neighbors == CBM_VERLET_TIME;
  c->nb_kind = c->full ? NB_FULL : (c->cl ? NB_CELL : NB_LIST);
  c->list_radius = m.skin;
  c->list_r2 = sq_threshold(m.skin);
  c->thr = m.rebuild_disp > 0 ? m.rebuild_disp : m.skin / 2;   // src/neighbors.jl:403
  choose_clv(c);
  HIPCHK(c, hipGetDevice(&c->device));
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(c, hipEventCreate(&c->ev0));
  HIPCHK(c, hipEventCreate(&c->ev1));
  if ((rc = setup_geometry(c))) return rc;
  if ((rc = setup_blocks(c))) return rc;
  HIPCHK(c, dalloc(&c->d_sims, c->S));
  HIPCHK(c, hipMemcpy(c->d_sims, c->h_sims.data(), sizeof(SimDev) * c->S, hipMemcpyHostToDevice));
  std::vector<SimState> st(c->S);
  for (int s = 0; s < c->S; ++s) {
    std::memset(&st[s], 0, sizeof(SimState));
    st[s].N = c->sims[s].n; st[s].n_tot = c->sims[s].n; st[s].id_max = c->sims[s].n;
  }
  HIPCHK(c, dalloc(&c->d_st, c->S));
  HIPCHK(c, hipMemcpy(c->d_st, st.data(), sizeof(SimState) * c->S, hipMemcpyHostToDevice));

  // columns
  for (int a = 0; a < c->dims; ++a) if ((rc = alloc_column(c, C_X + a, 8))) return rc;
  if ((rc = alloc_column(c, C_ID, 8))) return rc;
  if ((rc = alloc_column(c, C_SLOT, 4))) return rc;
  if (c->lagged) {
    for (int a = 0; a < c->dims; ++a) { if ((rc = alloc_column(c, C_FX + a, 8))) return rc; HIPCHK(c, dalloc(&c->Fnew[a], c->C)); }
  }
  if (c->growth) {
    if ((rc = alloc_column(c, C_VOL, 8))) return rc;
    if ((rc = alloc_column(c, C_GR, 8))) return rc;
    if ((rc = alloc_column(c, C_VB, 8))) return rc;
    HIPCHK(c, dalloc(&c->gr_new, c->C));
    HIPCHK(c, dalloc(&c->vb_new, c->C));
  }
  if (c->inertial) {
    for (int a = 0; a < c->dims; ++a) if ((rc = alloc_column(c, C_VX + a, 8))) return rc;
    if ((rc = alloc_column(c, C_RAD, 8))) return rc;
    if ((rc = alloc_column(c, C_MASS, 8))) return rc;
    if ((rc = alloc_column(c, C_FATE, 8))) return rc;
    std::vector<double> one((size_t)c->C, 1.0);
    std::vector<int64_t> fate1((size_t)c->C, 1);
    for (int b = 0; b < 2; ++b) {
      HIPCHK(c, hipMemcpy(c->col[C_RAD].buf[b], one.data(), sizeof(double) * c->C, hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->col[C_MASS].buf[b], one.data(), sizeof(double) * c->C, hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->col[C_FATE].buf[b], fate1.data(), sizeof(int64_t) * c->C, hipMemcpyHostToDevice));
    }
  }
  if (c->rods) {
    // DE velocities, angle and rate (0 until set); diameter, length and
    // mass 1 until set; the last stage's fx, fy, W and pressure
    const int cols[] = {C_VX, C_VY, C_TH, C_OM, C_DIAM, C_LEN, C_MASS, C_FX, C_FY, C_TQ, C_PR};
    for (int k : cols) if ((rc = alloc_column(c, k, 8))) return rc;
    if (m.rule == CBM_RULE_RODS_GROWTH && (rc = alloc_column(c, C_LT, 8))) return rc;
    std::vector<double> one((size_t)c->C, 1.0);
    for (int k : {C_DIAM, C_LEN, C_MASS})
      for (int b = 0; b < 2; ++b)
        HIPCHK(c, hipMemcpy(c->col[k].buf[b], one.data(), sizeof(double) * c->C, hipMemcpyHostToDevice));
  }
  // user agent parameters: read by a rule body, or DE variables of the law body
  if (m.rule == CBM_RULE_CUSTOM || m.user_de > 0)
    for (int k = 0; k < USER_PARAMS; ++k) if ((rc = alloc_column(c, C_U0 + k, 8))) return rc;
  // every DE variable's column exists before any stage addresses it
  for (int v = 0; v < c->nvar; ++v)
    if (!c->has_col[statecol_k(c, v)]) return fail(c, CBM_ERR_CONFIG, "DE variable without a column");
  if (c->clv && (rc = alloc_column(c, C_CCELL, 4))) return rc;
  if (c->rules) {
    if ((rc = alloc_column(c, C_LF, 8))) return rc;
    const int64_t nm = c->sims[0].n_max;
    HIPCHK(c, dalloc(&c->dead, c->C));
    HIPCHK(c, dalloc(&c->n_add_slot, nm + 1));
    HIPCHK(c, dalloc(&c->dead_slot, nm + 1));
    HIPCHK(c, dalloc(&c->off_add, nm + 1));
    HIPCHK(c, dalloc(&c->off_dead, nm + 1));
    HIPCHK(c, dalloc(&c->holes, nm + 1));
  }
  if (m.medium_n[0] > 0) {
    c->medium = true;
    c->med_nodes = 1;
    for (int a = 0; a < c->dims; ++a) c->med_nodes *= m.medium_n[a];
    HIPCHK(c, dalloc(&c->med_c[0], c->med_nodes));
    HIPCHK(c, dalloc(&c->med_c[1], c->med_nodes));
    HIPCHK(c, dalloc(&c->med_count, c->med_nodes));
    std::vector<double> c0((size_t)c->med_nodes, m.medium_c0);
    HIPCHK(c, hipMemcpy(c->med_c[0], c0.data(), sizeof(double) * c->med_nodes, hipMemcpyHostToDevice));
    if (m.medium_alg == CBM_MEDIUM_DGADI) {
      c->adi = true;
      HIPCHK(c, dalloc(&c->med_init, c->med_nodes));
      HIPCHK(c, dalloc(&c->med_tmp, c->med_nodes));
      HIPCHK(c, hipMemcpy(c->med_init, c0.data(), sizeof(double) * c->med_nodes, hipMemcpyHostToDevice));
      // diagonals -D r, 1 + 2 D r with r = dt / (dims dx^2) of the last axis
      // for every sweep (integrators.jl:850, 875-877 and the sweeps' column dims)
      const int la = c->dims - 1;
      const double dxl = (c->sims[0].box_hi[la] - c->sims[0].box_lo[la]) / (double)m.medium_n[la];
      const double r = m.dt / ((double)c->dims * (dxl * dxl));
      c->adi_sub = -m.medium_D * r;
      c->adi_main = 1 + 2 * m.medium_D * r;
      c->adi_sup = -m.medium_D * r;
      for (int a = 0; a < c->dims; ++a) {
        const int n = m.medium_n[a];
        std::vector<double> tab(2 * (size_t)n, 0.0);   // c*[i] at [i], m[i] at [n + i]
        tab[1] = c->adi_sup / c->adi_main;
        for (int i = 2; i <= n - 3; ++i) {
          const double mi = 1.0 / (c->adi_main - c->adi_sub * tab[i - 1]);
          tab[n + i] = mi;
          tab[i] = c->adi_sup * mi;
        }
        tab[n + n - 2] = 1.0 / (c->adi_main - c->adi_sub * tab[n - 3]);
        HIPCHK(c, dalloc(&c->adi_tab[a], 2 * (int64_t)n));
        HIPCHK(c, hipMemcpy(c->adi_tab[a], tab.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
      }
    }
  }
  // slot = local row, id = local row + 1 (id = 1:N, BASEPARAMETERS constants.jl:21)
  {
    std::vector<int32_t> slot(c->C);
    std::vector<int64_t> id(c->C);
    for (int s = 0; s < c->S; ++s)
      for (int64_t r = 0; r < c->sims[s].n_max; ++r) {
        slot[c->h_sims[s].row_off + r] = (int32_t)r;
        id[c->h_sims[s].row_off + r] = r + 1;
      }
    for (int b = 0; b < 2; ++b) {
      HIPCHK(c, hipMemcpy(c->col[C_SLOT].buf[b], slot.data(), sizeof(int32_t) * c->C, hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->col[C_ID].buf[b], id.data(), sizeof(int64_t) * c->C, hipMemcpyHostToDevice));
    }
  }
  // stage buffers
  int nst = (m.integrator == CBM_HEUN || m.integrator == CBM_EULER_HEUN) ? 1 : (m.integrator == CBM_RK4 ? 2 : 0);
  for (int b = 0; b < nst; ++b)
    for (int v = 0; v < c->nvar; ++v) HIPCHK(c, dalloc(&c->h[b][v], c->C));
  if (m.integrator == CBM_RK4 || m.integrator == CBM_EULER_HEUN)
    for (int v = 0; v < c->nvar; ++v) HIPCHK(c, dalloc(&c->kacc[v], c->C));

  // neighbor engine
  HIPCHK(c, dalloc(&c->key, c->C));
  HIPCHK(c, dalloc(&c->lin, c->C));
  {
    std::vector<int32_t> gb((size_t)c->NR);
    for (int s = 0; s < c->S; ++s) {
      const SimDev& sd = c->h_sims[s];
      for (int64_t l = 0; l < sd.ncells; ++l)
        gb[(size_t)(sd.rank_off + l)] = (int32_t)(sd.cell_off + c->h_crank[(size_t)(sd.rank_off + l)]);
    }
    HIPCHK(c, dalloc(&c->gbucket, c->NR));
    HIPCHK(c, hipMemcpy(c->gbucket, gb.data(), sizeof(int32_t) * c->NR, hipMemcpyHostToDevice));
    HIPCHK(c, dalloc(&c->cell_range, c->NR));
  }
  if (c->cl || c->verlet) {
    if ((rc = setup_bricks(c))) return rc;
    HIPCHK(c, dalloc(&c->brick_skip, c->C));
    HIPCHK(c, dalloc(&c->blk_flag, c->C / 64 + 2 * (int64_t)c->S + 1));   // row blocks at the smallest block size
  }
  c->h_crank.clear();
  c->h_crank.shrink_to_fit();
  HIPCHK(c, dalloc(&c->tmp, c->C));
  HIPCHK(c, dalloc(&c->cell_row, c->C));
  HIPCHK(c, dalloc(&c->count, c->NB));
  HIPCHK(c, dalloc(&c->cursor, c->NB));
  HIPCHK(c, dalloc(&c->start, c->NB + 1));
  {
    // the largest scan: the bucket table, or the rule decisions by slot
    int64_t longest = c->NB;
    for (int s = 0; s < c->S; ++s) longest = std::max<int64_t>(longest, c->sims[s].n_max + 1);
    c->scan_tiles = scan_blocks(longest);
    HIPCHK(c, dalloc(&c->scan_status, c->scan_tiles + 1));
  }
  HIPCHK(c, dalloc(&c->scan_ticket, 4));
  {
    unsigned int ep = 1;
    HIPCHK(c, hipMemcpy(c->scan_ticket + 2, &ep, sizeof(ep), hipMemcpyHostToDevice));
  }
  HIPCHK(c, dalloc(&c->flags, 8));
  HIPCHK(c, dalloc(&c->counters, 4));
  // cell-linked Heun / EulerHeun with a built-in law: stage 1 writes the
  // candidate lists that stage 2 walks (flags[4] says whether they are exact)
  c->wl = c->cl && !c->clv && (m.integrator == CBM_HEUN || m.integrator == CBM_EULER_HEUN) && !c->inertial &&
          (m.law == CBM_LAW_LINEAR || m.law == CBM_LAW_REP_ADH || m.law == CBM_LAW_AGGREGATION);
  if (c->wl) {
    c->wcap = 32;
    HIPCHK(c, dalloc(&c->list, (int64_t)c->wcap * c->C));
    HIPCHK(c, dalloc(&c->list_n, c->C));
  }
  if (c->verlet || c->clv) {
    if (c->verlet) c->cap = m.max_neighbors;
    for (int a = 0; a < c->dims; ++a) HIPCHK(c, dalloc(&c->xref[a], c->C));
    HIPCHK(c, dalloc(&c->list, (int64_t)c->cap * c->C));
    HIPCHK(c, dalloc(&c->list_n, c->C));
  }
  if (c->clv) {
    HIPCHK(c, dalloc(&c->emax, 2));   // the stage-1 maximum of |h - x_ref|^2, by step parity
    HIPCHK(c, hipMemset(c->emax, 0, 2 * sizeof(unsigned long long)));
    HIPCHK(c, dalloc(&c->wlist, (int64_t)c->cap * c->C));
    HIPCHK(c, dalloc(&c->wlist_n, c->C));
    HIPCHK(c, dalloc(&c->fix_rows, c->C));
    HIPCHK(c, dalloc(&c->fix_n, 2));
    HIPCHK(c, hipMemset(c->fix_n, 0, 2 * sizeof(int32_t)));
    HIPCHK(c, dalloc(&c->efin, 2));
    HIPCHK(c, hipMemset(c->efin, 0, 2 * sizeof(unsigned long long)));
    HIPCHK(c, hipHostMalloc((void**)&c->rec_host, sizeof(ClvRecord), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(c->rec_host, 0, sizeof(ClvRecord));
    c->rec_host->step = -1;
    HIPCHK(c, hipHostGetDevicePointer((void**)&c->rec_dev, c->rec_host, 0));
  }
  int64_t maxn = 1;
  for (int s = 0; s < c->S; ++s) maxn = std::max(maxn, c->sims[s].n_max);
  c->staging_bytes = maxn * 8;
  HIPCHK(c, hipMalloc(&c->d_staging, (size_t)c->staging_bytes));
  HIPCHK(c, hipDeviceSynchronize());
  return CBM_OK;
}

// ------------------------------------------------------------ step pieces
static PermCols perm_cols(cbm_ctx* c) {
  PermCols pc;
  std::memset(&pc, 0, sizeof(pc));
  for (int k = 0; k < C_NCOLS; ++k) {
    if (!c->has_col[k]) continue;
    pc.cur[pc.n] = c->col[k].cur();
    pc.alt[pc.n] = c->col[k].alt();
    pc.bytes[pc.n] = c->col[k].bytes;
    pc.n++;
  }
  return pc;
}

static RebuildArgs rebuild_args(cbm_ctx* c, const int32_t* flag, int32_t* flag_clear, bool tick) {
  RebuildArgs a;
  std::memset(&a, 0, sizeof(a));
  a.g = geom(c);
  a.flag = flag;
  a.flag_clear = flag_clear;
  a.rebuilds = c->counters;
  a.step_inc = tick ? c->counters + 1 : nullptr;
  a.gate_set = c->wl ? c->flags + 4 : nullptr;
  for (int d = 0; d < c->dims; ++d) a.pos[d] = (const double*)c->col[C_X + d].cur();
  a.dead = c->dead;
  a.slot = (const int32_t*)c->col[C_SLOT].cur();
  a.range = c->cell_range;
  a.nr = c->NR;
  a.key = c->key; a.lin = c->lin; a.gbucket = c->gbucket; a.count = c->count; a.cursor = c->cursor; a.start = c->start; a.tmp = c->tmp;
  a.cell_row = c->cell_row;
  a.dims = c->dims; a.periodic = c->m.periodic; a.full = c->full ? 1 : 0;
  a.geometric = (c->verlet || c->clv) ? 1 : 0;
  // the skin-list stages clear their own words; other modes keep none
  a.emax_clear = c->clv ? nullptr : c->emax;
  a.fix_clear = c->clv ? nullptr : c->fix_n;
  a.efin_clear = c->efin;
  return a;
}

// Sort agents into buckets and permute every column.  `flag` NULL runs
// unconditionally; otherwise every kernel exits early unless *flag != 0.
// copy_back: Verlet mode (build lists, copy the sorted rows back into the
// current buffers); otherwise the alternate buffers become current.
// Exclusive scan of n int32 counts (see k_scan); refuses a scan with more
// tiles than the status array holds.
static int run_scan(cbm_ctx* c, const int32_t* flag, int32_t* count, int32_t* cursor, int32_t* start, int64_t n) {
  if (scan_blocks(n) > c->scan_tiles)
    return fail(c, CBM_ERR_ARG, "scan longer than its status array");
  ScanArgs s;
  std::memset(&s, 0, sizeof(s));
  s.flag = flag; s.count = count; s.cursor = cursor; s.start = start; s.n = n;
  s.status = c->scan_status; s.ticket = c->scan_ticket; s.done = c->scan_ticket + 1;
  s.epoch = c->scan_ticket + 2; s.nblocks = scan_blocks(n);
  HIPCHK(c, launch_scan(c->stream, s));
  return CBM_OK;
}

static BrickSrc brick_src(cbm_ctx* c) {
  BrickSrc b;
  std::memset(&b, 0, sizeof(b));
  b.sims = c->d_sims;
  b.block_off = c->d_block_off;
  b.brick = c->bricks;
  b.bucket_start = c->start;
  b.cell_range = c->cell_range;
  for (int a = 0; a < 3; ++a) b.bdim[a] = c->bdim[a];
  b.cap = c->brick_cap;
  b.brick_skip = c->brick_skip;
  b.blk_flag = c->blk_flag;
  b.brick_over = (unsigned long long*)(c->counters + 2);
  return b;
}

static int rebuild_chain(cbm_ctx* c, const int32_t* flag, int32_t* flag_clear, bool tick, bool copy_back) {
  RebuildArgs a = rebuild_args(c, flag, flag_clear, tick);
  HIPCHK(c, launch_keys(c->nblocks, c->block, c->stream, a));
  int rc = run_scan(c, flag, c->count, c->cursor, c->start, c->NB);
  if (rc) return rc;
  HIPCHK(c, launch_scatter(c->nblocks, c->block, c->stream, a, !c->full));
  PermuteArgs pa;
  pa.rb = a;
  pa.cols = perm_cols(c);
  HIPCHK(c, launch_permute(c->nblocks, c->block, c->stream, pa));
  if (copy_back) {
    VerletArgs va;
    std::memset(&va, 0, sizeof(va));
    va.rb = a;
    va.cols = pa.cols;
    for (int d = 0; d < c->dims; ++d) { va.apos[d] = (const double*)c->col[C_X + d].alt(); va.xref[d] = c->xref[d]; }
    va.aslot = (const int32_t*)c->col[C_SLOT].alt();
    va.list = c->list; va.list_n = c->list_n; va.ldl = c->C; va.cap = c->cap;
    // a CellLinked skin list past its capacity is no error: its agent takes
    // the exact walk (flat_body), so the overflow word is one nobody reads
    va.overflow = c->flags + (c->clv ? 6 : 2);
    va.rl2 = c->list_r2;
    va.lds_r = verlet_lds_r(c);
    va.build = (c->verlet || c->clv) ? 1 : 0;
    if (va.build && c->bricks) {
      // brick tiles, then the rows of the bricks too dense for LDS
      va.bs = brick_src(c);
      HIPCHK(c, launch_verlet_brick(c->dims, c->m.periodic, c->nbricks, c->block, c->brick_lds, c->stream, va));
      va.fixup = 1;
    }
    HIPCHK(c, launch_verlet(c->dims, c->nblocks, c->block, c->stream, va));
  } else {
    for (int k = 0; k < C_NCOLS; ++k) if (c->has_col[k]) c->col[k].cidx ^= 1;
  }
  return CBM_OK;
}

// Forced (re)build before step c->step: computeNeighbors! of loadToPlatform!
static StageArgs stage_base(cbm_ctx* c);

static int force_rebuild(cbm_ctx* c) {
  if (c->verlet || c->clv) {
    int32_t* f = c->flags + ((c->step + 1) & 1);
    HIPCHK(c, launch_set_flag(c->stream, f, 1));
    int rc = rebuild_chain(c, f, c->flags + (c->step & 1), false, true);
    if (rc) return rc;
    if (c->vtime) c->t_next = c->t + c->m.dt_recompute;
    if (c->clv) {
      // the cells of the sorted rows; every final stage writes the next ones
      StageArgs a = stage_base(c);
      for (int d = 0; d < c->dims; ++d) a.S[d] = (const double*)c->col[C_X + d].cur();
      a.ccell_out = (int32_t*)c->col[C_CCELL].cur();
      HIPCHK(c, launch_clcell(c->dims, c->nblocks, c->block, c->stream, a));
      // records from this step on describe the new lists
      c->last_build_step = c->step - 1;
      c->rec_first = c->step;
    }
  } else {
    int rc = rebuild_chain(c, nullptr, nullptr, false, false);
    if (rc) return rc;
  }
  c->structure_valid = true;
  return CBM_OK;
}

static StageArgs stage_base(cbm_ctx* c) {
  StageArgs a;
  std::memset(&a, 0, sizeof(a));
  a.g = geom(c);
  if (c->lagged) for (int d = 0; d < c->dims; ++d) { a.F[d] = (const double*)c->col[C_FX + d].cur(); a.Fnew[d] = c->Fnew[d]; }
  for (int d = 0; d < c->dims; ++d) a.xref[d] = c->xref[d];
  a.id = (const int64_t*)c->col[C_ID].cur();
  a.list = c->list; a.list_n = c->list_n; a.ldl = c->C;
  a.cell_range = c->cell_range; a.cell_row = c->cell_row;
  a.brick = c->bricks;
  a.brick_skip = c->brick_skip;
  a.blk_flag = c->blk_flag;
  a.bucket_start = c->start;
  a.brick_over = (unsigned long long*)(c->counters + 2);
  for (int q = 0; q < 3; ++q) a.bdim[q] = c->bdim[q];
  a.brick_cap = c->brick_cap;
  a.brick2 = c->bricks2; a.nbricks2 = c->nbricks2; a.brick_lds2 = (int64_t)c->brick_lds2;
  for (int q = 0; q < 3; ++q) a.bdim2[q] = c->bdim2[q];
  a.brick_cap2 = c->brick_cap2;
  // flattened list stages: lists of at most FLAT_MAX_CAP entries
  {
    const int lcap = (c->verlet || c->clv) ? c->cap : (c->wl ? c->wcap : 0);
    const int chunk = c->m.walk >= 512 ? c->m.walk : 0;
    if (lcap > 0 && lcap <= FLAT_MAX_CAP && chunk > 0) { a.fchunk = chunk; a.fcap = lcap; }
  }
  if (c->clv) {
    a.ccell = (const int32_t*)c->col[C_CCELL].cur();
    a.cslot = (const int32_t*)c->col[C_SLOT].cur();
    a.clv_e1 = c->thr * (1.0 + 1e-6);
    a.clv_skin = c->clv_skin * (1.0 - 1e-9);
    a.clv_half = 0.5 * c->clv_skin * (1.0 + 1e-9);
    a.clv_half2 = 0.25 * c->clv_skin * c->clv_skin;
    a.fix_rows = c->fix_rows;
    a.fix_n = c->fix_n;
    a.nsims = c->S;
    a.lcap = c->cap;
    a.efin = c->efin;
    a.clv_ahead = CLV_AHEAD;
  }
  a.dt = c->m.dt; a.sqdt = std::sqrt(c->m.dt); a.drift = c->m.drift; a.noise = c->m.noise; a.thr2 = sq_threshold(c->thr);
  a.seed = c->m.seed;
  a.step_ptr = c->counters + 1;
  a.periodic = c->m.periodic;
  a.lagged = c->lagged ? 1 : 0;
  a.walls = (c->m.law == CBM_LAW_AGGREGATION) ? 1 : 0;
  a.nvar = c->nvar;
  a.write_fnew = c->lagged ? 1 : 0;
  if (c->growth) a.growth = (const double*)c->col[C_GR].cur();
  if (c->has_col[C_U0])
    for (int k = 0; k < USER_PARAMS; ++k) a.upar[k] = (const double*)c->col[C_U0 + k].cur();
  for (int v = 0; v < MAX_STATE; ++v) a.gacc[v] = c->gacc[v];
  if (c->inertial) {
    a.rad = (const double*)c->col[C_RAD].cur();
    a.mass = (const double*)c->col[C_MASS].cur();
    a.fate = (const int64_t*)c->col[C_FATE].cur();
  }
  if (c->rods) {
    a.mass = (const double*)c->col[C_MASS].cur();
    a.rod_d = (const double*)c->col[C_DIAM].cur();
    a.rod_l = (const double*)c->col[C_LEN].cur();
    const int outs[4] = {C_FX, C_FY, C_TQ, C_PR};
    for (int k = 0; k < 4; ++k) a.rod_out[k] = (double*)c->col[outs[k]].cur();
    a.rod_err = c->flags + 7;
    a.rod_growth = c->m.rule_p[0];
    a.rod_alpha = c->m.rule_p[3];
  }
  return a;
}

// Optional per-launch event pairs (cbm_profile): category = stage kind, or
// PROF_CHAIN for the neighbor update of one step.
constexpr int PROF_CHAIN = 8;
constexpr int PROF_STEP = 9;
constexpr int PROF_ORDER = 12;   // no launch of its own now: the first skin-list stage orders the lists
constexpr int PROF_NCAT = 13;
// profile category of a stage kind: 0..7 as the kind, EulerHeun stages 10, 11
static int prof_cat_of(int kind) { return kind == ST_EH1 ? 10 : (kind == ST_EH2 ? 11 : kind); }

static void prof_begin(cbm_ctx* c, int cat);
static void prof_end(cbm_ctx* c);

// One integrator-stage launch: the built-in law kernels, or the hipRTC
// module of a user law.
static hipError_t stage_launch(cbm_ctx* c, const StageArgs& a) {
  const int noise = (a.kind == ST_EM || a.kind == ST_EH1 || a.kind == ST_EH2) ? 1 : 0;
  const bool module = c->m.law == CBM_LAW_CUSTOM || (c->user_noise && noise);
  if (!module && a.gate) {
    // stage 2 of a listed cell-linked step: the list walk when stage 1's
    // lists are exact, else the brick tiles (each launch checks the gate)
    StageArgs b = a;
    b.gate_run = 1;
    hipError_t e = launch_stage(c->dims, NB_LIST, c->m.law, c->nblocks, c->block, c->stream, b, c->nbricks, c->brick_lds);
    if (e != hipSuccess) return e;
    b.gate_run = 0;
    return launch_stage(c->dims, NB_CELL, c->m.law, c->nblocks, c->block, c->stream, b, c->nbricks, c->brick_lds);
  }
  if (!module)
    return launch_stage(c->dims, c->nb_kind, c->m.law, c->nblocks, c->block, c->stream, a, c->nbricks, c->brick_lds);
  hipFunction_t f = c->user_fn[noise];
  if (!f) return hipErrorInvalidDeviceFunction;
  StageArgs copy = a;
  void* params[] = {(void*)&copy};
  if (c->user_nb == NB_CELL) {
    if (!c->user_brick[noise]) return hipErrorInvalidDeviceFunction;
    const hipError_t e = hipModuleLaunchKernel(c->user_brick[noise], (unsigned)c->nbricks, 1, 1, (unsigned)c->block,
                                               1, 1, (unsigned)c->brick_lds, c->stream, params, nullptr);
    if (e != hipSuccess) return e;
  }
  return hipModuleLaunchKernel(f, (unsigned)c->nblocks, 1, 1, (unsigned)c->block, 1, 1, 0, c->stream, params, nullptr);
}

// Cell-linked two-stage steps (c->wl): stage 1 writes the candidate lists
// and may clear flags[4]; stage 2 reads it to pick the list walk.
static void listed_stage(cbm_ctx* c, StageArgs& a, int stage) {
  a.wlist = nullptr; a.wlist_n = nullptr; a.lgate = nullptr; a.gate = nullptr; a.wcap = 0;
  if (!c->wl) return;
  if (stage == 1) {
    a.wlist = c->list; a.wlist_n = c->list_n; a.wcap = c->wcap; a.lgate = c->flags + 4;
  } else {
    a.gate = c->flags + 4;
  }
}

static int run_stage(cbm_ctx* c, StageArgs& a, int kind) {
  a.kind = kind;
  prof_begin(c, prof_cat_of(kind));
  HIPCHK(c, stage_launch(c, a));
  prof_end(c);
  c->last_launches++;
  return CBM_OK;
}

// Column of DE variable v: positions, then velocities (soft spheres) or vol.
static int statecol_k(const cbm_ctx* c, int v) {
  if (v < c->dims) return C_X + v;
  if (c->rods) return v < 4 ? C_VX + (v - 2) : (v < 6 ? C_TH + (v - 4) : C_LEN);
  if (c->inertial) return C_VX + (v - c->dims);
  return c->growth && v == c->dims ? C_VOL : C_U0 + (v - c->dims - (c->growth ? 1 : 0));
}

static double* statecol(cbm_ctx* c, int v, bool alt) {
  int k = statecol_k(c, v);
  return (double*)(alt ? c->col[k].alt() : c->col[k].cur());
}

static void flip_state(cbm_ctx* c) {
  for (int v = 0; v < c->nvar; ++v) c->col[statecol_k(c, v)].cidx ^= 1;
}

static int neighbor_update(cbm_ctx* c, int64_t s, bool disp);

static RuleArgs rule_args(cbm_ctx* c, int32_t* flag) {
  RuleArgs r;
  std::memset(&r, 0, sizeof(r));
  r.g = geom(c);
  r.rule = c->m.rule; r.dims = c->dims; r.lagged = c->lagged ? 1 : 0; r.growth = c->growth ? 1 : 0;
  r.periodic = c->m.periodic ? 1 : 0;
  for (int q = 0; q < 4; ++q) r.rp[q] = c->m.rule_p[q];
  r.dt = c->m.dt; r.seed = c->m.seed; r.step_ptr = c->counters + 1;
  if (c->rods && c->has_col[C_LT]) {
    const int rc[RC_N] = {C_VX, C_VY, C_TH, C_OM, C_LEN, C_DIAM, C_MASS, C_LT, C_FX, C_FY, C_TQ, C_PR};
    for (int j = 0; j < RC_N; ++j) {
      r.rcol[j] = (double*)c->col[rc[j]].cur();
      r.rcol_alt[j] = (double*)c->col[rc[j]].alt();
    }
  }
  for (int d = 0; d < c->dims; ++d) { r.pos[d] = (double*)c->col[C_X + d].cur(); r.pos_alt[d] = (double*)c->col[C_X + d].alt(); }
  if (c->growth) {
    r.vol = (double*)c->col[C_VOL].cur();
    r.vol_alt = (double*)c->col[C_VOL].alt();
    r.gr = (const double*)c->col[C_GR].cur(); r.gr_new = c->gr_new;
    r.vb = (const double*)c->col[C_VB].cur(); r.vb_new = c->vb_new;
  }
  if (c->lagged) for (int d = 0; d < c->dims; ++d) { r.F[d] = (double*)c->col[C_FX + d].cur(); r.Fnew[d] = c->Fnew[d]; }
  r.id = (int64_t*)c->col[C_ID].cur();
  r.lf = (int64_t*)c->col[C_LF].cur();
  if (c->has_col[C_U0])
    for (int k = 0; k < USER_PARAMS; ++k) {
      r.u[k] = (double*)c->col[C_U0 + k].cur();
      r.u_alt[k] = (double*)c->col[C_U0 + k].alt();
    }
  r.list = c->list; r.list_n = c->list_n; r.ldl = c->C;
  r.cell_row = c->cell_row; r.cell_range = c->cell_range;
  r.nb_kind = c->nb_kind;
  for (int d = 0; d < 3; ++d) r.nb_x[d] = c->rule_nb[d];
  r.nb_vol = c->rule_nb[3];
  for (int k = 0; k < USER_PARAMS; ++k) r.nb_u[k] = c->rule_nb[4 + k];
  r.slot = (int32_t*)c->col[C_SLOT].cur();
  r.dead = c->dead;
  r.n_add_slot = c->n_add_slot; r.dead_slot = c->dead_slot;
  r.off_add = c->off_add; r.off_dead = c->off_dead;
  r.holes = c->holes;
  r.err = c->flags + 3;
  r.rebuild_flag = flag;
  if (c->medium) {
    r.med_c = c->med_c[c->med_cur];
    r.med_add = c->med_add;
    for (int a = 0; a < 3; ++a) {
      r.med_n[a] = a < c->dims ? c->m.medium_n[a] : 1;
      r.med_lo[a] = c->sims[0].box_lo[a];
      r.med_hi[a] = c->sims[0].box_hi[a];
      r.med_dx[a] = a < c->dims ? (r.med_hi[a] - r.med_lo[a]) / (double)r.med_n[a] : 1.0;   // communityStructure.jl:274
    }
  }
  return r;
}

static int slot_scan(cbm_ctx* c, int32_t* in, int32_t* out, int64_t n) {
  return run_scan(c, nullptr, in, nullptr, out, n);
}

// agentStepRule!: decisions by slot, their prefixes, then the effects.
static int rule_phase(cbm_ctx* c, int32_t* flag) {
  RuleArgs r = rule_args(c, flag);
  const int64_t nm = c->sims[0].n_max;
  const bool user = c->m.rule == CBM_RULE_CUSTOM;
  if (user && !c->user_rule[0])
    return fail(c, CBM_ERR_CONFIG, "CBM_RULE_CUSTOM needs cbm_compile_model with a rule body before stepping");
  if (user && c->rule_nb[0]) HIPCHK(c, launch_rule_nb_copy(c->nblocks, c->block, c->stream, r, c->rule_nb));
  void* params[] = {(void*)&r};
  auto user_launch = [&](hipFunction_t f) {
    return hipModuleLaunchKernel(f, (unsigned)c->nblocks, 1, 1, (unsigned)c->block, 1, 1, 0, c->stream, params, nullptr);
  };
  HIPCHK(c, user ? user_launch(c->user_rule[0]) : launch_rule_decide(c->nblocks, c->block, c->stream, r));
  int rc;
  if ((rc = slot_scan(c, c->n_add_slot, c->off_add, nm))) return rc;
  if ((rc = slot_scan(c, c->dead_slot, c->off_dead, nm))) return rc;
  HIPCHK(c, user ? user_launch(c->user_rule[1]) : launch_rule_apply(c->nblocks, c->block, c->stream, r));
  return CBM_OK;
}

// update!: canonical compaction of the logical slots and the new counts.
static int compact_phase(cbm_ctx* c, int32_t* flag) {
  RuleArgs r = rule_args(c, flag);
  HIPCHK(c, launch_compact_holes(c->nblocks, c->block, c->stream, r));
  HIPCHK(c, launch_compact_slots(c->nblocks, c->block, c->stream, r));
  HIPCHK(c, launch_commit_counts(c->stream, r, c->S));
  return CBM_OK;
}

// DGADI sweeps after the reaction step (which wrote R = med_c[1 - cur]); the
// last sweep lands in C = med_c[cur].  Every sweep writes the interior lines
// of its destination, then the boundary nodes from the initial field.
static int adi_sweeps(cbm_ctx* c) {
  double* C = c->med_c[c->med_cur];
  double* R = c->med_c[1 - c->med_cur];
  double* T = c->med_tmp;
  const double* src[3];
  double* dst[3];
  if (c->dims == 1) { src[0] = R; dst[0] = C; }
  else if (c->dims == 2) { src[0] = R; dst[0] = T; src[1] = T; dst[1] = C; }
  else { src[0] = R; dst[0] = T; src[1] = T; dst[1] = R; src[2] = R; dst[2] = C; }
  int32_t n[3];
  for (int a = 0; a < 3; ++a) n[a] = a < c->dims ? c->m.medium_n[a] : 1;
  for (int ax = 0; ax < c->dims; ++ax) {
    AdiArgs a;
    std::memset(&a, 0, sizeof(a));
    a.src = src[ax]; a.dst = dst[ax];
    a.cst = c->adi_tab[ax]; a.mst = c->adi_tab[ax] + n[ax];
    for (int b = 0; b < 3; ++b) a.n[b] = n[b];
    a.dims = c->dims; a.axis = ax;
    a.sub = c->adi_sub; a.main = c->adi_main; a.sup = c->adi_sup;
    a.lines = 1;
    for (int b = 0; b < c->dims; ++b) if (b != ax) a.lines *= n[b] - 2;
    HIPCHK(c, launch_adi_sweep(c->stream, a));
    HIPCHK(c, launch_adi_border(c->stream, dst[ax], c->med_init, n, c->dims, c->med_nodes));
  }
  return CBM_OK;
}

// mediumStepDE!: agents per node, then one FTCS step of the field (or the
// reaction part and the ADI sweeps).
static int medium_phase(cbm_ctx* c) {
  MediumArgs md;
  std::memset(&md, 0, sizeof(md));
  md.g = geom(c);
  for (int d = 0; d < c->dims; ++d) md.pos[d] = (const double*)c->col[C_X + d].cur();
  md.c = c->med_c[c->med_cur];
  md.cn = c->med_c[1 - c->med_cur];
  md.count = c->med_count;
  md.dims = c->dims;
  for (int a = 0; a < 3; ++a) {
    md.n[a] = a < c->dims ? c->m.medium_n[a] : 1;
    md.lo[a] = c->sims[0].box_lo[a];
    md.hi[a] = c->sims[0].box_hi[a];
    md.dx[a] = a < c->dims ? (md.hi[a] - md.lo[a]) / (double)md.n[a] : 1.0;   // communityStructure.jl:274
  }
  md.D = c->m.medium_D; md.uptake = c->m.medium_consume; md.dt = c->m.dt;
  md.nodes = c->med_nodes;
  md.lap = c->adi ? 0 : 1;
  for (int k = 0; k < 6; ++k) { md.bmode[k] = c->m.medium_border[k]; md.bval[k] = c->m.medium_border_value[k]; }
  if (c->med_add) HIPCHK(c, launch_medium_inject(c->stream, c->med_c[c->med_cur], c->med_add, c->med_nodes));
  HIPCHK(c, launch_medium_count(c->nblocks, c->block, c->stream, md));
  if (c->user_medium) {
    void* params[] = {(void*)&md};
    HIPCHK(c, hipModuleLaunchKernel(c->user_medium, (unsigned)((md.nodes + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream,
                                    params, nullptr));
  } else {
    HIPCHK(c, launch_medium_step(c->stream, md));
  }
  if (c->adi) return adi_sweeps(c);
  c->med_cur ^= 1;
  return CBM_OK;
}

// Skin-list builds of eager steps.  The host stays at most CLV_HOST_LEAD
// steps ahead of the device: before step s it waits until the order kernel
// of step s - CLV_HOST_LEAD has written its record.  When a record written
// after the last build carries the cue (a final stage predicted an agent to
// pass half the skin within CLV_AHEAD step displacements, enough for the
// lead and one step and a half more), the build is enqueued at the end of
// step s.  A late build costs time, never exactness: stage 1 bounds every
// displacement by the reported maximum and sends agents whose list may miss
// a pair to the exact walk.
static void clv_schedule(cbm_ctx* c) {
  c->build_now = false;
  volatile ClvRecord* r = c->rec_host;
  const int64_t want = c->step - CLV_HOST_LEAD;
  if (want >= c->rec_first) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t spins = 1; r->step < want; ++spins)
      if (spins % 4096 == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10)) break;
  }
  if (r->step > c->last_build_step && r->soon != 0ull) c->build_now = true;
}

// modelStepRule! and modelStepDE! of every simulation (one thread each),
// committing the parameters the next step reads.
static int model_phase(cbm_ctx* c) {
  ModelArgs a;
  std::memset(&a, 0, sizeof(a));
  a.sims = c->d_sims;
  a.st = c->d_st;
  a.step_ptr = c->counters + 1;
  a.t_buf = c->model_t;
  a.dt = c->m.dt;
  a.seed = c->m.seed;
  a.S = c->S;
  void* params[] = {(void*)&a};
  HIPCHK(c, hipModuleLaunchKernel(c->model_fn, (unsigned)((c->S + 63) / 64), 1, 1, 64, 1, 1, 0, c->stream, params,
                                  nullptr));
  return CBM_OK;
}

static int step_once(cbm_ctx* c) {
  int rc;
  if (!c->structure_valid && (rc = force_rebuild(c))) return rc;
  const int64_t s = c->step;
  const bool clv_host = c->clv && !c->capturing;
  if (clv_host) clv_schedule(c);
  if (c->rules && (rc = rule_phase(c, c->flags + (s & 1)))) return rc;
  StageArgs a = stage_base(c);
  int32_t* flag_now = c->flags + (s & 1);
  const bool disp = (c->verlet && !c->vtime) || (c->clv && !clv_host);
  auto set_state = [&](double* const* S, double* const* U, double* const* out) {
    for (int v = 0; v < c->nvar; ++v) { a.S[v] = S[v]; a.U[v] = U[v]; a.out[v] = out[v]; a.kacc[v] = c->kacc[v]; }
  };
  double* cur[MAX_STATE] = {};
  double* alt[MAX_STATE] = {};
  for (int v = 0; v < c->nvar; ++v) { cur[v] = statecol(c, v, false); alt[v] = statecol(c, v, true); }
  a.flag = flag_now;
  // CellLinked on a skin list: the first stage orders the skin list into the
  // step's lists in cell-walk order (a.wlist), every stage walks them; the
  // final stage writes the next cells
  int32_t* ccell_next = c->clv ? (int32_t*)c->col[C_CCELL].alt() : nullptr;
  auto clv_args = [&](int first, int final_stage, int queue) {
    if (!c->clv) return;
    a.clv = 1;
    a.list = c->list; a.list_n = c->list_n;
    a.wlist = c->wlist; a.wlist_n = c->wlist_n;
    a.emax = c->emax + (s & 1);
    a.emax_next = c->emax + ((s + 1) & 1);
    a.clv_order = first;
    a.fix_stage = first ? 0 : 1;
    a.fix_q = queue;
    a.book = (first && clv_host) ? 1 : 0;
    a.rec = c->rec_dev;
    a.step_no = s;
    a.ccell_out = final_stage ? ccell_next : nullptr;
    a.tick = (final_stage && clv_host) ? c->counters + 1 : nullptr;
  };
  switch (c->m.integrator) {
    case CBM_EULER:
    case CBM_EM:
      set_state(cur, cur, alt);
      a.disp_check = disp ? 1 : 0;
      clv_args(1, 1, (int)(s & 1));
      if ((rc = run_stage(c, a, c->m.integrator == CBM_EULER ? ST_EULER : ST_EM))) return rc;
      flip_state(c);
      break;
    case CBM_HEUN:
      set_state(cur, cur, c->h[0]);
      a.disp_check = 0;
      listed_stage(c, a, 1);
      clv_args(1, 0, 0);
      if ((rc = run_stage(c, a, ST_HEUN1))) return rc;
      set_state(c->h[0], cur, cur);
      a.disp_check = disp ? 1 : 0;
      listed_stage(c, a, 2);
      clv_args(0, 1, 1);
      if ((rc = run_stage(c, a, ST_HEUN2))) return rc;
      break;
    case CBM_EULER_HEUN:
      set_state(cur, cur, c->h[0]);
      a.disp_check = 0;
      listed_stage(c, a, 1);
      if ((rc = run_stage(c, a, ST_EH1))) return rc;
      set_state(c->h[0], cur, cur);
      a.disp_check = disp ? 1 : 0;
      listed_stage(c, a, 2);
      if ((rc = run_stage(c, a, ST_EH2))) return rc;
      break;
    default:  // RK4
      a.disp_check = 0;
      set_state(cur, cur, c->h[0]);
      if ((rc = run_stage(c, a, ST_RK4_1))) return rc;
      set_state(c->h[0], cur, c->h[1]);
      if ((rc = run_stage(c, a, ST_RK4_2))) return rc;
      set_state(c->h[1], cur, c->h[0]);
      if ((rc = run_stage(c, a, ST_RK4_3))) return rc;
      set_state(c->h[0], cur, cur);
      a.disp_check = disp ? 1 : 0;
      if ((rc = run_stage(c, a, ST_RK4_4))) return rc;
      break;
  }
  // update!: commit the write-new buffers (read-old/write-new), then compact
  auto commit = [](Column& col, double*& fresh) {
    void* committed = col.buf[col.cidx];
    col.buf[col.cidx] = fresh;
    fresh = (double*)committed;
  };
  if (c->clv) c->col[C_CCELL].cidx ^= 1;
  if (c->lagged)
    for (int d = 0; d < c->dims; ++d) commit(c->col[C_FX + d], c->Fnew[d]);
  if (c->growth) {
    commit(c->col[C_GR], c->gr_new);
    commit(c->col[C_VB], c->vb_new);
  }
  if (c->medium && (rc = medium_phase(c))) return rc;   // mediumStepDE!
  if (c->model_fn && (rc = model_phase(c))) return rc;   // modelStepRule!, modelStepDE!
  if (c->rules && (rc = compact_phase(c, c->flags + (s & 1)))) return rc;
  c->t = c->t + c->m.dt;
  c->step++;
  // computeNeighbors!
  prof_begin(c, PROF_CHAIN);
  rc = neighbor_update(c, s, disp);
  prof_end(c);
  return rc;
}

static int neighbor_update(cbm_ctx* c, int64_t s, bool disp) {
  int rc;
  int32_t* flag_now = c->flags + (s & 1);
  if (c->cl && !c->clv) {
    if ((rc = rebuild_chain(c, nullptr, nullptr, true, false))) return rc;
  } else if (c->clv && !c->capturing) {
    // skin list, eager steps: the build the host scheduled, unconditional
    // (the step's last fix kernel has advanced the device step)
    if (c->build_now && s >= c->clv_dense_until) {
      c->clv_short = s - c->last_build_step <= CLV_DENSE_GAP ? c->clv_short + 1 : 0;
      if (c->clv_short >= 2) {
        c->clv_dense_until = s + CLV_DENSE_STEPS;
        c->clv_short = 0;
      }
    }
    if (c->build_now || s < c->clv_dense_until) {
      if ((rc = rebuild_chain(c, nullptr, nullptr, false, true))) return rc;
      c->last_build_step = s;
      c->build_now = false;
    }
  } else if (disp) {
    if ((rc = rebuild_chain(c, flag_now, c->flags + ((s + 1) & 1), true, true))) return rc;
  } else if (c->vtime && c->t_next <= c->t) {   // src/neighbors.jl:255
    int32_t* f = c->flags + (s & 1);
    HIPCHK(c, launch_set_flag(c->stream, f, 1));
    if ((rc = rebuild_chain(c, f, c->flags + ((s + 1) & 1), true, true))) return rc;
    c->t_next = c->t + c->m.dt_recompute;
  } else if (c->rules) {   // rebuild only when the population changed (device flag)
    if ((rc = rebuild_chain(c, flag_now, c->flags + ((s + 1) & 1), true, true))) return rc;
  } else {
    HIPCHK(c, launch_tick(c->stream, c->counters + 1));
  }
  return CBM_OK;
}

static int check_flags(cbm_ctx* c) {
  int32_t f[8];
  HIPCHK(c, hipMemcpyAsync(f, c->flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (f[2]) return fail(c, CBM_ERR_LIST_OVERFLOW, "Verlet list overflow: raise nMaxNeighbors");
  if (f[3] & 2) return fail(c, CBM_ERR_CONFIG, "an agentRule body added more than 2 agents for one agent in one step");
  if (f[3] & 4) return fail(c, CBM_ERR_CONFIG, "an agentRule body looped over neighbors without naming forNeighbors");
  if (f[3] & 8) return fail(c, CBM_ERR_CONFIG, "an agentRule body used the medium in a model without a medium");
  if (f[3]) return fail(c, CBM_ERR_CAPACITY, "agent capacity exceeded: raise preallocation (n_max)");
  if (f[7])
    return fail(c, CBM_ERR_NONFINITE, "rod contact with d_i < r_ij < (d_i + d_j)/2: the overlap d_i - r_ij is negative "
                                      "and (d_i - r_ij)^1.5 has no real value");
  return CBM_OK;
}

// ------------------------------------------------- allocation-free counts
// cbm_count_neighbors_noalloc writes its per-agent counts into this
// file-scope device array instead of memory allocated per call.  Every
// device the library runs on holds one copy (8 MiB) until the process ends.
// In host code `scratch` names a host-side stand-in, not the device copy:
// its device address is taken once per device with hipGetSymbolAddress at
// cbm_create (that device current) and only that address is used.
constexpr int64_t NOALLOC_MAX_AGENTS = 1 << 20;
static __device__ __align__(256) int64_t scratch[NOALLOC_MAX_AGENTS];

static std::mutex g_scratch_mu;
static std::vector<int64_t*> g_scratch_ptr;   // by device ordinal; nullptr until set up

// Device address of this device's scratch array (the current device is c->device).
static int scratch_setup(cbm_ctx* c) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= c->device) g_scratch_ptr.resize((size_t)c->device + 1, nullptr);
  if (g_scratch_ptr[(size_t)c->device]) return CBM_OK;
  int64_t* p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(c, CBM_ERR_HIP, std::string("scratch array address: ") + hipGetErrorString(e));
  g_scratch_ptr[(size_t)c->device] = p;
  return CBM_OK;
}

static int64_t* scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < (int)g_scratch_ptr.size() ? g_scratch_ptr[(size_t)device] : nullptr;
}

// ======================================================================= ABI
extern "C" {

const char* cbm_version(void) { return CBM_VERSION; }

int cbm_create(cbm_ctx** out, const cbm_model_desc* model, int32_t n_sims, const cbm_sim_desc* sims) {
  if (!out || !model || !sims || n_sims < 1) return CBM_ERR_ARG;
  *out = nullptr;
  cbm_ctx* c = new cbm_ctx();
  c->m = *model;
  c->S = n_sims;
  c->sims.assign(sims, sims + n_sims);
  int rc = init_ctx(c);
  if (rc == CBM_OK) rc = scratch_setup(c);
  *out = c;   // returned even on failure so cbm_last_error can explain
  return rc;
}

void cbm_destroy(cbm_ctx* c) { destroy_ctx(c); }

const char* cbm_last_error(const cbm_ctx* c) { return c ? c->err.c_str() : "null context"; }

// compileABM! for user bodies: each (HIP C++ defining struct UserLaw,
// UserNoise or UserRule, see include/cbm_hip.h) is compiled by hipRTC into
// the same skeletons as the built-in laws and rules, instantiated for this
// model's dimension, neighbor walk and boundary.
}  // extern "C"

// The hipRTC program of cbm_compile_model: the user bodies in the skeletons
// of a model of this dimension, neighbor walk (nb, one of NB_*) and box.
static std::string user_source(const cbm_model_desc& m, int nb_kind, bool growth, const char* law_body,
                               const char* noise_body, const char* rule_body, const char* medium_body) {
  const int dims = m.dims;
  const bool stages = law_body || noise_body;
  std::string src = "#define CBM_HIPRTC 1\n";
  if (m.user_de > 0)
    src += "#define CBM_USER_DE " + std::to_string(m.user_de) + "\n#define CBM_DE_BASE " +
           std::to_string(dims + (growth ? 1 : 0)) + "\n";
  src += "#include \"cbm_rule.h\"\n#include \"cbm_medium.h\"\nnamespace cbm {\n";
  if (law_body) { src += law_body; src += "\n"; }
  if (law_body && m.user_de > 0)
    src += "static_assert(has_de<UserLaw>::value || general_body<UserLaw>::value, "
           "\"user_de > 0 needs UserLaw::de, or UserLaw::pair and ode\");\n";
  if (noise_body) { src += noise_body; src += "\n"; }
  if (rule_body) { src += rule_body; src += "\n"; }
  if (medium_body) { src += medium_body; src += "\n"; }
  src += "}  // namespace cbm\n";
  if (medium_body)
    src += "extern \"C\" __global__ __launch_bounds__(256) void cbm_user_medium_step(cbm::MediumArgs m) {\n"
           "  cbm::medium_step_body<cbm::UserMedium>(m);\n}\n";
  if (rule_body)
    src += "extern \"C\" __global__ __launch_bounds__(256) void cbm_user_rule_decide(cbm::RuleArgs a) {\n"
           "  cbm::user_rule_decide<cbm::UserRule>(a);\n}\n"
           "extern \"C\" __global__ __launch_bounds__(256) void cbm_user_rule_apply(cbm::RuleArgs a) {\n"
           "  cbm::user_rule_apply<cbm::UserRule>(a);\n}\n";
  char line[256];
  if (law_body)
    src += "using cbm_law_t = cbm::UserLaw;\n"
           "extern \"C\" __device__ int cbm_law_general = cbm::general_body<cbm_law_t>::value ? 1 : 0;\n";
  else { std::snprintf(line, sizeof(line), "using cbm_law_t = cbm::PairLaw<%d>;\n", m.law); src += line; }
  src += noise_body ? "using cbm_noise_t = cbm::UserNoise;\n" : "using cbm_noise_t = cbm::ConstNoise;\n";
  // the same skeletons as the built-in laws: no neighbor walk without a law
  const bool walk = m.law != CBM_LAW_NONE;
  const int nb = walk ? nb_kind : NB_FULL;
  const int per = walk && m.periodic ? 1 : 0;
  char inst[512];
  for (int noise = 0; stages && noise < 2; ++noise) {
    std::snprintf(inst, sizeof(inst),
                  "extern \"C\" __global__ __launch_bounds__(256) void cbm_user_stage_%d(cbm::StageArgs a) {\n"
                  "  cbm::user_stage<%d, %d, cbm_law_t, %d, %d, cbm_noise_t>(a);\n}\n",
                  noise, dims, nb, per, noise);
    src += inst;
    if (nb == NB_CELL) {   // the brick-tiled cell-linked stage, run before the row kernel
      std::snprintf(inst, sizeof(inst),
                    "extern \"C\" __global__ __launch_bounds__(256) void cbm_user_brick_%d(cbm::StageArgs a) {\n"
                    "  extern __shared__ unsigned char lds[];\n"
                    "  cbm::user_brick<%d, cbm_law_t, %d, %d, cbm_noise_t>(a, lds);\n}\n",
                    noise, dims, per, noise);
      src += inst;
    }
  }
  return src;
}

// hipRTC of a user-body program for gfx950; the log on failure.
static hiprtcResult user_compile(const std::string& src, std::vector<char>* code, std::string* log) {
  const char* headers[] = {cbm_rtc_cbm_device_h, cbm_rtc_cbm_structs_h, cbm_rtc_cbm_stage_h, cbm_rtc_cbm_rule_h,
                           cbm_rtc_cbm_medium_h, cbm_rtc_cbm_model_h};
  const char* names[] = {"cbm_device.h", "cbm_structs.h", "cbm_stage.h", "cbm_rule.h", "cbm_medium.h", "cbm_model.h"};
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "cbm_user_bodies.hip", 6, headers, names);
  if (r != HIPRTC_SUCCESS) { *log = "hiprtcCreateProgram failed"; return r; }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math"};
  r = hiprtcCompileProgram(prog, 5, opts);
  size_t logn = 0;
  hiprtcGetProgramLogSize(prog, &logn);
  std::string l(logn + 1, '\0');
  if (logn) hiprtcGetProgramLog(prog, &l[0]);
  *log = l.c_str();
  if (r == HIPRTC_SUCCESS && code) {
    size_t code_n = 0;
    hiprtcGetCodeSize(prog, &code_n);
    code->resize(code_n);
    hiprtcGetCode(prog, code->data());
  }
  hiprtcDestroyProgram(&prog);
  return r;
}

extern "C" {

int cbm_compile_check(const cbm_model_desc* model, const char* law_body, const char* noise_body,
                      const char* rule_body, const char* medium_body, char* log, int64_t log_cap) {
  if (!model || (log && log_cap < 1)) return CBM_ERR_ARG;
  if (!law_body && !noise_body && !rule_body && !medium_body) return CBM_ERR_ARG;
  const cbm_model_desc& m = *model;
  if (m.dims < 1 || m.dims > 3 || m.neighbors < CBM_FULL || m.neighbors > CBM_CLVD) return CBM_ERR_CONFIG;
  const int nb = m.neighbors == CBM_FULL ? NB_FULL : (m.neighbors == CBM_CELL_LINKED ? NB_CELL : NB_LIST);
  const bool growth = m.rule == CBM_RULE_GROWTH_DIVISION || m.rule == CBM_RULE_CUSTOM;
  std::string lg;
  const hiprtcResult r = user_compile(user_source(m, nb, growth, law_body, noise_body, rule_body, medium_body),
                                      nullptr, &lg);
  if (log) {
    const size_t n = std::min(lg.size(), (size_t)log_cap - 1);
    std::memcpy(log, lg.data(), n);
    log[n] = '\0';
  }
  return r == HIPRTC_SUCCESS ? CBM_OK : CBM_ERR_COMPILE;
}

static int params_from_device(cbm_ctx* c, int32_t sim);

static std::string model_source(const char* body) {
  std::string src = "#define CBM_HIPRTC 1\n#include \"cbm_model.h\"\nnamespace cbm {\n";
  src += body;
  src += "\n}  // namespace cbm\n"
         "extern \"C\" __global__ __launch_bounds__(64) void cbm_user_model(cbm::ModelArgs a) {\n"
         "  cbm::model_step_body<cbm::UserModel>(a);\n}\n";
  return src;
}

int cbm_compile_model_scope_check(const char* body, char* log, int64_t log_cap) {
  if (!body || (log && log_cap < 1)) return CBM_ERR_ARG;
  std::string lg;
  const hiprtcResult r = user_compile(model_source(body), nullptr, &lg);
  if (log) {
    const size_t n = std::min(lg.size(), (size_t)log_cap - 1);
    std::memcpy(log, lg.data(), n);
    log[n] = '\0';
  }
  return r == HIPRTC_SUCCESS ? CBM_OK : CBM_ERR_COMPILE;
}

int cbm_compile_model_scope(cbm_ctx* c, const char* body) {
  if (!c || !body) return CBM_ERR_ARG;
  if (c->clv)
    return fail(c, CBM_ERR_CONFIG, "a model-scope body moves the law's parameters on the device, past what the "
                                   "cell-linked skin lists were sized for: create the context with walk = 1");
  std::vector<char> code;
  std::string log;
  if (user_compile(model_source(body), &code, &log) != HIPRTC_SUCCESS)
    return fail(c, CBM_ERR_COMPILE, std::string("hipRTC: ") + log);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < c->S; ++s) {
    int rc = params_from_device(c, s);
    if (rc) return rc;
  }
  if (c->model_mod) {
    (void)hipModuleUnload(c->model_mod);
    c->model_mod = nullptr;
    c->model_fn = nullptr;
  }
  HIPCHK(c, hipModuleLoadData(&c->model_mod, code.data()));
  HIPCHK(c, hipModuleGetFunction(&c->model_fn, c->model_mod, "cbm_user_model"));
  if (!c->model_t) HIPCHK(c, dalloc(&c->model_t, 2));
  // the current step's start time in its parity slot (the device step counter equals c->step)
  HIPCHK(c, hipMemcpy(c->model_t + (c->step & 1), &c->t, sizeof(double), hipMemcpyHostToDevice));
  return CBM_OK;
}

int cbm_compile_model(cbm_ctx* c, const char* law_body, const char* noise_body, const char* rule_body,
                      const char* medium_body) {
  if (!c) return CBM_ERR_ARG;
  const bool custom = c->m.law == CBM_LAW_CUSTOM;
  if (custom != (law_body != nullptr))
    return fail(c, CBM_ERR_CONFIG, "a law body is required with law CBM_LAW_CUSTOM and only then");
  if ((c->m.rule == CBM_RULE_CUSTOM) != (rule_body != nullptr))
    return fail(c, CBM_ERR_CONFIG, "a rule body is required with rule CBM_RULE_CUSTOM and only then");
  if (noise_body && c->m.integrator != CBM_EM && c->m.integrator != CBM_EULER_HEUN)
    return fail(c, CBM_ERR_CONFIG, "an agentSDE body needs the EM or EulerHeun integrator");
  if (medium_body && !c->medium) return fail(c, CBM_ERR_CONFIG, "a mediumODE body needs a medium (medium_n)");
  if (!law_body && !noise_body && !rule_body && !medium_body) return fail(c, CBM_ERR_ARG, "nothing to compile");
  std::vector<char> code;
  std::string log;
  const hiprtcResult r = user_compile(user_source(c->m, c->nb_kind, c->growth, law_body, noise_body, rule_body,
                                                  medium_body), &code, &log);
  if (r != HIPRTC_SUCCESS) return fail(c, CBM_ERR_COMPILE, std::string("hipRTC: ") + log);
  const bool stages = law_body || noise_body;
  const int nb = c->m.law != CBM_LAW_NONE ? c->nb_kind : NB_FULL;
  if (c->user_mod) {
    // steps queued on the stream may still run kernels of the old module
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipModuleUnload(c->user_mod);
    c->user_mod = nullptr;
  }
  for (int k = 0; k < 2; ++k) { c->user_fn[k] = nullptr; c->user_brick[k] = nullptr; c->user_rule[k] = nullptr; }
  c->user_medium = nullptr;
  HIPCHK(c, hipModuleLoadData(&c->user_mod, code.data()));
  if (law_body && c->lagged) {
    // a general agentODE body (UserLaw::ode) writes its own equations: there
    // is no force accumulator to lag
    hipDeviceptr_t gp = nullptr;
    size_t gb = 0;
    int general = 0;
    if (hipModuleGetGlobal(&gp, &gb, c->user_mod, "cbm_law_general") == hipSuccess && gb == sizeof(int))
      HIPCHK(c, hipMemcpyDtoH(&general, gp, sizeof(int)));
    if (general) {
      (void)hipModuleUnload(c->user_mod);
      c->user_mod = nullptr;
      return fail(c, CBM_ERR_CONFIG, "a general agentODE body (UserLaw::ode) takes no lagged forces");
    }
  }
  if (medium_body) HIPCHK(c, hipModuleGetFunction(&c->user_medium, c->user_mod, "cbm_user_medium_step"));
  if (rule_body) {
    HIPCHK(c, hipModuleGetFunction(&c->user_rule[0], c->user_mod, "cbm_user_rule_decide"));
    HIPCHK(c, hipModuleGetFunction(&c->user_rule[1], c->user_mod, "cbm_user_rule_apply"));
    // a body with a neighbor loop reads its neighbors from a copy taken
    // before the rule's passes
    if (std::strstr(rule_body, "forNeighbors"))
      for (int k = 0; k < 4 + USER_PARAMS; ++k)
        if (!c->rule_nb[k]) HIPCHK(c, dalloc(&c->rule_nb[k], c->C));
    // a body adding to the medium sums its additions per node
    if (std::strstr(rule_body, "medium_add") && c->medium && !c->med_add) {
      HIPCHK(c, dalloc(&c->med_add, c->med_nodes));
      HIPCHK(c, hipMemset(c->med_add, 0, sizeof(unsigned long long) * (size_t)c->med_nodes));
    }
  }
  if (stages) {
    HIPCHK(c, hipModuleGetFunction(&c->user_fn[0], c->user_mod, "cbm_user_stage_0"));
    HIPCHK(c, hipModuleGetFunction(&c->user_fn[1], c->user_mod, "cbm_user_stage_1"));
  }
  if (stages && nb == NB_CELL) {
    HIPCHK(c, hipModuleGetFunction(&c->user_brick[0], c->user_mod, "cbm_user_brick_0"));
    HIPCHK(c, hipModuleGetFunction(&c->user_brick[1], c->user_mod, "cbm_user_brick_1"));
  }
  c->user_nb = nb;
  c->user_noise = noise_body != nullptr;
  if (c->user_noise) {
    c->wl = false;   // the noisy stages go through the module, whose stage 2 walks the cells
    if (c->m.integrator == CBM_EULER_HEUN)
      for (int v = 0; v < c->nvar; ++v)
        if (!c->gacc[v]) HIPCHK(c, dalloc(&c->gacc[v], c->C));
  }
  return CBM_OK;
}

int cbm_compile(cbm_ctx* c, const char* body) {
  if (!c || !body) return CBM_ERR_ARG;
  return cbm_compile_model(c, body, nullptr, nullptr, nullptr);
}

static int field_col(cbm_ctx* c, int32_t field, int& k) {
  switch (field) {
    case CBM_F_X: case CBM_F_Y: case CBM_F_Z:
      if (field >= c->dims) return CBM_ERR_ARG;
      k = C_X + field; return CBM_OK;
    case CBM_F_ID: k = C_ID; return CBM_OK;
    case CBM_F_VOL: if (!c->growth) return CBM_ERR_ARG; k = C_VOL; return CBM_OK;
    case CBM_F_GROWTH: if (!c->growth) return CBM_ERR_ARG; k = C_GR; return CBM_OK;
    case CBM_F_VBIRTH: if (!c->growth) return CBM_ERR_ARG; k = C_VB; return CBM_OK;
    case CBM_F_LF: if (!c->rules) return CBM_ERR_ARG; k = C_LF; return CBM_OK;
    case CBM_F_FX: case CBM_F_FY: case CBM_F_FZ:
      if (!(c->lagged || c->rods) || field - CBM_F_FX >= c->dims) return CBM_ERR_ARG;
      k = C_FX + (field - CBM_F_FX); return CBM_OK;
    case CBM_F_VX: case CBM_F_VY: case CBM_F_VZ:
      if (!(c->inertial || c->rods) || field - CBM_F_VX >= c->dims) return CBM_ERR_ARG;
      k = C_VX + (field - CBM_F_VX); return CBM_OK;
    case CBM_F_RADIUS: if (!c->inertial) return CBM_ERR_ARG; k = C_RAD; return CBM_OK;
    case CBM_F_MASS: if (!(c->inertial || c->rods)) return CBM_ERR_ARG; k = C_MASS; return CBM_OK;
    case CBM_F_THETA: if (!c->rods) return CBM_ERR_ARG; k = C_TH; return CBM_OK;
    case CBM_F_OMEGA: if (!c->rods) return CBM_ERR_ARG; k = C_OM; return CBM_OK;
    case CBM_F_DIAMETER: if (!c->rods) return CBM_ERR_ARG; k = C_DIAM; return CBM_OK;
    case CBM_F_LENGTH: if (!c->rods) return CBM_ERR_ARG; k = C_LEN; return CBM_OK;
    case CBM_F_TORQUE: if (!c->rods) return CBM_ERR_ARG; k = C_TQ; return CBM_OK;
    case CBM_F_PRESSURE: if (!c->rods) return CBM_ERR_ARG; k = C_PR; return CBM_OK;
    case CBM_F_LTARGET: if (!c->has_col[C_LT]) return CBM_ERR_ARG; k = C_LT; return CBM_OK;
    case CBM_F_FATE: if (!c->inertial) return CBM_ERR_ARG; k = C_FATE; return CBM_OK;
    case CBM_F_U0: case CBM_F_U1: case CBM_F_U2: case CBM_F_U3:
      if (!c->has_col[C_U0]) return CBM_ERR_ARG;
      k = C_U0 + (field - CBM_F_U0); return CBM_OK;
    case CBM_F_U4: case CBM_F_U5: case CBM_F_U6: case CBM_F_U7:
      if (!c->has_col[C_U0]) return CBM_ERR_ARG;
      k = C_U4 + (field - CBM_F_U4); return CBM_OK;
    default: return CBM_ERR_ARG;
  }
}

static int xfer(cbm_ctx* c, int32_t field, void* host, int64_t n, int32_t sim, bool up) {
  if (!c) return CBM_ERR_ARG;
  if (sim < 0 || sim >= c->S) return fail(c, CBM_ERR_ARG, "simulation index out of range");
  int k;
  if (field_col(c, field, k)) return fail(c, CBM_ERR_ARG, "field not present in this model");
  if (n < 0 || n > c->sims[sim].n_max) return fail(c, CBM_ERR_ARG, "n exceeds capacity");
  if (n == 0) return CBM_OK;
  if (up && !(field == CBM_F_ID || field == CBM_F_LF || field == CBM_F_FATE))
    for (int64_t i = 0; i < n; ++i)
      if (!std::isfinite(((const double*)host)[i]))
        return fail(c, CBM_ERR_NONFINITE, "non-finite value in an uploaded agent field");
  if (up && field == CBM_F_FATE)
    for (int64_t i = 0; i < n; ++i) {
      const int64_t f = ((const int64_t*)host)[i];
      if (f < 1 || f > 2) return fail(c, CBM_ERR_ARG, "cellFate must be 1 or 2 (the f0 matrix is 2 x 2)");
    }
  const int bytes = c->col[k].bytes;
  // rows of this simulation only: a one-simulation geometry view
  XferArgs a;
  std::memset(&a, 0, sizeof(a));
  a.g = geom(c);
  a.slot = (const int32_t*)c->col[C_SLOT].cur();
  a.field = c->col[k].cur();
  a.staging = c->d_staging;
  a.bytes = bytes;
  a.to_device = up ? 1 : 0;
  a.n = n;
  a.sim = sim;
  if (up) HIPCHK(c, hipMemcpyAsync(c->d_staging, host, (size_t)n * bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_xfer(c->nblocks, c->block, c->stream, a));
  if (!up) {
    HIPCHK(c, hipMemcpyAsync(host, c->d_staging, (size_t)n * bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (k <= C_Z) c->structure_valid = false;
  }
  return CBM_OK;
}

int cbm_upload(cbm_ctx* c, int32_t field, const void* host, int64_t n, int32_t sim) {
  return xfer(c, field, const_cast<void*>(host), n, sim, true);
}

int cbm_download(cbm_ctx* c, int32_t field, void* host, int64_t n, int32_t sim) {
  int rc = xfer(c, field, host, n, sim, false);
  if (rc) return rc;
  return check_flags(c);
}

int cbm_set_n(cbm_ctx* c, int32_t sim, int64_t n) {
  if (!c || sim < 0 || sim >= c->S) return CBM_ERR_ARG;
  if (n < 0 || n > c->sims[sim].n_max) return fail(c, CBM_ERR_CAPACITY, "N exceeds capacity");
  SimState st;
  HIPCHK(c, hipMemcpyAsync(&st, c->d_st + sim, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  st.N = n; st.n_tot = n;
  if (st.id_max < n) st.id_max = n;
  HIPCHK(c, hipMemcpyAsync(c->d_st + sim, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->structure_valid = false;
  return CBM_OK;
}

int cbm_upload_medium(cbm_ctx* c, const double* host, int64_t n) {
  if (!c) return CBM_ERR_ARG;
  if (!c->medium) return fail(c, CBM_ERR_CONFIG, "model has no medium");
  if (!host || n != c->med_nodes) return fail(c, CBM_ERR_ARG, "medium upload needs prod(NMedium) values");
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(host[i])) return fail(c, CBM_ERR_NONFINITE, "non-finite value in the uploaded medium");
  HIPCHK(c, hipMemcpyAsync(c->med_c[c->med_cur], host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (c->adi)   // DGADI holds the boundary at this field
    HIPCHK(c, hipMemcpyAsync(c->med_init, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CBM_OK;
}

int cbm_download_medium(cbm_ctx* c, double* host, int64_t n) {
  if (!c) return CBM_ERR_ARG;
  if (!c->medium) return fail(c, CBM_ERR_CONFIG, "model has no medium");
  if (!host || n != c->med_nodes) return fail(c, CBM_ERR_ARG, "medium download needs prod(NMedium) values");
  HIPCHK(c, hipMemcpyAsync(host, c->med_c[c->med_cur], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CBM_OK;
}

int cbm_neighbors(cbm_ctx* c) {
  if (!c) return CBM_ERR_ARG;
  c->structure_valid = false;
  int rc = force_rebuild(c);
  if (rc) return rc;
  return CBM_OK;
}

int cbm_step(cbm_ctx* c, int64_t nsteps) {
  if (!c || nsteps < 0) return CBM_ERR_ARG;
  if (c->m.law == CBM_LAW_CUSTOM && !c->user_fn[0]) return fail(c, CBM_ERR_CONFIG, "law CBM_LAW_CUSTOM: call cbm_compile first");
  if (c->m.rule == CBM_RULE_CUSTOM && !c->user_rule[0])
    return fail(c, CBM_ERR_CONFIG, "rule CBM_RULE_CUSTOM: call cbm_compile_model with a rule body first");
  c->last_launches = 0;
  if (!c->structure_valid) { int rc = force_rebuild(c); if (rc) return rc; }
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int64_t k = 0; k < nsteps; ++k) {
    int rc = step_once(c);
    if (rc) return rc;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  return CBM_OK;
}

int cbm_step_graph(cbm_ctx* c, int64_t nsteps, int64_t replays) {
  if (!c || nsteps <= 0 || replays < 0) return CBM_ERR_ARG;
  if (c->vtime) return fail(c, CBM_ERR_CONFIG, "graph capture needs a device-side rebuild decision (not VerletTime)");
  if (nsteps % 2) return fail(c, CBM_ERR_ARG, "graph step count must be even (buffer parity)");
  if (!c->structure_valid) { int rc = force_rebuild(c); if (rc) return rc; }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int cidx0[C_NCOLS];
  for (int k = 0; k < C_NCOLS; ++k) cidx0[k] = c->col[k].cidx;
  void* fnew0[3] = {c->Fnew[0], c->Fnew[1], c->Fnew[2]};
  const int med0 = c->med_cur;
  void* grn0 = c->gr_new;
  void* vbn0 = c->vb_new;
  int64_t step0 = c->step;
  double t0 = c->t;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = CBM_OK;
  c->capturing = true;   // the device decides every build in a graph
  for (int64_t k = 0; k < nsteps && rc == CBM_OK; ++k) rc = step_once(c);
  c->capturing = false;
  hipError_t ce = hipStreamEndCapture(c->stream, &graph);
  if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
  if (ce != hipSuccess) return fail(c, CBM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
  bool same = true;
  for (int k = 0; k < C_NCOLS; ++k) same = same && (c->col[k].cidx == cidx0[k]);
  for (int d = 0; d < 3; ++d) same = same && (c->Fnew[d] == fnew0[d]);
  same = same && c->med_cur == med0 && (void*)c->gr_new == grn0 && (void*)c->vb_new == vbn0;
  if (!same) { (void)hipGraphDestroy(graph); return fail(c, CBM_ERR_CONFIG, "step buffers are not periodic over the captured steps"); }
  // capture advanced the host counters once; the replays do the real work
  c->step = step0; c->t = t0;
  hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ie != hipSuccess) { (void)hipGraphDestroy(graph); return fail(c, CBM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie)); }
  c->last_launches = 0;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int64_t r = 0; r < replays; ++r) {
    hipError_t le = hipGraphLaunch(exec, c->stream);
    if (le != hipSuccess) { (void)hipGraphExecDestroy(exec); (void)hipGraphDestroy(graph); return fail(c, CBM_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(le)); }
    for (int64_t k = 0; k < nsteps; ++k) { c->t = c->t + c->m.dt; c->step++; }
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  // eager steps after the replays: their records describe the lists the
  // replays left (the bound and the cue count from the last build)
  c->last_build_step = c->step - 1;
  c->rec_first = c->step;
  return CBM_OK;
}

int cbm_sync(cbm_ctx* c) {
  if (!c) return CBM_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_flags(c);
}

int cbm_get_n(cbm_ctx* c, int32_t sim, int64_t* n) {
  if (!c || !n || sim < 0 || sim >= c->S) return CBM_ERR_ARG;
  SimState st;
  HIPCHK(c, hipMemcpyAsync(&st, c->d_st + sim, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *n = st.N;
  return CBM_OK;
}

int cbm_get_id_max(cbm_ctx* c, int32_t sim, int64_t* id_max) {
  if (!c || !id_max || sim < 0 || sim >= c->S) return CBM_ERR_ARG;
  SimState st;
  HIPCHK(c, hipMemcpyAsync(&st, c->d_st + sim, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *id_max = st.id_max;
  return CBM_OK;
}

int cbm_get_t(cbm_ctx* c, double* t, int64_t* step) {
  if (!c) return CBM_ERR_ARG;
  if (t) *t = c->t;
  if (step) *step = c->step;
  return CBM_OK;
}

int cbm_get_bricks(cbm_ctx* c, int32_t* bdim, int32_t* cap, int64_t* nbricks, int64_t* overflows) {
  if (!c || !bdim || !cap || !nbricks || !overflows) return CBM_ERR_ARG;
  int64_t v[4];
  HIPCHK(c, hipMemcpyAsync(v, c->counters, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const bool tiled = c->bricks != nullptr;
  for (int a = 0; a < 3; ++a) bdim[a] = tiled ? c->bdim[a] : 0;
  *cap = tiled ? c->brick_cap : 0;
  *nbricks = tiled ? c->nbricks : 0;
  *overflows = v[2];
  return CBM_OK;
}

int cbm_get_sub_bricks(cbm_ctx* c, int32_t* bdim, int32_t* cap, int64_t* nbricks) {
  if (!c || !bdim || !cap || !nbricks) return CBM_ERR_ARG;
  const bool sub = c->bricks2 != nullptr;
  for (int a = 0; a < 3; ++a) bdim[a] = sub ? c->bdim2[a] : 0;
  *cap = sub ? c->brick_cap2 : 0;
  *nbricks = sub ? c->nbricks2 : 0;
  return CBM_OK;
}

// ------------------------------------------------------------- snapshots
static void snapshot_free(cbm_ctx* c) {
  for (void* p : c->snap_buf) if (p) (void)hipFree(p);
  c->snap_buf.clear();
  c->snap_fields.clear();
  if (c->snap_cnt) { (void)hipFree(c->snap_cnt); c->snap_cnt = nullptr; }
  if (c->snap_med) { (void)hipFree(c->snap_med); c->snap_med = nullptr; }
  c->snap_cap = 0; c->snap_count = 0;
  c->snap_t.clear();
}

int cbm_snapshot_config(cbm_ctx* c, const int32_t* fields, int32_t nfields, int32_t capacity) {
  if (!c || nfields < 0 || capacity < 0 || (nfields > 0 && !fields)) return CBM_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  snapshot_free(c);
  for (int32_t q = 0; q < nfields; ++q) {
    int k;
    if (field_col(c, fields[q], k)) return fail(c, CBM_ERR_ARG, "snapshot field not present in this model");
    if (c->col[k].bytes != 8) return fail(c, CBM_ERR_ARG, "snapshot fields are 8-byte columns");
  }
  if (capacity == 0) return CBM_OK;
  for (int32_t q = 0; q < nfields; ++q) {
    void* p = nullptr;
    HIPCHK(c, hipMalloc(&p, (size_t)capacity * c->C * 8));
    c->snap_buf.push_back(p);
    c->snap_fields.push_back(fields[q]);
  }
  HIPCHK(c, dalloc(&c->snap_cnt, (int64_t)capacity * c->S));
  if (c->medium) HIPCHK(c, dalloc(&c->snap_med, (int64_t)capacity * c->med_nodes));
  c->snap_cap = capacity;
  return CBM_OK;
}

// saveRAM!: queues the gathers of the current state as snapshot snap_count.
static int snapshot_take(cbm_ctx* c) {
  if (c->snap_count >= c->snap_cap) return fail(c, CBM_ERR_CAPACITY, "snapshot capacity exhausted");
  const int64_t k = c->snap_count;
  SnapArgs a;
  std::memset(&a, 0, sizeof(a));
  a.g = geom(c);
  a.slot = (const int32_t*)c->col[C_SLOT].cur();
  a.n_out = c->snap_cnt + k * c->S;
  for (size_t q = 0; q < c->snap_fields.size(); ++q) {
    int col;
    (void)field_col(c, c->snap_fields[q], col);
    a.field = (const uint64_t*)c->col[col].cur();
    a.out = (uint64_t*)c->snap_buf[q] + k * c->C;
    HIPCHK(c, launch_snapshot(c->nblocks, c->block, c->stream, a));
    a.n_out = nullptr;
  }
  if (c->snap_fields.empty()) HIPCHK(c, launch_snapshot(c->nblocks, c->block, c->stream, a));   // N only
  if (c->medium)
    HIPCHK(c, hipMemcpyAsync(c->snap_med + k * c->med_nodes, c->med_c[c->med_cur], sizeof(double) * c->med_nodes,
                             hipMemcpyDeviceToDevice, c->stream));
  c->snap_t.push_back(c->t);
  c->snap_count++;
  return CBM_OK;
}

int cbm_snapshot_take(cbm_ctx* c) {
  if (!c) return CBM_ERR_ARG;
  return snapshot_take(c);
}

int cbm_step_save(cbm_ctx* c, int64_t nsteps, int64_t save_each) {
  if (!c || nsteps < 0 || save_each < 1) return CBM_ERR_ARG;
  if (c->snap_count + nsteps / save_each > c->snap_cap)
    return fail(c, CBM_ERR_CAPACITY, "more snapshots than cbm_snapshot_config's capacity");
  for (int64_t i = 1; i <= nsteps; ++i) {
    int rc = cbm_step(c, 1);
    if (rc) return rc;
    if (i % save_each == 0 && (rc = snapshot_take(c))) return rc;
  }
  return CBM_OK;
}

int cbm_snapshot_count(cbm_ctx* c, int32_t* count) {
  if (!c || !count) return CBM_ERR_ARG;
  *count = c->snap_count;
  return CBM_OK;
}

int cbm_snapshot_get(cbm_ctx* c, int32_t k, int32_t field, int32_t sim, void* host, int64_t n_cap, int64_t* n_out,
                     double* t_out) {
  if (!c) return CBM_ERR_ARG;
  if (k < 0 || k >= c->snap_count) return fail(c, CBM_ERR_ARG, "snapshot index out of range");
  if (sim < 0 || sim >= c->S) return fail(c, CBM_ERR_ARG, "simulation index out of range");
  int q = -1;
  for (size_t f = 0; f < c->snap_fields.size(); ++f) if (c->snap_fields[f] == field) q = (int)f;
  if (q < 0) return fail(c, CBM_ERR_ARG, "field not in the snapshot configuration");
  int64_t n = 0;
  HIPCHK(c, hipMemcpyAsync(&n, c->snap_cnt + (int64_t)k * c->S + sim, sizeof(n), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n_out) *n_out = n;
  if (t_out) *t_out = c->snap_t[k];
  if (n > n_cap) return fail(c, CBM_ERR_ARG, "host buffer smaller than the snapshot's N");
  if (n > 0) {
    const uint64_t* src = (const uint64_t*)c->snap_buf[q] + (int64_t)k * c->C + c->h_sims[sim].row_off;
    HIPCHK(c, hipMemcpyAsync(host, src, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return check_flags(c);
}

int cbm_snapshot_get_medium(cbm_ctx* c, int32_t k, double* host, int64_t n) {
  if (!c) return CBM_ERR_ARG;
  if (!c->medium) return fail(c, CBM_ERR_CONFIG, "model has no medium");
  if (k < 0 || k >= c->snap_count) return fail(c, CBM_ERR_ARG, "snapshot index out of range");
  if (!host || n != c->med_nodes) return fail(c, CBM_ERR_ARG, "medium snapshot needs prod(NMedium) values");
  HIPCHK(c, hipMemcpyAsync(host, c->snap_med + (int64_t)k * c->med_nodes, sizeof(double) * n, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CBM_OK;
}

int cbm_get_rebuilds(cbm_ctx* c, int64_t* rebuilds) {
  if (!c || !rebuilds) return CBM_ERR_ARG;
  int64_t v[4];
  HIPCHK(c, hipMemcpyAsync(v, c->counters, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *rebuilds = v[0];
  return CBM_OK;
}

// Bound on every agent's distance from its skin-list reference position:
// half the skin (a little over), or the largest displacement the final
// stages reported past it since the last build (read from the device).
static int clv_bound(cbm_ctx* c, double* eb) {
  unsigned long long e = 0ull;
  HIPCHK(c, hipMemcpyAsync(&e, c->efin, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *eb = std::max(c->thr * (1.0 + 1e-6), std::sqrt(__builtin_bit_cast(double, e)) * (1.0 + 1e-9));
  return CBM_OK;
}

static int export_sets(cbm_ctx* c, int32_t sim, double radius, int32_t* counts, int64_t* lists, int64_t cap, int64_t* counts64) {
  if (!c || sim < 0 || sim >= c->S) return CBM_ERR_ARG;
  if (!c->structure_valid) { int rc = force_rebuild(c); if (rc) return rc; }
  int64_t N;
  int rc = cbm_get_n(c, sim, &N);
  if (rc) return rc;
  // Scratch for the export; every entry read back below is written by the
  // kernel, and the overflow word is cleared on the context stream (the
  // stream does not synchronise with the null stream).
  int32_t* d_counts = nullptr; int64_t* d_lists = nullptr; int64_t* d_c64 = nullptr;
  int32_t* d_ovf = c->flags + 4;
  int64_t nmax = c->sims[sim].n_max;
  if (counts) HIPCHK(c, hipMalloc((void**)&d_counts, sizeof(int32_t) * nmax));
  if (lists) HIPCHK(c, hipMalloc((void**)&d_lists, sizeof(int64_t) * nmax * std::max<int64_t>(cap, 1)));
  if (counts64) HIPCHK(c, hipMalloc((void**)&d_c64, sizeof(int64_t) * nmax));
  HIPCHK(c, launch_set_flag(c->stream, d_ovf, 0));
  ExportArgs e;
  std::memset(&e, 0, sizeof(e));
  e.s = stage_base(c);
  for (int d = 0; d < c->dims; ++d) e.s.S[d] = (const double*)c->col[C_X + d].cur();
  e.slot = (const int32_t*)c->col[C_SLOT].cur();
  e.counts = d_counts; e.lists = d_lists; e.counts64 = d_c64; e.cap = cap; e.radius = radius;
  e.take_all = (c->verlet && (lists || radius <= 0.0)) ? 1 : 0;
  e.clv = c->clv ? 1 : 0;
  if (c->clv && (rc = clv_bound(c, &e.e_bound))) return rc;
  e.overflow = d_ovf;
  e.only_sim = sim;
  int ovf = 0;
  hipError_t le = launch_export(c->dims, c->nb_kind, c->nblocks, c->block, c->stream, e);
  if (le == hipSuccess) le = hipStreamSynchronize(c->stream);
  if (le == hipSuccess && counts) le = hipMemcpyAsync(counts, d_counts, sizeof(int32_t) * N, hipMemcpyDeviceToHost, c->stream);
  if (le == hipSuccess && lists) le = hipMemcpyAsync(lists, d_lists, sizeof(int64_t) * N * cap, hipMemcpyDeviceToHost, c->stream);
  if (le == hipSuccess && counts64) le = hipMemcpyAsync(counts64, d_c64, sizeof(int64_t) * N, hipMemcpyDeviceToHost, c->stream);
  if (le == hipSuccess) le = hipMemcpyAsync(&ovf, d_ovf, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (le == hipSuccess) le = hipStreamSynchronize(c->stream);
  if (d_counts) (void)hipFree(d_counts); if (d_lists) (void)hipFree(d_lists); if (d_c64) (void)hipFree(d_c64);
  if (le != hipSuccess) return fail(c, CBM_ERR_HIP, std::string("neighbor export: ") + hipGetErrorString(le));
  if (ovf) return fail(c, CBM_ERR_LIST_OVERFLOW, "neighbor export capacity too small");
  return CBM_OK;
}

int cbm_get_neighbor_lists(cbm_ctx* c, int32_t sim, double radius, int32_t* counts, int64_t* lists, int64_t cap) {
  if (!counts || !lists || cap < 1) return CBM_ERR_ARG;
  return export_sets(c, sim, radius, counts, lists, cap, nullptr);
}

int cbm_count_neighbors(cbm_ctx* c, int32_t sim, double radius, int64_t* out) {
  if (!out) return CBM_ERR_ARG;
  return export_sets(c, sim, radius, nullptr, nullptr, 0, out);
}

int cbm_count_neighbors_noalloc(cbm_ctx* c, int32_t sim, double radius, int64_t* out) {
  if (!c || !out || sim < 0 || sim >= c->S) return CBM_ERR_ARG;
  int64_t* counts = scratch_of(c->device);
  if (!counts) return fail(c, CBM_ERR_HIP, "scratch array of this device not set up");
  int64_t N;
  int rc = cbm_get_n(c, sim, &N);
  if (rc) return rc;
  if (N > NOALLOC_MAX_AGENTS)
    return fail(c, CBM_ERR_ARG, "cbm_count_neighbors_noalloc: " + std::to_string(N) +
                                    " agents exceed its scratch array of 1048576; use cbm_count_neighbors");
  if (!c->structure_valid && (rc = force_rebuild(c))) return rc;
  // the kernel writes counts[slot] for every slot below N, so nothing left by
  // an earlier call is read; the overflow word is unused without lists
  ExportArgs e;
  std::memset(&e, 0, sizeof(e));
  e.s = stage_base(c);
  for (int d = 0; d < c->dims; ++d) e.s.S[d] = (const double*)c->col[C_X + d].cur();
  e.slot = (const int32_t*)c->col[C_SLOT].cur();
  e.counts64 = counts;
  e.radius = radius;
  e.take_all = (c->verlet && radius <= 0.0) ? 1 : 0;
  e.clv = c->clv ? 1 : 0;
  if (c->clv && (rc = clv_bound(c, &e.e_bound))) return rc;
  e.overflow = c->flags + 4;
  e.only_sim = sim;
  hipError_t le = launch_export(c->dims, c->nb_kind, c->nblocks, c->block, c->stream, e);
  if (le == hipSuccess && N > 0)
    le = hipMemcpyAsync(out, counts, sizeof(int64_t) * (size_t)N, hipMemcpyDeviceToHost, c->stream);
  if (le == hipSuccess) le = hipStreamSynchronize(c->stream);
  if (le != hipSuccess) return fail(c, CBM_ERR_HIP, std::string("neighbor counts: ") + hipGetErrorString(le));
  return CBM_OK;
}

// The parameters a model-scope body has moved on the device, into the host copies.
static int params_from_device(cbm_ctx* c, int32_t sim) {
  if (!c->model_fn) return CBM_OK;
  SimDev d;
  HIPCHK(c, hipMemcpyAsync(&d, c->d_sims + sim, sizeof(SimDev), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 8; ++k) { c->h_sims[sim].p[k] = d.p[k]; c->sims[sim].p[k] = d.p[k]; }
  return CBM_OK;
}

int cbm_get_params(cbm_ctx* c, int32_t sim, double* p, int32_t n) {
  if (!c || !p || sim < 0 || sim >= c->S || n < 1 || n > 8) return CBM_ERR_ARG;
  int rc = params_from_device(c, sim);
  if (rc) return rc;
  for (int k = 0; k < n; ++k) p[k] = c->h_sims[sim].p[k];
  return CBM_OK;
}

int cbm_set_params(cbm_ctx* c, int32_t sim, const double* p, int32_t n) {
  if (!c || !p || sim < 0 || sim >= c->S || n < 1 || n > 8) return CBM_ERR_ARG;
  {
    int rc = params_from_device(c, sim);   // the entries past n keep their device values
    if (rc) return rc;
  }
  if (c->clv) {
    double q[8];
    for (int k = 0; k < 8; ++k) q[k] = k < n ? p[k] : c->sims[sim].p[k];
    if (!(law_reach(c->m.law, q) <= c->clv_reach))
      return fail(c, CBM_ERR_CONFIG, "the law's reach grew past the one the cell-linked lists were sized for "
                                     "(create the context with the largest parameters, or walk = 1)");
  }
  for (int q = 0; q < n; ++q) { c->sims[sim].p[q] = p[q]; c->h_sims[sim].p[q] = p[q]; }
  HIPCHK(c, hipMemcpyAsync(c->d_sims + sim, &c->h_sims[sim], sizeof(SimDev), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CBM_OK;
}

// --------------------------------------------------------------- adjoint
// Neighbor (re)build that leaves the rows in place (lambda is indexed by row
// during the adjoint): keys, scan, scatter, then lists whose candidates are
// read through the bucket-ordered row table.  flag NULL = unconditional.
static int rebuild_in_place(cbm_ctx* c, double* const* X, const int32_t* flag,
                            int32_t* list = nullptr, int32_t* list_n = nullptr) {
  RebuildArgs a = rebuild_args(c, flag, nullptr, false);
  for (int d = 0; d < c->dims; ++d) a.pos[d] = X[d];
  a.dead = nullptr;
  HIPCHK(c, launch_keys(c->nblocks, c->block, c->stream, a));
  int rc = run_scan(c, flag, c->count, c->cursor, c->start, c->NB);
  if (rc) return rc;
  HIPCHK(c, launch_scatter(c->nblocks, c->block, c->stream, a, !c->full));
  VerletArgs va;
  std::memset(&va, 0, sizeof(va));
  va.rb = a;
  for (int d = 0; d < c->dims; ++d) { va.apos[d] = X[d]; va.xref[d] = c->xref[d]; }
  va.aslot = (const int32_t*)c->col[C_SLOT].cur();
  va.list = list ? list : c->list; va.list_n = list_n ? list_n : c->list_n;
  va.ldl = c->C; va.cap = c->cap;
  va.overflow = c->flags + 2;
  va.rl2 = c->list_r2;
  va.lds_r = verlet_lds_r(c);
  va.build = 1;
  va.nopermute = 1;
  HIPCHK(c, launch_verlet(c->dims, c->nblocks, c->block, c->stream, va));
  return CBM_OK;
}

// One overdamped-Euler step X -> Y with the displacement check against the
// list snapshot, then a conditional in-place rebuild at Y.
static int adj_forward_step(cbm_ctx* c, double* const* X, double* const* Y, int32_t* flag) {
  StageArgs a = stage_base(c);
  for (int v = 0; v < c->dims; ++v) { a.S[v] = X[v]; a.U[v] = X[v]; a.out[v] = Y[v]; }
  a.flag = flag;
  a.disp_check = 1;
  a.kind = ST_EULER;
  HIPCHK(c, stage_launch(c, a));
  int rc = rebuild_in_place(c, Y, flag);
  if (rc) return rc;
  HIPCHK(c, launch_set_flag(c->stream, flag, 0));
  return CBM_OK;
}

// One overdamped-Euler step X -> Y on the given lists, optionally with the
// displacement check into flag.
static int adj_stage(cbm_ctx* c, double* const* X, double* const* Y, const int32_t* list,
                     const int32_t* list_n, int32_t* flag) {
  StageArgs a = stage_base(c);
  for (int v = 0; v < c->dims; ++v) { a.S[v] = X[v]; a.U[v] = X[v]; a.out[v] = Y[v]; }
  a.list = list; a.list_n = list_n;
  a.flag = flag;
  a.disp_check = flag ? 1 : 0;
  a.kind = ST_EULER;
  HIPCHK(c, stage_launch(c, a));
  return CBM_OK;
}

// Status of the snapshot forward pass when its memory budget runs out.
constexpr int ADJ_NO_ROOM = 1;

int cbm_adjoint(cbm_ctx* c, const cbm_loss_desc* L, double* grad, double* loss) {
  if (!c || !L || !grad || !loss) return CBM_ERR_ARG;
  const cbm_model_desc& m = c->m;
  if (m.integrator != CBM_EULER || m.law != CBM_LAW_REP_ADH || !c->verlet || c->vtime || c->rules ||
      c->medium || c->lagged || m.drift != 0.0 || m.user_de > 0 || c->model_fn)
    return fail(c, CBM_ERR_CONFIG, "the adjoint covers overdamped Euler with the REP_ADH law on Verlet lists (no rules, medium, lag, drift or model scope)");
  const int64_t T = L->n_steps, K = L->checkpoint_every;
  const int S = L->n_saves;
  if (T < 1 || K < 1 || S < 1 || !L->save_steps) return fail(c, CBM_ERR_ARG, "bad loss description");
  std::vector<int> save_at((size_t)T + 1, -1);
  for (int q = 0; q < S; ++q) {
    const int64_t st = L->save_steps[q];
    if (st < 1 || st > T || (q && st <= L->save_steps[q - 1])) return fail(c, CBM_ERR_ARG, "save steps must be ascending in (0, n_steps]");
    save_at[(size_t)st] = q;
  }
  int rc;
  if (!c->structure_valid && (rc = force_rebuild(c))) return rc;
  const int D = c->dims;
  // per-sim sizes; the targets stay in the caller's layout (per sim, per
  // save, per axis, slot order) at per-sim bases
  std::vector<int64_t> n(c->S), base(c->S + 1, 0);
  std::vector<double> invn(c->S);
  for (int s = 0; s < c->S; ++s) {
    if ((rc = cbm_get_n(c, s, &n[s]))) return rc;
    if (n[s] < 1) return fail(c, CBM_ERR_ARG, "every simulation needs agents");
    base[s + 1] = base[s] + (int64_t)S * D * n[s];
    invn[s] = 1.0 / ((double)S * (double)n[s]);
  }
  if (!L->target && (c->adj_tgt_saves != S || c->adj_tgt_n != n))
    return fail(c, CBM_ERR_ARG, "no targets: the previous call's targets have another shape");
  const int64_t nck = (T + K - 1) / K;
  const size_t state = (size_t)D * c->C;
  hipError_t e = hipSuccess;
  auto grow = [&](cbm_ctx::DevBuf& b, size_t bytes) {
    if (e != hipSuccess || b.bytes >= bytes) return;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr; b.bytes = 0;
    e = hipMalloc(&b.p, bytes ? bytes : 8);
    if (e == hipSuccess) b.bytes = bytes;
  };
  grow(c->adj_tgt, sizeof(double) * (size_t)base[c->S]);
  grow(c->adj_tbase, sizeof(int64_t) * c->S);
  grow(c->adj_invn, sizeof(double) * c->S);
  grow(c->adj_ck, sizeof(double) * state * nck);
  grow(c->adj_seg, sizeof(double) * state * K);
  grow(c->adj_lam, sizeof(double) * state * 2);
  grow(c->adj_gpart, sizeof(double) * 3 * c->nblocks);
  grow(c->adj_lpart, sizeof(double) * c->nblocks);
  double* d_tgt = (double*)c->adj_tgt.p;
  double* d_invn = (double*)c->adj_invn.p;
  int64_t* d_base = (int64_t*)c->adj_tbase.p;
  double* ck = (double*)c->adj_ck.p;
  double* seg = (double*)c->adj_seg.p;
  double* lam = (double*)c->adj_lam.p;
  double* gpart = (double*)c->adj_gpart.p;
  double* lpart = (double*)c->adj_lpart.p;
  if (e == hipSuccess && L->target) {
    c->adj_tgt_saves = 0;
    e = hipMemcpyAsync(d_tgt, L->target, sizeof(double) * (size_t)base[c->S], hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) { c->adj_tgt_saves = S; c->adj_tgt_n = n; }
  }
  if (e == hipSuccess) e = hipMemcpyAsync(d_invn, invn.data(), sizeof(double) * c->S, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_base, base.data(), sizeof(int64_t) * c->S, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(gpart, 0, sizeof(double) * 3 * c->nblocks, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(lpart, 0, sizeof(double) * c->nblocks, c->stream);
  if (e != hipSuccess) return fail(c, CBM_ERR_HIP, std::string("adjoint buffers: ") + hipGetErrorString(e));
  auto ptrs = [&](double* b0, double** out3) { for (int a = 0; a < 3; ++a) out3[a] = a < D ? b0 + (size_t)a * c->C : nullptr; };
  int32_t* flag = c->flags + 5;
  double* X0[3];
  for (int a = 0; a < D; ++a) X0[a] = statecol(c, a, false);
  double* P[2][3];
  for (int a = 0; a < D; ++a) { P[0][a] = statecol(c, a, false); P[1][a] = statecol(c, a, true); }
  int cur = 0;
  LossArgs la;
  std::memset(&la, 0, sizeof(la));
  la.g = geom(c); la.slot = (const int32_t*)c->col[C_SLOT].cur(); la.target = d_tgt; la.tgt_base = d_base;
  la.dims = D; la.lpart = lpart;
  auto fwd_fail = [&](int code) { return code; };
  // ---- list snapshots.  Every list the forward pass builds is kept (while
  // half of the free device memory holds them), so the recompute and the
  // backward sweep run on exactly the forward's lists: no displacement
  // checks, no rebuilds.  Without room the lists are re-validated, and
  // rebuilt where needed, along the backward sweep instead.  The snapshot
  // buffers stay in the context for the next call.
  std::vector<int32_t*>& snap_l = c->snap_l;
  std::vector<int32_t*>& snap_n = c->snap_n;
  size_t used = 0;   // snapshots of this call
  std::vector<int32_t> epoch((size_t)T, 0);
  auto free_snaps = [&]() {
    for (int32_t* p : snap_l) (void)hipFree(p);
    for (int32_t* p : snap_n) (void)hipFree(p);
    snap_l.clear(); snap_n.clear();
  };
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  const size_t snap_bytes = sizeof(int32_t) * ((size_t)c->cap + 1) * (size_t)c->C;
  size_t budget = free_b / 2 + snap_l.size() * snap_bytes;
  if (const char* mb = std::getenv("CBM_ADJ_SNAPSHOT_MB")) budget = (size_t)std::strtoull(mb, nullptr, 10) << 20;

  auto new_snap = [&]() -> bool {
    if ((used + 1) * snap_bytes > budget) return false;
    if (used == snap_l.size()) {
      int32_t *l = nullptr, *nn = nullptr;
      if (hipMalloc((void**)&l, sizeof(int32_t) * (size_t)c->cap * (size_t)c->C) != hipSuccess) return false;
      if (hipMalloc((void**)&nn, sizeof(int32_t) * (size_t)c->C) != hipSuccess) { (void)hipFree(l); return false; }
      snap_l.push_back(l); snap_n.push_back(nn);
    }
    ++used;
    return true;
  };
  AdjointArgs aa;
  std::memset(&aa, 0, sizeof(aa));
  aa.g = geom(c); aa.ldl = c->C;
  aa.slot = (const int32_t*)c->col[C_SLOT].cur(); aa.target = d_tgt; aa.tgt_base = d_base;
  aa.periodic = m.periodic; aa.dt = m.dt; aa.inv_n = d_invn; aa.gpart = gpart;
  double* LAM[2][3];
  ptrs(lam, LAM[0]);
  ptrs(lam + state, LAM[1]);
  bool moved = false;   // P[0] no longer holds the initial state (it is in ck[0])
  auto run = [&](bool snaps) -> int {
    int rc2;
    cur = 0;
    int e = 0;
    HIPCHK(c, launch_set_flag(c->stream, flag, 0));
    // ---- forward with checkpoints, rows held in place
    if (snaps) {
      if (!new_snap()) return ADJ_NO_ROOM;
      if ((rc2 = rebuild_in_place(c, P[0], nullptr, snap_l[0], snap_n[0]))) return rc2;
    } else if ((rc2 = rebuild_in_place(c, P[0], nullptr))) {
      return rc2;
    }
    for (int64_t t = 0; t < T; ++t) {
      if (t % K == 0) {
        double* dst[3];
        ptrs(ck + state * (size_t)(t / K), dst);
        for (int a = 0; a < D; ++a)
          if (hipMemcpyAsync(dst[a], P[cur][a], sizeof(double) * c->C, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return fail(c, CBM_ERR_HIP, "checkpoint copy");
      }
      if (snaps) {
        epoch[(size_t)t] = e;
        if ((rc2 = adj_stage(c, P[cur], P[1 - cur], snap_l[(size_t)e], snap_n[(size_t)e], flag))) return rc2;
      } else if ((rc2 = adj_forward_step(c, P[cur], P[1 - cur], flag))) {
        return rc2;
      }
      cur ^= 1;
      moved = true;
      if (save_at[(size_t)t + 1] >= 0) {
        for (int a = 0; a < D; ++a) la.X[a] = P[cur][a];
        la.save = save_at[(size_t)t + 1];
        if (launch_loss(c->nblocks, c->block, c->stream, la) != hipSuccess) return fail(c, CBM_ERR_HIP, "loss");
      }
      if (snaps && t + 1 < T) {
        int32_t f = 0;
        HIPCHK(c, hipMemcpyAsync(&f, flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (f) {
          if (!new_snap()) return ADJ_NO_ROOM;
          ++e;
          HIPCHK(c, launch_set_flag(c->stream, flag, 0));
          if ((rc2 = rebuild_in_place(c, P[cur], nullptr, snap_l[(size_t)e], snap_n[(size_t)e]))) return rc2;
        }
      }
    }
    // ---- backward sweep, recomputing each checkpoint segment
    int lc = 0;
    aa.list = c->list; aa.list_n = c->list_n;
    for (int a = 0; a < D; ++a) aa.X[a] = P[cur][a];
    for (int a = 0; a < D; ++a) aa.lam_out[a] = LAM[lc][a];
    aa.save = save_at[(size_t)T];
    if (launch_adjoint_seed(D, c->nblocks, c->block, c->stream, aa) != hipSuccess) return fail(c, CBM_ERR_HIP, "adjoint seed");
    for (int64_t sgm = nck - 1; sgm >= 0; --sgm) {
      const int64_t t0 = sgm * K, t1 = std::min(t0 + K, T);
      double* Xk[64][3];
      const int64_t len = t1 - t0;
      for (int64_t k = 0; k < len; ++k) ptrs(seg + state * (size_t)k, Xk[k]);
      double* ckp[3];
      ptrs(ck + state * (size_t)sgm, ckp);
      for (int a = 0; a < D; ++a)
        if (hipMemcpyAsync(Xk[0][a], ckp[a], sizeof(double) * c->C, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
          return fail(c, CBM_ERR_HIP, "checkpoint restore");
      // Without snapshots the list only has to be valid for the positions it
      // is used with (every agent within the rebuild displacement of its
      // list snapshot), in either time direction: check, and rebuild only
      // when that fails.
      auto ensure_valid = [&](double* const* X) -> int {
        StageArgs ca = stage_base(c);
        for (int a = 0; a < D; ++a) ca.S[a] = X[a];
        ca.flag = flag;
        if (launch_disp_check(D, c->nblocks, c->block, c->stream, ca) != hipSuccess) return fail(c, CBM_ERR_HIP, "disp check");
        int r3 = rebuild_in_place(c, X, flag);
        if (r3) return r3;
        HIPCHK(c, launch_set_flag(c->stream, flag, 0));
        return CBM_OK;
      };
      if (snaps) {
        for (int64_t k = 0; k + 1 < len; ++k) {
          const size_t ep = (size_t)epoch[(size_t)(t0 + k)];
          if ((rc2 = adj_stage(c, Xk[k], Xk[k + 1], snap_l[ep], snap_n[ep], nullptr))) return rc2;
        }
      } else {
        if ((rc2 = ensure_valid(Xk[0]))) return rc2;
        for (int64_t k = 0; k + 1 < len; ++k)
          if ((rc2 = adj_forward_step(c, Xk[k], Xk[k + 1], flag))) return rc2;
      }
      for (int64_t t = t1 - 1; t >= t0; --t) {
        double* const* X = Xk[t - t0];
        if (snaps) {
          aa.list = snap_l[(size_t)epoch[(size_t)t]]; aa.list_n = snap_n[(size_t)epoch[(size_t)t]];
        } else if (t < t1 - 1 && (rc2 = ensure_valid(X))) {
          return rc2;
        }
        for (int a = 0; a < D; ++a) { aa.X[a] = X[a]; aa.lam[a] = LAM[lc][a]; aa.lam_out[a] = LAM[1 - lc][a]; }
        aa.save = (t >= 1) ? save_at[(size_t)t] : -1;
        if (launch_adjoint_step(D, c->nblocks, c->block, c->stream, aa) != hipSuccess) return fail(c, CBM_ERR_HIP, "adjoint step");
        lc ^= 1;
      }
    }
    return CBM_OK;
  };
  if (K > 64) return fwd_fail(fail(c, CBM_ERR_ARG, "checkpoint_every must be <= 64"));
  rc = run(true);
  if (rc == ADJ_NO_ROOM) {
    // start again without snapshots from the initial state
    free_snaps();
    used = 0;
    if (moved)
      for (int a = 0; a < D && e == hipSuccess; ++a)
        e = hipMemcpyAsync(P[0][a], ck + (size_t)a * c->C, sizeof(double) * c->C, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(lpart, 0, sizeof(double) * c->nblocks, c->stream);
    if (e != hipSuccess) return fwd_fail(fail(c, CBM_ERR_HIP, "adjoint restart"));
    rc = run(false);
  }
  if (rc) return fwd_fail(rc);
  // ---- per-simulation totals (blocks summed in order: deterministic)
  std::vector<double> hg((size_t)3 * c->nblocks), hl((size_t)c->nblocks);
  e = hipMemcpyAsync(hg.data(), gpart, sizeof(double) * hg.size(), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(hl.data(), lpart, sizeof(double) * hl.size(), hipMemcpyDeviceToHost, c->stream);
  // the community is left at its initial state
  for (int a = 0; a < D && e == hipSuccess; ++a) e = hipMemcpyAsync(X0[a], ck + (size_t)a * c->C, sizeof(double) * c->C, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, CBM_ERR_HIP, std::string("adjoint readback: ") + hipGetErrorString(e));
  for (int s = 0; s < c->S; ++s) { grad[3 * s] = grad[3 * s + 1] = grad[3 * s + 2] = 0.0; loss[s] = 0.0; }
  for (int b = 0; b < c->nblocks; ++b) {
    const int s = c->h_block_sim[b];
    for (int q = 0; q < 3; ++q) grad[3 * s + q] += hg[(size_t)3 * b + q];
    loss[s] += hl[(size_t)b];
  }
  for (int s = 0; s < c->S; ++s) loss[s] *= invn[s];
  c->structure_valid = false;
  return check_flags(c);
}

int cbm_set_block(cbm_ctx* c, int32_t block) {
  if (!c || (block != 64 && block != 128 && block != 256)) return CBM_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->block = block;
  return setup_blocks(c);
}

// Key of the tuning cache: everything that changes the stage kernel's work
// per row (the model family and the population scale).
static std::string tune_key(cbm_ctx* c) {
  int64_t n = 0;
  for (int s = 0; s < c->S; ++s) n += c->sims[s].n;
  int lg = 0;
  while (((int64_t)1 << (lg + 1)) <= n) ++lg;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "gfx950 d%d i%d nb%d law%d per%d lag%d rule%d sims%d n2^%d skin%.3g edge%.3g",
                c->dims, c->m.integrator, c->m.neighbors, c->m.law, c->m.periodic, c->lagged ? 1 : 0,
                c->m.rule, c->S, lg, c->m.skin, c->m.cell_edge);
  return buf;
}

int cbm_autotune(cbm_ctx* c, const char* cache_path, int32_t* block) {
  if (!c) return CBM_ERR_ARG;
  const std::string key = tune_key(c);
  if (cache_path) {   // "<key>\t<block>" lines
    FILE* f = std::fopen(cache_path, "r");
    if (f) {
      char line[512];
      int found = 0;
      while (std::fgets(line, sizeof(line), f)) {
        char* tab = std::strchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        if (key == line) found = std::atoi(tab + 1);
      }
      std::fclose(f);
      if (found == 64 || found == 128 || found == 256) {
        int rc = cbm_set_block(c, found);
        if (block) *block = c->block;
        return rc;
      }
    }
  }
  if (!c->structure_valid) { int rc = force_rebuild(c); if (rc) return rc; }
  // Time the first integrator stage written into scratch (the alternate
  // state buffers, or the stage buffer): the same gather and pair work as a
  // step, and no state changes.
  StageArgs a = stage_base(c);
  a.write_fnew = 0;
  for (int v = 0; v < c->nvar; ++v) {
    a.S[v] = statecol(c, v, false);
    a.U[v] = statecol(c, v, false);
    a.out[v] = c->h[0][v] ? c->h[0][v] : statecol(c, v, true);
  }
  a.kind = ST_HEUN1;
  const int cand[3] = {64, 128, 256};
  float best_ms = 1e30f;
  int best = c->block;
  for (int bs : cand) {
    int rc = cbm_set_block(c, bs);
    if (rc) return rc;
    a.g = geom(c);
    for (int w = 0; w < 2; ++w) HIPCHK(c, stage_launch(c, a));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < 10; ++r) HIPCHK(c, stage_launch(c, a));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (ms < best_ms) { best_ms = ms; best = bs; }
  }
  int rc = cbm_set_block(c, best);
  if (rc) return rc;
  if (cache_path) {
    FILE* f = std::fopen(cache_path, "a");
    if (f) { std::fprintf(f, "%s\t%d\n", key.c_str(), best); std::fclose(f); }
  }
  if (block) *block = best;
  return CBM_OK;
}

int cbm_profile(cbm_ctx* c, int64_t nsteps, double* avg_ms, int64_t* launches, int32_t ncat) {
  if (!c || nsteps < 1 || !avg_ms || !launches || ncat < PROF_NCAT) return CBM_ERR_ARG;
  if (!c->structure_valid) { int rc = force_rebuild(c); if (rc) return rc; }
  size_t need = (size_t)nsteps * 12 * 2;
  while (c->prof_ev.size() < need) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c->prof_ev.push_back(e);
  }
  c->prof_cat.assign(c->prof_ev.size() / 2 + 1, 0);
  c->prof_used = 0;
  c->profiling = true;
  int rc = CBM_OK;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int64_t k = 0; k < nsteps && rc == CBM_OK; ++k) rc = step_once(c);
  c->profiling = false;
  if (rc) return rc;
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int q = 0; q < ncat; ++q) { avg_ms[q] = 0.0; launches[q] = 0; }
  for (size_t p = 0; p + 1 < c->prof_used; p += 2) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->prof_ev[p], c->prof_ev[p + 1]));
    int cat = c->prof_cat[p / 2];
    avg_ms[cat] += ms;
    launches[cat] += 1;
  }
  for (int q = 0; q < PROF_NCAT; ++q) if (q != PROF_STEP && launches[q]) avg_ms[q] /= (double)launches[q];
  float tot = 0.f;
  HIPCHK(c, hipEventElapsedTime(&tot, c->ev0, c->ev1));
  avg_ms[PROF_STEP] = tot / (double)nsteps;
  launches[PROF_STEP] = nsteps;
  return check_flags(c);
}

int cbm_last_timing(cbm_ctx* c, double* ms_total, double* ms_force, int64_t* launches) {
  if (!c) return CBM_ERR_ARG;
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (ms_total) *ms_total = ms;
  if (ms_force) *ms_force = ms;
  if (launches) *launches = c->last_launches;
  return CBM_OK;
}

}  // extern "C"

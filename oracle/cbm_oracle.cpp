// cbm_oracle.cpp -- CPU restatement of the CellBasedModels.jl per-step agent
// update.  TEST INFRASTRUCTURE ONLY: this is the parity checker for the HIP
// library (tests/, __graft_entry__.smoke and bench.py's cpu_baseline leg);
// no product code path may load or call it.
//
// Everything is kept in the reference's logical slot order and follows the
// reference's serial CPU algorithms:
//   cell binning / counting sort      src/neighbors.jl:587-979
//   cellPos / cellPosNeigh            src/neighbors.jl:636-708
//   Full / Verlet list build          src/neighbors.jl:27-57, 92-165
//   Verlet displacement trigger       src/neighbors.jl:367-570
//   VerletTime trigger                src/neighbors.jl:240-281
//   euclidean                         src/metrics.jl:34-36
//   cellInMesh                        src/metrics.jl:18-20
//   Euler/Heun/RK4/EM                 src/integrators.jl:442-734
//   read-old / write-new (lag)        src/AgentStructure/agentStructure.jl:352-453
//   @addAgent / @removeAgent          src/AgentStructure/functionRule.jl:9-111
//   update! compaction + commit       src/CommunityStructure/update.jl:15-303
//   step! order                       src/CommunityStructure/step.jl:96-107
//   @∂2 medium stencil                src/AgentStructure/macros.jl:203-257
//   @mediumInside / @mediumBorder     src/AgentStructure/macros.jl:259-311
//   rods2D / rodForces                src/models/rods.jl:21-121
//   modelStepRule! / modelStepDE!     src/CommunityStructure/step.jl:21-30, 66-75, 96-107
//   rodIntersection and helpers       src/metrics.jl:138-246
// with the spec decisions the GPU path shares (exact floor quotients,
// periodic minimum image, canonical ascending compaction order, Philox noise,
// zero-flux ghost-mirror medium, border nodes assigned from the step-start
// field, d>0 guard on pair laws).
//
// Compiled with -ffp-contract=off: every expression below is evaluated in
// the written order with IEEE double rounding.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/cbm_hip.h"

namespace {

// ------------------------------------------------------------------ Philox
// Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3" (SC'11).  Multipliers and Weyl increments from the paper.
struct U4 { uint32_t v[4]; };
static inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32); lo = (uint32_t)p;
}
static U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c.v[0], hi0, lo0);
    mulhilo(0xCD9E8D57u, c.v[2], hi1, lo1);
    U4 o;
    o.v[0] = hi1 ^ c.v[1] ^ k0; o.v[1] = lo1;
    o.v[2] = hi0 ^ c.v[3] ^ k1; o.v[3] = lo0;
    c = o;
  }
  return c;
}
// 53-bit uniform in [0,1) from two 32-bit words.
static inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
static U4 draw(uint64_t seed, int64_t step, int64_t id, int32_t sim, uint32_t purpose) {
  U4 c; c.v[0] = (uint32_t)step; c.v[1] = (uint32_t)id; c.v[2] = (uint32_t)sim; c.v[3] = purpose;
  return philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}
// Box-Muller (Box & Muller 1958) on one Philox block -> two N(0,1).
static void normal2(const U4& r, double& z0, double& z1) {
  double u1 = 1.0 - u53(r.v[0], r.v[1]);
  double u2 = u53(r.v[2], r.v[3]);
  double rad = std::sqrt(-2.0 * std::log(u1));
  double th = 6.283185307179586 * u2;
  z0 = rad * std::cos(th);
  z1 = rad * std::sin(th);
}

// floor(a/b) of the exact real quotient, a >= 0, b > 0 (the reference's
// Int(a ÷ b) of src/neighbors.jl:636 and src/metrics.jl:19).
static inline int64_t floor_div(double a, double b) {
  double k = std::floor(a / b);
  double r = std::fma(-k, b, a);
  if (r < 0.0) k -= 1.0; else if (r >= b) k += 1.0;
  return (int64_t)k;
}

// ------------------------------------------------------------ dual numbers
// Forward-mode dual number with 3 tangents (k, r_eq, eps), used only for the
// adjoint cross-check (the reference Manifest pins ForwardDiff 0.10.36,
// which is what this stands in for).
struct D3 {
  double v, g[3];
  D3() : v(0) { g[0] = g[1] = g[2] = 0; }
  D3(double x) : v(x) { g[0] = g[1] = g[2] = 0; }
};
static inline D3 operator+(const D3& a, const D3& b) { D3 r; r.v = a.v + b.v; for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] + b.g[i]; return r; }
static inline D3 operator-(const D3& a, const D3& b) { D3 r; r.v = a.v - b.v; for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] - b.g[i]; return r; }
static inline D3 operator*(const D3& a, const D3& b) { D3 r; r.v = a.v * b.v; for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i]; return r; }
static inline D3 operator/(const D3& a, const D3& b) { D3 r; r.v = a.v / b.v; for (int i = 0; i < 3; ++i) r.g[i] = (a.g[i] * b.v - a.v * b.g[i]) / (b.v * b.v); return r; }
static inline D3 sqrt(const D3& a) { D3 r; r.v = std::sqrt(a.v); for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] / (2.0 * r.v); return r; }
static inline double val(const D3& a) { return a.v; }
static inline double val(double a) { return a; }
using std::sqrt;

// ------------------------------------------------------------------- model
struct Grid {
  int dims = 1;
  bool periodic = false;
  int64_t n[3] = {1, 1, 1};
  double edge[3] = {1, 1, 1};
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int64_t total() const { return n[0] * n[1] * n[2]; }
  // cellPos, src/neighbors.jl:636-638 (padded, clamped) or the periodic grid.
  bool geometric = false;   // Verlet builds: bin by position, then clamp
  int64_t axis_cell(int a, double x) const {
    if (geometric) {
      if (x < lo[a]) return 0;
      int64_t c = floor_div(x - lo[a], edge[a]) + (periodic ? 0 : 1);
      return c > n[a] - 1 ? n[a] - 1 : c;
    }
    if (periodic) {
      int64_t c = floor_div(x - lo[a], edge[a]);
      if (c < 0) c = 0;
      if (c > n[a] - 1) c = n[a] - 1;
      return c;
    }
    if (x > hi[a]) return n[a] - 1;
    if (x < lo[a]) return 0;
    return floor_div(x - lo[a], edge[a]) + 1;
  }
};

struct Orc {
  cbm_model_desc m;
  cbm_sim_desc s;
  int dims = 1;
  int64_t N = 0, nMax = 0, idMax = 0, step = 0;
  double t = 0;
  bool growth = false;
  bool inertial = false;                      // CBM_LAW_SOFT_SPHERES: velocities are DE variables
  bool rods = false;                          // CBM_LAW_RODS: x, y, vx, vy, theta, omega are DE variables
  int32_t sim = 0;                            // its index in a batched context: the noise streams' key
  std::vector<double> TH, OM, rodD, rodL;     // rods: angle, its rate, diameter, length
  std::vector<double> rodW, rodP;             // rods: torque and pressure of the last evaluation
  std::vector<double> LT;                     // rods with CBM_RULE_RODS_GROWTH: division length lTarget
  bool rods_growth = false;                   // ... and l is the seventh DE variable
  double noise_slope = 0.0;                   // state-dependent agentSDE g_a(x) = noise + slope x_a (positions)
  std::vector<double> X[3], F[3], Fnew[3];
  std::vector<double> V[3], rad, mass;        // velocities, radius, mass (soft spheres)
  std::vector<int64_t> fate;                  // cellFate, 1-based
  std::vector<double> vol, gr, vb, grNew, vbNew;
  std::vector<int64_t> id, lf;
  std::vector<double> U[8];                   // user agent parameters U0..U7
  std::vector<char> alive;
  void* rule_fn = nullptr;                    // CBM_RULE_CUSTOM body (orc_rule_fn)
  std::vector<double> ruleX[3], ruleVol, ruleU[8];   // the rule's neighbor view: values at the step's start
  void* medium_fn = nullptr;                  // user mediumODE body (orc_medium_fn), else built-in
  void* medium_assign = nullptr;              // its assignment branches (orc_medium_assign_fn), optional
  void* model_rule = nullptr;                 // model scope (orc_model_rule_fn, orc_model_ode_fn,
  void* model_ode = nullptr;                  // orc_model_noise_fn) and its DE parameters
  void* model_noise = nullptr;
  unsigned model_de = 0;
  void* law_pair = nullptr;                   // a general agentODE body (orc_law_pair_fn, orc_law_ode_fn)
  void* law_ode = nullptr;
  int law_sums = 0;
  // neighbor structures
  Grid g;
  std::vector<int64_t> cellOf, cellStart, cellCount, cellAgents;
  std::vector<int32_t> nbN;
  std::vector<int64_t> nbList;
  std::vector<double> Xref[3];
  double tNext = 0;
  int64_t rebuilds = 0;
  bool structure_valid = false;
  // medium
  int64_t mn[3] = {1, 1, 1};
  double mdx[3] = {1, 1, 1};
  std::vector<double> c, cnew;
  std::vector<double> cinit;                  // DGADI: the field at creation / last orc_set_medium
  std::vector<int64_t> medAdd;                // the rules' additions per node (32.32 fixed point)
  std::vector<int32_t> nodeCount;
  int err = 0;
  std::string msg;

  bool is_verlet() const {
    return m.neighbors == CBM_VERLET_DISPLACEMENT || m.neighbors == CBM_VERLET_TIME || m.neighbors == CBM_CLVD;
  }
  double rebuild_thr() const { return m.rebuild_disp > 0 ? m.rebuild_disp : m.skin / 2; }

  // Minimum-image component (periodic boxes only).
  inline double mi(int a, double d) const {
    if (!m.periodic) return d;
    double L = s.box_hi[a] - s.box_lo[a];
    if (d > 0.5 * L) d -= L; else if (d < -0.5 * L) d += L;
    return d;
  }
  template <class T> inline T mi(int a, T d) const {
    if (!m.periodic) return d;
    double L = s.box_hi[a] - s.box_lo[a];
    if (val(d) > 0.5 * L) d = d - T(L); else if (val(d) < -0.5 * L) d = d + T(L);
    return d;
  }
};

static void set_grid(Orc& o, Grid& g, double edge) {
  g.dims = o.dims;
  g.periodic = o.m.periodic != 0;
  for (int a = 0; a < 3; ++a) { g.n[a] = 1; g.edge[a] = 1; g.lo[a] = 0; g.hi[a] = 0; }
  for (int a = 0; a < o.dims; ++a) {
    g.lo[a] = o.s.box_lo[a]; g.hi[a] = o.s.box_hi[a];
    double L = g.hi[a] - g.lo[a];
    if (g.periodic) {
      g.n[a] = (int64_t)std::floor(L / edge);
      if (g.n[a] < 1) g.n[a] = 1;
      g.edge[a] = L / (double)g.n[a];
    } else {
      // nCells_ = ceil(L/edge + 2), src/neighbors.jl:604
      g.n[a] = (int64_t)std::ceil(L / edge + 2.0);
      g.edge[a] = edge;
    }
  }
}

static int64_t cell_of(const Orc& o, const Grid& g, int64_t i) {
  int64_t c[3] = {0, 0, 0};
  for (int a = 0; a < o.dims; ++a) c[a] = g.axis_cell(a, o.X[a][i]);
  return c[0] + g.n[0] * (c[1] + g.n[1] * c[2]);
}

// neighborsCellLinked{d}CPU!: zero counts, assignCells (serial), cumsum,
// sortAgentsInCells (serial, ascending agent index within a cell).
static void bin(Orc& o, const Grid& g) {
  int64_t nc = g.total();
  o.cellCount.assign(nc, 0);
  o.cellStart.assign(nc + 1, 0);
  o.cellOf.assign(o.N, 0);
  o.cellAgents.assign(o.N, 0);
  for (int64_t i = 0; i < o.N; ++i) { o.cellOf[i] = cell_of(o, g, i); o.cellCount[o.cellOf[i]]++; }
  for (int64_t k = 0; k < nc; ++k) o.cellStart[k + 1] = o.cellStart[k] + o.cellCount[k];
  std::vector<int64_t> cur(o.cellStart.begin(), o.cellStart.end() - 1);
  for (int64_t i = 0; i < o.N; ++i) o.cellAgents[cur[o.cellOf[i]]++] = i;
}

// cellPosNeigh, src/neighbors.jl:658-708 (0-based, x fastest; the correct
// nX stride in 2-D -- the reference's py*nY of :683 only differs when nX!=nY)
// or the periodic wrap.  Returns -1 outside a non-periodic grid.
static int64_t neigh_cell(const Grid& g, int64_t cell, int k) {
  int64_t p[3];
  p[0] = cell % g.n[0];
  p[1] = (cell / g.n[0]) % g.n[1];
  p[2] = cell / (g.n[0] * g.n[1]);
  int off[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
  for (int a = 0; a < 3; ++a) {
    if (a >= g.dims) { off[a] = 0; continue; }
    int64_t q = p[a] + off[a];
    if (g.periodic) { q = (q + g.n[a]) % g.n[a]; }
    else if (q < 0 || q >= g.n[a]) return -1;
    p[a] = q;
  }
  return p[0] + g.n[0] * (p[1] + g.n[1] * p[2]);
}

static int n_neigh_cells(const Orc& o) { int r = 1; for (int a = 0; a < o.dims; ++a) r *= 3; return r; }

// Enumerate the candidate neighbors of agent i in the order the reference
// loops visit them (neighborsLoop, src/neighbors.jl:65-76, 715-747, 37).
template <class Fn>
static void for_candidates(const Orc& o, int64_t i, Fn fn) {
  if (o.m.neighbors == CBM_FULL) {
    for (int64_t j = 0; j < o.N; ++j) if (j != i) fn(j);
  } else if (o.m.neighbors == CBM_CELL_LINKED) {
    int nk = n_neigh_cells(o);
    for (int k = 0; k < nk; ++k) {
      int64_t c = neigh_cell(o.g, o.cellOf[i], k);
      if (c < 0) continue;
      for (int64_t q = o.cellStart[c]; q < o.cellStart[c + 1]; ++q) {
        int64_t j = o.cellAgents[q];
        if (j != i) fn(j);
      }
    }
  } else {
    int32_t cnt = o.nbN[i];
    for (int32_t q = 0; q < cnt; ++q) fn(o.nbList[(int64_t)q * o.nMax + i]);
  }
}

template <class T>
static inline T dist(const Orc& o, const T* dx) {
  // euclidean, src/metrics.jl:34-36: sqrt((x1-x2)^2+(y1-y2)^2+(z1-z2)^2)
  T s = dx[0] * dx[0];
  if (o.dims > 1) s = s + dx[1] * dx[1];
  if (o.dims > 2) s = s + dx[2] * dx[2];
  return sqrt(s);
}

// Pair magnitude s(d) of the built-in laws; returns false when the pair does
// not interact.  p = law parameters (possibly duals).
template <class T>
static inline bool law_s(int law, const T* p, const T& d, T& s) {
  if (!(val(d) > 0.0)) return false;
  if (law == CBM_LAW_LINEAR) {
    if (val(d) < val(p[1])) { s = p[0] * (p[1] - d); return true; }
    return false;
  }
  if (law == CBM_LAW_REP_ADH) {
    // c = 2 r_eq, dc = f r_eq, a4 = 4 eps / (dc - c)^2 (well of depth eps)
    T c = T(2.0) * p[1];
    T dc = p[3] * p[1];
    if (val(d) < val(c)) { s = p[0] * (c - d); return true; }
    if (val(d) < val(dc)) {
      T w = dc - c;
      T a4 = (T(4.0) * p[2]) / (w * w);
      s = T(0.0) - a4 * (d - c) * (dc - d);
      return true;
    }
    return false;
  }
  if (law == CBM_LAW_AGGREGATION) {
    if (val(d) < val(p[0])) { s = p[1] * (p[0] - d); return true; }
    if (val(d) < val(p[2])) { s = T(0.0) - p[3] * (p[2] - d); return true; }
    return false;
  }
  return false;
}

// Soft-sphere pair term of src/models/softSpheres.jl:25-30 (2-D) and :64-71
// (3-D), in the reference's evaluation order: for 0 < d < mu r_ij, the
// force on i along axis a gains ((f0 (r_ij/d - 1)) (mu r_ij/d - 1) dx_a) / d.
static inline bool soft_s(const double* p, double d, double ri, double rj, int64_t fi, int64_t fj, double& s) {
  const double rij = ri + rj;
  if (!(d < p[1] * rij && d > 0)) return false;
  const double f0 = p[2 + 2 * (fi - 1) + (fj - 1)];
  s = f0 * (rij / d - 1) * (p[1] * rij / d - 1);
  return true;
}

// ------------------------------------------------- general agentODE bodies
// An agent as a general agentODE body sees it (functionDE.jl:13-116): the
// stage state of its position, of the vol variable and of the user DE
// variables, and the other user parameters as stored.
struct OrcLawAgent {
  int64_t row, id;
  double x[3];
  double u[8];
  double vol;
};
typedef void (*orc_law_pair_fn)(const OrcLawAgent* i, const OrcLawAgent* j, const double* dx, double d, double* s);
typedef void (*orc_law_ode_fn)(const OrcLawAgent* i, const double* s, double* dxdt, double* dudt);

// The body's derivative of agent i at stage state S: the sums over the
// neighbor candidates in loop order, then its equations; du holds the
// positions' and the user DE variables' derivatives at their DE indices
// (the caller adds vol's).
static void body_deriv(const Orc& o, int64_t i, const std::vector<double>* S, double* du) {
  const int d = o.dims;
  const int base = d + (o.growth ? 1 : 0);
  auto agent = [&](int64_t j) {
    OrcLawAgent r;
    r.row = j;
    r.id = o.id[j];
    for (int a = 0; a < 3; ++a) r.x[a] = a < d ? S[a][j] : 0.0;
    for (int k = 0; k < 8; ++k) r.u[k] = k < o.m.user_de ? S[base + k][j] : o.U[k][j];
    r.vol = o.growth ? S[d][j] : 1.0;
    return r;
  };
  const OrcLawAgent me = agent(i);
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for_candidates(o, i, [&](int64_t j) {
    const OrcLawAgent nb = agent(j);
    double dx[3] = {0, 0, 0};
    for (int a = 0; a < d; ++a) dx[a] = o.mi(a, S[a][i] - S[a][j]);
    ((orc_law_pair_fn)o.law_pair)(&me, &nb, dx, dist(o, dx), s);
  });
  double dxdt[3] = {0, 0, 0}, dudt[4] = {0, 0, 0, 0};
  ((orc_law_ode_fn)o.law_ode)(&me, s, dxdt, dudt);
  for (int a = 0; a < d; ++a) {
    double v = dxdt[a];
    if (o.m.drift != 0.0) v = v - o.m.drift * S[a][i];
    du[a] = v;
  }
  for (int k = 0; k < o.m.user_de; ++k) du[base + k] = dudt[k];
}

// ------------------------------------------------------------------ rods
// A rod's axis line: centre (x, y), length l, angle t.
struct Rod { double x, y, l, t; };

// intersection2lines, src/metrics.jl:148-166.
static void line_meet(const Rod& a, const Rod& b, double& px, double& py) {
  const double cxA = a.x - b.x, cyA = a.y - b.y;
  const double nrm = std::cos(a.t) * std::sin(b.t) - std::sin(a.t) * std::cos(b.t);
  if (std::fabs(nrm) > 0.0000001) {
    const double sc = (-std::sin(b.t) * cxA + std::cos(b.t) * cyA) / nrm;
    px = sc * std::cos(a.t) + a.x;
    py = sc * std::sin(a.t) + a.y;
  } else {
    px = (a.x + b.x) / 2;
    py = (a.y + b.y) / 2;
  }
}

// point2line, src/metrics.jl:175-183: the foot of (x1, y1) on the line
// through (x2, y2) at angle t2.
static void foot(double x1, double y1, double x2, double y2, double t2, double& ox, double& oy) {
  const double cx = x2 - x1, cy = y2 - y1;
  ox = (std::sin(t2) * cx - std::cos(t2) * cy) * std::sin(t2) + x1;
  oy = -(std::sin(t2) * cx - std::cos(t2) * cy) * std::cos(t2) + y1;
}

// pointInsideRod, src/metrics.jl:194-202.
static void inside(const Rod& r, double px, double py, double sep, double& ox, double& oy) {
  const double di = std::max(std::sqrt((r.x - px) * (r.x - px) + (r.y - py) * (r.y - py)), 0.00000001);
  const double dxi = (px - r.x) / di, dyi = (py - r.y) / di;
  ox = sep * std::min(di, r.l / 2) * dxi + r.x;
  oy = sep * std::min(di, r.l / 2) * dyi + r.y;
}

// rodIntersection, src/metrics.jl:210-246 (separation 0.99): the closest
// virtual spheres (x1a, y1a) on rod 1 and (x2a, y2a) on rod 2.
static void rod_intersection(const Rod& r1, const Rod& r2, double* out) {
  const double sep = 0.99;
  double x1a = r1.x, y1a = r1.y, x2a = r2.x, y2a = r2.y;
  double px, py;
  line_meet(r1, r2, px, py);
  const double di = std::sqrt((r1.x - px) * (r1.x - px) + (r1.y - py) * (r1.y - py));
  double dj = std::sqrt((r2.x - px) * (r2.x - px) + (r2.y - py) * (r2.y - py));
  const double nrm = std::cos(r1.t) * std::sin(r2.t) - std::sin(r1.t) * std::cos(r2.t);
  if (std::fabs(nrm) < 0.000001) {
    foot(px, py, r1.x, r1.y, r1.t, x1a, y1a);
    inside(r1, x1a, y1a, sep, x1a, y1a);
    foot(px, py, r2.x, r2.y, r2.t, x2a, y2a);
    inside(r2, x2a, y2a, sep, x2a, y2a);
  } else if (di <= r1.l / 2 && dj <= r2.l / 2) {
    inside(r1, px, py, sep, x1a, y1a);
    inside(r2, px, py, sep, x2a, y2a);
  } else if (di <= r1.l / 2 && dj > r2.l / 2) {
    inside(r2, px, py, sep, x2a, y2a);
    foot(x2a, y2a, r1.x, r1.y, r1.t, x1a, y1a);
  } else if (di > r1.l / 2 && dj <= r2.l / 2) {
    inside(r1, px, py, sep, x1a, y1a);
    foot(x1a, y1a, r2.x, r2.y, r2.t, x2a, y2a);
  } else {
    inside(r2, px, py, sep, x2a, y2a);
    foot(x2a, y2a, r1.x, r1.y, r1.t, x1a, y1a);
    dj = std::sqrt((r1.x - x1a) * (r1.x - x1a) + (r1.y - y1a) * (r1.y - y1a));
    if (dj > r1.l / 2) {
      inside(r1, px, py, sep, x1a, y1a);
      double x2b, y2b;
      foot(x1a, y1a, r2.x, r2.y, r2.t, x2b, y2b);
      dj = std::sqrt((r2.x - x2b) * (r2.x - x2b) + (r2.y - y2b) * (r2.y - y2b));
      if (dj < r2.l / 2) { x2a = x2b; y2a = y2b; }
    }
  }
  out[0] = x1a; out[1] = y1a; out[2] = x2a; out[3] = y2a;
}

// rodForces, src/models/rods.jl:21-59.  r = {x, y, d, l, theta, vx, vy, m}
// of each rod, k = {kn, gamma_n, gamma_t, mu_cc}.  Writes {Fx, Fy, W};
// returns -1 when the overlap d - r_ij of a contact is negative (the
// reference's delta^1.5 raises a DomainError there), else 0.
static int rod_forces(const double* r1, const double* r2, const double* k, double* out) {
  out[0] = 0.; out[1] = 0.; out[2] = 0.;
  double q[4];
  rod_intersection(Rod{r1[0], r1[1], r1[3], r1[4]}, Rod{r2[0], r2[1], r2[3], r2[4]}, q);
  const double rij = std::sqrt((q[0] - q[2]) * (q[0] - q[2]) + (q[1] - q[3]) * (q[1] - q[3]));
  if (rij > 0. && rij < (r1[2] + r2[2]) / 2) {
    const double delta = r1[2] - rij;
    if (delta < 0) return -1;
    const double me = r1[7] / 2;
    const double nx = (q[0] - q[2]) / rij, ny = (q[1] - q[3]) / rij;
    const double vx = r1[5] - r2[5], vy = r1[6] - r2[6];
    const double vn = nx * vx + ny * vy;
    const double fn = k[0] * std::pow(delta, 1.5) - k[1] * me * delta * vn;
    const double ft = -std::min(k[2] * me * std::pow(delta, .5), k[3] * fn);
    out[0] = fn * nx + ft * (vx - vn * nx);
    out[1] = fn * ny + ft * (vy - vn * ny);
    out[2] = (q[0] - r1[0]) * out[1] - (q[1] - r1[1]) * out[0];
  }
  return 0;
}

// The rods2D agentODE body (src/models/rods.jl:86-117) for agent i at stage
// state S (x, y, vx, vy, theta, omega): du of the six DE variables, and
// acc = {fx, fy, W, pressure}.  Returns -1 on a negative overlap.
static int rod_deriv(const Orc& o, int64_t i, const std::vector<double>* S, double* du, double* acc) {
  const double* p = o.s.p;
  double fx = 0, fy = 0, W = 0, pressure = 0;
  int bad = 0;
  // with growth the length is the DE variable S[6] (its stage value)
  const std::vector<double>& L = o.rods_growth ? S[6] : o.rodL;
  const double ri[8] = {S[0][i], S[1][i], o.rodD[i], L[i], S[4][i], S[2][i], S[3][i], o.mass[i]};
  for_candidates(o, i, [&](int64_t j) {
    const double rj[8] = {S[0][j], S[1][j], o.rodD[j], L[j], S[4][j], S[2][j], S[3][j], o.mass[j]};
    double f[3];
    if (rod_forces(ri, rj, p, f)) bad = -1;
    fx += f[0];
    fy += f[1];
    W += f[2];
    pressure += std::sqrt(f[0] * f[0] + f[1] * f[1]);
  });
  const double m = o.mass[i], d = o.rodD[i], l = L[i];
  const double vx = S[2][i], vy = S[3][i], w = S[5][i];
  du[0] = vx;
  du[1] = vy;
  du[2] = -p[5] * vx + fx / m;
  du[3] = -p[5] * vy + fy / m;
  du[4] = w;
  du[5] = W / (m * ((d + l) * (d + l)) / 12 + m * (d * d)) - p[6] * w;
  // rods2dGrowth: dt(l) = growth/(1+alpha*pressure) (src/models/rods.jl:137-139)
  if (o.rods_growth) du[6] = o.m.rule_p[0] / (1 + o.m.rule_p[3] * pressure);
  acc[0] = fx; acc[1] = fy; acc[2] = W; acc[3] = pressure;
  return bad;
}

// The agentODE body evaluated on stage state S for agent i: returns the
// derivative du (positions, then velocities for soft spheres) and the new
// force accumulator acc.
static void deriv(const Orc& o, int64_t i, const std::vector<double>* S, double* du, double* acc) {
  double a3[3] = {0, 0, 0};
  const double* p = o.s.p;
  int law = o.m.law;
  if (law == CBM_LAW_SOFT_SPHERES) {
    for_candidates(o, i, [&](int64_t j) {
      double dx[3] = {0, 0, 0};
      for (int a = 0; a < o.dims; ++a) dx[a] = o.mi(a, S[a][i] - S[a][j]);
      double d = dist(o, dx);
      double sv;
      if (soft_s(p, d, o.rad[i], o.rad[j], o.fate[i], o.fate[j], sv))
        for (int a = 0; a < o.dims; ++a) a3[a] = a3[a] + (sv * dx[a]) / d;
    });
  } else if (law != CBM_LAW_NONE) {
    for_candidates(o, i, [&](int64_t j) {
      double dx[3] = {0, 0, 0};
      for (int a = 0; a < o.dims; ++a) dx[a] = o.mi(a, S[a][i] - S[a][j]);
      double d = dist(o, dx);
      double sv;
      if (law_s(law, p, d, sv)) {
        if (law == CBM_LAW_AGGREGATION) {
          // docs/src/Aggregation.md:88-98: dirx = (x - x[i2])/d; fx += s*dirx
          for (int a = 0; a < o.dims; ++a) a3[a] = a3[a] + sv * (dx[a] / d);
        } else {
          const double f = sv / d;          // built-in laws: F = (s/d) (x_i - x_j)
          for (int a = 0; a < o.dims; ++a) a3[a] = a3[a] + f * dx[a];
        }
      }
    });
    if (law == CBM_LAW_AGGREGATION && p[4] != 0.0) {
      // wall push of docs/src/Aggregation.md:100-112
      for (int a = 0; a < o.dims; ++a) {
        double x = S[a][i];
        if (x < o.s.box_lo[a] + p[0] / 2) a3[a] = a3[a] + p[1];
        else if (x > o.s.box_hi[a] - p[0] / 2) a3[a] = a3[a] - p[1];
      }
    }
  }
  if (o.inertial) {
    // dt(v) = -b*v/m + f/m, dt(x) = v (src/models/softSpheres.jl:32-35)
    const double b = p[0], mi = o.mass[i];
    for (int a = 0; a < o.dims; ++a) {
      acc[a] = a3[a];
      const double f = o.m.lagged ? o.F[a][i] : a3[a];
      const double v = S[o.dims + a][i];
      du[a] = v;
      du[o.dims + a] = -b * v / mi + f / mi;
    }
    return;
  }
  for (int a = 0; a < o.dims; ++a) {
    acc[a] = a3[a];
    double v = o.m.lagged ? o.F[a][i] : a3[a];
    if (o.m.drift != 0.0) v = v - o.m.drift * S[a][i];
    du[a] = v;
  }
}

// ---------------------------------------------------------------- Verlet
static void build_verlet(Orc& o) {
  int cap = o.m.max_neighbors;
  o.nbN.assign(o.nMax, 0);
  o.nbList.assign((int64_t)cap * o.nMax, -1);
  double rl = o.m.skin;
  // Candidate enumeration through a scratch grid of edge >= skin; lists are
  // then ordered by ascending slot, identical to the reference brute force
  // `for i2_ in 1:N` (src/neighbors.jl:114-126).
  // The reference's cellPos clamps every out-of-box agent into the outermost
  // padding cell, which can separate an agent just outside the box from a
  // neighbor two cells away; the list build therefore bins geometrically
  // (then clamps), which keeps every pair within rl in adjacent cells.
  Grid g; set_grid(o, g, std::max(rl, 1e-300));
  g.geometric = true;
  bool brute = (o.N < 64);
  for (int a = 0; a < o.dims; ++a) if (g.n[a] < 3) brute = true;
  std::vector<int64_t> cOf, cSt, cAg;
  if (!brute) {
    int64_t nc = g.total();
    std::vector<int64_t> cnt(nc, 0);
    cOf.assign(o.N, 0); cSt.assign(nc + 1, 0); cAg.assign(o.N, 0);
    for (int64_t i = 0; i < o.N; ++i) { cOf[i] = cell_of(o, g, i); cnt[cOf[i]]++; }
    for (int64_t k = 0; k < nc; ++k) cSt[k + 1] = cSt[k] + cnt[k];
    std::vector<int64_t> cur(cSt.begin(), cSt.end() - 1);
    for (int64_t i = 0; i < o.N; ++i) cAg[cur[cOf[i]]++] = i;
  }
  int nk = n_neigh_cells(o);
  int overflow = 0;
#pragma omp parallel for schedule(static) reduction(| : overflow)
  for (int64_t i = 0; i < o.N; ++i) {
    std::vector<int64_t> cand;
    auto test = [&](int64_t j) {
      if (j == i) return;
      double dx[3] = {0, 0, 0};
      for (int a = 0; a < o.dims; ++a) dx[a] = o.mi(a, o.X[a][i] - o.X[a][j]);
      if (dist(o, dx) < rl) cand.push_back(j);
    };
    if (brute) { for (int64_t j = 0; j < o.N; ++j) test(j); }
    else {
      for (int k = 0; k < nk; ++k) {
        int64_t c = neigh_cell(g, cOf[i], k);
        if (c < 0) continue;
        for (int64_t q = cSt[c]; q < cSt[c + 1]; ++q) test(cAg[q]);
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    }
    if ((int64_t)cand.size() > cap) { overflow = 1; cand.resize(cap); }
    o.nbN[i] = (int32_t)cand.size();
    for (size_t q = 0; q < cand.size(); ++q) o.nbList[(int64_t)q * o.nMax + i] = cand[q];
  }
  if (overflow) { o.err = CBM_ERR_LIST_OVERFLOW; o.msg = "Verlet list overflow"; }
  for (int a = 0; a < o.dims; ++a)
    for (int64_t i = 0; i < o.N; ++i) o.Xref[a][i] = o.X[a][i];
  o.rebuilds++;
}

static void compute_neighbors(Orc& o, bool force) {
  if (o.m.neighbors == CBM_FULL) { o.structure_valid = true; return; }
  if (o.m.neighbors == CBM_CELL_LINKED) { bin(o, o.g); o.rebuilds++; o.structure_valid = true; return; }
  bool rebuild = force || !o.structure_valid;
  if (!rebuild) {
    if (o.m.neighbors == CBM_VERLET_TIME) {
      rebuild = (o.tNext <= o.t);   // src/neighbors.jl:255
    } else {
      double thr = o.rebuild_thr();
      for (int64_t i = 0; i < o.N && !rebuild; ++i) {
        double dx[3] = {0, 0, 0};
        for (int a = 0; a < o.dims; ++a) dx[a] = o.mi(a, o.X[a][i] - o.Xref[a][i]);
        if (dist(o, dx) >= thr) rebuild = true;   // src/neighbors.jl:403
      }
    }
  }
  if (rebuild) {
    build_verlet(o);
    if (o.m.neighbors == CBM_VERLET_TIME) o.tNext = o.t + o.m.dt_recompute;
  }
  o.structure_valid = true;
}

// --------------------------------------------------------------- medium
// A user mediumODE body on the CPU: dt(c) at one node.
struct OrcMediumNode {
  double c, lap[3], x[3];
  int64_t idx[3];
  int32_t agents, dims;
  double dt;
  double nb[3][2];       // the neighbors' step-start values per axis, below / above (the node past the edge)
  int32_t border[3];     // -1 / +1 on the lower / upper border of that axis, else 0
  int32_t pad;
};
typedef double (*orc_medium_fn)(const OrcMediumNode* node);
// Optionally claims a node for assignment (the `p = ...` branches of a
// mediumODE body): returns nonzero and sets *v.
typedef int (*orc_medium_assign_fn)(const OrcMediumNode* node, double* v);

// The value a node on a side with a Dirichlet or copy border is assigned:
// the first such side in the order x-, x+, y-, y+, z-, z+ that the node lies
// on (`@mediumBorder(axis, side)` branches of an if/elseif chain,
// macros.jl:284-311); a copy takes the inward neighbor's step-start value
// (the reference's reflective idiom `p = p[2, i2_]`, docs/src/Usage_Agent.md
// :397-403).  Returns false for a node on no such side.
static bool border_value(const Orc& o, int64_t k, const int64_t* p, const int64_t* stride, double* v) {
  for (int a = 0; a < o.dims; ++a)
    for (int side = 0; side < 2; ++side) {
      const bool on = side == 0 ? p[a] == 0 : p[a] == o.mn[a] - 1;
      const int32_t mode = o.m.medium_border[2 * a + side];
      if (!on || mode == CBM_BORDER_MIRROR) continue;
      if (mode == CBM_BORDER_DIRICHLET) { *v = o.m.medium_border_value[2 * a + side]; return true; }
      // copy: the inward neighbor (the node itself on a one-node axis)
      const int64_t q = side == 0 ? p[a] + 1 : p[a] - 1;
      *v = (q >= 0 && q < o.mn[a]) ? o.c[k + (q - p[a]) * stride[a]] : o.c[k];
      return true;
    }
  return false;
}

// Douglas-Gunn ADI (src/integrators.jl:864-935): after the reaction step,
// one implicit sweep per axis with the reference's tridiagonal solver
// (:141-428): interior nodes of every line of interior nodes, diagonals
// -D r, 1 + 2 D r, -D r with r = dt / (dims dx^2) of the last axis (the
// sweeps all read the coefficient column `dims`), c*[i] recomputed per line;
// every node a sweep does not solve takes its value from u_new, which the
// reference never writes there, i.e. the initial field.
static void adi_sweeps(Orc& o) {
  const int dims = o.dims;
  const int la = dims - 1;
  const double dxl = (o.s.box_hi[la] - o.s.box_lo[la]) / (double)o.mn[la];
  const double r = o.m.dt / ((double)dims * (dxl * dxl));
  const double sub = -o.m.medium_D * r, main = 1 + 2 * o.m.medium_D * r, sup = -o.m.medium_D * r;
  const int64_t stride[3] = {1, o.mn[0], o.mn[0] * o.mn[1]};
  const int64_t nn = o.mn[0] * o.mn[1] * o.mn[2];
  std::vector<double> cst, dstar;
  for (int ax = 0; ax < dims; ++ax) {
    std::vector<double> xnew = o.cinit;
    const std::vector<double>& x = o.c;
    const int64_t n = o.mn[ax], s = stride[ax];
    cst.assign(n, 0.0);
    dstar.assign(n, 0.0);
    for (int64_t k = 0; k < nn; ++k) {
      const int64_t p[3] = {k % o.mn[0], (k / o.mn[0]) % o.mn[1], k / (o.mn[0] * o.mn[1])};
      bool line = p[ax] == 0;
      for (int b = 0; b < dims; ++b)
        if (b != ax) line = line && p[b] >= 1 && p[b] <= o.mn[b] - 2;
      if (!line) continue;
      cst[1] = sup / main;
      dstar[1] = (x[k + s] - sub * x[k]) / main;
      for (int64_t i = 2; i <= n - 3; ++i) {
        const double m = 1.0 / (main - sub * cst[i - 1]);
        cst[i] = sup * m;
        dstar[i] = (x[k + i * s] - sub * dstar[i - 1]) * m;
      }
      const double m = 1.0 / (main - sub * cst[n - 3]);
      dstar[n - 2] = ((x[k + (n - 2) * s] - sup * x[k + (n - 1) * s]) - sub * dstar[n - 3]) * m;
      xnew[k + (n - 2) * s] = dstar[n - 2];
      for (int64_t i = n - 3; i >= 1; --i) xnew[k + i * s] = dstar[i] - cst[i] * xnew[k + (i + 1) * s];
    }
    o.c.swap(xnew);
  }
}

// The medium node of a position: cellInMesh (src/metrics.jl:18-20, 0-based;
// x == xMax clamped), per axis.
static int64_t medium_node(const Orc& o, const double* x) {
  int64_t q[3] = {0, 0, 0};
  for (int a = 0; a < o.dims; ++a) {
    int64_t v;
    if (x[a] > o.s.box_hi[a]) v = o.mn[a] - 1;
    else if (x[a] < o.s.box_lo[a]) v = 0;
    else v = floor_div(x[a] - o.s.box_lo[a], o.mdx[a]);
    if (v > o.mn[a] - 1) v = o.mn[a] - 1;
    q[a] = v;
  }
  return q[0] + o.mn[0] * (q[1] + o.mn[1] * q[2]);
}

static void medium_step(Orc& o, int64_t nAgents) {
  const bool adi = o.m.medium_alg == CBM_MEDIUM_DGADI;   // reaction part here, diffusion in the sweeps
  int64_t nn = o.mn[0] * o.mn[1] * o.mn[2];
  // the rules' additions of this step land first
  for (int64_t k = 0; k < nn && !o.medAdd.empty(); ++k)
    if (o.medAdd[k] != 0) {
      o.c[k] = o.c[k] + (double)o.medAdd[k] * (1.0 / 4294967296.0);
      o.medAdd[k] = 0;
    }
  o.nodeCount.assign(nn, 0);
  for (int64_t i = 0; i < nAgents; ++i) {
    int64_t q[3] = {0, 0, 0};
    for (int a = 0; a < o.dims; ++a) {
      double x = o.X[a][i];
      // cellInMesh, src/metrics.jl:18-20 (0-based; x == xMax clamped)
      int64_t v;
      if (x > o.s.box_hi[a]) v = o.mn[a] - 1;
      else if (x < o.s.box_lo[a]) v = 0;
      else v = floor_div(x - o.s.box_lo[a], o.mdx[a]);
      if (v > o.mn[a] - 1) v = o.mn[a] - 1;
      q[a] = v;
    }
    o.nodeCount[q[0] + o.mn[0] * (q[1] + o.mn[1] * q[2])]++;
  }
  o.cnew.resize(nn);
  const double D = o.m.medium_D, dt = o.m.dt, u = o.m.medium_consume;
  if (o.medium_fn) {   // a user body (test code's restatement), serially
    orc_medium_fn fn = (orc_medium_fn)o.medium_fn;
    orc_medium_assign_fn asg = (orc_medium_assign_fn)o.medium_assign;
    for (int64_t k = 0; k < nn; ++k) {
      OrcMediumNode nd;
      std::memset(&nd, 0, sizeof(nd));
      int64_t p[3] = {k % o.mn[0], (k / o.mn[0]) % o.mn[1], k / (o.mn[0] * o.mn[1])};
      int64_t stride[3] = {1, o.mn[0], o.mn[0] * o.mn[1]};
      nd.c = o.c[k]; nd.agents = o.nodeCount[k]; nd.dims = o.dims; nd.dt = dt;
      for (int a = 0; a < 3; ++a) {
        nd.idx[a] = p[a]; nd.lap[a] = 0.0; nd.x[a] = 0.0;
        if (a >= o.dims) continue;
        double cp = (p[a] + 1 < o.mn[a]) ? o.c[k + stride[a]] : nd.c;
        double cm = (p[a] > 0) ? o.c[k - stride[a]] : nd.c;
        nd.nb[a][0] = cm;
        nd.nb[a][1] = cp;
        nd.border[a] = p[a] == 0 ? -1 : (p[a] == o.mn[a] - 1 ? 1 : 0);
        if (!adi) nd.lap[a] = ((cp - 2 * nd.c) + cm) / (o.mdx[a] * o.mdx[a]);
        nd.x[a] = o.s.box_lo[a] + ((double)p[a] + 0.5) * o.mdx[a];
      }
      double v;
      if (border_value(o, k, p, stride, &v)) o.cnew[k] = v;
      else if (asg && asg(&nd, &v)) o.cnew[k] = v;
      else o.cnew[k] = nd.c + dt * fn(&nd);
    }
    o.c.swap(o.cnew);
    if (adi) adi_sweeps(o);
    return;
  }
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nn; ++k) {
    int64_t p[3] = {k % o.mn[0], (k / o.mn[0]) % o.mn[1], k / (o.mn[0] * o.mn[1])};
    double cc = o.c[k];
    double lap = 0.0;
    int64_t stride[3] = {1, o.mn[0], o.mn[0] * o.mn[1]};
    for (int a = 0; a < o.dims && !adi; ++a) {
      // zero-flux ghost mirror: c[-1] = c[0], c[n] = c[n-1]
      double cp = (p[a] + 1 < o.mn[a]) ? o.c[k + stride[a]] : cc;
      double cm = (p[a] > 0) ? o.c[k - stride[a]] : cc;
      double l = ((cp - 2 * cc) + cm) / (o.mdx[a] * o.mdx[a]);   // @∂2, macros.jl:203-257
      lap = (a == 0) ? l : lap + l;
    }
    double f = D * lap - (u * cc) * (double)o.nodeCount[k];
    double v;
    o.cnew[k] = border_value(o, k, p, stride, &v) ? v : cc + dt * f;
  }
  o.c.swap(o.cnew);
  if (adi) adi_sweeps(o);
}

// ------------------------------------------------------------------ rules
struct Births { std::vector<int64_t> parent; };

// A CBM_RULE_CUSTOM body on the CPU: one agent's values at the start of the
// step (old), its new values (nw, initialised to old) and up to two added
// agents (add, each initialised to old by the caller); sets *n_add, *dead.
struct OrcRuleRec {
  double x[3]; double vol, growth, vbirth; double u[8]; int64_t lf;
  double medium;   // old: the medium at the agent's node as the step started
  double madd;     // nw: the body's additions to that node, in call order
};
typedef void (*orc_rule_fn)(int64_t id, int64_t step, const OrcRuleRec* old, OrcRuleRec* nw, OrcRuleRec* add,
                            int32_t* n_add, int32_t* dead, int64_t slot);

// The generated rule loop of functionRule.jl:121-160 around a custom body,
// serially in ascending slot order: the new values are written in place (DE
// state, committed growth/vbirth through the write-new buffers), @addAgent
// appends at N + NAdd_ with id idMax + 1 + NAdd_ (functionRule.jl:59-69),
// @removeAgent pushes the slot on the hole list (:98-101).
static void apply_custom_rule(Orc& o, std::vector<int64_t>& holes, int64_t& nAdd) {
  orc_rule_fn fn = (orc_rule_fn)o.rule_fn;
  if (!fn) { o.err = CBM_ERR_CONFIG; o.msg = "no custom rule"; return; }
  const int64_t N = o.N;
  // neighbors are read as they were when the loop started (the reference's
  // reads of x[i2] see the committed values; writes go to x__)
  for (int a = 0; a < 3; ++a) o.ruleX[a] = o.X[a];
  o.ruleVol = o.vol;
  for (int k = 0; k < 8; ++k) o.ruleU[k] = o.U[k];
  const bool med = o.m.medium_n[0] > 0;
  for (int64_t i = 0; i < N; ++i) {
    OrcRuleRec old;
    for (int a = 0; a < 3; ++a) old.x[a] = a < o.dims ? o.X[a][i] : 0.0;
    old.vol = o.vol[i]; old.growth = o.gr[i]; old.vbirth = o.vb[i];
    for (int k = 0; k < 8; ++k) old.u[k] = o.U[k][i];
    old.lf = o.lf[i];
    const int64_t node = med ? medium_node(o, old.x) : 0;
    old.medium = med ? o.c[node] : 0.0;
    old.madd = 0.0;
    OrcRuleRec nw = old, add[2] = {old, old};
    int32_t na = 0, dead = 0;
    fn(o.id[i], o.step, &old, &nw, add, &na, &dead, i);
    if (na > 2) { o.err = CBM_ERR_CONFIG; o.msg = "more than 2 additions"; na = 2; }
    // `p += v` of the body on its node: summed per node in 32.32 fixed
    // point (exact in any order), added to the field before its update
    if (nw.madd != 0.0) {
      if (!med) { o.err = CBM_ERR_CONFIG; o.msg = "medium access in a model without a medium"; }
      else o.medAdd[node] += std::llrint(nw.madd * 4294967296.0);
    }
    for (int a = 0; a < o.dims; ++a) o.X[a][i] = nw.x[a];
    o.vol[i] = nw.vol; o.grNew[i] = nw.growth; o.vbNew[i] = nw.vbirth;
    for (int k = 0; k < 8; ++k) o.U[k][i] = nw.u[k];
    o.lf[i] = nw.lf;
    if (dead) { holes.push_back(i); o.alive[i] = 0; }
    for (int k = 0; k < na; ++k) {
      const int64_t j = N + nAdd;
      if (j >= o.nMax) { o.err = CBM_ERR_CAPACITY; o.msg = "agent capacity exceeded"; continue; }
      const OrcRuleRec& d = add[k];
      for (int a = 0; a < o.dims; ++a) { o.X[a][j] = d.x[a]; o.F[a][j] = o.F[a][i]; o.Fnew[a][j] = o.F[a][i]; }
      o.vol[j] = d.vol; o.gr[j] = d.growth; o.grNew[j] = d.growth; o.vb[j] = d.vbirth; o.vbNew[j] = d.vbirth;
      for (int q = 0; q < 8; ++q) o.U[q][j] = d.u[q];
      o.lf[j] = d.lf;
      o.id[j] = o.idMax + 1 + nAdd; o.alive[j] = 1;
      nAdd++;
    }
  }
}

// The built-in agentRule bodies, executed serially in ascending slot order
// (the single-thread order that makes @removeAgent's hole list ascending).
static void apply_rule(Orc& o, std::vector<int64_t>& holes, int64_t& nAdd) {
  holes.clear(); nAdd = 0;
  if (o.m.rule == CBM_RULE_NONE) return;
  if (o.m.rule == CBM_RULE_CUSTOM) { apply_custom_rule(o, holes, nAdd); return; }
  int64_t N = o.N;
  for (int64_t i = 0; i < N; ++i) {
    if (o.m.rule == CBM_RULE_PARITY) {
      double xold = o.X[0][i];
      o.X[0][i] = xold + 1.0;
      if (o.id[i] % 2 == 1) {
        holes.push_back(i); o.alive[i] = 0;
      } else if (o.id[i] % 2 == 0) {
        int64_t nw = N + nAdd;
        int64_t newId = o.idMax + 1 + nAdd;
        if (nw >= o.nMax) { o.err = CBM_ERR_CAPACITY; o.msg = "agent capacity exceeded"; continue; }
        nAdd++;
        for (int a = 0; a < o.dims; ++a) { o.X[a][nw] = o.X[a][i]; o.F[a][nw] = o.F[a][i]; o.Fnew[a][nw] = o.F[a][i]; }
        o.X[0][nw] = xold;
        o.id[nw] = newId; o.lf[nw] = o.id[i]; o.alive[nw] = 1;
        if (o.growth) { o.vol[nw] = o.vol[i]; o.grNew[nw] = o.gr[i]; o.vbNew[nw] = o.vb[i]; o.gr[nw] = o.gr[i]; o.vb[nw] = o.vb[i]; }
      }
    } else if (o.m.rule == CBM_RULE_DEATH) {
      U4 r = draw(o.m.seed, o.step, o.id[i], o.sim, 0x400u);
      if (u53(r.v[0], r.v[1]) < o.m.rule_p[0] * o.m.dt) { holes.push_back(i); o.alive[i] = 0; }
    } else if (o.m.rule == CBM_RULE_RODS_GROWTH) {
      // rods2dGrowth's agentRule (src/models/rods.jl:141-160): a rod past
      // its lTarget is removed and two daughters are appended
      if (!(o.rodL[i] > o.LT[i])) continue;
      holes.push_back(i); o.alive[i] = 0;
      const U4 q0 = draw(o.m.seed, o.step, o.id[i], o.sim, 0x700u);
      const U4 q1 = draw(o.m.seed, o.step, o.id[i], o.sim, 0x701u);
      const double ww = (0.1 - (-0.1)) * u53(q0.v[0], q0.v[1]) + (-0.1);      // uniform(-.1, .1)
      const double l0 = o.m.rule_p[2] - o.m.rule_p[1], l1 = o.m.rule_p[2] + o.m.rule_p[1];
      const double lt[2] = {(l1 - l0) * u53(q0.v[2], q0.v[3]) + l0, (l1 - l0) * u53(q1.v[0], q1.v[1]) + l0};
      const double x = o.X[0][i], y = o.X[1][i], th = o.TH[i], om = o.OM[i], l = o.rodL[i], d = o.rodD[i];
      for (int k = 0; k < 2; ++k) {
        const int64_t nw = N + nAdd;
        if (nw >= o.nMax) { o.err = CBM_ERR_CAPACITY; o.msg = "agent capacity exceeded"; continue; }
        const int64_t newId = o.idMax + 1 + nAdd;
        nAdd++;
        for (int a = 0; a < 2; ++a) { o.V[a][nw] = o.V[a][i]; o.F[a][nw] = o.F[a][i]; o.Fnew[a][nw] = o.F[a][i]; }
        o.TH[nw] = th; o.rodD[nw] = d; o.mass[nw] = o.mass[i]; o.rodW[nw] = o.rodW[i]; o.rodP[nw] = o.rodP[i];
        // x = (l+d)/4*cos(theta)+x, the second daughter -(l+d)/4*cos(theta)+x
        const double q = k ? -((l + d) / 4) : (l + d) / 4;
        o.X[0][nw] = q * std::cos(th) + x;
        o.X[1][nw] = q * std::sin(th) + y;
        o.rodL[nw] = (l - d) / 2;
        o.OM[nw] = k ? om - ww : om + ww;
        o.LT[nw] = lt[k];
        o.id[nw] = newId; o.lf[nw] = o.lf[i]; o.alive[nw] = 1;
      }
    } else if (o.m.rule == CBM_RULE_GROWTH_DIVISION) {
      if (!(o.vol[i] >= o.m.rule_p[3] * o.vb[i])) continue;
      int64_t nw = N + nAdd;
      if (nw >= o.nMax) { o.err = CBM_ERR_CAPACITY; o.msg = "agent capacity exceeded"; continue; }
      int64_t newId = o.idMax + 1 + nAdd;
      nAdd++;
      U4 ra = draw(o.m.seed, o.step, o.id[i], o.sim, 0x200u);
      U4 rg = draw(o.m.seed, o.step, o.id[i], o.sim, 0x300u);
      double ax[3] = {0, 0, 0};
      if (o.dims == 1) ax[0] = (u53(ra.v[0], ra.v[1]) < 0.5) ? -1.0 : 1.0;
      else if (o.dims == 2) {
        double th = 6.283185307179586 * u53(ra.v[0], ra.v[1]);
        ax[0] = std::cos(th); ax[1] = std::sin(th);
      } else {
        double z = 2.0 * u53(ra.v[0], ra.v[1]) - 1.0;
        double ph = 6.283185307179586 * u53(ra.v[2], ra.v[3]);
        double rr = std::sqrt(1.0 - z * z);
        ax[0] = rr * std::cos(ph); ax[1] = rr * std::sin(ph); ax[2] = z;
      }
      double split = o.m.rule_p[2];
      double half = o.vol[i] / 2.0;
      double gm = o.m.rule_p[0], gs = o.m.rule_p[1];
      double g1 = gm * (1.0 + gs * (2.0 * u53(rg.v[0], rg.v[1]) - 1.0));
      double g2 = gm * (1.0 + gs * (2.0 * u53(rg.v[2], rg.v[3]) - 1.0));
      for (int a = 0; a < o.dims; ++a) {
        double x = o.X[a][i];
        o.X[a][nw] = x + split * ax[a];
        o.X[a][i] = x - split * ax[a];
        o.F[a][nw] = o.F[a][i]; o.Fnew[a][nw] = o.F[a][i];
      }
      o.vol[i] = half; o.vol[nw] = half;
      o.vbNew[i] = half; o.vbNew[nw] = half; o.vb[nw] = half;
      o.grNew[i] = g1; o.grNew[nw] = g2; o.gr[nw] = g2;
      o.id[nw] = newId; o.lf[nw] = o.lf[i]; o.alive[nw] = 1;
    }
  }
}

// update!: canonical compaction (src/CommunityStructure/update.jl:15-119),
// ascending holes, survivors scanned from the tail.
static void compact(Orc& o, const std::vector<int64_t>& holes, int64_t nAdd) {
  int64_t R = (int64_t)holes.size();
  int64_t top = o.N + nAdd;
  if (R > 0) {
    std::vector<int64_t> repos;
    int64_t last = top - 1;
    while ((int64_t)repos.size() < R && last >= 0) { if (o.alive[last]) repos.push_back(last); last--; }
    for (int64_t m = 0; m < (int64_t)repos.size(); ++m) {
      int64_t dst = holes[m], src = repos[m];
      if (!(dst < src)) continue;
      for (int a = 0; a < o.dims; ++a) { o.X[a][dst] = o.X[a][src]; o.F[a][dst] = o.F[a][src]; o.Fnew[a][dst] = o.Fnew[a][src]; }
      o.id[dst] = o.id[src]; o.lf[dst] = o.lf[src];
      for (int k = 0; k < 8; ++k) o.U[k][dst] = o.U[k][src];
      if (o.growth) { o.vol[dst] = o.vol[src]; o.gr[dst] = o.gr[src]; o.vb[dst] = o.vb[src]; o.grNew[dst] = o.grNew[src]; o.vbNew[dst] = o.vbNew[src]; }
      if (o.rods) {
        for (int a = 0; a < 2; ++a) o.V[a][dst] = o.V[a][src];
        o.TH[dst] = o.TH[src]; o.OM[dst] = o.OM[src]; o.rodD[dst] = o.rodD[src]; o.rodL[dst] = o.rodL[src];
        o.mass[dst] = o.mass[src]; o.LT[dst] = o.LT[src]; o.rodW[dst] = o.rodW[src]; o.rodP[dst] = o.rodP[src];
      }
    }
  }
  o.N = top - R;
  o.idMax += nAdd;
  for (int64_t i = 0; i < o.N; ++i) o.alive[i] = 1;
}

// ---------------------------------------------------------------- stepping
static void de_step(Orc& o, int64_t nA) {
  const int d = o.dims;
  const double dt = o.m.dt;
  // DE variables: positions (+ vol), or positions then velocities, or the
  // rods' x, y, vx, vy, theta, omega
  const bool body = o.law_ode != nullptr;
  const int ude = body ? o.m.user_de : 0;
  const int base = d + (o.growth ? 1 : 0);
  const int nv = o.rods ? (o.rods_growth ? 7 : 6) : (o.inertial ? 2 * d : base + ude);
  std::vector<double> U[8];
  for (int a = 0; a < d; ++a) U[a] = o.X[a];
  if (o.inertial || o.rods) for (int a = 0; a < d; ++a) U[d + a] = o.V[a];
  if (o.rods) { U[4] = o.TH; U[5] = o.OM; }
  if (o.rods_growth) U[6] = o.rodL;
  if (o.growth) U[d] = o.vol;
  for (int k = 0; k < ude; ++k) U[base + k] = o.U[k];
  std::vector<char> bad(o.rods ? nA : 0, 0);
  auto eval = [&](const std::vector<double>* S, std::vector<double>* K) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nA; ++i) {
      double du[8], acc[4];
      if (body) {
        body_deriv(o, i, S, du);
        for (int a = 0; a < d; ++a) K[a][i] = du[a];
        for (int k = 0; k < ude; ++k) K[base + k][i] = du[base + k];
        if (o.growth) K[d][i] = o.gr[i] * S[d][i];
        continue;
      }
      if (o.rods) {
        // every evaluation leaves its fx, fy, W and pressure in the agent
        if (rod_deriv(o, i, S, du, acc)) bad[i] = 1;
        for (int v = 0; v < nv; ++v) K[v][i] = du[v];
        o.Fnew[0][i] = acc[0]; o.Fnew[1][i] = acc[1]; o.rodW[i] = acc[2]; o.rodP[i] = acc[3];
        continue;
      }
      deriv(o, i, S, du, acc);
      for (int a = 0; a < d; ++a) { K[a][i] = du[a]; o.Fnew[a][i] = acc[a]; }
      if (o.inertial) for (int a = 0; a < d; ++a) K[d + a][i] = du[d + a];
      if (o.growth) K[d][i] = o.gr[i] * S[d][i];
    }
  };
  std::vector<double> K1[8], K2[8], K3[8], K4[8], H[8];
  for (int v = 0; v < nv; ++v) { K1[v].assign(nA, 0); H[v] = U[v]; }
  int integ = o.m.integrator;
  if (integ == CBM_EULER) {
    eval(U, K1);
    for (int v = 0; v < nv; ++v) for (int64_t i = 0; i < nA; ++i) U[v][i] = U[v][i] + K1[v][i] * dt;
  } else if (integ == CBM_HEUN) {
    eval(U, K1);
    for (int v = 0; v < nv; ++v) for (int64_t i = 0; i < nA; ++i) H[v][i] = U[v][i] + K1[v][i] * dt;
    for (int v = 0; v < nv; ++v) K2[v].assign(nA, 0);
    eval(H, K2);
    for (int v = 0; v < nv; ++v) for (int64_t i = 0; i < nA; ++i) U[v][i] = (H[v][i] + (U[v][i] + K2[v][i] * dt)) / 2;
  } else if (integ == CBM_RK4) {
    for (int v = 0; v < nv; ++v) { K2[v].assign(nA, 0); K3[v].assign(nA, 0); K4[v].assign(nA, 0); }
    eval(U, K1);
    for (int v = 0; v < nv; ++v) for (int64_t i = 0; i < nA; ++i) H[v][i] = U[v][i] + K1[v][i] * dt / 2;
    eval(H, K2);
    for (int v = 0; v < nv; ++v) for (int64_t i = 0; i < nA; ++i) H[v][i] = U[v][i] + K2[v][i] * dt / 2;
    eval(H, K3);
    for (int v = 0; v < nv; ++v) for (int64_t i = 0; i < nA; ++i) H[v][i] = U[v][i] + K3[v][i] * dt;
    eval(H, K4);
    for (int v = 0; v < nv; ++v)
      for (int64_t i = 0; i < nA; ++i)
        U[v][i] = U[v][i] + (K1[v][i] + 2 * K2[v][i] + 2 * K3[v][i] + K4[v][i]) * dt / 6;
  } else if (integ == CBM_EULER_HEUN) {
    // src/integrators.jl:775-807 (CPU branch): predictor with noise z1, then
    // the average of both stages with a fresh draw z2 (reference defect D8)
    eval(U, K1);
    const double sq = std::sqrt(dt);
    const double gsig = o.m.noise;
    const double slope = o.noise_slope;
    std::vector<double> G1[3];                 // g of the predictor (state-dependent noise)
    for (int a = 0; a < d; ++a) G1[a].assign(nA, gsig);
    for (int64_t i = 0; i < nA; ++i) {
      double z[4];
      U4 r0 = draw(o.m.seed, o.step, o.id[i], o.sim, 0x100u);
      normal2(r0, z[0], z[1]);
      if (d > 2) { U4 r1 = draw(o.m.seed, o.step, o.id[i], o.sim, 0x101u); normal2(r1, z[2], z[3]); }
      for (int a = 0; a < d; ++a) {
        if (slope != 0.0) G1[a][i] = gsig + slope * U[a][i];
        H[a][i] = (U[a][i] + K1[a][i] * dt) + G1[a][i] * z[a] * sq;
      }
      for (int v = d; v < nv; ++v) H[v][i] = U[v][i] + K1[v][i] * dt;    // no noise off the positions
    }
    for (int v = 0; v < nv; ++v) K2[v].assign(nA, 0);
    eval(H, K2);
    for (int64_t i = 0; i < nA; ++i) {
      double z[4];
      U4 r0 = draw(o.m.seed, o.step, o.id[i], o.sim, 0x500u);
      normal2(r0, z[0], z[1]);
      if (d > 2) { U4 r1 = draw(o.m.seed, o.step, o.id[i], o.sim, 0x501u); normal2(r1, z[2], z[3]); }
      for (int a = 0; a < d; ++a) {
        // (g(u) + g(h)) of the reference's corrector (src/integrators.jl:801-803)
        const double g2 = slope != 0.0 ? gsig + slope * H[a][i] : gsig;
        U[a][i] = U[a][i] + ((K1[a][i] + K2[a][i]) * dt + (G1[a][i] + g2) * z[a] * sq) / 2;
      }
      for (int v = d; v < nv; ++v) U[v][i] = U[v][i] + ((K1[v][i] + K2[v][i]) * dt) / 2;
    }
  } else {  // EM
    eval(U, K1);
    double sq = std::sqrt(dt);
    double gsig = o.m.noise;
    const double slope = o.noise_slope;
    for (int64_t i = 0; i < nA; ++i) {
      double z[4];
      U4 r0 = draw(o.m.seed, o.step, o.id[i], o.sim, 0x100u);
      normal2(r0, z[0], z[1]);
      if (d > 2) { U4 r1 = draw(o.m.seed, o.step, o.id[i], o.sim, 0x101u); normal2(r1, z[2], z[3]); }
      for (int a = 0; a < d; ++a) {
        const double g = slope != 0.0 ? gsig + slope * U[a][i] : gsig;   // g at u (src/integrators.jl:709-727)
        U[a][i] = (U[a][i] + K1[a][i] * dt) + g * z[a] * sq;
      }
      for (int v = d; v < nv; ++v) U[v][i] = U[v][i] + K1[v][i] * dt;
    }
  }
  for (int a = 0; a < d; ++a) std::copy(U[a].begin(), U[a].begin() + nA, o.X[a].begin());
  if (o.inertial || o.rods) for (int a = 0; a < d; ++a) std::copy(U[d + a].begin(), U[d + a].begin() + nA, o.V[a].begin());
  if (o.rods) {
    std::copy(U[4].begin(), U[4].begin() + nA, o.TH.begin());
    std::copy(U[5].begin(), U[5].begin() + nA, o.OM.begin());
    if (o.rods_growth) std::copy(U[6].begin(), U[6].begin() + nA, o.rodL.begin());
    if (std::find(bad.begin(), bad.end(), 1) != bad.end()) {
      o.err = CBM_ERR_NONFINITE;
      o.msg = "rod contact with a negative overlap d_i - r_ij";
    }
  }
  if (o.growth) std::copy(U[d].begin(), U[d].begin() + nA, o.vol.begin());
  for (int k = 0; k < ude; ++k) std::copy(U[base + k].begin(), U[base + k].begin() + nA, o.U[k].begin());
}

static void wrap_positions(Orc& o, int64_t n) {
  if (!o.m.periodic) return;
  for (int a = 0; a < o.dims; ++a) {
    double lo = o.s.box_lo[a], L = o.s.box_hi[a] - o.s.box_lo[a];
    for (int64_t i = 0; i < n; ++i) {
      double x = o.X[a][i];
      if (x >= lo + L) x = x - L; else if (x < lo) x = x + L;
      o.X[a][i] = x;
    }
  }
}

// ------------------------------------------------------------ model scope
// modelStepRule! then modelStepDE! (step.jl:96-107): the rule writes new
// values from the step-start ones, the DE parameters step by Euler or, with
// a noise term, Euler-Maruyama (the default modelAlg, agentStructure.jl:
// 276-280); the new values are committed at once, after every agent and
// medium update of the step has read the old ones.
struct OrcModelScope {
  double t, dt;
  int64_t N, step;
  int32_t sim, pad;
  double p[8];
  double pn[8];
};
typedef void (*orc_model_rule_fn)(OrcModelScope* m);
typedef void (*orc_model_ode_fn)(const OrcModelScope* m, double* dp);

static void model_step(Orc& o) {
  if (!o.model_rule && !o.model_ode) return;
  OrcModelScope m;
  m.t = o.t; m.dt = o.m.dt; m.N = o.N; m.step = o.step; m.sim = o.sim; m.pad = 0;
  for (int k = 0; k < 8; ++k) { m.p[k] = o.s.p[k]; m.pn[k] = m.p[k]; }
  if (o.model_rule) ((orc_model_rule_fn)o.model_rule)(&m);
  if (o.model_ode) {
    double dp[8] = {0, 0, 0, 0, 0, 0, 0, 0}, g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    ((orc_model_ode_fn)o.model_ode)(&m, dp);
    if (o.model_noise) ((orc_model_ode_fn)o.model_noise)(&m, g);
    const double sq = std::sqrt(m.dt);
    for (int q = 0; q < 4; ++q) {
      double z[2] = {0, 0};
      if (o.model_noise && ((o.model_de >> (2 * q)) & 3u))
        normal2(draw(o.m.seed, o.step, 0, o.sim, 0x600u + (uint32_t)q), z[0], z[1]);
      for (int h = 0; h < 2; ++h) {
        const int k = 2 * q + h;
        if (!((o.model_de >> k) & 1u)) continue;
        double v = m.p[k] + dp[k] * m.dt;
        if (o.model_noise) v = v + g[k] * z[h] * sq;
        m.pn[k] = v;
      }
    }
  }
  for (int k = 0; k < 8; ++k) o.s.p[k] = m.pn[k];
}

static void one_step(Orc& o) {
  std::vector<int64_t> holes; int64_t nAdd = 0;
  apply_rule(o, holes, nAdd);                 // agentStepRule!
  int64_t nA = o.N;                            // DE loop runs over 1:N
  de_step(o, nA);                              // agentStepDE!
  wrap_positions(o, nA);                       // periodic wrap after the final stage: the medium sees it
  if (o.m.medium_n[0] > 0) medium_step(o, nA); // mediumStepDE!
  model_step(o);                               // modelStepRule!, modelStepDE!
  // update!: commit the write-new buffers, then compact
  for (int a = 0; a < o.dims; ++a) for (int64_t i = 0; i < nA; ++i) o.F[a][i] = o.Fnew[a][i];
  if (o.growth) for (int64_t i = 0; i < o.N + nAdd; ++i) { o.gr[i] = o.grNew[i]; o.vb[i] = o.vbNew[i]; }
  bool changed = (nAdd > 0) || !holes.empty();
  compact(o, holes, nAdd);
  o.t = o.t + o.m.dt;
  o.step++;
  wrap_positions(o, o.N);                      // new agents placed outside the box
  compute_neighbors(o, changed);               // computeNeighbors!
}

static int field_is_int(int f) { return f == CBM_F_ID || f == CBM_F_LF || f == CBM_F_FATE; }

// -------------------------------------------------------------- adjoint
// Overdamped Euler trajectory loss for the REP_ADH law (model-scope
// parameters theta = (k, r_eq, eps), p[3] = dc factor), evaluated in type T.
// The neighbor set of each step is the Verlet list of the oracle, rebuilt
// with the usual trigger on the value part of the state.
template <class T>
static T traj_loss(Orc& o, const T* theta, int64_t n_steps, int32_t n_saves,
                   const int64_t* save_steps, const double* target) {
  const int d = o.dims;
  const int64_t N = o.N;
  std::vector<T> X[3], K[3];
  for (int a = 0; a < d; ++a) { X[a].resize(N); K[a].resize(N); for (int64_t i = 0; i < N; ++i) X[a][i] = T(o.X[a][i]); }
  T P[4] = {theta[0], theta[1], theta[2], T(o.s.p[3])};
  T loss(0.0);
  int32_t si = 0;
  compute_neighbors(o, true);
  for (int64_t st = 1; st <= n_steps; ++st) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      T acc[3] = {T(0.0), T(0.0), T(0.0)};
      for (int32_t q = 0; q < o.nbN[i]; ++q) {
        int64_t j = o.nbList[(int64_t)q * o.nMax + i];
        T dx[3] = {T(0.0), T(0.0), T(0.0)};
        for (int a = 0; a < d; ++a) dx[a] = o.mi(a, X[a][i] - X[a][j]);
        T dd = dist(o, dx);
        T sv;
        if (law_s(CBM_LAW_REP_ADH, P, dd, sv)) {
          T f = sv / dd;
          for (int a = 0; a < d; ++a) acc[a] = acc[a] + f * dx[a];
        }
      }
      for (int a = 0; a < d; ++a) K[a][i] = acc[a];
    }
    for (int a = 0; a < d; ++a) for (int64_t i = 0; i < N; ++i) X[a][i] = X[a][i] + K[a][i] * T(o.m.dt);
    // keep the value part in the oracle state for the rebuild trigger
    for (int a = 0; a < d; ++a) for (int64_t i = 0; i < N; ++i) o.X[a][i] = val(X[a][i]);
    o.t = o.t + o.m.dt; o.step++;
    compute_neighbors(o, false);
    if (si < n_saves && save_steps[si] == st) {
      const double* tg = target + (int64_t)si * d * N;
      T part(0.0);
      for (int a = 0; a < d; ++a)
        for (int64_t i = 0; i < N; ++i) { T e = X[a][i] - T(tg[(int64_t)a * N + i]); part = part + e * e; }
      loss = loss + part;
      si++;
    }
  }
  return loss / T((double)n_saves * (double)N);
}

}  // namespace

// ================================================================== C API
extern "C" {

void* orc_create(const cbm_model_desc* m, const cbm_sim_desc* s) {
  Orc* o = new Orc();
  o->m = *m; o->s = *s;
  o->dims = m->dims;
  o->N = s->n; o->nMax = std::max<int64_t>(s->n_max, s->n);
  o->idMax = s->n;
  o->growth = (m->rule == CBM_RULE_GROWTH_DIVISION || m->rule == CBM_RULE_CUSTOM);
  o->inertial = (m->law == CBM_LAW_SOFT_SPHERES);
  if (o->inertial && m->rule != CBM_RULE_NONE) { o->err = CBM_ERR_CONFIG; o->msg = "soft spheres take no rules"; }
  o->rods = (m->law == CBM_LAW_RODS);
  o->rods_growth = o->rods && m->rule == CBM_RULE_RODS_GROWTH;
  if ((m->rule == CBM_RULE_RODS_GROWTH) != o->rods_growth) { o->err = CBM_ERR_CONFIG; o->msg = "CBM_RULE_RODS_GROWTH needs the rod model"; }
  if (o->rods && (m->dims != 2 || (m->rule != CBM_RULE_NONE && !o->rods_growth) || m->periodic || m->lagged || m->medium_n[0] > 0 ||
                  m->drift != 0.0 || !(m->integrator == CBM_EULER || m->integrator == CBM_HEUN || m->integrator == CBM_RK4))) {
    o->err = CBM_ERR_CONFIG;
    o->msg = "the rod model is 2-D, without rules, periodic box, lagged forces, medium or drift; Euler, Heun or RK4";
  }
  for (int a = 0; a < 3; ++a) {
    o->X[a].assign(o->nMax, 0); o->F[a].assign(o->nMax, 0); o->Fnew[a].assign(o->nMax, 0); o->Xref[a].assign(o->nMax, 0);
    o->V[a].assign(o->nMax, 0);
  }
  o->rad.assign(o->nMax, 1.0); o->mass.assign(o->nMax, 1.0); o->fate.assign(o->nMax, 1);
  o->TH.assign(o->nMax, 0.0); o->OM.assign(o->nMax, 0.0); o->rodD.assign(o->nMax, 1.0); o->rodL.assign(o->nMax, 1.0);
  o->rodW.assign(o->nMax, 0.0); o->rodP.assign(o->nMax, 0.0); o->LT.assign(o->nMax, 0.0);
  o->id.resize(o->nMax); o->lf.assign(o->nMax, 0); o->alive.assign(o->nMax, 1);
  for (int64_t i = 0; i < o->nMax; ++i) o->id[i] = i + 1;
  o->vol.assign(o->nMax, 1.0); o->gr.assign(o->nMax, 0.0); o->vb.assign(o->nMax, 1.0);
  o->grNew.assign(o->nMax, 0.0); o->vbNew.assign(o->nMax, 1.0);
  for (int k = 0; k < 8; ++k) o->U[k].assign(o->nMax, 0.0);
  double edge = m->cell_edge;
  if (m->neighbors == CBM_CELL_LINKED) {
    set_grid(*o, o->g, edge);
    if (o->g.periodic)
      for (int a = 0; a < o->dims; ++a)
        if (o->g.n[a] < 3) { o->err = CBM_ERR_CONFIG; o->msg = "periodic cell grid needs >= 3 cells per axis"; }
  }
  if (m->medium_n[0] > 0) {
    for (int a = 0; a < 3; ++a) { o->mn[a] = 1; o->mdx[a] = 1; }
    for (int a = 0; a < o->dims; ++a) {
      o->mn[a] = m->medium_n[a];
      o->mdx[a] = (s->box_hi[a] - s->box_lo[a]) / (double)m->medium_n[a];   // communityStructure.jl:274
    }
    o->c.assign(o->mn[0] * o->mn[1] * o->mn[2], m->medium_c0);
    o->medAdd.assign(o->c.size(), 0);
    o->cinit = o->c;
    if (m->medium_alg == CBM_MEDIUM_DGADI)
      for (int a = 0; a < o->dims; ++a)
        if (o->mn[a] < 4) { o->err = CBM_ERR_CONFIG; o->msg = "DGADI needs NMedium >= 4 on every axis"; }
  }
  return o;
}

void orc_destroy(void* h) { delete (Orc*)h; }

// The rod geometry and contact force alone, for their pins: r = {x, y, l,
// theta} per rod -> the two contact points; r = {x, y, d, l, theta, vx, vy,
// m} per rod and k = {kn, gamma_n, gamma_t, mu_cc} -> {Fx, Fy, W}.
void orc_rod_intersection(const double* r1, const double* r2, double* out4) {
  rod_intersection(Rod{r1[0], r1[1], r1[2], r1[3]}, Rod{r2[0], r2[1], r2[2], r2[3]}, out4);
}
int orc_rod_forces(const double* r1, const double* r2, const double* k, double* out3) {
  return rod_forces(r1, r2, k, out3);
}

int orc_set(void* h, int field, const void* data, int64_t n) {
  Orc& o = *(Orc*)h;
  if (n > o.nMax) return CBM_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) {
    double v = field_is_int(field) ? 0.0 : ((const double*)data)[i];
    int64_t iv = field_is_int(field) ? ((const int64_t*)data)[i] : 0;
    switch (field) {
      case CBM_F_X: case CBM_F_Y: case CBM_F_Z: o.X[field][i] = v; break;
      case CBM_F_FX: case CBM_F_FY: case CBM_F_FZ: o.F[field - CBM_F_FX][i] = v; o.Fnew[field - CBM_F_FX][i] = v; break;
      case CBM_F_VOL: o.vol[i] = v; break;
      case CBM_F_GROWTH: o.gr[i] = v; o.grNew[i] = v; break;
      case CBM_F_VBIRTH: o.vb[i] = v; o.vbNew[i] = v; break;
      case CBM_F_ID: o.id[i] = iv; break;
      case CBM_F_LF: o.lf[i] = iv; break;
      case CBM_F_VX: case CBM_F_VY: case CBM_F_VZ: o.V[field - CBM_F_VX][i] = v; break;
      case CBM_F_RADIUS: o.rad[i] = v; break;
      case CBM_F_MASS: o.mass[i] = v; break;
      case CBM_F_FATE: if (iv < 1 || iv > 2) return CBM_ERR_ARG; o.fate[i] = iv; break;
      case CBM_F_U0: case CBM_F_U1: case CBM_F_U2: case CBM_F_U3: o.U[field - CBM_F_U0][i] = v; break;
      case CBM_F_U4: case CBM_F_U5: case CBM_F_U6: case CBM_F_U7: o.U[4 + field - CBM_F_U4][i] = v; break;
      case CBM_F_THETA: o.TH[i] = v; break;
      case CBM_F_OMEGA: o.OM[i] = v; break;
      case CBM_F_DIAMETER: o.rodD[i] = v; break;
      case CBM_F_LENGTH: o.rodL[i] = v; break;
      case CBM_F_TORQUE: o.rodW[i] = v; break;
      case CBM_F_PRESSURE: o.rodP[i] = v; break;
      case CBM_F_LTARGET: o.LT[i] = v; break;
      default: return CBM_ERR_ARG;
    }
  }
  o.structure_valid = false;
  return 0;
}

int orc_get(void* h, int field, void* out, int64_t n) {
  Orc& o = *(Orc*)h;
  if (n > o.nMax) return CBM_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) {
    double* dv = (double*)out; int64_t* iv = (int64_t*)out;
    switch (field) {
      case CBM_F_X: case CBM_F_Y: case CBM_F_Z: dv[i] = o.X[field][i]; break;
      case CBM_F_FX: case CBM_F_FY: case CBM_F_FZ: dv[i] = o.F[field - CBM_F_FX][i]; break;
      case CBM_F_VOL: dv[i] = o.vol[i]; break;
      case CBM_F_GROWTH: dv[i] = o.gr[i]; break;
      case CBM_F_VBIRTH: dv[i] = o.vb[i]; break;
      case CBM_F_ID: iv[i] = o.id[i]; break;
      case CBM_F_LF: iv[i] = o.lf[i]; break;
      case CBM_F_VX: case CBM_F_VY: case CBM_F_VZ: dv[i] = o.V[field - CBM_F_VX][i]; break;
      case CBM_F_RADIUS: dv[i] = o.rad[i]; break;
      case CBM_F_MASS: dv[i] = o.mass[i]; break;
      case CBM_F_FATE: iv[i] = o.fate[i]; break;
      case CBM_F_U0: case CBM_F_U1: case CBM_F_U2: case CBM_F_U3: dv[i] = o.U[field - CBM_F_U0][i]; break;
      case CBM_F_U4: case CBM_F_U5: case CBM_F_U6: case CBM_F_U7: dv[i] = o.U[4 + field - CBM_F_U4][i]; break;
      case CBM_F_THETA: dv[i] = o.TH[i]; break;
      case CBM_F_OMEGA: dv[i] = o.OM[i]; break;
      case CBM_F_DIAMETER: dv[i] = o.rodD[i]; break;
      case CBM_F_LENGTH: dv[i] = o.rodL[i]; break;
      case CBM_F_TORQUE: dv[i] = o.rodW[i]; break;
      case CBM_F_PRESSURE: dv[i] = o.rodP[i]; break;
      case CBM_F_LTARGET: dv[i] = o.LT[i]; break;
      default: return CBM_ERR_ARG;
    }
  }
  return 0;
}

int64_t orc_n(void* h) { return ((Orc*)h)->N; }
int64_t orc_id_max(void* h) { return ((Orc*)h)->idMax; }
int64_t orc_rebuilds(void* h) { return ((Orc*)h)->rebuilds; }
double orc_t(void* h) { return ((Orc*)h)->t; }
int orc_error(void* h) { return ((Orc*)h)->err; }

// loadToPlatform! -> first computeNeighbors!
int orc_init(void* h) {
  Orc& o = *(Orc*)h;
  compute_neighbors(o, true);
  return o.err;
}

int orc_step(void* h, int64_t nsteps) {
  Orc& o = *(Orc*)h;
  if (!o.structure_valid) compute_neighbors(o, true);
  for (int64_t k = 0; k < nsteps; ++k) one_step(o);
  return o.err;
}

// Neighbor set of agent i (slots), for sorted-set comparisons.  Verlet: the
// list; CellLinked / Full: the candidates within `radius`.
int64_t orc_neighbors_of(void* h, int64_t i, double radius, int64_t* out, int64_t cap) {
  Orc& o = *(Orc*)h;
  int64_t k = 0;
  if (o.is_verlet()) {
    for (int32_t q = 0; q < o.nbN[i]; ++q) { if (k < cap) out[k] = o.nbList[(int64_t)q * o.nMax + i]; k++; }
    return k;
  }
  for_candidates(o, i, [&](int64_t j) {
    double dx[3] = {0, 0, 0};
    for (int a = 0; a < o.dims; ++a) dx[a] = o.mi(a, o.X[a][i] - o.X[a][j]);
    if (dist(o, dx) < radius) { if (k < cap) out[k] = j; k++; }
  });
  return k;
}

// agentRule of the neighbor KAT (test/testCommunity.jl:375-383).
int orc_count_neighbors(void* h, double radius, int64_t* out) {
  Orc& o = *(Orc*)h;
  for (int64_t i = 0; i < o.N; ++i) {
    int64_t c = 0;
    for_candidates(o, i, [&](int64_t j) {
      double dx[3] = {0, 0, 0};
      for (int a = 0; a < o.dims; ++a) dx[a] = o.mi(a, o.X[a][i] - o.X[a][j]);
      if (dist(o, dx) < radius) c++;
    });
    out[i] = c;
  }
  return 0;
}

int orc_get_medium(void* h, double* out, int64_t n) {
  Orc& o = *(Orc*)h;
  if (n > (int64_t)o.c.size()) return CBM_ERR_ARG;
  std::memcpy(out, o.c.data(), sizeof(double) * n);
  return 0;
}
int orc_set_medium(void* h, const double* in, int64_t n) {
  Orc& o = *(Orc*)h;
  if (n > (int64_t)o.c.size()) return CBM_ERR_ARG;
  std::memcpy(o.c.data(), in, sizeof(double) * n);
  o.cinit = o.c;
  return 0;
}

// Philox known-answer access for tests: out[4] = philox4x32-10(ctr, key).
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  U4 c; for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  U4 r = philox(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}
// CBM_RULE_CUSTOM: the rule body, called once per agent and step (test code
// passes a restatement of the device body it checks).
void orc_set_rule(void* h, void* fn) { ((Orc*)h)->rule_fn = fn; }
// A custom rule's neighbor loop: the candidates of slot i in the structure's
// walk order, and their values as the rule's loop started (field 0..2
// positions, 3 vol, 4..11 the user parameters, 12 id).
int64_t orc_rule_neighbors(void* h, int64_t i, int64_t* out, int64_t cap) {
  Orc& o = *(Orc*)h;
  int64_t k = 0;
  for_candidates(o, i, [&](int64_t j) { if (k < cap) out[k] = j; k++; });
  return k;
}
double orc_rule_value(void* h, int field, int64_t j) {
  Orc& o = *(Orc*)h;
  if (field < 3) return o.ruleX[field][j];
  if (field == 3) return o.ruleVol[j];
  if (field < 12) return o.ruleU[field - 4][j];
  return (double)o.id[j];
}

// A user mediumODE body in place of the built-in diffusion with uptake.
void orc_set_medium_body(void* h, void* fn) { ((Orc*)h)->medium_fn = fn; }
// Its assignment branches (orc_medium_assign_fn; NULL: none).
void orc_set_medium_assign(void* h, void* fn) { ((Orc*)h)->medium_assign = fn; }

// A general agentODE body: pair (orc_law_pair_fn) adds to `sums` accumulators
// per neighbor, ode (orc_law_ode_fn) gives dt of the positions and of the
// model's user_de user variables (U0..).  Returns CBM_ERR_ARG for sums
// outside 1..8.
// A model-scope body: rule (orc_model_rule_fn), ode and noise
// (orc_model_ode_fn: dp or g), any of them NULL; de = the DE parameters.
void orc_set_model_body(void* h, void* rule, void* ode, void* noise, unsigned de) {
  Orc& o = *(Orc*)h;
  o.model_rule = rule;
  o.model_ode = ode;
  o.model_noise = noise;
  o.model_de = de;
}

// The simulation's index among a batched context's simulations, which keys
// every random draw (Philox counter word `sim`).
void orc_set_sim(void* h, int32_t sim) { ((Orc*)h)->sim = sim; }

double orc_param(void* h, int k) { return (k >= 0 && k < 8) ? ((Orc*)h)->s.p[k] : 0.0; }

int orc_set_law_body(void* h, void* pair, void* ode, int sums) {
  if (sums < 1 || sums > 8) return CBM_ERR_ARG;
  Orc& o = *(Orc*)h;
  o.law_pair = pair;
  o.law_ode = ode;
  o.law_sums = sums;
  return 0;
}

void orc_normals(uint64_t seed, int64_t step, int64_t id, int32_t sim, uint32_t purpose, double* z2) {
  U4 r = draw(seed, step, id, sim, purpose);
  normal2(r, z2[0], z2[1]);
}
int64_t orc_floor_div(double a, double b) { return floor_div(a, b); }

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// OpenMP threads of the oracle's parallel loops (the CPU baseline at 1 thread).
// A state-dependent agentSDE term for the checks of user noise bodies:
// g_a(x) = noise + slope * x_a on the positions.
void orc_set_noise_slope(Orc* o, double slope) { if (o) o->noise_slope = slope; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// loss with double arithmetic; theta = (k, r_eq, eps).  Restores positions.
double orc_traj_loss(void* h, const double* theta, int64_t n_steps, int32_t n_saves,
                     const int64_t* save_steps, const double* target) {
  Orc& o = *(Orc*)h;
  std::vector<double> keep[3];
  for (int a = 0; a < o.dims; ++a) keep[a] = o.X[a];
  double t0 = o.t; int64_t s0 = o.step;
  double r = traj_loss<double>(o, theta, n_steps, n_saves, save_steps, target);
  for (int a = 0; a < o.dims; ++a) o.X[a] = keep[a];
  o.t = t0; o.step = s0; o.structure_valid = false;
  return r;
}

// loss and forward-mode gradient (dual numbers).  Restores positions.
double orc_traj_grad_dual(void* h, const double* theta, int64_t n_steps, int32_t n_saves,
                          const int64_t* save_steps, const double* target, double* grad) {
  Orc& o = *(Orc*)h;
  std::vector<double> keep[3];
  for (int a = 0; a < o.dims; ++a) keep[a] = o.X[a];
  double t0 = o.t; int64_t s0 = o.step;
  D3 th[3];
  for (int q = 0; q < 3; ++q) { th[q] = D3(theta[q]); th[q].g[q] = 1.0; }
  D3 r = traj_loss<D3>(o, th, n_steps, n_saves, save_steps, target);
  for (int q = 0; q < 3; ++q) grad[q] = r.g[q];
  for (int a = 0; a < o.dims; ++a) o.X[a] = keep[a];
  o.t = t0; o.step = s0; o.structure_valid = false;
  return r.v;
}

}  // extern "C"

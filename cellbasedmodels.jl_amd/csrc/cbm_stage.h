// cbm_stage.h -- the integrator-stage kernel body and the row/cell
// geometry helpers it needs.  Self-contained (with cbm_device.h and
// cbm_structs.h) so hipRTC can compile it around a user interaction body.
#ifndef CBM_STAGE_H
#define CBM_STAGE_H

#include "cbm_device.h"
#include "cbm_structs.h"

#ifndef CBM_LIST_CHUNK
#define CBM_LIST_CHUNK 2
#endif

namespace cbm {

// ------------------------------------------------------------ row geometry
struct RowIdx { int sim; int64_t local; int64_t row; };

// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// (own L2) serves one contiguous range of rows, i.e. one compact region of
// the Morton-ordered agents, and neighbor gathers stay in that L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, idx = b >> 3;
  return x < r ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// Row of this thread for block bx of a launch whose first nb blocks serve
// the rows.
__device__ __forceinline__ RowIdx row_index_of(const RowGeom& g, int bx, int nb) {
  RowIdx r;
  const int b = xcd_tile(bx, nb);
  r.sim = g.block_sim[b];
  r.local = (int64_t)(b - g.block_off[r.sim]) * blockDim.x + threadIdx.x;
  r.row = g.sims[r.sim].row_off + r.local;
  return r;
}

__device__ __forceinline__ RowIdx row_index(const RowGeom& g) {
  RowIdx r;
  const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  r.sim = g.block_sim[b];
  r.local = (int64_t)(b - g.block_off[r.sim]) * blockDim.x + threadIdx.x;
  r.row = g.sims[r.sim].row_off + r.local;
  return r;
}

__device__ __forceinline__ double min_image(double d, double L) {
  if (d > 0.5 * L) d -= L; else if (d < -0.5 * L) d += L;
  return d;
}

// k-th of the 3^D neighbor cells (cellPosNeigh order, x fastest) as a local
// cell index, or -1 outside a non-periodic grid.
template <int D>
__device__ __forceinline__ int32_t neigh_cell(const SimDev& sd, int32_t cx, int32_t cy, int32_t cz, int k, int periodic) {
  int32_t q[3] = {cx, cy, cz};
  const int off[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    int32_t v = q[a] + off[a];
    if (periodic) { v = v < 0 ? v + sd.gn[a] : (v >= sd.gn[a] ? v - sd.gn[a] : v); }
    else if (v < 0 || v >= sd.gn[a]) return -1;
    q[a] = v;
  }
  return q[0] + sd.gn[0] * (q[1] + sd.gn[1] * q[2]);
}

// Cell of x on the model's grid along axis a: cellPos (src/neighbors.jl:
// 636-638, 0-based) on the padded grid, clamped on the periodic one.
__device__ __forceinline__ int32_t cl_axis(const SimDev& sd, int a, double x, int periodic) {
  if (periodic) {
    int64_t c = floor_div(x - sd.lo[a], sd.cedge[a]);
    c = c < 0 ? 0 : (c > sd.cgn[a] - 1 ? sd.cgn[a] - 1 : c);
    return (int32_t)c;
  }
  if (x > sd.hi[a]) return sd.cgn[a] - 1;
  if (x < sd.lo[a]) return 0;
  return (int32_t)(floor_div(x - sd.lo[a], sd.cedge[a]) + 1);
}

// The cell packed as 10 bits per axis (grids of at most 1023 cells per axis).
template <int D>
__device__ __forceinline__ int32_t cl_pack(const SimDev& sd, const double* x, int periodic) {
  int32_t v = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) v |= cl_axis(sd, a, x[a], periodic) << (10 * a);
  return v;
}

// Position of cell pj among the 3^D neighbor cells of pi (cellPosNeigh
// order), or 27 when pj is not one of them.
template <int D, int PER>
__device__ __forceinline__ int cl_key(const SimDev& sd, int32_t pi, int32_t pj) {
  int k = 0, w = 1;
  bool adj = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    int d = ((pj >> (10 * a)) & 1023) - ((pi >> (10 * a)) & 1023);
    if (PER) d = d == sd.cgn[a] - 1 ? -1 : (d == 1 - sd.cgn[a] ? 1 : d);
    adj = adj && d >= -1 && d <= 1;
    k += (d + 1) * w;
    w *= 3;
  }
  return adj ? k : 27;
}

__host__ __device__ constexpr int pow3(int d) { return d == 1 ? 3 : (d == 2 ? 9 : 27); }

// A law with a per-agent term: `void self(const double* x, double* f) const`
// adds f (D values, zeroed before the call) to dt(x) at the stage state x,
// the agentODE body's terms that are not pair forces.
template <class L, class = void> struct has_self { static constexpr bool value = false; };
template <class L> struct has_self<L, decltype((void)&L::self)> { static constexpr bool value = true; };
template <class L, class = void> struct has_de { static constexpr bool value = false; };
template <class L> struct has_de<L, decltype((void)&L::de)> { static constexpr bool value = true; };

// A program around a body with user DE variables defines CBM_USER_DE (their
// count) and CBM_DE_BASE (the DE index of the first: dims, + 1 with vol).
#ifndef CBM_USER_DE
#define CBM_USER_DE 0
#define CBM_DE_BASE 0
#endif

// A general agentODE body (src/AgentStructure/functionDE.jl:13-116): a
// UserLaw with
//   static constexpr int sums = K;     // neighbor sums, 1..8
//   __device__ void pair(const LawAgent& i, const LawAgent& j, const double* dx, double d, double* s) const;
//   __device__ void ode(const LawAgent& i, const double* s, double* dxdt, double* dudt) const;
// The skeleton zeroes s[0..K), calls pair for every neighbor candidate j of
// agent i in reference loop order (dx = x_i - x_j, minimum image in a
// periodic box; d = |dx|), each call adding its terms to s in whatever order
// the body writes, and then ode with the sums: dxdt[0..D) = dt of the
// positions, dudt[k] = dt of user DE variable k (both zeroed before the call).
// The agents are seen at the stage state.
template <class L, class = void> struct general_body { static constexpr bool value = false; };
template <class L> struct general_body<L, decltype((void)&L::ode)> { static constexpr bool value = true; };
template <class L, class = void> struct law_sums { static constexpr int value = 1; };
template <class L> struct law_sums<L, decltype((void)L::sums)> { static constexpr int value = L::sums; };
// the Aggregation wall push, for laws that declare it
template <class L, class = void> struct law_walls { static constexpr bool value = false; };
template <class L> struct law_walls<L, decltype((void)L::walls)> {
  static constexpr bool value = L::active && L::walls;
};

// One agent of a pair as a general agentODE body sees it.
struct LawAgent {
  int64_t row, id;
  double x[3];              // position at the stage state
  double uv[USER_PARAMS];   // user parameters U0..U7: DE variables at the stage state, others as stored
  double volv;              // vol at the stage state (1 without the vol variable)
  __device__ double u(int k) const { return uv[k]; }
  __device__ double vol() const { return volv; }
};

template <int D, class Args>
__device__ __forceinline__ LawAgent law_agent(const Args& a, int64_t row, const double* x) {
  LawAgent r;
  r.row = row;
  r.id = a.id[row];
#pragma unroll
  for (int c = 0; c < 3; ++c) r.x[c] = c < D ? x[c] : 0.0;
#pragma unroll
  for (int k = 0; k < USER_PARAMS; ++k)
    r.uv[k] = k < CBM_USER_DE ? a.S[CBM_DE_BASE + k][row] : (a.upar[k] ? a.upar[k][row] : 0.0);
  r.volv = a.growth ? a.S[D][row] : 1.0;
  return r;
}

// Walk the candidate neighbors of row i in reference loop order (Args: the
// stage or rule arguments, with the list and cell tables of the last
// neighbor update).
template <int D, int NB, class Args, class Fn>
__device__ __forceinline__ void for_candidates(const Args& a, const SimDev& sd, int64_t N, int64_t i, Fn&& fn) {
  if (NB == NB_LIST) {
    const int32_t cnt = a.list_n[i];
    for (int32_t q = 0; q < cnt; ++q) fn((int64_t)a.list[(int64_t)q * a.ldl + i]);
  } else if (NB == NB_FULL) {
    const int64_t j0 = sd.row_off, j1 = sd.row_off + N;
    for (int64_t j = j0; j < j1; ++j) if (j != i) fn(j);
  } else {
    const int32_t c = a.cell_row[i];
    const int32_t cx = c % sd.gn[0];
    const int32_t cy = (c / sd.gn[0]) % sd.gn[1];
    const int32_t cz = c / (sd.gn[0] * sd.gn[1]);
    const CellRange* __restrict__ cr = a.cell_range + sd.rank_off;
    for (int k = 0; k < pow3(D); ++k) {
      const int32_t nc = neigh_cell<D>(sd, cx, cy, cz, k, a.periodic);
      if (nc < 0) continue;
      const CellRange r = cr[nc];
      for (int32_t j = r.s; j < r.e; ++j) if (j != i) fn((int64_t)j);
    }
  }
}

// Periodic box: a coordinate brought back into [lo, lo + L) (one period at most).
__device__ __forceinline__ double wrap_box(double x, double lo, double L) {
  if (x >= lo + L) return x - L;
  if (x < lo) return x + L;
  return x;
}

// ----------------------------------------------------------- stage kernel
// Minimum image (periodic boxes): the same comparisons as the oracle, as selects.
__device__ __forceinline__ double wrap_dx(double d, double L, double hL) {
  return d > hL ? d - L : (d < -hL ? d + L : d);
}

// One neighbor at position xj: the agentODE interaction body, added to acc
// in the reference's per-axis order.  live = false (list padding, or the
// agent itself) leaves acc unchanged.
template <int D, class Law, int PER, bool SOFT>
__device__ __forceinline__ void pair_into(const Law& law, const double* xi, const double* xj, const double* L,
                                          const double* hL, double rad_i, double rad_j, int64_t fate_i,
                                          int64_t fate_j, bool live, double* acc) {
  double dx[3];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    dx[c] = xi[c] - xj[c];
    if (PER) dx[c] = wrap_dx(dx[c], L[c], hL[c]);
  }
  double s2 = dx[0] * dx[0];
  if (D > 1) s2 = s2 + dx[1] * dx[1];
  if (D > 2) s2 = s2 + dx[2] * dx[2];
  const double d = sqrt(s2);
  double s;
  bool on;
  if constexpr (SOFT) on = law.pair(d, rad_i, rad_j, fate_i, fate_j, s) && live;
  else on = law(d, s) && live;
  if constexpr (SOFT) {
    // ((s dx) / d) per axis, the reference's order (softSpheres.jl:28-29)
#pragma unroll
    for (int c = 0; c < D; ++c) { const double v = acc[c] + (s * dx[c]) / d; acc[c] = on ? v : acc[c]; }
  } else if (Law::per_axis_division) {
#pragma unroll
    for (int c = 0; c < D; ++c) { const double v = acc[c] + s * (dx[c] / d); acc[c] = on ? v : acc[c]; }
  } else {
    // F = (s/d) (x_i - x_j).  An inactive pair adds 0*dx = +-0, which
    // leaves acc unchanged: acc starts at +0 and a sum is -0 only when
    // both terms are, so acc is never -0.
    const double f = s / d;
    const double fe = on ? f : 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = acc[c] + fe * dx[c];
  }
}

template <int D, int PER>
__device__ __forceinline__ void box_halves(const SimDev& sd, double* L, double* hL) {
#pragma unroll
  for (int c = 0; c < 3; ++c) { L[c] = 0.0; hL[c] = 0.0; }
  if (PER) {
#pragma unroll
    for (int c = 0; c < D; ++c) { L[c] = sd.L[c]; hL[c] = 0.5 * sd.L[c]; }
  }
}

// Pair sum of row i over its candidates in global memory (Verlet list,
// all rows, or the cell-linked walk), in reference loop order.
template <int D, int NB, class Law, int PER>
__device__ __forceinline__ void gather_global(const StageArgs& a, const SimDev& sd, int64_t N, int64_t i,
                                              const double* xi, double* acc) {
  constexpr bool SOFT = same_type<Law, PairLaw<LAW_SOFT_SPHERES>>::value;
  const Law law(sd.p);
  const double rad_i = SOFT ? a.rad[i] : 0.0;
  const int64_t fate_i = SOFT ? a.fate[i] : 0;
  double L[3], hL[3];
  box_halves<D, PER>(sd, L, hL);
  auto body = [&](const double* xj, int64_t j, bool live) {
    pair_into<D, Law, PER, SOFT>(law, xi, xj, L, hL, rad_i, SOFT ? a.rad[j] : 0.0, fate_i,
                                 SOFT ? a.fate[j] : 0, live, acc);
  };
  if (NB == NB_LIST) {
    // Verlet list in chunks of LCH gathers issued together (memory-level
    // parallelism per lane), branch-free: slots past the count read the
    // agent itself (d = 0, never active) and add nothing.  The list
    // indices of the next chunk are loaded before the current chunk's
    // positions are used.
    constexpr int LCH = CBM_LIST_CHUNK;
    const int32_t cnt = a.list_n[i];
    const int64_t ldl = a.ldl;
    const int32_t* __restrict__ lp = a.list + i;     // entry q at lp[q * ldl]
    int32_t jn[LCH];
#pragma unroll
    for (int u = 0; u < LCH; ++u) jn[u] = (u < cnt) ? lp[(int64_t)u * ldl] : (int32_t)i;
    for (int32_t q0 = 0; q0 < cnt; q0 += LCH) {
      double px[LCH][3];
      int32_t jc[LCH];
#pragma unroll
      for (int u = 0; u < LCH; ++u) {
        jc[u] = jn[u];
#pragma unroll
        for (int c = 0; c < D; ++c) px[u][c] = a.S[c][jc[u]];
      }
      lp += LCH * ldl;
#pragma unroll
      for (int u = 0; u < LCH; ++u) jn[u] = (q0 + LCH + u < cnt) ? lp[(int64_t)u * ldl] : (int32_t)i;
#pragma unroll
      for (int u = 0; u < LCH; ++u) body(px[u], jc[u], q0 + u < cnt);
    }
  } else {
    for_candidates<D, NB>(a, sd, N, i, [&](int64_t j) {
      double xj[3];
#pragma unroll
      for (int c = 0; c < D; ++c) xj[c] = a.S[c][j];
      body(xj, j, true);
    });
  }
}

// The rest of the stage for row i once the pair sum acc is known: walls,
// the derivative of every DE variable, the integrator update, periodic
// wrap, the output write and the displacement check.
// Noise amplitude of the agentSDE part.  ConstNoise is the model's constant g
// on the positions (StageArgs::noise).  A user body (hipRTC) defines
//   struct UserNoise {
//     __device__ explicit UserNoise(const double* p);          // the simulation's p[8]
//     __device__ double operator()(int var, const double* u) const;
//   };
// giving g of DE variable var at the agent's state u (positions, then the
// other DE variables), evaluated at the stage state as the reference's g
// (src/integrators.jl:709-727 EM, :775-803 EulerHeun).
struct ConstNoise {
  __device__ explicit ConstNoise(const double*) {}
  __device__ double operator()(int, const double*) const { return 0.0; }
};

template <int D, class Law, int NOISE, bool PRE = false, class Noise = ConstNoise>
__device__ __forceinline__ void stage_finish(const StageArgs& a, const SimDev& sd, int sim, int64_t i,
                                             const double* xi, double* acc, const double* upre,
                                             double disp2 = -1.0, double* e2_out = nullptr) {
  constexpr bool SOFT = same_type<Law, PairLaw<LAW_SOFT_SPHERES>>::value;
  constexpr bool RODS = same_type<Law, RodLaw>::value;
  // u of variable c (with PRE, the caller loaded the positions' u early into
  // upre), and the stage state, whose positions are xi
  const double u0 = upre[0], u1 = upre[1], u2 = upre[2];   // values, so no pointer is selected
  auto Uc = [&](int c) -> double {
    if (PRE && c < D) return c == 0 ? u0 : (c == 1 ? u1 : u2);
    return a.U[c][i];
  };
  auto Sc = [&](int c) { return c < D ? xi[c] : a.S[c][i]; };
  if (law_walls<Law>::value && a.walls && sd.p[4] != 0.0) {   // p[4]: the walls switch
    const double rRep = sd.p[0], fRep = sd.p[1];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double x = xi[c];
      if (x < sd.lo[c] + rRep / 2) acc[c] = acc[c] + fRep;
      else if (x > sd.hi[c] - rRep / 2) acc[c] = acc[c] - fRep;
    }
  }

  // derivative of each DE variable
  constexpr bool PAIRB = general_body<Law>::value;
  constexpr bool UDE = CBM_USER_DE > 0 && (has_de<Law>::value || PAIRB);
  constexpr int NS = (UDE || RODS) ? MAX_STATE : BUILTIN_STATE;
  double du[NS];
  double dudt[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (PAIRB) {
    // a general agentODE body: its equations from the neighbor sums acc
    const LawAgent me = law_agent<D>(a, i, xi);
    double dxdt[3] = {0.0, 0.0, 0.0};
    const Law law(sd.p);
    law.ode(me, acc, dxdt, dudt);
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double v = dxdt[c];
      if (a.drift != 0.0) v = v - a.drift * xi[c];
      du[c] = v;
    }
  } else if constexpr (RODS) {
    // acc = {fx, fy, W, pressure}; the DE variables x, y, vx, vy, theta,
    // omega (src/models/rods.jl:111-116)
    const double beta = sd.p[5], bw = sd.p[6];
    const double m = a.mass[i], d = a.rod_d[i], l = (a.nvar == 7 ? a.S[6] : a.rod_l)[i];
    const double vx = a.S[2][i], vy = a.S[3][i], om = a.S[5][i];
    du[0] = vx;
    du[1] = vy;
    du[2] = -beta * vx + acc[0] / m;
    du[3] = -beta * vy + acc[1] / m;
    du[4] = om;
    du[5] = acc[2] / (m * ((d + l) * (d + l)) / 12 + m * (d * d)) - bw * om;
    // rods2dGrowth (src/models/rods.jl:137-139): the length grows, slowed by
    // this evaluation's pressure
    if (a.nvar == 7) du[6] = a.rod_growth / (1 + a.rod_alpha * acc[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) a.rod_out[k][i] = acc[k];
  } else if constexpr (SOFT) {
    // dt(v) = -b*v/m + f/m, dt(x) = v (src/models/softSpheres.jl:32-35)
    const double b = sd.p[0], m = a.mass[i];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double f = a.lagged ? a.F[c][i] : acc[c];
      const double v = a.S[D + c][i];
      du[c] = v;
      du[D + c] = -b * v / m + f / m;
      if (a.write_fnew) a.Fnew[c][i] = acc[c];
    }
  } else {
    double fs[3] = {0.0, 0.0, 0.0};
    if constexpr (has_self<Law>::value) {
      double xs[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < D; ++c) xs[c] = xi[c];
      const Law law(sd.p);
      law.self(xs, fs);
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double v = a.lagged ? a.F[c][i] : acc[c];
      if (a.drift != 0.0) v = v - a.drift * xi[c];
      if constexpr (has_self<Law>::value) v = v + fs[c];
      du[c] = v;
      if (a.write_fnew) a.Fnew[c][i] = acc[c];
    }
  }
  const bool grow = UDE ? CBM_DE_BASE > D : !SOFT && !RODS && a.nvar > D;
  if (grow) du[D] = a.growth[i] * a.S[D][i];
  if constexpr (UDE) {
    // dt of the user DE variables at the stage state
    if constexpr (PAIRB) {
#pragma unroll
      for (int k = 0; k < CBM_USER_DE; ++k) du[CBM_DE_BASE + k] = dudt[k];
    } else {
      double xs[3] = {0.0, 0.0, 0.0}, us[4] = {0.0, 0.0, 0.0, 0.0}, ds[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < D; ++c) xs[c] = xi[c];
#pragma unroll
      for (int k = 0; k < CBM_USER_DE; ++k) us[k] = a.S[CBM_DE_BASE + k][i];
      const Law law(sd.p);
      law.de(xs, us, ds);
#pragma unroll
      for (int k = 0; k < CBM_USER_DE; ++k) du[CBM_DE_BASE + k] = ds[k];
    }
  }

  const double dt = a.dt;
  const int nv = UDE ? CBM_DE_BASE + CBM_USER_DE : (RODS ? a.nvar : (SOFT ? 2 * D : (grow ? D + 1 : D)));
  double res[NS];
  if (NOISE && !same_type<Noise, ConstNoise>::value) {
    // a user agentSDE body: g of every DE variable at the stage state, one
    // normal per variable (pairs keyed pur, pur + 1, pur + 2); the
    // EulerHeun corrector averages g of both stages, g of stage 1 kept in gacc
    const int64_t id = a.id[i];
    const int64_t step = *a.step_ptr;
    const uint32_t pur = a.kind == ST_EH2 ? 0x500u : 0x100u;
    double z[NS], us[NS];
#pragma unroll
    for (int k = 0; k < NS / 2; ++k) {
      z[2 * k] = 0.0;
      z[2 * k + 1] = 0.0;
      if (2 * k < nv) normal2(draw(a.seed, step, id, sim, pur + (uint32_t)k), z[2 * k], z[2 * k + 1]);
    }
#pragma unroll
    for (int v = 0; v < NS; ++v) us[v] = v < nv ? Sc(v) : 0.0;
    const Noise gfn(sd.p);
#pragma unroll
    for (int v = 0; v < NS; ++v) {
      if (v >= nv) continue;
      const double g = gfn(v, us);
      const double u = v < D ? Uc(v) : a.U[v][i];
      if (a.kind == ST_EH2) {
        res[v] = u + ((a.kacc[v][i] + du[v]) * dt + (a.gacc[v][i] + g) * z[v] * a.sqdt) / 2;
      } else {
        res[v] = (u + du[v] * dt) + g * z[v] * a.sqdt;
        if (a.kind == ST_EH1) { a.kacc[v][i] = du[v]; a.gacc[v][i] = g; }
      }
    }
  } else if (NOISE) {
    // EM and the EulerHeun predictor: (u + f dt) + (g xi) sqrt(dt), xi keyed
    // 0x100/0x101; the EulerHeun corrector draws fresh noise (0x500/0x501)
    double z[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t id = a.id[i];
    const int64_t step = *a.step_ptr;
    const uint32_t pur = a.kind == ST_EH2 ? 0x500u : 0x100u;
    normal2(draw(a.seed, step, id, sim, pur), z[0], z[1]);
    if (D > 2) normal2(draw(a.seed, step, id, sim, pur + 1u), z[2], z[3]);
    if (a.kind == ST_EH2) {
#pragma unroll
      for (int c = 0; c < D; ++c)
        res[c] = Uc(c) + ((a.kacc[c][i] + du[c]) * dt + (a.noise + a.noise) * z[c] * a.sqdt) / 2;
      for (int v = D; v < nv; ++v) res[v] = a.U[v][i] + ((a.kacc[v][i] + du[v]) * dt) / 2;
    } else {
#pragma unroll
      for (int c = 0; c < D; ++c) res[c] = (Uc(c) + du[c] * dt) + a.noise * z[c] * a.sqdt;
      for (int v = D; v < nv; ++v) res[v] = a.U[v][i] + du[v] * dt;   // no noise off the positions
      if (a.kind == ST_EH1)
        for (int c = 0; c < nv; ++c) a.kacc[c][i] = du[c];
    }
  } else {
    // variable c from its u; positions first (c known at compile time, so
    // an early-loaded u stays in registers), then the other DE variables
    auto upd = [&](int c, double u) -> double {
      switch (a.kind) {
        case ST_EULER:
        case ST_HEUN1:
          return u + du[c] * dt;
        case ST_HEUN2:
          return (Sc(c) + (u + du[c] * dt)) / 2;
        case ST_RK4_1:
          a.kacc[c][i] = du[c];
          return u + du[c] * dt / 2;
        case ST_RK4_2:
          a.kacc[c][i] = a.kacc[c][i] + 2 * du[c];
          return u + du[c] * dt / 2;
        case ST_RK4_3:
          a.kacc[c][i] = a.kacc[c][i] + 2 * du[c];
          return u + du[c] * dt;
        default:  // ST_RK4_4
          return u + (a.kacc[c][i] + du[c]) * dt / 6;
      }
    };
#pragma unroll
    for (int c = 0; c < D; ++c) res[c] = upd(c, Uc(c));
    for (int v = D; v < nv; ++v) res[v] = upd(v, a.U[v][i]);
  }

  const bool final_stage = a.kind == ST_EULER || a.kind == ST_EM || a.kind == ST_HEUN2 ||
                           a.kind == ST_RK4_4 || a.kind == ST_EH2;
  if (final_stage && a.periodic) {
#pragma unroll
    for (int c = 0; c < D; ++c) res[c] = wrap_box(res[c], sd.lo[c], sd.L[c]);
  }
  for (int c = 0; c < nv; ++c) a.out[c][i] = res[c];
  if (a.clv) {
    // CellLinked on a skin list: the next step's cells, or this stage's
    // distance from the list's reference positions
    if (final_stage && a.ccell_out) a.ccell_out[i] = cl_pack<D>(sd, res, a.periodic);
    if (!final_stage && e2_out) {
      double dx[3];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        dx[c] = res[c] - a.xref[c][i];
        if (a.periodic) dx[c] = min_image(dx[c], sd.L[c]);
      }
      double s2 = dx[0] * dx[0];
      if (D > 1) s2 = s2 + dx[1] * dx[1];
      if (D > 2) s2 = s2 + dx[2] * dx[2];
      *e2_out = s2;
    }
  }
  if (disp2 >= 0.0) {
    // the stage-2 lists stay exact while no agent moves more than half their margin
    double s2 = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) { const double d = res[c] - Uc(c); s2 = s2 + d * d; }
    if (!(s2 <= disp2)) *a.lgate = 0;
  }

  if (final_stage && (a.disp_check || (a.clv && a.efin))) {
    double dx[3];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dx[c] = res[c] - a.xref[c][i];
      if (a.periodic) dx[c] = min_image(dx[c], sd.L[c]);
    }
    double s2 = dx[0] * dx[0];
    if (D > 1) s2 = s2 + dx[1] * dx[1];
    if (D > 2) s2 = s2 + dx[2] * dx[2];
    double st = 0.0;   // this step's displacement, squared
    if (a.clv) {
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double d = res[c] - Uc(c);
        if (a.periodic) d = min_image(d, sd.L[c]);
        st = st + d * d;
      }
    }
    if (a.disp_check) {
      if (s2 > a.thr2) *a.flag = 1;   // sqrt(s2) >= thr, src/neighbors.jl:403 (>= skin/2)
      // CellLinked on a skin list rebuilds one step early: when this step's
      // displacement, repeated 1.5 times, would carry the agent past the
      // threshold, so that the next step's stages rarely leave half the skin
      if (a.clv && !(sqrt(s2) + 1.5 * sqrt(st) <= a.clv_half)) *a.flag = 1;
    }
    if (a.clv && a.efin) {
      // the displacement bound of the next step's stage 1 (agents past half
      // the skin; a value at or below the running maximum adds nothing), and
      // the host's cue to schedule a build: an agent that clv_ahead more
      // displacements like this step's would carry past half the skin
      if (s2 > a.clv_half2) {
        const unsigned long long e = (unsigned long long)__double_as_longlong(s2);
        if (e > *(volatile unsigned long long*)a.efin) atomicMax(a.efin, e);
      }
      if (!(sqrt(s2) + a.clv_ahead * sqrt(st) <= a.clv_half)) a.efin[1] = 1ull;
    }
  }
}

// One integrator stage for every agent row: the agentODE body over the
// neighbor candidates (Law = the pair law functor) fused with the stage
// update.  Built-in laws instantiate it in libcbm_hip.so; user laws are
// compiled around it with hipRTC.
template <int D, int NB, class Law, int PER, int NOISE, class Noise = ConstNoise>
__device__ __forceinline__ void stage_body(const StageArgs& a) {
  if (a.gate && *a.gate != a.gate_run) return;
  if (NB == NB_CELL) {
    // cell-linked: only blocks holding rows of bricks too dense for LDS
    const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    if (!a.blk_flag[b]) return;
    __syncthreads();
    if (threadIdx.x == 0) a.blk_flag[b] = 0;
  }
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  const int64_t i = ri.row;
  // cell-linked: the brick kernel has done every row of a brick that fit in
  // LDS; the others are marked, and the mark is cleared here
  if (NB == NB_CELL) {
    if (ri.local >= sd.n_max || !a.brick_skip[i]) return;
    a.brick_skip[i] = 0;
  }
  if (ri.local >= N) return;

  double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) xi[c] = a.S[c][i];
  double acc[3] = {0.0, 0.0, 0.0};
  if (Law::active) gather_global<D, NB, Law, PER>(a, sd, N, i, xi, acc);
  stage_finish<D, Law, NOISE, false, Noise>(a, sd, ri.sim, i, xi, acc, xi);
}

// One integrator stage of the rod model for every agent row: the contact
// sums over the neighbor candidates in reference loop order (pairs out of
// contact add exactly nothing: the reference adds +0 to sums that start at
// +0), then the stage update.  2-D, no periodic box.
template <int NB>
__device__ __forceinline__ void rod_stage_body(const StageArgs& a) {
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  const int64_t i = ri.row;
  if (ri.local >= N) return;
  const double kn = sd.p[0], gn = sd.p[1], gt = sd.p[2], mucc = sd.p[3];
  const double* __restrict__ lcol = a.nvar == 7 ? a.S[6] : a.rod_l;   // l is a DE variable with growth
  auto rod_at = [&](int64_t j) {
    RodState r;
    r.a.x = a.S[0][j];
    r.a.y = a.S[1][j];
    r.vx = a.S[2][j];
    r.vy = a.S[3][j];
    sincos(a.S[4][j], &r.a.s, &r.a.c);
    r.a.l = lcol[j];
    r.d = a.rod_d[j];
    r.m = a.mass[j];
    return r;
  };
  const RodState me = rod_at(i);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  bool domain = false;
  for_candidates<2, NB>(a, sd, N, i, [&](int64_t j) {
    const RodState other = rod_at(j);
    double fx, fy, w;
    if (rod_force(me, other, kn, gn, gt, mucc, fx, fy, w, domain)) {
      acc[0] = acc[0] + fx;
      acc[1] = acc[1] + fy;
      acc[2] = acc[2] + w;
      acc[3] = acc[3] + sqrt(fx * fx + fy * fy);
    }
  });
  if (domain) *a.rod_err = 1;
  double xi[3] = {me.a.x, me.a.y, 0.0};
  stage_finish<2, RodLaw, 0>(a, sd, ri.sim, i, xi, acc, xi);
}

// One integrator stage of a general agentODE body (UserLaw::ode) for every agent
// row: the body's neighbor sums over the candidates in reference loop order
// (Full, the cell-linked walk or the Verlet list), then its equations.
template <int D, int NB, class Law, int PER, int NOISE, class Noise>
__device__ __forceinline__ void pair_stage_body(const StageArgs& a) {
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  const int64_t i = ri.row;
  if (ri.local >= N) return;
  double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) xi[c] = a.S[c][i];
  const LawAgent me = law_agent<D>(a, i, xi);
  constexpr int K = law_sums<Law>::value;
  static_assert(K >= 1 && K <= 8, "UserLaw::sums is 1..8");
  double s[K];
#pragma unroll
  for (int k = 0; k < K; ++k) s[k] = 0.0;
  double L[3], hL[3];
  box_halves<D, PER>(sd, L, hL);
  const Law law(sd.p);
  for_candidates<D, NB>(a, sd, N, i, [&](int64_t j) {
    double xj[3] = {0.0, 0.0, 0.0}, dx[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < D; ++c) {
      xj[c] = a.S[c][j];
      dx[c] = xi[c] - xj[c];
      if (PER) dx[c] = wrap_dx(dx[c], L[c], hL[c]);
    }
    double s2 = dx[0] * dx[0];
    if (D > 1) s2 = s2 + dx[1] * dx[1];
    if (D > 2) s2 = s2 + dx[2] * dx[2];
    const LawAgent nb = law_agent<D>(a, j, xj);
    law.pair(me, nb, dx, sqrt(s2), s);
  });
  stage_finish<D, Law, NOISE, false, Noise>(a, sd, ri.sim, i, xi, s, xi);
}

// The stage and brick kernels of a hipRTC module: a general agentODE body
// walks its candidates row by row (no brick tiles: its brick kernel returns
// at once and the row kernel takes every row); a pair law takes the
// built-in skeletons.
template <int D, int NB, class Law, int PER, int NOISE, class Noise>
__device__ __forceinline__ void user_stage(const StageArgs& a) {
  if constexpr (general_body<Law>::value) pair_stage_body<D, NB, Law, PER, NOISE, Noise>(a);
  else stage_body<D, NB, Law, PER, NOISE, Noise>(a);
}

// ------------------------------------------------ flattened list stage
// The pair term of one list entry: what pair_into adds to the accumulator,
// returned instead (+0 for an inactive pair; the accumulator never holds
// -0, so adding +0 or -0 leaves it unchanged and the sums equal pair_into's).
template <int D, class Law, int PER>
__device__ __forceinline__ void pair_term(const Law& law, const double* xi, const double* xj, const double* L,
                                          const double* hL, double* t) {
  double dx[3];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    dx[c] = xi[c] - xj[c];
    if (PER) dx[c] = wrap_dx(dx[c], L[c], hL[c]);
  }
  double s2 = dx[0] * dx[0];
  if (D > 1) s2 = s2 + dx[1] * dx[1];
  if (D > 2) s2 = s2 + dx[2] * dx[2];
  const double d = sqrt(s2);
  double s;
  const bool on = law(d, s);
  if (Law::per_axis_division) {
#pragma unroll
    for (int c = 0; c < D; ++c) t[c] = on ? s * (dx[c] / d) : 0.0;
  } else {
    const double f = s / d;
    const double fe = on ? f : 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) t[c] = fe * dx[c];
  }
}

// Exclusive block scan of one int32 per thread; the total lands in *total.
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t* wsum, int32_t* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = (blockDim.x + 63) >> 6;
  int32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63 || tid == (int)blockDim.x - 1) wsum[w] = incl;
  __syncthreads();
  int32_t before = 0, tot = 0;
  for (int q = 0; q < nw; ++q) {
    const int32_t ws = wsum[q];
    if (q < w) before += ws;
    tot += ws;
  }
  if (tid == 0) *total = tot;
  return before + incl - v;
}

// ------------------------------------------------ cell-linked brick tiles
// The cell-linked stage on LDS tiles.  Block b serves one brick of
// bdim[0] x bdim[1] x bdim[2] grid cells, a Morton-aligned block (its edges
// are powers of two with x >= y >= z >= x/2), so that the rows of its cells
// are one contiguous range.  The agents of the brick and of its one-cell
// halo are staged in LDS in halo-cell order (x fastest): the three
// x-adjacent cells of a neighbor column are then one contiguous LDS run, and
// each home agent walks its 3^D cells there in the reference order
// (cellPosNeigh, x fastest; the rows of a cell by slot).  A brick whose halo
// holds more than brick_cap agents marks its rows in brick_skip and leaves
// them to the row-parallel stage kernel launched after it, which walks
// global memory in the same order, so the result does not depend on the
// path.

// Reach of a built-in law: a pair can be active only when d < reach (all of
// them also need d > 0).  Laws without it (user bodies) are not filtered.
template <class L, class = void> struct has_reach { static constexpr bool value = false; };
template <class L> struct has_reach<L, decltype((void)&L::reach)> { static constexpr bool value = true; };

// A bound on d^2 that keeps every pair with d = sqrt(d2) < reach: sqrt is
// correctly rounded, so d < reach implies d2 < reach^2 (1 + 2^-51), and the
// product below exceeds that.  A NaN or infinite reach keeps every pair.
__device__ __forceinline__ double reach2_bound(double reach) {
  const double r2 = (reach * reach) * (1.0 + 0x1p-44);
  return (r2 == r2) ? r2 : __builtin_inf();
}

// Order-preserving int64 key of a double (for LDS atomicMax).
__device__ __forceinline__ long long order_key(double x) {
  const long long b = __double_as_longlong(x);
  return b < 0 ? (long long)(~(unsigned long long)b | 0x8000000000000000ull) : b;
}
__device__ __forceinline__ double order_val(long long k) {
  return __longlong_as_double(k < 0 ? (long long)(~(unsigned long long)k | 0x8000000000000000ull) : k);
}


// One brick's LDS tile.  The caller points hdr, hs, hrow, owner and home_li
// into its LDS; brick_stage fills the rest.
struct BrickTile {
  int sim;
  int32_t o[3];                   // origin cell
  int B[3], HX, HY, H;            // brick edge, halo extent, halo cells
  int32_t r0, r1, T;              // home rows [r0, r1); staged agents
  int32_t* hdr;                   // BRICK_HDR int32 words of scratch
  int32_t* hs;                    // LDS offset of halo cell h's agents, hs[H] = T
  int32_t* hrow;                  // first global row of halo cell h
  uint16_t* owner;                // halo cell of each staged agent, | 0x8000 at its home cell
  uint16_t* home_li;              // LDS index of home row r0 + k
};

// What brick_stage copies per agent besides the positions: an optional
// double column (radius; its block maximum is returned, +inf if any is NaN)
// and an optional integer column read as int64 (fate) or int32 (slot).
enum { BX_NONE = 0, BX_I64 = 1, BX_I32 = 2 };
struct BrickCols {
  const double* S[3];       // positions, S[c][row]
  double* pos[3];           // their LDS arrays
  const double* xd;         // double column (or NULL)
  double* xd_lds;
  const void* xi;           // integer column
  int32_t* xi_lds;
};

// Stages brick b: the rows and counts of its halo cells (wrapped on a
// periodic grid, empty outside a bounded one), their exclusive scan, then
// the columns of every halo agent, U4 agents per thread in flight, at its
// LDS index q (halo-cell order, x fastest).  Returns 0 for an empty brick, 2
// for one whose halo exceeds cap (its rows are then marked for the
// row-parallel kernel), 1 when staged.
template <int D, int PER, int U4, bool XD, int XI>
__device__ __forceinline__ int brick_stage(const BrickSrc& src, int b, BrickTile& t, const BrickCols& cols,
                                           double& xd_max) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int32_t* bt = src.brick + BRICK_WORDS * (int64_t)b;
  t.sim = bt[0];
  t.o[0] = bt[1]; t.o[1] = bt[2]; t.o[2] = bt[3];
  const SimDev& sd = src.sims[t.sim];
  // the brick's cells hold consecutive buckets [b0, b1): its rows are one range
  t.r0 = src.bucket_start[bt[4]];
  t.r1 = src.bucket_start[bt[5]];
  if (t.r0 >= t.r1) return 0;
  // a sub-brick whose main brick was staged: nothing left here (a main
  // brick marks all its rows or none, and this brick lies inside one)
  if (src.sub && !src.brick_skip[t.r0]) return 0;
  t.B[0] = src.bdim[0]; t.B[1] = D > 1 ? src.bdim[1] : 1; t.B[2] = D > 2 ? src.bdim[2] : 1;
  t.HX = t.B[0] + 2;
  t.HY = D > 1 ? t.B[1] + 2 : 1;
  t.H = brick_halo_cells(D, t.B[0], t.B[1], t.B[2]);
  const int HX = t.HX, HY = t.HY, H = t.H;
  int32_t* hs = t.hs;
  int32_t* hrow = t.hrow;
  const CellRange* __restrict__ cr = src.cell_range + sd.rank_off;

  // halo cell h -> (hx, hy, hz), x fastest.  h + 0.5 lies at least 0.5 from
  // a multiple of the divisor, far beyond the float product's rounding for
  // h < 2^16, so the truncated products are the exact quotients.
  const float rHX = 1.0f / (float)HX, rHXY = 1.0f / (float)(HX * HY);
  auto halo_cell = [&](int h, bool& home) {
    const int hz = (int)(((float)h + 0.5f) * rHXY);
    const int rem = h - hz * HX * HY;
    const int hy = (int)(((float)rem + 0.5f) * rHX);
    const int hh[3] = {rem - hy * HX, hy, hz};
    bool ok = true;
    home = true;
    int32_t q[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < D; ++e) {
      int32_t v = t.o[e] + hh[e] - 1;
      home = home && hh[e] >= 1 && hh[e] <= t.B[e] && v < sd.gn[e];
      if (PER) {
        if (v == -1) v = sd.gn[e] - 1;
        else if (v == sd.gn[e]) v = 0;
        else if (v > sd.gn[e]) ok = false;
      } else {
        ok = ok && v >= 0 && v < sd.gn[e];
      }
      q[e] = v;
    }
    return ok ? q[0] + sd.gn[0] * (q[1] + sd.gn[1] * q[2]) : -1;
  };
  // 1. rows and counts of the halo cells, U4 cells per thread in flight
  for (int h0 = tid; h0 < H; h0 += U4 * nt) {
    CellRange r[U4];
#pragma unroll
    for (int u = 0; u < U4; ++u) {
      bool home;
      const int h = h0 + u * nt;
      const int32_t lin = h < H ? halo_cell(h, home) : -1;
      r[u] = lin >= 0 ? cr[lin] : CellRange{0, 0};
    }
#pragma unroll
    for (int u = 0; u < U4; ++u) {
      const int h = h0 + u * nt;
      if (h < H) { hs[h] = r[u].e - r[u].s; hrow[h] = r[u].s; }
    }
  }
  __syncthreads();

  // 2. exclusive scan of the halo counts (a contiguous chunk per thread)
  {
    const int K = (H + nt - 1) / nt;
    const int c0 = min(tid * K, H), c1 = min(c0 + K, H);
    int32_t sum = 0;
    for (int h = c0; h < c1; ++h) sum += hs[h];
    const int lane = tid & 63, w = tid >> 6, nw = (nt + 63) >> 6;
    int32_t incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t y = __shfl_up(incl, off, 64);
      if (lane >= off) incl += y;
    }
    if (lane == 63) t.hdr[4 + w] = incl;
    __syncthreads();
    int32_t run = incl - sum, tot = 0;
    for (int q = 0; q < nw; ++q) {
      const int32_t ws = t.hdr[4 + q];
      if (q < w) run += ws;
      tot += ws;
    }
    for (int h = c0; h < c1; ++h) { const int32_t v = hs[h]; hs[h] = run; run += v; }
    if (tid == 0) { hs[H] = tot; t.hdr[2] = tot; }
  }
  __syncthreads();
  t.T = t.hdr[2];
  const int32_t r0 = t.r0, r1 = t.r1;
  if (t.T > src.cap) {
    // too dense for LDS: the row-parallel kernel launched after this one
    // takes the brick's rows
    if (tid == 0 && !src.sub) atomicAdd(src.brick_over, 1ull);
    for (int64_t i = (int64_t)r0 + tid; i < (int64_t)r1; i += nt) src.brick_skip[i] = 1;
    // wake the row-kernel blocks holding these rows (same block size)
    const int64_t q0 = src.block_off[t.sim] + ((int64_t)r0 - sd.row_off) / nt;
    const int64_t q1 = src.block_off[t.sim] + ((int64_t)r1 - 1 - sd.row_off) / nt;
    for (int64_t q = q0 + tid; q <= q1; q += nt) src.blk_flag[q] = 1;
    return 2;
  }

  // 3. stage the halo agents; each home row notes the LDS index of its copy
  //    at its home cell
  for (int h = tid; h < H; h += nt) {
    bool home;
    (void)halo_cell(h, home);
    const uint16_t tag = (uint16_t)(h | (home ? 0x8000 : 0));
    for (int32_t q = hs[h]; q < hs[h + 1]; ++q) t.owner[q] = tag;
  }
  __syncthreads();
  const int32_t T = t.T;
  double xm = -__builtin_inf();
  bool xnan = false;
  for (int32_t q0 = tid; q0 < T; q0 += U4 * nt) {
    int64_t row[U4];
    int tag[U4];
    double v[U4][3], vd[U4];
    int32_t vi[U4];
#pragma unroll
    for (int u = 0; u < U4; ++u) {
      const int32_t q = q0 + u * nt;
      tag[u] = q < T ? t.owner[q] : 0;
      const int h = tag[u] & 0x7FFF;
      row[u] = q < T ? (int64_t)hrow[h] + (q - hs[h]) : -1;
      const int64_t rr = row[u] < 0 ? (int64_t)r0 : row[u];
#pragma unroll
      for (int c = 0; c < D; ++c) v[u][c] = cols.S[c][rr];
      if (XD) vd[u] = cols.xd[rr];
      if (XI == BX_I64) vi[u] = (int32_t)((const int64_t*)cols.xi)[rr];
      if (XI == BX_I32) vi[u] = ((const int32_t*)cols.xi)[rr];
    }
#pragma unroll
    for (int u = 0; u < U4; ++u) {
      const int32_t q = q0 + u * nt;
      if (q >= T) continue;
#pragma unroll
      for (int c = 0; c < D; ++c) cols.pos[c][q] = v[u][c];
      if (XD) {
        cols.xd_lds[q] = vd[u];
        xm = vd[u] > xm ? vd[u] : xm;
        xnan = xnan || vd[u] != vd[u];
      }
      if (XI != BX_NONE) cols.xi_lds[q] = vi[u];
      if (tag[u] & 0x8000) t.home_li[row[u] - r0] = (uint16_t)q;
    }
  }
  if (XD) {
    // block maximum of the double column through an order-preserving key
    long long* key = (long long*)(t.hdr + 12);
    if (tid == 0) *key = order_key(-__builtin_inf());
    __syncthreads();
    atomicMax(key, xnan ? order_key(__builtin_inf()) : order_key(xm));
  }
  __syncthreads();
  if (XD) xd_max = order_val(*(const long long*)(t.hdr + 12));
  return 1;
}

// Candidates of a home agent in halo cell h: its 3^(D-1) runs of three
// x-adjacent cells (z outer, then y); run (ry, rz) covers cells hb .. hb + 2.
template <int D>
__device__ __forceinline__ int brick_run0(const BrickTile& t, int h) {
  return h - 1 - (D > 1 ? t.HX : 0) - (D > 2 ? t.HX * t.HY : 0);
}
template <int D>
__device__ __forceinline__ int32_t brick_ncand(const BrickTile& t, int h) {
  const int hb0 = brick_run0<D>(t, h);
  int32_t n = 0;
#pragma unroll
  for (int rz = 0; rz < (D > 2 ? 3 : 1); ++rz)
#pragma unroll
    for (int ry = 0; ry < (D > 1 ? 3 : 1); ++ry) {
      const int hb = hb0 + rz * t.HX * t.HY + ry * t.HX;
      n += t.hs[hb + 3] - t.hs[hb];
    }
  return n;
}
// Visits the candidates of a home agent in halo cell h in walk order, as one
// flattened sequence, so the lanes of a wavefront stay in step.
template <int D, class Visit>
__device__ __forceinline__ void brick_walk(const BrickTile& t, int h, Visit&& visit) {
  const int32_t ncand = brick_ncand<D>(t, h);
  int ry = 0;
  int hb = brick_run0<D>(t, h);
  int32_t q = t.hs[hb], qe = t.hs[hb + 3];
  for (int32_t k = 0; k < ncand; ++k) {
    while (q == qe) {   // the next non-empty run; one exists while k < ncand
      if (++ry == 3) { ry = 0; hb += t.HX * t.HY - 2 * t.HX; } else { hb += t.HX; }
      q = t.hs[hb];
      qe = t.hs[hb + 3];
    }
    visit(q);
    ++q;
  }
}

template <int D, class Law, int PER, int NOISE, class Noise = ConstNoise>
__device__ __forceinline__ void brick_body(const StageArgs& a, unsigned char* lds) {
  constexpr bool SOFT = same_type<Law, PairLaw<LAW_SOFT_SPHERES>>::value;
  // Built-in laws walk the candidates twice: a cheap pass keeps the LDS
  // indices of the pairs within the law's reach (in walk order, up to
  // BRICK_LIST per lane in an LDS queue), and the full pair body runs on
  // those only.  Every pair left out is inactive, so the sum is unchanged.
  constexpr bool FILTER = SOFT || has_reach<Law>::value;
  constexpr int U4 = 4;               // loads issued together per thread while staging
  // stage-2 lists (built-in laws without per-agent radii)
  constexpr bool WL = FILTER && !SOFT;
  if (a.gate && *a.gate != a.gate_run) return;
  const bool wl = WL && a.wlist != nullptr;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int cap = a.brick_cap;
  BrickSrc src;
  src.sims = a.g.sims; src.block_off = a.g.block_off; src.brick = a.brick; src.bucket_start = a.bucket_start;
  src.cell_range = a.cell_range; src.cap = cap;
  for (int e = 0; e < 3; ++e) src.bdim[e] = a.bdim[e];
  src.brick_skip = a.brick_skip; src.blk_flag = a.blk_flag; src.brick_over = a.brick_over;
  src.sub = a.brick_sub;
  const int H = brick_halo_cells(D, a.bdim[0], D > 1 ? a.bdim[1] : 1, D > 2 ? a.bdim[2] : 1);

  BrickTile t;
  t.hdr = (int32_t*)lds;
  double* pos[3];
  unsigned char* p = lds + 4 * BRICK_HDR;
#pragma unroll
  for (int c = 0; c < D; ++c) { pos[c] = (double*)p; p += (size_t)8 * cap; }
  double* rad = (double*)p;
  if (SOFT) p += (size_t)8 * cap;
  int32_t* fate = (int32_t*)p;
  if (SOFT) p += (size_t)4 * cap;
  t.hs = (int32_t*)p;
  p += (size_t)4 * (H + 1);
  t.hrow = (int32_t*)p;
  p += (size_t)4 * H;
  uint16_t* lst = (uint16_t*)p;       // per-lane queue of kept candidates, entry u at [u * nt + tid]
  p += (size_t)2 * BRICK_LIST * BRICK_MAX_BLOCK;
  t.owner = (uint16_t*)p;
  p += (size_t)2 * cap;
  t.home_li = (uint16_t*)p;

  // staged columns: positions (and radius and fate for soft spheres)
  BrickCols cols;
#pragma unroll
  for (int c = 0; c < 3; ++c) { cols.S[c] = a.S[c]; cols.pos[c] = pos[c]; }
  cols.xd = a.rad;
  cols.xd_lds = rad;
  cols.xi = a.fate;
  cols.xi_lds = fate;
  double rmax = 0.0;
  const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int staged = brick_stage<D, PER, U4, SOFT, SOFT ? BX_I64 : BX_NONE>(src, b, t, cols, rmax);
  if (staged == 2 && wl && tid == 0) *a.lgate = 0;   // its rows get no list
  if (staged != 1) return;
  const int sim = t.sim;
  const SimDev& sd = a.g.sims[sim];
  const int64_t N = a.g.st[sim].N;
  const int32_t* hs = t.hs;
  const int32_t* hrow = t.hrow;
  const uint16_t* owner = t.owner;
  const uint16_t* home_li = t.home_li;

  // 4. the home agents, one per thread
  const int32_t nh = t.r1 - t.r0;
  const Law law(sd.p);
  double L[3], hL[3];
  box_halves<D, PER>(sd, L, hL);
  double reach2 = __builtin_inf();
  double disp2 = -1.0;
  if constexpr (FILTER && !SOFT) {
    reach2 = reach2_bound(law.reach());
    if (wl) {
      // lists to CL_LIST_GROW x reach: valid while displacements stay under
      // half the margin (with a little slack for rounding)
      reach2 = reach2_bound(law.reach() * CL_LIST_GROW);
      const double half = 0.5 * (CL_LIST_GROW - 1.0) * law.reach() * (1.0 - 1e-9);
      disp2 = half > 0.0 ? half * half : 0.0;
    }
  }
  for (int32_t k = tid; k < nh; k += nt) {
    const int32_t li = home_li[k];
    const int h = owner[li] & 0x7FFF;
    const int64_t i = (int64_t)hrow[h] + (li - hs[h]);
    if (a.brick_sub) a.brick_skip[i] = 0;   // done here, not by the row kernel
    if (i - sd.row_off >= N) continue;
    double xi[3] = {0.0, 0.0, 0.0};
    double acc[3] = {0.0, 0.0, 0.0};
    double upre[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < D; ++c) upre[c] = a.U[c][i];   // consumed after the walk
    {
#pragma unroll
      for (int c = 0; c < D; ++c) xi[c] = pos[c][li];
      const double rad_i = SOFT ? rad[li] : 0.0;
      const int64_t fate_i = SOFT ? fate[li] : 0;
      double keep2 = reach2;
      if constexpr (SOFT) keep2 = law.mu >= 0.0 ? reach2_bound(law.mu * (rad_i + rmax)) : __builtin_inf();
      auto walk = [&](auto&& visit) { brick_walk<D>(t, h, visit); };
      auto body = [&](int32_t q, bool live) {
        double xj[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < D; ++c) xj[c] = pos[c][q];
        pair_into<D, Law, PER, SOFT>(law, xi, xj, L, hL, rad_i, SOFT ? rad[q] : 0.0, fate_i,
                                     SOFT ? (int64_t)fate[q] : 0, live, acc);
      };
      // a sub-brick walks every candidate once: its dense cells would
      // overflow the per-lane queue of kept candidates and walk twice
      if (FILTER && !a.brick_sub) {
        auto near = [&](int32_t q) {
          double dx[3];
#pragma unroll
          for (int c = 0; c < D; ++c) {
            dx[c] = xi[c] - pos[c][q];
            if (PER) dx[c] = wrap_dx(dx[c], L[c], hL[c]);
          }
          double s2 = dx[0] * dx[0];
          if (D > 1) s2 = s2 + dx[1] * dx[1];
          if (D > 2) s2 = s2 + dx[2] * dx[2];
          return s2 <= keep2 && q != li;
        };
        auto row_of = [&](int32_t q) {
          const int hq = owner[q] & 0x7FFF;
          return (int32_t)(hrow[hq] + (q - hs[hq]));
        };
        // pass 1: LDS indices of the candidates within reach, in walk order
        int nq = 0;
        walk([&](int32_t q) {
          if (near(q)) {
            if (nq < BRICK_LIST) lst[nq * nt + tid] = (uint16_t)q;
            ++nq;
          }
        });
        // pass 2: the pair body on those (and their rows as the stage-2
        // list); a lane whose queue overflowed walks every candidate again
        if (nq <= BRICK_LIST) {
          for (int u = 0; u < nq; ++u) {
            const int32_t q = lst[u * nt + tid];
            body(q, true);
            if (wl) a.wlist[(int64_t)u * a.ldl + i] = row_of(q);
          }
          if (wl) a.wlist_n[i] = nq;
        } else {
          int32_t cnt = 0;
          walk([&](int32_t q) {
            body(q, q != li);
            if (wl && near(q)) {
              if (cnt < a.wcap) a.wlist[(int64_t)cnt * a.ldl + i] = row_of(q);
              ++cnt;
            }
          });
          if (wl) {
            a.wlist_n[i] = cnt < a.wcap ? cnt : a.wcap;
            if (cnt > a.wcap) *a.lgate = 0;
          }
        }
      } else {
        walk([&](int32_t q) { body(q, q != li); });
      }
    }
    stage_finish<D, Law, NOISE, true, Noise>(a, sd, sim, i, xi, acc, upre, disp2);
  }
}

// The brick kernel of a hipRTC module: nothing for a general agentODE body
// (its row kernel walks every row).
template <int D, class Law, int PER, int NOISE, class Noise>
__device__ __forceinline__ void user_brick(const StageArgs& a, unsigned char* lds) {
  if constexpr (!general_body<Law>::value) brick_body<D, Law, PER, NOISE, Noise>(a, lds);
}

// An element at a 32-bit byte offset from a base that all lanes share (a
// kernel argument): the access takes the base from scalar registers and the
// offset from one vector register, with no 64-bit address arithmetic.
template <class T>
__device__ __forceinline__ T ld_b(const T* base, uint32_t off) {
  return *(const T*)((const char*)base + off);
}
template <class T>
__device__ __forceinline__ void st_b(T* base, uint32_t off, T v) {
  *(T*)((char*)base + off) = v;
}

// Batcher's odd-even merge sort of N keys held in registers (every index is
// a compile-time constant once the loops unroll).
template <int N, class T>
__device__ __forceinline__ void sort_net(T* k) {
#pragma unroll
  for (int p = 1; p < N; p <<= 1)
#pragma unroll
    for (int kk = p; kk >= 1; kk >>= 1)
#pragma unroll
      for (int j = kk % p; j + kk < N; j += 2 * kk)
#pragma unroll
        for (int i = 0; i < kk && i + j + kk < N; ++i)
          if ((i + j) / (2 * p) == (i + j + kk) / (2 * p)) {
            const T x = k[i + j], y = k[i + j + kk];
            k[i + j] = x < y ? x : y;
            k[i + j + kk] = x < y ? y : x;
          }
}

// ------------------------------------------ CellLinked on a skin list
// CellLinked (src/neighbors.jl:587-747) walks, for agent i, the 3^D cells
// around i's cell in cellPosNeigh order (x fastest) and each cell's agents by
// slot.  Its active pairs (d < the law's reach) therefore come in the order
// (k, slot j), k the position of j's cell among i's neighbor cells.  With a
// skin list (radius reach + skin, rebuilt when an agent has moved skin/2)
// holding every pair within reach, the step needs no cell sort: each agent's
// list entries are ordered by (k, slot) from the agents' packed cells of the
// step start.

// Every candidate j of row i in the reference's CellLinked order whose
// stage position lies within sqrt(bound2) of xi (fn(j, s2)), found without
// the skin list: repeated minimum selection of the key (k, slot) over the
// list-build grid cells within R of xi, where R bounds the distance of every
// such j's position at the last list build.  Exact for any displacement;
// used for the few agents whose list may miss a pair, and for exports.
template <int D, int PER, class Fn>
__device__ void clv_each(const StageArgs& a, const SimDev& sd, int64_t i, const double* xi, double bound2, double R,
                         Fn&& fn) {
  double L[3], hL[3];
  box_halves<D, PER>(sd, L, hL);
  const int32_t ci = a.ccell[i];
  int32_t c0[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
  for (int e = 0; e < D; ++e) {
    const double x = xi[e];
    int64_t c = x < sd.lo[e] ? 0 : floor_div(x - sd.lo[e], sd.gedge[e]) + (PER ? 0 : 1);
    c = c < 0 ? 0 : c;
    c0[e] = (int32_t)(c > sd.gn[e] - 1 ? sd.gn[e] - 1 : c);
    const double q = R / sd.gedge[e];
    const int32_t m = (q == q && q < (double)sd.gn[e]) ? (int32_t)ceil(q) : sd.gn[e];
    if (PER && 2 * m + 1 >= sd.gn[e]) { lo[e] = 0; hi[e] = sd.gn[e] - 1; c0[e] = 0; }
    else { lo[e] = -m; hi[e] = m; }
  }
  const CellRange* __restrict__ cr = a.cell_range + sd.rank_off;
  unsigned long long from = 0;
  for (;;) {
    unsigned long long best = ~0ull;
    int64_t bj = -1;
    double bs2 = 0.0;
    for (int32_t oz = D > 2 ? lo[2] : 0; oz <= (D > 2 ? hi[2] : 0); ++oz)
      for (int32_t oy = D > 1 ? lo[1] : 0; oy <= (D > 1 ? hi[1] : 0); ++oy)
        for (int32_t ox = lo[0]; ox <= hi[0]; ++ox) {
          const int32_t o[3] = {ox, oy, oz};
          int32_t q[3] = {0, 0, 0};
          bool ok = true;
#pragma unroll
          for (int e = 0; e < D; ++e) {
            int32_t v = c0[e] + o[e];
            if (PER) v = ((v % sd.gn[e]) + sd.gn[e]) % sd.gn[e];
            else ok = ok && v >= 0 && v < sd.gn[e];
            q[e] = v;
          }
          if (!ok) continue;
          const CellRange r = cr[q[0] + sd.gn[0] * (q[1] + sd.gn[1] * q[2])];
          for (int32_t j = r.s; j < r.e; ++j) {
            if (j == i) continue;
            const int k = cl_key<D, PER>(sd, ci, a.ccell[j]);
            if (k == 27) continue;
            double dx[3];
#pragma unroll
            for (int c = 0; c < D; ++c) {
              dx[c] = xi[c] - a.S[c][j];
              if (PER) dx[c] = wrap_dx(dx[c], L[c], hL[c]);
            }
            double s2 = dx[0] * dx[0];
            if (D > 1) s2 = s2 + dx[1] * dx[1];
            if (D > 2) s2 = s2 + dx[2] * dx[2];
            if (!(s2 <= bound2)) continue;
            const unsigned long long key = ((unsigned long long)k << 32) | (uint32_t)a.cslot[j];
            if (key >= from && key < best) { best = key; bj = j; bs2 = s2; }
          }
        }
    if (bj < 0) return;
    fn(bj, bs2);
    from = best + 1;
  }
}

// The pair sum of row i by clv_each within the law's reach.
template <int D, class Law, int PER>
__device__ void clv_walk(const StageArgs& a, const SimDev& sd, int64_t i, const double* xi, double E, double* acc) {
  const Law law(sd.p);
  double L[3], hL[3];
  box_halves<D, PER>(sd, L, hL);
  const double reach = law.reach();
  double s[3] = {0.0, 0.0, 0.0};
  clv_each<D, PER>(a, sd, i, xi, reach2_bound(reach), reach + E, [&](int64_t j, double) {
    double xj[3] = {0.0, 0.0, 0.0}, t[3];
#pragma unroll
    for (int c = 0; c < D; ++c) xj[c] = a.S[c][j];
    pair_term<D, Law, PER>(law, xi, xj, L, hL, t);
#pragma unroll
    for (int c = 0; c < D; ++c) s[c] = s[c] + t[c];
  });
#pragma unroll
  for (int c = 0; c < D; ++c) acc[c] = s[c];
}

// CellLinked on a skin list, one agent per lane.  Per step:
//   clv_order_row (inside the step's first stage): each agent's skin-list
//     entries inside its 3^D neighbor cells are put in CellLinked order --
//     cell k among the neighbor cells, then list position (= slot order) --
//     and written as the step's list (a.wlist): lists of at most 16 entries
//     by a sorting network over keys (k, position, row) in registers, longer
//     ones by insertion of the keys (k << 6 | position) into the lane's LDS
//     queue (stride blockDim, a.lcap entries); a list at capacity (possibly
//     truncated) is marked -1.
//   clv_stage_body: every stage walks that list.  An agent whose list may
//     miss a pair -- marked -1, or whose distance e from its reference
//     position plus the largest, E, over all agents exceeds the skin --
//     writes nothing and is queued for
//   clv_fix_body, which sums its pairs by clv_walk and finishes its stage.
// Stage 1 of a two-stage integrator records max e^2 for stage 2, the final
// stage for the next step's stage 1 (only agents past half the skin report;
// E is at least half the skin).

// One lane's queued agent (a.fix_rows), counted per queue in a.fix_n[fix_q].
__device__ __forceinline__ void clv_defer(const StageArgs& a, int64_t i) {
  const int32_t q = atomicAdd(a.fix_n + a.fix_q, 1);
  a.fix_rows[q] = (int32_t)i;
}

// e^2 of a stage-1 output past half the skin, into the device maximum.
__device__ __forceinline__ void clv_report(const StageArgs& a, double e2) {
  if (a.emax && !(e2 <= a.clv_half2)) atomicMax(a.emax, (unsigned long long)__double_as_longlong(e2));
}

// The ordered list of an agent with n <= W skin-list entries (see
// clv_order_row); returns its length.  The list holds lcap >= 16 entries
// per row, so all W loads are in bounds; entries past the length are
// replaced by the agent itself and sorted last.
template <int W, int D, int PER>
__device__ __forceinline__ int32_t clv_order_small(const StageArgs& a, const SimDev& sd, int64_t i, int32_t n,
                                                   int32_t ci) {
  uint32_t key[W];
  const uint32_t i4 = (uint32_t)i * 4u, l4 = (uint32_t)a.ldl * 4u;
#pragma unroll
  for (int u = 0; u < W; ++u) {
    const int32_t raw = ld_b(a.list, i4 + (uint32_t)u * l4);
    key[u] = (uint32_t)(u < n ? raw : (int32_t)i);
  }
  // all W gathers are issued before the first key: the empty asm uses each
  // loaded value unconditionally, so the compiler cannot move a load into
  // the (u < n) branch behind its own wait
  int32_t cj[W];
#pragma unroll
  for (int u = 0; u < W; ++u) cj[u] = ld_b(a.ccell, key[u] * 4u);
#pragma unroll
  for (int u = 0; u < W; ++u) __asm__ volatile("" ::"v"(cj[u]));
  int32_t kept = 0;
#pragma unroll
  for (int u = 0; u < W; ++u) {
    const int k = u < n ? cl_key<D, PER>(sd, ci, cj[u]) : 27;   // 27: outside the 3^D cells, sorts last
    key[u] |= ((uint32_t)k << 27) | ((uint32_t)u << 23);
    kept += k < 27 ? 1 : 0;
  }
  sort_net<W>(key);
#pragma unroll
  for (int v = 0; v < W; ++v)
    if (v < kept) st_b(a.wlist, i4 + (uint32_t)v * l4, (int32_t)(key[v] & 0x7FFFFFu));
  return kept;
}

// Row i's skin-list entries (a.list, a.list_n) in CellLinked order, written
// to a.wlist; returns their count, or -1 for a list at capacity (possibly
// truncated).  LDS: blockDim * a.lcap 16-bit keys (lists over 16 entries).
template <int D, int PER>
__device__ __forceinline__ int32_t clv_order_row(const StageArgs& a, const SimDev& sd, int64_t i,
                                                 unsigned char* lds) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int32_t* __restrict__ lp = a.list + i;       // entry q at lp[q * ldl]
  const int64_t ldl = a.ldl;
  const int32_t n = a.list_n[i];
  if (n >= a.lcap) return -1;
  const int32_t ci = a.ccell[i];
  // lists of at most 16 entries: every list read and cell gather in flight,
  // and the keys (k << 27 | position << 23 | row) sorted by a network in
  // registers; the row rides in the key (the set-up allows this form below
  // 2^23 rows), and 32-bit byte offsets address from the kernel-argument
  // bases.  (A form for 8 entries beside it gains nothing: the lanes of a
  // wave rarely all hold lists that short.)
  if (n <= 16) return clv_order_small<16, D, PER>(a, sd, i, n, ci);
  uint16_t* qk = (uint16_t*)lds + tid;                // this lane's queue, entry u at qk[u * nt]
  int32_t kept = 0;
  auto insert = [&](uint16_t key) {
    int32_t pos = kept;
    while (pos > 0) {
      const uint16_t prev = qk[(pos - 1) * nt];
      if (prev < key) break;
      qk[pos * nt] = prev;
      --pos;
    }
    qk[pos * nt] = key;
    ++kept;
  };
  for (int32_t q0 = 0; q0 < n; q0 += 4) {
    int32_t jj[4], cc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jj[u] = q0 + u < n ? lp[(int64_t)(q0 + u) * ldl] : (int32_t)i;
#pragma unroll
    for (int u = 0; u < 4; ++u) cc[u] = a.ccell[jj[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (q0 + u >= n) continue;
      const int k = cl_key<D, PER>(sd, ci, cc[u]);
      if (k < 27) insert((uint16_t)((k << 6) | (q0 + u)));
    }
  }
  for (int32_t v0 = 0; v0 < kept; v0 += 4) {
    int32_t jj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jj[u] = v0 + u < kept ? lp[(int64_t)(qk[(v0 + u) * nt] & 63) * ldl] : 0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (v0 + u < kept) a.wlist[(int64_t)(v0 + u) * ldl + i] = jj[u];
  }
  return kept;
}

template <int D, class Law, int PER, int NOISE>
__device__ __forceinline__ void clv_stage_body(const StageArgs& a, unsigned char* lds) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the step's bookkeeping, where no other kernel is using the words: the
    // other queue (its fix kernel is done, the next user is a later stage),
    // and in the first stage the next step's stage-1 maximum
    a.fix_n[1 - a.fix_q] = 0;
    if (a.clv_order) {
      *a.emax_next = 0ull;
      if (a.book) {
        // the record the host reads (system-scope stores into mapped host
        // memory, the step last)
        __hip_atomic_store(&a.rec->soon, a.efin[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&a.rec->e2max, a.efin[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&a.rec->step, a.step_no, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  if (ri.local >= a.g.st[ri.sim].N) return;
  const int64_t i = ri.row;
  double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) xi[c] = a.S[c][i];
  int32_t n;
  if (a.clv_order) {
    // the step's first stage: is the agent within reach of every pair its
    // list must hold?  Then its entries in CellLinked order, written for
    // the later stages and walked below
    if (a.efin) {
      // every agent is within E0 of its reference position, E0 the largest
      // reported displacement (agents past half the skin report) or a
      // little over half the skin; with no report all pairs are in the lists
      const unsigned long long e0 = *a.efin;
      if (e0 != 0ull) {
        const double E0 = fmax(sqrt(__longlong_as_double((long long)e0)), a.clv_e1);
        double s2 = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double d = xi[c] - a.xref[c][i];
          if (PER) d = min_image(d, sd.L[c]);
          s2 = s2 + d * d;
        }
        if (!(sqrt(s2) + E0 <= a.clv_skin)) { a.wlist_n[i] = -1; clv_defer(a, i); return; }
      }
    }
    n = clv_order_row<D, PER>(a, sd, i, lds);
    a.wlist_n[i] = n;
  } else {
    n = a.wlist_n[i];
  }
  if (n < 0) { clv_defer(a, i); return; }
  if (a.fix_stage == 1 && *a.emax != 0ull) {
    // stage 2: could a pair active here lie outside the list?  (With no
    // stage-1 report every agent is within half the skin: none can.)
    double dx[3];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dx[c] = xi[c] - a.xref[c][i];
      if (PER) dx[c] = min_image(dx[c], sd.L[c]);
    }
    double s2 = dx[0] * dx[0];
    if (D > 1) s2 = s2 + dx[1] * dx[1];
    if (D > 2) s2 = s2 + dx[2] * dx[2];
    const double E = fmax(sqrt(__longlong_as_double((long long)*a.emax)), a.clv_half);
    if (!(sqrt(s2) + E <= a.clv_skin)) { clv_defer(a, i); return; }
  }
  double upre[3] = {0.0, 0.0, 0.0}, acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) upre[c] = a.U[c][i];
  const Law law(sd.p);
  double L[3], hL[3];
  box_halves<D, PER>(sd, L, hL);
  const int32_t* __restrict__ lp = a.wlist + i;       // entry q at lp[q * ldl]
  const int64_t ldl = a.ldl;
  int32_t jn[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) jn[u] = u < n ? lp[(int64_t)u * ldl] : (int32_t)i;
  for (int32_t q0 = 0; q0 < n; q0 += 2) {
    int32_t jc[2];
    double xj[2][3];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      jc[u] = jn[u];
#pragma unroll
      for (int c = 0; c < D; ++c) xj[u][c] = a.S[c][jc[u]];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) jn[u] = q0 + 2 + u < n ? lp[(int64_t)(q0 + 2 + u) * ldl] : (int32_t)i;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      double t[3];
      pair_term<D, Law, PER>(law, xi, xj[u], L, hL, t);
      const bool v = q0 + u < n;
#pragma unroll
      for (int c = 0; c < D; ++c) acc[c] = acc[c] + (v ? t[c] : 0.0);
    }
  }
  double e2 = 0.0;
  stage_finish<D, Law, NOISE, true>(a, sd, ri.sim, i, xi, acc, upre, -1.0, &e2);
  clv_report(a, e2);
}

template <int D, class Law, int PER, int NOISE>
__device__ void clv_fix_body(const StageArgs& a) {
  const int32_t cnt = a.fix_n[a.fix_q];
  for (int32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
    const int64_t i = a.fix_rows[q];
    int sim = 0;
    while (sim + 1 < a.nsims && a.g.sims[sim + 1].row_off <= i) ++sim;
    const SimDev& sd = a.g.sims[sim];
    double xi[3] = {0.0, 0.0, 0.0}, upre[3] = {0.0, 0.0, 0.0}, acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < D; ++c) { xi[c] = a.S[c][i]; upre[c] = a.U[c][i]; }
    const double E = a.fix_stage == 1 ? fmax(sqrt(__longlong_as_double((long long)*a.emax)), a.clv_half)
                                      : fmax(a.efin ? sqrt(__longlong_as_double((long long)*a.efin)) : 0.0, a.clv_e1);
    clv_walk<D, Law, PER>(a, sd, i, xi, E, acc);
    double e2 = 0.0;
    stage_finish<D, Law, NOISE, true>(a, sd, sim, i, xi, acc, upre, -1.0, &e2);
    clv_report(a, e2);
  }
}

// One integrator stage over the Verlet lists with the pair work spread
// evenly over the block's lanes.  The entries of the block's agents are
// numbered agent-major (a block-wide scan of the list lengths) and processed
// in rounds of whole agents holding about `fchunk` entries: every lane
// evaluates entries of the round -- gather, square root, division and law
// of one pair per lane step -- and stores the term in LDS; then each
// agent's thread adds its terms in list order (the order of the per-agent
// walk, so the sums are bit-identical) and finishes the stage.  A wavefront
// thus no longer runs to its longest list.  Lists are at most FLAT_MAX_CAP
// long.
template <int D, class Law, int PER, int NOISE>
__device__ __forceinline__ void flat_body(const StageArgs& a, unsigned char* lds) {
  if (a.gate && *a.gate != a.gate_run) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int chunk = a.fchunk;
  const int ch = chunk + a.fcap;
  double* term[3];
  unsigned char* p = lds;
#pragma unroll
  for (int c = 0; c < D; ++c) { term[c] = (double*)p; p += (size_t)8 * ch; }
  double* hx[3];
#pragma unroll
  for (int c = 0; c < D; ++c) { hx[c] = (double*)p; p += (size_t)8 * BRICK_MAX_BLOCK; }
  int32_t* offs = (int32_t*)p;
  p += (size_t)4 * (BRICK_MAX_BLOCK + 1);
  int32_t* rfirst = (int32_t*)p;
  p += (size_t)4 * (FLAT_MAX_ROUNDS + 1);
  int32_t* wsum = (int32_t*)p;
  p += 4 * 8;
  uint8_t* owner = (uint8_t*)p;

  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  const int64_t i = ri.row;
  const int64_t row0 = i - tid;                  // the block's rows are consecutive
  const bool live = ri.local < N;
  double xi[3] = {0.0, 0.0, 0.0}, upre[3] = {0.0, 0.0, 0.0};
  int32_t n = 0;
  if (live) {
#pragma unroll
    for (int c = 0; c < D; ++c) xi[c] = a.S[c][i];
    n = a.list_n[i];
#pragma unroll
    for (int c = 0; c < D; ++c) upre[c] = a.U[c][i];   // consumed by the stage update
  }
#pragma unroll
  for (int c = 0; c < D; ++c) hx[c][tid] = xi[c];
  for (int r = tid; r <= FLAT_MAX_ROUNDS; r += nt) rfirst[r] = 0x7FFFFFFF;
  const int32_t off = block_excl_scan(n, wsum, wsum + 7);
  offs[tid] = off;
  // round of an agent: the chunk its first entry falls in; a round's entries
  // run from its first agent's first entry to the next round's
  const int myr = n > 0 ? off / chunk : -1;
  __syncthreads();
  if (n > 0) atomicMin(&rfirst[myr], off);
  const int32_t P = wsum[7];
  __syncthreads();

  const Law law(sd.p);
  double L[3], hL[3];
  box_halves<D, PER>(sd, L, hL);
  const int32_t* __restrict__ list = a.list;
  const int64_t ldl = a.ldl;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int r = 0; r < FLAT_MAX_ROUNDS; ++r) {
    const int32_t lo = rfirst[r];
    if (lo >= P) break;
    const int32_t nx = rfirst[r + 1];
    const int32_t hi = nx < P ? nx : P;
    if (myr == r)
      for (int32_t k = 0; k < n; ++k) owner[off - lo + k] = (uint8_t)tid;
    __syncthreads();
    // every lane evaluates entries of the round, two in flight
    for (int32_t p0 = lo + tid; p0 < hi; p0 += 2 * nt) {
      int32_t pp[2], jj[2], tt[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        pp[u] = p0 + u * nt;
        const bool v = pp[u] < hi;
        tt[u] = v ? owner[pp[u] - lo] : tid;
        const int32_t q = pp[u] - offs[tt[u]];
        jj[u] = v ? list[(int64_t)q * ldl + row0 + tt[u]] : (int32_t)i;
      }
      double xj[2][3];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int c = 0; c < D; ++c) xj[u][c] = a.S[c][jj[u]];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if (pp[u] >= hi) continue;
        double xo[3] = {0.0, 0.0, 0.0}, t[3];
#pragma unroll
        for (int c = 0; c < D; ++c) xo[c] = hx[c][tt[u]];
        pair_term<D, Law, PER>(law, xo, xj[u], L, hL, t);
#pragma unroll
        for (int c = 0; c < D; ++c) term[c][pp[u] - lo] = t[c];
      }
    }
    __syncthreads();
    if (myr == r)
      for (int32_t k = off - lo; k < off - lo + n; ++k)
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] = acc[c] + term[c][k];
    __syncthreads();
  }
  if (!live) return;
  stage_finish<D, Law, NOISE, true>(a, sd, ri.sim, i, xi, acc, upre);
}

}  // namespace cbm

#endif  // CBM_STAGE_H

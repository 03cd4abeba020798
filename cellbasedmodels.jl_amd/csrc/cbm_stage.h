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

__host__ __device__ constexpr int pow3(int d) { return d == 1 ? 3 : (d == 2 ? 9 : 27); }

// Walk the candidate neighbors of row i in reference loop order.
template <int D, int NB, class Fn>
__device__ __forceinline__ void for_candidates(const StageArgs& a, const SimDev& sd, int64_t N, int64_t i, Fn&& fn) {
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

// ----------------------------------------------------------- stage kernel
// Minimum image (periodic boxes): the same comparisons as the oracle, as selects.
__device__ __forceinline__ double wrap_dx(double d, double L, double hL) {
  return d > hL ? d - L : (d < -hL ? d + L : d);
}

// One integrator stage for every agent row: the agentODE body over the
// neighbor candidates (Law = the pair law functor) fused with the stage
// update.  Built-in laws instantiate it in libcbm_hip.so; user laws are
// compiled around it with hipRTC.
template <int D, int NB, class Law, int PER, int NOISE>
__device__ __forceinline__ void stage_body(const StageArgs& a) {
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  if (ri.local >= N) return;
  const int64_t i = ri.row;

  double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) xi[c] = a.S[c][i];
  double acc[3] = {0.0, 0.0, 0.0};
  // the inertial soft-sphere model: per-agent radius and fate in the pair
  // term, velocities as DE variables
  constexpr bool SOFT = same_type<Law, PairLaw<LAW_SOFT_SPHERES>>::value;

  if (Law::active) {
    const Law law(sd.p);
    const double rad_i = SOFT ? a.rad[i] : 0.0;
    const int64_t fate_i = SOFT ? a.fate[i] : 0;
    double L[3] = {0.0, 0.0, 0.0}, hL[3] = {0.0, 0.0, 0.0};
    if (PER) {
#pragma unroll
      for (int c = 0; c < D; ++c) { L[c] = sd.L[c]; hL[c] = 0.5 * sd.L[c]; }
    }
    // one neighbor at position xj (the agentODE interaction body)
    auto body = [&](const double* xj, int64_t j, bool live) {
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
      if constexpr (SOFT) on = law.pair(d, rad_i, a.rad[j], fate_i, a.fate[j], s) && live;
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
    if (Law::walls && a.walls) {
      const double rRep = sd.p[0], fRep = sd.p[1];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const double x = xi[c];
        if (x < sd.lo[c] + rRep / 2) acc[c] = acc[c] + fRep;
        else if (x > sd.hi[c] - rRep / 2) acc[c] = acc[c] - fRep;
      }
    }
  }

  // derivative of each DE variable
  double du[MAX_STATE];
  if constexpr (SOFT) {
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
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double v = a.lagged ? a.F[c][i] : acc[c];
      if (a.drift != 0.0) v = v - a.drift * xi[c];
      du[c] = v;
      if (a.write_fnew) a.Fnew[c][i] = acc[c];
    }
  }
  const bool grow = !SOFT && a.nvar > D;
  if (grow) du[D] = a.growth[i] * a.S[D][i];

  const double dt = a.dt;
  const int nv = SOFT ? 2 * D : (grow ? D + 1 : D);
  double res[MAX_STATE];
  if (NOISE) {
    // EM and the EulerHeun predictor: (u + f dt) + (g xi) sqrt(dt), xi keyed
    // 0x100/0x101; the EulerHeun corrector draws fresh noise (0x500/0x501)
    double z[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t id = a.id[i];
    const int64_t step = *a.step_ptr;
    const uint32_t pur = a.kind == ST_EH2 ? 0x500u : 0x100u;
    normal2(draw(a.seed, step, id, ri.sim, pur), z[0], z[1]);
    if (D > 2) normal2(draw(a.seed, step, id, ri.sim, pur + 1u), z[2], z[3]);
    if (a.kind == ST_EH2) {
#pragma unroll
      for (int c = 0; c < D; ++c)
        res[c] = a.U[c][i] + ((a.kacc[c][i] + du[c]) * dt + (a.noise + a.noise) * z[c] * a.sqdt) / 2;
      for (int v = D; v < nv; ++v) res[v] = a.U[v][i] + ((a.kacc[v][i] + du[v]) * dt) / 2;
    } else {
#pragma unroll
      for (int c = 0; c < D; ++c) res[c] = (a.U[c][i] + du[c] * dt) + a.noise * z[c] * a.sqdt;
      for (int v = D; v < nv; ++v) res[v] = a.U[v][i] + du[v] * dt;   // no noise off the positions
      if (a.kind == ST_EH1)
        for (int c = 0; c < nv; ++c) a.kacc[c][i] = du[c];
    }
  } else {
    switch (a.kind) {
      case ST_EULER:
        for (int c = 0; c < nv; ++c) res[c] = a.U[c][i] + du[c] * dt;
        break;
      case ST_HEUN1:
        for (int c = 0; c < nv; ++c) res[c] = a.U[c][i] + du[c] * dt;
        break;
      case ST_HEUN2:
        for (int c = 0; c < nv; ++c) res[c] = (a.S[c][i] + (a.U[c][i] + du[c] * dt)) / 2;
        break;
      case ST_RK4_1:
        for (int c = 0; c < nv; ++c) { a.kacc[c][i] = du[c]; res[c] = a.U[c][i] + du[c] * dt / 2; }
        break;
      case ST_RK4_2:
        for (int c = 0; c < nv; ++c) { a.kacc[c][i] = a.kacc[c][i] + 2 * du[c]; res[c] = a.U[c][i] + du[c] * dt / 2; }
        break;
      case ST_RK4_3:
        for (int c = 0; c < nv; ++c) { a.kacc[c][i] = a.kacc[c][i] + 2 * du[c]; res[c] = a.U[c][i] + du[c] * dt; }
        break;
      default:  // ST_RK4_4
        for (int c = 0; c < nv; ++c) res[c] = a.U[c][i] + (a.kacc[c][i] + du[c]) * dt / 6;
        break;
    }
  }

  const bool final_stage = a.kind == ST_EULER || a.kind == ST_EM || a.kind == ST_HEUN2 ||
                           a.kind == ST_RK4_4 || a.kind == ST_EH2;
  if (final_stage && a.periodic) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double x = res[c];
      if (x >= sd.lo[c] + sd.L[c]) x = x - sd.L[c];
      else if (x < sd.lo[c]) x = x + sd.L[c];
      res[c] = x;
    }
  }
  for (int c = 0; c < nv; ++c) a.out[c][i] = res[c];

  if (final_stage && a.disp_check) {
    double dx[3];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dx[c] = res[c] - a.xref[c][i];
      if (a.periodic) dx[c] = min_image(dx[c], sd.L[c]);
    }
    double s2 = dx[0] * dx[0];
    if (D > 1) s2 = s2 + dx[1] * dx[1];
    if (D > 2) s2 = s2 + dx[2] * dx[2];
    if (s2 > a.thr2) *a.flag = 1;   // sqrt(s2) >= thr, src/neighbors.jl:403 (>= skin/2)
  }
}

}  // namespace cbm

#endif  // CBM_STAGE_H

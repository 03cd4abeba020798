// cbm_medium.h -- the FTCS kernel skeleton of the medium (mediumODE,
// src/AgentStructure/functionDE.jl:13-116 with the medium loop of
// auxiliar.jl:168-172,184-210), around the built-in body (diffusion with
// agent uptake) or a user body (struct UserMedium, compiled with hipRTC).
// Each node gets c' = c + dt * f(node), f from the body; the second
// derivatives use ghost nodes mirroring the grid's edge nodes (zero flux).
// A node on a side with a Dirichlet or copy border, or one the body's
// optional `bool assign(const MediumNode&, double& v)` claims, is assigned
// v instead (the `@mediumBorder` branches of a mediumODE body,
// src/AgentStructure/macros.jl:284-311, with `@mediumInside()` nodes
// integrating: docs/src/Usage_Agent.md:397-403).  Every read is of the
// step-start field.
#ifndef CBM_MEDIUM_H
#define CBM_MEDIUM_H

#include "cbm_stage.h"

namespace cbm {

// What a medium body sees at one node.
struct MediumNode {
  double c;               // concentration
  double lap[3];          // @∂2 per axis: ((c[+1] - 2c) + c[-1]) / dx^2 (0 past dims)
  double x[3];            // node centre
  int64_t idx[3];         // node index per axis, 0-based
  int32_t agents;         // agents whose position falls in the node (cellInMesh)
  int32_t dims;
  double dt;
  double nb[3][2];        // the neighbors' values along each axis, [a][0] below, [a][1] above
                          // (the node itself past the grid's edge; 0 past dims)
  int32_t border[3];      // -1: a lower border node of that axis (index 0), +1: an upper
                          // one, 0: neither (@mediumBorder(axis, side))
  // @mediumInside(): on no border of any axis
  __device__ bool inside() const {
    bool in = border[0] == 0;
    if (dims > 1) in = in && border[1] == 0;
    if (dims > 2) in = in && border[2] == 0;
    return in;
  }
  // the Laplacian, summed over the axes in order
  __device__ double laplacian() const {
    double l = lap[0];
    if (dims > 1) l = l + lap[1];
    if (dims > 2) l = l + lap[2];
    return l;
  }
};

// The built-in body: dt(c) = D lap(c) - (uptake c) n_agents.
struct BuiltinMedium {
  double D, uptake;
  __device__ BuiltinMedium(double D_, double uptake_, const double*) : D(D_), uptake(uptake_) {}
  __device__ double operator()(const MediumNode& m) const { return D * m.laplacian() - (uptake * m.c) * (double)m.agents; }
};

template <class B, class = void> struct has_assign { static constexpr bool value = false; };
template <class B> struct has_assign<B, decltype((void)&B::assign)> { static constexpr bool value = true; };

template <class Body>
__device__ __forceinline__ void medium_step_body(const MediumArgs& m) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m.nodes) return;
  const int64_t nx = m.n[0], ny = m.n[1];
  const int64_t p[3] = {k % nx, (k / nx) % ny, k / (nx * ny)};
  const int64_t stride[3] = {1, nx, nx * ny};
  MediumNode nd;
  nd.c = m.c[k];
  nd.dims = m.dims;
  nd.dt = m.dt;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    nd.idx[a] = p[a];
    nd.lap[a] = 0.0;
    nd.x[a] = 0.0;
    nd.nb[a][0] = 0.0;
    nd.nb[a][1] = 0.0;
    nd.border[a] = 0;
    if (a < m.dims) {
      const double cp = (p[a] + 1 < m.n[a]) ? m.c[k + stride[a]] : nd.c;
      const double cm = (p[a] > 0) ? m.c[k - stride[a]] : nd.c;
      nd.nb[a][0] = cm;
      nd.nb[a][1] = cp;
      if (m.lap) nd.lap[a] = ((cp - 2 * nd.c) + cm) / (m.dx[a] * m.dx[a]);
      nd.x[a] = m.lo[a] + ((double)p[a] + 0.5) * m.dx[a];
      nd.border[a] = p[a] == 0 ? -1 : (p[a] == m.n[a] - 1 ? 1 : 0);
    }
  }
  nd.agents = m.count[k];
  m.count[k] = 0;
  // the first side (x-, x+, y-, ...) with a Dirichlet or copy border that
  // the node lies on assigns it; a copy takes the inward neighbor.  Every
  // index is a constant once the loops unroll (no kernel-argument or node
  // array is indexed at run time)
  bool assigned = false;
  double av = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const bool on = a < m.dims && (side == 0 ? p[a] == 0 : p[a] == m.n[a] - 1);
      const int32_t mode = m.bmode[2 * a + side];
      if (!assigned && on && mode != 0) {
        av = mode == 1 ? m.bval[2 * a + side] : nd.nb[a][1 - side];
        assigned = true;
      }
    }
  if (assigned) {
    m.cn[k] = av;
    return;
  }
  const Body body(m.D, m.uptake, m.g.sims[0].p);
  if constexpr (has_assign<Body>::value) {
    double v = 0.0;
    if (body.assign(nd, v)) { m.cn[k] = v; return; }
  }
  const double f = body(nd);
  m.cn[k] = nd.c + m.dt * f;
}

}  // namespace cbm

#endif  // CBM_MEDIUM_H

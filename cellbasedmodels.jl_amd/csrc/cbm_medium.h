// cbm_medium.h -- the FTCS kernel skeleton of the medium (mediumODE,
// src/AgentStructure/functionDE.jl:13-116 with the medium loop of
// auxiliar.jl:168-172,184-210), around the built-in body (diffusion with
// agent uptake) or a user body (struct UserMedium, compiled with hipRTC).
// Each node gets c' = c + dt * f(node), f from the body; the second
// derivatives use ghost nodes mirroring the boundary node (zero flux).
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
    if (a < m.dims) {
      if (m.lap) {
        const double cp = (p[a] + 1 < m.n[a]) ? m.c[k + stride[a]] : nd.c;
        const double cm = (p[a] > 0) ? m.c[k - stride[a]] : nd.c;
        nd.lap[a] = ((cp - 2 * nd.c) + cm) / (m.dx[a] * m.dx[a]);
      }
      nd.x[a] = m.lo[a] + ((double)p[a] + 0.5) * m.dx[a];
    }
  }
  nd.agents = m.count[k];
  m.count[k] = 0;
  const Body body(m.D, m.uptake, m.g.sims[0].p);
  const double f = body(nd);
  m.cn[k] = nd.c + m.dt * f;
}

}  // namespace cbm

#endif  // CBM_MEDIUM_H

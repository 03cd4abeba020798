// cbm_model.h -- the model scope of a step: modelStepRule! and modelStepDE!
// (src/CommunityStructure/step.jl:21-30, :66-75, run after the medium
// step, src/CommunityStructure/step.jl:96-107), compiled with hipRTC around
// a user body.  The model variables are the simulation's parameters p[8],
// which every agent and medium body reads.
#ifndef CBM_MODEL_H
#define CBM_MODEL_H

#include "cbm_device.h"
#include "cbm_structs.h"

namespace cbm {

// What a UserModel body receives for one simulation: the values at the
// start of the step (p, read) and their new values (pn, initialised to p),
// committed at the step's end as update! commits the write-new copies
// (src/CommunityStructure/update.jl), so the agents and the medium of this
// step saw p and those of the next step see pn.
struct ModelScope {
  double t, dt;                 // the step's start time and the time step
  int64_t N, step;              // agents of the simulation, step number
  int32_t sim, pad;
  double p[8];
  double pn[8];
};

// A UserModel body, inside namespace cbm:
//   struct UserModel {
//     __device__ void rule(ModelScope& m) const;                  // modelRule: writes m.pn
//     static constexpr unsigned de = MASK;                        // modelODE: p[k] with bit k set
//     __device__ void ode(const ModelScope& m, double* dp) const; // dt(p[k]) at the step start
//     __device__ void noise(const ModelScope& m, double* g) const; // modelSDE: g of p[k], optional
//   };
// each member optional.  The DE parameters step by Euler, or by
// Euler-Maruyama with noise (the reference's default modelAlg,
// src/AgentStructure/agentStructure.jl:276-280): pn[k] = (p[k] + dp[k] dt)
// + g[k] z_k sqrt(dt), z from Philox keyed (seed, step, 0, sim, 0x600 + k/2);
// a DE parameter takes the integrated value over a rule's write.
template <class M, class = void> struct model_has_rule { static constexpr bool value = false; };
template <class M> struct model_has_rule<M, decltype((void)&M::rule)> { static constexpr bool value = true; };
template <class M, class = void> struct model_has_ode { static constexpr bool value = false; };
template <class M> struct model_has_ode<M, decltype((void)&M::ode)> { static constexpr bool value = true; };
template <class M, class = void> struct model_has_noise { static constexpr bool value = false; };
template <class M> struct model_has_noise<M, decltype((void)&M::noise)> { static constexpr bool value = true; };

template <class M>
__device__ __forceinline__ void model_step_body(const ModelArgs& a) {
  const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= a.S) return;
  const int64_t step = *a.step_ptr;
  ModelScope m;
  m.t = a.t_buf[step & 1];
  m.dt = a.dt;
  m.N = a.st[s].N;
  m.step = step;
  m.sim = s;
  m.pad = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    m.p[k] = a.sims[s].p[k];
    m.pn[k] = m.p[k];
  }
  const M body;
  if constexpr (model_has_rule<M>::value) body.rule(m);
  if constexpr (model_has_ode<M>::value) {
    double dp[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    body.ode(m, dp);
    double g[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if constexpr (model_has_noise<M>::value) body.noise(m, g);
    const double sq = sqrt(m.dt);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double z0 = 0.0, z1 = 0.0;
      if (model_has_noise<M>::value && ((M::de >> (2 * q)) & 3u))
        normal2(draw(a.seed, step, 0, s, 0x600u + (uint32_t)q), z0, z1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = 2 * q + h;
        if ((M::de >> k) & 1u) {
          double v = m.p[k] + dp[k] * m.dt;
          if (model_has_noise<M>::value) v = v + g[k] * (h ? z1 : z0) * sq;
          m.pn[k] = v;
        }
      }
    }
  }
  // update!: the new values, and the next step's start time
#pragma unroll
  for (int k = 0; k < 8; ++k) a.sims[s].p[k] = m.pn[k];
  if (s == 0) a.t_buf[(step + 1) & 1] = m.t + m.dt;
}

}  // namespace cbm

#endif  // CBM_MODEL_H

"""Model-scope bodies (UserModel: modelRule, modelODE and modelSDE of
src/AgentStructure/agentStructure.jl:43-45) for the tests, each as HIP C++
(compiled with hipRTC by cbm_compile_model_scope) and as its CPU
restatement for the oracle's model-scope callbacks.

COMBINED: a counter p3 += 1, the step's start time into p4, N into p5 and
p7 += p6 (rule); p0 relaxing to 2 (the linear law's stiffness k: agents
feel it from the next step), p2 growing at rate 1 and p6 an
Ornstein-Uhlenbeck parameter with noise 0.3 (ode + noise, Euler-Maruyama).

KAT: the model part of test/testCommunity.jl:35-200 on p: gi (p2) += 1 and
ga[1,2] (p3) = 7 by the rule, dt(gf) = 1 for gf = p1, which the agents'
dt(x) = gf reads (SELF_GF, a law with a per-agent term)."""

COMBINED = """
struct UserModel {
  static constexpr unsigned de = (1u << 0) | (1u << 2) | (1u << 6);
  __device__ void rule(ModelScope& m) const {
    m.pn[3] = m.p[3] + 1.0;
    m.pn[4] = m.t;
    m.pn[5] = (double)m.N;
    m.pn[7] = m.p[7] + m.p[6];
  }
  __device__ void ode(const ModelScope& m, double* dp) const {
    dp[0] = 0.5 * (2.0 - m.p[0]);
    dp[2] = 1.0;
    dp[6] = -m.p[6];
  }
  __device__ void noise(const ModelScope&, double* g) const { g[6] = 0.3; }
};
"""


def combined():
    def rule(m):
        m.pn[3] = m.p[3] + 1.0
        m.pn[4] = m.t
        m.pn[5] = float(m.N)
        m.pn[7] = m.p[7] + m.p[6]

    def ode(m, dp):
        dp[0] = 0.5 * (2.0 - m.p[0])
        dp[2] = 1.0
        dp[6] = -m.p[6]

    def noise(m, g):
        g[6] = 0.3

    return dict(rule=rule, ode=ode, noise=noise, de=(1 << 0) | (1 << 2) | (1 << 6))


KAT = """
struct UserModel {
  static constexpr unsigned de = 1u << 1;
  __device__ void rule(ModelScope& m) const {
    m.pn[2] = m.p[2] + 1.0;
    m.pn[3] = 7.0;
  }
  __device__ void ode(const ModelScope&, double* dp) const { dp[1] = 1.0; }
};
"""


def kat():
    def rule(m):
        m.pn[2] = m.p[2] + 1.0
        m.pn[3] = 7.0

    def ode(m, dp):
        dp[1] = 1.0

    return dict(rule=rule, ode=ode, de=1 << 1)


SELF_GF = """
struct UserLaw {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double gf;
  __device__ explicit UserLaw(const double* p) : gf(p[1]) {}
  __device__ bool operator()(double, double& s) const { s = 0.0; return false; }
  __device__ void self(const double*, double* f) const { f[0] = gf; }
};
"""

"""General agentODE bodies (UserLaw::pair + UserLaw::ode) for the tests, each
as HIP C++ (compiled with hipRTC into the library's stage skeleton) and as
its CPU restatement for the oracle's general-body callback, with the same
operations in the same order.

SOFT2 / SOFT3 restate src/models/softSpheres.jl:19-36 (2-D) and :58-77
(3-D) over user parameters instead of the built-in soft-sphere columns:
velocities are user DE variables, the radius (and in 2-D the cellFate, as
1.0 or 2.0) user parameters, the mass the law parameter p[6].

CROWD sums into a non-position DE variable: a linear repulsion reaching
R (r_i + r_j), r in U1, and a crowding level q (U0, a DE variable) with
dt(q) = sum_j w_ij q_j - c q + 1, w_ij = 1 - d/(R (r_i + r_j)), reading the
neighbors' q at the stage state."""

SOFT2 = """
struct UserLaw {
  static constexpr int sums = 2;
  double b, mu, m;
  const double* f0;
  __device__ explicit UserLaw(const double* p) : b(p[0]), mu(p[1]), m(p[6]), f0(p + 2) {}
  __device__ void pair(const LawAgent& i, const LawAgent& j, const double* dx, double d, double* s) const {
    const double rij = i.u(2) + j.u(2);
    if (d < mu * rij && d > 0) {
      const double f = f0[2 * ((int)i.u(3) - 1) + ((int)j.u(3) - 1)];
      const double sv = f * (rij / d - 1) * (mu * rij / d - 1);
      s[0] = s[0] + (sv * dx[0]) / d;
      s[1] = s[1] + (sv * dx[1]) / d;
    }
  }
  __device__ void ode(const LawAgent& i, const double* s, double* dxdt, double* dudt) const {
    dxdt[0] = i.u(0);
    dxdt[1] = i.u(1);
    dudt[0] = -b * i.u(0) / m + s[0] / m;
    dudt[1] = -b * i.u(1) / m + s[1] / m;
  }
};
"""


def soft2(p):
    b, mu, f0, m = p[0], p[1], p[2:6], p[6]

    def pair(i, j, dx, d, s):
        rij = i.u[2] + j.u[2]
        if d < mu * rij and d > 0:
            f = f0[2 * (int(i.u[3]) - 1) + (int(j.u[3]) - 1)]
            sv = f * (rij / d - 1) * (mu * rij / d - 1)
            s[0] = s[0] + (sv * dx[0]) / d
            s[1] = s[1] + (sv * dx[1]) / d

    def ode(i, s, dxdt, dudt):
        dxdt[0] = i.u[0]
        dxdt[1] = i.u[1]
        dudt[0] = -b * i.u[0] / m + s[0] / m
        dudt[1] = -b * i.u[1] / m + s[1] / m

    return pair, ode, 2


SOFT3 = """
struct UserLaw {
  static constexpr int sums = 3;
  double b, mu, f, m;
  __device__ explicit UserLaw(const double* p) : b(p[0]), mu(p[1]), f(p[2]), m(p[6]) {}
  __device__ void pair(const LawAgent& i, const LawAgent& j, const double* dx, double d, double* s) const {
    const double rij = i.u(3) + j.u(3);
    if (d < mu * rij && d > 0) {
      const double sv = f * (rij / d - 1) * (mu * rij / d - 1);
      for (int c = 0; c < 3; ++c) s[c] = s[c] + (sv * dx[c]) / d;
    }
  }
  __device__ void ode(const LawAgent& i, const double* s, double* dxdt, double* dudt) const {
    for (int c = 0; c < 3; ++c) {
      dxdt[c] = i.u(c);
      dudt[c] = -b * i.u(c) / m + s[c] / m;
    }
  }
};
"""


def soft3(p):
    b, mu, f, m = p[0], p[1], p[2], p[6]

    def pair(i, j, dx, d, s):
        rij = i.u[3] + j.u[3]
        if d < mu * rij and d > 0:
            sv = f * (rij / d - 1) * (mu * rij / d - 1)
            for c in range(3):
                s[c] = s[c] + (sv * dx[c]) / d

    def ode(i, s, dxdt, dudt):
        for c in range(3):
            dxdt[c] = i.u[c]
            dudt[c] = -b * i.u[c] / m + s[c] / m

    return pair, ode, 3


CROWD = """
struct UserLaw {
  static constexpr int sums = 4;
  double k, R, c;
  __device__ explicit UserLaw(const double* p) : k(p[0]), R(p[1]), c(p[2]) {}
  __device__ void pair(const LawAgent& i, const LawAgent& j, const double* dx, double d, double* s) const {
    const double reach = R * (i.u(1) + j.u(1));
    if (d > 0 && d < reach) {
      const double w = 1 - d / reach;
      const double f = k * w / d;
      s[0] = s[0] + f * dx[0];
      s[1] = s[1] + f * dx[1];
      s[2] = s[2] + f * dx[2];
      s[3] = s[3] + w * j.u(0);
    }
  }
  __device__ void ode(const LawAgent& i, const double* s, double* dxdt, double* dudt) const {
    dxdt[0] = s[0];
    dxdt[1] = s[1];
    dxdt[2] = s[2];
    dudt[0] = s[3] - c * i.u(0) + 1.0;
  }
};
"""


def crowd(p):
    k, R, c = p[0], p[1], p[2]

    def pair(i, j, dx, d, s):
        reach = R * (i.u[1] + j.u[1])
        if d > 0 and d < reach:
            w = 1 - d / reach
            f = k * w / d
            s[0] = s[0] + f * dx[0]
            s[1] = s[1] + f * dx[1]
            s[2] = s[2] + f * dx[2]
            s[3] = s[3] + w * j.u[0]

    def ode(i, s, dxdt, dudt):
        dxdt[0] = s[0]
        dxdt[1] = s[1]
        dxdt[2] = s[2]
        dudt[0] = s[3] - c * i.u[0] + 1.0

    return pair, ode, 4

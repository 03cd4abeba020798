"""mediumODE bodies for the hipRTC medium tests, each as HIP C++ (struct
UserMedium, compiled into the library's FTCS skeleton) and as its CPU
restatement on oracle.MediumNode (the oracle's user-medium callback)."""

# the built-in law: diffusion with agent uptake
BUILTIN = """
struct UserMedium {
  double D, k;
  __device__ UserMedium(double D_, double consume, const double*) : D(D_), k(consume) {}
  __device__ double operator()(const MediumNode& m) const { return D * m.laplacian() - (k * m.c) * (double)m.agents; }
};
"""


def builtin(D, k):
    return lambda m: D * m.laplacian() - (k * m.c) * float(m.agents)


# agents secrete (p[6] per agent), the field decays (p[7]) and a source grows along x
SECRETE = """
struct UserMedium {
  double D, prod, decay;
  __device__ UserMedium(double D_, double, const double* p) : D(D_), prod(p[6]), decay(p[7]) {}
  __device__ double operator()(const MediumNode& m) const {
    return D * m.laplacian() + prod * (double)m.agents - decay * m.c + 0.01 * m.x[0];
  }
};
"""


def secrete(D, prod, decay):
    return lambda m: D * m.laplacian() + prod * float(m.agents) - decay * m.c + 0.01 * m.x[0]


# the node geometry itself: dt(c) = x + 1000 idx[1]
GEOMETRY = """
struct UserMedium {
  __device__ UserMedium(double, double, const double*) {}
  __device__ double operator()(const MediumNode& m) const { return m.x[0] + 1000.0 * (double)m.idx[1]; }
};
"""


def geometry(m):
    return m.x[0] + 1000.0 * float(m.idx[1])

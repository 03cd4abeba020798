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


# the reference's mediumODE idiom (docs/src/Usage_Agent.md:390-403): inside
# nodes diffuse, the x- border reflects (p = p[2, i2_]), the x+ border
# absorbs (p = 0); other nodes (the y borders) diffuse through the mirror
IDIOM = """
struct UserMedium {
  double D;
  __device__ UserMedium(double D_, double, const double*) : D(D_) {}
  __device__ double operator()(const MediumNode& m) const { return D * m.laplacian(); }
  __device__ bool assign(const MediumNode& m, double& v) const {
    if (m.inside()) return false;
    if (m.border[0] == -1) { v = m.nb[0][1]; return true; }
    if (m.border[0] == 1) { v = 0.0; return true; }
    return false;
  }
};
"""


def idiom(D):
    def body(m):
        return D * m.laplacian()

    def assign(m):
        if m.inside():
            return None
        if m.border[0] == -1:
            return m.nb[0][1]
        if m.border[0] == 1:
            return 0.0
        return None
    return body, assign

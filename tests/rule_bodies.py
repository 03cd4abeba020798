"""agentRule bodies for the hipRTC rule tests, each as HIP C++ (struct
UserRule, compiled into the library's rule skeleton) and as its CPU
restatement on oracle.RuleAgent (the oracle's CBM_RULE_CUSTOM callback)."""
import math

# test/testCommunity.jl:241-316: x = x + 1; odd ids removed, even ids add one (lf = id)
PARITY = """
struct UserRule {
  __device__ UserRule(const double*, const double*) {}
  __device__ void operator()(RuleAgent& a) const {
    a.nw.x[0] = a.old.x[0] + 1.0;
    if (a.id % 2 == 1) a.remove();
    else { RuleRec d = a.child(); d.lf = a.id; a.add(d); }
  }
};
"""


def parity(a):
    a.nw.x[0] = a.old.x[0] + 1.0
    if a.id % 2 == 1:
        a.remove()
    else:
        d = a.child()
        d.lf = a.id
        a.add(d)


# test/testCommunity.jl:319-356: CBMDistributions drawn into agent parameters
RANDOM = """
struct UserRule {
  __device__ UserRule(const double*, const double*) {}
  __device__ void operator()(RuleAgent& a) const {
    a.nw.u[0] = a.normal(1.0, 2.0);
    a.nw.u[1] = a.uniform(1.0, 2.0);
    a.nw.u[2] = a.exponential(3.0);
  }
};
"""


def random_draws(a):
    a.nw.u[0] = a.normal(1.0, 2.0)
    a.nw.u[1] = a.uniform(1.0, 2.0)
    a.nw.u[2] = a.exponential(3.0)


# examples/Bacteries.ipynb agentRule: at vol >= ratio * vbirth, two new
# agents +-split along a random axis, each with half the volume and a fresh
# growth rate, and the parent removed; u[0] counts generations
DIVIDE_2D = """
struct UserRule {
  double ratio, split, gm, gs;
  __device__ UserRule(const double* rp, const double*) : ratio(rp[0]), split(rp[1]), gm(rp[2]), gs(rp[3]) {}
  __device__ void operator()(RuleAgent& a) const {
    if (!(a.old.vol >= ratio * a.old.vbirth)) return;
    const double th = a.uniform(0.0, 6.283185307179586);
    const double c = cos(th), s = sin(th);
    const double half = a.old.vol / 2.0;
    for (int k = 0; k < 2; ++k) {
      RuleRec d = a.child();
      const double sg = k == 0 ? 1.0 : -1.0;
      d.x[0] = a.old.x[0] + sg * split * c;
      d.x[1] = a.old.x[1] + sg * split * s;
      d.vol = half;
      d.vbirth = half;
      d.growth = a.uniform(gm * (1.0 - gs), gm * (1.0 + gs));
      d.lf = a.id;
      d.u[0] = a.old.u[0] + 1.0;
      a.add(d);
    }
    a.remove();
  }
};
"""


def divide_2d(rp):
    ratio, split, gm, gs = rp

    def body(a):
        if not (a.old.vol >= ratio * a.old.vbirth):
            return
        th = a.uniform(0.0, 6.283185307179586)
        c, s = math.cos(th), math.sin(th)
        half = a.old.vol / 2.0
        for k in range(2):
            d = a.child()
            sg = 1.0 if k == 0 else -1.0
            d.x[0] = a.old.x[0] + sg * split * c
            d.x[1] = a.old.x[1] + sg * split * s
            d.vol = half
            d.vbirth = half
            d.growth = a.uniform(gm * (1.0 - gs), gm * (1.0 + gs))
            d.lf = a.id
            d.u[0] = a.old.u[0] + 1.0
            a.add(d)
        a.remove()
    return body


# three @addAgent calls: one more than a rule may make per agent and step
THREE_ADDS = """
struct UserRule {
  __device__ UserRule(const double*, const double*) {}
  __device__ void operator()(RuleAgent& a) const {
    if (a.id == 1) { a.add(a.child()); a.add(a.child()); a.add(a.child()); }
  }
};
"""


# testCommunity.jl:361-415: nn = number of neighbors closer than rp[0], by
# @loopOverNeighbors, into u[0]
COUNT_NEIGHBORS = """
struct UserRule {
  double r;
  __device__ UserRule(const double* rp, const double*) : r(rp[0]) {}
  __device__ void operator()(RuleAgent& a) const {
    double nn = 0.0;
    a.forNeighbors([&](const RuleNeighbor& n) { if (n.d < r) nn += 1.0; });
    a.nw.u[0] = nn;
  }
};
"""


def count_neighbors(r):
    def body(a):
        nn = 0.0
        for n in a.neighbors():
            if n.d < r:
                nn += 1.0
        a.nw.u[0] = nn
    return body


# contact inhibition: an agent with fewer than rp[1] neighbors closer than
# rp[0] adds one at +0.3 along x; u[1] = distance to its nearest neighbor,
# u[2] = the neighbors' u[3] summed in walk order; every agent is pushed
# 0.01 * (its neighbor count) along x
INHIBIT = """
struct UserRule {
  double r, nmax;
  __device__ UserRule(const double* rp, const double*) : r(rp[0]), nmax(rp[1]) {}
  __device__ void operator()(RuleAgent& a) const {
    double nn = 0.0, dmin = 1e9, su = 0.0;
    a.forNeighbors([&](const RuleNeighbor& n) {
      if (n.d < r) { nn += 1.0; su = su + n.u(3); }
      dmin = n.d < dmin ? n.d : dmin;
    });
    a.nw.u[0] = nn;
    a.nw.u[1] = dmin;
    a.nw.u[2] = su;
    a.nw.x[0] = a.old.x[0] + 0.01 * nn;
    if (nn < nmax) {
      RuleRec d = a.child();
      d.x[0] = a.old.x[0] + 0.3;
      d.u[3] = a.old.u[3] + 1.0;
      d.lf = a.id;
      a.add(d);
    }
  }
};
"""


def inhibit(rp):
    r, nmax = rp[0], rp[1]

    def body(a):
        nn, dmin, su = 0.0, 1e9, 0.0
        for n in a.neighbors():
            if n.d < r:
                nn += 1.0
                su = su + n.u(3)
            dmin = n.d if n.d < dmin else dmin
        a.nw.u[0] = nn
        a.nw.u[1] = dmin
        a.nw.u[2] = su
        a.nw.x[0] = a.old.x[0] + 0.01 * nn
        if nn < nmax:
            d = a.child()
            d.x[0] = a.old.x[0] + 0.3
            d.u[3] = a.old.u[3] + 1.0
            d.lf = a.id
            a.add(d)
    return body


# the docs' secretion idiom (Usage_Agent.md:390-395: `p += secrete*dt/dx*dy`
# in agentRule, p a medium variable at the agent's node): each agent adds
# rp[0] * dt to its node and keeps in u[0] what the node held as the step
# started (a medium read in agent code, auxiliar.jl:353-365)
SECRETE = """
struct UserRule {
  double s;
  __device__ UserRule(const double* rp, const double*) : s(rp[0]) {}
  __device__ void operator()(RuleAgent& a) const {
    a.nw.u[0] = a.medium();
    a.medium_add(s * a.dt);
  }
};
"""


def secrete(s):
    def body(a):
        a.nw.u[0] = a.medium()
        a.medium_add(s * a.dt)
    return body


# eight user parameters: u[7] counts the agent's steps, u[5] doubles, u[6]
# sums the neighbors' u[4] read before the rule; every third agent adds a
# daughter carrying u[6] + 1 in its u[7]
EIGHT = """
struct UserRule {
  __device__ UserRule(const double*, const double*) {}
  __device__ void operator()(RuleAgent& a) const {
    a.nw.u[7] = a.old.u[7] + 1.0;
    a.nw.u[5] = a.old.u[5] * 2.0;
    double s = 0.0;
    a.forNeighbors([&](const RuleNeighbor& n) { if (n.d < 1.5) s = s + n.u(4); });
    a.nw.u[6] = s;
    if (a.id % 3 == 0 && a.step == 0) {
      RuleRec d = a.child();
      d.u[7] = s + 1.0;
      d.x[0] = a.old.x[0] + 0.25;
      a.add(d);
    }
  }
};
"""


def eight(a):
    a.nw.u[7] = a.old.u[7] + 1.0
    a.nw.u[5] = a.old.u[5] * 2.0
    s = 0.0
    for n in a.neighbors():
        if n.d < 1.5:
            s = s + n.u(4)
    a.nw.u[6] = s
    if a.id % 3 == 0 and a.step == 0:
        d = a.child()
        d.u[7] = s + 1.0
        d.x[0] = a.old.x[0] + 0.25
        a.add(d)

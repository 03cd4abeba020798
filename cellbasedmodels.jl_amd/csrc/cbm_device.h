// cbm_device.h -- device-side building blocks of the agent step kernels
// (gfx950, FP64).  Self-contained so that the same text can be compiled by
// hipcc into libcbm_hip.so and by hipRTC around a user interaction body.
//
// Arithmetic is written in the exact evaluation order of the reference
// expressions it restates; the library is compiled with -ffp-contract=off so
// no multiply-add is fused and results are bit-identical to an IEEE host
// evaluation of the same expressions.
#ifndef CBM_DEVICE_H
#define CBM_DEVICE_H

#ifndef CBM_HIPRTC
#include <hip/hip_runtime.h>
#include <stdint.h>
#else
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

namespace cbm {

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon et al., SC'11): counter-based, so every agent draws
// its own noise from (seed, step, id, sim, purpose) with no RNG state.
struct u4 { uint32_t x, y, z, w; };

__device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    u4 o;
    o.x = hi1 ^ c.y ^ k0; o.y = lo1;
    o.z = hi0 ^ c.w ^ k1; o.w = lo0;
    c = o;
  }
  return c;
}

__device__ __forceinline__ u4 draw(uint64_t seed, int64_t step, int64_t id, int32_t sim, uint32_t purpose) {
  u4 c; c.x = (uint32_t)step; c.y = (uint32_t)id; c.z = (uint32_t)sim; c.w = purpose;
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// 53-bit uniform in [0,1).
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Box-Muller transform (Box & Muller, 1958) of one Philox block.
__device__ __forceinline__ void normal2(u4 r, double& z0, double& z1) {
  double u1 = 1.0 - u53(r.x, r.y);
  double u2 = u53(r.z, r.w);
  double rad = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586 * u2;
  double s, c;
  sincos(th, &s, &c);
  z0 = rad * c;
  z1 = rad * s;
}

// floor of the exact real quotient a/b for a >= 0, b > 0: the reference's
// Int(a ÷ b) (src/neighbors.jl:636, src/metrics.jl:19).  The fused
// multiply-add gives the exact residual a - k*b, which corrects the rounded
// quotient by one where a/b rounded across an integer.
__device__ __forceinline__ int64_t floor_div(double a, double b) {
  double k = floor(a / b);
  double r = fma(-k, b, a);
  if (r < 0.0) k -= 1.0; else if (r >= b) k += 1.0;
  return (int64_t)k;
}

// ------------------------------------------------------------- pair laws
// Pair magnitude s(d) of the built-in agentODE bodies with their constants
// hoisted out of the pair loop.  operator() returns whether the pair
// interacts (coincident agents, d == 0, never do) and always sets s, so the
// caller can update its accumulators with selects instead of branches.
enum { LAW_NONE = 0, LAW_LINEAR = 1, LAW_REP_ADH = 2, LAW_AGGREGATION = 3, LAW_CUSTOM = 4, LAW_SOFT_SPHERES = 5,
       LAW_RODS = 6 };

template <class A, class B> struct same_type { static constexpr bool value = false; };
template <class A> struct same_type<A, A> { static constexpr bool value = true; };

template <int LAW> struct PairLaw;

// s = k (c - d) for d < c
// Traits of a law functor: active (has a pair body), per_axis_division
// (force s*(dx/d) per axis instead of (s/d)*dx), walls (the Aggregation
// box-wall push).
template <> struct PairLaw<LAW_LINEAR> {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double k, c;
  __device__ __forceinline__ explicit PairLaw(const double* p) : k(p[0]), c(p[1]) {}
  __device__ __forceinline__ bool operator()(double d, double& s) const {
    s = k * (c - d);
    return d > 0.0 && d < c;
  }
  __device__ __forceinline__ double reach() const { return c; }
};

// c = 2 r_eq, dc = f r_eq, a4 = 4 eps / (dc - c)^2:
//   s = k (c - d)                      d < c
//   s = 0 - (a4 (d - c)) (dc - d)      c <= d < dc   (well of depth eps)
template <> struct PairLaw<LAW_REP_ADH> {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double k, c, dc, a4;
  __device__ __forceinline__ explicit PairLaw(const double* p) : k(p[0]) {
    c = 2.0 * p[1];
    dc = p[3] * p[1];
    const double w = dc - c;
    a4 = (4.0 * p[2]) / (w * w);
  }
  __device__ __forceinline__ bool operator()(double d, double& s) const {
    const double rep = k * (c - d);
    const double adh = 0.0 - a4 * (d - c) * (dc - d);
    s = d < c ? rep : adh;
    return d > 0.0 && d < dc;
  }
  __device__ __forceinline__ double reach() const { return dc; }
};

// docs/src/Aggregation.md:88-98:
//   s = fRep (rRep - d)  d < rRep;  s = -fAtr (rAtr - d)  rRep <= d < rAtr
template <> struct PairLaw<LAW_AGGREGATION> {
  static constexpr bool active = true, per_axis_division = true, walls = true;   // dirx = (x - x[i2])/d
  double rRep, fRep, rAtr, fAtr;
  __device__ __forceinline__ explicit PairLaw(const double* p) : rRep(p[0]), fRep(p[1]), rAtr(p[2]), fAtr(p[3]) {}
  __device__ __forceinline__ bool operator()(double d, double& s) const {
    const double rep = fRep * (rRep - d);
    const double atr = 0.0 - fAtr * (rAtr - d);
    s = d < rRep ? rep : atr;
    return d > 0.0 && (d < rRep || d < rAtr);
  }
  __device__ __forceinline__ double reach() const { return fmax(rRep, rAtr); }
};

// Derivatives of phi(d) = s(d)/d of PairLaw<LAW_REP_ADH> for the adjoint:
// dphi/dd and dphi/d(k, r_eq, eps), with dc = f r_eq moving with r_eq.
//   d < c:        s = k (c - d)              ds/dd = -k, ds/dk = c - d, ds/dr = 2k
//   c <= d < dc:  s = -a4 (d - c)(dc - d)    ds/dd = -a4 (dc + c - 2d)
//                 ds/deps = -(4/w^2)(d - c)(dc - d)
//                 ds/dr   = (2 a4/r)(d - c)(dc - d) + 2 a4 (dc - d) - f a4 (d - c)
struct RepAdhGrad {
  double k, r, f, c, dc, a4, e4, a4r, fa4;
  __device__ __forceinline__ explicit RepAdhGrad(const double* p) : k(p[0]), r(p[1]), f(p[3]) {
    c = 2.0 * r;
    dc = f * r;
    const double w = dc - c;
    const double w2 = w * w;
    a4 = (4.0 * p[2]) / w2;
    e4 = 4.0 / w2;
    a4r = 2.0 * a4 / r;
    fa4 = f * a4;
  }
  // Branch-free: returns whether the pair interacts and fills phi = s/d,
  // dphi/dd / d and dphi/dtheta, with id = 1/d.
  __device__ __forceinline__ bool operator()(double d, double id, double& phi, double& dphid,
                                             double* dth) const {
    const double u = d - c, v = dc - d;
    const double uv = u * v;
    const bool rep = d < c;
    const double s = rep ? k * (c - d) : 0.0 - a4 * uv;
    const double sd = rep ? -k : -a4 * (dc + c - 2.0 * d);
    const double sk = rep ? c - d : 0.0;
    const double sr = rep ? 2.0 * k : a4r * uv + 2.0 * a4 * v - fa4 * u;
    const double se = rep ? 0.0 : -e4 * uv;
    phi = s * id;
    dphid = (sd - phi) * id * id;
    dth[0] = sk * id; dth[1] = sr * id; dth[2] = se * id;
    return d > 0.0 && d < dc;
  }
};

// src/models/softSpheres.jl:25-30 (2-D), :64-71 (3-D): with r_ij = r_i + r_j,
// for 0 < d < mu r_ij the force on i along each axis gains
// ((f0[fate_i][fate_j] (r_ij/d - 1)) (mu r_ij/d - 1) (x_i - x_j)) / d.
// p = {b, mu, f0[1][1], f0[1][2], f0[2][1], f0[2][2]}.  The stage skeleton
// handles this law apart (per-agent radius and fate, inertial variables).
template <> struct PairLaw<LAW_SOFT_SPHERES> {
  static constexpr bool active = true, per_axis_division = false, walls = false;
  double mu;
  const double* f0;
  __device__ __forceinline__ explicit PairLaw(const double* p) : mu(p[1]), f0(p + 2) {}
  __device__ __forceinline__ bool pair(double d, double ri, double rj, int64_t fi, int64_t fj, double& s) const {
    const double rij = ri + rj;
    const double f = f0[2 * (fi - 1) + (fj - 1)];
    s = f * (rij / d - 1) * (mu * rij / d - 1);
    return d < mu * rij && d > 0;
  }
  __device__ __forceinline__ bool operator()(double, double& s) const { s = 0.0; return false; }
};

// ------------------------------------------------------------------ rods
// The 2-D rod model (src/models/rods.jl:21-124).  A rod is the segment of
// length l through its centre (x, y) at angle theta, thickened to diameter
// d.  Two rods touch when the closest "virtual spheres" on their axes
// (rodIntersection, src/metrics.jl:199-246) are closer than the mean
// diameter; the stage skeleton handles this law apart (rod_stage_body).
struct RodLaw {
  static constexpr bool active = true, per_axis_division = false, walls = false;
};

// A rod's axis: centre, length, cos and sin of the angle (each evaluated
// once; the reference recomputes them per call, to the same values).
struct RodAxis { double x, y, l, c, s; };

// intersection2lines (src/metrics.jl:148-166): the meeting point of the two
// axis lines, or the midpoint of the centres when |sin(t2 - t1)| <= 1e-7.
__device__ __forceinline__ void rod_meet(const RodAxis& p, const RodAxis& q, double& px, double& py) {
  const double cx = p.x - q.x, cy = p.y - q.y;
  const double nrm = p.c * q.s - p.s * q.c;
  if (fabs(nrm) > 0.0000001) {
    const double k = (-q.s * cx + q.c * cy) / nrm;
    px = k * p.c + p.x;
    py = k * p.s + p.y;
  } else {
    px = (p.x + q.x) / 2;
    py = (p.y + q.y) / 2;
  }
}

// point2line (src/metrics.jl:175-183): (ax, ay) projected onto q's axis line.
__device__ __forceinline__ void rod_project(double ax, double ay, const RodAxis& q, double& ox, double& oy) {
  const double cx = q.x - ax, cy = q.y - ay;
  const double k = q.s * cx - q.c * cy;
  ox = k * q.s + ax;
  oy = -k * q.c + ay;
}

// pointInsideRod (src/metrics.jl:194-202) with separation 0.99: a point on
// q's axis pulled to 0.99 of its distance from the centre, at most l/2.
__device__ __forceinline__ void rod_clamp(const RodAxis& q, double ax, double ay, double& ox, double& oy) {
  const double ex = q.x - ax, ey = q.y - ay;
  const double di = fmax(sqrt(ex * ex + ey * ey), 0.00000001);
  const double ux = (ax - q.x) / di, uy = (ay - q.y) / di;
  const double r = 0.99 * fmin(di, q.l / 2);
  ox = r * ux + q.x;
  oy = r * uy + q.y;
}

// rodIntersection (src/metrics.jl:210-246): the contact points (ax, ay) on p
// and (bx, by) on q, branch for branch.
__device__ __forceinline__ void rod_closest(const RodAxis& p, const RodAxis& q, double& ax, double& ay, double& bx,
                                            double& by) {
  double px, py;
  rod_meet(p, q, px, py);
  const double di = sqrt((p.x - px) * (p.x - px) + (p.y - py) * (p.y - py));
  double dj = sqrt((q.x - px) * (q.x - px) + (q.y - py) * (q.y - py));
  const double nrm = p.c * q.s - p.s * q.c;
  if (fabs(nrm) < 0.000001) {
    rod_project(px, py, p, ax, ay);
    rod_clamp(p, ax, ay, ax, ay);
    rod_project(px, py, q, bx, by);
    rod_clamp(q, bx, by, bx, by);
  } else if (di <= p.l / 2 && dj <= q.l / 2) {
    rod_clamp(p, px, py, ax, ay);
    rod_clamp(q, px, py, bx, by);
  } else if (di <= p.l / 2 && dj > q.l / 2) {
    rod_clamp(q, px, py, bx, by);
    rod_project(bx, by, p, ax, ay);
  } else if (di > p.l / 2 && dj <= q.l / 2) {
    rod_clamp(p, px, py, ax, ay);
    rod_project(ax, ay, q, bx, by);
  } else {
    rod_clamp(q, px, py, bx, by);
    rod_project(bx, by, p, ax, ay);
    dj = sqrt((p.x - ax) * (p.x - ax) + (p.y - ay) * (p.y - ay));
    if (dj > p.l / 2) {
      rod_clamp(p, px, py, ax, ay);
      double tx, ty;
      rod_project(ax, ay, q, tx, ty);
      dj = sqrt((q.x - tx) * (q.x - tx) + (q.y - ty) * (q.y - ty));
      if (dj < q.l / 2) { bx = tx; by = ty; }
    }
  }
}

// One rod of a pair: axis, diameter, mass, velocity.
struct RodState { RodAxis a; double d, m, vx, vy; };

// rodForces (src/models/rods.jl:21-59): the force (fx, fy) and torque w on
// rod i from rod j; false (and nothing set) without contact.  The overlap
// uses d_i, the effective mass m_i / 2.  An overlap below zero (d_i < r_ij <
// (d_i + d_j)/2) sets *domain: the reference's delta^1.5 has no real value.
__device__ __forceinline__ bool rod_force(const RodState& i, const RodState& j, double kn, double gn, double gt,
                                          double mucc, double& fx, double& fy, double& w, bool& domain) {
  double ax, ay, bx, by;
  rod_closest(i.a, j.a, ax, ay, bx, by);
  const double rij = sqrt((ax - bx) * (ax - bx) + (ay - by) * (ay - by));
  if (!(rij > 0.0 && rij < (i.d + j.d) / 2)) return false;
  const double del = i.d - rij;
  if (del < 0.0) domain = true;
  const double me = i.m / 2;
  const double nx = (ax - bx) / rij, ny = (ay - by) / rij;
  const double vx = i.vx - j.vx, vy = i.vy - j.vy;
  const double vn = nx * vx + ny * vy;
  const double fn = kn * pow(del, 1.5) - gn * me * del * vn;
  const double ft = -fmin(gt * me * pow(del, 0.5), mucc * fn);
  fx = fn * nx + ft * (vx - vn * nx);
  fy = fn * ny + ft * (vy - vn * ny);
  w = (ax - i.a.x) * fy - (ay - i.a.y) * fx;
  return true;
}

template <> struct PairLaw<LAW_NONE> {
  static constexpr bool active = false, per_axis_division = false, walls = false;
  __device__ __forceinline__ explicit PairLaw(const double*) {}
  __device__ __forceinline__ bool operator()(double, double& s) const { s = 0.0; return false; }
};

}  // namespace cbm

#endif  // CBM_DEVICE_H

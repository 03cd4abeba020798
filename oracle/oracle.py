"""ctypes wrapper of liboracle.so -- the CPU restatement of the reference
semantics.  TEST INFRASTRUCTURE ONLY: imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
checker, never as the product path.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "orc_create": (P, [P, P]),
            "orc_destroy": (None, [P]),
            "orc_set": (C.c_int, [P, C.c_int, P, C.c_int64]),
            "orc_get": (C.c_int, [P, C.c_int, P, C.c_int64]),
            "orc_n": (C.c_int64, [P]),
            "orc_id_max": (C.c_int64, [P]),
            "orc_rebuilds": (C.c_int64, [P]),
            "orc_t": (C.c_double, [P]),
            "orc_error": (C.c_int, [P]),
            "orc_init": (C.c_int, [P]),
            "orc_step": (C.c_int, [P, C.c_int64]),
            "orc_neighbors_of": (C.c_int64, [P, C.c_int64, C.c_double, P, C.c_int64]),
            "orc_count_neighbors": (C.c_int, [P, C.c_double, P]),
            "orc_get_medium": (C.c_int, [P, P, C.c_int64]),
            "orc_set_medium": (C.c_int, [P, P, C.c_int64]),
            "orc_philox": (None, [P, P, P]),
            "orc_normals": (None, [C.c_uint64, C.c_int64, C.c_int64, C.c_int32, C.c_uint32, P]),
            "orc_floor_div": (C.c_int64, [C.c_double, C.c_double]),
            "orc_threads": (C.c_int, []),
            "orc_set_threads": (None, [C.c_int]),
            "orc_set_noise_slope": (None, [P, C.c_double]),
            "orc_set_rule": (None, [P, P]),
            "orc_rule_neighbors": (C.c_int64, [P, C.c_int64, P, C.c_int64]),
            "orc_rule_value": (C.c_double, [P, C.c_int, C.c_int64]),
            "orc_set_medium_body": (None, [P, P]),
            "orc_set_medium_assign": (None, [P, P]),
            "orc_traj_loss": (C.c_double, [P, P, C.c_int64, C.c_int32, P, P]),
            "orc_traj_grad_dual": (C.c_double, [P, P, C.c_int64, C.c_int32, P, P, P]),
            "orc_rod_intersection": (None, [P, P, P]),
            "orc_set_law_body": (C.c_int, [P, P, P, C.c_int]),
            "orc_set_model_body": (None, [P, P, P, P, C.c_uint]),
            "orc_param": (C.c_double, [P, C.c_int]),
            "orc_set_sim": (None, [P, C.c_int32]),
            "orc_rod_forces": (C.c_int, [P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


INT_FIELDS = (3, 10, 16)   # id, lf, cellFate (int64)


class RuleRec(C.Structure):
    """One agent's values as a CBM_RULE_CUSTOM body sees them (OrcRuleRec)."""
    _fields_ = [("x", C.c_double * 3), ("vol", C.c_double), ("growth", C.c_double), ("vbirth", C.c_double),
                ("u", C.c_double * 8), ("lf", C.c_int64), ("medium", C.c_double), ("madd", C.c_double)]


class LawAgent(C.Structure):
    """One agent as a general agentODE body sees it (OrcLawAgent): stage-state
    position, user parameters u[0..8) (DE variables at the stage state), vol."""
    _fields_ = [("row", C.c_int64), ("id", C.c_int64), ("x", C.c_double * 3), ("u", C.c_double * 8),
                ("vol", C.c_double)]


LAW_PAIR_FN = C.CFUNCTYPE(None, C.POINTER(LawAgent), C.POINTER(LawAgent), C.POINTER(C.c_double), C.c_double,
                          C.POINTER(C.c_double))
LAW_ODE_FN = C.CFUNCTYPE(None, C.POINTER(LawAgent), C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.POINTER(C.c_double))


class ModelScope(C.Structure):
    """What a model-scope body sees (OrcModelScope): t, dt, N, step, sim, the
    step-start parameters p[8] and their new values pn[8]."""
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("N", C.c_int64), ("step", C.c_int64), ("sim", C.c_int32),
                ("pad", C.c_int32), ("p", C.c_double * 8), ("pn", C.c_double * 8)]


MODEL_RULE_FN = C.CFUNCTYPE(None, C.POINTER(ModelScope))
MODEL_ODE_FN = C.CFUNCTYPE(None, C.POINTER(ModelScope), C.POINTER(C.c_double))


RULE_FN = C.CFUNCTYPE(None, C.c_int64, C.c_int64, C.POINTER(RuleRec), C.POINTER(RuleRec), C.POINTER(RuleRec),
                      C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64)


class RuleNeighbor:
    """A neighbor in a rule's loop (values as the loop started)."""

    def __init__(self, oracle, ag, j):
        lib_, h = oracle.lib, oracle.h
        self.row = j
        self.id = int(lib_.orc_rule_value(h, 12, j))
        self.x = [lib_.orc_rule_value(h, c, j) if c < ag.dims else 0.0 for c in range(3)]
        self.dx = [0.0, 0.0, 0.0]
        s2 = 0.0
        for c in range(ag.dims):
            e = ag.old.x[c] - self.x[c]
            if oracle.model.periodic:
                L = oracle.sim.box_hi[c] - oracle.sim.box_lo[c]
                e = e - L if e > 0.5 * L else (e + L if e < -0.5 * L else e)
            self.dx[c] = e
            s2 = e * e if c == 0 else s2 + e * e
        self.d = math.sqrt(s2)
        self._o, self._j = oracle, j

    def vol(self):
        return self._o.lib.orc_rule_value(self._o.h, 3, self._j)

    def u(self, k):
        return self._o.lib.orc_rule_value(self._o.h, 4 + k, self._j)
RULE_PURPOSE = 0x600


class MediumNode(C.Structure):
    """What a mediumODE body sees at one node (OrcMediumNode)."""
    _fields_ = [("c", C.c_double), ("lap", C.c_double * 3), ("x", C.c_double * 3), ("idx", C.c_int64 * 3),
                ("agents", C.c_int32), ("dims", C.c_int32), ("dt", C.c_double),
                ("nb", (C.c_double * 2) * 3), ("border", C.c_int32 * 3), ("pad", C.c_int32)]

    def inside(self):
        """@mediumInside(): on no border of any axis."""
        return all(self.border[a] == 0 for a in range(self.dims))

    def laplacian(self):
        lap = self.lap[0]
        if self.dims > 1:
            lap = lap + self.lap[1]
        if self.dims > 2:
            lap = lap + self.lap[2]
        return lap


MEDIUM_FN = C.CFUNCTYPE(C.c_double, C.POINTER(MediumNode))
MEDIUM_ASSIGN_FN = C.CFUNCTYPE(C.c_int, C.POINTER(MediumNode), C.POINTER(C.c_double))


def u53(a, b):
    return ((a >> 5) * 67108864.0 + (b >> 6)) * (1.0 / 9007199254740992.0)


class RuleAgent:
    """CPU restatement of a rule body's view of its agent (old, nw, child(),
    add(), remove(), and the CBMDistributions of src/random.jl:1-23 on the agent's
    k-th Philox draw of the step, purpose RULE_PURPOSE + k)."""

    def __init__(self, seed, id, step, dims, dt, old, nw, add):
        self.seed, self.id, self.step, self.dims, self.dt = seed, id, step, dims, dt
        self.old, self.nw, self._add = old, nw, add
        self.n_add, self.dead, self._k = 0, 0, 0

    def child(self):
        return RuleRec.from_buffer_copy(self.old)

    def add(self, d):
        if self.n_add >= 2:
            raise RuntimeError("more than 2 additions")
        C.memmove(C.addressof(self._add[self.n_add]), C.addressof(d), C.sizeof(RuleRec))
        self.n_add += 1

    def remove(self):
        self.dead = 1

    def medium(self):
        """The medium at the agent's node as the step started."""
        return self.old.medium

    def medium_add(self, v):
        """`p += v` on the agent's node (summed per node in 32.32 fixed point)."""
        self.nw.madd = self.nw.madd + v

    def _words(self):
        ctr = np.array([self.step & 0xFFFFFFFF, self.id & 0xFFFFFFFF, 0, RULE_PURPOSE + self._k], dtype=np.uint32)
        key = np.array([self.seed & 0xFFFFFFFF, self.seed >> 32], dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        lib().orc_philox(ctr.ctypes.data, key.ctypes.data, out.ctypes.data)
        self._k += 1
        return [int(w) for w in out]

    def rand(self):
        w = self._words()
        return u53(w[0], w[1])

    def uniform(self, l0, l1):
        return (l1 - l0) * self.rand() + l0

    def normal(self, mu, sigma):
        z = np.zeros(2)
        lib().orc_normals(self.seed, self.step, self.id, 0, RULE_PURPOSE + self._k, z.ctypes.data)
        self._k += 1
        return sigma * float(z[0]) + mu

    def exponential(self, theta):
        return -math.log(1.0 - self.rand()) * theta


class Oracle:
    """One simulation in logical slot order (the reference's CPU platform)."""

    def __init__(self, model, sim):
        self.lib = lib()
        self.model = model
        self.sim = sim
        self.h = self.lib.orc_create(C.addressof(model), C.addressof(sim))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def set(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.int64 if field in INT_FIELDS else np.float64)
        rc = self.lib.orc_set(self.h, field, v.ctypes.data, v.size)
        assert rc == 0, rc

    def get(self, field, n=None):
        n = self.n if n is None else n
        v = np.empty(n, dtype=np.int64 if field in INT_FIELDS else np.float64)
        rc = self.lib.orc_get(self.h, field, v.ctypes.data, n)
        assert rc == 0, rc
        return v

    @property
    def n(self):
        return self.lib.orc_n(self.h)

    @property
    def id_max(self):
        return self.lib.orc_id_max(self.h)

    @property
    def t(self):
        return self.lib.orc_t(self.h)

    def set_rule(self, body):
        """CBM_RULE_CUSTOM: body(agent) restates the device rule on a RuleAgent."""
        m = self.model

        def cb(id_, step, old, nw, add, n_add, dead, slot):
            ag = RuleAgent(int(m.seed), int(id_), int(step), int(m.dims), float(m.dt), old.contents, nw.contents,
                           [add[0], add[1]])
            ag.neighbors = lambda: self._rule_neighbors(ag, int(slot))
            body(ag)
            n_add[0] = ag.n_add
            dead[0] = ag.dead

        self._rule_cb = RULE_FN(cb)
        self.lib.orc_set_rule(self.h, C.cast(self._rule_cb, C.c_void_p))

    def set_model_body(self, rule=None, ode=None, noise=None, de=0):
        """A model-scope body: rule(m) writes m.pn; ode(m, dp) and noise(m, g)
        give dt and g of the parameters whose bit is set in `de`."""
        none = C.c_void_p(None)
        self._model_cb = (MODEL_RULE_FN(lambda m: rule(m.contents)) if rule else None,
                          MODEL_ODE_FN(lambda m, dp: ode(m.contents, dp)) if ode else None,
                          MODEL_ODE_FN(lambda m, g: noise(m.contents, g)) if noise else None)
        ptrs = [C.cast(f, C.c_void_p) if f is not None else none for f in self._model_cb]
        self.lib.orc_set_model_body(self.h, *ptrs, int(de))

    def set_sim(self, sim):
        """This simulation's index in a batched context (keys its random draws)."""
        self.lib.orc_set_sim(self.h, int(sim))

    def params(self):
        return np.array([self.lib.orc_param(self.h, k) for k in range(8)])

    def set_law_body(self, pair, ode, sums):
        """A general agentODE body: pair(i, j, dx, d, s) adds neighbor j's terms
        to the sums s[0..sums); ode(i, s, dxdt, dudt) writes dt of the
        positions and of the user DE variables (LawAgent records)."""
        self._law_cb = (LAW_PAIR_FN(lambda i, j, dx, d, s: pair(i.contents, j.contents, dx, d, s)),
                        LAW_ODE_FN(lambda i, s, dxdt, dudt: ode(i.contents, s, dxdt, dudt)))
        rc = self.lib.orc_set_law_body(self.h, C.cast(self._law_cb[0], C.c_void_p),
                                       C.cast(self._law_cb[1], C.c_void_p), int(sums))
        assert rc == 0, rc

    def set_medium_body(self, body, assign=None):
        """A user mediumODE body: body(node) returns dt(c) at a MediumNode;
        assign(node), if given, returns the value a node is assigned (its
        `p = ...` branches) or None."""
        self._medium_cb = MEDIUM_FN(lambda nd: float(body(nd.contents)))
        self.lib.orc_set_medium_body(self.h, C.cast(self._medium_cb, C.c_void_p))
        if assign is not None:
            def cb(nd, v):
                r = assign(nd.contents)
                if r is None:
                    return 0
                v[0] = float(r)
                return 1
            self._medium_assign_cb = MEDIUM_ASSIGN_FN(cb)
            self.lib.orc_set_medium_assign(self.h, C.cast(self._medium_assign_cb, C.c_void_p))

    def _rule_neighbors(self, ag, slot):
        out = np.zeros(4096, dtype=np.int64)
        k = self.lib.orc_rule_neighbors(self.h, slot, out.ctypes.data, out.size)
        assert k <= out.size
        return [RuleNeighbor(self, ag, int(j)) for j in out[:k]]

    def set_noise_slope(self, slope):
        """State-dependent agentSDE term g_a(x) = noise + slope * x_a on the positions."""
        self.lib.orc_set_noise_slope(self.h, float(slope))

    @property
    def rebuilds(self):
        return self.lib.orc_rebuilds(self.h)

    def error(self):
        return self.lib.orc_error(self.h)

    def init(self):
        return self.lib.orc_init(self.h)

    def step(self, k=1):
        return self.lib.orc_step(self.h, int(k))

    def neighbors_of(self, i, radius, cap=4096):
        out = np.empty(cap, dtype=np.int64)
        k = self.lib.orc_neighbors_of(self.h, int(i), float(radius), out.ctypes.data, cap)
        return out[: min(k, cap)]

    def neighbor_sets(self, radius, cap=4096):
        return [sorted(self.neighbors_of(i, radius, cap).tolist()) for i in range(self.n)]

    def count_neighbors(self, radius):
        out = np.empty(max(self.n, 1), dtype=np.int64)
        self.lib.orc_count_neighbors(self.h, float(radius), out.ctypes.data)
        return out[: self.n]

    def medium(self, n):
        v = np.empty(n)
        assert self.lib.orc_get_medium(self.h, v.ctypes.data, n) == 0
        return v

    def set_medium(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert self.lib.orc_set_medium(self.h, v.ctypes.data, v.size) == 0

    def traj_loss(self, theta, n_steps, save_steps, target):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        ss = np.ascontiguousarray(save_steps, dtype=np.int64)
        tg = np.ascontiguousarray(target, dtype=np.float64)
        return self.lib.orc_traj_loss(self.h, th.ctypes.data, int(n_steps), ss.size, ss.ctypes.data, tg.ctypes.data)

    def traj_grad_dual(self, theta, n_steps, save_steps, target):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        ss = np.ascontiguousarray(save_steps, dtype=np.int64)
        tg = np.ascontiguousarray(target, dtype=np.float64)
        g = np.zeros(3)
        loss = self.lib.orc_traj_grad_dual(self.h, th.ctypes.data, int(n_steps), ss.size, ss.ctypes.data,
                                           tg.ctypes.data, g.ctypes.data)
        return loss, g


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def normals(seed, step, id_, sim, purpose):
    z = np.zeros(2)
    lib().orc_normals(seed, step, id_, sim, purpose, z.ctypes.data)
    return z


def floor_div(a, b):
    return lib().orc_floor_div(float(a), float(b))


def threads():
    return lib().orc_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))


def rod_intersection(r1, r2):
    """Contact points of two rods r = (x, y, l, theta): (x1, y1, x2, y2)."""
    a = np.ascontiguousarray(r1, dtype=np.float64)
    b = np.ascontiguousarray(r2, dtype=np.float64)
    o = np.zeros(4)
    lib().orc_rod_intersection(a.ctypes.data, b.ctypes.data, o.ctypes.data)
    return o


def rod_forces(r1, r2, k):
    """rodForces on rod 1 from rod 2, r = (x, y, d, l, theta, vx, vy, m) and
    k = (kn, gamma_n, gamma_t, mu_cc): ((Fx, Fy, W), status)."""
    a = np.ascontiguousarray(r1, dtype=np.float64)
    b = np.ascontiguousarray(r2, dtype=np.float64)
    kk = np.ascontiguousarray(k, dtype=np.float64)
    o = np.zeros(3)
    rc = lib().orc_rod_forces(a.ctypes.data, b.ctypes.data, kk.ctypes.data, o.ctypes.data)
    return o, rc

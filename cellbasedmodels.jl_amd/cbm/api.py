"""ABM / Community / evolve_ over the HIP engine.

Mirrors the reference's user-facing seams:
  platform      src/platforms.jl:1-88          -> HIP()
  neighbors     src/neighbors.jl (CBMNeighbors) -> CBMNeighbors.*
  integrators   src/integrators.jl              -> CBMIntegrators.*
  models        src/models/softSpheres.jl, docs/src/Aggregation.md:62-123 -> CBMModels.*
  ABM           src/AgentStructure/agentStructure.jl:131-304
  Community     src/CommunityStructure/communityStructure.jl:71-187
  loadToPlatform!/bringFromPlatform!  communityStructure.jl:606-711
  step!/evolve!  src/CommunityStructure/step.jl:96-155
Errors from the library raise CBMError, as the reference raises `error()`.
"""
import ctypes as C

import numpy as np

from . import _abi as A


# ------------------------------------------------------------------ platform
class HIP:
    """The MI355X platform (the reference's CPU()/GPU() seam, src/platforms.jl:9-65)."""

    def __repr__(self):
        return "HIP()"


def version():
    return A.lib().cbm_version().decode()


# ----------------------------------------------------------------- neighbors
class CBMNeighbors:
    """Neighbor algorithms of src/neighbors.jl with the reference's constructor keywords."""

    class Full:
        kind = A.FULL

        def _fill(self, m):
            pass

    class CellLinked:
        kind = A.CELL_LINKED

        def __init__(self, cellEdge):
            self.cellEdge = float(cellEdge)

        def _fill(self, m):
            m.cell_edge = self.cellEdge

    class VerletDisplacement:
        """List radius `skin` (the reference's meaning, src/neighbors.jl:385); the list is
        rebuilt when an agent moved `rebuildDisplacement` (default skin/2, :403)."""
        kind = A.VERLET_DISPLACEMENT

        def __init__(self, skin, nMaxNeighbors, rebuildDisplacement=None):
            self.skin = float(skin)
            self.nMaxNeighbors = int(nMaxNeighbors)
            self.rebuildDisplacement = rebuildDisplacement

        def _fill(self, m):
            m.skin = self.skin
            m.max_neighbors = self.nMaxNeighbors
            m.rebuild_disp = float(self.rebuildDisplacement) if self.rebuildDisplacement else 0.0

    class VerletTime:
        kind = A.VERLET_TIME

        def __init__(self, skin, dtNeighborRecompute, nMaxNeighbors):
            self.skin = float(skin)
            self.dtNeighborRecompute = float(dtNeighborRecompute)
            self.nMaxNeighbors = int(nMaxNeighbors)

        def _fill(self, m):
            m.skin = self.skin
            m.dt_recompute = self.dtNeighborRecompute
            m.max_neighbors = self.nMaxNeighbors

    class CLVD:
        kind = A.CLVD

        def __init__(self, skin, nMaxNeighbors, cellEdge, rebuildDisplacement=None):
            self.skin = float(skin)
            self.nMaxNeighbors = int(nMaxNeighbors)
            self.cellEdge = float(cellEdge)
            self.rebuildDisplacement = rebuildDisplacement

        def _fill(self, m):
            m.skin = self.skin
            m.max_neighbors = self.nMaxNeighbors
            m.cell_edge = self.cellEdge
            m.rebuild_disp = float(self.rebuildDisplacement) if self.rebuildDisplacement else 0.0


# --------------------------------------------------------------- integrators
class CBMIntegrators:
    """src/integrators.jl:465 (Euler), :516 (Heun), :583 (RungeKutta4), :683 (EM),
    :744 (EulerHeun)."""

    class Euler:
        kind = A.EULER

    class Heun:
        kind = A.HEUN

    class RungeKutta4:
        kind = A.RK4

    class EM:
        kind = A.EM

    class EulerHeun:
        kind = A.EULER_HEUN


# -------------------------------------------------------------------- models
class _Law:
    def __init__(self, law, p, walls=0.0):
        self.law = law
        self.p = list(p) + [0.0] * (8 - len(p))


class CBMModels:
    """Built-in agentODE interaction bodies (pair force s(d)*(x_i-x_j)/d)."""

    @staticmethod
    def noInteraction():
        return _Law(A.LAW_NONE, [])

    @staticmethod
    def linearRepulsion(k, contact):
        """F = k (contact - d) u for d < contact (softSpheres idiom, src/models/softSpheres.jl:19-36)."""
        return _Law(A.LAW_LINEAR, [k, contact])

    @staticmethod
    def repulsionAdhesion(k, r, eps, cutoffFactor=2.4):
        """Repulsion k(2r-d) for d<2r and a continuous adhesion well of depth eps
        out to cutoffFactor*r (the Aggregation shape of docs/src/Aggregation.md:92-98
        made continuous at both breakpoints)."""
        return _Law(A.LAW_REP_ADH, [k, r, eps, cutoffFactor])

    @staticmethod
    def custom(source, p=()):
        """A user interaction body in HIP C++: `struct UserLaw` with a constructor
        from the 8 law parameters and `bool operator()(double d, double& s)`
        returning the pair magnitude (force (s/d)(x_i - x_j)); compiled with
        hipRTC into the stage kernel (the reference evals its agentODE quote
        block, src/AgentStructure/functionDE.jl:110)."""
        law = _Law(A.LAW_CUSTOM, p)
        law.source = source
        return law

    @staticmethod
    def softSpheres(b, mu, f0):
        """CBMModels.softSpheres2D/3D (src/models/softSpheres.jl:1-81): inertial
        dynamics dt(v) = -b v/m + F/m, dt(x) = v with the pair force
        f0[cellFate_i][cellFate_j] (r_ij/d - 1)(mu r_ij/d - 1)(x_i - x_j)/d for
        d < mu r_ij (r_ij = r_i + r_j); f0 is 2 x 2 (cellFate 1 or 2).
        Per-agent fields: vx, vy, vz, r, m, cellFate."""
        f = np.asarray(f0, dtype=np.float64).reshape(2, 2)
        return _Law(A.LAW_SOFT_SPHERES, [b, mu, f[0, 0], f[0, 1], f[1, 0], f[1, 1]])

    @staticmethod
    def rods2D(kn, gamma_n, gamma_t, mu_cc, mu_cw, beta, beta_omega):
        """CBMModels.rods2D (src/models/rods.jl:61-121): 2-D rods of diameter d,
        length l and mass m at angle theta; the contact force and torque of
        rodForces summed over the neighbors into fx, fy, W and pressure, with
        dt(v) = -beta v + f/m, dt(theta) = omega and
        dt(omega) = W/(m (d + l)^2/12 + m d^2) - beta_omega omega.
        Per-agent fields: vx, vy, theta, ω, d, l, m (inputs) and fx, fy, W,
        pressure (the last evaluation's sums).  mu_cw is carried, unused, as
        in the reference."""
        return _Law(A.LAW_RODS, [kn, gamma_n, gamma_t, mu_cc, mu_cw, beta, beta_omega])

    @staticmethod
    def aggregation(rRep, fRep, rAtr, fAtr, walls=True):
        """The Aggregation model of docs/src/Aggregation.md:62-123 (with its wall push)."""
        return _Law(A.LAW_AGGREGATION, [rRep, fRep, rAtr, fAtr, 1.0 if walls else 0.0])


class CBMRules:
    """Built-in agentRule bodies (see enum cbm_rule in include/cbm_hip.h)."""

    @staticmethod
    def rodsGrowth(growth, sigma_lTarget, lMax, alpha):
        """rods2dGrowth (src/models/rods.jl:123-164) with CBMModels.rods2D: l
        grows at growth/(1 + alpha pressure); past lTarget a rod becomes two
        daughters of length (l - d)/2.  Per-agent field: lTarget."""
        return (A.RULE_RODS_GROWTH, [growth, sigma_lTarget, lMax, alpha])

    @staticmethod
    def growthDivision(g_mean=0.01, g_spread=0.1, split=0.5, ratio=2.0):
        return (A.RULE_GROWTH_DIVISION, [g_mean, g_spread, split, ratio])

    @staticmethod
    def parity():
        return (A.RULE_PARITY, [0.0, 0.0, 0.0, 0.0])

    @staticmethod
    def death(rate):
        return (A.RULE_DEATH, [rate, 0.0, 0.0, 0.0])


# ----------------------------------------------------------------------- ABM
class ABM:
    """Model description: ABM(dims; agentODE, agentSDE, agentRule, neighborsAlg, agentAlg, platform).

    agentODE    a CBMModels law (the interaction body); `drift` adds dt(x) = -drift*x
    agentSDE    diagonal noise amplitude g (EM and EulerHeun): dt(x) = g in the SDE part;
                a number (g on the positions) or a HIP C++ body defining
                struct UserNoise (g of each DE variable at the agent's state),
                compiled with hipRTC by loadToPlatform_
    agentRule   a built-in rule (CBMRules.*) or a HIP C++ body defining struct
                UserRule (optionally as (body, rule_p)), compiled with hipRTC
    agent       user agent parameters {name: float} (at most 8) for a UserRule
                body: Community attributes, u[agent::name] in the body
    deVariables names of user agent parameters that are DE variables of a custom
                agentODE body (UserLaw::de; the dt(var) lines of functionDE.jl)
    medium      {"NMedium": [...], "D", "consume", "c0"} for the built-in
                diffusion with uptake; "mediumODE": a HIP C++ body defining
                struct UserMedium (dt(c) at a node) replaces that law;
                "mediumAlg": "FTCS" (default) or "DGADI" (Douglas-Gunn ADI);
                "borders": one entry per axis side (x-, x+, y-, y+, z-, z+;
                missing ones "mirror"): "mirror" (zero flux), "copy" (the
                reference's `p = p[2, i2_]` idiom: the inward neighbor's
                value) or a number (a Dirichlet value, `p = v`)
    lagged      dt(x) reads the previous step's committed force accumulator, the
                reference's read-old/write-new idiom (agentStructure.jl:352-453)
    periodic    periodic box with minimum-image distances
    modelScope  a HIP C++ body defining struct UserModel: the modelRule,
                modelODE and modelSDE of the reference over the simulation's
                parameters p[8] (cbm_compile_model_scope), run once per step
                after the medium; com.ctx.params() reads them back
    """

    def __init__(self, dims, agentODE=None, agentSDE=0.0, agentRule=None, drift=0.0,
                 neighborsAlg=None, agentAlg=None, platform=None, periodic=False,
                 lagged=False, seed=0, medium=None, agent=None, deVariables=(), modelScope=None):
        if dims not in (1, 2, 3):
            raise ValueError("dims must be 1, 2 or 3")
        self.dims = dims
        self.modelScope = modelScope
        self.agentODE = agentODE if agentODE is not None else CBMModels.noInteraction()
        self.agentSDE = agentSDE if isinstance(agentSDE, str) else float(agentSDE)
        if isinstance(agentRule, str):
            agentRule = (agentRule, [0.0] * 4)
        if agentRule is not None and isinstance(agentRule[0], str):
            self.rule_src = agentRule[0]
            agentRule = (A.RULE_CUSTOM, list(agentRule[1]))
        else:
            self.rule_src = None
        self.agentRule = agentRule
        agent = dict(agent or {})
        if len(agent) > A.USER_PARAMS:
            raise ValueError(f"at most {A.USER_PARAMS} user agent parameters")
        de = tuple(deVariables)
        if len(set(de)) != len(de):
            raise ValueError("deVariables names a parameter more than once")
        if any(k not in agent for k in de):
            raise ValueError("deVariables must name user agent parameters")
        if len(de) > 4:
            raise ValueError("at most 4 deVariables (U0..U3)")
        if de and self.agentODE.law != A.LAW_CUSTOM:
            raise ValueError("deVariables need a custom agentODE body defining UserLaw::de")
        if agent and self.rule_src is None and not de:
            raise ValueError("user agent parameters need an agentRule body or deVariables")
        agent = dict([(k, agent[k]) for k in de] + [(k, v) for k, v in agent.items() if k not in de])
        self.user_de = len(de)
        for name, typ in agent.items():
            if typ is not float or name in _FIELDS or not name.isidentifier():
                raise ValueError(f"user agent parameter {name!r} must be a new Float64 name")
        self.user_fields = {name: A.user_field(k) for k, name in enumerate(agent)}
        self.drift = float(drift)
        self.neighbors = neighborsAlg if neighborsAlg is not None else CBMNeighbors.Full()
        self.agentAlg = agentAlg if agentAlg is not None else CBMIntegrators.Euler()
        self.platform = platform if platform is not None else HIP()
        if not isinstance(self.platform, HIP):
            raise ValueError("this package implements the HIP() platform")
        self.periodic = bool(periodic)
        self.lagged = bool(lagged)
        self.seed = int(seed)
        self.medium = medium

    def model_desc(self, dt):
        m = A.ModelDesc()
        m.dims = self.dims
        m.integrator = self.agentAlg.kind
        m.neighbors = self.neighbors.kind
        self.neighbors._fill(m)
        m.law = self.agentODE.law
        m.periodic = 1 if self.periodic else 0
        m.lagged = 1 if self.lagged else 0
        m.dt = float(dt)
        m.drift = self.drift
        m.noise = 0.0 if isinstance(self.agentSDE, str) else self.agentSDE
        m.seed = self.seed
        m.user_de = self.user_de
        if self.agentRule is not None:
            m.rule = self.agentRule[0]
            for q, v in enumerate(self.agentRule[1]):
                m.rule_p[q] = v
        if self.medium is not None:
            for a, n in enumerate(self.medium["NMedium"]):
                m.medium_n[a] = int(n)
            m.medium_D = float(self.medium.get("D", 0.0))
            m.medium_consume = float(self.medium.get("consume", 0.0))
            m.medium_c0 = float(self.medium.get("c0", 0.0))
            m.medium_alg = {"FTCS": A.MEDIUM_FTCS, "DGADI": A.MEDIUM_DGADI}[self.medium.get("mediumAlg", "FTCS")]
            borders = list(self.medium.get("borders", ()))
            if len(borders) > 2 * self.dims:
                raise ValueError(f"borders: at most {2 * self.dims} sides for a {self.dims}-D medium")
            for k, b in enumerate(borders):
                if isinstance(b, str):
                    if b not in ("mirror", "copy"):
                        raise ValueError(f"unknown medium border {b!r}")
                    m.medium_border[k] = A.BORDER_COPY if b == "copy" else A.BORDER_MIRROR
                else:
                    m.medium_border[k] = A.BORDER_DIRICHLET
                    m.medium_border_value[k] = float(b)
        return m


def sim_desc(abm, N, n_max, simBox, law_p=None):
    s = A.SimDesc()
    s.n = int(N)
    s.n_max = int(max(n_max, N, 1))
    box = np.asarray(simBox, dtype=np.float64).reshape(-1, 2)
    for a in range(abm.dims):
        s.box_lo[a] = box[a, 0]
        s.box_hi[a] = box[a, 1]
    p = law_p if law_p is not None else abm.agentODE.p
    for q, v in enumerate(p[:8]):
        s.p[q] = float(v)
    return s


_FIELDS = {"x": A.F_X, "y": A.F_Y, "z": A.F_Z, "id": A.F_ID, "fx": A.F_FX, "fy": A.F_FY,
           "fz": A.F_FZ, "vol": A.F_VOL, "growth": A.F_GROWTH, "vbirth": A.F_VBIRTH, "lf": A.F_LF,
           "vx": A.F_VX, "vy": A.F_VY, "vz": A.F_VZ, "r": A.F_RADIUS, "m": A.F_MASS, "cellFate": A.F_FATE,
           "theta": A.F_THETA, "ω": A.F_OMEGA, "d": A.F_DIAMETER, "l": A.F_LENGTH, "W": A.F_TORQUE,
           "pressure": A.F_PRESSURE, "lTarget": A.F_LTARGET}


class Context:
    """Owns one cbm_ctx (a set of simulations on the current device)."""

    def __init__(self, model, sims):
        self.lib = A.lib()
        arr = (A.SimDesc * len(sims))(*sims)
        h = C.c_void_p()
        rc = self.lib.cbm_create(C.byref(h), C.byref(model), len(sims), arr)
        self.h = h
        if rc != A.OK:
            msg = self.lib.cbm_last_error(h).decode() if h else ""
            if h:
                self.lib.cbm_destroy(h)
                self.h = None
            raise A.CBMError(rc, msg)
        self.model = model
        self.sims = list(sims)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        A.check(rc, self.h)

    def upload(self, field, values, sim=0):
        isint = field in A.INT_FIELDS
        v = np.ascontiguousarray(values, dtype=np.int64 if isint else np.float64)
        self._chk(self.lib.cbm_upload(self.h, field, v.ctypes.data, v.size, sim))

    def download(self, field, n, sim=0):
        isint = field in A.INT_FIELDS
        v = np.empty(int(n), dtype=np.int64 if isint else np.float64)
        self._chk(self.lib.cbm_download(self.h, field, v.ctypes.data, v.size, sim))
        return v

    def set_n(self, n, sim=0):
        self._chk(self.lib.cbm_set_n(self.h, sim, int(n)))

    def neighbors(self):
        self._chk(self.lib.cbm_neighbors(self.h))

    def step(self, n=1):
        self._chk(self.lib.cbm_step(self.h, int(n)))

    def step_graph(self, n, replays):
        self._chk(self.lib.cbm_step_graph(self.h, int(n), int(replays)))

    def sync(self):
        self._chk(self.lib.cbm_sync(self.h))

    def n(self, sim=0):
        v = C.c_int64()
        self._chk(self.lib.cbm_get_n(self.h, sim, C.byref(v)))
        return v.value

    def id_max(self, sim=0):
        v = C.c_int64()
        self._chk(self.lib.cbm_get_id_max(self.h, sim, C.byref(v)))
        return v.value

    def time(self):
        t = C.c_double()
        s = C.c_int64()
        self._chk(self.lib.cbm_get_t(self.h, C.byref(t), C.byref(s)))
        return t.value, s.value

    def rebuilds(self):
        v = C.c_int64()
        self._chk(self.lib.cbm_get_rebuilds(self.h, C.byref(v)))
        return v.value

    def bricks(self):
        """CellLinked LDS tiling: {'bdim', 'cap', 'nbricks', 'overflows'}."""
        bd = np.zeros(3, dtype=np.int32)
        cap, nb, over = C.c_int32(), C.c_int64(), C.c_int64()
        self._chk(self.lib.cbm_get_bricks(self.h, bd.ctypes.data, C.byref(cap), C.byref(nb), C.byref(over)))
        return {"bdim": bd.tolist(), "cap": cap.value, "nbricks": nb.value, "overflows": over.value}

    def sub_bricks(self):
        """The sub-brick table of a CellLinked context with rules: {'bdim', 'cap', 'nbricks'}."""
        bd = np.zeros(3, dtype=np.int32)
        cap, nb = C.c_int32(), C.c_int64()
        self._chk(self.lib.cbm_get_sub_bricks(self.h, bd.ctypes.data, C.byref(cap), C.byref(nb)))
        return {"bdim": bd.tolist(), "cap": cap.value, "nbricks": nb.value}

    def neighbor_lists(self, radius, cap, sim=0):
        n = self.n(sim)
        counts = np.zeros(max(n, 1), dtype=np.int32)
        lists = np.full(max(n, 1) * cap, -1, dtype=np.int64)
        self._chk(self.lib.cbm_get_neighbor_lists(self.h, sim, float(radius), counts.ctypes.data,
                                                  lists.ctypes.data, int(cap)))
        return counts[:n], lists[: n * cap].reshape(n, cap)

    def count_neighbors(self, radius, sim=0):
        n = self.n(sim)
        out = np.zeros(max(n, 1), dtype=np.int64)
        self._chk(self.lib.cbm_count_neighbors(self.h, sim, float(radius), out.ctypes.data))
        return out[:n]

    def last_timing(self):
        tot, force, nl = C.c_double(), C.c_double(), C.c_int64()
        self._chk(self.lib.cbm_last_timing(self.h, C.byref(tot), C.byref(force), C.byref(nl)))
        return tot.value, force.value, nl.value

    def compile(self, source):
        """Compile a user interaction body (struct UserLaw, HIP C++) with hipRTC."""
        self._chk(self.lib.cbm_compile(self.h, source.encode()))

    def compile_model(self, law=None, noise=None, rule=None, medium=None):
        """Compile user agentODE (struct UserLaw), agentSDE (struct UserNoise),
        agentRule (struct UserRule) and mediumODE (struct UserMedium) bodies, all of
        the model's in one call."""
        enc = lambda s: s.encode() if s is not None else None
        self._chk(self.lib.cbm_compile_model(self.h, enc(law), enc(noise), enc(rule), enc(medium)))

    # evolve!(saveEach) snapshots kept on the device (cbm_snapshot_*)
    def snapshot_config(self, fields, capacity):
        f = np.ascontiguousarray(fields, dtype=np.int32)
        self._chk(self.lib.cbm_snapshot_config(self.h, f.ctypes.data, f.size, int(capacity)))

    def snapshot_take(self):
        self._chk(self.lib.cbm_snapshot_take(self.h))

    def step_save(self, n, save_each):
        self._chk(self.lib.cbm_step_save(self.h, int(n), int(save_each)))

    def snapshot_count(self):
        k = C.c_int32()
        self._chk(self.lib.cbm_snapshot_count(self.h, C.byref(k)))
        return k.value

    def snapshot(self, k, field, sim=0):
        """(values in slot order, N, t) of snapshot k."""
        cap = int(self.sims[sim].n_max)
        v = np.empty(max(cap, 1), dtype=np.int64 if field in A.INT_FIELDS else np.float64)
        n, t = C.c_int64(), C.c_double()
        self._chk(self.lib.cbm_snapshot_get(self.h, int(k), int(field), int(sim), v.ctypes.data, cap,
                                            C.byref(n), C.byref(t)))
        return v[: n.value].copy(), n.value, t.value

    def snapshot_medium(self, k, nodes):
        v = np.empty(int(nodes))
        self._chk(self.lib.cbm_snapshot_get_medium(self.h, int(k), v.ctypes.data, v.size))
        return v

    def set_params(self, p, sim=0):
        v = np.ascontiguousarray(p, dtype=np.float64)
        self._chk(self.lib.cbm_set_params(self.h, sim, v.ctypes.data, v.size))

    def params(self, sim=0, n=8):
        """The model-scope parameters p[0..n) as the model-scope body has left them."""
        v = np.zeros(int(n))
        self._chk(self.lib.cbm_get_params(self.h, sim, v.ctypes.data, v.size))
        return v

    def compile_model_scope(self, body):
        """A modelRule / modelODE / modelSDE body (struct UserModel), run once per
        simulation and step after the medium step (cbm_compile_model_scope)."""
        self._chk(self.lib.cbm_compile_model_scope(self.h, body.encode()))

    def adjoint(self, n_steps, checkpoint_every, save_steps, target):
        """Loss and d(loss)/d(k, r_eq, eps) per simulation (cbm_adjoint).

        target: per simulation, an array [n_saves, dims, N] in logical slot order,
        or None to reuse the targets of the previous call (already on the device)."""
        ss = np.ascontiguousarray(save_steps, dtype=np.int64)
        L = A.LossDesc()
        L.n_steps, L.checkpoint_every, L.n_saves = int(n_steps), int(checkpoint_every), ss.size
        L.save_steps = ss.ctypes.data_as(C.POINTER(C.c_int64))
        if target is None:
            L.target = None
        else:
            tg = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.float64).ravel() for t in target]))
            L.target = tg.ctypes.data_as(C.POINTER(C.c_double))
        S = len(self.sims)
        grad = np.zeros(3 * S)
        loss = np.zeros(S)
        self._chk(self.lib.cbm_adjoint(self.h, C.byref(L), grad.ctypes.data, loss.ctypes.data))
        return loss, grad.reshape(S, 3)

    def profile(self, nsteps):
        """Per-kernel average device times (ms) over nsteps steps, see cbm_profile."""
        ms = np.zeros(16)
        nl = np.zeros(16, dtype=np.int64)
        self._chk(self.lib.cbm_profile(self.h, int(nsteps), ms.ctypes.data, nl.ctypes.data, 16))
        return ms, nl

    def set_block(self, block):
        self._chk(self.lib.cbm_set_block(self.h, int(block)))


# ----------------------------------------------------------------- Community
class Community:
    """State of one simulation: Community(abm; N, dt, simBox, ...) of the reference.

    Agent parameters are numpy arrays in the reference's slot order (``com.x``,
    ``com.id``, ...); they live on the device between loadToPlatform_ and
    bringFromPlatform_, exactly as in the reference.
    """

    def __init__(self, abm, N, dt, simBox, lawParameters=None, **params):
        self.abm = abm
        self.N = int(N)
        self.dt = float(dt)
        self.t = 0.0
        self.simBox = np.asarray(simBox, dtype=np.float64).reshape(-1, 2)[: abm.dims]
        self.law_p = lawParameters
        self.fields = dict(_FIELDS, **abm.user_fields)
        self.params = {}
        for k in ("x", "y", "z")[: abm.dims]:
            self.params[k] = np.zeros(self.N)
        self.params["id"] = np.arange(1, self.N + 1, dtype=np.int64)
        if abm.lagged:
            for k in ("fx", "fy", "fz")[: abm.dims]:
                self.params[k] = np.zeros(self.N)
        if abm.agentODE.law == A.LAW_SOFT_SPHERES:
            for k in ("vx", "vy", "vz")[: abm.dims]:
                self.params[k] = np.zeros(self.N)
            self.params["r"] = np.ones(self.N)
            self.params["m"] = np.ones(self.N)
            self.params["cellFate"] = np.ones(self.N, dtype=np.int64)
        if abm.agentODE.law == A.LAW_RODS:
            for k in ("vx", "vy", "theta", "ω", "fx", "fy", "W", "pressure"):
                self.params[k] = np.zeros(self.N)
            for k in ("d", "l", "m"):
                self.params[k] = np.ones(self.N)
            if abm.agentRule is not None and abm.agentRule[0] == A.RULE_RODS_GROWTH:
                self.params["lTarget"] = np.zeros(self.N)
        if abm.rule_src is not None:
            self.params["vol"] = np.ones(self.N)
            self.params["growth"] = np.zeros(self.N)
            self.params["vbirth"] = np.ones(self.N)
            self.params["lf"] = np.zeros(self.N, dtype=np.int64)
            for k in abm.user_fields:
                self.params[k] = np.zeros(self.N)
        for k, v in params.items():
            self[k] = v
        self.ctx = None
        self.pastTimes = []

    def __getitem__(self, k):
        return self.params[k]

    def __setitem__(self, k, v):
        if k not in self.fields:
            raise KeyError(f"{k} is not an agent parameter of this model")
        dt = np.int64 if self.fields[k] in A.INT_FIELDS else np.float64
        arr = np.broadcast_to(np.asarray(v, dtype=dt), (self.N,)).copy()
        self.params[k] = arr

    def __getattr__(self, k):
        p = self.__dict__.get("params")
        if p is not None and k in p:
            return p[k]
        raise AttributeError(k)

    def __setattr__(self, k, v):
        if "params" in self.__dict__ and k in self.__dict__["fields"]:
            self[k] = v
        else:
            object.__setattr__(self, k, v)


def loadToPlatform_(com, preallocateAgents=0):
    """communityStructure.jl:606-655: allocate on the device (capacity N + preallocateAgents),
    upload every agent parameter and compute the neighbor structure."""
    abm = com.abm
    model = abm.model_desc(com.dt)
    sim = sim_desc(abm, com.N, com.N + int(preallocateAgents), com.simBox.reshape(-1).tolist(), com.law_p)
    com.ctx = Context(model, [sim])
    # user parameter names as indices of RuleAgent u[] and of UserLaw::de's u/du
    names = "".join(f"constexpr int {n} = {k};\n" for k, n in enumerate(abm.user_fields))
    names = "namespace agent {\n" + names + "}\n" if names else ""
    law_src = names + abm.agentODE.source if abm.agentODE.law == A.LAW_CUSTOM else None
    noise_src = abm.agentSDE if isinstance(abm.agentSDE, str) else None
    rule_src = None
    if abm.rule_src is not None:   # one definition of the names per program
        rule_src = (names if law_src is None else "") + abm.rule_src
    medium_src = abm.medium.get("mediumODE") if abm.medium is not None else None
    if law_src is not None or noise_src is not None or rule_src is not None or medium_src is not None:
        com.ctx.compile_model(law_src, noise_src, rule_src, medium_src)
    if abm.modelScope is not None:
        com.ctx.compile_model_scope(abm.modelScope)
    for k, arr in com.params.items():
        com.ctx.upload(com.fields[k], arr)
    com.ctx.neighbors()
    return com


def bringFromPlatform_(com):
    """communityStructure.jl:663-711: download every agent parameter (slot order, 1:N)."""
    ctx = com.ctx
    n = ctx.n()
    object.__setattr__(com, "N", n)
    com.t, _ = ctx.time()
    for k in list(com.params):
        com.params[k] = ctx.download(com.fields[k], n)
    return com


def computeNeighbors_(com):
    """src/neighbors.jl:18."""
    com.ctx.neighbors()


def step_(com, n=1):
    """step! (src/CommunityStructure/step.jl:96-107), n times, device-resident."""
    com.ctx.step(n)
    com.t, _ = com.ctx.time()


def evolve_(com, steps, saveEach=None, preallocateAgents=0, saveCurrentState=False):
    """evolve! (step.jl:124-155): loads if needed, runs `steps` steps and, after
    every step i with i % saveEach == 0, saves the community into com.pastTimes
    (the reference's saveRAM!, IO.jl:6-31).  The snapshots are taken on the
    device while the steps run and read back at the end."""
    if com.ctx is None:
        loadToPlatform_(com, preallocateAgents)
    ctx = com.ctx
    if not saveEach and not saveCurrentState:
        step_(com, steps)
    else:
        names = list(com.params)
        every = int(saveEach) if saveEach else steps + 1
        ctx.snapshot_config([com.fields[k] for k in names], steps // every + (1 if saveCurrentState else 0))
        if saveCurrentState:
            ctx.snapshot_take()
        ctx.step_save(steps, every)
        com.t, _ = ctx.time()
        med = com.abm.medium is not None
        nodes = int(np.prod(com.abm.medium["NMedium"])) if med else 0
        for k in range(ctx.snapshot_count()):
            snap = {}
            for name in names:
                snap[name], n, t = ctx.snapshot(k, com.fields[name])
            snap["t"], snap["N"] = t, n
            if med:
                snap["medium"] = ctx.snapshot_medium(k, nodes)
            com.pastTimes.append(snap)
        ctx.snapshot_config([], 0)
    bringFromPlatform_(com)
    return com


def getParameter(com, syms):
    """communityStructure.jl getParameter over the saved snapshots."""
    return {s: [snap[s] for snap in com.pastTimes] for s in syms}

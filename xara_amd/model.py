"""Model builder mirroring the reference's command interface for the
state-determination path.

The methods carry the names and argument meaning of the reference's Tcl /
OpenSeesPy commands (``node``, ``fix``, ``mass``, ``nDMaterial``,
``uniaxialMaterial``, ``section``, ``geomTransf``, ``element``, ``load``,
``system``, ``numberer``, ``test``, ``integrator``, ``analyze``,
``nodeDisp``), plus bulk variants (``nodes``, ``elements``...) for large
meshes and the state-determination primitives of the C-ABI (``update``,
``form_tangent``, ``form_unbalance``, ``commit`` ...).  Everything runs in
the gfx950 library through include/xara_gpu.h.
"""
import ctypes as C

import numpy as np

from . import _lib

OUT = {"stress": 1, "tangent": 2, "strain": 3, "plastic": 4}
OPS = {"update": 1, "tangent": 2, "unbalance": 3, "spmv": 4, "copy": 5}
SYSTEMS = {"gpupcg": 1, "pcg": 1, "gpublockpcg": 2, "blockpcg": 2, "dense": 3,
           "profilespd": 3, "umfpack": 3, "fullgeneral": 3}
NUMBERERS = {"rcm": 1, "plain": 2}
TESTS = {"normunbalance": 1, "normdispincr": 2, "energyincr": 3}


class XaraError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _d(a):
    return a.ctypes.data_as(_lib.P_D)


def _i(a):
    return a.ctypes.data_as(_lib.P_I)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


class Model:
    """One model on one GPU (``model basic -ndm ndm -ndf ndf``)."""

    def __init__(self, ndm, ndf=None, device=0, _binding=None):
        self._b = _binding if _binding is not None else _lib.load()
        ndf = ndf if ndf is not None else {2: 2, 3: 3}[ndm]
        h = C.c_void_p()
        self._check(self._b.create(ndm, ndf, device, C.byref(h)))
        self._h = h
        self.ndm, self.ndf = ndm, ndf
        self._node_tags = None
        self._mat_tags = {"Steel02": set(), "Concrete02": set()}

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc):
        if rc < 0:
            raise XaraError(rc, self._b.error())
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._b.destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- model definition -------------------------------------------------
    def node(self, tag, *crd):
        self.nodes([tag], [crd])

    def nodes(self, tags, crds):
        tags = _i32(tags)
        crds = _f64(crds).reshape(len(tags), self.ndm)
        self._check(self._b.nodes(self._h, len(tags), _i(tags), _d(crds)))
        known = self._node_tags if self._node_tags is not None else np.zeros(0, np.int32)
        self._node_tags = np.union1d(known, tags)  # ascending: the order of nodal results

    def fix(self, tag, *flags):
        self.fix_many([tag], [flags])

    def fix_many(self, tags, flags):
        tags = _i32(tags)
        flags = _i32(flags).reshape(len(tags), self.ndf)
        self._check(self._b.fix(self._h, len(tags), _i(tags), _i(flags)))

    def mass(self, tag, *m):
        self.masses([tag], [m])

    def masses(self, tags, m):
        tags = _i32(tags)
        m = _f64(m).reshape(len(tags), self.ndf)
        self._check(self._b.mass(self._h, len(tags), _i(tags), _d(m)))

    def nDMaterial(self, kind, tag, *args):
        if kind != "J2Plasticity":
            raise XaraError(-1, f"nDMaterial {kind} is not on the GPU path")
        K, G, sig0, sigInf, delta, H = args[:6]
        eta = args[6] if len(args) > 6 else 0.0
        self._check(self._b.nd_j2plasticity(self._h, tag, K, G, sig0, sigInf, delta, H, eta))

    def uniaxialMaterial(self, kind, tag, *args):
        if kind == "Steel02":
            # defaults of TclCommand_newFedeasSteel (fedeas.cpp:264-270)
            fy, E, b = args[:3]
            R0, cR1, cR2 = (list(args[3:6]) + [15.0, 0.925, 0.15][len(args[3:6]):])
            a1, a2, a3, a4 = (list(args[6:10]) + [0.0, 1.0, 0.0, 1.0][len(args[6:10]):])
            sig0 = args[10] if len(args) > 10 else 0.0
            self._check(self._b.uniaxial_steel02(self._h, tag, fy, E, b, R0, cR1, cR2, a1, a2, a3, a4, sig0))
            self._mat_tags["Steel02"].add(tag)
        elif kind == "Concrete02":
            fc, epsc0, fcu, epscu = args[:4]
            if len(args) >= 7:
                rat, ft, Ets = args[4:7]
            else:  # Concrete02.cpp:134-138
                rat, ft, Ets = 0.1, abs(0.1 * fc), 0.1 * fc / epsc0
            self._check(self._b.uniaxial_concrete02(self._h, tag, fc, epsc0, fcu, epscu, rat, ft, Ets))
            self._mat_tags["Concrete02"].add(tag)
        elif kind == "Elastic":
            self._check(self._b.uniaxial_elastic(self._h, tag, args[0]))
        else:
            raise XaraError(-1, f"uniaxialMaterial {kind} is not on the GPU path")

    STATE_KINDS = {"nodes": 1, "J2": 2, "Steel02": 3, "Concrete02": 4, "domain": 5}

    def get_state(self, kind):
        """Committed state of one kind in the reference's sendSelf layout
        (see include/xara_gpu.h: XG_STATE_*), as a flat float64 array."""
        k = self.STATE_KINDS[kind]
        n = self._b.state_size(self._h, k)
        self._check(n)
        out = np.zeros(n)
        self._check(self._b.get_state(self._h, k, _d(out)))
        return out

    def set_state(self, kind, values):
        """Write a committed state (trial state set equal to it)."""
        k = self.STATE_KINDS[kind]
        v = _f64(values)
        n = self._b.state_size(self._h, k)
        self._check(n)
        if v.size != n:
            raise XaraError(-1, f"state {kind}: {v.size} values given, {n} expected")
        self._check(self._b.set_state(self._h, k, _d(v)))

    def save_state(self):
        """All committed state kinds (checkpoint)."""
        return {k: self.get_state(k) for k in self.STATE_KINDS}

    def load_state(self, state):
        for k, v in state.items():
            self.set_state(k, v)

    def material_tags(self, kind):
        """Tags of the defined materials of one kind, ascending (the order of
        per-material results such as the adjoint gradients)."""
        return sorted(self._mat_tags[kind])

    def section(self, kind, tag, fibers, GJ=None, centroid=True):
        """``section Fiber tag <-GJ GJ> {fiber y z A mat ...}``; fibers is a
        sequence of (y, z, A, matTag) (z ignored in 2D)."""
        if kind != "Fiber":
            raise XaraError(-1, f"section {kind} is not on the GPU path")
        f = np.asarray(fibers, dtype=np.float64).reshape(-1, 4)
        y, z, A = _f64(f[:, 0]), _f64(f[:, 1]), _f64(f[:, 2])
        mats = _i32(f[:, 3])
        gj = -1.0 if GJ is None else float(GJ)
        self._check(self._b.section_fiber(self._h, tag, len(mats), _i(mats), _d(y), _d(z), _d(A), gj,
                                          1 if centroid else 0))

    def geomTransf(self, kind, tag, *vecxz):
        if kind != "Linear":
            raise XaraError(-1, f"geomTransf {kind} is not on the GPU path")
        v = _f64(vecxz if vecxz else [0.0, 0.0, 0.0])
        self._check(self._b.geom_transf_linear(self._h, tag, _d(v)))

    def element(self, kind, tag, *args):
        if kind in ("stdBrick", "brick"):
            nodes, rest = args[:8], args[8:]
            self.elements("stdBrick", [tag], [nodes], *rest)
        elif kind in ("quad", "FourNodeQuad"):
            nodes, thk, ptype, mat = args[:4], args[4], args[5], args[6]
            if ptype != "PlaneStrain":
                raise XaraError(-1, "quad: only PlaneStrain is on the GPU path")
            self.elements("quad", [tag], [nodes], thk, mat, *args[7:])
        elif kind == "dispBeamColumn":
            i, j, transf = args[:3]
            if isinstance(args[3], str):  # tag i j transf integration sec nip
                sec, nip = args[4], args[5]
            else:  # legacy: tag i j nip sec transf
                nip, sec, transf = args[2], args[3], args[4]
            self.elements("dispBeamColumn", [tag], [(i, j)], transf, sec, nip)
        elif kind == "forceBeamColumn":  # tag i j transf "Lobatto" sec nip <max_iter tol>
            i, j, transf, integ, sec, nip = args[:6]
            if integ != "Lobatto":
                raise XaraError(-1, f"forceBeamColumn: {integ} integration is not on the GPU path")
            self.elements("forceBeamColumn", [tag], [(i, j)], transf, sec, nip, *args[6:8])
        else:
            raise XaraError(-1, f"element {kind} is not on the GPU path")

    def elements(self, kind, tags, conn, *args):
        tags = _i32(tags)
        n = len(tags)
        if kind == "stdBrick":
            conn = _i32(conn).reshape(n, 8)
            mat = int(args[0])
            b = list(args[1:4]) + [0.0] * (3 - len(args[1:4]))
            self._check(self._b.elements_brick(self._h, n, _i(tags), _i(conn), mat, *b))
        elif kind == "quad":
            conn = _i32(conn).reshape(n, 4)
            thk, mat = float(args[0]), int(args[1])
            pressure, rho, b1, b2 = (list(args[2:6]) + [0.0] * 4)[:4]
            self._check(self._b.elements_quad(self._h, n, _i(tags), _i(conn), thk, mat, pressure, rho, b1, b2))
        elif kind == "dispBeamColumn":
            conn = _i32(conn).reshape(n, 2)
            transf, sec, nip = int(args[0]), int(args[1]), int(args[2])
            rho = float(args[3]) if len(args) > 3 else 0.0
            self._check(self._b.elements_dispbeam(self._h, n, _i(tags), _i(conn), transf, sec, nip, rho))
        elif kind == "forceBeamColumn":
            conn = _i32(conn).reshape(n, 2)
            transf, sec, nip = int(args[0]), int(args[1]), int(args[2])
            max_iter = int(args[3]) if len(args) > 3 else 10
            tol = float(args[4]) if len(args) > 4 else 1e-12
            self._check(self._b.elements_forcebeam(self._h, n, _i(tags), _i(conn), transf, sec, nip, max_iter, tol))
        else:
            raise XaraError(-1, f"element {kind} is not on the GPU path")

    def load(self, tag, *P):
        self.loads([tag], [P])

    def loads(self, tags, P):
        tags = _i32(tags)
        P = _f64(P).reshape(len(tags), self.ndf)
        self._check(self._b.nodal_loads(self._h, len(tags), _i(tags), _d(P)))

    def uniformExcitation(self, dof, accel, dt, factor=1.0):
        """``pattern UniformExcitation tag dof -accel {Path -dt dt -values ...}``
        (dof 1-based, translational)."""
        a = _f64(accel)
        self._check(self._b.uniform_excitation(self._h, int(dof) - 1, len(a), float(dt), _d(a), float(factor)))

    def rayleigh(self, alphaM, betaK=0.0, betaKinit=0.0, betaKcomm=0.0):
        self._check(self._b.rayleigh(self._h, alphaM, betaK, betaKinit, betaKcomm))

    # -- analysis objects -------------------------------------------------
    def system(self, kind, rtol=0.0, maxit=0):
        self._check(self._b.system(self._h, SYSTEMS[kind.lower()], rtol, maxit))

    def spmv_config(self, fmt="auto", lanes=0):
        """SpMV layout of the PCG: "auto" (node-blocked column indices when
        every node has all DOFs free or none), "csr", or "sym" (upper node
        blocks of a symmetric matrix, 2 or 3 DOFs per node); lanes per row
        or block row (0 = from the mean length)."""
        self._check(self._b.spmv_config(self._h, {"auto": 0, "csr": 1, "sym": 2}[fmt], int(lanes)))

    def cg_config(self, chunk=16, graph=True, fuse=True, small=True, warm=True, wide=True):
        """PCG issue pattern: iterations issued between convergence reads
        (even), whether a full chunk is replayed from a captured graph,
        whether small systems run the alpha / beta reductions inside the
        vector kernels (the iterates do not depend on these three), whether
        systems that fit one workgroup are solved by it in one launch (same
        iteration, sums in another order), and whether the first Newton
        iteration of a step (and DisplacementControl's reference-load solve)
        starts from the previous step's solution instead of zero (the same
        tolerance, fewer iterations), and whether systems of more than 1,024
        equations with a band of at most 1,024 are preconditioned by their
        exact tiled band factor (with small)."""
        flags = (int(bool(graph)) | 2 * int(bool(fuse)) | 4 * int(bool(small)) | 8 * int(bool(warm))
                 | 16 * int(bool(wide)))
        self._check(self._b.cg_config(self._h, int(chunk), flags))

    def numberer(self, kind):
        self._check(self._b.numberer(self._h, NUMBERERS[kind.lower()]))

    def test(self, kind, tol, maxit):
        self._check(self._b.test(self._h, TESTS[kind.lower()], tol, maxit))

    def integrator(self, kind, *args):
        if kind == "LoadControl":
            self._check(self._b.integrator_load_control(self._h, float(args[0])))
        elif kind == "DisplacementControl":
            self._check(self._b.integrator_disp_control(self._h, int(args[0]), int(args[1]), float(args[2])))
        elif kind == "Newmark":
            self._check(self._b.integrator_newmark(self._h, float(args[0]), float(args[1])))
        else:
            raise XaraError(-1, f"integrator {kind} is not on the GPU path")

    def analyze(self, n, dt=0.0):
        return self._check(self._b.analyze(self._h, n, dt))

    # -- setup and numbering ---------------------------------------------
    def setup(self):
        self._check(self._b.setup(self._h))

    @property
    def neq(self):
        return self._b.num_eqn(self._h)

    @property
    def nnz(self):
        return self._b.csr_nnz(self._h)

    @property
    def num_nodes(self):
        return self._b.num_nodes(self._h)

    def node_eqns(self):
        out = np.zeros((self.num_nodes, self.ndf), np.int32)
        self._check(self._b.node_eqns(self._h, _i(out)))
        return out

    def csr_pattern(self):
        rp = np.zeros(self.neq + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        self._check(self._b.csr_pattern(self._h, _i(rp), _i(ci)))
        return rp, ci

    # -- state determination ---------------------------------------------
    def set_trial_disp(self, U):
        U = _f64(U).reshape(self.num_nodes, self.ndf)
        self._check(self._b.set_trial_disp(self._h, _d(U)))

    def trial_disp(self):
        U = np.zeros((self.num_nodes, self.ndf))
        self._check(self._b.get_trial_disp(self._h, _d(U)))
        return U

    def incr_disp(self, dU):
        dU = _f64(dU)
        self._check(self._b.incr_disp(self._h, _d(dU)))

    def update(self):
        return self._check(self._b.update(self._h))

    def form_tangent(self, values=True):
        if not values:
            self._check(self._b.form_tangent(self._h, None))
            return None
        v = np.zeros(self.nnz)
        self._check(self._b.form_tangent(self._h, _d(v)))
        return v

    def form_unbalance(self):
        B = np.zeros(self.neq)
        self._check(self._b.form_unbalance(self._h, _d(B)))
        return B

    def solve(self, B):
        B = _f64(B)
        X = np.zeros(self.neq)
        self._check(self._b.solve(self._h, _d(B), _d(X)))
        return X

    def spmv(self, x):
        x = _f64(x)
        y = np.zeros(self.neq)
        self._check(self._b.spmv(self._h, _d(x), _d(y)))
        return y

    def commit(self):
        self._check(self._b.commit(self._h))

    def revert(self):
        self._check(self._b.revert_to_last_commit(self._h))

    def revert_to_start(self):
        self._check(self._b.revert_to_start(self._h))

    def set_load_factor(self, lam):
        self._check(self._b.set_load_factor(self._h, lam))

    @property
    def load_factor(self):
        return self._b.get_load_factor(self._h)

    def point_data(self, what, npts_total, max_width=36):
        out = np.zeros(npts_total * max_width)
        w = C.c_int(0)
        self._check(self._b.get_point_data(self._h, OUT[what], _d(out), C.byref(w)))
        return out[: npts_total * w.value].reshape(npts_total, w.value)

    def element_tangent(self, tag):
        K = np.zeros(24 * 24)
        nd = C.c_int(0)
        self._check(self._b.element_tangent(self._h, tag, _d(K), C.byref(nd)))
        n = nd.value
        return K[: n * n].reshape(n, n)

    def element_resisting_force(self, tag):
        R = np.zeros(24)
        nd = C.c_int(0)
        self._check(self._b.element_resisting_force(self._h, tag, _d(R), C.byref(nd)))
        return R[: nd.value].copy()

    def counters(self):
        c = _lib.Counters()
        self._check(self._b.get_counters(self._h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in c._fields_}

    def time_op(self, op, reps):
        ms = C.c_double(0)
        self._check(self._b.time_op(self._h, OPS[op], reps, C.byref(ms)))
        return ms.value

    # -- results ------------------------------------------------------------
    def _tags(self):
        if self._node_tags is None:
            raise XaraError(-2, "node tags unknown; call set_node_tags")
        return self._node_tags

    def nodeDisp(self, tag, dof=None, node_tags=None):
        tags = node_tags if node_tags is not None else self._tags()
        U = self.trial_disp()
        i = int(np.searchsorted(tags, tag))
        return U[i] if dof is None else U[i, dof - 1]

    def set_node_tags(self, tags):
        self._node_tags = np.sort(np.asarray(tags))

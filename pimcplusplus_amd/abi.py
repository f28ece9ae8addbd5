"""ctypes mirror of include/pimc_gpu.h and a thin numpy-facing wrapper.

`PimcLib` wraps any shared library that exports the pimc_* C ABI.  The product
loads the HIP library (`pimcplusplus_amd.native.load`); nothing here loads a
library by itself.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

PIMC_ABI_VERSION = 2
MAX_ACTIVE = 4
MAX_LEVELS = 8
MAX_SPECIES = 4

OK, ERR_ARG, ERR_STATE, ERR_DEVICE, ERR_NOMEM, ERR_UNSUPPORTED = range(6)
OLDMODE, NEWMODE, BOTHMODES = 0, 1, 2
FP64, MIXED = 0, 1
GRID_LINEAR, GRID_LOG = 0, 1

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)


class DavidFitDesc(C.Structure):
    _fields_ = [
        ("grid_type", C.c_int32),
        ("num_points", C.c_int32),
        ("grid_start", C.c_double),
        ("grid_end", C.c_double),
        ("num_tau", C.c_int32),
        ("num_cols", C.c_int32),
        ("num_terms", C.c_int32),
        ("reserved", C.c_int32),
        ("smallest_tau", C.c_double),
        ("desired_tau", C.c_double),
        ("y", _f64p),
        ("dy", _f64p),
    ]


class CoulombBCFitDesc(C.Structure):
    _fields_ = [
        ("q_grid_type", C.c_int32),
        ("q_points", C.c_int32),
        ("q_start", C.c_double),
        ("q_end", C.c_double),
        ("t_grid_type", C.c_int32),
        ("t_points", C.c_int32),
        ("t_start", C.c_double),
        ("t_end", C.c_double),
        ("num_betas", C.c_int32),
        ("reserved", C.c_int32),
        ("smallest_beta", C.c_double),
        ("z1z2", C.c_double),
        ("umat", _f64p),
        ("dumat", _f64p),
    ]


class ReplicaDesc(C.Structure):
    _fields_ = [
        ("num_particles", C.c_int32),
        ("num_time_slices", C.c_int32),
        ("num_species", C.c_int32),
        ("num_images", C.c_int32),
        ("box", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
        ("tau", C.c_double),
        ("species_count", _i32p),
        ("lambda_", _f64p),
        ("pair_fit", _i32p),
        ("k_cutoff", C.c_double),
        ("use_long_range", C.c_int32),
        ("use_background", C.c_int32),
        ("rng_stream", C.c_uint32),
        ("reserved", C.c_int32),
    ]


class ActionReq(C.Structure):
    _fields_ = [
        ("replica", C.c_int32),
        ("mode", C.c_int32),
        ("slice1", C.c_int32),
        ("slice2", C.c_int32),
        ("level", C.c_int32),
        ("nptcls", C.c_int32),
        ("ptcls", C.c_int32 * MAX_ACTIVE),
    ]


class RangeReq(C.Structure):
    _fields_ = [
        ("replica", C.c_int32),
        ("mode", C.c_int32),
        ("slice1", C.c_int32),
        ("slice2", C.c_int32),
        ("level", C.c_int32),
        ("reserved", C.c_int32),
    ]


ACTION_SHORT_RANGE, ACTION_KINETIC, ACTION_LONG_RANGE = 0, 1, 2


class BisectionDesc(C.Structure):
    _fields_ = [
        ("num_levels", C.c_int32),
        ("lowest_level", C.c_int32),
        ("species", C.c_int32),
        ("reserved", C.c_int32),
        ("seed", C.c_uint64),
    ]


class MoveRecord(C.Structure):
    _fields_ = [
        ("slice1", C.c_int32),
        ("ptcl", C.c_int32),
        ("accepted", C.c_int32),
        ("num_stages", C.c_int32),
        ("old_action", C.c_double * MAX_LEVELS),
        ("new_action", C.c_double * MAX_LEVELS),
        ("log_sample_ratio", C.c_double * MAX_LEVELS),
        ("log_u", C.c_double * MAX_LEVELS),
        ("cycle_length", C.c_int32),
        ("cycle", C.c_int32 * MAX_ACTIVE),
        ("num_cycles", C.c_int32),
        ("forward_t", C.c_double),
        ("reverse_t", C.c_double),
        ("accept_prob", C.c_double),
        ("psi", C.c_double),
    ]


class BisectionStats(C.Structure):
    _fields_ = [
        ("moves", C.c_int64),
        ("accepted", C.c_int64),
        ("stages", C.c_int64),
        ("pair_evals", C.c_int64),
        ("k_evals", C.c_int64),
        ("device_ms", C.c_double),
    ]


class BisectionPlan(C.Structure):
    _fields_ = [("block_threads", C.c_int32), ("link_split", C.c_int32), ("persistent", C.c_int32),
                ("cluster", C.c_int32), ("trial_ms", C.c_double)]


class Energy(C.Structure):
    _fields_ = [("kinetic", C.c_double), ("du_short", C.c_double), ("du_long", C.c_double),
                ("v_short", C.c_double), ("v_long", C.c_double), ("total", C.c_double)]


# name -> (restype, argtypes); every pimc_* entry point of include/pimc_gpu.h
_CTX = C.c_void_p
SIGNATURES = {
    "pimc_abi_version": (C.c_int, []),
    "pimc_last_error": (C.c_char_p, []),
    "pimc_ctx_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "pimc_ctx_destroy": (C.c_int, [_CTX]),
    "pimc_ctx_synchronize": (C.c_int, [_CTX]),
    "pimc_fit_create_david": (C.c_int, [_CTX, C.POINTER(DavidFitDesc), _i32p]),
    "pimc_fit_eval_u": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int64, _f64p, _f64p, _f64p, C.c_int32, _f64p]),
    "pimc_fit_create_coulomb_bc": (C.c_int, [_CTX, C.POINTER(CoulombBCFitDesc), _i32p]),
    "pimc_replica_create": (C.c_int, [_CTX, C.POINTER(ReplicaDesc), _i32p]),
    "pimc_replica_set_ulong": (
        C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32, _f64p,
                  _f64p]),
    "pimc_replica_set_vlong": (
        C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, _f64p, C.c_double, C.c_double, C.c_double, C.c_double,
                  C.c_double, C.c_int32, _f64p]),
    "pimc_replica_num_kvecs": (C.c_int, [_CTX, C.c_int32, _i32p]),
    "pimc_replica_kvecs": (C.c_int, [_CTX, C.c_int32, _f64p, _i32p]),
    "pimc_replica_set_long_range": (
        C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _f64p, _f64p, _f64p, _f64p,
                  _f64p, _f64p, _f64p, _f64p]),
    "pimc_estimator_energy": (C.c_int, [_CTX, _i32p, C.c_int32, C.c_int32, C.POINTER(Energy)]),
    "pimc_action_d_dbeta": (C.c_int, [_CTX, C.c_int32, C.POINTER(RangeReq), C.c_int32, _f64p]),
    "pimc_estimator_pair_correlation": (
        C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32,
                  C.POINTER(C.c_int64)]),
    "pimc_path_set": (C.c_int, [_CTX, C.c_int32, C.c_int32, _f64p]),
    "pimc_path_get": (C.c_int, [_CTX, C.c_int32, C.c_int32, _f64p]),
    "pimc_path_set_beads": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _f64p]),
    "pimc_permutation_get": (C.c_int, [_CTX, C.c_int32, C.c_int32, _i32p]),
    "pimc_accept_copy": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, _i32p, C.c_int32]),
    "pimc_reject_copy": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, _i32p, C.c_int32]),
    "pimc_move_join": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32]),
    "pimc_replica_join": (C.c_int, [_CTX, C.c_int32, _i32p]),
    "pimc_put_in_box": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int64, _f64p]),
    "pimc_permute_htable": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                      _f32p]),
    "pimc_cycle_apply": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _i32p]),
    "pimc_replica_set_permute": (C.c_int, [_CTX, C.c_int32, C.c_int32, _f64p, C.c_double]),
    "pimc_rho_k_update_all": (C.c_int, [_CTX, C.c_int32]),
    "pimc_rho_k_get": (C.c_int, [_CTX, C.c_int32, C.c_int32, _f64p]),
    "pimc_action_short_range": (C.c_int, [_CTX, C.POINTER(ActionReq), C.c_int32, C.c_int32, _f64p]),
    "pimc_action_short_range_noalloc": (C.c_int, [_CTX, C.POINTER(ActionReq), C.c_int32, C.c_int32, _f64p]),
    "pimc_action_kinetic": (C.c_int, [_CTX, C.POINTER(ActionReq), C.c_int32, _f64p]),
    "pimc_action_long_range": (C.c_int, [_CTX, C.POINTER(ActionReq), C.c_int32, C.c_int32, _f64p]),
    "pimc_dist_disp": (C.c_int, [_CTX, C.c_int32, C.c_int32, C.c_int64, _i32p, _f64p, _i32p]),
    "pimc_bisection_run": (
        C.c_int,
        [_CTX, _i32p, C.c_int32, C.POINTER(BisectionDesc), C.c_int64, C.c_int32, C.c_int32,
         C.POINTER(MoveRecord), C.POINTER(BisectionStats)]),
    "pimc_bisection_stats_take": (C.c_int, [_CTX, _i32p, C.c_int32, C.POINTER(BisectionStats)]),
    "pimc_bisection_set_plan": (C.c_int, [_CTX, C.POINTER(BisectionPlan)]),
    "pimc_bisection_get_plan": (C.c_int, [_CTX, C.POINTER(BisectionPlan)]),
    "pimc_bisection_tune": (
        C.c_int,
        [_CTX, _i32p, C.c_int32, C.POINTER(BisectionDesc), C.c_int64, C.c_int32, C.c_int32,
         C.POINTER(BisectionPlan)]),
    "pimc_sweep_graph_create": (
        C.c_int,
        [_CTX, _i32p, C.c_int32, C.POINTER(BisectionDesc), C.c_int64, C.c_int32, C.c_int32,
         C.POINTER(C.c_int32)]),
    "pimc_sweep_graph_launch": (C.c_int, [_CTX, C.c_int32, C.c_int32]),
    "pimc_sweep_graph_destroy": (C.c_int, [_CTX, C.c_int32]),
}


class PimcError(RuntimeError):
    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed with status {code}: {msg}")
        self.code = code


def bind(lib: C.CDLL) -> C.CDLL:
    """Declares argument and return types of every pimc_* symbol on `lib`."""
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


@dataclass
class ReplicaInfo:
    num_particles: int
    num_time_slices: int
    num_species: int
    nk: int


class PimcLib:
    """Numpy-facing wrapper of one library that exports the pimc_* ABI.

    One instance owns one context.  Methods raise PimcError on a non-zero status.
    """

    def __init__(self, lib: C.CDLL, device: int = 0):
        self.lib = bind(lib)
        if self.lib.pimc_abi_version() != PIMC_ABI_VERSION:
            raise RuntimeError("pimc ABI version mismatch")
        ctx = C.c_void_p()
        self._check("pimc_ctx_create", self.lib.pimc_ctx_create(device, C.byref(ctx)))
        self.ctx = ctx
        self.replicas: dict[int, ReplicaInfo] = {}

    # -- plumbing --------------------------------------------------------
    def _check(self, name: str, rc: int):
        if rc != OK:
            msg = self.lib.pimc_last_error()
            raise PimcError(name, rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.pimc_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self._check("pimc_ctx_synchronize", self.lib.pimc_ctx_synchronize(self.ctx))

    # -- fits --------------------------------------------------------------
    def fit_create_david(self, grid_type: int, grid_start: float, grid_end: float, y: np.ndarray,
                         num_terms: int, smallest_tau: float, desired_tau: float, dy=None) -> int:
        """y is [num_tau][num_points][num_cols] (column 0 potential, 1 u00, 2.. u_kj);
        dy the beta-derivative table in the same layout, or None."""
        y = _f64(y)
        if y.ndim != 3:
            raise ValueError("y must be [num_tau][num_points][num_cols]")
        dyp = None
        if dy is not None:
            dy = _f64(dy)
            if dy.shape != y.shape:
                raise ValueError("dy must have the shape of y")
            dyp = _p(dy, _f64p)
        d = DavidFitDesc(grid_type=grid_type, num_points=y.shape[1], grid_start=grid_start,
                         grid_end=grid_end, num_tau=y.shape[0], num_cols=y.shape[2],
                         num_terms=num_terms, reserved=0, smallest_tau=smallest_tau,
                         desired_tau=desired_tau, y=_p(y, _f64p), dy=dyp)
        fid = C.c_int32()
        self._check("pimc_fit_create_david", self.lib.pimc_fit_create_david(self.ctx, C.byref(d), C.byref(fid)))
        return fid.value

    def fit_create_coulomb_bc(self, q_grid: tuple, t_grid: tuple, umat: np.ndarray, smallest_beta: float,
                              z1z2: float, dumat=None) -> int:
        """q_grid, t_grid = (grid_type, start, end); umat (and dumat, dU/dbeta, or None) is [num_betas][nq][nt]."""
        u = _f64(umat)
        if u.ndim != 3:
            raise ValueError("umat must be [num_betas][nq][nt]")
        du = None
        if dumat is not None:
            du = _f64(dumat)
            if du.shape != u.shape:
                raise ValueError("dumat must have the shape of umat")
        d = CoulombBCFitDesc(q_grid_type=q_grid[0], q_points=u.shape[1], q_start=q_grid[1], q_end=q_grid[2],
                             t_grid_type=t_grid[0], t_points=u.shape[2], t_start=t_grid[1], t_end=t_grid[2],
                             num_betas=u.shape[0], reserved=0, smallest_beta=smallest_beta, z1z2=z1z2,
                             umat=_p(u, _f64p), dumat=_p(du, _f64p) if du is not None else None)
        fid = C.c_int32()
        self._check("pimc_fit_create_coulomb_bc",
                    self.lib.pimc_fit_create_coulomb_bc(self.ctx, C.byref(d), C.byref(fid)))
        return fid.value

    def set_ulong(self, rep: int, sa: int, sb: int, grid_start: float, grid_end: float, coefs: np.ndarray,
                  dcoefs=None):
        """coefs [num_levels][num_points][6] = {Y, B, C, D, E, F} per knot; dcoefs: dUlong, or None."""
        c = _f64(coefs)
        if c.ndim != 3 or c.shape[2] != 6:
            raise ValueError("coefs must be [num_levels][num_points][6]")
        dc = None
        if dcoefs is not None:
            dc = _f64(dcoefs)
            if dc.shape != c.shape:
                raise ValueError("dcoefs must have the shape of coefs")
        self._check("pimc_replica_set_ulong", self.lib.pimc_replica_set_ulong(
            self.ctx, rep, sa, sb, c.shape[0], grid_start, grid_end, c.shape[1], _p(c, _f64p),
            _p(dc, _f64p) if dc is not None else None))

    def set_vlong(self, rep: int, sa: int, sb: int, vlong_k, vlong_r0: float, vshort_k0: float, vlong_k0: float,
                  grid_start: float, grid_end: float, coefs: np.ndarray):
        """The pair's long-range potential: Vlong_k [nk], the three constants, and Vlong(r) as
        quintic data coefs [num_points][6] on LinearGrid(grid_start, grid_end, num_points)."""
        vk = _f64(vlong_k)
        c = _f64(coefs)
        if c.ndim != 2 or c.shape[1] != 6:
            raise ValueError("coefs must be [num_points][6]")
        self._check("pimc_replica_set_vlong", self.lib.pimc_replica_set_vlong(
            self.ctx, rep, sa, sb, _p(vk, _f64p), vlong_r0, vshort_k0, vlong_k0, grid_start, grid_end, c.shape[0],
            _p(c, _f64p)))

    def fit_eval_u(self, fit: int, level: int, q, z, s2, precision: int = FP64) -> np.ndarray:
        q, z, s2 = _f64(q), _f64(z), _f64(s2)
        n = q.size
        if z.size != n or s2.size != n:
            raise ValueError("q, z, s2 sizes differ")
        out = np.empty(n, dtype=np.float64)
        self._check("pimc_fit_eval_u", self.lib.pimc_fit_eval_u(
            self.ctx, fit, level, n, _p(q, _f64p), _p(z, _f64p), _p(s2, _f64p), precision, _p(out, _f64p)))
        return out

    # -- replicas ----------------------------------------------------------
    def replica_create(self, num_particles: int, num_time_slices: int, box, tau: float,
                       species_count, lambdas, pair_fit, num_images: int = 0,
                       periodic=(1, 1, 1), k_cutoff: float = 0.0, use_long_range: bool = False,
                       use_background: bool = False, rng_stream: int = 0) -> int:
        sc = _i32(species_count)
        lam = _f64(lambdas)
        pf = _i32(pair_fit).reshape(-1)
        ns = sc.size
        if lam.size != ns or pf.size != ns * ns:
            raise ValueError("species arrays have inconsistent sizes")
        d = ReplicaDesc()
        d.num_particles = num_particles
        d.num_time_slices = num_time_slices
        d.num_species = ns
        d.num_images = num_images
        for i in range(3):
            d.box[i] = float(box[i])
            d.periodic[i] = int(periodic[i])
        d.tau = tau
        d.species_count = _p(sc, _i32p)
        d.lambda_ = _p(lam, _f64p)
        d.pair_fit = _p(pf, _i32p)
        d.k_cutoff = k_cutoff
        d.use_long_range = int(use_long_range)
        d.use_background = int(use_background)
        d.rng_stream = rng_stream
        rid = C.c_int32()
        self._check("pimc_replica_create", self.lib.pimc_replica_create(self.ctx, C.byref(d), C.byref(rid)))
        nk = C.c_int32()
        self._check("pimc_replica_num_kvecs", self.lib.pimc_replica_num_kvecs(self.ctx, rid.value, C.byref(nk)))
        self.replicas[rid.value] = ReplicaInfo(num_particles, num_time_slices, ns, nk.value)
        return rid.value

    def kvecs(self, rep: int):
        info = self.replicas[rep]
        kv = np.empty((info.nk, 3), dtype=np.float64)
        ki = np.empty((info.nk, 3), dtype=np.int32)
        self._check("pimc_replica_kvecs", self.lib.pimc_replica_kvecs(self.ctx, rep, _p(kv, _f64p), _p(ki, _i32p)))
        return kv, ki

    def set_long_range(self, rep: int, sa: int, sb: int, ulong_k, ulong_r0, ushort_k0, ulong_k0,
                       dulong_k=None, dulong_r0=None, dushort_k0=None, dulong_k0=None):
        uk = _f64(ulong_k)
        r0, sk0, lk0 = _f64(ulong_r0), _f64(ushort_k0), _f64(ulong_k0)
        nl = r0.size
        if uk.shape != (nl, self.replicas[rep].nk) or sk0.size != nl or lk0.size != nl:
            raise ValueError("long-range table shapes")
        d = [dulong_k, dulong_r0, dushort_k0, dulong_k0]
        if any(x is not None for x in d):
            d = [_f64(x) for x in d]
            if d[0].shape != uk.shape or any(x.size != nl for x in d[1:]):
                raise ValueError("long-range derivative table shapes")
        dp = [_p(x, _f64p) if x is not None else None for x in d]
        self._check("pimc_replica_set_long_range", self.lib.pimc_replica_set_long_range(
            self.ctx, rep, sa, sb, nl, _p(uk, _f64p), _p(r0, _f64p), _p(sk0, _f64p), _p(lk0, _f64p), *dp))

    def energy(self, replicas, mode: int = OLDMODE) -> list:
        reps = _i32(replicas)
        out = (Energy * max(1, reps.size))()
        self._check("pimc_estimator_energy", self.lib.pimc_estimator_energy(
            self.ctx, _p(reps, _i32p), reps.size, mode, out))
        return [out[i] for i in range(reps.size)]

    def d_dbeta(self, action: int, reqs) -> np.ndarray:
        """ActionBaseClass::d_dBeta of one action (ACTION_SHORT_RANGE, ACTION_KINETIC or
        ACTION_LONG_RANGE) for requests (replica, mode, slice1, slice2, level)."""
        reqs = list(reqs)
        arr = (RangeReq * max(1, len(reqs)))()
        for i, (rep, mode, s1, s2, level) in enumerate(reqs):
            arr[i].replica, arr[i].mode, arr[i].slice1, arr[i].slice2, arr[i].level = rep, mode, s1, s2, level
        out = np.zeros(max(1, len(reqs)), dtype=np.float64)
        self._check("pimc_action_d_dbeta", self.lib.pimc_action_d_dbeta(self.ctx, action, arr, len(reqs),
                                                                        _p(out, _f64p)))
        return out[:len(reqs)]

    def pair_correlation(self, rep: int, s1: int, s2: int, start: float, end: float, num_points: int,
                         mode: int = OLDMODE, hist=None) -> np.ndarray:
        h = np.zeros(num_points - 1, dtype=np.int64) if hist is None else hist
        self._check("pimc_estimator_pair_correlation", self.lib.pimc_estimator_pair_correlation(
            self.ctx, rep, mode, s1, s2, start, end, num_points, h.ctypes.data_as(C.POINTER(C.c_int64))))
        return h

    def _path_shape(self, rep: int):
        info = self.replicas[rep]
        return (info.num_time_slices + 1, info.num_particles, 3)

    def path_set(self, rep: int, beads, mode: int = BOTHMODES):
        b = _f64(beads)
        if b.shape != self._path_shape(rep):
            raise ValueError(f"beads must have shape {self._path_shape(rep)}")
        self._check("pimc_path_set", self.lib.pimc_path_set(self.ctx, rep, mode, _p(b, _f64p)))

    def path_get(self, rep: int, mode: int = OLDMODE) -> np.ndarray:
        out = np.empty(self._path_shape(rep), dtype=np.float64)
        self._check("pimc_path_get", self.lib.pimc_path_get(self.ctx, rep, mode, _p(out, _f64p)))
        return out

    def path_set_beads(self, rep: int, mode: int, ptcl: int, slice1: int, xyz):
        x = _f64(xyz).reshape(-1, 3)
        slice2 = slice1 + x.shape[0] - 1
        self._check("pimc_path_set_beads", self.lib.pimc_path_set_beads(
            self.ctx, rep, mode, ptcl, slice1, slice2, _p(x, _f64p)))

    def permutation(self, rep: int, mode: int = OLDMODE) -> np.ndarray:
        out = np.empty(self.replicas[rep].num_particles, dtype=np.int32)
        self._check("pimc_permutation_get", self.lib.pimc_permutation_get(self.ctx, rep, mode, _p(out, _i32p)))
        return out

    def accept_copy(self, rep: int, slice1: int, slice2: int, ptcls):
        p = _i32(ptcls)
        self._check("pimc_accept_copy", self.lib.pimc_accept_copy(self.ctx, rep, slice1, slice2, _p(p, _i32p), p.size))

    def reject_copy(self, rep: int, slice1: int, slice2: int, ptcls):
        p = _i32(ptcls)
        self._check("pimc_reject_copy", self.lib.pimc_reject_copy(self.ctx, rep, slice1, slice2, _p(p, _i32p), p.size))

    def move_join(self, rep: int, old_join: int, new_join: int):
        self._check("pimc_move_join", self.lib.pimc_move_join(self.ctx, rep, old_join, new_join))

    def join(self, rep: int) -> int:
        """The replica's current join slice (PathDataClass::Join)."""
        out = C.c_int32()
        self._check("pimc_replica_join", self.lib.pimc_replica_join(self.ctx, rep, C.byref(out)))
        return out.value

    def put_in_box(self, rep: int, v, fast: bool = False) -> np.ndarray:
        """PathClass::PutInBox (fast=False) or PutInBoxFast (fast=True) of displacements [n, 3]."""
        x = np.array(_f64(v).reshape(-1, 3), dtype=np.float64, order="C", copy=True)
        self._check("pimc_put_in_box", self.lib.pimc_put_in_box(self.ctx, rep, 1 if fast else 0, x.shape[0],
                                                                _p(x, _f64p)))
        return x

    def permute_htable(self, rep: int, species: int, slice1: int, slice2: int, epsilon: float,
                       n: int, mode: int = OLDMODE) -> np.ndarray:
        """PermuteTableClass::ConstructHTable of `species` (n particles): float [n, n]."""
        out = np.zeros((n, n), dtype=np.float32)
        self._check("pimc_permute_htable", self.lib.pimc_permute_htable(
            self.ctx, rep, mode, species, slice1, slice2, epsilon, out.ctypes.data_as(_f32p)))
        return out

    def cycle_apply(self, rep: int, first_ptcl: int, slice_: int, cycle):
        """CycleClass::Apply on the NEW copy at one slice."""
        c = _i32(cycle)
        self._check("pimc_cycle_apply", self.lib.pimc_cycle_apply(self.ctx, rep, first_ptcl, slice_, c.size,
                                                                  _p(c, _i32p)))

    def set_permute(self, rep: int, species: int, gamma=None, epsilon: float = 0.0):
        """PermuteType TABLE for the bisection moves of `species` (gamma: the 4 cycle weights),
        or no permutations (gamma None)."""
        g = None if gamma is None else _f64(gamma)
        if g is not None and g.size != 4:
            raise ValueError("gamma holds the weights of 1- to 4-cycles")
        self._check("pimc_replica_set_permute", self.lib.pimc_replica_set_permute(
            self.ctx, rep, species, _p(g, _f64p) if g is not None else None, epsilon))

    def rho_k_update_all(self, rep: int):
        self._check("pimc_rho_k_update_all", self.lib.pimc_rho_k_update_all(self.ctx, rep))

    def rho_k(self, rep: int, mode: int = OLDMODE) -> np.ndarray:
        info = self.replicas[rep]
        out = np.empty((info.num_time_slices + 1, info.num_species, info.nk, 2), dtype=np.float64)
        self._check("pimc_rho_k_get", self.lib.pimc_rho_k_get(self.ctx, rep, mode, _p(out, _f64p)))
        return out[..., 0] + 1j * out[..., 1]

    # -- actions -----------------------------------------------------------
    @staticmethod
    def make_reqs(reqs) -> C.Array:
        """reqs: iterable of (replica, mode, slice1, slice2, level, ptcls)."""
        reqs = list(reqs)
        arr = (ActionReq * max(1, len(reqs)))()
        for i, (rep, mode, s1, s2, level, ptcls) in enumerate(reqs):
            ptcls = list(ptcls)
            if not 1 <= len(ptcls) <= MAX_ACTIVE:
                raise ValueError("1..4 active particles per request")
            arr[i].replica, arr[i].mode, arr[i].slice1, arr[i].slice2 = rep, mode, s1, s2
            arr[i].level, arr[i].nptcls = level, len(ptcls)
            for j, p in enumerate(ptcls):
                arr[i].ptcls[j] = p
        return arr, len(reqs)

    def action_short_range(self, reqs, precision: int = FP64) -> np.ndarray:
        arr, n = self.make_reqs(reqs)
        out = np.zeros(max(1, n), dtype=np.float64)
        self._check("pimc_action_short_range", self.lib.pimc_action_short_range(
            self.ctx, arr, n, precision, _p(out, _f64p)))
        return out[:n]

    def action_short_range_noalloc(self, reqs, precision: int = FP64) -> np.ndarray:
        """pimc_action_short_range_noalloc: the same results, no device allocation."""
        arr, n = self.make_reqs(reqs)
        out = np.zeros(max(1, n), dtype=np.float64)
        self._check("pimc_action_short_range_noalloc", self.lib.pimc_action_short_range_noalloc(
            self.ctx, arr, n, precision, _p(out, _f64p)))
        return out[:n]

    def action_kinetic(self, reqs) -> np.ndarray:
        arr, n = self.make_reqs(reqs)
        out = np.zeros(max(1, n), dtype=np.float64)
        self._check("pimc_action_kinetic", self.lib.pimc_action_kinetic(self.ctx, arr, n, _p(out, _f64p)))
        return out[:n]

    def action_long_range(self, reqs, precision: int = FP64) -> np.ndarray:
        arr, n = self.make_reqs(reqs)
        out = np.zeros(max(1, n), dtype=np.float64)
        self._check("pimc_action_long_range", self.lib.pimc_action_long_range(
            self.ctx, arr, n, precision, _p(out, _f64p)))
        return out[:n]

    def dist_disp(self, rep: int, mode: int, idx):
        idx = _i32(idx).reshape(-1, 4)
        n = idx.shape[0]
        out = np.empty((n, 8), dtype=np.float64)
        img = np.empty((n, 6), dtype=np.int32)
        self._check("pimc_dist_disp", self.lib.pimc_dist_disp(
            self.ctx, rep, mode, n, _p(idx, _i32p), _p(out, _f64p), _p(img, _i32p)))
        return out, img

    # -- bisection ---------------------------------------------------------
    def bisection_run(self, replicas, num_levels: int, seed: int, first_move: int, num_moves: int,
                      precision: int = FP64, species: int = 0, lowest_level: int = 0,
                      records: bool = True, wait: bool = True):
        """Runs the moves; returns (records or None, stats).  wait=False (records must be
        False) queues the moves and returns (None, None); read the counters with
        bisection_stats_take."""
        reps = _i32(replicas)
        d = BisectionDesc(num_levels=num_levels, lowest_level=lowest_level, species=species,
                          reserved=0, seed=seed)
        if not wait and records:
            raise ValueError("records need wait=True")
        st = BisectionStats() if wait else None
        rec = (MoveRecord * (reps.size * num_moves))() if records and num_moves > 0 else None
        self._check("pimc_bisection_run", self.lib.pimc_bisection_run(
            self.ctx, _p(reps, _i32p), reps.size, C.byref(d), first_move, num_moves, precision,
            rec, C.byref(st) if st is not None else None))
        return (records_to_numpy(rec, reps.size, num_moves) if rec is not None else None), st

    def set_plan(self, block_threads: int = 256, link_split: int = 1, persistent: bool = False, cluster: int = 1):
        p = BisectionPlan(block_threads=block_threads, link_split=link_split, persistent=int(persistent),
                          cluster=cluster, trial_ms=0.0)
        self._check("pimc_bisection_set_plan", self.lib.pimc_bisection_set_plan(self.ctx, C.byref(p)))

    def get_plan(self) -> BisectionPlan:
        p = BisectionPlan()
        self._check("pimc_bisection_get_plan", self.lib.pimc_bisection_get_plan(self.ctx, C.byref(p)))
        return p

    def bisection_tune(self, replicas, num_levels: int, seed: int, first_move: int, trial_moves: int,
                       precision: int = FP64, species: int = 0, lowest_level: int = 0) -> BisectionPlan:
        reps = _i32(replicas)
        d = BisectionDesc(num_levels=num_levels, lowest_level=lowest_level, species=species, reserved=0, seed=seed)
        p = BisectionPlan()
        self._check("pimc_bisection_tune", self.lib.pimc_bisection_tune(
            self.ctx, _p(reps, _i32p), reps.size, C.byref(d), first_move, trial_moves, precision, C.byref(p)))
        return p

    def sweep_graph_create(self, replicas, num_levels: int, seed: int, first_move: int, num_moves: int,
                           precision: int = FP64, species: int = 0, lowest_level: int = 0) -> int:
        reps = _i32(replicas)
        d = BisectionDesc(num_levels=num_levels, lowest_level=lowest_level, species=species, reserved=0, seed=seed)
        gid = C.c_int32()
        self._check("pimc_sweep_graph_create", self.lib.pimc_sweep_graph_create(
            self.ctx, _p(reps, _i32p), reps.size, C.byref(d), first_move, num_moves, precision, C.byref(gid)))
        return gid.value

    def sweep_graph_launch(self, graph: int, num_sweeps: int = 1):
        self._check("pimc_sweep_graph_launch", self.lib.pimc_sweep_graph_launch(self.ctx, graph, num_sweeps))

    def sweep_graph_destroy(self, graph: int):
        self._check("pimc_sweep_graph_destroy", self.lib.pimc_sweep_graph_destroy(self.ctx, graph))

    def bisection_stats_take(self, replicas) -> BisectionStats:
        reps = _i32(replicas)
        st = BisectionStats()
        self._check("pimc_bisection_stats_take", self.lib.pimc_bisection_stats_take(
            self.ctx, _p(reps, _i32p), reps.size, C.byref(st)))
        return st


def records_to_numpy(rec, nrep: int, nmoves: int) -> dict:
    """Converts a MoveRecord array to a dict of numpy arrays shaped [nrep, nmoves, ...]."""
    raw = np.ctypeslib.as_array(C.cast(rec, C.POINTER(C.c_uint8)), shape=(C.sizeof(rec),)).copy()
    dt = np.dtype([("slice1", "<i4"), ("ptcl", "<i4"), ("accepted", "<i4"), ("num_stages", "<i4"),
                   ("old_action", "<f8", (MAX_LEVELS,)), ("new_action", "<f8", (MAX_LEVELS,)),
                   ("log_sample_ratio", "<f8", (MAX_LEVELS,)), ("log_u", "<f8", (MAX_LEVELS,)),
                   ("cycle_length", "<i4"), ("cycle", "<i4", (MAX_ACTIVE,)), ("num_cycles", "<i4"),
                   ("forward_t", "<f8"), ("reverse_t", "<f8"), ("accept_prob", "<f8"), ("psi", "<f8")])
    assert dt.itemsize == C.sizeof(MoveRecord)
    arr = raw.view(dt).reshape(nrep, nmoves)
    return {name: np.array(arr[name]) for name in dt.names}

"""ctypes binding of the p4g C ABI (include/p4g.h) and of the fish and
obstacle drivers (p4pdes_amd/fish/fish.h, p4pdes_amd/obstacle/obstacle.h).  The compute path is the HIP library; there is no
CPU fallback: loading fails loudly if the built libraries are missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")
LIBP4G = os.path.join(LIBDIR, "libp4g.so")
LIBFISH = os.path.join(LIBDIR, "libp4g_fish.so")
LIBOBST = os.path.join(LIBDIR, "libp4g_obstacle.so")
LIBMIN = os.path.join(LIBDIR, "libp4g_minimal.so")

# kernel classes (enum in p4g.h)
K_STENCIL_DOT, K_CG_UPDATE, K_RESID_ZG, K_RESTRICT, K_PROLONG, K_SMOOTH, K_COARSE, \
    K_REDUCE, K_APPLY, K_CG_RESID, K_SMOOTH_ZG, K_SMOOTH_LAST, K_OTHER, K_COUNT = range(14)
KERNEL_NAMES = ["stencil_dot", "cg_update", "resid_zg", "restrict", "prolong", "smooth",
                "coarse", "reduce", "apply", "cg_resid", "smooth_zg", "smooth_last", "other"]

REASONS = {2: "CONVERGED_RTOL", 3: "CONVERGED_ATOL", 4: "CONVERGED_ITS", -3: "DIVERGED_ITS",
           -4: "DIVERGED_DTOL", -5: "DIVERGED_BREAKDOWN", -8: "DIVERGED_INDEFINITE_PC", -10: "DIVERGED_INDEFINITE_MAT",
           -9: "DIVERGED_NANORINF"}


class Report(ctypes.Structure):
    """p4g_Report."""
    _fields_ = [(n, ctypes.c_int) for n in ("its", "reason", "inner_its", "nlevels")] + \
               [(n, ctypes.c_double) for n in ("rnorm0", "rnorm", "true_relres", "fnorm0", "fnorm1",
                                               "setup_ms", "solve_ms", "emax_fine")] + \
               [("history", ctypes.c_double * 64)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_ if n != "history"}
        d["history"] = list(self.history[:max(1, self.its + 1)])
        return d


class FishResult(ctypes.Structure):
    """fish_result (same layout as the oracle's)."""
    _fields_ = [(n, ctypes.c_int) for n in
                ("dim", "mx", "my", "mz", "its", "reason", "inner_its", "nlevels")] + \
               [(n, ctypes.c_double) for n in
                ("fnorm0", "fnorm1", "rnorm0", "rnorm", "true_relres", "errinf", "err2h",
                 "setup_ms", "solve_ms", "emax_fine")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_p4g = None
class VIReport(ctypes.Structure):
    """p4g_VIReport of include/p4g.h."""
    _fields_ = [(n, ctypes.c_int) for n in ("its", "reason", "last_ksp_its", "ksp_its_total")] + \
               [("active", ctypes.c_long)] + \
               [(n, ctypes.c_double) for n in ("fnorm0", "fnorm", "setup_ms", "solve_ms")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ObstacleResult(ctypes.Structure):
    """obstacle_result (same layout as the oracle's)."""
    _fields_ = [(n, ctypes.c_int) for n in
                ("mx", "my", "its", "reason", "last_ksp_its", "ksp_its_total", "stages")] + \
               [("stage_its", ctypes.c_int * 16), ("active", ctypes.c_long)] + \
               [(n, ctypes.c_double) for n in
                ("fnorm0", "fnorm", "error1", "errorinf", "areaerr", "setup_ms", "solve_ms")]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["stage_its"] = list(self.stage_its)[:self.stages]
        return d


class MinimalResult(ctypes.Structure):
    """minimal_result (same layout as the oracle's)."""
    _fields_ = [(n, ctypes.c_int) for n in
                ("mx", "my", "its", "reason", "last_ksp_its", "ksp_its_total", "stages")] + \
               [("stage_its", ctypes.c_int * 16)] + \
               [(n, ctypes.c_double) for n in ("fnorm0", "fnorm", "errinf", "area", "setup_ms", "solve_ms")]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["stage_its"] = list(self.stage_its)[:self.stages]
        return d


_fish = None
_obst = None
_min = None


def lib():
    """libp4g.so (raises if it has not been built)."""
    global _p4g
    if _p4g is None:
        if not os.path.exists(LIBP4G):
            raise RuntimeError("libp4g.so is not built: run __graft_entry__.build() or "
                               "python -m p4pdes_amd.build")
        L = ctypes.CDLL(LIBP4G, mode=ctypes.RTLD_GLOBAL)
        vp, i, d, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
        L.p4g_last_error.restype = cp
        L.p4g_version.restype = cp
        L.p4g_device_malloc.restype = vp
        L.p4g_device_malloc.argtypes = [ctypes.c_size_t]
        L.p4g_device_free.argtypes = [vp]
        L.p4g_memcpy_htod.argtypes = [vp, vp, ctypes.c_size_t]
        L.p4g_memcpy_dtoh.argtypes = [vp, vp, ctypes.c_size_t]
        L.p4g_memcpy_dtod.argtypes = [vp, vp, ctypes.c_size_t]
        L.p4g_da_create.argtypes = [i, i, i, i, d, d, d, ctypes.POINTER(vp)]
        L.p4g_da_set_from_options.argtypes = [vp, cp]
        L.p4g_da_set_cell_coefficient.argtypes = [vp, vp]
        L.p4g_da_destroy.argtypes = [ctypes.POINTER(vp)]
        L.p4g_solver_create.argtypes = [vp, cp, ctypes.POINTER(vp)]
        L.p4g_solver_setup.argtypes = [vp]
        L.p4g_solver_destroy.argtypes = [ctypes.POINTER(vp)]
        L.p4g_ksp_solve.argtypes = [vp, vp, vp, ctypes.POINTER(Report)]
        L.p4g_ksp_solve_device.argtypes = [vp, vp, vp, ctypes.POINTER(Report)]
        L.p4g_pc_apply.argtypes = [vp, vp, vp]
        L.p4g_pc_apply_device.argtypes = [vp, vp, vp]
        L.p4g_pc_apply_device_noalloc.argtypes = [vp, vp, vp]
        L.p4g_mat_mult.argtypes = [vp, vp, vp]
        L.p4g_mat_mult_device.argtypes = [vp, vp, vp]
        L.p4g_solver_get_report.argtypes = [vp, ctypes.POINTER(Report)]
        L.p4g_solver_set_timing.argtypes = [vp, i, i]
        L.p4g_solver_kernel_stats.argtypes = [vp, i, i, ctypes.POINTER(ctypes.c_long),
                                              ctypes.POINTER(d), ctypes.POINTER(d)]
        L.p4g_solver_reset_stats.argtypes = [vp]
        L.p4g_device_count.argtypes = [ctypes.POINTER(i)]
        L.p4g_vec_view_binary.argtypes = [cp, ctypes.c_long, vp, i]
        L.p4g_vec_load_binary.argtypes = [cp, i, ctypes.POINTER(ctypes.c_long), vp]
        L.p4g_da_mat_view_binary.argtypes = [vp, cp, i]
        L.p4g_vi_solve.argtypes = [vp, vp, vp, vp, vp, ctypes.POINTER(VIReport)]
        L.p4g_vi_get_history.argtypes = [vp, vp, vp, vp, i, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.DMDASetUniformCoordinates.argtypes = [vp] + [d] * 6
        _p4g = L
    return _p4g


def fishlib():
    global _fish
    if _fish is None:
        lib()
        if not os.path.exists(LIBFISH):
            raise RuntimeError("libp4g_fish.so is not built")
        L = ctypes.CDLL(LIBFISH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.fish_run.argtypes = [i, vp, ctypes.POINTER(FishResult), vp, ctypes.c_char_p, i]
        L.fish_write_operator.argtypes = [i, vp, ctypes.c_char_p]
        L.fish_session_create.argtypes = [i, vp, ctypes.POINTER(vp)]
        L.fish_session_solve.argtypes = [vp, ctypes.POINTER(Report)]
        L.fish_session_errors.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(d)]
        L.fish_session_solution.argtypes = [vp, vp]
        L.fish_session_solver.argtypes = [vp]
        L.fish_session_solver.restype = vp
        L.fish_session_sizes.argtypes = [vp] + [ctypes.POINTER(i)] * 4
        L.fish_session_rhs_device.argtypes = [vp]
        L.fish_session_rhs_device.restype = vp
        L.fish_session_destroy.argtypes = [ctypes.POINTER(vp)]
        _fish = L
    return _fish


def obstaclelib():
    global _obst
    if _obst is None:
        lib()
        if not os.path.exists(LIBOBST):
            raise RuntimeError("libp4g_obstacle.so is not built")
        L = ctypes.CDLL(LIBOBST)
        L.obstacle_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ObstacleResult),
                                   ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        _obst = L
    return _obst


def minimallib():
    global _min
    if _min is None:
        lib()
        if not os.path.exists(LIBMIN):
            raise RuntimeError("libp4g_minimal.so is not built")
        L = ctypes.CDLL(LIBMIN)
        L.minimal_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(MinimalResult),
                                  ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        _min = L
    return _min


def minimal_run(args, want_u=False):
    """Run minimal on the GPU.  Returns (rc, text, MinimalResult, u or None)."""
    argc, argv = _argv(args)
    res = MinimalResult()
    buf = ctypes.create_string_buffer(1 << 18)
    u = None
    uptr = None
    if want_u:
        mx, my = obstacle_shape(args)
        u = np.zeros(mx * my)
        uptr = u.ctypes.data
    rc = minimallib().minimal_run(argc, argv, ctypes.byref(res), uptr, buf, len(buf))
    return rc, buf.value.decode(), res, u


def obstacle_shape(args):
    """(mx, my) of the final grid obstacle builds for these options."""
    if isinstance(args, str):
        args = args.split()
    def opt(k, dflt):
        return int(args[args.index(k) + 1]) if k in args else dflt
    r = opt("-da_refine", 0) + opt("-snes_grid_sequence", 0)
    return (opt("-da_grid_x", 3) - 1) * 2 ** r + 1, (opt("-da_grid_y", 3) - 1) * 2 ** r + 1


def obstacle_run(args, want_u=False):
    """Run obstacle on the GPU.  Returns (rc, text, ObstacleResult, u or None)."""
    argc, argv = _argv(args)
    res = ObstacleResult()
    buf = ctypes.create_string_buffer(1 << 18)
    u = None
    uptr = None
    if want_u:
        mx, my = obstacle_shape(args)
        u = np.zeros(mx * my)
        uptr = u.ctypes.data
    rc = obstaclelib().obstacle_run(argc, argv, ctypes.byref(res), uptr, buf, len(buf))
    return rc, buf.value.decode(), res, u


def last_error():
    return lib().p4g_last_error().decode()


def device_count():
    n = ctypes.c_int(0)
    lib().p4g_device_count(ctypes.byref(n))
    return n.value


class P4GError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("p4g error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc:
        raise P4GError(rc, last_error())


def _argv(args):
    if isinstance(args, str):
        args = args.split()
    arr = (ctypes.c_char_p * max(1, len(args)))(*[a.encode() for a in args])
    return len(args), arr


def grid_shape(args):
    """(dim, mx, my, mz) of the grid fish builds for these options."""
    if isinstance(args, str):
        args = args.split()

    def opt(k, dflt):
        return args[len(args) - 1 - args[::-1].index(k) + 1] if k in args else dflt
    dim = int(opt("-fsh_dim", 2))
    r = int(opt("-da_refine", 0))
    m = [(int(opt("-da_grid_" + c, 3)) - 1) * 2 ** r + 1 for c in "xyz"]
    if dim < 3:
        m[2] = 1
    if dim < 2:
        m[1] = 1
    return dim, m[0], m[1], m[2]


def vec_view_binary(path, x, append=False):
    """Write x as a PETSc binary Vec (VecView on a binary viewer)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    check(lib().p4g_vec_view_binary(str(path).encode(), x.size, x.ctypes.data, 1 if append else 0))


def vec_load_binary(path, index=0):
    """The index-th object of a PETSc binary file, which must be a Vec."""
    n = ctypes.c_long(0)
    check(lib().p4g_vec_load_binary(str(path).encode(), index, ctypes.byref(n), None))
    x = np.empty(n.value)
    check(lib().p4g_vec_load_binary(str(path).encode(), index, ctypes.byref(n), x.ctypes.data))
    return x


def write_operator(args, path):
    """Set up fish for `args` (host only) and write its operator as a PETSc binary AIJ matrix."""
    argc, argv = _argv(args)
    return fishlib().fish_write_operator(argc, argv, str(path).encode())


def fish_run(args, want_u=False):
    """Run fish on the GPU.  Returns (rc, text, FishResult, u or None)."""
    argc, argv = _argv(args)
    res = FishResult()
    buf = ctypes.create_string_buffer(1 << 16)
    u = None
    uptr = None
    if want_u:
        _, mx, my, mz = grid_shape(args)
        u = np.zeros(mx * my * mz)
        uptr = u.ctypes.data
    rc = fishlib().fish_run(argc, argv, ctypes.byref(res), uptr, buf, len(buf))
    return rc, buf.value.decode(), res, u


class Session:
    """One fish problem resident on the GPU (F(u0) and set-up done once)."""

    def __init__(self, args):
        self._argc, self._argv = _argv(args)
        h = ctypes.c_void_p()
        check(fishlib().fish_session_create(self._argc, self._argv, ctypes.byref(h)))
        self.h = h
        dims = [ctypes.c_int() for _ in range(4)]
        fishlib().fish_session_sizes(h, *[ctypes.byref(x) for x in dims])
        self.dim, self.mx, self.my, self.mz = [x.value for x in dims]
        self.n = self.mx * self.my * self.mz
        self.solver = fishlib().fish_session_solver(h)

    def solve(self):
        rep = Report()
        check(fishlib().fish_session_solve(self.h, ctypes.byref(rep)))
        return rep

    def errors(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(fishlib().fish_session_errors(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def solution(self):
        u = np.zeros(self.n)
        check(fishlib().fish_session_solution(self.h, u.ctypes.data))
        return u

    def report(self):
        rep = Report()
        check(lib().p4g_solver_get_report(self.solver, ctypes.byref(rep)))
        return rep

    def set_timing(self, cls, on=True, fine_only=False):
        """HIP-event timing of a kernel class: every launch, or with
        fine_only only the launches on the finest level."""
        check(lib().p4g_solver_set_timing(self.solver, cls, (2 if fine_only else 1) if on else 0))

    def reset_stats(self):
        check(lib().p4g_solver_reset_stats(self.solver))

    def kernel_stats(self, cls, fine_only=False):
        n, ms, by = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
        check(lib().p4g_solver_kernel_stats(self.solver, cls, 1 if fine_only else 0,
                                            ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)))
        return n.value, ms.value, by.value

    def close(self):
        if getattr(self, "h", None):
            fishlib().fish_session_destroy(ctypes.byref(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """Direct use of the solver ABI on a grid with a cell coefficient or a
    Jacobian callback set through the fish driver (tests use Session)."""

    def __init__(self, session):
        self.s = session.solver
        self.n = session.n

    def pc_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        check(lib().p4g_pc_apply(self.s, r.ctypes.data, z.ctypes.data))
        return z

    def mat_mult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        check(lib().p4g_mat_mult(self.s, x.ctypes.data, y.ctypes.data))
        return y

    def ksp_solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        rep = Report()
        check(lib().p4g_ksp_solve(self.s, b.ctypes.data, x.ctypes.data, ctypes.byref(rep)))
        return x, rep


class ServerStats(ctypes.Structure):
    """p4g_ServerStats."""
    _fields_ = [(n, ctypes.c_long) for n in ("submitted", "completed", "rejected", "active", "queued",
                                             "rounds", "graph_captures")] + \
               [(n, ctypes.c_double) for n in ("p50_ms", "p99_ms", "max_ms", "mean_ms", "queue_p50_ms",
                                               "wall_s", "dof_per_s", "peak_bytes")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def _server_sigs(L):
    vp = ctypes.c_void_p
    L.p4g_server_create.argtypes = [ctypes.c_size_t, ctypes.c_char_p, ctypes.POINTER(vp)]
    L.p4g_server_request_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    L.p4g_server_submit.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ctypes.c_long)]
    L.p4g_server_poll.argtypes = [vp, ctypes.c_long, ctypes.POINTER(ctypes.c_int)]
    L.p4g_server_wait.argtypes = [vp, ctypes.c_long, ctypes.POINTER(Report)]
    L.p4g_server_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.p4g_server_destroy.argtypes = [ctypes.POINTER(vp)]


class Server:
    """Continuous-batching solve server (p4g_server_*).  Requests reference a
    set-up solver (e.g. a Session's) and a host right-hand side."""

    def __init__(self, hbm_budget_bytes, options=""):
        L = lib()
        _server_sigs(L)
        h = ctypes.c_void_p()
        check(L.p4g_server_create(int(hbm_budget_bytes), options.encode(), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    @staticmethod
    def request_bytes(solver):
        L = lib()
        _server_sigs(L)
        b = ctypes.c_size_t()
        check(L.p4g_server_request_bytes(solver, ctypes.byref(b)))
        return b.value

    def submit(self, solver, b, x=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        if x is None:
            x = np.zeros_like(b)
        i = ctypes.c_long()
        check(lib().p4g_server_submit(self.h, solver, b.ctypes.data, x.ctypes.data, ctypes.byref(i)))
        self._keep[i.value] = (b, x)
        return i.value

    def poll(self, rid):
        s = ctypes.c_int()
        check(lib().p4g_server_poll(self.h, rid, ctypes.byref(s)))
        return s.value

    def wait(self, rid):
        rep = Report()
        check(lib().p4g_server_wait(self.h, rid, ctypes.byref(rep)))
        b, x = self._keep.pop(rid)
        return x, rep

    def stats(self):
        st = ServerStats()
        check(lib().p4g_server_stats(self.h, ctypes.byref(st)))
        return st

    def close(self):
        if getattr(self, "h", None):
            lib().p4g_server_destroy(ctypes.byref(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

"""CPU ORACLE -- test infrastructure only.

ctypes front end of oracle/_build/liboracle.so (the C restatement of the
reference's fish/poissonfunctions path and PETSc solver semantics).  Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class FishResult(ctypes.Structure):
    """Mirror of fish_result in oracle/oracle.h."""
    _fields_ = [(n, ctypes.c_int) for n in
                ("dim", "mx", "my", "mz", "its", "reason", "inner_its", "nlevels")] + \
               [(n, ctypes.c_double) for n in
                ("fnorm0", "fnorm1", "rnorm0", "rnorm", "true_relres", "errinf", "err2h",
                 "setup_ms", "solve_ms", "emax_fine")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_fish_run.restype = ctypes.c_int
        _lib.oracle_fish_time.restype = ctypes.c_int
        _lib.oracle_noisy.restype = ctypes.c_double
        _lib.oracle_noisy.argtypes = [ctypes.c_int64]
        _lib.oracle_normal.restype = ctypes.c_double
        _lib.oracle_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        _lib.oracle_varcoef_k.restype = ctypes.c_double
        _lib.oracle_varcoef_k.argtypes = [ctypes.c_double] * 3 + [ctypes.c_uint64]
    return _lib


def _argv(args):
    if isinstance(args, str):
        args = args.split()
    arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    return len(args), arr


def grid_shape(args):
    """(dim, mx, my, mz) that fish builds for these options."""
    if isinstance(args, str):
        args = args.split()
    def opt(k, d):
        return args[args.index(k) + 1] if k in args else d
    dim = int(opt("-fsh_dim", 2))
    r = int(opt("-da_refine", 0))
    m = [int(opt("-da_grid_" + c, 3)) for c in "xyz"]
    m = [(q - 1) * 2 ** r + 1 for q in m]
    if dim < 3:
        m[2] = 1
    if dim < 2:
        m[1] = 1
    return dim, m[0], m[1], m[2]


def fish_run(args, want_u=False):
    """Run the oracle as fish would run with `args`.

    Returns (rc, text, FishResult, u or None)."""
    argc, argv = _argv(args)
    res = FishResult()
    buf = ctypes.create_string_buffer(1 << 16)
    u = None
    uptr = None
    if want_u:
        _, mx, my, mz = grid_shape(args)
        u = np.zeros(mx * my * mz)
        uptr = u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = lib().oracle_fish_run(argc, argv, ctypes.byref(res), uptr, buf, len(buf))
    return rc, buf.value.decode(), res, u


def fish_time(args, nthreads=1, warmup=1, nsolves=3):
    """Median-ready list of solve times (ms) of the linear solve fish builds."""
    argc, argv = _argv(args)
    times = (ctypes.c_double * nsolves)()
    res = FishResult()
    rc = lib().oracle_fish_time(argc, argv, nthreads, warmup, nsolves, times, ctypes.byref(res))
    if rc:
        raise RuntimeError("oracle_fish_time failed: %d" % rc)
    return list(times), res


def build_native(outdir, timeout=300):
    """Build the CPU-baseline copy of the oracle for this machine's
    instruction set (-march=native) into outdir; returns its path."""
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native", "NATIVE=" + outdir],
                   check=True, timeout=timeout)
    return os.path.join(outdir, "liboracle.so")


def fish_time_sample(args, nthreads, warmup, nsolves, sample_its, libpath=None):
    """oracle_fish_time_sample: solve times (ms) on nthreads threads, the
    set-up time, and one single-thread solve capped at sample_its iterations.
    libpath selects another build of the oracle (e.g. build_native's)."""
    L = ctypes.CDLL(libpath) if libpath else lib()
    f = L.oracle_fish_time_sample
    f.restype = ctypes.c_int
    argc, argv = _argv(args)
    times = (ctypes.c_double * nsolves)()
    res = FishResult()
    setup = ctypes.c_double()
    sms = ctypes.c_double()
    sdone = ctypes.c_int()
    rc = f(argc, argv, nthreads, warmup, nsolves, times, ctypes.byref(res), ctypes.byref(setup),
           sample_its, ctypes.byref(sms), ctypes.byref(sdone))
    if rc:
        raise RuntimeError("oracle_fish_time_sample failed: %d" % rc)
    return {"times_ms": list(times), "res": res, "setup_ms": setup.value,
            "sample_ms": sms.value, "sample_its": sdone.value}


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def poisson_apply(dim, mx, my, mz, x, L=(1.0, 1.0, 1.0), c=(1.0, 1.0, 1.0)):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    f = lib().oracle_poisson_apply
    f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 6 + [ctypes.c_void_p] * 2
    f(dim, mx, my, mz, *L, *c, x.ctypes.data, y.ctypes.data)
    return y


def interp_add(dim, mxc, myc, mzc, xc, xf):
    f = lib().oracle_interp_add
    f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 2
    f(dim, mxc, myc, mzc, xc.ctypes.data, xf.ctypes.data)
    return xf


def restrict(dim, mxc, myc, mzc, rf):
    bc = np.zeros(mxc * max(myc, 1) * max(mzc, 1))
    f = lib().oracle_restrict
    f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 2
    f(dim, mxc, myc, mzc, rf.ctypes.data, bc.ctypes.data)
    return bc


def varcoef_apply(mx, my, mz, kcell, x, L=(1.0, 1.0, 1.0)):
    y = np.zeros_like(x)
    f = lib().oracle_varcoef_apply
    f.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_void_p] * 3
    f(mx, my, mz, *L, kcell.ctypes.data, x.ctypes.data, y.ctypes.data)
    return y


def _vec_call(fname, args, x):
    argc, argv = _argv(args)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    f = getattr(lib(), fname)
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    rc = f(argc, argv, x.ctypes.data, y.ctypes.data)
    if rc:
        raise RuntimeError("%s failed: %d" % (fname, rc))
    return y


def pc_apply(args, r):
    """z = B r with the preconditioner fish's KSP uses for `args`."""
    return _vec_call("oracle_pc_apply", args, r)


def mat_mult(args, x):
    """y = A x with the operator fish assembles for `args`."""
    return _vec_call("oracle_mat_mult", args, x)


def rhs(args):
    """F(u0): the right-hand side of fish's linear solve."""
    argc, argv = _argv(args)
    _, mx, my, mz = grid_shape(args)
    F = np.zeros(mx * my * mz)
    f = lib().oracle_rhs
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rc = f(argc, argv, F.ctypes.data)
    if rc:
        raise RuntimeError("oracle_rhs failed: %d" % rc)
    return F


class ObstacleResult(ctypes.Structure):
    """Mirror of obstacle_result in oracle/oracle.h."""
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
    """Mirror of minimal_result in oracle/oracle.h."""
    _fields_ = [(n, ctypes.c_int) for n in
                ("mx", "my", "its", "reason", "last_ksp_its", "ksp_its_total", "stages")] + \
               [("stage_its", ctypes.c_int * 16)] + \
               [(n, ctypes.c_double) for n in ("fnorm0", "fnorm", "errinf", "area", "setup_ms", "solve_ms")]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["stage_its"] = list(self.stage_its)[:self.stages]
        return d


def minimal_run(args, want_u=False):
    """Run the oracle as c/ch7/minimal.c would run with `args`.

    Returns (rc, text, MinimalResult, u or None)."""
    argc, argv = _argv(args)
    res = MinimalResult()
    buf = ctypes.create_string_buffer(1 << 18)
    u = None
    uptr = None
    if want_u:
        mx, my = obstacle_shape(args)
        u = np.zeros(mx * my)
        uptr = u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = lib().oracle_minimal_run(argc, argv, ctypes.byref(res), uptr, buf, len(buf))
    return rc, buf.value.decode(), res, u


def minimal_residual(mx, my, problem, q, H, c, u):
    """F(u) of c/ch7/minimal.c (problem 0 tent, 1 catenoid)."""
    F = np.zeros(mx * my)
    dp = ctypes.POINTER(ctypes.c_double)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lib().oracle_minimal_residual(mx, my, problem, ctypes.c_double(q), ctypes.c_double(H), ctypes.c_double(c),
                                  u.ctypes.data_as(dp), F.ctypes.data_as(dp))
    return F


def minimal_fdjac(mx, my, problem, q, H, c, u):
    """The colored finite-difference Jacobian, shape (mx*my, 9)."""
    J = np.zeros((mx * my, 9))
    dp = ctypes.POINTER(ctypes.c_double)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lib().oracle_minimal_fdjac(mx, my, problem, ctypes.c_double(q), ctypes.c_double(H), ctypes.c_double(c),
                               u.ctypes.data_as(dp), J.ctypes.data_as(dp))
    return J


def minimal_mffd(mx, my, problem, q, H, c, u, a):
    """y = J(u) a by -snes_mf_operator's matrix-free finite difference."""
    y = np.zeros(mx * my)
    dp = ctypes.POINTER(ctypes.c_double)
    u = np.ascontiguousarray(u, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    lib().oracle_minimal_mffd(mx, my, problem, ctypes.c_double(q), ctypes.c_double(H), ctypes.c_double(c),
                              u.ctypes.data_as(dp), a.ctypes.data_as(dp), y.ctypes.data_as(dp))
    return y


def obstacle_shape(args):
    """(mx, my) of the final grid obstacle builds for these options."""
    if isinstance(args, str):
        args = args.split()
    def opt(k, d):
        return int(args[args.index(k) + 1]) if k in args else d
    r = opt("-da_refine", 0) + opt("-snes_grid_sequence", 0)
    return (opt("-da_grid_x", 3) - 1) * 2 ** r + 1, (opt("-da_grid_y", 3) - 1) * 2 ** r + 1


def obstacle_run(args, want_u=False):
    """Run the oracle as c/ch12/obstacle.c would run with `args`.

    Returns (rc, text, ObstacleResult, u or None)."""
    argc, argv = _argv(args)
    res = ObstacleResult()
    buf = ctypes.create_string_buffer(1 << 18)
    u = None
    uptr = None
    if want_u:
        mx, my = obstacle_shape(args)
        u = np.zeros(mx * my)
        uptr = u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = lib().oracle_obstacle_run(argc, argv, ctypes.byref(res), uptr, buf, len(buf))
    return rc, buf.value.decode(), res, u

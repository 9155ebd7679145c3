"""Python face of the CPU oracle (fdtd_oracle.c) and the numpy restatements.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (virs_amd/) never
imports it.

Reference constants are computed here in float32 the way
src/simulator/simulator_gpu.cu:195-198 and :147 compute them, so the oracle
and the kernels run one discrete scheme.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VIRS_ORACLE_SO selects another build of the oracle (the sanitized one)
_SO = os.environ.get("VIRS_ORACLE_SO") or os.path.join(_HERE, "_build", "libvirs_oracle.so")
_lib = None

f32 = np.float32


class OracleDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("c2dt2", C.c_float), ("ih2", C.c_float), ("kconst", C.c_float),
        ("flags", C.c_void_p),
        ("sponge_w", C.c_int32),
        ("spg_a", C.c_void_p), ("spg_b", C.c_void_p), ("spg_g", C.c_void_p),
        ("nb", C.c_int64), ("bidx", C.c_void_p), ("bmat", C.c_void_p),
        ("nmat", C.c_int32), ("coeffs", C.c_void_p), ("state_f32", C.c_int32),
        ("hard_sig", C.c_void_p), ("hard_len", C.c_int64),
        ("nsoft", C.c_int32), ("soft_idx", C.c_void_p), ("soft_sig", C.c_void_p),
        ("soft_len", C.c_int64),
        ("nrcv", C.c_int32), ("rcv_idx", C.c_void_p),
        ("init_p", C.c_void_p), ("init_q", C.c_void_p),
        ("state_io", C.c_void_p),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = C.CDLL(_SO)
        dp = C.POINTER(C.c_double)
        L.oracle_run_f32.restype = C.c_int
        L.oracle_run_f32.argtypes = [C.POINTER(OracleDesc), C.c_int64, C.c_int64, dp, dp, dp, dp, dp]
        L.oracle_run_f64.restype = C.c_int
        L.oracle_run_f64.argtypes = L.oracle_run_f32.argtypes
        L.oracle_adjoint_f64.restype = C.c_int
        L.oracle_adjoint_f64.argtypes = [C.POINTER(OracleDesc), C.c_int64, dp, dp]
        L.oracle_rms_continue.restype = C.c_int
        L.oracle_rms_continue.argtypes = [C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_float),
                                          C.POINTER(C.c_float), C.c_float]
        fpp = C.POINTER(C.POINTER(C.c_float))
        L.oracle_refcpu_steps.restype = None
        L.oracle_refcpu_steps.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_float,
                                          C.c_int32, fpp, fpp, fpp]
        L.oracle_run_s16.restype = C.c_int
        L.oracle_run_s16.argtypes = [C.POINTER(OracleDesc), C.c_int, C.c_int, C.c_int64, dp, dp, dp,
                                     C.POINTER(C.c_int32)]
        L.oracle_round16.restype = None
        L.oracle_round16.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_set_threads.restype = None
        L.oracle_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def set_threads(n):
    lib().oracle_set_threads(int(n))


def reference_constants(h, c=343.0, dt=0.0):
    """(dt, c2dt2, ih2, kconst) in float32 as simulator_gpu.cu:195-198,147."""
    h, c = f32(h), f32(c)
    dt = f32(dt) if dt else f32(0.5) * h / (c * np.sqrt(f32(3.0)))
    c2dt2 = c * c * dt * dt
    ih2 = f32(1.0) / (h * h)
    k = (c * dt) / h * f32(0.95)
    return dt, c2dt2, ih2, k


def dt_for_lambda2(lam2, h, c=343.0):
    """Time step giving Courant number^2 = lam2 (float32)."""
    return f32(np.sqrt(lam2) * float(h) / float(c))


def sponge_tables(width, sigma):
    """(a, b, g) per distance to the domain edge; the centred damping update
    R = (2P - (1-s)Q + L)/(1+s) with s = sigma ((W-d)/W)^2 / 2."""
    a = np.empty(width, f32)
    b = np.empty(width, f32)
    g = np.empty(width, f32)
    sigma = float(np.float32(sigma))  # the library takes sigma as a float
    for k in range(width):
        t = (width - k) / width
        s = 0.5 * sigma * t * t
        a[k] = 2.0 / (1.0 + s)
        b[k] = (1.0 - s) / (1.0 + s)
        g[k] = 1.0 / (1.0 + s)
    return a, b, g


class Problem:
    """Everything one oracle run needs, held as numpy arrays."""

    def __init__(self, flags, h, c=343.0, dt=0.0, bidx=None, bmat=None, coeffs=None,
                 state_f32=False, hard_sig=None, soft_idx=(), soft_sig=None, rcv_idx=(),
                 sponge_width=0, sponge_sigma=0.0, init_p=None, init_q=None):
        self.flags = np.ascontiguousarray(flags, dtype=np.uint8)
        assert self.flags.ndim == 3
        self.nz, self.ny, self.nx = self.flags.shape
        self.dt, self.c2dt2, self.ih2, self.kconst = reference_constants(h, c, dt)
        self.bidx = np.ascontiguousarray(np.zeros(0, np.uint32) if bidx is None else bidx, np.uint32)
        self.bmat = None if bmat is None else np.ascontiguousarray(bmat, np.uint8)
        self.coeffs = np.ascontiguousarray(
            np.zeros((1, 5)) if coeffs is None else np.asarray(coeffs, np.float64).reshape(-1, 5))
        self.state_f32 = int(bool(state_f32))
        self.hard_sig = None if hard_sig is None else np.ascontiguousarray(hard_sig, np.float64)
        self.soft_idx = np.ascontiguousarray(np.asarray(soft_idx, np.int64).reshape(-1))
        ns = len(self.soft_idx)
        self.soft_sig = np.ascontiguousarray(
            np.zeros((ns, 1)) if soft_sig is None else np.asarray(soft_sig, np.float64).reshape(ns, -1))
        self.rcv_idx = np.ascontiguousarray(np.asarray(rcv_idx, np.int64).reshape(-1))
        self.sponge_width = int(sponge_width)
        self.spg = sponge_tables(self.sponge_width, sponge_sigma) if sponge_width else (
            np.zeros(1, f32), np.zeros(1, f32), np.zeros(1, f32))
        self.init_p = None if init_p is None else np.ascontiguousarray(init_p, np.float64)
        self.init_q = None if init_q is None else np.ascontiguousarray(init_q, np.float64)
        # optional [nb*4] filter state carried between chunked FP32/FP64 runs
        self.state_io = None

    def desc(self):
        d = OracleDesc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.c2dt2, d.ih2, d.kconst = float(self.c2dt2), float(self.ih2), float(self.kconst)
        d.flags = self.flags.ctypes.data
        d.sponge_w = self.sponge_width
        d.spg_a, d.spg_b, d.spg_g = (x.ctypes.data for x in self.spg)
        d.nb = len(self.bidx)
        d.bidx = self.bidx.ctypes.data
        d.bmat = None if self.bmat is None else self.bmat.ctypes.data
        d.nmat = len(self.coeffs)
        d.coeffs = self.coeffs.ctypes.data
        d.state_f32 = self.state_f32
        d.hard_sig = None if self.hard_sig is None else self.hard_sig.ctypes.data
        d.hard_len = 0 if self.hard_sig is None else len(self.hard_sig)
        d.nsoft = len(self.soft_idx)
        d.soft_idx = self.soft_idx.ctypes.data
        d.soft_sig = self.soft_sig.ctypes.data
        d.soft_len = self.soft_sig.shape[1]
        d.nrcv = len(self.rcv_idx)
        d.rcv_idx = self.rcv_idx.ctypes.data
        d.init_p = None if self.init_p is None else self.init_p.ctypes.data
        d.init_q = None if self.init_q is None else self.init_q.ctypes.data
        d.state_io = None if self.state_io is None else self.state_io.ctypes.data
        self._keep = d
        return d

    def run(self, nsteps, precision="f32", n0=0, want_fields=False, want_trace=False):
        """Returns dict(ir=[nrcv][nsteps] float64, p=..., q=..., trace_d, trace_y)."""
        L = lib()
        d = self.desc()
        ir = np.zeros((max(1, len(self.rcv_idx)), nsteps), np.float64)
        N = self.nx * self.ny * self.nz
        fp = np.zeros(N) if want_fields else None
        fq = np.zeros(N) if want_fields else None
        nb = len(self.bidx)
        td = np.zeros((nsteps, max(1, nb))) if want_trace else None
        ty = np.zeros((nsteps, max(1, nb))) if want_trace else None
        dp = C.POINTER(C.c_double)
        ptr = lambda a: None if a is None else a.ctypes.data_as(dp)
        fn = L.oracle_run_f32 if precision == "f32" else L.oracle_run_f64
        rc = fn(C.byref(d), n0, nsteps, ptr(ir), ptr(fp), ptr(fq), ptr(td), ptr(ty))
        if rc != 0:
            raise RuntimeError(f"oracle run failed: {rc}")
        out = {"ir": ir[: len(self.rcv_idx)]}
        if want_fields:
            out["p"] = fp.reshape(self.nz, self.ny, self.nx)
            out["q"] = fq.reshape(self.nz, self.ny, self.nx)
        if want_trace:
            out["trace_d"], out["trace_y"] = td, ty
        return out

    def run16(self, nsteps, fmt="fp16", stochastic=True, want_fields=False):
        """16-bit storage run (scaled FP16 or BF16 fields, FP32 arithmetic),
        the kernels' storage rules restated; from zero fields."""
        L = lib()
        d = self.desc()
        ir = np.zeros((max(1, len(self.rcv_idx)), nsteps), np.float64)
        N = self.nx * self.ny * self.nz
        fp = np.zeros(N) if want_fields else None
        fq = np.zeros(N) if want_fields else None
        ex = np.zeros(nsteps, np.int32)
        dp = C.POINTER(C.c_double)
        ptr = lambda a: None if a is None else a.ctypes.data_as(dp)
        rc = L.oracle_run_s16(C.byref(d), int(fmt == "bf16"), int(bool(stochastic)), nsteps, ptr(ir),
                              ptr(fp), ptr(fq), ex.ctypes.data_as(C.POINTER(C.c_int32)))
        if rc != 0:
            raise RuntimeError(f"oracle 16-bit run failed: {rc}")
        out = {"ir": ir[: len(self.rcv_idx)], "exp": ex}
        if want_fields:
            out["p"] = fp.reshape(self.nz, self.ny, self.nx)
            out["q"] = fq.reshape(self.nz, self.ny, self.nx)
        return out

    def adjoint(self, nsteps, dl_dir):
        L = lib()
        d = self.desc()
        g = np.zeros((len(self.coeffs), 5))
        dl = np.ascontiguousarray(np.asarray(dl_dir, np.float64).reshape(len(self.rcv_idx), nsteps))
        dp = C.POINTER(C.c_double)
        rc = L.oracle_adjoint_f64(C.byref(d), nsteps, dl.ctypes.data_as(dp), g.ctypes.data_as(dp))
        if rc != 0:
            raise RuntimeError(f"oracle adjoint failed: {rc}")
        return g


def round16(x, fmt="fp16", rbits=None):
    """Storage rounding of the 16-bit runner: (uint16 patterns, float32 values).
    rbits (16 random bits per value) selects stochastic rounding."""
    L = lib()
    x = np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))
    out = np.zeros(x.size, np.uint16)
    back = np.zeros(x.size, np.float32)
    rb = None
    if rbits is not None:
        rb = np.ascontiguousarray(np.broadcast_to(np.asarray(rbits, np.uint32), x.shape))
    L.oracle_round16(x.ctypes.data, None if rb is None else rb.ctypes.data, x.size,
                     int(fmt == "bf16"), out.ctypes.data, back.ctypes.data)
    return out, back


def rms_continue_sequence(ir, target=0.0009):
    """Applies the RMS stop rule step by step; returns the first stopping step or -1."""
    L = lib()
    ir = np.ascontiguousarray(ir, np.float32)
    st = C.c_float(0.0)
    pc = C.c_float(0.0)
    p = ir.ctypes.data_as(C.POINTER(C.c_float))
    for s in range(len(ir)):
        if not L.oracle_rms_continue(p, s, C.byref(st), C.byref(pc), target):
            return s
    return -1


def layer_image(p_curr, flags, layer):
    """VectorSpace::layerToImage (src/simulator/vectorspace.cpp:77-148) as the
    reference writes it: loops over the padded (even) width and height, the
    padded cells written black at their padded-stride position, the others at
    k*width + i.  Pure-Python loops in float32, for small grids.  Returns the
    byte buffer as uint8 [H, W, 3] and raises IndexError for a layer outside
    [0, Ny) (the reference throws std::out_of_range, :79-80)."""
    flags = np.asarray(flags, np.uint8)
    nz, ny, nx = flags.shape
    if not 0 <= layer < ny:
        raise IndexError("layerToImage: layer index out of range")
    p = np.asarray(p_curr, np.float32).reshape(-1)
    fl = flags.reshape(-1)
    width, height = nx, nz
    ow = width + 1 if width % 2 else width
    oh = height + 1 if height % 2 else height
    scale = f32(1500.0)
    px = [0] * (ow * oh * 3)

    def clamp_byte(v):  # std::clamp(v, 0, 255) then the unsigned char conversion
        v = f32(v)
        v = f32(0) if v < 0 else (f32(255) if v > 255 else v)
        return int(v)

    for k in range(oh):
        row = k * nx * ny + layer * nx
        for i in range(ow):
            if i >= width or k >= height:
                px[(k * ow + i) * 3: (k * ow + i) * 3 + 3] = [0, 0, 0]
                continue
            v = p[row + i]
            flag = int(fl[row + i])
            red = green = blue = 0
            if flag == 1:
                green = 100
            if flag == 2:
                green = 255
            if v >= f32(0):
                red = clamp_byte(v * scale)
            else:
                blue = clamp_byte(-v * scale)
            j = (k * width + i) * 3
            px[j: j + 3] = [red, green, blue]
    return np.array(px, np.uint8).reshape(oh, ow, 3)


def refcpu_steps(flags, h, nsteps, p=None, q=None):
    """The reference's own CPU step (simulator_cpu.cpp:55-81) run nsteps."""
    L = lib()
    flags = np.ascontiguousarray(flags, np.uint8)
    nz, ny, nx = flags.shape
    bufs = [np.zeros(flags.size, np.float32) for _ in range(3)]
    if p is not None:
        bufs[0][:] = np.asarray(p, np.float32).reshape(-1)
    if q is not None:
        bufs[1][:] = np.asarray(q, np.float32).reshape(-1)
    fptr = C.POINTER(C.c_float)
    ptrs = [fptr(C.c_float.from_address(b.ctypes.data)) for b in bufs]
    P, Q, R = (C.pointer(x) for x in ptrs)
    addr = {b.ctypes.data: b for b in bufs}
    L.oracle_refcpu_steps(nx, ny, nz, flags.ctypes.data, float(h), int(nsteps), P, Q, R)
    newest = addr[C.cast(P.contents, C.c_void_p).value]
    return newest.reshape(nz, ny, nx)

"""Python host mirror of the reference's FDTD backend, over libvirs_hip.so.

The methods follow the reference's dispatch layer
(src/simulator/simulator_dispatch.h:4-25) one to one:

    fdtd_setup                 -> FDTD(...)
    updateGPUFromGrid          -> upload_flags / write_field
    uploadNormalsToGPU         -> upload_normals
    uploadBoundaryIndicesToGPU -> upload_boundary (with allocFilterStates)
    allocFilterCoeffs +
    uploadFilterCoeffsToGPU    -> set_materials
    fdtd_start_simulation      -> start
    fdtd_step                  -> step (n steps per call)
    updateCurrentGridFromGPU   -> field
    fdtd_cleanup               -> close

Errors raise FDTDError with the library's message; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _lib

PRECISIONS = {"fp32": 0, "fp16": 1, "bf16": 2, "fp64": 3}
BYTES = {"fp32": 4, "fp16": 2, "bf16": 2, "fp64": 8}


class FDTDError(RuntimeError):
    pass


def dt_for_lambda2(lam2, h, c=343.0):
    """Time step (float32) for Courant number^2 = lam2."""
    return np.float32(np.sqrt(lam2) * float(np.float32(h)) / float(c))


class FDTD:
    def __init__(self, dims, h, dt=0.0, lam2=None, c=343.0, precision="fp32", sponge_width=0,
                 sponge_sigma=0.0, device=0, use_graph=False, dense_blocks=0, fuse_boundary=False,
                 rounding="stochastic", vector_width=0):
        self.L = _lib.hip()
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {list(PRECISIONS)}")
        d = _lib.VirsDesc()
        self.L.virs_desc_default(C.byref(d))
        d.nx, d.ny, d.nz = (int(v) for v in dims)
        d.h = float(h)
        d.c = float(c)
        if lam2 is not None:
            dt = dt_for_lambda2(lam2, h, c)
        d.dt = float(dt)
        d.precision = PRECISIONS[precision]
        d.sponge_width = int(sponge_width)
        d.sponge_sigma = float(sponge_sigma)
        d.device = int(device)
        d.use_graph = int(bool(use_graph))
        d.dense_blocks = int(dense_blocks)
        d.fuse_boundary = int(bool(fuse_boundary))
        if rounding not in ("nearest", "stochastic"):
            raise ValueError("rounding must be 'nearest' or 'stochastic'")
        d.rounding = 1 if rounding == "stochastic" else 0
        d.vector_width = int(vector_width)
        self.desc = d
        self.dims = tuple(int(v) for v in dims)
        self.n = self.dims[0] * self.dims[1] * self.dims[2]
        self.precision = precision
        self.nrcv = 0
        self.nmat = 0
        self._ctx = C.c_void_p()
        rc = self.L.virs_create(C.byref(d), C.byref(self._ctx))
        if rc != 0:
            self._ctx = None
            raise FDTDError(f"virs_create failed ({rc})")

    # -- plumbing ---------------------------------------------------------
    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.virs_last_error(self._ctx)
            raise FDTDError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if self._ctx:
            self.L.virs_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def constants(self):
        """(dt, c2dt2, ih2, kconst) as the kernels use them (float32)."""
        v = [C.c_float() for _ in range(4)]
        self._chk(self.L.virs_get_constants(self._ctx, *(C.byref(x) for x in v)), "constants")
        return tuple(np.float32(x.value) for x in v)

    # -- grid -------------------------------------------------------------
    def upload_flags(self, flags):
        f = np.ascontiguousarray(flags, np.uint8).reshape(-1)
        if f.size != self.n:
            raise ValueError("flags size does not match the grid")
        self._chk(self.L.virs_upload_flags(self._ctx, f.ctypes.data_as(_lib.c_u8p)), "upload flags")

    def upload_normals(self, normals):
        f = np.ascontiguousarray(normals, np.uint8).reshape(-1)
        if f.size != self.n:
            raise ValueError("normals size does not match the grid")
        self._chk(self.L.virs_upload_normals(self._ctx, f.ctypes.data_as(_lib.c_u8p)), "upload normals")

    def upload_boundary(self, bidx, bmat=None):
        b = np.ascontiguousarray(bidx, np.uint32)
        m = None if bmat is None else np.ascontiguousarray(bmat, np.uint8)
        self._chk(self.L.virs_upload_boundary(self._ctx, b.ctypes.data_as(_lib.c_u32p),
                                              None if m is None else m.ctypes.data_as(_lib.c_u8p),
                                              len(b)), "upload boundary")

    def set_materials(self, coeffs):
        c = np.ascontiguousarray(np.asarray(coeffs, np.float64).reshape(-1, 5))
        self._chk(self.L.virs_set_materials(self._ctx, c.ctypes.data_as(_lib.c_dp), len(c)), "materials")
        self.nmat = len(c)

    def upload_room(self, room, with_normals=True):
        self.upload_flags(room.flags)
        if with_normals:
            self.upload_normals(room.normals)
        self.upload_boundary(room.bidx, room.bmat)

    # -- sources / receivers ---------------------------------------------
    def set_hard_source(self, sig=None):
        if sig is None:
            self._chk(self.L.virs_set_hard_source(self._ctx, None, 0), "hard source")
        else:
            s = np.ascontiguousarray(sig, np.float64)
            self._chk(self.L.virs_set_hard_source(self._ctx, s.ctypes.data_as(_lib.c_dp), len(s)), "hard source")

    def set_soft_sources(self, idx, sig):
        i = np.ascontiguousarray(np.asarray(idx, np.int64).reshape(-1))
        s = np.ascontiguousarray(np.asarray(sig, np.float64).reshape(len(i), -1)) if len(i) else np.zeros((0, 1))
        self._chk(self.L.virs_set_soft_sources(self._ctx, i.ctypes.data_as(_lib.c_i64p),
                                               s.ctypes.data_as(_lib.c_dp), len(i), s.shape[1]), "soft sources")

    def set_receivers(self, idx):
        i = np.ascontiguousarray(np.asarray(idx, np.int64).reshape(-1))
        self._chk(self.L.virs_set_receivers(self._ctx, i.ctypes.data_as(_lib.c_i64p), len(i)), "receivers")
        self.nrcv = len(i)

    def set_early_stop(self, enable=True, target_rms=0.0009):
        self._chk(self.L.virs_set_early_stop(self._ctx, int(enable), float(target_rms)), "early stop")

    # -- time loop ----------------------------------------------------------
    def start(self, nsteps):
        self._chk(self.L.virs_start_simulation(self._ctx, int(nsteps)), "start")

    def step(self, nsteps=1):
        """Runs nsteps; returns False once the RMS stop fired."""
        kg = C.c_int32(1)
        self._chk(self.L.virs_step(self._ctx, int(nsteps), C.byref(kg)), "step")
        return bool(kg.value)

    def synchronize(self):
        self._chk(self.L.virs_synchronize(self._ctx), "synchronize")

    @property
    def steps_done(self):
        return int(self.L.virs_steps_done(self._ctx))

    @property
    def percent_to_target_rms(self):
        return float(self.L.virs_percent_to_target_rms(self._ctx))

    def receivers(self, f64=False):
        k = self.steps_done
        if f64:
            out = np.zeros((self.nrcv, k), np.float64)
            self._chk(self.L.virs_read_receivers_f64(self._ctx, out.ctypes.data_as(_lib.c_dp)), "read receivers")
        else:
            out = np.zeros((self.nrcv, k), np.float32)
            self._chk(self.L.virs_read_receivers(self._ctx, out.ctypes.data_as(_lib.c_fp)), "read receivers")
        return out

    def field(self, which=0):
        """p^n (which 0) or p^{n-1} (which 1) as float64 [nz, ny, nx]."""
        out = np.zeros(self.n, np.float64)
        self._chk(self.L.virs_read_field_f64(self._ctx, int(which), out.ctypes.data_as(_lib.c_dp)), "read field")
        nx, ny, nz = self.dims
        return out.reshape(nz, ny, nx)

    def plane_request(self, layer):
        """Queues a snapshot of the XZ plane y = layer of p^n; returns its ticket."""
        t = C.c_int32(-1)
        self._chk(self.L.virs_plane_request(self._ctx, int(layer), C.byref(t)), "plane request")
        return t.value

    def plane_fetch(self, ticket):
        """The snapshot of `ticket` as float32 [nz, nx] (waits for it alone)."""
        nx, ny, nz = self.dims
        out = np.zeros((nz, nx), np.float32)
        self._chk(self.L.virs_plane_fetch(self._ctx, int(ticket), out.ctypes.data_as(_lib.c_fp)), "plane fetch")
        return out

    def write_field(self, which, values):
        v = np.ascontiguousarray(np.asarray(values, np.float64).reshape(-1))
        self._chk(self.L.virs_write_field_f64(self._ctx, int(which), v.ctypes.data_as(_lib.c_dp)), "write field")

    def field_exponent(self, which=0):
        e = C.c_int32(0)
        self._chk(self.L.virs_field_exponent(self._ctx, int(which), C.byref(e)), "field exponent")
        return e.value

    def memory_bytes(self):
        b = C.c_int64(0)
        self._chk(self.L.virs_memory_bytes(self._ctx, C.byref(b)), "memory")
        return b.value

    def time_steps(self, nsteps, nsample=32):
        """Times nsteps steps as virs_step runs them, after a sampling pass of
        nsample single steps that gives the mean launch durations.  Returns a
        dict: total_ms (the timed region), dense_ms and boundary_ms (mean
        launch durations over the sampling pass), and what the timed region
        launched (graph_launches, graph_steps, tail_graph_steps,
        single_steps)."""
        t = _lib.VirsTiming()
        self._chk(self.L.virs_time_steps(self._ctx, int(nsteps), int(nsample), C.byref(t)), "time steps")
        return {k: getattr(t, k) for k, _ in t._fields_}

    # -- adjoint -----------------------------------------------------------
    def forward_record(self, nsteps, ncheckpoints):
        self._rec_steps = 0
        self._chk(self.L.virs_forward_record(self._ctx, int(nsteps), int(ncheckpoints)), "forward record")
        self._rec_steps = int(nsteps)

    def adjoint(self, dl_dir):
        """dL/dIR [nrcv, recorded steps] -> gradient [nmat, 5]."""
        d = np.ascontiguousarray(np.asarray(dl_dir, np.float32))
        want = (self.nrcv, getattr(self, "_rec_steps", 0))
        if want[1] and d.shape != want:
            raise ValueError(f"dl_dir has shape {d.shape}, expected {want} (receivers, recorded steps)")
        g = np.zeros((max(self.nmat, 1), 5), np.float64)
        self._chk(self.L.virs_adjoint(self._ctx, d.ctypes.data_as(_lib.c_fp), g.ctypes.data_as(_lib.c_dp)), "adjoint")
        self._rec_steps = 0
        return g[: self.nmat]


def setup(prep, precision="fp32", use_graph=False, dense_blocks=0, device=0, sponge=True,
          fuse_boundary=False, rounding="stochastic", vector_width=0):
    """A ready-to-run context for a rooms.Prepared workload."""
    sc = prep.scenario
    f = FDTD(prep.room.dims, sc.h, dt=prep.dt, precision=precision,
             sponge_width=sc.sponge_width if sponge else 0, sponge_sigma=sc.sponge_sigma,
             device=device, use_graph=use_graph, dense_blocks=dense_blocks,
             fuse_boundary=fuse_boundary, rounding=rounding, vector_width=vector_width)
    f.upload_room(prep.room)
    f.set_materials(sc.coeffs)
    f.set_soft_sources(prep.src_idx, prep.src_sig)
    f.set_receivers(prep.rcv_idx)
    f.start(sc.steps)
    return f

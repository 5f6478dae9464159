"""ctypes binding of ``libpyfwi_hip.so`` (declared in include/pyfwi_hip.h).

There is no fallback: if the library is missing or no GPU is visible, the
calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpyfwi_hip.so")

MEM_HOST = 0
MEM_DEVICE = 1
MODE_PYFWI = 0
MODE_EXACT = 1
CKPT_AUTO = -1
F_PRECOND = 1
REC_LINE = 0
REC_TABLE = 1
REC_VZ_FROM_VX = 1

(K_FWD_VEL, K_FWD_STR, K_REV_STR, K_REV_VEL, K_ADJ_STR, K_ADJ_VEL, K_EXA_STR, K_EXA_VEL,
 K_FWD_STEP, K_REV_STEP, K_FWD_LOOP, K_BWD_LOOP, K_FWD_SPARSE, K_BWD_SPARSE,
 K_ADJ_STEP, K_FWD_TWO) = range(16)
KERNEL_NAMES = ["fwd_vel", "fwd_str", "rev_str", "rev_vel", "adj_str", "adj_vel_grad",
                "exact_str", "exact_vel", "fwd_step", "rev_step", None, None,
                "fwd_sparse", "bwd_sparse", "adj_step", "fwd_two"]
# bits of set_tuning's `fused`
TUNE_FUSED_STEPS, TUNE_ADJ_FUSED, TUNE_FWD_TWO = 1, 2, 4
TIMING_LAUNCHES, TIMING_LOOPS = 1, 2

# exported symbols, in header order (checked by the CPU test suite)
EXPORTS = [
    "pyfwi_version", "pyfwi_last_error", "pyfwi_plan_create", "pyfwi_plan_destroy",
    "pyfwi_plan_set_stream", "pyfwi_set_receivers", "pyfwi_set_model", "pyfwi_forward",
    "pyfwi_forward_noalloc", "pyfwi_gradient",
    "pyfwi_misfit_gradient", "pyfwi_set_record_filter", "pyfwi_plan_set_tuning", "pyfwi_plan_set_graphs",
    "pyfwi_graph_stats", "pyfwi_get_snapshot", "pyfwi_last_passes", "pyfwi_last_order",
    "pyfwi_synchronize", "pyfwi_plan_set_timing", "pyfwi_kernel_timing",
]

# status codes (include/pyfwi_hip.h)
OK, EINVAL, ENOMEM, EHIP, ESTATE, ENONFINITE, ECFL = range(7)


class Grid(ctypes.Structure):
    _fields_ = [("tnz", ctypes.c_int32), ("tnx", ctypes.c_int32),
                ("npml", ctypes.c_int32), ("sdo", ctypes.c_int32),
                ("c", ctypes.c_double * 4), ("dt", ctypes.c_double),
                ("dx", ctypes.c_double), ("dz", ctypes.c_double)]


class Shot(ctypes.Structure):
    _fields_ = [("src_x", ctypes.c_int32), ("src_z", ctypes.c_int32),
                ("rec_z", ctypes.c_int32), ("rec_x0", ctypes.c_int32),
                ("rec_dx", ctypes.c_int32), ("n_rec", ctypes.c_int32),
                ("rec_stride", ctypes.c_int32), ("nt", ctypes.c_int32),
                ("dt_scale", ctypes.c_int32), ("n_samples", ctypes.c_int32),
                ("src_offset", ctypes.c_int64), ("rec_offset", ctypes.c_int64),
                ("rec_layout", ctypes.c_int32), ("rec_flags", ctypes.c_int32),
                ("rec_table", ctypes.c_int64)]


class NativeError(RuntimeError):
    """Error reported by libpyfwi_hip (status code + message)."""

    def __init__(self, code, msg):
        super().__init__(f"libpyfwi_hip error {code}: {msg}")
        self.code = code


_lib = None


def load(path=None):
    """Load and prototype the library once; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("PYFWI_LIB") or LIB_PATH  # PYFWI_LIB: another build
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (hipcc, gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i32 = ctypes.c_int32
    fp = ctypes.c_void_p  # float pointers are passed as raw addresses
    lib.pyfwi_version.restype = i32
    lib.pyfwi_last_error.restype = ctypes.c_char_p
    lib.pyfwi_plan_create.argtypes = [ctypes.POINTER(Grid), P, ctypes.POINTER(P)]
    lib.pyfwi_plan_destroy.argtypes = [P]
    lib.pyfwi_plan_set_stream.argtypes = [P, P]
    lib.pyfwi_set_receivers.argtypes = [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
    lib.pyfwi_set_model.argtypes = [P, fp, fp, fp, fp, fp, i32]
    lib.pyfwi_forward.argtypes = [P, ctypes.POINTER(Shot), i32, fp, fp,
                                  ctypes.POINTER(ctypes.c_int32), i32, i32, i32]
    lib.pyfwi_forward_noalloc.argtypes = [P, ctypes.POINTER(Shot), i32, fp, fp]
    lib.pyfwi_gradient.argtypes = [P, ctypes.POINTER(Shot), i32, fp, fp, fp, i32, i32, i32, i32]
    lib.pyfwi_plan_set_tuning.argtypes = [P, i32, i32, i32]
    lib.pyfwi_misfit_gradient.argtypes = [P, ctypes.POINTER(Shot), i32, fp, fp, fp, fp,
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_int32), i32, i32, i32, i32]
    lib.pyfwi_get_snapshot.argtypes = [P, i32, i32, fp, i32]
    lib.pyfwi_plan_set_graphs.argtypes = [P, i32]
    lib.pyfwi_set_record_filter.argtypes = [P, ctypes.POINTER(ctypes.c_double), i32, i32]
    lib.pyfwi_graph_stats.argtypes = [P, ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64)]
    lib.pyfwi_last_passes.argtypes = [P, ctypes.POINTER(ctypes.c_int32), i32,
                                      ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32)]
    lib.pyfwi_last_order.argtypes = [P, ctypes.POINTER(ctypes.c_int32), i32,
                                     ctypes.POINTER(ctypes.c_int32)]
    lib.pyfwi_synchronize.argtypes = [P]
    lib.pyfwi_plan_set_timing.argtypes = [P, i32]
    lib.pyfwi_kernel_timing.argtypes = [P, i32, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int64)]
    for name in EXPORTS:
        if name not in ("pyfwi_version", "pyfwi_last_error"):
            getattr(lib, name).restype = i32
    _lib = lib
    return lib


def check(code):
    if code != 0:
        msg = _lib.pyfwi_last_error().decode(errors="replace")
        raise NativeError(code, msg)


def ptr(a, mem=None):
    """Raw address of a contiguous numpy array or torch tensor.  With `mem`
    given, the buffer must live where `mem` says: host memory (a numpy
    array or a CPU tensor) for MEM_HOST, a GPU tensor for MEM_DEVICE -- the
    library takes the address as it is, and a kernel that reads host memory
    through it faults the GPU."""
    if isinstance(a, np.ndarray):
        if mem == MEM_DEVICE:
            raise ValueError("a host (numpy) array was passed with MEM_DEVICE")
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    on_gpu = bool(getattr(a, "is_cuda", False))
    if mem == MEM_DEVICE and not on_gpu:
        raise ValueError("a host tensor was passed with MEM_DEVICE")
    if mem == MEM_HOST and on_gpu:
        raise ValueError("a device tensor was passed with MEM_HOST")
    if hasattr(a, "is_contiguous") and not a.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return a.data_ptr()


def numel(a):
    """Number of float elements of a numpy array or torch tensor."""
    return int(a.size) if isinstance(a, np.ndarray) else int(a.numel())


def extents(shots):
    """Floats the source table and the record buffers must hold for a shot
    table: max(src_offset + 5*ceil(nt/dt_scale)) and
    max(rec_offset + 5*n_samples*rec_stride) (src_extent / rec_extent of the
    library)."""
    src = rec = 0
    for sh in shots:
        ninj = -(-int(sh.nt) // max(1, int(sh.dt_scale)))
        src = max(src, int(sh.src_offset) + 5 * ninj)
        rec = max(rec, int(sh.rec_offset) + 5 * int(sh.n_samples) * int(sh.rec_stride))
    return src, rec


def _need(a, n, what):
    """The kernels read and write a caller buffer as far as the shot table
    says; a shorter buffer would be read or written past its end."""
    if a is not None and numel(a) < n:
        raise ValueError(f"{what} holds {numel(a)} floats; the call needs {n}")


class Plan:
    """Owner of one native plan (one grid on the current GPU)."""

    def __init__(self, grid, stream=None):
        self.lib = load()
        self.grid = grid
        h = ctypes.c_void_p()
        check(self.lib.pyfwi_plan_create(ctypes.byref(grid), stream, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.pyfwi_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        check(self.lib.pyfwi_plan_set_stream(self.h, stream))

    def set_receivers(self, table):
        """Receiver table (host int32 cells row*tnx + column, or -1) read by
        REC_TABLE shots; see pyfwi_set_receivers."""
        t = np.ascontiguousarray(table, np.int32)
        check(self.lib.pyfwi_set_receivers(
            self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if t.size else None,
            int(t.size)))

    @property
    def cells(self):
        return int(self.grid.tnz) * int(self.grid.tnx)

    def set_model(self, lam, mu, buoy, dxp, dzp, mem):
        for a, n, what in ((lam, self.cells, "lam"), (mu, self.cells, "mu"),
                           (buoy, self.cells, "buoyancy"), (dxp, int(self.grid.tnx), "dx profile"),
                           (dzp, int(self.grid.tnz), "dz profile")):
            _need(a, n, what)
        check(self.lib.pyfwi_set_model(self.h, ptr(lam, mem), ptr(mu, mem), ptr(buoy, mem),
                                       ptr(dxp, mem), ptr(dzp, mem), mem))

    def forward(self, shots, src, seis, chp, ckpt_every, mem):
        ns, nr = extents(shots)
        _need(src, ns, "source table")
        _need(seis, nr, "records")
        arr = (Shot * len(shots))(*shots)
        chp = np.ascontiguousarray(chp, np.int32) if chp is not None else np.zeros(0, np.int32)
        cp = chp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if chp.size else None
        check(self.lib.pyfwi_forward(self.h, arr, len(shots), ptr(src, mem), ptr(seis, mem), cp,
                                     int(chp.size), int(ckpt_every), mem))

    def forward_noalloc(self, shots, src, seis):
        """Records of all shots from device buffers without any allocation
        (see pyfwi_forward_noalloc)."""
        ns, nr = extents(shots)
        _need(src, ns, "source table")
        _need(seis, nr, "records")
        arr = (Shot * len(shots))(*shots)
        check(self.lib.pyfwi_forward_noalloc(self.h, arr, len(shots), ptr(src, MEM_DEVICE),
                                             ptr(seis, MEM_DEVICE)))

    def gradient(self, shots, src, res, grad, mode, ckpt_every, flags, mem):
        ns, nr = extents(shots)
        _need(src, ns, "source table")
        _need(res, nr, "residuals")
        _need(grad, 6 * self.cells, "gradient")
        arr = (Shot * len(shots))(*shots)
        check(self.lib.pyfwi_gradient(self.h, arr, len(shots),
                                      ptr(src, mem) if src is not None else None,
                                      ptr(res, mem), ptr(grad, mem), mode, int(ckpt_every), flags,
                                      mem))

    def misfit_gradient(self, shots, src, obs, seis, grad, chp, comp_mask, flags, mem):
        """Returns the L2 misfit; see pyfwi_misfit_gradient."""
        ns, nr = extents(shots)
        _need(src, ns, "source table")
        _need(obs, nr, "observed records")
        _need(seis, nr, "records")
        _need(grad, 6 * self.cells, "gradient")
        arr = (Shot * len(shots))(*shots)
        chp = np.ascontiguousarray(chp, np.int32)
        out = ctypes.c_double()
        check(self.lib.pyfwi_misfit_gradient(
            self.h, arr, len(shots), ptr(src, mem), ptr(obs, mem),
            ptr(seis, mem) if seis is not None else None, ptr(grad, mem), ctypes.byref(out), chp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            int(chp.size), int(comp_mask), int(flags), mem))
        return out.value

    def set_tuning(self, shots_per_pass=0, fused=0, strip_rows=0):
        """fused: bits TUNE_FUSED_STEPS, TUNE_ADJ_FUSED, TUNE_FWD_TWO (True
        means TUNE_FUSED_STEPS)."""
        check(self.lib.pyfwi_plan_set_tuning(self.h, int(shots_per_pass), int(fused),
                                             int(strip_rows)))

    def set_record_filter(self, taps, pad=512):
        """Lowpass of pyfwi_misfit_gradient's misfit (taps k[0..L], float64),
        or none with taps=None; see pyfwi_set_record_filter."""
        if taps is None:
            check(self.lib.pyfwi_set_record_filter(self.h, None, -1, 0))
            return
        t = np.ascontiguousarray(taps, np.float64)
        check(self.lib.pyfwi_set_record_filter(
            self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(t.size) - 1, int(pad)))

    def set_graphs(self, enable):
        """Replay the time loops as captured hipGraphs (see pyfwi_plan_set_graphs)."""
        check(self.lib.pyfwi_plan_set_graphs(self.h, 1 if enable else 0))

    def graph_stats(self):
        """(loops captured, loops replayed) since the plan was created."""
        c, r = ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.pyfwi_graph_stats(self.h, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def snapshot(self, shot, slot, out, mem):
        _need(out, 5 * self.cells, "snapshot output")
        check(self.lib.pyfwi_get_snapshot(self.h, shot, slot, ptr(out, mem), mem))

    def last_passes(self, with_interval=False):
        """Shots per pass of the last forward / gradient / misfit call (and,
        with_interval, the exact mode's checkpoint interval)."""
        cap = 4096
        buf = (ctypes.c_int32 * cap)()
        n, k = ctypes.c_int32(), ctypes.c_int32()
        check(self.lib.pyfwi_last_passes(self.h, buf, cap, ctypes.byref(n), ctypes.byref(k)))
        if n.value > cap:
            buf = (ctypes.c_int32 * n.value)()
            check(self.lib.pyfwi_last_passes(self.h, buf, n.value, ctypes.byref(n),
                                             ctypes.byref(k)))
        sizes = [int(buf[j]) for j in range(n.value)]
        return (sizes, int(k.value)) if with_interval else sizes

    def last_order(self):
        """The caller's shot indices of the last call in the order its passes
        ran them (the passes take consecutive runs of it)."""
        n = ctypes.c_int32()
        check(self.lib.pyfwi_last_order(self.h, None, 0, ctypes.byref(n)))
        buf = (ctypes.c_int32 * max(n.value, 1))()
        check(self.lib.pyfwi_last_order(self.h, buf, n.value, ctypes.byref(n)))
        return [int(buf[j]) for j in range(n.value)]

    def synchronize(self):
        check(self.lib.pyfwi_synchronize(self.h))

    def set_timing(self, mode):
        """0 off, TIMING_LAUNCHES (or True) every stencil launch, TIMING_LOOPS
        the time loops (see pyfwi_plan_set_timing)."""
        check(self.lib.pyfwi_plan_set_timing(self.h, int(mode)))

    def kernel_timing(self, kind):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(self.lib.pyfwi_kernel_timing(self.h, kind, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

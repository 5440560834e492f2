"""ctypes binding of libqpd.so, the HIP/gfx950 library of the diffraction path.

The library is built in-tree (``qpop_diffraction/lib/libqpd.so``; ``QPD_LIB``
overrides the path).  Device buffers are PyTorch-ROCm tensors passed by
``data_ptr()``; the stream is ``torch.cuda.current_stream().cuda_stream``.
Nothing here falls back to the CPU: a missing library or GPU raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("QPD_LIB", Path(__file__).resolve().parent / "lib" / "libqpd.so"))

# Entry points declared in include/qpd.h
EXPORTS = (
    "qpd_plan_create", "qpd_plan_create_opts", "qpd_plan_info", "qpd_plan_pass_bits", "qpd_plan_destroy",
    "qpd_workspace_bytes",
    "qpd_launches_per_volume",
    "qpd_plan_schedules", "qpd_plan_schedule_name", "qpd_plan_set_schedule", "qpd_plan_get_schedule",
    "qpd_fast_lengths", "qpd_intensity", "qpd_intensity_pass", "qpd_intensity_noalloc", "qpd_pass_bytes_per_cell",
    "qpd_fft3d",
    "qpd_fft_length_supported", "qpd_ingest_fields", "qpd_project", "qpd_reverse_axes", "qpd_last_error",
    "qpd_version",
)

# QPD_NOALLOC_SCRATCH_BYTES: the static per-GPU scratch of qpd_intensity_noalloc
NOALLOC_SCRATCH_BYTES = 64 << 20
NO_WORKSPACE = object()  # _check_inputs: the call brings no workspace

ROLES = {"source": 0, "contiguous": 1, "strided": 2, "intensity": 3, "strided1": 4, "split": 5,
         "strided_split": 6, "intensity_split": 7, "strided_groups1": 8, "strided_groups2": 9,
         "groups1_split": 10, "groups2_split": 11, "source_wave": 12}

c_double3 = ctypes.c_double * 3


class Grid(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int64 * 3), ("axis", ctypes.c_int * 3), ("h", c_double3)]


class Crystal(ctypes.Structure):
    _fields_ = [
        ("n_atom", ctypes.c_int), ("n_phase", ctypes.c_int), ("n_struc", ctypes.c_int),
        ("R", ctypes.POINTER(ctypes.c_double)), ("f_re", ctypes.POINTER(ctypes.c_double)),
        ("f_im", ctypes.POINTER(ctypes.c_double)), ("su", ctypes.POINTER(ctypes.c_double)),
        ("G", c_double3), ("qC", c_double3), ("V", ctypes.c_double),
    ]


class Fields(ctypes.Structure):
    _fields_ = [("P", ctypes.c_void_p), ("U", ctypes.c_void_p), ("phase", ctypes.c_void_p),
                ("fraction", ctypes.c_void_p), ("region", ctypes.c_void_p)]


class Ingest(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int64 * 3), ("nz_full", ctypes.c_int64), ("z0", ctypes.c_int64),
                ("P", ctypes.c_void_p * 3), ("P_bits", ctypes.c_int), ("P_div", ctypes.c_double),
                ("U", ctypes.c_void_p * 3), ("U_bits", ctypes.c_int), ("U_div", ctypes.c_double),
                ("material", ctypes.c_void_p), ("n_material", ctypes.c_int), ("dc", ctypes.c_void_p),
                ("phase_of", ctypes.c_void_p), ("dz", ctypes.c_double)]


class QPDError(RuntimeError):
    """A libqpd call returned an error code."""


_lib = None


def load_library():
    """Load libqpd.so once; raise ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"libqpd.so not found at {LIB_PATH}; build it with `python __graft_entry__.py build` "
                          "or `python q-pop-diffraction_amd/build_lib.py`")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    lib.qpd_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Grid), ctypes.POINTER(Crystal), i32]
    lib.qpd_plan_create.restype = i32
    lib.qpd_plan_create_opts.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Grid), ctypes.POINTER(Crystal), i32, i32,
                                         i32]
    lib.qpd_plan_create_opts.restype = i32
    lib.qpd_plan_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    lib.qpd_plan_info.restype = i32
    lib.qpd_plan_pass_bits.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    lib.qpd_plan_pass_bits.restype = i32
    lib.qpd_plan_destroy.argtypes = [vp]
    lib.qpd_plan_destroy.restype = i32
    lib.qpd_workspace_bytes.argtypes = [vp]
    lib.qpd_workspace_bytes.restype = ctypes.c_size_t
    lib.qpd_launches_per_volume.argtypes = [vp]
    lib.qpd_launches_per_volume.restype = i32
    lib.qpd_plan_schedules.argtypes = [vp, i32]
    lib.qpd_plan_schedules.restype = i32
    lib.qpd_plan_schedule_name.argtypes = [vp, i32, i32, ctypes.c_char_p, i32]
    lib.qpd_plan_schedule_name.restype = i32
    lib.qpd_plan_set_schedule.argtypes = [vp, i32, i32]
    lib.qpd_plan_set_schedule.restype = i32
    lib.qpd_plan_get_schedule.argtypes = [vp, i32]
    lib.qpd_plan_get_schedule.restype = i32
    lib.qpd_fast_lengths.argtypes = [i32, ctypes.POINTER(i64), i32]
    lib.qpd_fast_lengths.restype = i32
    lib.qpd_intensity.argtypes = [vp, ctypes.POINTER(Fields), vp, vp, vp]
    lib.qpd_intensity.restype = i32
    lib.qpd_intensity_pass.argtypes = [vp, ctypes.POINTER(Fields), vp, vp, i32, vp]
    lib.qpd_intensity_pass.restype = i32
    lib.qpd_intensity_noalloc.argtypes = [vp, ctypes.POINTER(Fields), vp, vp]
    lib.qpd_intensity_noalloc.restype = i32
    lib.qpd_pass_bytes_per_cell.argtypes = [vp, ctypes.POINTER(Fields), i32]
    lib.qpd_pass_bytes_per_cell.restype = ctypes.c_double
    lib.qpd_fft3d.argtypes = [ctypes.POINTER(i64), i32, vp, i32, vp]
    lib.qpd_fft3d.restype = i32
    lib.qpd_ingest_fields.argtypes = [ctypes.POINTER(Ingest), i32, vp, vp, vp, vp]
    lib.qpd_ingest_fields.restype = i32
    lib.qpd_project.argtypes = [ctypes.POINTER(i64), vp, vp, vp, vp, vp]
    lib.qpd_project.restype = i32
    lib.qpd_reverse_axes.argtypes = [ctypes.POINTER(i64), i32, i32, vp, vp, vp]
    lib.qpd_reverse_axes.restype = i32
    lib.qpd_fft_length_supported.argtypes = [i64]
    lib.qpd_fft_length_supported.restype = i32
    lib.qpd_last_error.argtypes = []
    lib.qpd_last_error.restype = ctypes.c_char_p
    lib.qpd_version.argtypes = []
    lib.qpd_version.restype = i32
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load_library().qpd_last_error().decode(errors="replace")
        raise QPDError(f"libqpd error {rc}: {msg}")


def _dptr(a):
    return ctypes.cast(a.ctypes.data, ctypes.POINTER(ctypes.c_double))


ORDERS = {"auto": 0, "memory": 1, "swapped": 2}
ARITHS = {"auto": 0, "fp32": 1, "mixed": 2, "fp64": 3}


class Plan:
    """A libqpd plan: one grid layout, one crystal, one precision.

    dims are the memory extents (dims[0] fastest) and axis[d] the physical axis
    (0 = x, 1 = y, 2 = z) of memory dim d; h the x, y, z spacings in nm.
    order ("auto", "memory", "swapped") is the order of the two strided transforms,
    arith ("auto", "fp32", "mixed", "fp64") the arithmetic of a precision-32 plan: FP32
    transforms and spectra, FP64 transforms with FP32 spectra between passes, or FP64
    throughout (inputs widened); "auto" takes "mixed" where its error estimate allows.
    """

    def __init__(self, dims, axis, h, R, f, su, G, qC, V, precision=32, order="auto", arith="auto"):
        lib = load_library()
        self.dims = tuple(int(d) for d in dims)
        self.axis = tuple(int(a) for a in axis)
        self.precision = int(precision)
        self.n = int(np.prod(self.dims))
        self._R = np.ascontiguousarray(R, dtype=np.float64)
        f = np.asarray(f, dtype=complex)
        self._fre = np.ascontiguousarray(f.real, dtype=np.float64)
        self._fim = np.ascontiguousarray(f.imag, dtype=np.float64)
        self._su = np.ascontiguousarray(su, dtype=np.float64)
        n_atom = self._R.shape[0]
        n_phase = f.shape[0]
        if self._su.shape != (3, 3, n_phase, n_atom):
            raise ValueError(f"su must be (3, 3, nPhase, nAtom), got {self._su.shape}")
        self.n_phase = n_phase
        self.n_atom = n_atom
        grid = Grid((ctypes.c_int64 * 3)(*self.dims), (ctypes.c_int * 3)(*self.axis), c_double3(*h))
        cr = Crystal(n_atom, n_phase, 3, _dptr(self._R), _dptr(self._fre), _dptr(self._fim), _dptr(self._su),
                     c_double3(*G), c_double3(*qC), float(V))
        if order not in ORDERS or arith not in ARITHS:
            raise ValueError(f"order must be one of {list(ORDERS)} and arith one of {list(ARITHS)}")
        handle = ctypes.c_void_p()
        check(lib.qpd_plan_create_opts(ctypes.byref(handle), ctypes.byref(grid), ctypes.byref(cr), self.precision,
                                       ORDERS[order], ARITHS[arith]))
        self._h = handle
        self.workspace_bytes = int(lib.qpd_workspace_bytes(handle))
        self.launches = int(lib.qpd_launches_per_volume(handle))
        pd = (ctypes.c_int * 3)()
        cbits, sbits = ctypes.c_int(), ctypes.c_int()
        est = ctypes.c_double()
        check(lib.qpd_plan_info(handle, pd, ctypes.byref(cbits), ctypes.byref(sbits), ctypes.byref(est)))
        self.pass_dims = tuple(int(v) for v in pd)  # memory dim transformed by each pass
        self.compute_bits = int(cbits.value)        # precision of the transforms
        self.storage_bits = int(sbits.value)        # precision of the spectra between passes
        self.dlog10_estimate = float(est.value)     # the error estimate of that arithmetic
        pb = (ctypes.c_int * 3)()
        check(lib.qpd_plan_pass_bits(handle, pb))
        self.pass_bits = tuple(int(v) for v in pb)   # bits of each pass's transforms, in pass order

    def _check_inputs(self, P, U, phase, I, work, fraction, region, which=None):
        """Shapes, dtypes, device and contiguity of the buffers a call reads or writes (every
        pass, or pass ``which`` alone), before any device access."""
        import torch
        real = torch.float32 if self.precision == 32 else torch.float64
        n = self.n

        def need(t, name, dtype, numel, optional=False):
            if t is None:
                if optional:
                    return
                raise ValueError(f"{name} is required")
            if not isinstance(t, torch.Tensor) or not t.is_cuda:
                raise ValueError(f"{name} must be a CUDA tensor")
            if t.dtype != dtype:
                raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
            if not t.is_contiguous():
                raise ValueError(f"{name} must be contiguous")
            if numel is not None and t.numel() != numel:
                raise ValueError(f"{name} must hold {numel} values, got {t.numel()}")

        if which in (None, 0):
            need(P, "P", real, 3 * n)
            need(U, "U", torch.float64, 3 * n)
            need(phase, "phase", torch.int8, n)
            need(fraction, "fraction", real, self.n_phase * n, optional=True)
            need(region, "region", real, n, optional=True)
        else:
            P = U = phase = fraction = region = None
        if which in (None, 2):
            need(I, "I", torch.float64, n)
        else:
            I = None
        if work is not NO_WORKSPACE:
            if not isinstance(work, torch.Tensor) or not work.is_cuda or not work.is_contiguous():
                raise ValueError("work must be a contiguous CUDA tensor")
            if work.numel() * work.element_size() < self.workspace_bytes:
                raise ValueError(f"work holds {work.numel() * work.element_size()} bytes, the plan needs "
                                 f"{self.workspace_bytes}")
        else:
            work = None
        devs = {t.device for t in (P, U, phase, fraction, region, I, work) if t is not None}
        if len(devs) != 1:
            raise ValueError(f"buffers are on several devices: {sorted(str(d) for d in devs)}")

    def intensity(self, P, U, phase, I, work, fraction=None, region=None, stream=None):
        """Enqueue one volume: device tensors in, fftshifted |A|^2 (float64) into I."""
        self._check_inputs(P, U, phase, I, work, fraction, region)
        fields = self._fields(P, U, phase, fraction, region)
        check(load_library().qpd_intensity(self._h, ctypes.byref(fields), I.data_ptr(), work.data_ptr(),
                                           stream if stream is not None else current_stream()))

    def intensity_noalloc(self, P, U, phase, I, fraction=None, region=None, stream=None):
        """intensity() without a workspace: the library's static per-GPU scratch
        (NOALLOC_SCRATCH_BYTES) holds the intermediate spectra and nothing is allocated.
        Calls that overlap on one GPU share that scratch and race; keep them on one stream."""
        self._check_inputs(P, U, phase, I, NO_WORKSPACE, fraction, region)
        if self.workspace_bytes > NOALLOC_SCRATCH_BYTES:
            raise ValueError(f"the plan needs {self.workspace_bytes} bytes of workspace; the no-allocation "
                             f"entry point holds at most {NOALLOC_SCRATCH_BYTES}")
        fields = self._fields(P, U, phase, fraction, region)
        check(load_library().qpd_intensity_noalloc(self._h, ctypes.byref(fields), I.data_ptr(),
                                                   stream if stream is not None else current_stream()))

    def _fields(self, P, U, phase, fraction, region):
        return Fields(P.data_ptr(), U.data_ptr(), phase.data_ptr(),
                      fraction.data_ptr() if fraction is not None else None,
                      region.data_ptr() if region is not None else None)

    def run_pass(self, which, P, U, phase, I, work, fraction=None, region=None, stream=None):
        """Enqueue one of the three passes only (for timing each kernel)."""
        self._check_inputs(P, U, phase, I, work, fraction, region, which=int(which))
        fields = self._fields(P, U, phase, fraction, region)
        check(load_library().qpd_intensity_pass(self._h, ctypes.byref(fields), I.data_ptr(), work.data_ptr(),
                                                int(which), stream if stream is not None else current_stream()))

    def pass_bytes_per_cell(self, which, P, U, phase, fraction=None, region=None):
        fields = self._fields(P, U, phase, fraction, region)
        return float(load_library().qpd_pass_bytes_per_cell(self._h, ctypes.byref(fields), int(which)))

    def schedules(self, which):
        """Names of the launch schedules pass `which` can run (index 0 is the default)."""
        lib = load_library()
        n = lib.qpd_plan_schedules(self._h, int(which))
        if n < 0:
            check(n)
        names = []
        for i in range(n):
            buf = ctypes.create_string_buffer(256)
            lib.qpd_plan_schedule_name(self._h, int(which), i, buf, 256)
            names.append(buf.value.decode())
        return names

    def schedule(self, which):
        rc = load_library().qpd_plan_get_schedule(self._h, int(which))
        if rc < 0:
            check(rc)
        return rc

    def set_schedule(self, which, index):
        check(load_library().qpd_plan_set_schedule(self._h, int(which), int(index)))
        self.launches = int(load_library().qpd_launches_per_volume(self._h))

    def close(self):
        if getattr(self, "_h", None):
            load_library().qpd_plan_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def current_stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def fft3d(x, precision=None, stream=None):
    """In-place forward unnormalised 3-D DFT of a complex device tensor (..., d2, d1, d0)."""
    import torch
    if x.dtype not in (torch.complex64, torch.complex128) or not x.is_contiguous() or not x.is_cuda:
        raise TypeError("fft3d needs a contiguous complex64/complex128 CUDA tensor")
    prec = 32 if x.dtype == torch.complex64 else 64
    if precision is not None and precision != prec:
        raise TypeError("precision does not match the tensor dtype")
    d2, d1, d0 = x.shape[-3:]
    nf = int(np.prod(x.shape[:-3])) if x.dim() > 3 else 1
    dims = (ctypes.c_int64 * 3)(d0, d1, d2)
    check(load_library().qpd_fft3d(dims, prec, ctypes.c_void_p(x.data_ptr()), nf,
                                   stream if stream is not None else current_stream()))
    return x


def fast_lengths(role):
    """Transform lengths with a compile-time FP32 kernel of this build for a role (name or number)."""
    lib = load_library()
    r = ROLES[role] if isinstance(role, str) else int(role)
    n = lib.qpd_fast_lengths(r, None, 0)
    if n < 0:
        check(n)
    buf = (ctypes.c_int64 * max(1, n))()
    lib.qpd_fast_lengths(r, buf, n)
    return [int(buf[i]) for i in range(n)]


def reverse_axes(t, dims, stream=None):
    """A new device tensor holding t's fields with the axis order reversed.

    t: contiguous CUDA tensor of n_comp x dims[0] x dims[1] x dims[2] elements (dims[0]
    fastest; 1-, 4- or 8-byte elements, any leading shape); the result has t's shape and
    dims[2] fastest.
    """
    import torch
    if not t.is_cuda or not t.is_contiguous():
        raise TypeError("reverse_axes needs a contiguous CUDA tensor")
    d0, d1, d2 = (int(v) for v in dims)
    n = d0 * d1 * d2
    if t.numel() % n:
        raise ValueError(f"{t.numel()} elements are not whole fields of {d0} x {d1} x {d2}")
    out = torch.empty_like(t)
    check(load_library().qpd_reverse_axes((ctypes.c_int64 * 3)(d0, d1, d2), t.numel() // n, t.element_size(),
                                          ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                          stream if stream is not None else current_stream()))
    return out


def project(A, dims, which=(0, 1, 2), stream=None):
    """Sums of a device float64 array (dims[0] fastest) over the memory dims in ``which``.

    Returns {d: tensor}: the sum over dim d, shaped by the two other dims (slower one first).
    """
    import torch
    if A.dtype != torch.float64 or not A.is_cuda or not A.is_contiguous():
        raise TypeError("project needs a contiguous float64 CUDA tensor")
    d0, d1, d2 = (int(v) for v in dims)
    shapes = {0: (d2, d1), 1: (d2, d0), 2: (d1, d0)}
    out = {d: torch.empty(shapes[d], dtype=torch.float64, device=A.device) for d in which}
    ptr = lambda d: ctypes.c_void_p(out[d].data_ptr()) if d in out else None
    check(load_library().qpd_project((ctypes.c_int64 * 3)(d0, d1, d2), ctypes.c_void_p(A.data_ptr()), ptr(0), ptr(1),
                                     ptr(2), stream if stream is not None else current_stream()))
    return out

"""Python binding of libsmip_mi355x.so (the MI355X speckle-masking engine).

A thin ctypes layer over the C ABI in include/smip_mi355x.h; every call runs
on the GPU through the HIP kernels of speckle-stacker_amd/csrc.  There is no
CPU fallback: if the library or a GPU is missing, calls raise.

Arrays use the layout of smip::Array2 (row-major, x fastest), i.e. numpy
arrays of shape (nrows, ncols); complex values are complex128 unless the
bispectrum is complex64.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(PKG, "lib")

OK = 0
ERR_INVALID_ARGUMENT, ERR_OUT_OF_RANGE, ERR_RUNTIME = 1, 2, 3
ERR_BAD_FUNCTION_CALL, ERR_LOGIC, ERR_DOMAIN, ERR_HIP, ERR_NO_MEMORY = 4, 5, 6, 7, 8

DTYPE_U8, DTYPE_U16, DTYPE_F32, DTYPE_F64, DTYPE_C64, DTYPE_C128 = 1, 2, 3, 4, 5, 6
OP_ADD, OP_SUB, OP_MUL, OP_DIV = 0, 1, 2, 3
PRECISION_FP64, PRECISION_MIXED = 0, 1

_NP_DTYPES = {np.dtype(np.uint8): DTYPE_U8, np.dtype(np.uint16): DTYPE_U16,
              np.dtype(np.float32): DTYPE_F32, np.dtype(np.float64): DTYPE_F64}


class SmipError(RuntimeError):
    """A failing ABI call; ``code`` maps to the reference's exception class."""

    EXC = {ERR_INVALID_ARGUMENT: "invalid_argument", ERR_OUT_OF_RANGE: "out_of_range",
           ERR_RUNTIME: "runtime_error", ERR_BAD_FUNCTION_CALL: "bad_function_call",
           ERR_LOGIC: "logic_error", ERR_DOMAIN: "domain_error", ERR_HIP: "hip_error",
           ERR_NO_MEMORY: "no_memory"}

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{self.EXC.get(code, code)}] {msg}")
        self.code = code


class BispecInfo(C.Structure):
    _fields_ = [("dims", C.c_uint64 * 4), ("sizes", C.c_uint64 * 4), ("base_sizes", C.c_uint64 * 4),
                ("base_size", C.c_uint64), ("totalsize", C.c_uint64),
                ("min_indices", C.c_int32 * 4), ("max_indices", C.c_int32 * 4), ("dtype", C.c_int32)]


class JobParams(C.Structure):
    _fields_ = [("ncols", C.c_uint64), ("nrows", C.c_uint64), ("bdepth", C.c_uint64),
                ("reco_radius", C.c_double), ("frame_dtype", C.c_int32), ("precision", C.c_int32),
                ("saa", C.c_int32), ("frames_per_batch", C.c_int32), ("vy_tile", C.c_int32)]


class JobStats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("accumulate_ms", C.c_double), ("saa_ms", C.c_double),
                ("reco_ms", C.c_double), ("synth_ms", C.c_double), ("footprint", C.c_uint64),
                ("reco_calls", C.c_uint64), ("cells_per_frame", C.c_uint64),
                ("triples_per_frame", C.c_uint64), ("tile_ms", C.c_double), ("tile_launches", C.c_uint64),
                ("graph_launches", C.c_uint64)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_active", C.c_int32), ("workers", C.c_int32)]


class ServerJobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("footprint", C.c_uint64), ("frames_in", C.c_uint64),
                ("frames_done", C.c_uint64), ("queue_ms", C.c_double), ("latency_ms", C.c_double),
                ("tail_ms", C.c_double), ("error", C.c_char * 256)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("failed", C.c_uint64), ("waiting", C.c_uint64), ("active", C.c_uint64),
                ("bytes_in_use", C.c_uint64), ("budget", C.c_uint64), ("frames", C.c_uint64),
                ("finalize_groups", C.c_uint64), ("finalized_in_groups", C.c_uint64)]


class TuneResult(C.Structure):
    _fields_ = [("vy_tile", C.c_int32), ("frames_per_batch", C.c_int32), ("frames_per_s", C.c_double),
                ("tile_tflops", C.c_double), ("candidates", C.c_int32)]


JOB_WAITING, JOB_ACTIVE, JOB_DONE, JOB_FAILED = 0, 1, 2, 3

# name -> (restype, argtypes); every function in include/smip_mi355x.h
_P = C.c_void_p
_SIGS = {
    "smip_last_error": (C.c_char_p, []),
    "smip_version": (C.c_int, [_P, _P, _P]),
    "smip_ctx_create": (C.c_int, [C.c_int32, _P]),
    "smip_ctx_destroy": (C.c_int, [_P]),
    "smip_ctx_synchronize": (C.c_int, [_P]),
    "smip_ctx_mem_info": (C.c_int, [_P, _P, _P]),
    "smip_bispec_create": (C.c_int, [_P, _P, C.c_int32, _P]),
    "smip_bispec_destroy": (C.c_int, [_P]),
    "smip_bispec_info_get": (C.c_int, [_P, _P]),
    "smip_bispec_accumulate_from_fft": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64]),
    "smip_bispec_get_elements": (C.c_int, [_P, _P, C.c_uint64, _P]),
    "smip_bispec_get_elements_noalloc": (C.c_int, [_P, _P, C.c_uint64, _P]),
    "smip_bispec_put_element": (C.c_int, [_P, _P, _P]),
    "smip_bispec_calc_offset": (C.c_uint64, [_P, _P]),
    "smip_bispec_calc_indices": (C.c_int, [_P, C.c_uint64, _P]),
    "smip_bispec_download": (C.c_int, [_P, _P]),
    "smip_bispec_upload": (C.c_int, [_P, _P]),
    "smip_bispec_write_file": (C.c_int, [_P, C.c_char_p]),
    "smip_bispec_read_file": (C.c_int, [_P, C.c_char_p]),
    "smip_bispec_scalar_op": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double]),
    "smip_bispec_array_op": (C.c_int, [_P, _P, C.c_int32]),
    "smip_bispec_assign": (C.c_int, [_P, _P]),
    "smip_reconstruct_phases": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_double, _P, _P, _P]),
    "smip_calc_phase": (C.c_int, [_P, C.c_uint64, C.c_uint64, _P, _P, _P, C.c_int32, C.c_int32]),
    "smip_next_reco_index": (None, [_P, _P, _P, _P]),
    "smip_xcorr_create": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64, _P]),
    "smip_xcorr_destroy": (C.c_int, [_P]),
    "smip_xcorr_correlate": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64]),
    "smip_xcorr_get_correlation": (C.c_int, [_P, _P]),
    "smip_xcorr_get_displacement": (C.c_int, [_P, _P]),
    "smip_xcorr_displacement": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64, _P]),
    "smip_general_hamming": (C.c_int, [C.c_uint64, C.c_uint64, C.c_double, C.c_double, _P]),
    "smip_frames_create": (C.c_int, [_P, C.c_int32, C.c_uint64, C.c_uint64, C.c_uint64, _P]),
    "smip_frames_destroy": (C.c_int, [_P]),
    "smip_frames_upload": (C.c_int, [_P, C.c_uint64, C.c_uint64, _P]),
    "smip_job_footprint": (C.c_int, [_P, _P]),
    "smip_job_create": (C.c_int, [_P, _P, _P]),
    "smip_job_destroy": (C.c_int, [_P]),
    "smip_job_add_frames": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64]),
    "smip_job_add_frames_many": (C.c_int, [_P, _P, _P, _P, C.c_uint64]),
    "smip_job_add_frames_host": (C.c_int, [_P, _P, C.c_uint64]),
    "smip_job_push_frames_host": (C.c_int, [_P, _P, C.c_uint64]),
    "smip_job_finalize": (C.c_int, [_P]),
    "smip_job_finalize_many": (C.c_int, [_P, C.c_uint64, _P]),
    "smip_job_synchronize": (C.c_int, [_P]),
    "smip_job_get_stats": (C.c_int, [_P, _P]),
    "smip_job_bispectrum": (C.c_int, [_P, _P]),
    "smip_job_get_powerspec": (C.c_int, [_P, _P]),
    "smip_job_get_sum_image": (C.c_int, [_P, _P]),
    "smip_job_get_shifts": (C.c_int, [_P, _P, C.c_uint64]),
    "smip_job_get_phases": (C.c_int, [_P, _P, _P, _P]),
    "smip_job_get_image": (C.c_int, [_P, _P]),
    "smip_job_resolve_params": (C.c_int, [_P, _P]),
    "smip_autotune": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, _P]),
    "smip_autotune_forget": (C.c_int, []),
    "smip_server_create": (C.c_int, [_P, _P, _P]),
    "smip_server_destroy": (C.c_int, [_P]),
    "smip_server_submit": (C.c_int, [_P, _P, _P]),
    "smip_server_push": (C.c_int, [_P, C.c_uint64, _P, C.c_uint64]),
    "smip_server_close": (C.c_int, [_P, C.c_uint64]),
    "smip_server_wait": (C.c_int, [_P, C.c_uint64, C.c_double, _P]),
    "smip_server_job": (C.c_int, [_P, C.c_uint64, _P]),
    "smip_server_get_job_info": (C.c_int, [_P, C.c_uint64, _P]),
    "smip_server_release": (C.c_int, [_P, C.c_uint64]),
    "smip_server_get_stats": (C.c_int, [_P, _P]),
}

_LIB: C.CDLL | None = None


def library_path() -> str:
    return os.path.join(LIBDIR, "libsmip_mi355x.so")


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise SmipError(ERR_RUNTIME, f"{path} is not built (run speckle-stacker_amd/build.py)")
        L = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int) -> None:
    if rc != OK:
        raise SmipError(rc, lib().smip_last_error().decode(errors="replace"))


# --------------------------------------------------------------------------
class Context:
    """One GPU (smip_ctx).

    The library refuses to destroy a context while objects made on it are
    alive, so close() first closes this context's live servers, then its jobs,
    then its other objects."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().smip_ctx_create(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()

    def _adopt(self, obj) -> None:
        self._children.add(obj)

    def close(self):
        if self.h:
            order = {Server: 0, Job: 1}
            for obj in sorted(list(self._children), key=lambda o: order.get(type(o), 2)):
                obj.close()
            _check(lib().smip_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().smip_ctx_synchronize(self.h))

    def mem_info(self):
        f, t = C.c_uint64(), C.c_uint64()
        _check(lib().smip_ctx_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value


class Bispectrum:
    """smip::Bispectrum<T> on the device (include/bispectrum.h:28-125)."""

    def __init__(self, ctx: Context, dims=(0, 0, 0, 0), dtype=np.complex128, _handle=None, _owner=None):
        self.ctx = ctx
        self._owner = _owner
        if _handle is not None:
            self.h = _handle
            self.borrowed = True
        else:
            d = (C.c_uint64 * 4)(*[int(x) for x in dims])
            h = C.c_void_p()
            code = DTYPE_C64 if np.dtype(dtype) == np.complex64 else DTYPE_C128
            _check(lib().smip_bispec_create(ctx.h, d, code, C.byref(h)))
            self.h = h
            self.borrowed = False
            ctx._adopt(self)

    def close(self):
        if self.h and not self.borrowed:
            lib().smip_bispec_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> BispecInfo:
        i = BispecInfo()
        _check(lib().smip_bispec_info_get(self.h, C.byref(i)))
        return i

    @property
    def dtype(self):
        return np.complex64 if self.info().dtype == DTYPE_C64 else np.complex128

    @property
    def base_size(self) -> int:
        return int(self.info().base_size)

    def accumulate_from_fft(self, fft: np.ndarray) -> None:
        a = np.ascontiguousarray(fft, dtype=np.complex128)
        _check(lib().smip_bispec_accumulate_from_fft(self.h, _p(a), a.shape[1], a.shape[0]))

    def get_elements(self, idx) -> np.ndarray:
        a = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 4))
        out = np.zeros(a.shape[0], dtype=np.complex128)
        _check(lib().smip_bispec_get_elements(self.h, _p(a), a.shape[0], _p(out)))
        return out

    def get_element(self, idx) -> complex:
        return complex(self.get_elements([idx])[0])

    def put_element(self, idx, value: complex) -> None:
        a = np.asarray(idx, dtype=np.int32)
        v = np.array([value.real, value.imag])
        _check(lib().smip_bispec_put_element(self.h, _p(a), _p(v)))

    def calc_offset(self, idx) -> int:
        a = np.asarray(idx, dtype=np.int32)
        return int(lib().smip_bispec_calc_offset(self.h, _p(a)))

    def calc_indices(self, addr: int):
        out = np.zeros(4, dtype=np.int32)
        _check(lib().smip_bispec_calc_indices(self.h, addr, _p(out)))
        return tuple(int(x) for x in out)

    def download(self) -> np.ndarray:
        out = np.zeros(self.base_size, dtype=self.dtype)
        _check(lib().smip_bispec_download(self.h, _p(out)))
        return out

    def upload(self, data: np.ndarray) -> None:
        a = np.ascontiguousarray(data, dtype=self.dtype)
        if a.size != self.base_size:
            raise SmipError(ERR_INVALID_ARGUMENT, "size mismatch")
        _check(lib().smip_bispec_upload(self.h, _p(a)))

    def write_to_file(self, path: str) -> None:
        _check(lib().smip_bispec_write_file(self.h, path.encode()))

    def read_from_file(self, path: str) -> None:
        _check(lib().smip_bispec_read_file(self.h, path.encode()))

    def scalar_op(self, op: int, value: complex) -> None:
        _check(lib().smip_bispec_scalar_op(self.h, op, value.real, value.imag))

    def array_op(self, other: "Bispectrum", op: int) -> None:
        _check(lib().smip_bispec_array_op(self.h, other.h, op))

    def assign(self, other: "Bispectrum") -> None:
        _check(lib().smip_bispec_assign(self.h, other.h))

    def calc_phase(self, phases: np.ndarray, flags: np.ndarray, consistency: np.ndarray, w) -> None:
        """One calc_phase step on (phases, flags, consistency), updated in place."""
        for a, dt in ((phases, np.complex128), (flags, np.uint8), (consistency, np.float64)):
            if a.dtype != dt or not a.flags.c_contiguous:
                raise SmipError(ERR_INVALID_ARGUMENT, "calc_phase state arrays must be contiguous")
        ys, xs = phases.shape
        _check(lib().smip_calc_phase(self.h, xs, ys, _p(phases), _p(flags), _p(consistency), int(w[0]), int(w[1])))

    def reconstruct_phases(self, xsize: int, ysize: int, reco_radius: float):
        ph = np.zeros((ysize, xsize), dtype=np.complex128)
        fl = np.zeros((ysize, xsize), dtype=np.uint8)
        co = np.zeros((ysize, xsize), dtype=np.float64)
        _check(lib().smip_reconstruct_phases(self.h, xsize, ysize, reco_radius, _p(ph), _p(fl), _p(co)))
        return ph, fl, co


def next_reco_index(r: float, phi: float, i: int, j: int):
    rr, pp, ii, jj = C.c_double(r), C.c_double(phi), C.c_int32(i), C.c_int32(j)
    lib().smip_next_reco_index(C.byref(rr), C.byref(pp), C.byref(ii), C.byref(jj))
    return rr.value, pp.value, ii.value, jj.value


def general_hamming(xsize: int, ysize: int, aperture: float, alpha: float) -> np.ndarray:
    out = np.zeros((ysize, xsize))
    _check(lib().smip_general_hamming(xsize, ysize, aperture, alpha, _p(out)))
    return out


class CrossCorrelation:
    """smip::CrossCorrelation<double> (include/crosscorrel.h:43-189)."""

    def __init__(self, ctx: Context, ref: np.ndarray):
        a = np.ascontiguousarray(ref, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().smip_xcorr_create(ctx.h, _p(a), a.shape[1], a.shape[0], C.byref(h)))
        self.h = h
        self.shape = a.shape
        ctx._adopt(self)

    def close(self):
        if self.h:
            lib().smip_xcorr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def correlate(self, frame: np.ndarray) -> None:
        a = np.ascontiguousarray(frame, dtype=np.float64)
        _check(lib().smip_xcorr_correlate(self.h, _p(a), a.shape[1], a.shape[0]))

    def get_correlation_array(self) -> np.ndarray:
        out = np.zeros(self.shape)
        _check(lib().smip_xcorr_get_correlation(self.h, _p(out)))
        return out

    def get_displacement(self):
        s = np.zeros(2, dtype=np.int32)
        _check(lib().smip_xcorr_get_displacement(self.h, _p(s)))
        return int(s[0]), int(s[1])

    def __call__(self, frame: np.ndarray):
        a = np.ascontiguousarray(frame, dtype=np.float64)
        s = np.zeros(2, dtype=np.int32)
        _check(lib().smip_xcorr_displacement(self.h, _p(a), a.shape[1], a.shape[0], _p(s)))
        return int(s[0]), int(s[1])


class Frames:
    """A device-resident frame stack."""

    def __init__(self, ctx: Context, frames: np.ndarray | None = None, dtype=None, shape=None):
        if frames is not None:
            frames = np.ascontiguousarray(frames)
            dtype, shape = frames.dtype, frames.shape
        n, nrows, ncols = shape
        h = C.c_void_p()
        _check(lib().smip_frames_create(ctx.h, _NP_DTYPES[np.dtype(dtype)], ncols, nrows, n, C.byref(h)))
        self.h = h
        ctx._adopt(self)
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        if frames is not None:
            self.upload(0, frames)

    def upload(self, first: int, frames: np.ndarray) -> None:
        a = np.ascontiguousarray(frames, dtype=self.dtype)
        _check(lib().smip_frames_upload(self.h, first, a.shape[0], _p(a)))

    def close(self):
        if self.h:
            lib().smip_frames_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class JobResult:
    bispectrum: np.ndarray | None
    powerspec: np.ndarray
    sum_image: np.ndarray | None
    shifts: np.ndarray | None
    phases: np.ndarray
    flags: np.ndarray
    consistency: np.ndarray
    image: np.ndarray
    stats: dict


def job_params(ncols, nrows, bdepth, reco_radius, frame_dtype, precision=PRECISION_MIXED, saa=True,
               frames_per_batch=0, vy_tile=0) -> JobParams:
    return JobParams(ncols, nrows, bdepth, float(reco_radius), _NP_DTYPES[np.dtype(frame_dtype)], precision,
                     int(bool(saa)), frames_per_batch, vy_tile)


def finalize_many(jobs: list) -> None:
    """Job.finalize for several jobs at once; their phase recoveries run as one group."""
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    st = (C.c_int32 * len(jobs))()
    _check(lib().smip_job_finalize_many(arr, len(jobs), st))


def add_frames_many(jobs: list, stacks: list, firsts=None, counts=None) -> None:
    """Job.add_frames for several jobs at once: one launch per stage for the current
    batch of every job (smip_job_add_frames_many)."""
    n = len(jobs)
    firsts = [0] * n if firsts is None else list(firsts)
    counts = [s.shape[0] - f for s, f in zip(stacks, firsts)] if counts is None else list(counts)
    ja = (C.c_void_p * n)(*[j.h for j in jobs])
    sa = (C.c_void_p * n)(*[s.h for s in stacks])
    fa = (C.c_uint64 * n)(*firsts)
    ca = (C.c_uint64 * n)(*counts)
    _check(lib().smip_job_add_frames_many(ja, sa, fa, ca, n))


def resolve_params(params: JobParams) -> JobParams:
    """The parameters a job is created with (tuned or built-in defaults filled in)."""
    out = JobParams()
    _check(lib().smip_job_resolve_params(C.byref(params), C.byref(out)))
    return out


def autotune(ctx: Context, ncols: int, nrows: int, bdepth: int, max_bytes: int = 0) -> dict:
    """Times the batched kernel's tile shapes and batch sizes; later jobs of this
    geometry use the choice."""
    r = TuneResult()
    _check(lib().smip_autotune(ctx.h, ncols, nrows, bdepth, max_bytes, C.byref(r)))
    return {k: getattr(r, k) for k, _ in TuneResult._fields_}


def autotune_forget() -> None:
    _check(lib().smip_autotune_forget())


def job_footprint(params: JobParams) -> int:
    b = C.c_uint64()
    _check(lib().smip_job_footprint(C.byref(params), C.byref(b)))
    return b.value


class Job:
    """One speckle-masking reconstruction (smip-cli/main.cpp:242-376)."""

    def __init__(self, ctx: Context, params: JobParams, _handle=None):
        self.borrowed = _handle is not None
        if _handle is None:
            h = C.c_void_p()
            _check(lib().smip_job_create(ctx.h, C.byref(params), C.byref(h)))
            _handle = h
            ctx._adopt(self)
        self.h = _handle
        self.ctx = ctx
        self.p = params

    def close(self):
        if self.h and not self.borrowed:
            lib().smip_job_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_frames(self, frames: Frames, first: int = 0, count: int | None = None) -> None:
        if count is None:
            count = frames.shape[0] - first
        _check(lib().smip_job_add_frames(self.h, frames.h, first, count))

    def add_frames_host(self, frames: np.ndarray) -> None:
        a = np.ascontiguousarray(frames)
        _check(lib().smip_job_add_frames_host(self.h, _p(a), a.shape[0]))

    def push_frames_host(self, frames: np.ndarray) -> None:
        """Streaming form: batch boundaries follow the frame sequence, not the calls."""
        a = np.ascontiguousarray(frames)
        _check(lib().smip_job_push_frames_host(self.h, _p(a), a.shape[0]))

    def finalize(self) -> None:
        _check(lib().smip_job_finalize(self.h))

    def synchronize(self) -> None:
        _check(lib().smip_job_synchronize(self.h))

    def stats(self) -> dict:
        s = JobStats()
        _check(lib().smip_job_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in JobStats._fields_}

    def bispectrum(self) -> Bispectrum:
        h = C.c_void_p()
        _check(lib().smip_job_bispectrum(self.h, C.byref(h)))
        return Bispectrum(self.ctx, _handle=h, _owner=self)

    def powerspec(self) -> np.ndarray:
        out = np.zeros((self.p.nrows, self.p.ncols))
        _check(lib().smip_job_get_powerspec(self.h, _p(out)))
        return out

    def sum_image(self) -> np.ndarray:
        out = np.zeros((self.p.nrows, self.p.ncols))
        _check(lib().smip_job_get_sum_image(self.h, _p(out)))
        return out

    def shifts(self, count: int) -> np.ndarray:
        out = np.zeros((count, 2), dtype=np.int32)
        _check(lib().smip_job_get_shifts(self.h, _p(out), count))
        return out

    def phases(self):
        n = (self.p.nrows, self.p.ncols)
        ph = np.zeros(n, dtype=np.complex128)
        fl = np.zeros(n, dtype=np.uint8)
        co = np.zeros(n)
        _check(lib().smip_job_get_phases(self.h, _p(ph), _p(fl), _p(co)))
        return ph, fl, co

    def image(self) -> np.ndarray:
        out = np.zeros((self.p.nrows, self.p.ncols), dtype=np.complex128)
        _check(lib().smip_job_get_image(self.h, _p(out)))
        return out

    def result(self, with_bispectrum: bool = True) -> JobResult:
        ph, fl, co = self.phases()
        st = self.stats()
        return JobResult(self.bispectrum().download() if with_bispectrum else None, self.powerspec(),
                         self.sum_image() if self.p.saa else None,
                         self.shifts(int(st["frames"])) if self.p.saa else None,
                         ph, fl, co, self.image(), st)


class Server:
    """The continuous-batching request loop (smip_server_*): jobs are admitted by
    device footprint as slots and memory free up, stream frames in chunks and
    overlap on the GPU."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_active: int = 0, workers: int = 0):
        h = C.c_void_p()
        prm = ServerParams(hbm_budget, max_active, workers)
        _check(lib().smip_server_create(ctx.h, C.byref(prm), C.byref(h)))
        self.h = h
        self.ctx = ctx
        ctx._adopt(self)
        self.params: dict[int, JobParams] = {}

    def close(self):
        if self.h:
            lib().smip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, params: JobParams) -> int:
        jid = C.c_uint64()
        _check(lib().smip_server_submit(self.h, C.byref(params), C.byref(jid)))
        self.params[jid.value] = params
        return jid.value

    def push(self, jid: int, frames: np.ndarray) -> None:
        a = np.ascontiguousarray(frames)
        _check(lib().smip_server_push(self.h, jid, _p(a), a.shape[0]))

    def close_job(self, jid: int) -> None:
        _check(lib().smip_server_close(self.h, jid))

    def wait(self, jid: int, timeout: float = -1.0) -> int:
        st = C.c_int32()
        _check(lib().smip_server_wait(self.h, jid, timeout, C.byref(st)))
        return st.value

    def job(self, jid: int) -> Job:
        h = C.c_void_p()
        _check(lib().smip_server_job(self.h, jid, C.byref(h)))
        return Job(self.ctx, self.params[jid], _handle=h)

    def info(self, jid: int) -> dict:
        i = ServerJobInfo()
        _check(lib().smip_server_get_job_info(self.h, jid, C.byref(i)))
        d = {k: getattr(i, k) for k, _ in ServerJobInfo._fields_}
        d["error"] = i.error.decode()
        return d

    def release(self, jid: int) -> None:
        _check(lib().smip_server_release(self.h, jid))
        self.params.pop(jid, None)

    def stats(self) -> dict:
        s = ServerStats()
        _check(lib().smip_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}


def run_pipeline(ctx: Context, frames: np.ndarray, bdepth: int, reco_radius: float,
                 precision: int = PRECISION_MIXED, saa: bool = True, frames_per_batch: int = 0,
                 vy_tile: int = 0, with_bispectrum: bool = True) -> JobResult:
    """The smip-cli processing of a (n, nrows, ncols) stack on the GPU."""
    n, nrows, ncols = frames.shape
    p = job_params(ncols, nrows, bdepth, reco_radius, frames.dtype, precision, saa, frames_per_batch, vy_tile)
    job = Job(ctx, p)
    try:
        dev = Frames(ctx, frames)
        job.add_frames(dev)
        job.finalize()
        return job.result(with_bispectrum)
    finally:
        job.close()


# --------------------------------------------------------------------------
# synthetic speckle frame stacks (libsmip_synth.so, host only)
# --------------------------------------------------------------------------
class SynthParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("ncols", C.c_int32), ("nrows", C.c_int32), ("dtype", C.c_int32),
                ("object", C.c_int32), ("pupil_diameter", C.c_double), ("d_over_r0", C.c_double),
                ("photons", C.c_double), ("read_noise", C.c_double), ("clip_max", C.c_double)]


OBJ_BINARY, OBJ_TRIPLE, OBJ_EXTENDED = 0, 1, 2

# The workloads of the measured configurations (frame count, frame side, sample
# type, object, pupil, D/r0, photons, read noise, clip, b, p).
WORKLOADS = {
    "binary128": dict(n=40, N=128, dtype=np.uint8, object=OBJ_BINARY, pupil=32, d_r0=10, photons=2e4,
                      read_noise=2.0, clip=255, b=32, p=64, seed=0x5EC1E001),
    "triple256": dict(n=2000, N=256, dtype=np.uint16, object=OBJ_TRIPLE, pupil=64, d_r0=20, photons=1e5,
                      read_noise=2.0, clip=4095, b=48, p=128, seed=0x5EC1E002),
    "extended512": dict(n=10000, N=512, dtype=np.float32, object=OBJ_EXTENDED, pupil=128, d_r0=30,
                        photons=1e6, read_noise=2.0, clip=0, b=64, p=256, seed=0x5EC1E003),
}

_SYNTH: C.CDLL | None = None


def _synth_lib() -> C.CDLL:
    global _SYNTH
    if _SYNTH is None:
        path = os.path.join(LIBDIR, "libsmip_synth.so")
        if not os.path.exists(path):
            raise SmipError(ERR_RUNTIME, f"{path} is not built (run speckle-stacker_amd/build.py)")
        L = C.CDLL(path)
        L.smip_synth_frames.argtypes = [C.POINTER(SynthParams), C.c_uint64, C.c_uint64, C.c_void_p, C.c_int32]
        L.smip_synth_object.argtypes = [C.POINTER(SynthParams), C.c_void_p]
        _SYNTH = L
    return _SYNTH


def synth_params(name: str | None = None, **kw) -> SynthParams:
    w = dict(WORKLOADS[name]) if name else {}
    w.update(kw)
    code = _NP_DTYPES[np.dtype(w["dtype"])]
    return SynthParams(int(w["seed"]), int(w["N"]), int(w["N"]), code, int(w["object"]), float(w["pupil"]),
                       float(w["d_r0"]), float(w["photons"]), float(w["read_noise"]), float(w["clip"]))


def synth_frames(params: SynthParams, first: int, count: int, threads: int | None = None) -> np.ndarray:
    """Frames [first, first+count) of the deterministic stack described by params."""
    dt = {DTYPE_U8: np.uint8, DTYPE_U16: np.uint16, DTYPE_F32: np.float32}[params.dtype]
    out = np.zeros((count, params.nrows, params.ncols), dtype=dt)
    t = threads if threads else min(16, os.cpu_count() or 1)
    rc = _synth_lib().smip_synth_frames(C.byref(params), first, count, _p(out), t)
    if rc != 0:
        raise SmipError(ERR_INVALID_ARGUMENT, "invalid synthetic frame parameters")
    return out


def synth_object(params: SynthParams) -> np.ndarray:
    out = np.zeros((params.nrows, params.ncols))
    if _synth_lib().smip_synth_object(C.byref(params), _p(out)) != 0:
        raise SmipError(ERR_INVALID_ARGUMENT, "invalid synthetic object parameters")
    return out

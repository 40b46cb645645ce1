"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker; the product path never does.

Arrays follow the oracle's conventions: complex data is ``complex128`` (or
``complex64`` for the CLI's ``Bispectrum<complex<float>>``), 2-D arrays are
``(nrows, ncols)`` row-major with x fastest, which is the layout of
``smip::Array2`` (include/array2.h:290-308).

Parity unpinned (stated in smip_oracle.c's header as well): get_element on the
aliased T3/T9 lookups, the NextRecoIndex visit order beyond its invariants, and
displacement maxima that tie to the last bit of the FFT rounding.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

OK, INVALID_ARGUMENT, OUT_OF_RANGE, RUNTIME_ERROR = 0, 1, 2, 3
BAD_FUNCTION_CALL, LOGIC_ERROR, DOMAIN_ERROR = 4, 5, 6


class OracleError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"oracle {where} failed with status {code}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [
        ("dims", C.c_uint64 * 4),
        ("sizes", C.c_uint64 * 4),
        ("base_sizes", C.c_uint64 * 4),
        ("base_size", C.c_uint64),
        ("totalsize", C.c_uint64),
        ("min_idx", C.c_int32 * 4),
        ("max_idx", C.c_int32 * 4),
    ]


class RecoStats(C.Structure):
    _fields_ = [
        ("visits", C.c_int64),
        ("calls", C.c_int64),
        ("lookups", C.c_int64),
        ("aliased", C.c_int64),
        ("flagged", C.c_int64),
    ]


class PipelineOut(C.Structure):
    _fields_ = [
        ("bispectrum", C.c_void_p),
        ("powerspec", C.c_void_p),
        ("sum_image", C.c_void_p),
        ("shifts", C.c_void_p),
        ("phases", C.c_void_p),
        ("flags", C.c_void_p),
        ("consistency", C.c_void_p),
        ("image", C.c_void_p),
        ("stats", RecoStats),
    ]


def _build_if_missing(name: str) -> str:
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return path


_LIBS: dict[str, C.CDLL] = {}


def lib(fast: bool = False) -> C.CDLL:
    name = "libsmip_oracle_fast.so" if fast else "libsmip_oracle.so"
    if name in _LIBS:
        return _LIBS[name]
    L = C.CDLL(_build_if_missing(name))
    P = C.c_void_p
    L.so_desc_init.argtypes = [C.POINTER(Desc), C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]
    L.so_desc_init_empty.argtypes = [C.POINTER(Desc)]
    L.so_calc_offset.argtypes = [C.POINTER(Desc), P]
    L.so_calc_offset.restype = C.c_uint64
    L.so_calc_indices.argtypes = [C.POINTER(Desc), C.c_uint64, P]
    L.so_lookup.argtypes = [C.POINTER(Desc), P, P, P, P]
    L.so_get_element.argtypes = [C.POINTER(Desc), P, C.c_int, P, P]
    L.so_put_element.argtypes = [C.POINTER(Desc), P, C.c_int, P, P]
    L.so_accumulate_f64.argtypes = [C.POINTER(Desc), P, P, C.c_uint64, C.c_uint64]
    L.so_accumulate_f32.argtypes = [C.POINTER(Desc), P, P, C.c_uint64, C.c_uint64]
    L.so_count_triples.argtypes = [C.POINTER(Desc), C.c_uint64, C.c_uint64]
    L.so_count_triples.restype = C.c_uint64
    L.so_fft2_c2c.argtypes = [P, C.c_int64, C.c_int64, C.c_int]
    L.so_fft2_r2c.argtypes = [P, P, C.c_int64, C.c_int64]
    L.so_fft2_c2r.argtypes = [P, P, C.c_int64, C.c_int64]
    L.so_next_reco_index.argtypes = [P, P, P, P]
    L.so_spiral_visits.argtypes = [C.c_double, P, C.c_int64]
    L.so_spiral_visits.restype = C.c_int64
    L.so_reconstruct_phases.argtypes = [C.POINTER(Desc), P, C.c_int, C.c_uint64, C.c_uint64,
                                        C.c_double, P, P, P, C.POINTER(RecoStats)]
    L.so_reconstruct_phases_alias.argtypes = [C.POINTER(Desc), P, C.c_int, C.c_uint64, C.c_uint64,
                                              C.c_double, P, P, P, C.POINTER(RecoStats), C.c_int]
    L.so_xcorr.argtypes = [P, P, C.c_int64, C.c_int64, P]
    L.so_displacement.argtypes = [P, C.c_int64, C.c_int64, P]
    L.so_shifted.argtypes = [P, C.c_int64, C.c_int64, C.c_int32, C.c_int32, P]
    L.so_general_hamming.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, P]
    L.so_pipeline.argtypes = [P, C.c_int64, C.c_int64, C.c_int64, C.c_uint64, C.c_double,
                              C.c_int, C.POINTER(PipelineOut)]
    L.so_power_and_bispectrum.argtypes = [P, C.c_int64, C.c_int64, C.c_int64, C.c_uint64,
                                          C.c_int, P, P]
    L.so_saa.argtypes = [P, C.c_int64, C.c_int64, C.c_int64, P, P]
    L.so_synthesize.argtypes = [P, P, C.c_int64, C.c_int64, C.c_double, P]
    _LIBS[name] = L
    return L


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int, where: str) -> None:
    if rc != OK:
        raise OracleError(rc, where)


# --------------------------------------------------------------------------
# Bispectrum descriptor and element access (include/bispectrum.h)
# --------------------------------------------------------------------------
def desc(ncols: int, nrows: int, b0: int, b1: int | None = None) -> Desc:
    d = Desc()
    lib().so_desc_init(C.byref(d), ncols, nrows, b0, b0 if b1 is None else b1)
    return d


def empty_desc() -> Desc:
    d = Desc()
    lib().so_desc_init_empty(C.byref(d))
    return d


def calc_offset(d: Desc, idx) -> int:
    a = np.asarray(idx, dtype=np.int32)
    return int(lib().so_calc_offset(C.byref(d), _p(a)))


def calc_indices(d: Desc, addr: int) -> tuple[int, int, int, int]:
    out = np.zeros(4, dtype=np.int32)
    lib().so_calc_indices(C.byref(d), addr, _p(out))
    return tuple(int(x) for x in out)


def lookup(d: Desc, idx):
    """(address, conjugate, aliased) of get_element's lookup; raises on out_of_range."""
    a = np.asarray(idx, dtype=np.int32)
    addr = C.c_uint64()
    conj = C.c_int()
    alias = C.c_int()
    rc = lib().so_lookup(C.byref(d), _p(a), C.byref(addr), C.byref(conj), C.byref(alias))
    _check(rc, "get_element")
    return int(addr.value), bool(conj.value), bool(alias.value)


def get_element(d: Desc, data: np.ndarray, idx) -> complex:
    a = np.asarray(idx, dtype=np.int32)
    out = np.zeros(2)
    rc = lib().so_get_element(C.byref(d), _p(data), int(data.dtype == np.complex64), _p(a), _p(out))
    _check(rc, "get_element")
    return complex(out[0], out[1])


def put_element(d: Desc, data: np.ndarray, idx, value: complex) -> None:
    a = np.asarray(idx, dtype=np.int32)
    v = np.array([value.real, value.imag])
    rc = lib().so_put_element(C.byref(d), _p(data), int(data.dtype == np.complex64), _p(a), _p(v))
    _check(rc, "put_element")


def new_bispectrum(d: Desc, dtype=np.complex128) -> np.ndarray:
    return np.zeros(int(d.base_size), dtype=dtype)


def accumulate_from_fft(d: Desc, bispec: np.ndarray, fft: np.ndarray) -> None:
    """Bispectrum<T>::accumulate_from_fft (include/bispectrum.h:356-409).

    fft: (nrows, ncols) complex128 in FFTW output order."""
    fft = np.ascontiguousarray(fft, dtype=np.complex128)
    nrows, ncols = fft.shape
    if bispec.dtype == np.complex64:
        rc = lib().so_accumulate_f32(C.byref(d), _p(bispec), _p(fft), ncols, nrows)
    else:
        rc = lib().so_accumulate_f64(C.byref(d), _p(bispec), _p(fft), ncols, nrows)
    _check(rc, "accumulate_from_fft")


def count_triples(d: Desc, ncols: int, nrows: int) -> int:
    return int(lib().so_count_triples(C.byref(d), ncols, nrows))


# --------------------------------------------------------------------------
# FFT (FFTW conventions)
# --------------------------------------------------------------------------
def fft2(x: np.ndarray, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    nrows, ncols = a.shape
    lib().so_fft2_c2c(_p(a), nrows, ncols, sign)
    return a


def rfft2(x: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.float64)
    nrows, ncols = a.shape
    out = np.zeros((nrows, ncols // 2 + 1), dtype=np.complex128)
    lib().so_fft2_r2c(_p(a), _p(out), nrows, ncols)
    return out


def irfft2(h: np.ndarray, ncols: int) -> np.ndarray:
    a = np.ascontiguousarray(h, dtype=np.complex128)
    nrows = a.shape[0]
    out = np.zeros((nrows, ncols), dtype=np.float64)
    lib().so_fft2_c2r(_p(a), _p(out), nrows, ncols)
    return out


# --------------------------------------------------------------------------
# Phase reconstruction (include/phasereco.h, src/phasereco.cpp)
# --------------------------------------------------------------------------
def next_reco_index(r: float, phi: float, i: int, j: int):
    rr, pp = C.c_double(r), C.c_double(phi)
    ii, jj = C.c_int32(i), C.c_int32(j)
    lib().so_next_reco_index(C.byref(rr), C.byref(pp), C.byref(ii), C.byref(jj))
    return rr.value, pp.value, ii.value, jj.value


def spiral_visits(reco_radius: float) -> np.ndarray:
    n = int(lib().so_spiral_visits(reco_radius, None, 0))
    out = np.zeros((n, 2), dtype=np.int32)
    lib().so_spiral_visits(reco_radius, _p(out), n)
    return out


@dataclass
class Reco:
    phases: np.ndarray       # (ysize, xsize) complex128
    flags: np.ndarray        # (ysize, xsize) uint8
    consistency: np.ndarray  # (ysize, xsize) float64
    stats: dict


def reconstruct_phases(d: Desc, bispec: np.ndarray, xsize: int, ysize: int, reco_radius: float,
                       alias_mode: int = 0) -> Reco:
    """alias_mode 1 leaves the aliased (T3/T9) lookups out of each mean: a
    sensitivity probe, not the reference (alias_mode 0)."""
    ph = np.zeros((ysize, xsize), dtype=np.complex128)
    fl = np.zeros((ysize, xsize), dtype=np.uint8)
    co = np.zeros((ysize, xsize), dtype=np.float64)
    st = RecoStats()
    rc = lib().so_reconstruct_phases_alias(C.byref(d), _p(bispec), int(bispec.dtype == np.complex64),
                                           xsize, ysize, reco_radius, _p(ph), _p(fl), _p(co), C.byref(st),
                                           int(alias_mode))
    _check(rc, "reconstruct_phases")
    stats = {k: int(getattr(st, k)) for k, _ in RecoStats._fields_}
    return Reco(ph, fl, co, stats)


# --------------------------------------------------------------------------
# Cross-correlation, shift-and-add, window, synthesis
# --------------------------------------------------------------------------
def xcorr(ref: np.ndarray, frame: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(ref, dtype=np.float64)
    f = np.ascontiguousarray(frame, dtype=np.float64)
    if r.shape != f.shape:
        raise OracleError(INVALID_ARGUMENT, "xcorr")
    out = np.zeros_like(r)
    lib().so_xcorr(_p(r), _p(f), r.shape[1], r.shape[0], _p(out))
    return out


def displacement(corr: np.ndarray) -> tuple[int, int]:
    c = np.ascontiguousarray(corr, dtype=np.float64)
    s = np.zeros(2, dtype=np.int32)
    lib().so_displacement(_p(c), c.shape[1], c.shape[0], _p(s))
    return int(s[0]), int(s[1])


def shifted(a: np.ndarray, dx: int, dy: int) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros_like(x)
    rc = lib().so_shifted(_p(x), x.shape[1], x.shape[0], dx, dy, _p(out))
    _check(rc, "shifted")
    return out


def general_hamming(xsize: int, ysize: int, aperture: float, alpha: float) -> np.ndarray:
    out = np.zeros((ysize, xsize), dtype=np.float64)
    rc = lib().so_general_hamming(xsize, ysize, aperture, alpha, _p(out))
    _check(rc, "general_hamming")
    return out


def hann(xsize: int, ysize: int, aperture: float) -> np.ndarray:
    return general_hamming(xsize, ysize, aperture, 0.5)


@dataclass
class Pipeline:
    bispectrum: np.ndarray
    powerspec: np.ndarray
    sum_image: np.ndarray
    shifts: np.ndarray
    phases: np.ndarray
    flags: np.ndarray
    consistency: np.ndarray
    image: np.ndarray
    stats: dict


def pipeline(frames: np.ndarray, bdepth: int, reco_radius: float, bispec_f32: bool = False,
             fast: bool = False) -> Pipeline:
    """smip-cli's processing (smip-cli/main.cpp:242-376) on a (n, nrows, ncols) stack."""
    fr = np.ascontiguousarray(frames, dtype=np.float64)
    n, nrows, ncols = fr.shape
    d = desc(ncols, nrows, bdepth)
    B = new_bispectrum(d, np.complex64 if bispec_f32 else np.complex128)
    P = np.zeros((nrows, ncols))
    S = np.zeros((nrows, ncols))
    sh = np.zeros((n, 2), dtype=np.int32)
    ph = np.zeros((nrows, ncols), dtype=np.complex128)
    fl = np.zeros((nrows, ncols), dtype=np.uint8)
    co = np.zeros((nrows, ncols))
    im = np.zeros((nrows, ncols), dtype=np.complex128)
    out = PipelineOut(_p(B), _p(P), _p(S), _p(sh), _p(ph), _p(fl), _p(co), _p(im), RecoStats())
    rc = lib(fast).so_pipeline(_p(fr), n, ncols, nrows, bdepth, reco_radius, int(bispec_f32), C.byref(out))
    _check(rc, "pipeline")
    stats = {k: int(getattr(out.stats, k)) for k, _ in RecoStats._fields_}
    return Pipeline(B, P, S, sh, ph, fl, co, im, stats)


def power_and_bispectrum(frames: np.ndarray, bdepth: int, bispec_f32: bool = False, fast: bool = False):
    fr = np.ascontiguousarray(frames, dtype=np.float64)
    n, nrows, ncols = fr.shape
    d = desc(ncols, nrows, bdepth)
    B = new_bispectrum(d, np.complex64 if bispec_f32 else np.complex128)
    P = np.zeros((nrows, ncols))
    rc = lib(fast).so_power_and_bispectrum(_p(fr), n, ncols, nrows, bdepth, int(bispec_f32), _p(B), _p(P))
    _check(rc, "power_and_bispectrum")
    return B, P


def saa(frames: np.ndarray):
    fr = np.ascontiguousarray(frames, dtype=np.float64)
    n, nrows, ncols = fr.shape
    sh = np.zeros((n, 2), dtype=np.int32)
    S = np.zeros((nrows, ncols))
    rc = lib().so_saa(_p(fr), n, ncols, nrows, _p(sh), _p(S))
    _check(rc, "saa")
    return sh, S


def synthesize(powerspec: np.ndarray, phases: np.ndarray, reco_radius: float) -> np.ndarray:
    P = np.ascontiguousarray(powerspec, dtype=np.float64)
    ph = np.ascontiguousarray(phases, dtype=np.complex128)
    nrows, ncols = P.shape
    out = np.zeros((nrows, ncols), dtype=np.complex128)
    rc = lib().so_synthesize(_p(P), _p(ph), ncols, nrows, reco_radius, _p(out))
    _check(rc, "synthesize")
    return out

"""ctypes wrapper of the oracle (CPU restatement of the reference path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

from replace_amd.netlist import (Netlist, RpEvalOut, RpIterRecord, RpNbInfo, RpNesterovState, RpNesterovStats,
                                 nesterov_state_buffers, nesterov_state_values,
                                 RpParams)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librp_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_poisson.so")

POISSON_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                              ctypes.POINTER(ctypes.c_float))

_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE])


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.ro_create.restype = vp
        L.ro_create.argtypes = [ctypes.POINTER(ctypes.c_byte), ctypes.POINTER(RpParams),
                                ctypes.POINTER(ctypes.c_int)]
        L.ro_destroy.argtypes = [vp]
        L.ro_set_poisson.argtypes = [vp, vp, vp]
        L.ro_initial_place.argtypes = [vp, ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int)]
        L.ro_nesterov_setup.argtypes = [vp, ctypes.POINTER(RpNbInfo)]
        L.ro_get_gcells.argtypes = [vp] + [vp] * 8
        L.ro_get_nonplace.argtypes = [vp, vp]
        L.ro_eval_gradient.argtypes = [vp, vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.POINTER(RpEvalOut)]
        L.ro_eval_poisson.argtypes = [vp, vp, vp, vp, vp]
        L.ro_nesterov_place.argtypes = [vp, ctypes.POINTER(RpNesterovStats)]
        L.ro_nesterov_init.argtypes = [vp]
        L.ro_nesterov_iterate.argtypes = [vp, ctypes.c_int, ctypes.POINTER(RpNesterovStats)]
        L.ro_get_state.argtypes = [vp, ctypes.POINTER(RpNesterovState)]
        L.ro_get_stage_times.argtypes = [vp, vp]
        L.ro_set_state.argtypes = [vp, ctypes.POINTER(RpNesterovState)]
        L.ro_get_iter_records.argtypes = [vp, ctypes.POINTER(RpIterRecord), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int32)]
        L.ro_get_positions.argtypes = [vp, vp, vp]
        L.ro_set_positions.argtypes = [vp, vp, vp]
        L.ro_inst_hpwl.restype = ctypes.c_int64
        L.ro_inst_hpwl.argtypes = [vp]
        L.ro_num_insts.argtypes = [vp]
        L.ro_mt19937.argtypes = [ctypes.c_uint32, vp, ctypes.c_int]
        L.ro_fast_exp.restype = ctypes.c_float
        L.ro_fast_exp.argtypes = [ctypes.c_float]
        L.ro_poisson_direct64.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, vp, vp, vp, vp]
        L.ro_poisson_fft64.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, vp, vp, vp, vp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_lib() -> Optional[ctypes.CDLL]:
    """The reference's own Poisson solver built into oracle/_ref (or None)."""
    global _ref
    if _ref is None and ref_available():
        _ref = ctypes.CDLL(REF_PATH)
        _ref.ref_poisson.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
    return _ref


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def poisson_direct64(rho: np.ndarray, bsx: int, bsy: int):
    ny, nx = rho.shape
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    out = [np.zeros((ny, nx), np.float64) for _ in range(3)]
    lib().ro_poisson_direct64(None, nx, ny, bsx, bsy, _p(rho), *(_p(o) for o in out))
    return out


def poisson_fft64(rho: np.ndarray, bsx: int, bsy: int):
    ny, nx = rho.shape
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    out = [np.zeros((ny, nx), np.float32) for _ in range(3)]
    lib().ro_poisson_fft64(None, nx, ny, bsx, bsy, _p(rho), *(_p(o) for o in out))
    return out


def poisson_reference(rho: np.ndarray, bsx: int, bsy: int):
    r = ref_lib()
    if r is None:
        return None
    ny, nx = rho.shape
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    out = [np.zeros((ny, nx), np.float32) for _ in range(3)]
    r.ref_poisson(None, nx, ny, bsx, bsy, _p(rho), *(_p(o) for o in out))
    return out


def mt19937(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().ro_mt19937(seed, _p(out), n)
    return out


def fast_exp(a: float) -> float:
    return lib().ro_fast_exp(a)


class Oracle:
    """One placement problem on the CPU restatement."""

    def __init__(self, netlist: Netlist, params: RpParams, poisson: str = "auto"):
        self._nl = netlist
        self._cnl = netlist.to_c()
        self.params = params
        code = ctypes.c_int(0)
        L = lib()
        self.h = L.ro_create(ctypes.cast(ctypes.byref(self._cnl), ctypes.POINTER(ctypes.c_byte)),
                             ctypes.byref(params), ctypes.byref(code))
        self.create_code = code.value
        self.poisson = "fft64"
        if poisson in ("auto", "reference") and ref_available():
            r = ref_lib()
            L.ro_set_poisson(self.h, ctypes.cast(r.ref_poisson, ctypes.c_void_p), None)
            self.poisson = "reference"
        elif poisson == "reference":
            raise RuntimeError("reference Poisson build (oracle/_ref) not present")
        self.info = None

    def close(self):
        if self.h:
            lib().ro_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initial_place(self, cap: int = 64):
        err = np.zeros(cap, np.float32)
        hp = np.zeros(cap, np.int64)
        n = ctypes.c_int(0)
        lib().ro_initial_place(self.h, err.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               hp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap,
                               ctypes.byref(n))
        return err[: n.value], hp[: n.value]

    def setup(self) -> RpNbInfo:
        info = RpNbInfo()
        r = lib().ro_nesterov_setup(self.h, ctypes.byref(info))
        if r:
            raise RuntimeError(f"nesterov setup failed with code {r}")
        self.info = info
        return info

    def gcells(self):
        n = self.info.num_gcells
        d = dict(dhalf_x=np.zeros(n, np.int32), dhalf_y=np.zeros(n, np.int32),
                 scale=np.zeros(n, np.float32), dx=np.zeros(n, np.int32), dy=np.zeros(n, np.int32),
                 init_cx=np.zeros(n, np.float32), init_cy=np.zeros(n, np.float32),
                 kind=np.zeros(n, np.uint8))
        lib().ro_get_gcells(self.h, *(_p(d[k]) for k in ("dhalf_x", "dhalf_y", "scale", "dx",
                                                         "dy", "init_cx", "init_cy", "kind")))
        return d

    def nonplace(self):
        out = np.zeros(self.info.bin_cnt_x * self.info.bin_cnt_y, np.int64)
        lib().ro_get_nonplace(self.h, _p(out))
        return out

    def eval_gradient(self, cx, cy, lam, wlx, wly):
        from replace_amd.capi import alloc_eval_out
        out, bufs = alloc_eval_out(self.info)
        cx = np.ascontiguousarray(cx, np.float32)
        cy = np.ascontiguousarray(cy, np.float32)
        r = lib().ro_eval_gradient(self.h, _p(cx), _p(cy), lam, wlx, wly, ctypes.byref(out))
        if r:
            raise RuntimeError(f"oracle eval failed {r}")
        return bufs

    def nesterov_place(self):
        st = RpNesterovStats()
        code = lib().ro_nesterov_place(self.h, ctypes.byref(st))
        return code, st

    def nesterov_init(self):
        return lib().ro_nesterov_init(self.h)

    def nesterov_iterate(self, n):
        st = RpNesterovStats()
        code = lib().ro_nesterov_iterate(self.h, n, ctypes.byref(st))
        return code, st

    STAGES = ("step", "scatter", "poisson", "wa", "grad", "step_length")

    def stage_times(self) -> dict:
        """Seconds per Nesterov-loop stage so far (the restatement's own
        instrumentation; see ro_get_stage_times)."""
        out = np.zeros(len(self.STAGES), np.float64)
        lib().ro_get_stage_times(self.h, _p(out))
        return dict(zip(self.STAGES, out.tolist()))

    def get_state(self) -> dict:
        st, bufs = nesterov_state_buffers(self.info.num_gcells)
        r = lib().ro_get_state(self.h, ctypes.byref(st))
        if r:
            raise RuntimeError(f"oracle get_state failed {r}")
        return nesterov_state_values(st, bufs)

    def set_state(self, values: dict) -> None:
        st, _bufs = nesterov_state_buffers(self.info.num_gcells, values)
        r = lib().ro_set_state(self.h, ctypes.byref(st))
        if r:
            raise RuntimeError(f"oracle set_state failed {r}")

    def records(self):
        cnt = ctypes.c_int32(0)
        lib().ro_get_iter_records(self.h, None, 0, ctypes.byref(cnt))
        arr = (RpIterRecord * max(cnt.value, 1))()
        lib().ro_get_iter_records(self.h, arr, cnt.value, ctypes.byref(cnt))
        return [(r.overflow, r.hpwl, r.step_length, r.density_penalty, r.num_back_track)
                for r in arr[: cnt.value]]

    def positions(self):
        n = lib().ro_num_insts(self.h)
        lx = np.zeros(n, np.int32)
        ly = np.zeros(n, np.int32)
        lib().ro_get_positions(self.h, _p(lx), _p(ly))
        return lx, ly

    def inst_hpwl(self) -> int:
        return lib().ro_inst_hpwl(self.h)

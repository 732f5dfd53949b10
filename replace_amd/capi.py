"""ctypes binding of the product C ABI (include/replace/rp_capi.h).

``libreplace_amd.so`` (built by replace_amd/build.py with hipcc for gfx950)
is the only compute path: if it is missing this module raises, there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from .netlist import (Netlist, RpEvalOut, RpEvalRecord, RpIpStats, RpIterRecord, RpJobResult, RpNbInfo, RpNesterovState, RpPbInfo,
                      RpNesterovStats, RpParams, RpServerOpts, RpServerStats, RpTileChoice, nesterov_state_buffers,
                      nesterov_state_values, params_default)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libreplace_amd.so")

_lib = None

# return codes of include/replace/rp_capi.h
RP_OK = 0
RP_ERR_ARG = -1
RP_ERR_HIP = -2
RP_ERR_NOMEM = -3
RP_ERR_STATE = -4
RP_ERR_NO_DEVICE = -5


class ReplaceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rp error {code}: {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run python -c 'import __graft_entry__ as g; g.build()'")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.rp_params_default.argtypes = [ctypes.POINTER(RpParams)]
        L.rp_create.argtypes = [vp, ctypes.POINTER(RpParams), ctypes.POINTER(vp)]
        L.rp_initial_place.argtypes = [vp, ctypes.POINTER(RpIpStats)]
        L.rp_nesterov_place.argtypes = [vp, ctypes.POINTER(RpNesterovStats)]
        L.rp_nesterov_init.argtypes = [vp]
        L.rp_get_pb_info.argtypes = [vp, ctypes.POINTER(RpPbInfo)]
        L.rp_get_eval_records.argtypes = [vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.rp_nesterov_iterate.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(RpNesterovStats)]
        L.rp_nesterov_get_state.argtypes = [vp, ctypes.POINTER(RpNesterovState)]
        L.rp_nesterov_set_state.argtypes = [vp, ctypes.POINTER(RpNesterovState)]
        L.rp_get_iter_records.argtypes = [vp, ctypes.POINTER(RpIterRecord), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int32)]
        L.rp_get_positions.argtypes = [vp, vp, vp]
        L.rp_get_gcell_centres.argtypes = [vp, vp, vp]
        L.rp_set_positions.argtypes = [vp, vp, vp]
        L.rp_destroy.argtypes = [vp]
        L.rp_last_error.restype = ctypes.c_char_p
        L.rp_last_error.argtypes = [vp]
        L.rp_nesterov_setup.argtypes = [vp, ctypes.POINTER(RpNbInfo)]
        L.rp_get_gcells.argtypes = [vp] + [vp] * 8
        L.rp_get_nonplace.argtypes = [vp, vp]
        L.rp_eval_gradient.argtypes = [vp, vp, vp, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.POINTER(RpEvalOut)]
        L.rp_eval_poisson.argtypes = [vp, vp, vp, vp, vp]
        L.rp_set_kernel_timing.argtypes = [vp, ctypes.c_int]
        L.rp_get_kernel_timing.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int64)]
        L.rp_get_ip_records.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.rp_poisson_solve_noalloc.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               vp, vp, vp, vp]
        L.rp_server_opts_default.argtypes = [ctypes.POINTER(RpServerOpts)]
        L.rp_server_tuning.argtypes = [vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.rp_server_set_tiles.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.rp_server_create.argtypes = [ctypes.POINTER(RpServerOpts), ctypes.POINTER(vp)]
        L.rp_server_submit.argtypes = [vp, vp, ctypes.POINTER(RpParams), ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_int32)]
        L.rp_server_run.argtypes = [vp]
        L.rp_server_job.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(RpJobResult)]
        L.rp_server_job_positions.argtypes = [vp, ctypes.c_int32, vp, vp]
        L.rp_server_job_records.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(RpIterRecord),
                                            ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.rp_server_get_stats.argtypes = [vp, ctypes.POINTER(RpServerStats)]
        L.rp_server_last_error.restype = ctypes.c_char_p
        L.rp_server_last_error.argtypes = [vp]
        L.rp_server_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def poisson_solve_noalloc(density_ptr: int, nx: int, ny: int, bin_size_x: int, bin_size_y: int,
                          phi_ptr: Optional[int], ex_ptr: int, ey_ptr: int, stream: Optional[int] = None) -> int:
    """rp_poisson_solve_noalloc on device buffers given by address (queued on
    `stream`, a hipStream_t handle; not waited for).  Returns the C code."""
    return lib().rp_poisson_solve_noalloc(ctypes.c_void_p(density_ptr), nx, ny, bin_size_x, bin_size_y,
                                          ctypes.c_void_p(phi_ptr) if phi_ptr else None, ctypes.c_void_p(ex_ptr),
                                          ctypes.c_void_p(ey_ptr), ctypes.c_void_p(stream) if stream else None)


def alloc_eval_out(info: RpNbInfo):
    """Host buffers for every ``rp_eval_out`` field, plus the C view."""
    nb = info.bin_cnt_x * info.bin_cnt_y
    ng, nn = info.num_gcells, info.num_gnets
    bufs = dict(
        bin_inst=np.zeros(nb, np.int64), bin_filler=np.zeros(nb, np.int64),
        density=np.zeros(nb, np.float32), phi=np.zeros(nb, np.float32),
        ex=np.zeros(nb, np.float32), ey=np.zeros(nb, np.float32),
        overflow_area=np.zeros(1, np.int64), net_sums=np.zeros(8 * nn, np.float32),
        net_box=np.zeros(4 * nn, np.int32), wl_grad=np.zeros(2 * ng, np.float32),
        dens_grad=np.zeros(2 * ng, np.float32), sum_grad=np.zeros(2 * ng, np.float32),
        sums=np.zeros(3, np.float64), hpwl=np.zeros(1, np.int64),
    )
    out = RpEvalOut()
    for name, ctype in RpEvalOut._fields_:
        setattr(out, name, bufs[name].ctypes.data_as(ctype))
    return out, bufs


class Placer:
    """One placement context on the GPU (owns an ``rp_ctx``)."""

    def __init__(self, netlist: Netlist, params: Optional[RpParams] = None):
        self._nl = netlist
        self._cnl = netlist.to_c()
        self.params = params if params is not None else params_default()
        h = ctypes.c_void_p()
        code = lib().rp_create(ctypes.byref(self._cnl), ctypes.byref(self.params),
                               ctypes.byref(h))
        self.h = h
        self.create_code = code
        if code < 0:
            raise ReplaceError(code, self.last_error())
        self.info: Optional[RpNbInfo] = None

    def last_error(self) -> str:
        m = lib().rp_last_error(self.h)
        return m.decode() if m else ""

    def _check(self, r: int) -> int:
        if r < 0:
            raise ReplaceError(r, self.last_error())
        return r

    def close(self) -> None:
        if self.h:
            lib().rp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initial_place(self) -> RpIpStats:
        st = RpIpStats()
        self._check(lib().rp_initial_place(self.h, ctypes.byref(st)))
        return st

    def setup(self) -> RpNbInfo:
        info = RpNbInfo()
        r = self._check(lib().rp_nesterov_setup(self.h, ctypes.byref(info)))
        if r:
            raise ReplaceError(r, "nesterov setup")
        self.info = info
        return info

    def gcells(self):
        n = self.info.num_gcells
        d = dict(dhalf_x=np.zeros(n, np.int32), dhalf_y=np.zeros(n, np.int32),
                 scale=np.zeros(n, np.float32), dx=np.zeros(n, np.int32), dy=np.zeros(n, np.int32),
                 init_cx=np.zeros(n, np.float32), init_cy=np.zeros(n, np.float32),
                 kind=np.zeros(n, np.uint8))
        self._check(lib().rp_get_gcells(self.h, *(_p(d[k]) for k in (
            "dhalf_x", "dhalf_y", "scale", "dx", "dy", "init_cx", "init_cy", "kind"))))
        return d

    def nonplace(self):
        out = np.zeros(self.info.bin_cnt_x * self.info.bin_cnt_y, np.int64)
        self._check(lib().rp_get_nonplace(self.h, _p(out)))
        return out

    def eval_gradient(self, cx, cy, lam, wlx, wly):
        out, bufs = alloc_eval_out(self.info)
        cx = np.ascontiguousarray(cx, np.float32)
        cy = np.ascontiguousarray(cy, np.float32)
        self._check(lib().rp_eval_gradient(self.h, _p(cx), _p(cy), lam, wlx, wly,
                                           ctypes.byref(out)))
        return bufs

    def eval_poisson(self, density):
        nb = self.info.bin_cnt_x * self.info.bin_cnt_y
        d = np.ascontiguousarray(density, np.float32).reshape(-1)
        outs = [np.zeros(nb, np.float32) for _ in range(3)]
        self._check(lib().rp_eval_poisson(self.h, _p(d), *(_p(o) for o in outs)))
        return outs

    def nesterov_place(self):
        st = RpNesterovStats()
        code = self._check(lib().rp_nesterov_place(self.h, ctypes.byref(st)))
        return code, st

    def pb_info(self) -> RpPbInfo:
        info = RpPbInfo()
        self._check(lib().rp_get_pb_info(self.h, ctypes.byref(info)))
        return info

    def eval_records(self):
        """Verbose-3 trace: one dict per gradient evaluation of the last run."""
        n = ctypes.c_int32(0)
        self._check(lib().rp_get_eval_records(self.h, None, 0, ctypes.byref(n)))
        arr = (RpEvalRecord * max(1, n.value))()
        self._check(lib().rp_get_eval_records(self.h, arr, n.value, ctypes.byref(n)))
        return [{k: getattr(r, k) for k, _ in RpEvalRecord._fields_} for r in arr[: n.value]]

    def nesterov_init(self) -> int:
        return self._check(lib().rp_nesterov_init(self.h))

    def nesterov_iterate(self, n: int):
        st = RpNesterovStats()
        code = self._check(lib().rp_nesterov_iterate(self.h, int(n), ctypes.byref(st)))
        return code, st

    def get_state(self) -> dict:
        """The loop state between two iterations (rp_nesterov_state) as a dict."""
        st, bufs = nesterov_state_buffers(self.info.num_gcells)
        self._check(lib().rp_nesterov_get_state(self.h, ctypes.byref(st)))
        return nesterov_state_values(st, bufs)

    def set_state(self, values: dict) -> None:
        st, _bufs = nesterov_state_buffers(self.info.num_gcells, values)
        self._check(lib().rp_nesterov_set_state(self.h, ctypes.byref(st)))

    def records(self):
        cnt = ctypes.c_int32(0)
        self._check(lib().rp_get_iter_records(self.h, None, 0, ctypes.byref(cnt)))
        arr = (RpIterRecord * max(cnt.value, 1))()
        self._check(lib().rp_get_iter_records(self.h, arr, cnt.value, ctypes.byref(cnt)))
        return [(r.overflow, r.hpwl, r.step_length, r.density_penalty, r.num_back_track)
                for r in arr[: cnt.value]]

    KERNELS = ("step", "scatter", "pin_pos", "binfin_rowdct", "mfma_poisson", "col_spectral", "row_inverse",
               "wa", "grad", "control", "sort_refresh")

    def set_kernel_timing(self, on: bool) -> None:
        self._check(lib().rp_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing(self):
        """{kernel: (total_ms, launches)} of the slots run since timing was enabled."""
        tot = np.zeros(len(self.KERNELS), np.float64)
        n = ctypes.c_int64(0)
        self._check(lib().rp_get_kernel_timing(self.h, _p(tot), ctypes.byref(n)))
        return {k: (float(t), int(n.value)) for k, t in zip(self.KERNELS, tot)}

    def ip_records(self):
        n = ctypes.c_int32(0)
        self._check(lib().rp_get_ip_records(self.h, None, None, 0, ctypes.byref(n)))
        err = np.zeros(max(1, n.value), np.float32)
        hp = np.zeros(max(1, n.value), np.int64)
        self._check(lib().rp_get_ip_records(self.h, _p(err), _p(hp), n.value, ctypes.byref(n)))
        return err[: n.value], hp[: n.value]

    def gcell_centres(self):
        """Density centres of every gcell (reference gcell order) of the last
        evaluated iterate."""
        n = self.info.num_gcells
        cx = np.zeros(n, np.float32)
        cy = np.zeros(n, np.float32)
        self._check(lib().rp_get_gcell_centres(self.h, _p(cx), _p(cy)))
        return cx, cy

    def positions(self):
        n = self._nl.num_insts
        lx = np.zeros(n, np.int32)
        ly = np.zeros(n, np.int32)
        self._check(lib().rp_get_positions(self.h, _p(lx), _p(ly)))
        return lx, ly


class Server:
    """Placement service: many jobs in continuous batches on one GPU."""

    def __init__(self, hbm_budget_bytes=None, max_active=None, setup_threads=None,
                 steps_per_poll=None, use_graph=None, autotune=None, batched_ip=None):
        o = RpServerOpts()
        lib().rp_server_opts_default(ctypes.byref(o))
        for k, v in dict(hbm_budget_bytes=hbm_budget_bytes, max_active=max_active,
                         setup_threads=setup_threads, steps_per_poll=steps_per_poll,
                         use_graph=use_graph, autotune=autotune, batched_ip=batched_ip).items():
            if v is not None:
                setattr(o, k, int(v))
        self.opts = o
        h = ctypes.c_void_p()
        r = lib().rp_server_create(ctypes.byref(o), ctypes.byref(h))
        self.h = h
        if r < 0:
            raise ReplaceError(r, self.last_error())
        self._nl = {}

    def last_error(self) -> str:
        m = lib().rp_server_last_error(self.h)
        return m.decode() if m else ""

    def submit(self, netlist: Netlist, params: RpParams, arrival_s: float = 0.0) -> int:
        jid = ctypes.c_int32(0)
        cnl = netlist.to_c()
        r = lib().rp_server_submit(self.h, ctypes.byref(cnl), ctypes.byref(params),
                                   float(arrival_s), ctypes.byref(jid))
        if r < 0:
            raise ReplaceError(r, self.last_error())
        self._nl[jid.value] = netlist
        return jid.value

    def run(self) -> None:
        r = lib().rp_server_run(self.h)
        if r < 0:
            raise ReplaceError(r, self.last_error())

    def result(self, job: int) -> RpJobResult:
        out = RpJobResult()
        lib().rp_server_job(self.h, job, ctypes.byref(out))
        return out

    def positions(self, job: int):
        n = self._nl[job].num_insts
        lx = np.zeros(n, np.int32)
        ly = np.zeros(n, np.int32)
        r = lib().rp_server_job_positions(self.h, job, _p(lx), _p(ly))
        if r < 0:
            raise ReplaceError(r, self.last_error())
        return lx, ly

    def records(self, job: int):
        cnt = ctypes.c_int32(0)
        lib().rp_server_job_records(self.h, job, None, 0, ctypes.byref(cnt))
        arr = (RpIterRecord * max(cnt.value, 1))()
        lib().rp_server_job_records(self.h, job, arr, cnt.value, ctypes.byref(cnt))
        return [(r.overflow, r.hpwl, r.step_length, r.density_penalty, r.num_back_track)
                for r in arr[: cnt.value]]

    def set_tiles(self, bucket: int, scatter_cells: int, wa_pins: int) -> None:
        r = lib().rp_server_set_tiles(self.h, bucket, scatter_cells, wa_pins)
        if r < 0:
            raise ReplaceError(r, self.last_error())

    def tuning(self):
        """Tile shape per job size bucket: list of dicts (rp_tile_choice)."""
        arr = (RpTileChoice * 8)()
        n = ctypes.c_int32(0)
        r = lib().rp_server_tuning(self.h, arr, 8, ctypes.byref(n))
        if r < 0:
            raise ReplaceError(r, self.last_error())
        return [dict(bucket=t.bucket, tuned=t.tuned, jobs_sampled=t.jobs_sampled,
                     scatter_cells=t.scatter_cells, wa_pins=t.wa_pins,
                     scatter_us=[round(v, 2) for v in t.scatter_us], wa_us=[round(v, 2) for v in t.wa_us])
                for t in arr[: n.value]]

    def stats(self) -> RpServerStats:
        st = RpServerStats()
        lib().rp_server_get_stats(self.h, ctypes.byref(st))
        return st

    def close(self) -> None:
        if self.h:
            lib().rp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Host SoA netlist: the data the C ABI's ``rp_netlist`` borrows.

It mirrors what ``PlacerBase::init`` reads out of OpenDB (reference
src/placerBase.cpp:605-758): rows (core area), instances with their fixed
status, signal nets with their pins in net-major order (ITerms first, then
BTerms), ITerm offsets from the master centre and BTerm pin boxes.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

INT_MAX = 2**31 - 1
INT_MIN = -(2**31)

_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class RpNetlist(ctypes.Structure):
    """ctypes mirror of ``rp_netlist`` (include/replace/rp_capi.h)."""

    _fields_ = [
        ("dbu_per_micron", ctypes.c_int32),
        ("die_lx", ctypes.c_int32),
        ("die_ly", ctypes.c_int32),
        ("die_ux", ctypes.c_int32),
        ("die_uy", ctypes.c_int32),
        ("site_w", ctypes.c_int32),
        ("site_h", ctypes.c_int32),
        ("num_rows", ctypes.c_int32),
        ("row_box", _i32p),
        ("num_insts", ctypes.c_int32),
        ("inst_box", _i32p),
        ("inst_fixed", _u8p),
        ("num_nets", ctypes.c_int32),
        ("net_pin_start", _i32p),
        ("num_pins", ctypes.c_int32),
        ("pin_inst", _i32p),
        ("pin_off", _i32p),
        ("pin_box", _i32p),
        ("inst_pin_start", _i32p),
        ("inst_pins", _i32p),
    ]


class RpParams(ctypes.Structure):
    """ctypes mirror of ``rp_params``."""

    _fields_ = [
        ("initial_place_max_iter", ctypes.c_int32),
        ("initial_place_min_diff_length", ctypes.c_int32),
        ("initial_place_max_solver_iter", ctypes.c_int32),
        ("initial_place_max_fanout", ctypes.c_int32),
        ("initial_place_net_weight_scale", ctypes.c_float),
        ("nesterov_place_max_iter", ctypes.c_int32),
        ("bin_grid_cnt_x", ctypes.c_int32),
        ("bin_grid_cnt_y", ctypes.c_int32),
        ("target_overflow", ctypes.c_float),
        ("target_density", ctypes.c_float),
        ("init_density_penalty", ctypes.c_float),
        ("init_wire_length_coef", ctypes.c_float),
        ("min_phi_coef", ctypes.c_float),
        ("max_phi_coef", ctypes.c_float),
        ("reference_hpwl", ctypes.c_float),
        ("incremental_place_mode", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("max_back_track", ctypes.c_int32),
        ("min_preconditioner", ctypes.c_float),
        ("initial_prev_coordi_update_coef", ctypes.c_float),
        ("min_wire_length_force_bar", ctypes.c_float),
        ("poisson_transform", ctypes.c_int32),
    ]


class RpIterRecord(ctypes.Structure):
    _fields_ = [
        ("overflow", ctypes.c_float),
        ("step_length", ctypes.c_float),
        ("density_penalty", ctypes.c_float),
        ("num_back_track", ctypes.c_int32),
        ("hpwl", ctypes.c_int64),
    ]


class RpNesterovStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("evals", ctypes.c_int32),
        ("finished", ctypes.c_int32),
        ("code", ctypes.c_int32),
        ("final_overflow", ctypes.c_float),
        ("final_step_length", ctypes.c_float),
        ("final_density_penalty", ctypes.c_float),
        ("diverged", ctypes.c_int32),
        ("final_hpwl", ctypes.c_int64),
        ("seconds", ctypes.c_double),
    ]


class RpNesterovState(ctypes.Structure):
    """ctypes mirror of ``rp_nesterov_state``."""

    _fields_ = [
        ("cur_slp", _f32p),
        ("cur_slp_grad", _f32p),
        ("prev_slp", _f32p),
        ("prev_slp_grad", _f32p),
        ("cur_coord", _f32p),
        ("step_length", ctypes.c_float),
        ("density_penalty", ctypes.c_float),
        ("wl_coef_x", ctypes.c_float),
        ("wl_coef_y", ctypes.c_float),
        ("base_wl_coef", ctypes.c_float),
        ("cur_a", ctypes.c_float),
        ("sum_overflow", ctypes.c_float),
        ("min_sum_overflow", ctypes.c_float),
        ("hpwl_with_min_sum_overflow", ctypes.c_float),
        ("max_phi_coef", ctypes.c_float),
        ("is_max_phi_coef_changed", ctypes.c_int32),
        ("iter", ctypes.c_int32),
        ("prev_hpwl", ctypes.c_int64),
    ]

    ARRAYS = ("cur_slp", "cur_slp_grad", "prev_slp", "prev_slp_grad", "cur_coord")
    SCALARS = ("step_length", "density_penalty", "wl_coef_x", "wl_coef_y", "base_wl_coef", "cur_a",
               "sum_overflow", "min_sum_overflow", "hpwl_with_min_sum_overflow", "max_phi_coef",
               "is_max_phi_coef_changed", "iter", "prev_hpwl")


def nesterov_state_buffers(num_gcells: int, values: Optional[dict] = None):
    """An ``rp_nesterov_state`` whose arrays point into numpy buffers
    (2 * num_gcells floats each, x/y interleaved), filled from ``values``
    when given.  Returns (state, dict of the buffers and scalars)."""
    st = RpNesterovState()
    d = {}
    for k in RpNesterovState.ARRAYS:
        a = np.zeros(2 * num_gcells, np.float32)
        if values is not None:
            a[:] = values[k]
        d[k] = a
        setattr(st, k, a.ctypes.data_as(_f32p))
    if values is not None:
        for k in RpNesterovState.SCALARS:
            setattr(st, k, values[k])
    return st, d


def nesterov_state_values(st: "RpNesterovState", bufs: dict) -> dict:
    """The filled state as a dict (arrays copied)."""
    out = {k: bufs[k].copy() for k in RpNesterovState.ARRAYS}
    for k in RpNesterovState.SCALARS:
        out[k] = getattr(st, k)
    return out


class RpIpStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("last_error", ctypes.c_float),
        ("last_hpwl", ctypes.c_int64),
        ("seconds", ctypes.c_double),
    ]


class RpNbInfo(ctypes.Structure):
    _fields_ = [
        ("num_gcells", ctypes.c_int32),
        ("num_fillers", ctypes.c_int32),
        ("num_place_insts", ctypes.c_int32),
        ("num_gnets", ctypes.c_int32),
        ("num_gpins", ctypes.c_int32),
        ("bin_cnt_x", ctypes.c_int32),
        ("bin_cnt_y", ctypes.c_int32),
        ("bin_size_x", ctypes.c_int32),
        ("bin_size_y", ctypes.c_int32),
        ("core_lx", ctypes.c_int32),
        ("core_ly", ctypes.c_int32),
        ("core_ux", ctypes.c_int32),
        ("core_uy", ctypes.c_int32),
        ("filler_dx", ctypes.c_int32),
        ("filler_dy", ctypes.c_int32),
        ("std_insts_area", ctypes.c_int64),
        ("macro_insts_area", ctypes.c_int64),
        ("place_insts_area", ctypes.c_int64),
        ("nonplace_insts_area", ctypes.c_int64),
        ("target_density", ctypes.c_float),
        ("core_area", ctypes.c_int64),
        ("white_space_area", ctypes.c_int64),
        ("movable_area", ctypes.c_int64),
        ("total_filler_area", ctypes.c_int64),
        ("average_place_inst_area", ctypes.c_int64),
        ("ideal_bin_area", ctypes.c_int64),
        ("ideal_bin_count", ctypes.c_int64),
        ("total_bin_area", ctypes.c_int64),
    ]


class RpPbInfo(ctypes.Structure):
    """ctypes mirror of ``rp_pb_info``."""

    _fields_ = [(k, ctypes.c_int32) for k in (
        "dbu_per_micron", "site_w", "site_h", "die_lx", "die_ly", "die_ux", "die_uy",
        "core_lx", "core_ly", "core_ux", "core_uy", "num_insts", "num_place_insts", "num_fixed_insts",
        "num_dummy_insts", "num_nets", "num_pins", "num_unplaced_ports")] + [
        (k, ctypes.c_int64) for k in ("core_area", "nonplace_insts_area", "place_insts_area", "std_insts_area",
                                      "macro_insts_area")] + [("util", ctypes.c_float)]


class RpEvalRecord(ctypes.Structure):
    """ctypes mirror of ``rp_eval_record``."""

    _fields_ = [("iter", ctypes.c_int32), ("trial", ctypes.c_int32), ("diverged", ctypes.c_int32)] + [
        (k, ctypes.c_float) for k in ("density_penalty", "wl_coef", "wl_grad_sum", "density_grad_sum", "grad_sum",
                                      "coordinate_distance", "gradient_distance", "new_step_length",
                                      "gradient_norm", "sum_phi", "overflow")] + [("hpwl", ctypes.c_int64)]




class RpEvalOut(ctypes.Structure):
    _fields_ = [
        ("bin_inst", _i64p),
        ("bin_filler", _i64p),
        ("density", _f32p),
        ("phi", _f32p),
        ("ex", _f32p),
        ("ey", _f32p),
        ("overflow_area", _i64p),
        ("net_sums", _f32p),
        ("net_box", _i32p),
        ("wl_grad", _f32p),
        ("dens_grad", _f32p),
        ("sum_grad", _f32p),
        ("sums", _f64p),
        ("hpwl", _i64p),
    ]


def params_default(**overrides) -> RpParams:
    """``rp_params_default``: the Replace constructor defaults (src/replace.cpp:13-33)."""
    p = RpParams(
        initial_place_max_iter=20,
        initial_place_min_diff_length=1500,
        initial_place_max_solver_iter=100,
        initial_place_max_fanout=200,
        initial_place_net_weight_scale=800.0,
        nesterov_place_max_iter=2000,
        bin_grid_cnt_x=0,
        bin_grid_cnt_y=0,
        target_overflow=0.1,
        target_density=1.0,
        init_density_penalty=0.00008,
        init_wire_length_coef=0.25,
        min_phi_coef=0.95,
        max_phi_coef=1.05,
        reference_hpwl=446000000.0,
        incremental_place_mode=0,
        verbose=0,
        max_back_track=10,
        min_preconditioner=1.0,
        initial_prev_coordi_update_coef=100.0,
        min_wire_length_force_bar=-300.0,
        poisson_transform=0,
    )
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(ctype))
    return a.ctypes.data_as(ctypes.POINTER(ctype))


@dataclass
class Netlist:
    """Structure-of-arrays netlist (int32 DBU)."""

    dbu_per_micron: int
    die: tuple
    site_w: int
    site_h: int
    row_box: np.ndarray            # (R, 4)
    inst_box: np.ndarray           # (N, 4)
    inst_fixed: np.ndarray         # (N,) uint8
    net_pin_start: np.ndarray      # (Nn+1,)
    pin_inst: np.ndarray           # (Np,)  -1 = BTerm
    pin_off: np.ndarray            # (Np, 2)
    pin_box: Optional[np.ndarray] = None   # (Np, 4) BTerm boxes
    inst_pin_start: Optional[np.ndarray] = None
    inst_pins: Optional[np.ndarray] = None
    pin_names: Optional[list] = None       # per pin: the port name of a BTerm, "" for an ITerm
    name: str = "netlist"
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        self.row_box = np.ascontiguousarray(self.row_box, dtype=np.int32).reshape(-1, 4)
        self.inst_box = np.ascontiguousarray(self.inst_box, dtype=np.int32).reshape(-1, 4)
        self.inst_fixed = np.ascontiguousarray(self.inst_fixed, dtype=np.uint8)
        self.net_pin_start = np.ascontiguousarray(self.net_pin_start, dtype=np.int32)
        self.pin_inst = np.ascontiguousarray(self.pin_inst, dtype=np.int32)
        self.pin_off = np.ascontiguousarray(self.pin_off, dtype=np.int32).reshape(-1, 2)
        if self.pin_box is not None:
            self.pin_box = np.ascontiguousarray(self.pin_box, dtype=np.int32).reshape(-1, 4)
        if self.inst_pin_start is not None:
            self.inst_pin_start = np.ascontiguousarray(self.inst_pin_start, dtype=np.int32)
            self.inst_pins = np.ascontiguousarray(self.inst_pins, dtype=np.int32)

    @property
    def num_insts(self) -> int:
        return int(self.inst_box.shape[0])

    @property
    def num_nets(self) -> int:
        return int(self.net_pin_start.shape[0] - 1)

    @property
    def num_pins(self) -> int:
        return int(self.pin_inst.shape[0])

    def to_c(self) -> RpNetlist:
        """Build the borrowed C view (arrays stay owned by this object)."""
        nl = RpNetlist()
        nl.dbu_per_micron = int(self.dbu_per_micron)
        nl.die_lx, nl.die_ly, nl.die_ux, nl.die_uy = (int(v) for v in self.die)
        nl.site_w = int(self.site_w)
        nl.site_h = int(self.site_h)
        nl.num_rows = int(self.row_box.shape[0])
        nl.row_box = _ptr(self.row_box, ctypes.c_int32)
        nl.num_insts = self.num_insts
        nl.inst_box = _ptr(self.inst_box, ctypes.c_int32)
        nl.inst_fixed = _ptr(self.inst_fixed, ctypes.c_uint8)
        nl.num_nets = self.num_nets
        nl.net_pin_start = _ptr(self.net_pin_start, ctypes.c_int32)
        nl.num_pins = self.num_pins
        nl.pin_inst = _ptr(self.pin_inst, ctypes.c_int32)
        nl.pin_off = _ptr(self.pin_off, ctypes.c_int32)
        nl.pin_box = _ptr(self.pin_box, ctypes.c_int32)
        nl.inst_pin_start = _ptr(self.inst_pin_start, ctypes.c_int32)
        nl.inst_pins = _ptr(self.inst_pins, ctypes.c_int32)
        return nl

    def save(self, path: str) -> None:
        arrs = dict(
            meta=np.array([self.dbu_per_micron, *self.die, self.site_w, self.site_h], dtype=np.int64),
            row_box=self.row_box,
            inst_box=self.inst_box,
            inst_fixed=self.inst_fixed,
            net_pin_start=self.net_pin_start,
            pin_inst=self.pin_inst,
            pin_off=self.pin_off,
        )
        if self.pin_box is not None:
            arrs["pin_box"] = self.pin_box
        if self.inst_pin_start is not None:
            arrs["inst_pin_start"] = self.inst_pin_start
            arrs["inst_pins"] = self.inst_pins
        if self.pin_names is not None:
            arrs["pin_names"] = np.array(self.pin_names, dtype=np.str_)
        np.savez_compressed(path, **arrs)

    @staticmethod
    def load(path: str, name: Optional[str] = None) -> "Netlist":
        z = np.load(path, allow_pickle=False)
        m = z["meta"]
        return Netlist(
            dbu_per_micron=int(m[0]),
            die=tuple(int(v) for v in m[1:5]),
            site_w=int(m[5]),
            site_h=int(m[6]),
            row_box=z["row_box"],
            inst_box=z["inst_box"],
            inst_fixed=z["inst_fixed"],
            net_pin_start=z["net_pin_start"],
            pin_inst=z["pin_inst"],
            pin_off=z["pin_off"],
            pin_box=z["pin_box"] if "pin_box" in z.files else None,
            inst_pin_start=z["inst_pin_start"] if "inst_pin_start" in z.files else None,
            inst_pins=z["inst_pins"] if "inst_pins" in z.files else None,
            pin_names=[str(v) for v in z["pin_names"]] if "pin_names" in z.files else None,
            name=name or path,
        )

    def write_bin(self, path: str) -> None:
        """Flat little-endian int32 file read by the C++ driver (replace_cli)."""
        npin = self.num_pins
        pin_box = self.pin_box if self.pin_box is not None else np.zeros((0, 4), np.int32)
        ips = self.inst_pin_start if self.inst_pin_start is not None else np.zeros(0, np.int32)
        ipn = self.inst_pins if self.inst_pins is not None else np.zeros(0, np.int32)
        # version 2: a block of NUL-terminated pin names (ports) follows
        names = b"".join((n or "").encode() + b"\0" for n in self.pin_names) if self.pin_names else b""
        head = np.array(
            [0x52504E4C, 2, self.dbu_per_micron, *self.die, self.site_w, self.site_h,
             self.row_box.shape[0], self.num_insts, self.num_nets, npin,
             pin_box.shape[0], ips.shape[0], ipn.shape[0], len(names)],
            dtype=np.int32,
        )
        with open(path, "wb") as f:
            for a in (head, self.row_box, self.inst_box, self.inst_fixed.astype(np.int32),
                      self.net_pin_start, self.pin_inst, self.pin_off, pin_box, ips, ipn):
                f.write(np.ascontiguousarray(a, dtype=np.int32).tobytes())
            f.write(names)


class RpServerOpts(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_int64),
        ("max_active", ctypes.c_int32),
        ("setup_threads", ctypes.c_int32),
        ("steps_per_poll", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("batched_ip", ctypes.c_int32),
    ]


class RpTileChoice(ctypes.Structure):
    """ctypes mirror of ``rp_tile_choice``."""

    _fields_ = [
        ("bucket", ctypes.c_int32),
        ("tuned", ctypes.c_int32),
        ("jobs_sampled", ctypes.c_int32),
        ("scatter_cells", ctypes.c_int32),
        ("wa_pins", ctypes.c_int32),
        ("scatter_us", ctypes.c_float * 3),
        ("wa_us", ctypes.c_float * 3),
    ]


class RpJobResult(ctypes.Structure):
    _fields_ = [
        ("stats", RpNesterovStats),
        ("code", ctypes.c_int32),
        ("done", ctypes.c_int32),
        ("arrival_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("start_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
    ]


class RpServerStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64),
        ("rebuilds", ctypes.c_int32),
        ("wall_s", ctypes.c_double),
    ]

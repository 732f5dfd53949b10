"""Shared helpers of the test suite (fixture loading, trace formatting)."""
import json
import os

import numpy as np

from replace_amd.netlist import Netlist, params_default

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def traces():
    with open(os.path.join(GOLDEN, "traces.json")) as f:
        return json.load(f)


def design(name):
    """(netlist, params) of a reference regression design, flags of its .tcl."""
    t = traces()[name]
    nl = Netlist.load(os.path.join(GOLDEN, name + ".npz"), name=name)
    return nl, params_default(**t["flags"])


def nesterov_lines(records, finished=None):
    """[NesterovSolve] lines as the reference prints them (src/nesterovPlace.cpp:421-423, 460)."""
    out = []
    for i, r in enumerate(records):
        if i == 0 or (i + 1) % 10 == 0:
            out.append(f"[NesterovSolve] Iter: {i + 1} overflow: {r[0]:g} HPWL: {r[1]}")
    if finished is not None:
        out.append(f"[NesterovSolve] Finished with Overflow: {finished:g}")
    return out


def golden_lines(name):
    t = traces()[name]
    out = [f"[NesterovSolve] Iter: {i} overflow: {o} HPWL: {h}" for i, o, h in t["nesterov"]]
    if t["finished"] is not None:
        out.append(f"[NesterovSolve] Finished with Overflow: {t['finished']}")
    return out


def rel_err(a, b):
    """max |a-b| / max |b| (normwise relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(b).max()
    return float(np.abs(a - b).max() / (d if d > 0 else 1.0))


def clamped_random_centres(gc, info, seed):
    """Random density centres inside the core, clamped so every density box
    stays in it (the state a Nesterov step leaves after its layout clamp)."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(info.core_lx, info.core_ux, info.num_gcells).astype(np.float32)
    cy = rng.uniform(info.core_ly, info.core_uy, info.num_gcells).astype(np.float32)
    cx = np.clip(cx, info.core_lx + gc["dhalf_x"], info.core_ux - gc["dhalf_x"]).astype(np.float32)
    cy = np.clip(cy, info.core_ly + gc["dhalf_y"], info.core_uy - gc["dhalf_y"]).astype(np.float32)
    return cx, cy


def bin_areas(info):
    """Bin areas of BinGrid::initBins (src/nesterovBase.cpp:674-711): regular
    bins of binSize, the last column / row truncated at the core edge
    (y-major, j * Nx + i)."""
    nx, ny = info.bin_cnt_x, info.bin_cnt_y
    lx = info.core_lx + np.arange(nx, dtype=np.int64) * info.bin_size_x
    ly = info.core_ly + np.arange(ny, dtype=np.int64) * info.bin_size_y
    w = np.minimum(lx + info.bin_size_x, info.core_ux) - lx
    h = np.minimum(ly + info.bin_size_y, info.core_uy) - ly
    return (h[:, None] * w[None, :]).reshape(-1)


def overflow_exact(bin_inst, nonplace, info):
    """Exact sum of the per-bin overflow terms max(0, float(inst) +
    float(nonPlace) - float(binArea * td)) (src/nesterovBase.cpp:812-816),
    each term in float as the reference computes it, truncated to int64."""
    at = (bin_areas(info).astype(np.float32) * np.float32(info.target_density)).astype(np.float32)
    t = (bin_inst.astype(np.float32) + nonplace.astype(np.float32)).astype(np.float32) - at
    t = np.maximum(np.float32(0), t.astype(np.float32))
    return int(np.sum(t.astype(np.float64)))


def check_eval(a, b, lam, info=None, nonplace=None):
    """One gradient evaluation of the HIP path (a) against the oracle (b):
    integer work and the float sums kept in the reference's order bit-exact,
    the Poisson solve and what follows it within 1e-4 relative, the gradient
    norms (FP64 here, sequential float in the oracle) within the float
    summation bound."""
    assert np.array_equal(a["bin_inst"], b["bin_inst"])
    assert np.array_equal(a["bin_filler"], b["bin_filler"])
    assert np.array_equal(a["density"], b["density"])
    # overflow area (src/nesterovBase.cpp:800-816): the reference adds each
    # bin's float term to an int64 through float (int64 += float), so its
    # running sum carries up to one rounding of the running value per bin:
    # within nb (u S + 1) of the exact sum, u = 2^-24.  The HIP path sums
    # the same float terms in FP64, so it is checked against their exact sum
    # (from its own bins, which equal the oracle's) and against the oracle's
    # value within that bound.
    nb = a["bin_inst"].size
    ov_b = int(b["overflow_area"][0])
    assert abs(int(a["overflow_area"][0]) - ov_b) <= max(64, nb * (2.0 ** -24 * abs(ov_b) + 1)), (
        a["overflow_area"], b["overflow_area"])
    if info is not None:
        exact = overflow_exact(a["bin_inst"], nonplace, info)
        assert abs(int(a["overflow_area"][0]) - exact) <= 1 + 1e-12 * exact, (a["overflow_area"], exact)
    assert np.array_equal(a["net_box"], b["net_box"])
    assert int(a["hpwl"][0]) == int(b["hpwl"][0])
    assert np.array_equal(a["net_sums"], b["net_sums"])
    assert np.array_equal(a["wl_grad"], b["wl_grad"])
    for k in ("phi", "ex", "ey", "dens_grad"):
        assert rel_err(a[k], b[k]) < 1e-4, (k, rel_err(a[k], b[k]))
    if lam > 0:
        assert rel_err(a["sum_grad"], b["sum_grad"]) < 1e-4
    else:
        assert np.array_equal(a["sum_grad"], b["sum_grad"])
    # sum |wl| and sum |dens| (src/nesterovPlace.cpp:225-235): the oracle sums
    # the n cells' nonnegative terms sequentially in float, as the reference
    # does, the HIP path in FP64, so the two differ by at most the sequential
    # summation bound (n - 1) u / (1 - (n - 1) u), u = 2^-24 (Higham, Accuracy
    # and Stability of Numerical Algorithms, 2nd ed., section 4.2), besides the
    # 1e-4 of the field for the density term
    n = a["wl_grad"].size // 2
    gam = (n - 1) * 2.0 ** -24 / (1 - (n - 1) * 2.0 ** -24)
    assert abs(a["sums"][0] - b["sums"][0]) <= max(1e-5, gam) * abs(b["sums"][0]), (a["sums"], b["sums"])
    assert abs(a["sums"][1] - b["sums"][1]) <= (1e-4 + gam) * abs(b["sums"][1]), (a["sums"], b["sums"])


def hpwl_at(rec, ovf):
    """HPWL of the first record at or below overflow ovf (else the last)."""
    for r in rec:
        if r[0] <= ovf:
            return r[1]
    return rec[-1][1]


def assert_tracks_trace(name, rec, iterations, final_overflow, ip):
    """A run's records against the reference's golden [NesterovSolve] lines:
    the first 50 iterations close (2e-3 overflow / 0.2 % HPWL; 5e-3 / 0.5 %
    for runs seeded by the parity-unpinned BiCGSTAB iterates), the finish
    within +-12 iterations, HPWL within 0.5 % at equal overflow."""
    t = traces()[name]
    for it, ovf, hpwl in t["nesterov"]:
        if it > 50 or it - 1 >= len(rec):
            break
        assert abs(rec[it - 1][0] - float(ovf)) < (5e-3 if ip else 2e-3)
        assert abs(rec[it - 1][1] - hpwl) <= (5e-3 if ip else 2e-3) * hpwl
    last_it, last_o, last_h = t["nesterov"][-1]
    assert abs(iterations - last_it) <= 12
    assert abs(hpwl_at(rec, float(last_o)) - last_h) <= 0.005 * last_h
    assert final_overflow <= 0.1


class DeviceArray:
    """A device buffer for the tests of entry points that take device
    pointers, allocated through the HIP runtime the library itself uses
    (libamdhip64.so.7, resolved to the copy already loaded with
    libreplace_amd.so), so no second runtime enters the process."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes
            from replace_amd.capi import lib
            lib()  # the library first: its runtime is the one reused below
            h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
            vp = ctypes.c_void_p
            h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
            h.hipFree.argtypes = [vp]
            h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host: np.ndarray):
        import ctypes
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        p = ctypes.c_void_p()
        assert self.hip().hipMalloc(ctypes.byref(p), max(1, self.nbytes)) == 0
        self.ptr = p.value
        assert self.hip().hipMemcpy(ctypes.c_void_p(self.ptr), self.host.ctypes.data_as(ctypes.c_void_p),
                                    self.nbytes, 1) == 0  # host to device

    def get(self) -> np.ndarray:
        import ctypes
        out = np.empty_like(self.host)
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                    self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            import ctypes
            self.hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

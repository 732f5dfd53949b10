"""Minimal LEF/DEF reader producing the placer's SoA ``Netlist``.

It reads what ``PlacerBase::init`` takes from OpenDB (reference
src/placerBase.cpp:605-758, 330-406) for the reference's regression designs:
LEF SITE and MACRO (SIZE, PIN shapes: RECT / POLYGON bboxes, pin USE), DEF
DIEAREA, ROW, COMPONENTS (+ PLACED / FIXED / COVER / UNPLACED), PINS
(+ LAYER box, placement, orientation) and NETS (signal nets; POWER, GROUND
and RESET nets are skipped as src/placerBase.cpp:686-693 does).

Conventions the reference inherits from OpenDB that this reader restates:
  * ITerm offset = centre of the bbox of all pin shapes minus the master
    centre, orientation ignored (src/placerBase.cpp:330-376).
  * BTerm position = centre of its pin box after the DEF orientation
    transform; an unplaced port has no box (src/placerBase.cpp:381-406).
  * Row bbox = origin + (numSites-1)*step + site size.
  * ``net_iterm_order`` selects whether a net's ITerms are listed in DEF
    order ("def") or reversed ("reverse"); ``inst_iterm_order`` likewise for
    an instance's ITerms against the LEF pin order.
"""
from __future__ import annotations

import re
from typing import Dict, List, Tuple

import numpy as np

from .netlist import INT_MAX, INT_MIN, Netlist

_TOKEN = re.compile(r'"[^"]*"|\S+')


def _tokens(text: str) -> List[str]:
    out = []
    for line in text.splitlines():
        h = line.find("#")
        if h >= 0:
            line = line[:h]
        out.extend(_TOKEN.findall(line))
    return out


class LefMacro:
    __slots__ = ("name", "w", "h", "pins", "pin_order", "pin_use")

    def __init__(self, name):
        self.name = name
        self.w = 0
        self.h = 0
        self.pins: Dict[str, Tuple[int, int, int, int]] = {}
        self.pin_order: List[str] = []
        self.pin_use: Dict[str, str] = {}


class Lef:
    def __init__(self):
        self.dbu = 2000
        self.sites: Dict[str, Tuple[int, int]] = {}
        self.macros: Dict[str, LefMacro] = {}


def _um(v: str, dbu: int) -> int:
    return int(round(float(v) * dbu))


def read_lef(path: str, dbu: int = 2000) -> Lef:
    lef = Lef()
    lef.dbu = dbu
    toks = _tokens(open(path).read())
    i, n = 0, len(toks)
    while i < n:
        t = toks[i]
        if t == "SITE" and i + 1 < n:
            name = toks[i + 1]
            i += 2
            w = h = 0
            while i < n and not (toks[i] == "END" and i + 1 < n and toks[i + 1] == name):
                if toks[i] == "SIZE":
                    w = _um(toks[i + 1], dbu)
                    h = _um(toks[i + 3], dbu)
                    i += 4
                    continue
                i += 1
            lef.sites[name] = (w, h)
            i += 2
        elif t == "MACRO" and i + 1 < n:
            m = LefMacro(toks[i + 1])
            i += 2
            while i < n and not (toks[i] == "END" and i + 1 < n and toks[i + 1] == m.name):
                if toks[i] == "SIZE":
                    m.w = _um(toks[i + 1], dbu)
                    m.h = _um(toks[i + 3], dbu)
                    i += 4
                elif toks[i] == "PIN":
                    pname = toks[i + 1]
                    i += 2
                    lx = ly = INT_MAX
                    ux = uy = INT_MIN
                    use = "SIGNAL"
                    while i < n and not (toks[i] == "END" and i + 1 < n and toks[i + 1] == pname):
                        tk = toks[i]
                        if tk == "USE":
                            use = toks[i + 1].upper()
                            i += 2
                            continue
                        if tk in ("RECT", "POLYGON"):
                            j = i + 1
                            if j < n and toks[j] == "MASK":
                                j += 2
                            coords = []
                            while j < n and toks[j] != ";":
                                coords.append(_um(toks[j], dbu))
                                j += 1
                            xs, ys = coords[0::2], coords[1::2]
                            lx, ly = min(lx, min(xs)), min(ly, min(ys))
                            ux, uy = max(ux, max(xs)), max(uy, max(ys))
                            i = j + 1
                            continue
                        i += 1
                    m.pins[pname] = (lx, ly, ux, uy)
                    m.pin_order.append(pname)
                    m.pin_use[pname] = use
                    i += 2
                elif toks[i] == "OBS":
                    while i < n and toks[i] != "END":
                        i += 1
                    i += 1
                else:
                    i += 1
            lef.macros[m.name] = m
            i += 2
        else:
            i += 1
    return lef


_ORIENT = {
    "N": lambda x, y: (x, y),
    "W": lambda x, y: (-y, x),
    "S": lambda x, y: (-x, -y),
    "E": lambda x, y: (y, -x),
    "FN": lambda x, y: (-x, y),
    "FW": lambda x, y: (-y, -x),
    "FS": lambda x, y: (x, -y),
    "FE": lambda x, y: (y, x),
}


def read_def(path: str, lef: Lef, net_iterm_order: str = "def",
             inst_iterm_order: str = "lef", name: str = None) -> Netlist:
    toks = _tokens(open(path).read())
    n = len(toks)
    dbu = lef.dbu
    die = (0, 0, 0, 0)
    rows = []
    comp_names: List[str] = []
    comp_master: List[str] = []
    comp_box = []
    comp_fixed = []
    bterm: Dict[str, Tuple[int, int, int, int]] = {}
    nets: List[Tuple[List[Tuple[str, str]], str]] = []
    i = 0
    while i < n:
        t = toks[i]
        if t == "UNITS" and toks[i + 1] == "DISTANCE":
            dbu = int(toks[i + 3])
            i += 5
        elif t == "DIEAREA":
            nums = []
            j = i + 1
            while toks[j] != ";":
                if toks[j] not in ("(", ")"):
                    nums.append(int(toks[j]))
                j += 1
            xs, ys = nums[0::2], nums[1::2]
            die = (min(xs), min(ys), max(xs), max(ys))
            i = j + 1
        elif t == "ROW":
            site = toks[i + 2]
            x, y = int(toks[i + 3]), int(toks[i + 4])
            orient = toks[i + 5]
            nx = ny = 1
            sx = sy = 0
            j = i + 6
            while toks[j] != ";":
                if toks[j] == "DO":
                    nx, ny = int(toks[j + 1]), int(toks[j + 3])
                    if toks[j + 4] == "STEP":
                        sx, sy = int(toks[j + 5]), int(toks[j + 6])
                    j += 7
                    continue
                j += 1
            sw, sh = lef.sites[site]
            if orient in ("W", "E", "FW", "FE"):
                sw, sh = sh, sw
            ux = x + (nx - 1) * sx + sw
            uy = y + (ny - 1) * sy + sh
            rows.append((x, y, ux, uy, sw, sh))
            i = j + 1
        elif t == "COMPONENTS":
            i += 3
            while toks[i] != "END":
                if toks[i] != "-":
                    i += 1
                    continue
                cname, master = toks[i + 1], toks[i + 2]
                j = i + 3
                status = "UNPLACED"
                x = y = 0
                orient = "N"
                while toks[j] != ";":
                    if toks[j] == "+" and toks[j + 1] in ("PLACED", "FIXED", "COVER"):
                        status = toks[j + 1]
                        x, y = int(toks[j + 3]), int(toks[j + 4])
                        orient = toks[j + 6]
                        j += 7
                        continue
                    if toks[j] == "+" and toks[j + 1] == "UNPLACED":
                        status = "UNPLACED"
                    j += 1
                m = lef.macros[master]
                w, h = m.w, m.h
                if orient in ("W", "E", "FW", "FE"):
                    w, h = h, w
                comp_names.append(cname)
                comp_master.append(master)
                comp_box.append((x, y, x + w, y + h))
                comp_fixed.append(1 if status in ("FIXED", "COVER") else 0)
                i = j + 1
            i += 2
        elif t == "PINS":
            i += 3
            while toks[i] != "END":
                if toks[i] != "-":
                    i += 1
                    continue
                pname = toks[i + 1]
                j = i + 2
                rect = None
                place = None
                orient = "N"
                while toks[j] != ";":
                    if toks[j] == "LAYER":
                        k = j + 2
                        nums = []
                        while len(nums) < 4:
                            if toks[k] not in ("(", ")"):
                                nums.append(int(toks[k]))
                            k += 1
                        rect = nums
                        j = k
                        continue
                    if toks[j] in ("PLACED", "FIXED", "COVER") and toks[j + 1] == "(":
                        place = (int(toks[j + 2]), int(toks[j + 3]))
                        orient = toks[j + 5]
                        j += 6
                        continue
                    j += 1
                if rect is not None and place is not None:
                    f = _ORIENT.get(orient, _ORIENT["N"])
                    a = f(rect[0], rect[1])
                    b = f(rect[2], rect[3])
                    bterm[pname] = (min(a[0], b[0]) + place[0], min(a[1], b[1]) + place[1],
                                    max(a[0], b[0]) + place[0], max(a[1], b[1]) + place[1])
                else:
                    bterm[pname] = None
                i = j + 1
            i += 2
        elif t == "NETS":
            i += 3
            while toks[i] != "END":
                if toks[i] != "-":
                    i += 1
                    continue
                j = i + 2
                conns = []
                use = "SIGNAL"
                in_conn = True
                while toks[j] != ";":
                    if toks[j] == "+":
                        in_conn = False
                        if toks[j + 1] == "USE":
                            use = toks[j + 2].upper()
                        j += 1
                        continue
                    if in_conn and toks[j] == "(":
                        conns.append((toks[j + 1], toks[j + 2]))
                        j += 4
                        continue
                    j += 1
                nets.append((conns, use))
                i = j + 1
            i += 2
        elif t == "SPECIALNETS":
            while not (toks[i] == "END" and toks[i + 1] == "SPECIALNETS"):
                i += 1
            i += 2
        else:
            i += 1

    comp_index = {c: k for k, c in enumerate(comp_names)}
    pin_inst, pin_off, pin_box, pin_name = [], [], [], []
    net_start = [0]
    inst_term: Dict[Tuple[int, str], int] = {}
    for conns, use in nets:
        if use in ("POWER", "GROUND", "RESET"):
            continue
        iterms = [(a, b) for a, b in conns if a != "PIN"]
        bterms = [b for a, b in conns if a == "PIN"]
        if net_iterm_order == "reverse":
            iterms = iterms[::-1]
        for comp, term in iterms:
            k = comp_index[comp]
            m = lef.macros[comp_master[k]]
            lx, ly, ux, uy = m.pins.get(term, (INT_MAX, INT_MAX, INT_MIN, INT_MIN))
            if lx == INT_MAX:
                ox = oy = 0
            else:
                ox = (lx + ux) // 2 - m.w // 2 if (lx + ux) >= 0 else -((-(lx + ux)) // 2) - m.w // 2
                oy = (ly + uy) // 2 - m.h // 2 if (ly + uy) >= 0 else -((-(ly + uy)) // 2) - m.h // 2
            inst_term[(k, term)] = len(pin_inst)
            pin_inst.append(k)
            pin_off.append((ox, oy))
            pin_box.append((0, 0, 0, 0))
            pin_name.append("")
        for b in bterms:
            bx = bterm.get(b)
            pin_inst.append(-1)
            pin_off.append((0, 0))
            pin_box.append(bx if bx is not None else (INT_MAX, INT_MAX, INT_MIN, INT_MIN))
            pin_name.append(b)
        net_start.append(len(pin_inst))
    # per-instance ITerm order: master pin order restricted to connected pins
    ips = [0]
    ipl: List[int] = []
    for k in range(len(comp_names)):
        m = lef.macros[comp_master[k]]
        order = m.pin_order if inst_iterm_order == "lef" else m.pin_order[::-1]
        for term in order:
            q = inst_term.get((k, term))
            if q is not None:
                ipl.append(q)
        ips.append(len(ipl))
    site_w, site_h = rows[0][4], rows[0][5]
    # the die PlacerBase reads is OpenDB's block bounding box
    # (dbBlock::getBBox, src/placerBase.cpp:616-617): the DIEAREA together
    # with every instance and every placed port shape
    bx = [die] + [tuple(b) for b in comp_box] + [b for b in pin_box if b[0] != INT_MAX and b[0] <= b[2]]
    bx = [b for b in bx if b[0] < b[2] or b[1] < b[3]]
    die = (min(b[0] for b in bx), min(b[1] for b in bx), max(b[2] for b in bx), max(b[3] for b in bx))
    return Netlist(
        dbu_per_micron=dbu,
        die=die,
        site_w=site_w,
        site_h=site_h,
        row_box=np.array([r[:4] for r in rows], dtype=np.int32),
        inst_box=np.array(comp_box, dtype=np.int32).reshape(-1, 4),
        inst_fixed=np.array(comp_fixed, dtype=np.uint8),
        net_pin_start=np.array(net_start, dtype=np.int32),
        pin_inst=np.array(pin_inst, dtype=np.int32),
        pin_off=np.array(pin_off, dtype=np.int32).reshape(-1, 2),
        pin_box=np.array(pin_box, dtype=np.int64).clip(INT_MIN, INT_MAX).astype(np.int32).reshape(-1, 4),
        inst_pin_start=np.array(ips, dtype=np.int32),
        inst_pins=np.array(ipl, dtype=np.int32),
        pin_names=pin_name,
        name=name or path,
    )


def write_def(src_def: str, out_path: str, lx, ly) -> None:
    """Writes ``src_def`` with its COMPONENTS section replaced by the placed
    lower-left corners ``lx``, ``ly`` (one per component, in the component
    order ``read_def`` returns), in the form OpenDB's write_def gives it in
    the reference's golden result files (test/*.defok):

        COMPONENTS n ;
            - name master + PLACED ( x y ) ORIENT ;
        END COMPONENTS

    Fixed (FIXED / COVER) components keep their status and position; every
    other component is written PLACED at its new corner with its orientation
    (N when the source leaves it unplaced).  The rest of the file is copied.
    """
    text = open(src_def).read()
    a = re.search(r"^COMPONENTS\s+\d+\s*;", text, re.M)
    b = re.search(r"^END COMPONENTS", text, re.M)
    if not a or not b:
        raise ValueError(f"{src_def}: no COMPONENTS section")
    body = text[a.end():b.start()]
    entries = [e for e in (s.strip() for s in body.split(";")) if e.startswith("-")]
    if len(entries) != len(lx) or len(entries) != len(ly):
        raise ValueError(f"{len(entries)} components, {len(lx)} positions")
    lines = [f"COMPONENTS {len(entries)} ;"]
    for k, e in enumerate(entries):
        toks = e.split()
        name, master = toks[1], toks[2]
        status, orient, pos = None, "N", None
        for j, t in enumerate(toks):
            if t == "+" and j + 1 < len(toks) and toks[j + 1] in ("PLACED", "FIXED", "COVER"):
                status = toks[j + 1]
                pos = (int(toks[j + 3]), int(toks[j + 4]))
                orient = toks[j + 6]
        if status in ("FIXED", "COVER"):
            lines.append(f"    - {name} {master} + {status} ( {pos[0]} {pos[1]} ) {orient} ;")
        else:
            lines.append(f"    - {name} {master} + PLACED ( {int(lx[k])} {int(ly[k])} ) {orient} ;")
    out = text[:a.start()] + "\n".join(lines) + "\n" + text[b.start():]
    with open(out_path, "w") as f:
        f.write(out)


def components_section(path: str) -> List[str]:
    """The COMPONENTS entries of a DEF, one normalised line each."""
    text = open(path).read()
    a = re.search(r"^COMPONENTS\s+\d+\s*;", text, re.M)
    b = re.search(r"^END COMPONENTS", text, re.M)
    body = text[a.end():b.start()]
    return [" ".join(e.split()) for e in body.split(";") if e.strip().startswith("-")]

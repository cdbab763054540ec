"""FlurryPP input files: the same keys, defaults and derived quantities as
input::readInputFile / input::nonDimensionalize (src/input.cpp:257-596).

File format (fileReader, src/input.cpp:76-250): one ``key value`` pair per
line; the first whitespace-separated token is the key, the next token is the
value, anything after it is ignored; the first matching line wins;
``mesh_bound`` may repeat (``mesh_bound <gmsh name> <bc>``).  A required key
that is missing raises ``InputError`` (the reference exits).
"""
from __future__ import annotations

import math
from pathlib import Path


class InputError(ValueError):
    pass


def _parse_lines(text: str) -> list[list[str]]:
    rows = []
    for line in text.splitlines():
        tok = line.split()
        if tok:
            rows.append(tok)
    return rows


def _conv(tok: str, typ):
    """Mimic `stringstream >> value` for int, float and str."""
    if typ is str:
        return tok
    if typ is int:
        s = tok.strip()
        i = 0
        if i < len(s) and s[i] in "+-":
            i += 1
        j = i
        while j < len(s) and s[j].isdigit():
            j += 1
        if j == i:
            raise ValueError(tok)
        return int(s[:j])
    if typ is float:
        # longest prefix that parses as a float
        for end in range(len(tok), 0, -1):
            try:
                return float(tok[:end])
            except ValueError:
                continue
        raise ValueError(tok)
    if typ is bool:
        return bool(_conv(tok, int))
    raise TypeError(typ)


class Options:
    def __init__(self, rows: list[list[str]]):
        self.rows = rows

    def get(self, key: str, typ, default=None, required=False):
        for r in self.rows:
            if r[0] == key:
                if len(r) < 2:
                    if required:
                        raise InputError(f"Required option not set: {key}")
                    return default
                try:
                    return _conv(r[1], typ)
                except ValueError:
                    if required:
                        raise InputError(f"Required option not set: {key}")
                    return default
        if required:
            raise InputError(f"Required option not found: {key}")
        return default

    def get_map(self, key: str) -> dict[str, str]:
        out = {}
        for r in self.rows:
            if r[0] == key:
                if len(r) < 3:
                    raise InputError(f"Required option not set: {key}")
                out[r[1]] = r[2]
        if not out:
            raise InputError(f"Required option not found: {key}")
        return out


class InputParams:
    """Run-time options of one simulation (the reference's `input` class)."""

    def __init__(self):
        self.time = 0.0
        self.rkTime = 0.0
        self.iter = 0
        self.meshBounds: dict[str, str] = {}

    # -- reading -----------------------------------------------------------
    @classmethod
    def from_file(cls, path: str | Path) -> "InputParams":
        return cls.from_text(Path(path).read_text())

    @classmethod
    def from_dict(cls, d: dict) -> "InputParams":
        lines = []
        for k, v in d.items():
            if k == "mesh_bound":
                for a, b in dict(v).items():
                    lines.append(f"mesh_bound {a} {b}")
            else:
                lines.append(f"{k} {v}")
        return cls.from_text("\n".join(lines))

    @classmethod
    def from_text(cls, text: str) -> "InputParams":
        p = cls()
        o = Options(_parse_lines(text))
        g = o.get
        p.equation = g("equation", int, 1)
        p.icType = g("icType", int, 0)
        p.nDims = g("nDims", int, required=True)
        if p.equation == 0:
            p.advectVx = g("advectVx", float, 1.0)
            p.advectVy = g("advectVy", float, 1.0)
            p.advectVz = g("advectVz", float, 0.0)
            p.diffD = g("diffD", float, 1.0)
            p.lambda_ = g("lambda", float, 1.0)
            p.nFields = 1
        elif p.equation == 1:
            p.gamma = g("gamma", float, 1.4)
            p.rhoBound = g("rhoBound", float, 1.0)
            p.uBound = g("uBound", float, 0.2)
            p.vBound = g("vBound", float, 0.0)
            p.wBound = g("wBound", float, 0.0)
            p.pBound = g("pBound", float, 0.7142857143)
            p.calcEntropySensor = g("entropySensor", int, 0)
            if p.icType == 0:
                p.rhoIC = g("rhoIC", float, p.rhoBound)
                p.vxIC = g("vxIC", float, p.uBound)
                p.vyIC = g("vyIC", float, p.vBound)
                p.vzIC = g("vzIC", float, p.wBound)
                p.pIC = g("pIC", float, p.pBound)
            elif p.icType == 2:
                p.vortexAngle = g("vortexAngle", float, math.atan(0.5))
            p.nFields = 4 if p.nDims == 2 else 5
        p.timeType = g("timeType", int, 4)
        p.dtType = g("dtType", int, 0)
        p.iterMax = g("iterMax", int, required=True)
        if p.dtType != 0:
            p.CFL = g("CFL", float, required=True)
            p.maxTime = g("maxTime", float, required=True)
            p.dt = 0.0
        else:
            p.dt = g("dt", float, required=True)
            p.maxTime = p.iterMax * p.dt
        p.viscous = g("viscous", int, 0)
        p.motion = g("motion", int, 0)
        p.order = g("order", int, 3)
        p.riemannType = g("riemannType", int, 0)
        p.testCase = g("testCase", int, 0)
        if p.viscous:
            p.penFact = g("LDG_penFact", float, 0.0)
            p.tau = g("LDG_tau", float, 1.0)
            if p.equation == 1:
                p.Re = g("Re", float, required=True)
                p.Lref = g("Lref", float, 1.0)
                p.muGas = g("muGas", float, 1.827e-5)
                p.prandtl = g("prandtl", float, 0.72)
                p.SGas = g("SGas", float, 120.0)
                p.TGas = g("TGas", float, 291.15)
                p.RGas = g("RGas", float, 286.9)
                p.fixVis = g("fixVis", int, 0)
                p.TBound = g("TBound", float, 300.0)
                p.TWall = g("TWall", float, 300.0)
                p.MachBound = g("MachBound", float, required=True)
                p.nxBound = g("nxBound", float, 1.0)
                p.nyBound = g("nyBound", float, 0.0)
                p.nzBound = g("nzBound", float, 0.0)
                p.MachWall = g("MachWall", float, required=True)
                p.nxWall = g("nxWall", float, 1.0)
                p.nyWall = g("nyWall", float, 0.0)
                p.nzWall = g("nzWall", float, 0.0)
        else:
            p.penFact, p.tau = 0.0, 1.0
        p.restart = g("restart", int, 0)
        if p.restart:
            p.restartIter = g("restartIter", int, required=True)
        p.meshType = g("meshType", int, required=True)
        if p.meshType == 1:
            p.nx = g("nx", int, 10)
            p.ny = g("ny", int, 10)
            p.nz = g("nz", int, 10)
            p.xmin = g("xmin", float, -10.0)
            p.xmax = g("xmax", float, 10.0)
            p.ymin = g("ymin", float, -10.0)
            p.ymax = g("ymax", float, 10.0)
            p.zmin = g("zmin", float, -10.0)
            p.zmax = g("zmax", float, 10.0)
            for side in ("Top", "Bottom", "Left", "Right", "Front", "Back"):
                setattr(p, f"create_bc{side}", g(f"create_bc{side}", str, "periodic"))
        else:
            if p.meshType == 0:
                p.meshFileName = g("meshFileName", str, required=True)
            elif p.meshType == 2:
                raise InputError("overset meshes (meshType 2) are outside this library")
            p.meshBounds = {a.lower(): b.lower() for a, b in o.get_map("mesh_bound").items()}
        p.periodicDX = g("periodicDX", float, math.inf)
        p.periodicDY = g("periodicDY", float, math.inf)
        p.periodicDZ = g("periodicDZ", float, math.inf)
        p.periodicTol = g("periodicTol", float, 1e-6)
        p.monitorResFreq = g("monitorResFreq", int, 10)
        if p.monitorResFreq < 0:
            p.monitorResFreq = 2**31 - 1
        p.monitorErrFreq = g("monitorErrFreq", int, -1)
        p.errorNorm = g("errorNorm", int, 1)
        p.quadOrder = g("quadOrder", int, 10)
        p.resType = g("resType", int, 2)
        p.plotType = g("plotType", int, 1)
        p.plotFreq = g("plotFreq", int, 100)
        p.dataFileName = g("dataFileName", str, "simData")
        p.sptsTypeQuad = g("spts_type_quad", str, "Legendre")
        p.vcjhSchemeQuad = g("vcjhSchemeQuad", int, 0)
        p.scFlag = g("shockCapture", int, 0)
        p.squeeze = g("squeeze", int, 0)
        p.PMG = g("PMG", int, 0)
        if p.PMG:
            p.lowOrder = g("lowOrder", int, 0)
            p.smoothSteps = g("smoothSteps", int, 1)
        p.HMG = g("HMG", int, 0)
        p.initIter = p.restartIter if p.restart else 0
        p.iter = p.initIter

        rk = {0: ([0.0], [1.0]), 2: ([0.0, 1.0], [0.5, 0.5]),
              4: ([0.0, 0.5, 0.5, 1.0], [1 / 6, 1 / 3, 1 / 3, 1 / 6]),
              5: ([0.0, 1 / 4, 1 / 3, 1 / 2, 1.0], [0.0])}
        if p.timeType in rk:
            p.RKa, p.RKb = rk[p.timeType]
        elif p.timeType == 6:
            p.RKa = [0.0, 0.153, 0.442, 0.931, 1.0] if p.order <= 2 else [0.0, 0.158, 0.477, 1.0, 1.0]
            p.RKb = [0.0]
        else:
            raise InputError("Time-Stepping type not supported.")
        p.nRKSteps = {0: 1, 2: 2, 4: 4, 5: 4, 6: 4}[p.timeType]
        if p.timeType == 4:
            p.RKb = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]

        if p.equation == 1:
            if p.nDims == 2:
                p.wBound = 0.0
            p.Uinf = math.sqrt(p.uBound ** 2 + p.vBound ** 2 + p.wBound ** 2)
            if p.viscous:
                p.non_dimensionalize()
            else:
                p.mu_inf = p.rt_inf = p.c_sth = 0.0
                p.prandtl = getattr(p, "prandtl", 0.72)
                p.fixVis = 1
                p.TWall = getattr(p, "TWall", 300.0)
                p.RGas = getattr(p, "RGas", 286.9)
                p.uWall = p.vWall = p.wWall = 0.0
        return p

    # src/input.cpp:529-596
    def non_dimensionalize(self) -> None:
        p = self
        if p.nDims == 2:
            p.nzBound = 0.0
        nMag = math.sqrt(p.nxBound ** 2 + p.nyBound ** 2 + p.nzBound ** 2)
        p.nxBound /= nMag
        p.nyBound /= nMag
        p.nzBound /= nMag
        Tref = p.TBound
        UBound = p.MachBound * math.sqrt(p.gamma * p.RGas * p.TBound)
        p.uBound, p.vBound, p.wBound = UBound * p.nxBound, UBound * p.nyBound, UBound * p.nzBound
        p.muBound = p.muGas * (p.TBound / p.TGas) ** 1.5 * ((p.TGas + p.SGas) / (p.TBound + p.SGas))
        p.rhoBound = p.muBound * p.Re / (UBound * p.Lref)
        p.pBound = p.rhoBound * p.RGas * p.TBound
        UWall = p.MachWall * math.sqrt(p.gamma * p.RGas * p.TWall)
        p.uWall, p.vWall, p.wWall = UWall * p.nxWall, UWall * p.nyWall, UWall * p.nzWall
        rhoRef = p.rhoBound
        pRef = rhoRef * UBound * UBound
        muRef = rhoRef * UBound * p.Lref
        p.c_sth = p.SGas / p.TGas
        p.mu_inf = p.muGas / muRef
        p.rt_inf = p.TGas * p.RGas / (UBound * UBound)
        p.RGas = p.RGas * p.TBound / (UBound * UBound)
        p.rhoBound = 1.0
        p.uBound /= UBound
        p.vBound /= UBound
        p.wBound /= UBound
        p.pBound = p.pBound / pRef
        p.TWall = p.TWall / Tref
        p.uWall /= UBound
        p.vWall /= UBound
        p.wWall /= UBound
        p.Uinf = math.sqrt(p.uBound ** 2 + p.vBound ** 2 + p.wBound ** 2)
        p.rhoIC, p.vxIC, p.vyIC, p.vzIC, p.pIC = p.rhoBound, p.uBound, p.vBound, p.wBound, p.pBound

"""Algorithmic work of each libflurry_hip kernel (FP64 fused path), for roofline fractions.

Bytes count every array a kernel must read or write once at its storage
precision (FP64 state, metrics and fluxes; int32 face indices); flops count
the arithmetic of the operator contractions only (flux formulas are a few
hundred flops per point and never bound these kernels).  Counts are per
launch of the kernel on a mesh with E elements and nFF face flux points of
which nFFb lie on boundaries.
"""
from __future__ import annotations

from dataclasses import dataclass

B8, B4 = 8, 4  # bytes per double / int32


@dataclass
class Work:
    bytes: float
    flops: float


def fused_work(D: int, P: int, E: int, nFF: int, nFFb: int, viscous: bool, nRK: int, time_type: int) -> dict:
    """Per-launch algorithmic bytes/flops of the fused-path kernels."""
    N, P1 = D + 2, P + 1
    S = P1 ** D
    F = 2 * D * P1 ** (D - 1)
    nFFi = nFF - nFFb
    out = {}
    # face_inviscid: read UL, UR (interior), normal, 3 int32 per face flux
    # point; write Fn once per pair (at L), the wave speed (+ dUc L (+R) in
    # viscous runs)
    rd = nFF * (N + D + 3 * B4 / B8) + nFFi * N
    wr = nFF * (N + 1)
    if viscous:
        wr += nFF * N + nFFi * N
    out["fused_face_inviscid"] = Work((rd + wr) * B8, 0.0)
    # volume
    rd = E * (N * S + D * D * S)
    wr = E * (N * S + N * F)
    fl = E * N * (D * S * 2 * P1 + F * D * 2 * P1)   # divergence + normal-flux extrapolation
    if viscous:
        rd += E * (S + N * F)                        # detJ, dUc
        rd += E * F * (D + B4 / B8)                  # face normal, side/partner code
        wr += E * N * F + nFFb * D * N               # side of the LDG flux; boundary gradient
        fl += E * N * (D * S * 2 * P1 + D * S * 2 * 2 * D + (D + 1) * F * 2 * P1)
        out["fused_volume_viscous"] = Work((rd + wr) * B8, fl)
    else:
        # inviscid runs: the volume kernel also applies the RK update (below);
        # divF and the discontinuous normal flux stay on chip, the common flux
        # Fn is read instead (with dA and the side code)
        vol_rd = rd + E * N * F + E * F * (1 + B4 / B8)
        vol_wr = wr - E * (N * S + N * F)
    if viscous and nFFb:
        # boundary LDG flux: U, normal, gradient in; the point's flux part out
        rd = nFFb * (N + D + D * N + 3 * B4 / B8)
        wr = nFFb * N
        out["fused_face_boundary"] = Work((rd + wr) * B8, 0.0)
    # rk update, averaged over the stages of one step
    per_stage = []
    for st in range(nRK):
        rd = E * (N * S + 2 * N * F + S)             # divF, Fn, disFn, detJ
        if viscous:
            rd += E * F * (2 * N + 1 + B4 / B8)      # both sides' LDG parts, dA, partner code
        wr = E * (N * S + N * F)                     # U, U at the faces
        if time_type >= 5:
            rd += E * N * S                          # U0
        elif nRK == 1:
            rd += E * N * S                          # U
        elif st == 0:
            rd += E * N * S                          # U (the step's U0: stages before the last write a stage copy)
            wr += E * N * S                          # Uacc
        elif st == nRK - 1:
            rd += E * N * S                          # Uacc
        else:
            rd += 2 * E * N * S                      # U0, Uacc
            wr += E * N * S                          # Uacc
        # (divF_spts[0] is stored by the last step of an update call only, not
        # per step: not counted)
        per_stage.append((rd + wr) * B8)
    rk = Work(sum(per_stage) / len(per_stage), E * N * S * 2 * 2 * D)
    if viscous:
        out["fused_rk_update"] = rk
    else:
        # the RK part without its reads of divF, Fn and disFn
        out["fused_volume_rk"] = Work((vol_rd + vol_wr) * B8 + rk.bytes - E * (N * S + 2 * N * F) * B8, fl + rk.flops)
    return out


def mixed_work(D: int, P: int, E: int, nFF: int, nFFb: int, viscous: bool, nRK: int, time_type: int) -> dict:
    """Per-launch algorithmic bytes of the mixed-path kernels: FP32 face and
    volume intermediates, fluxes, U at the flux points, stage copy of U and
    stored stage residuals, FP32 copies of the geometry; FP64 state, stage-0
    copy, detJ and wave speeds.  Flops as for the FP64 kernels."""
    N, P1 = D + 2, P + 1
    S = P1 ** D
    F = 2 * D * P1 ** (D - 1)
    nFFi = nFF - nFFb
    f4, d8, i4 = 4, 8, 4
    fp = fused_work(D, P, E, nFF, nFFb, viscous, nRK, time_type)
    out = {}
    # face: 3 ints, UL, UR, normal in; Fn once per pair (at L), wave speed
    # (FP64) [+ dUc L (+R)] out
    rd = nFF * (3 * i4 + f4 * (N + D)) + nFFi * f4 * N
    wr = nFF * (f4 * N + d8)
    if viscous:
        wr += (nFF + nFFi) * f4 * N
    out["mixed_face_inviscid"] = Work(rd + wr, 0.0)
    # volume: U (stage copy), metric terms (inviscid runs: with dA and the
    # side codes for the common flux)
    rd = E * f4 * (N * S + D * D * S) + (0 if viscous else E * F * (f4 + i4))
    if viscous:
        # 1/detJ, dUc, face normal, dA, side code; divF, this side's flux jump
        # (with its LDG part), the LDG part, boundary gradient out
        rd += E * (f4 * (S + N * F + D * F + F) + i4 * F)
        wr = E * f4 * (N * S + 2 * N * F) + nFFb * f4 * D * N
        out["mixed_volume_viscous"] = Work(rd + wr, fp["fused_volume_viscous"].flops)
        if nFFb:
            # U, normal, gradient, 3 ints, dA in; the point's flux jump updated
            out["mixed_face_boundary"] = Work(nFFb * (f4 * (N + D + D * N + 1) + 3 * i4 + 2 * f4 * N), 0.0)
    # classical RK stages keep their residuals in FP32 (the form without a
    # multigrid source); the FP64 state, written at the last stage only, is
    # also the step's U0; Jameson-form stages read it and write the state last
    per_stage = []
    for st in range(nRK):
        rd_ = E * (f4 * N * S + d8 * S)                # divF, detJ
        if viscous:
            rd_ += E * (f4 * 3 * N * F + f4 * F + i4 * F)  # Fn, flux jump, partner's LDG part, dA, code
        else:
            rd_ += E * f4 * N * F                      # Fn (the volume kernel forms the jump on chip)
        wr_ = E * (f4 * N * F + f4 * N * S)            # U at the faces, stage copy of U
        if time_type >= 5:
            rd_ += E * d8 * N * S                      # U0
            if st == nRK - 1:
                wr_ += E * d8 * N * S                  # U
        elif nRK == 1:
            rd_ += E * d8 * N * S
            wr_ += E * d8 * N * S
        elif st == 0:
            rd_ += E * d8 * N * S                      # U (the step's U0 until the last stage writes it)
            wr_ += E * f4 * N * S                      # residual of stage 0
        elif st == nRK - 1:
            rd_ += E * (d8 * N * S + f4 * N * S * (nRK - 1))  # U0, stored residuals
            wr_ += E * d8 * N * S                      # U
        else:
            rd_ += E * d8 * N * S                      # U0
            wr_ += E * f4 * N * S                      # residual of this stage
        per_stage.append(rd_ + wr_)
    rk = sum(per_stage) / len(per_stage)
    if viscous:
        out["mixed_rk_update"] = Work(rk, fp["fused_rk_update"].flops)
    else:
        out["mixed_volume_rk"] = Work(rd + rk - E * f4 * N * S, fp["fused_volume_rk"].flops)
    return out


def hbm_bound(w: Work, peak_tf: float = 78.6, peak_bw: float = 8.0) -> bool:
    """True when the kernel's intensity is under the FP64 ridge point."""
    if w.bytes == 0:
        return False
    return (w.flops / w.bytes) < (peak_tf / peak_bw)

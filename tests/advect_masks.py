"""Where an advection gradient is not fixed by the mathematics alone, for the
parity tests of the MacCormack adjoint against the FP64 oracle
(oracle/advect.py, VelocityGrid.warp semantics, data_structures.py:1690-1737).

Two kinds of points are non-differentiable, and fp32 (the device, like the
reference's kernels) and fp64 (the oracle) may land on different sides:

* A sample position i -+ dt*v whose fp32 sum rounds onto an integer along an
  axis: the trilinear derivative of the reference's corner rule (floor and
  ceil coincide there, sampling.hpp:514-527) is 0 along that axis, while the
  oracle's fp64 position is a hair off the integer.  Only that component of
  the cell's velocity gradient is affected (the other derivatives and every
  value are continuous there).  At 128^3 about 400 of 6M components.
* A branch choice within rounding of its switch: MC_SMOOTH (the corrected
  value within a relative 1e-5 of the corners' min or max) and NEGATIVE (the
  value within rounding of 0).  There the oracle takes the device's choice,
  read off its outputs: its unclamped MacCormack value against the min / max
  of the corners (exact in either precision), or the sign of its unclamped
  value.  Either choice is a valid one-sided derivative, and the device's
  backward must follow its own forward.  Without the device's results (the
  full-size fixtures), the cells whose gradient depends on a tie's choice are
  left out instead.  Exact ties on plateaus (all values involved equal, as in
  clipped densities) are evaluated exactly in both precisions and are not
  ties of rounding.
"""
import numpy as np

from oracle import advect as A
from oracle import sampling as S

REL_TIE = 1e-5


def near_integer_components(vc, dt=1.0):
    """bool [D,H,W,3]: velocity-gradient components whose fp32 sample
    positions (forward i - dt*v and back-step i + dt*v) may round onto an
    integer along that axis."""
    i = A._idx(vc.shape[:3])
    out = np.zeros(vc.shape, dtype=bool)
    for p in (i - dt * vc, i + dt * vc):
        out |= np.abs(p - np.round(p)) <= 4 * 2.0 ** -24 * np.maximum(np.abs(p), 1.0)
    return out


def faces_of_cells(comp_mask):
    """Per MAC component k, the faces beside a cell whose component k is in
    ``comp_mask`` [D,H,W,3] -> ([D,H,W+1,1], [D,H+1,W,1], [D+1,H,W,1])."""
    D, H, W, _ = comp_mask.shape
    fx = np.zeros((D, H, W + 1, 1), bool)
    fy = np.zeros((D, H + 1, W, 1), bool)
    fz = np.zeros((D + 1, H, W, 1), bool)
    m = comp_mask
    fx[:, :, 1:, 0] |= m[..., 0]
    fx[:, :, :-1, 0] |= m[..., 0]
    fy[:, 1:, :, 0] |= m[..., 1]
    fy[:, :-1, :, 0] |= m[..., 1]
    fz[1:, :, :, 0] |= m[..., 2]
    fz[:-1, :, :, 0] |= m[..., 2]
    return fx, fy, fz


def branch_ties(rho, vc, parts, order, clamp, boundary, dt=1.0):
    """bool [D,H,W,C]: branch choices within rounding of their switch."""
    clamp = clamp.upper()
    if clamp == "MC_SMOOTH":
        corr, mn, mx, fwd, back = parts["corr"], parts["mn"], parts["mx"], parts["fwd"], parts["back"]
        scale = np.maximum.reduce([np.abs(fwd), np.abs(rho), np.abs(back), np.abs(mn), np.abs(mx)])
        # on a plateau (every value involved equal: clipped densities) fp32
        # evaluates the tie exactly as fp64 does, so it is no tie of rounding
        flat = (fwd == back) & (back == rho) & (mn == mx) & (mx == fwd)
        return (np.minimum(np.abs(corr - mn), np.abs(corr - mx)) <= REL_TIE * scale) & ~flat
    if clamp == "NEGATIVE":
        i = A._idx(rho.shape[:3])
        p = i - dt * vc
        corner = np.maximum(np.abs(S.sample_minmax(rho, p, boundary, False)),
                            np.abs(S.sample_minmax(rho, p, boundary, True)))
        if order == 2:
            scale = np.maximum.reduce([np.abs(parts["fwd"]), np.abs(rho), np.abs(parts["back"]), corner])
            value = parts["corr"]
        else:
            scale, value = corner, parts["fwd"]
        return np.abs(value) <= REL_TIE * scale
    return np.zeros(rho.shape, dtype=bool)


def oracle_adjoint(rho, vc, g, order, clamp, boundary, gpu=None, dt=1.0):
    """The oracle's (g_rho, g_v) with the device's branch choices at ties,
    and the cells to compare: (g_rho, g_v, rho_keep [D,H,W], vel_keep
    [D,H,W,3], info).  ``gpu``: the device's results for the same inputs,
    {"sl": order 1 without clamp, "corr": order 2 without clamp, "out": the
    tested warp} as fp64 arrays [D,H,W,C]; without them every tie counts as
    undecided and its influence is left out)."""
    clamp = clamp.upper()
    _, parts = A.warp_forward(rho, vc, dt, order, clamp, boundary, return_parts=True)
    ties = branch_ties(rho, vc, parts, order, clamp, boundary, dt)
    decisions = {}
    undecided = np.zeros(rho.shape, dtype=bool)
    if clamp == "MC_SMOOTH":
        if gpu is not None:
            # the device's choice: its unclamped value against the corners'
            # min / max (exact in either precision); where the clamped result
            # shows it (SL and corrected values differ), it must agree
            dev_oob = (gpu["corr"] < parts["mn"]) | (gpu["corr"] > parts["mx"])
            shown = gpu["sl"] != gpu["corr"]
            assert np.array_equal((gpu["out"] == gpu["sl"])[shown], dev_oob[shown]), \
                "the clamped result does not follow the device's own revert decision"
            decisions["oob"] = np.where(ties, dev_oob, parts["oob"])
        else:
            decisions["oob"] = parts["oob"]
            undecided = ties
    elif clamp == "NEGATIVE":
        if gpu is not None:
            value = gpu["corr"] if order == 2 else gpu["sl"]
            decisions["neg_pass"] = np.where(ties, value >= 0.0, parts["neg_pass"])
        else:
            decisions["neg_pass"] = parts["neg_pass"]
            undecided = ties
    g_rho, g_v = A.warp_backward(rho, vc, g, dt, order, clamp, boundary, decisions=decisions)
    rho_keep = np.ones(rho.shape[:3], dtype=bool)
    vel_keep = ~near_integer_components(vc, dt)
    if np.any(undecided):
        # the gradient of g at the undecided cells, under either choice
        gu = np.where(undecided, g, 0.0)
        key = "oob" if clamp == "MC_SMOOTH" else "neg_pass"
        flipped = dict(decisions, **{key: np.where(undecided, ~decisions[key], decisions[key])})
        for dec in (decisions, flipped):
            r, v = A.warp_backward(rho, vc, gu, dt, order, clamp, boundary, decisions=dec)
            rho_keep &= ~np.any(r != 0.0, axis=-1)
            vel_keep &= ~(np.any(v != 0.0, axis=-1, keepdims=True))
    info = dict(ties=int(np.sum(np.any(ties, -1))), undecided=int(np.sum(np.any(undecided, -1))),
                rho_left_out=int(np.sum(~rho_keep)), vel_left_out=int(np.sum(~vel_keep)))
    return g_rho, g_v, rho_keep, vel_keep, info


def rel_on(got, exp, keep):
    """max |got - exp| over the kept entries / max |exp| over all."""
    scale = max(float(np.max(np.abs(exp))), 1e-30)
    keep = np.broadcast_to(keep, got.shape)
    return float(np.max(np.abs(got - exp)[keep], initial=0.0)) / scale

"""Algorithmic cost model of the chaos-game step, for roofline reporting.

The iterator does no HBM-heavy or matrix work: its ceiling is VALU issue.
The unit is one VALU issue slot per lane: simple FP32/integer ops count 1,
quarter-rate ops (v_mul_lo_u32 and the transcendental class: sin, cos,
sqrt, rcp, exp, log, atan2 kernel) count 4, and a division counts as one
rcp plus a multiply. Uniform work hoisted to the host (cos/sin of the
rotation, cumulative weights, the linearised palette) costs nothing.

Peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz = 39.3e12 slots/s, which is
MI355X's published 157.3 TFLOP/s FP32 divided by the 4 FLOP a packed FMA
retires per slot.
"""

PEAK_SLOTS = 256 * 4 * 16 * 2.4e9  # 39.3e12
Q = 4  # quarter-rate op

RNG_DRAW = 2 * Q + 10  # two v_mul_lo_u32, shifts/xors/adds, cvt and scale
# variation bodies (src/variations/defs/basic.rs, advanced.rs), slots
BODY = {
    0: 0,                                # linear
    1: 2 * Q,                            # sinusoidal: 2 sin
    2: 3 + 1 + 2 * (Q + 1),              # spherical: r2, +eps, 2 divisions
    3: 3 + 2 * Q + 6,                    # swirl: r2, sin, cos, 4 mul 2 add
    4: 3 + Q + 1 + Q + 7,                # horseshoe: r, sqrt, rcp, products
    5: 3 + Q + Q + (Q + 1) + 1,          # polar: atan2, sqrt, /pi, r - 1
    8: 3 + Q + Q + (Q + 1) + 1 + 2 * Q + 2,   # disc
    10: 3 + 1 + (Q + 1),                 # hyperbolic
    13: 3 + Q + Q + Q + RNG_DRAW + 3 + 2 * Q + 2,  # julia
    15: 2 * (3 + (Q + 1) + Q + 2),       # waves
    24: 3 + 2 * Q + Q + RNG_DRAW + 3 + (Q + 2) + 2 * Q + 2,  # julian: pow = log+exp
    27: 3 + 2 + Q + 2,                   # bubble
    47: 14 + Q + 6,                      # curl
}
SELECT_PER_XFORM = 2  # compare + select
AFFINE = 8
VAR_SUM = 4           # w*vx + w*vy accumulated
STEP_FIXED = 4 + 2 + 2  # bad-value test, colour blend, fuse
PLOT = 22 + 6 + 4       # world_to_pixel, palette texel, splat address


def xform_slots(fd, xf):
    s = AFFINE + (AFFINE if xf.post_enabled else 0)
    for k in range(fd.n_vars):
        if xf.var_weight[k] != 0.0:
            s += VAR_SUM + BODY[fd.var_id[k]]
    return s


def slots_per_iteration(fd, plot_fraction=1.0):
    """Expected VALU slots of one iteration of flame `fd`: transforms taken
    with their selection probabilities; the final chain and the plot paid on
    `plot_fraction` of iterations (those past the fuse)."""
    n = fd.n_xforms
    total_w = sum(fd.xforms[i].weight for i in range(n)) or 1.0
    s = 2 * RNG_DRAW + 1 + SELECT_PER_XFORM * n + 1 + STEP_FIXED
    s += sum(fd.xforms[i].weight / total_w * xform_slots(fd, fd.xforms[i]) for i in range(n))
    plot = PLOT + sum(xform_slots(fd, fd.finals[i]) for i in range(fd.n_finals))
    return s + plot_fraction * plot


def iterate_roofline(fd, samples, seconds, burn_in=20, ipt=256):
    """(achieved slots/s, fraction of peak) for `samples` iterations in `seconds`."""
    frac_plot = max(ipt - burn_in, 0) / ipt
    slots = slots_per_iteration(fd, frac_plot)
    achieved = samples * slots / seconds
    return achieved, achieved / PEAK_SLOTS, slots

"""Precision study at a 10 km origin offset (street scene).

The generator's street is 200 m long, so a 5 km path is covered by five
street segments whose origins sit at x = 10.0, 11.25, 12.5, 13.75 and 15 km
(y = 2 km, z = 50 m).  At these magnitudes one fp32 ulp is 0.98 mm.

Three arms are compared scan by scan, every scan starting from the FP64
oracle's prior (teacher forced):
  fp64     the oracle (the reference's algorithm in FP64);
  local    the registration path: compensated fp32 per point in a frame
           anchored at the scan's pose voxel, fp64 map moments;
  world32  the same kernels with the observe pass in plain fp32 world
           coordinates (obs_arith = 1): fl(R) p + fl(t), floorf key,
           residual against the fp32 world plane mean.
Reported per arm: per-scan pose error, relative error of the residual sum over
slots valid in both, valid-flag flips, keys that differ from the fp64 key, and
the plane-normal angle error of the final map.  Insertion uses the reference's
fp32 transform in every arm (as the reference does), so the maps differ only
through the poses.  The local arm must stay within the registration
tolerances (1e-4 m, 1e-5 rad); the world32 figures are reported, not bounded.
"""
import sys

import numpy as np
import pytest

from mapcmp import state_err
from oracle import Oracle
from voxelmapplus_fastlio2_amd import Context, LIOConfig, Registrar, initial_covariance, state_vector
from voxelmapplus_fastlio2_amd.synth import QDIAG, SEED0, STREET, Sequence

pytestmark = pytest.mark.gpu

NPTS = 30000
NSCANS = 16
SEGMENTS = [(10000.0 + 1250.0 * i, 2000.0, 50.0) for i in range(5)]


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def residual_cmp(vo, ro, v, r):
    both = vo & v
    so = float(np.sum(ro[both].astype(np.float64)))
    sa = float(np.sum(r[both].astype(np.float64)))
    denom = float(np.sum(np.abs(ro[both].astype(np.float64))))
    return abs(sa - so) / max(denom, 1e-300), int((vo != v).sum()), float(np.abs(ro[both] - r[both]).max())


def normal_angles(ref, got):
    """Angle between plane normals of voxels that are planes in both maps."""
    ka = {tuple(k): i for i, k in enumerate(ref["keys"])}
    ang = []
    for j, k in enumerate(got["keys"]):
        i = ka.get(tuple(k))
        if i is None or not (ref["ints"][i, 1] & 2) or not (got["ints"][j, 1] & 2):
            continue
        na, nb = ref["dbl"][i, 3:6], got["dbl"][j, 3:6]
        ang.append(np.arctan2(np.linalg.norm(np.cross(na, nb)), abs(float(na @ nb))))
    return np.array(ang)


def run_segment(ctxs, off, seed):
    seq = Sequence(STREET, seed, NPTS, origin_offset=off)
    cfg = LIOConfig(max_scan_points=NPTS, map_capacity=2000000)
    R, t = seq.pose(0)
    s0, P0 = state_vector(R, t), initial_covariance()
    x0, y0, z0 = seq.scan(0)
    orc = Oracle(cfg)
    orc.set_state(s0, P0)
    orc.build(x0, y0, z0)
    arms = {}
    for name, ctx in ctxs.items():
        g = Registrar(ctx, cfg)
        g.set_state(s0, P0)
        g.set_points(x0, y0, z0)
        g.build()
        arms[name] = g
    out = {name: {"pos": [], "rot": [], "rsum": [], "flips": [], "rmax": [], "keydiff": [], "fallback": []}
           for name in arms}
    for k in range(1, NSCANS):
        orc.apply_motion(*seq.motion(k), QDIAG)
        sp, Pp = orc.get_state()
        x, y, z = seq.scan(k)
        orc.set_points(x, y, z)
        orc.observe(sp)
        vo, ro, _ = orc.export_slots(NPTS)
        orc.register(x, y, z)
        so, _ = orc.get_state()
        for name, g in arms.items():
            g.set_state(sp, Pp)
            g.set_points(x, y, z)
            _, _, ost = g.observe(sp)
            v, r, _ = g.export_slots(NPTS)
            rel, flips, rmax = residual_cmp(vo, ro, v, r)
            g.register()
            sg, _ = g.get_state()
            ep, er = state_err(so, sg)
            o = out[name]
            o["pos"].append(ep)
            o["rot"].append(er)
            o["rsum"].append(rel)
            o["flips"].append(flips)
            o["rmax"].append(rmax)
            o["keydiff"].append(ost.key_fallback if name == "world32" else 0)
            o["fallback"].append(ost.key_fallback if name == "local" else 0)
    ref = orc.export_map()
    for name, g in arms.items():
        out[name]["normal"] = normal_angles(ref, g.export_map())
        g.close()
    orc.close()
    seq.close()
    return out


def test_precision_study_10km():
    ctxs = {"local": Context(0), "world32": Context(0)}
    ctxs["world32"].set_param("obs_arith", 1)
    agg = {n: {} for n in ctxs}
    try:
        for si, off in enumerate(SEGMENTS):
            res = run_segment(ctxs, off, SEED0 + 500 + si)
            for name, o in res.items():
                a = agg[name]
                for key, vals in o.items():
                    a.setdefault(key, []).extend(list(vals))
                log(f"segment x0={off[0]:.0f} m {name:8s}: pos max {max(o['pos']):.2e} m, rot max "
                    f"{max(o['rot']):.2e} rad, residual-sum rel max {max(o['rsum']):.2e}, flips "
                    f"{sum(o['flips'])}, key diffs {sum(o['keydiff'])}, normal max "
                    f"{o['normal'].max() if len(o['normal']) else 0.0:.2e} rad")
    finally:
        for c in ctxs.values():
            c.close()
    log("| arm | pose err max (m / rad) | residual-sum rel err (max / median) | valid flips per scan "
        "| key diffs per scan | normal angle err (max / p99, rad) |")
    for name in ("local", "world32"):
        a = agg[name]
        nrm = np.array(a["normal"])
        scans = len(a["pos"])
        log(f"| {name} | {max(a['pos']):.2e} / {max(a['rot']):.2e} | {max(a['rsum']):.2e} / "
            f"{np.median(a['rsum']):.2e} | {sum(a['flips']) / scans:.1f} | {sum(a['keydiff']) / scans:.1f} | "
            f"{nrm.max():.2e} / {np.percentile(nrm, 99):.2e} |")
    loc = agg["local"]
    assert max(loc["pos"]) <= 1e-4 and max(loc["rot"]) <= 1e-5
    assert max(loc["rsum"]) < 1e-4
    assert sum(loc["keydiff"]) == 0

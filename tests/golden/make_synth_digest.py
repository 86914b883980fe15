#!/usr/bin/env python3
"""Writes synth_digest.json: SHA-256 digests of a few scans and poses of the
seeded room / street sequences, so that every machine that runs the tests is
shown to generate bit-identical inputs (the parity tests and the benchmark
depend on it).  Run from the repository root after building the generator:

    python tests/golden/make_synth_digest.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from voxelmapplus_fastlio2_amd.synth import ROOM, SEED0, STREET, ImuSequence, Sequence  # noqa: E402

CASES = [("room", ROOM, SEED0, 5000, (0, 1, 7, 150)), ("street", STREET, SEED0, 6000, (0, 1, 300))]
# raw-scan packages (points, per-point times, IMU samples, sweep bounds) of
# the IMU sequences, in order from the first package
IMU_CASES = [("room_imu", ROOM, SEED0, 5000, 4), ("street_imu", STREET, SEED0, 6000, 3)]


def digest(scene, seed, n, ks):
    seq = Sequence(scene, seed, n)
    out = {}
    for k in ks:
        h = hashlib.sha256()
        for a in seq.scan(k):
            h.update(a.tobytes())
        R, t = seq.pose(k)
        h.update(R.tobytes())
        h.update(t.tobytes())
        dR, dp = seq.motion(k) if k > 0 else (R * 0, t * 0)
        h.update(dR.tobytes())
        h.update(dp.tobytes())
        out[str(k)] = h.hexdigest()
    seq.close()
    return out


def digest_imu(scene, seed, n, npkg):
    seq = ImuSequence(scene, seed, n)
    out = {}
    for k in range(npkg):
        x, y, z, t, (it, acc, gyro), cs, ce = seq.package(k)
        h = hashlib.sha256()
        for a in (x, y, z, t, it, acc, gyro, np.array([cs, ce])):
            h.update(np.ascontiguousarray(a).tobytes())
        out[str(k)] = h.hexdigest()
    seq.close()
    return out


def main():
    res = {name: {"seed": seed, "n": n, "digests": digest(sc, seed, n, ks)} for name, sc, seed, n, ks in CASES}
    for name, sc, seed, n, npkg in IMU_CASES:
        res[name] = {"seed": seed, "n": n, "digests": digest_imu(sc, seed, n, npkg)}
    with open(os.path.join(HERE, "synth_digest.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()

"""FP32-vs-FP64 drift report of the AVBD step.

The CPU oracle runs the same scene in float and in double (colour order,
portable trig); the GPU step equals the float run bit for bit, so the table
is the GPU path's drift against an FP64 evaluation of the same algorithm.

    python tests/fp64_drift.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from oracle import pyoracle as po  # noqa: E402


def drift(scene, steps):
    """Largest and median position difference (m) and largest rotation
    difference (1 - |<q32, q64>|) over dynamic bodies after `steps`."""
    o32 = po.OracleEngine(scene, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE, precision=32)
    o64 = po.OracleEngine(scene, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE, precision=64)
    o32.step_inner(steps)
    o64.step_inner(steps)
    a = o32.state().astype(np.float64)
    b = o64.state_f64()
    dyn = scene["body_type"] == 1
    dp = np.linalg.norm(a[dyn, 0:3] - b[dyn, 0:3], axis=1)
    dq = np.abs(1.0 - np.abs(np.sum(a[dyn, 3:7] * b[dyn, 3:7], axis=1)))
    return {"pos_max": float(dp.max()), "pos_median": float(np.median(dp)), "rot_max": float(dq.max())}


def main():
    from ornis_amd import scenes
    from scenes_ref import snapshot_scene, stack_scene
    cases = [("reference snapshot scene (9 bodies)", snapshot_scene(), 120),
             ("4-box stack", stack_scene(4), 300),
             ("200-box random drop", scenes.box_drop_scene(200, np.random.Generator(np.random.PCG64(7))), 150),
             ("210-cube pyramid", scenes.box_pyramid(), 300)]
    print("| scene | steps | max |dx| (m) | median |dx| (m) | max rotation diff |")
    print("|---|---:|---:|---:|---:|")
    for name, sc, steps in cases:
        r = drift(sc, steps)
        print(f"| {name} | {steps} | {r['pos_max']:.3g} | {r['pos_median']:.3g} | {r['rot_max']:.3g} |")


if __name__ == "__main__":
    main()

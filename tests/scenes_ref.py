"""Test scenes taken from the reference's own tests."""
import numpy as np

from ornis_amd.scenes import SceneBuilder


def snapshot_scene():
    """crates/physics/tests/avbd_engine.rs:1558-1603: floor, 4-box stack,
    dropped sphere, fast box, ball-jointed pendulum (9 bodies)."""
    sb = SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (10.0, 1.0, 10.0), 0.0)
    for i in range(4):
        y = float(np.float32(0.5) + np.float32(i) * np.float32(1.02))
        sb.box((0.0, y, 0.0), (0.5, 0.5, 0.5), 1.0)
    sb.sphere((3.0, 6.0, 0.0), 0.5, 1.0)
    sb.box((-3.0, 8.0, 0.0), (0.4, 0.4, 0.4), 1.0, vel=(0.0, -30.0, 0.0))
    a = sb.box((6.0, 3.0, 0.0), (0.5, 0.5, 0.5), 0.0)
    b = sb.box((6.0, 1.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    sb.ball_joint(a, b, (0.0, -1.0, 0.0), (0.0, 1.0, 0.0))
    return sb.build()


def stack_scene(n=4):
    """crates/physics/tests/avbd_engine.rs:31-62: a 4-box stack on a floor."""
    sb = SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (5.0, 1.0, 5.0), 0.0)
    for i in range(n):
        sb.box((0.0, float(np.float32(0.5) + np.float32(i) * np.float32(1.02)), 0.0), (0.5, 0.5, 0.5), 1.0)
    return sb.build()


def read_golden(path):
    lines = open(path).read().splitlines()
    rows = [np.array([int(w, 16) for w in ln.split()], dtype=np.uint32) for ln in lines[1:]]
    return lines[0], np.stack(rows).view(np.float32)

"""The ridge centerline method (runCircleFittingAndRidgeTraversal,
tube-segmentation.cpp:1260-1350; runRidgeTraversal,
ridgeTraversalCenterlineExtraction.cpp:37-369) on the benchmark's C1 workload:
64^3, 12 tubes r ~ U(1, 4), 1.0 on 0.0 plus N(0, 0.1^2), through the C++
drop-in's run() on a .mhd file (mode white, blur 0.5, GVF mu 0.1 x 250 in the
reference's default 16-bit mode, radii 1..5 step 0.5).

The walk has no oracle restatement; the centerline is scored against the
tree's known axes as the reference's own validation does against its ground
truth (tests/tubeValidation.cpp:17-137): the mean distance from the
extracted centerline voxels to the nearest axis, and the fraction of axis
points (every 0.5 voxel along segments of radius >= 1.5, away from the
volume's border) with a centerline voxel within 2 voxels.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "test_dropin")


def write_mhd(path, vol):
    Z, Y, X = vol.shape
    raw = os.path.splitext(path)[0] + ".raw"
    np.ascontiguousarray(vol, np.float32).tofile(raw)
    with open(path, "w") as f:
        f.write(f"ObjectType = Image\nNDims = 3\nDimSize = {X} {Y} {Z}\nElementSpacing = 1 1 1\n"
                f"ElementType = MET_FLOAT\nElementDataFile = {os.path.basename(raw)}\n")


def segment_distance(P, A, B):
    """Distance of points P (n, 3; x, y, z) to segment AB."""
    AB = B - A
    t = np.clip(((P - A) @ AB) / max(float(AB @ AB), 1e-12), 0.0, 1.0)
    return np.linalg.norm(P - (A + t[:, None] * AB), axis=1)


def test_ridge_on_c1(tmp_path):
    from scipy.spatial import cKDTree
    from tsf import synth
    vol, segs = synth.config1()
    mhd = str(tmp_path / "c1.mhd")
    write_mhd(mhd, vol)
    out = str(tmp_path / "centerline.raw")
    r = subprocess.run([EXE, "ridge_volume", mhd, out], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASS" in r.stdout
    mask = np.fromfile(out, np.uint8).reshape(vol.shape)
    z, y, x = np.nonzero(mask)
    P = np.stack([x, y, z], 1).astype(np.float64)
    assert len(P) > 20
    d = np.min(np.stack([segment_distance(P, A, B) for A, B, _, _ in segs]), axis=0)
    # axis points of the tubes the method is set up for (radius >= 1.5),
    # 4 voxels or more from the border
    pts = []
    for A, B, rad, _ in segs:
        if rad < 1.5:
            continue
        n = max(2, int(np.linalg.norm(B - A) / 0.5))
        pts.append(A + np.linspace(0, 1, n)[:, None] * (B - A))
    Q = np.concatenate(pts)
    Q = Q[np.all((Q >= 4) & (Q <= np.array(vol.shape[::-1]) - 5), axis=1)]
    near, _ = cKDTree(P).query(Q, distance_upper_bound=2.0)
    covered = float(np.mean(np.isfinite(near)))
    print(f"C1 ridge: {len(P)} centerline voxels, mean distance to the axes {d.mean():.3f}, "
          f"within 1.5 voxels {np.mean(d <= 1.5):.3f}; axis coverage {covered:.3f} of {len(Q)} points")
    assert d.mean() < 1.5 and covered > 0.5

"""Segmentation on the device against the oracle, voxel for voxel: the
inverse-gradient region growing (with its closing) and the sphere
segmentation, on a synthetic tube field, a GVF field of rendered tubes, and
the 16-bit storage of that field."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tube_field(shape, peak):
    z, y, x = np.mgrid[:shape[0], :shape[1], :shape[2]].astype(np.float32)
    dy, dz = y - (shape[1] - 1) / 2, z - (shape[0] - 1) / 2
    rho = np.sqrt(dy * dy + dz * dz)
    m = np.where(rho <= peak, rho, np.maximum(0.0, 2 * peak - rho))
    F = np.zeros((4,) + shape, np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        F[1] = np.where(rho > 0, -m * dy / rho, 0.0)
        F[2] = np.where(rho > 0, -m * dz / rho, 0.0)
    w = np.sqrt(F[1] ** 2 + F[2] ** 2)
    F[3] = np.where(w > 0, w, 1.0)
    return F


def run_device(tsf, cl, radius, F, storage=None):
    fld = tsf.Field.from_host(F) if storage is None else tsf.Field.from_host(F, storage[0], storage[1])
    out = tsf.DeviceArray(cl.shape, np.uint8)
    passes = tsf.segment_inverse_gradient(tsf.DeviceArray.from_host(cl), tsf.DeviceArray.from_host(radius), fld, out)
    return out.download(), passes


@pytest.mark.parametrize("nx", [32, 31], ids=["words", "bytes"])
def test_segment_tube_vs_oracle(tsf, orc, nx):
    """X = 32 takes the closing on 4-voxel words, X = 31 the byte passes."""
    shape = (28, 28, nx)
    F = tube_field(shape, 6.0)
    cl = np.zeros(shape, np.uint8)
    cl[13:15, 13:15, 2:30] = 1
    r = np.full(shape, 2.0, np.float32)
    got, passes = run_device(tsf, cl, r, F)
    exp, opasses = orc.segment_inverse_gradient(cl, r, F)
    print(f"tube: {int(got.sum())} voxels, {passes} passes (oracle {opasses})")
    assert passes == opasses and np.array_equal(got, exp)


@pytest.mark.parametrize("seed", [31, 32])
def test_segment_gvf_of_tubes_vs_oracle(tsf, orc, seed):
    # the GVF field of rendered tubes and a centerline from their skeleton
    from tsf import synth
    shape = (40, 44, 48)
    vol, _ = synth.tube_volume(shape, 5, 1.5, 3.5, seed, noise=0.02)
    F = orc.gvf_euler(orc.vector_field(orc.blur(vol, 0.5, "f32"), 0.2, 1, "f32"), 0.05, 40, "f32")
    t, rad = orc.tdf_circle(F, 1.0, 4.0, 0.5, "f32")
    cl = (t > np.quantile(t, 0.985)).astype(np.uint8)
    got, passes = run_device(tsf, cl, rad, F)
    exp, opasses = orc.segment_inverse_gradient(cl, rad, F)
    iou = (got & (vol > 0.5)).sum() / max((got | (vol > 0.5)).sum(), 1)
    print(f"tubes {seed}: {int(cl.sum())} centerline voxels -> {int(got.sum())} segmented, {passes} passes, "
          f"IoU with the tubes {iou:.3f}")
    assert passes == opasses and np.array_equal(got, exp)
    # the sphere segmentation of the same centerline
    s = tsf.DeviceArray(shape, np.uint8)
    tsf.segment_spheres(tsf.DeviceArray.from_host(cl), tsf.DeviceArray.from_host(rad), s)
    assert np.array_equal(s.download(), orc.segment_spheres(cl, rad))


def test_segment_reads_16bit_fields(tsf, orc):
    # an SN16 field: the device decodes the codes; the oracle gets the decoded values
    shape = (28, 28, 32)
    F = tube_field(shape, 6.0)
    F[:3] /= 8.0
    F[3] = np.where(np.sqrt((F[:3] ** 2).sum(0)) > 0, np.sqrt((F[:3] ** 2).sum(0)), 1.0)
    Fq = orc.sn16_round(F)
    cl = np.zeros(shape, np.uint8)
    cl[13:15, 13:15, 2:30] = 1
    r = np.full(shape, 2.0, np.float32)
    got, passes = run_device(tsf, cl, r, Fq, storage=(tsf.STORAGE_SN16, 1.0))
    exp, opasses = orc.segment_inverse_gradient(cl, r, Fq)
    assert passes == opasses and np.array_equal(got, exp)

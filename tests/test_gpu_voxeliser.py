"""Device sphere voxeliser (mplbm_place_spheres) against the host
place_sphere restatement (src/multiphase_lbm.cu:893-927), which
tests/test_harness.py pins by brute-force recounts and the reference's
767-sphere positions file.  Masks and solid numbers must be identical:
the float-stepped sample points, the dist2 < R^2 test and nearest() are the
same IEEE operations on both sides, and the last sphere of the list that
covers a cell names it."""
import os

import numpy as np
import pytest

from multiphase_lbm_amd import harness as H

pytestmark = pytest.mark.gpu

import multiphase_lbm_amd as mp  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _host(shape, spheres, first_id, mask=None, sn=None):
    m = np.ones(shape, np.uint8) if mask is None else mask.copy()
    s = np.zeros(shape, np.int64) if sn is None else sn.copy()
    for k, (x, y, z, r) in enumerate(np.asarray(spheres, np.float32)):
        H.place_sphere(m, float(x), float(y), float(z), float(r), s, first_id + k)
    return m, s


def _device(shape, spheres, first_id, mask=None, sn=None):
    m = np.ones(shape, np.uint8) if mask is None else mask.copy()
    s = np.zeros(shape, np.int64) if sn is None else sn.copy()
    mp.place_spheres(m, spheres, first_id, s)
    return m, s


@pytest.mark.parametrize("shape,nsph,rmin,rmax,seed", [((40, 36, 28), 60, 1.0, 7.5, 1), ((17, 23, 64), 25, 0.3, 12.0, 2),
                                                       ((96, 64, 80), 400, 2.0, 9.0, 3)])
def test_random_spheres_match_host(shape, nsph, rmin, rmax, seed):
    # non-integer centres anywhere in the box (spheres cross the periodic
    # faces), radii from below one cell, heavy overlap
    rng = np.random.default_rng(seed)
    sph = np.column_stack([rng.uniform(0, 1, (nsph, 3)) * np.array(shape), rng.uniform(rmin, rmax, nsph)])
    sph = sph.astype(np.float32)
    print(f"case voxeliser random {shape} n={nsph}", flush=True)
    hm, hs = _host(shape, sph, 5)
    dm, ds = _device(shape, sph, 5)
    assert np.array_equal(hm, dm)
    assert np.array_equal(hs, ds)
    assert (dm == 0).any() and ds.max() > 5


def test_existing_geometry_and_numbers_are_kept_or_overwritten():
    # a second list over an earlier one: covered cells take the new numbers,
    # the rest keep theirs
    shape = (30, 30, 30)
    rng = np.random.default_rng(8)
    a = np.column_stack([rng.uniform(0, 30, (20, 3)), rng.uniform(2, 5, 20)]).astype(np.float32)
    b = np.column_stack([rng.uniform(0, 30, (15, 3)), rng.uniform(2, 5, 15)]).astype(np.float32)
    hm, hs = _host(shape, a, 1)
    hm2, hs2 = _host(shape, b, 21, hm, hs)
    dm2, ds2 = _device(shape, b, 21, hm, hs)
    assert np.array_equal(hm2, dm2) and np.array_equal(hs2, ds2)
    assert ((hs2 > 0) & (hs2 <= 20)).any() and (hs2 > 20).any()


def test_edge_spheres():
    # empty list, a zero radius, a sphere bigger than the box (cells beyond
    # one periodic wrap are skipped), integer centres (sample points on the
    # rounding boundary of nearest()), negative centres
    shape = (12, 10, 14)
    print("case voxeliser edges", flush=True)
    m, s = _device(shape, np.zeros((0, 4), np.float32), 1)
    assert (m == 1).all() and (s == 0).all()
    sph = np.array([[5.0, 5.0, 5.0, 0.0], [6.0, 5.0, 7.0, 9.5], [0.5, 0.5, 0.5, 2.0], [-1.25, 3.5, 13.75, 3.0],
                    [11.5, 9.5, 13.5, 1.0]], np.float32)
    for k in range(len(sph)):
        hm, hs = _host(shape, sph[k:k + 1], 3)
        dm, ds = _device(shape, sph[k:k + 1], 3)
        assert np.array_equal(hm, dm) and np.array_equal(hs, ds), k
    hm, hs = _host(shape, sph, 1)
    dm, ds = _device(shape, sph, 1)
    assert np.array_equal(hm, dm) and np.array_equal(hs, ds)


def test_positions_file_on_device():
    # the reference's 767-sphere positions file (res 1 and 2)
    for res, n in ((1.0, 200), (2.0, 400)):
        shape = (n, n, n)
        print(f"case voxeliser positions res={res}", flush=True)
        hm, hs, k1 = H.read_positions(os.path.join(GOLDEN, "example_positions_file"), shape, res=res)
        dm, ds, k2 = H.read_positions(os.path.join(GOLDEN, "example_positions_file"), shape, res=res, device=True)
        assert k1 == k2 == 767
        assert np.array_equal(hm, dm) and np.array_equal(hs, ds)


INPUT = """device_voxeliser {dev}
result_folder {folder}
statPeriod 1
lx 40
ly 36
lz 30
rho0 0.7
IniPerturbRate 0.0285714
G -5.0
resize
generate_solids 30 2.0 6.0 0 40 0 36 0 30
read_positions {positions}
step_max 1
vtkPeriod 1
kernel compact
"""


def test_driver_geometry_tokens_on_device(tmp_path):
    """generate_solids (rand() radii and centres) and read_positions through
    the driver: the VTK Solids / Zones blocks match the host voxeliser's."""
    pos = tmp_path / "pos.txt"
    rng = np.random.default_rng(4)
    H.write_positions(pos, np.column_stack([rng.uniform(0, 30, (12, 3)), rng.uniform(1.5, 4.0, 12)]))
    out = {}
    for dev in (0, 1):
        folder = tmp_path / f"dev{dev}"
        folder.mkdir()
        inp = tmp_path / f"in{dev}.txt"
        inp.write_text(INPUT.format(folder=folder, positions=pos, dev=dev))
        print(f"case voxeliser driver device_voxeliser={dev}", flush=True)
        assert H.load().mplbm_host_run(str(inp).encode()) == 0, H.load().mplbm_host_last_error()
        out[dev] = (folder / "result000000.vtk").read_bytes()
    assert out[0] == out[1]

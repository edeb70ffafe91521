"""Host harness (libmplbm_host.so): input parser on the reference's example
files, sphere voxeliser, positions loader, srand(1) density
initialisation, VTK round trip and cluster statistics (CPU only)."""
import ctypes
import math
import os

import numpy as np
import pytest

from multiphase_lbm_amd import harness as H

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_parse_example_txt():
    d = H.parse_input(os.path.join(GOLDEN, "example.txt"))
    assert d["result_folder"] == "./TEST_RESULT"
    assert d["statPeriod"] == 10 and d["num_threads"] == 8
    assert d["save_vtk_ascii"] == 0  # the vtkPeriod line is commented out
    assert (d["lx"], d["ly"], d["lz"]) == (200, 200, 200)  # lxSample/lySample/lzSample
    assert d["rho0"] == 0.2 and d["IniPerturbRate"] == 1.0 and d["rhos"] == 0.35 and d["tau"] == 1.0
    assert d["G"] == -1.0  # "!G -6.12" is a comment; the constructor default stays
    assert d["carn_star"] == 1 and d["T_Tc"] == 0.7
    assert d["rhol_spinodal"] == 0.2725 and d["rhog_spinodal"] == 0.0484
    assert d["rho_inject_period"] == 100 and d["rho_increment"] == 0.005 and d["inject_density"] == 1
    assert d["inject_type"] == 0 and d["inject_position"] == 1.0
    assert d["erosion"] == 1 and d["target_density"] == 0.49
    assert d["step_max"] == 60000000
    assert d["unknown_tokens"] == 0


def test_parse_input_inject_txt():
    d = H.parse_input(os.path.join(GOLDEN, "input_inject.txt"))
    assert d["result_folder"] == "./hyteresis_res2"
    assert d["vtkPeriod"] == 5000 and d["save_vtk_ascii"] == 1
    assert d["read_densities"] == 1 and d["densities_filename"] == "./hyteresis_res2_/result000002.vtk"
    assert d["rho0"] == 0.15 and d["rho_inject_period"] == 200 and d["rho_increment"] == 0.0015
    assert d["hysteresis"] == 1 and d["max_saturation"] == 0.95 and d["min_saturation"] == 0.3
    assert d["saturation_cycle"] == 0  # the reference's input never changes the injection direction
    assert d["carn_star"] == 1


def test_parse_extension_tokens(tmp_path):
    p = tmp_path / "in.txt"
    p.write_text("lx 32 ly 16 lz 8 # sizes\nstorage fp16s\nkernel twopass\npressure_x 2.0 0.15\n"
                 "body_force 1e-6 0 0\nbogus_token\ncheckpointPeriod 500\nread_checkpoint run/ck.mplbm\n"
                 "host_statistics 1\nrounding stochastic\nshift node\n")
    d = H.parse_input(p)
    assert (d["lx"], d["ly"], d["lz"]) == (32, 16, 8)
    assert d["storage"] == 2 and d["kernel"] == 2
    assert d["face_type"][:2] == [1, 1] and d["face_rho"][:2] == [2.0, 0.15]
    assert d["body_force"] == [1e-6, 0.0, 0.0]
    assert d["unknown_tokens"] == 1
    assert d["checkpointPeriod"] == 500 and d["read_checkpoint"] == 1
    assert d["checkpoint_filename"] == "run/ck.mplbm" and d["host_statistics"] == 1
    assert d["rounding"] == 1 and d["shift"] == 1


def test_place_sphere_integer_centre_is_the_open_ball():
    """With an integer centre and radius the sample points c - R + k are the
    integer lattice, so the marked voxels are exactly {v : |v - c|^2 < R^2}."""
    shape = (20, 20, 20)
    X, Y, Z = np.meshgrid(*[np.arange(20)] * 3, indexing="ij")
    for R in (1.0, 4.0, 6.0):
        m = np.ones(shape, np.uint8)
        H.place_sphere(m, 10.0, 9.0, 11.0, R)
        want = (X - 10) ** 2 + (Y - 9) ** 2 + (Z - 11) ** 2 < R * R
        assert np.array_equal(m == 0, want), R


def test_place_sphere_fractional_radius_brute_force():
    """Sample points c - R + k (exact in float for these dyadic values) that
    fall strictly inside the sphere mark voxel floor(s + 0.5)."""
    shape = (20, 20, 20)
    c = np.array([10.25, 9.5, 11.75])
    R = 3.5
    m = np.ones(shape, np.uint8)
    H.place_sphere(m, *c, R)
    want = np.zeros(shape, bool)
    k = np.arange(0, int(2 * R) + 1)
    for sx in c[0] - R + k:
        for sy in c[1] - R + k:
            for sz in c[2] - R + k:
                if (sx - c[0]) ** 2 + (sy - c[1]) ** 2 + (sz - c[2]) ** 2 < R * R:
                    want[int(np.floor(sx + 0.5)) % 20, int(np.floor(sy + 0.5)) % 20, int(np.floor(sz + 0.5)) % 20] = True
    assert np.array_equal(m == 0, want)


def test_place_sphere_wraps_once_and_numbers_solids():
    shape = (16, 16, 16)
    m = np.ones(shape, np.uint8)
    sn = np.zeros(shape, np.int64)
    H.place_sphere(m, 0.0, 0.0, 15.0, 3.0, solid_number=sn, solid_id=7)
    # the ball around a corner appears in all eight corners of the periodic box
    for x in (0, 15):
        for y in (0, 15):
            for z in (0, 15):
                assert m[x, y, z] == 0 and sn[x, y, z] == 7
    assert (m == 0).sum() == ((np.arange(-3, 4)[:, None, None] ** 2 + np.arange(-3, 4)[None, :, None] ** 2 +
                               np.arange(-3, 4)[None, None, :] ** 2) < 9).sum()


def test_place_sphere_volume_converges():
    shape = (64, 64, 64)
    m = np.ones(shape, np.uint8)
    H.place_sphere(m, 31.37, 30.81, 32.29, 20.0)
    v = (m == 0).sum()
    assert abs(v - 4 / 3 * math.pi * 20 ** 3) / v < 0.01


def test_positions_file_packing_density():
    """The reference's 767-sphere file (radii 2.5-7.5 in a 200^3 box,
    porosity 0.5 by its name) voxelises to a packing density near 0.5."""
    shape = (200, 200, 200)
    m, sn, n = H.read_positions(os.path.join(GOLDEN, "example_positions_file"), shape, res=1.0)
    assert n == 767
    pd = (m == 0).mean()
    assert abs(pd - 0.5) < 0.01
    assert sn.max() == 767 and sn.min() == 0


def test_rho_init_follows_srand1_sequence():
    """resize_gpu draws rand() after srand(1) in z, y, x loop order
    (src/multiphase_lbm.cu:134-144)."""
    shape = (3, 2, 2)
    rho = H.rho_init(shape, 0.7, 0.02)
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    seq = [libc.rand() for _ in range(12)]
    k = 0
    for z in range(2):
        for y in range(2):
            for x in range(3):
                r = np.float32(seq[k]) / np.float32(2147483647)
                want = np.float32(0.7 * (1.0 + float(np.float32(0.02)) * (float(r) - 0.5)))
                want = np.float32(float(np.float32(0.7)) * (1.0 + float(np.float32(0.02)) * (float(r) - 0.5)))
                assert rho[x, y, z] == want
                k += 1


def test_vtk_roundtrip(tmp_path):
    shape = (5, 4, 3)
    rng = np.random.default_rng(0)
    fluid = (rng.uniform(size=shape) > 0.2).astype(np.uint8)
    rho = rng.uniform(0.1, 2.0, shape).astype(np.float32)
    p = rng.uniform(0.0, 0.1, shape).astype(np.float32)
    path = tmp_path / "result000000.vtk"
    H.write_vtk(path, fluid, rho, p)
    text = path.read_text().splitlines()
    assert text[0] == "# vtk DataFile Version 2.0" and text[3] == "DATASET RECTILINEAR_GRID"
    assert text[4] == "DIMENSIONS 5 4 3"
    assert text[6].split()[1] == "2.0000e-01"  # x coordinate i/lx
    back, n = H.read_densities(path, shape)
    assert n == 60
    want = np.where(fluid == 1, np.array([float(f"{v:.4e}") for v in rho.ravel()]).reshape(shape), 0.0)
    assert np.array_equal(back, want.astype(np.float32))


def test_vtr_binary_roundtrip(tmp_path):
    """The binary VTR file holds exactly the arrays, in x-fastest point
    order, with write_vtk's coordinates; its XML is well formed."""
    import xml.etree.ElementTree as ET
    shape = (5, 4, 3)
    rng = np.random.default_rng(1)
    fluid = (rng.uniform(size=shape) > 0.2).astype(np.uint8)
    rho = rng.uniform(0.1, 2.0, shape).astype(np.float32)
    p = rng.uniform(0.0, 0.1, shape).astype(np.float32)
    liq = rng.integers(0, 5, shape).astype(np.int64)
    gas = rng.integers(0, 3, shape).astype(np.int64)
    sn = np.where(fluid == 0, rng.integers(1, 9, shape), 0).astype(np.int64)
    path = tmp_path / "result000000.vtr"
    H.write_vtr(path, fluid, rho, p, liq, gas, sn)
    d = H.read_vtr(path)
    assert np.array_equal(d["Density"], np.where(fluid == 1, rho, 0).astype(np.float32))
    assert np.array_equal(d["Pressure"], np.where(fluid == 1, p, 0).astype(np.float32))
    assert np.array_equal(d["Zones"], (fluid == 0).astype(np.int32))
    assert np.array_equal(d["LiqClusters"], liq) and np.array_equal(d["GasClusters"], gas)
    assert np.array_equal(d["Solids"], sn)
    pas = np.float32(1.0 / shape[0])
    assert np.array_equal(d["x"], np.arange(5, dtype=np.float32) * pas)
    assert np.array_equal(d["z"], np.arange(3, dtype=np.float32) * pas)
    raw = path.read_bytes()
    head = raw[:raw.index(b"<AppendedData")] + b"</VTKFile>"
    root = ET.fromstring(head.replace(b"</VTKFile>", b"") + b"</VTKFile>")
    assert root.attrib["type"] == "RectilinearGrid" and root.attrib["header_type"] == "UInt64"
    assert raw.rstrip().endswith(b"</VTKFile>")


def test_cluster_stats_hand_built():
    shape = (12, 12, 12)
    fluid = np.ones(shape, np.uint8)
    fluid[0, :, :] = 0  # a solid wall
    rho = np.full(shape, 0.1, np.float32)
    rho[3:6, 3:6, 3:6] = 1.5      # 27-node liquid cluster touching nothing solid
    rho[1:3, 8:10, 8:10] = 1.5    # 8-node cluster touching the wall at x = 0
    rho[9, 1, 1] = 1.5            # single node: removed (< 7 nodes)
    p = np.where(rho > 1, 0.05, 0.01).astype(np.float32)
    sn = np.zeros(shape, np.int64)
    sn[0, :, :] = 4
    s, liq, gas = H.cluster_stats(fluid, rho, p, sn, rhol=1.0, rhog=0.2, solid_count=4, want_labels=True)
    assert s["liq_cluster_count"] == 2
    assert s["liquid_count"] == 27 + 8 + 1 and s["fluid_count"] == 11 * 144
    assert s["gas_cluster_count"] == 1
    # clusters are numbered in x-major scan order (src/multiphase_lbm.cu:322-338)
    assert liq[9, 1, 1] == 0 and liq[2, 8, 8] == 1 and liq[4, 4, 4] == 2
    assert abs(s["saturation"] - float(np.float32(36) / np.float32(11 * 144))) < 1e-12
    assert abs(s["avg_liq_press"] - 0.05) < 1e-7 and abs(s["avg_vap_rho"] - 0.1) < 1e-7
    # only cluster 2 touches solid 4: order = 1 solid / 1 cluster-with-solids
    assert s["cluster_order"] == 1.0 and s["wet_coord_num"] == 0.25
    assert s["min_rho"] == np.float32(0.1) and s["max_rho"] == np.float32(1.5)


def test_single_cluster_radius_across_periodic_boundary():
    shape = (16, 16, 16)
    fluid = np.ones(shape, np.uint8)
    rho = np.full(shape, 0.1, np.float32)
    X, Y, Z = np.meshgrid(*[np.arange(16)] * 3, indexing="ij")
    dx = np.minimum(np.abs(X - 0), 16 - np.abs(X - 0))
    ball = dx ** 2 + (Y - 8) ** 2 + (Z - 8) ** 2 <= 9
    rho[ball] = 1.5
    s = H.cluster_stats(fluid, rho, rho, rhol=1.0, rhog=0.2)
    assert s["liq_cluster_count"] == 1
    assert abs(s["cluster_radius"] - 3.0) < 1.01


def test_generate_packing_targets_porosity():
    m, sph = H.generate_packing((64, 64, 64), 7, 10 ** 6, 3.0, 6.0, 0.45)
    assert abs(m.mean() - 0.45) < 0.01
    m2, sph2 = H.generate_packing((64, 64, 64), 7, 40, 3.0, 6.0, 0.0, nonoverlap=True)
    assert len(sph2) == 40
    c = sph2[:, :3]
    r = sph2[:, 3]
    for i in range(40):
        d = np.abs(c - c[i])
        d = np.minimum(d, 64 - d)
        dist = np.sqrt((d ** 2).sum(1))
        dist[i] = np.inf
        assert np.all(dist >= r + r[i])
    # reproducible from the seed
    m3, sph3 = H.generate_packing((64, 64, 64), 7, 40, 3.0, 6.0, 0.0, nonoverlap=True)
    assert np.array_equal(sph2, sph3) and np.array_equal(m2, m3)


def test_positions_written_and_read_back(tmp_path):
    m, sph = H.generate_packing((32, 32, 32), 3, 10, 2.0, 4.0, 0.0, nonoverlap=True)
    p = tmp_path / "pos.txt"
    H.write_positions(p, sph)
    m2, _, n = H.read_positions(p, (32, 32, 32))
    assert n == len(sph) and np.array_equal(m, m2)


def _write_sample(path, vox):
    """A read_sample voxel file: header 'lx ly lz', then one id per line in
    [x][y][z] order with z fastest, 1 = solid."""
    fx, fy, fz = vox.shape
    with open(path, "w") as f:
        f.write(f"{fx} {fy} {fz}\n")
        f.write("\n".join(str(int(v)) for v in vox.reshape(-1)))
        f.write("\n")


@pytest.mark.parametrize("file_shape,shape", [((6, 5, 7), (6, 5, 7)), ((12, 10, 14), (6, 5, 7)),
                                              ((9, 9, 12), (3, 3, 4))])
def test_read_sample_downsampling(tmp_path, file_shape, shape):
    """read_sample (src/multiphase_lbm.cu:750-806): ids 1 = solid; every
    voxel whose x, y and z are multiples of lx_file / lx is kept, in file
    order, then laid out over the domain [x][y][z]."""
    rng = np.random.default_rng(sum(file_shape))
    vox = (rng.random(file_shape) < 0.4).astype(np.int64)
    p = tmp_path / "sample.txt"
    _write_sample(p, vox)
    m = H.read_sample(p, shape)
    ds = file_shape[0] // shape[0]
    kept = vox[::ds, ::ds, ::ds]
    assert kept.shape == shape
    assert np.array_equal(m, (1 - kept).astype(np.uint8))


def test_read_sample_refusals(tmp_path):
    p = tmp_path / "short.txt"
    _write_sample(p, np.zeros((4, 4, 4), np.int64))
    with pytest.raises(RuntimeError, match="smaller than the domain"):
        H.read_sample(p, (8, 8, 8))
    q = tmp_path / "trunc.txt"
    q.write_text("4 4 4\n0\n1\n")
    with pytest.raises(RuntimeError, match="kept voxels"):
        H.read_sample(q, (4, 4, 4))
    with pytest.raises(RuntimeError, match="Cannot open"):
        H.read_sample(tmp_path / "missing.txt", (4, 4, 4))


# --- erosion (erode_solid_touch_liquid, src/multiphase_lbm.cu:842-890) -----

EROSION_INPUT = """result_folder {folder}
lx 24
ly 20
lz 16
rho0 0.7
IniPerturbRate 0.0285714
resize
read_positions {positions}
{extra}
"""

# three spheres x, y, z, r (read_positions computes the packing density that
# the erosion count starts from, :684-691; place_sphere does not)
EROSION_SPHERES = [(8, 9, 7, 6.5), (18, 6, 10, 5), (3, 17, 2, 4)]

EX, EY, EZ = (np.array(v) for v in (
    [0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0],
    [0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1],
    [0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1]))


def _touching(fluid):
    """find_solid_touch_liquid (:810-838): a solid node with a fluid node at
    x - e_i for some i = 1..18, periodic."""
    up_fluid = np.zeros_like(fluid)
    for i in range(1, 19):
        up_fluid |= np.roll(fluid, shift=(EX[i], EY[i], EZ[i]), axis=(0, 1, 2))
    return up_fluid & ~fluid


def _erode_replay(fluid, target):
    """Brute-force erosion: the reference's float count, whole layers while
    the count exceeds the touching solids (recounted after each layer), then
    one random pass in z, y, x order on glibc rand() after the density
    initialisation's srand(1) and one draw per cell (:134-144).  Returns the
    mask, the number eroded and the number of whole layers."""
    fluid = fluid.copy()
    solids = int((~fluid).sum())
    pd = np.float32(solids) / np.float32(fluid.size)
    count = int((np.float32(1) - np.float32(target) / pd) * np.float32(solids))
    touch = _touching(fluid)
    done, layers = 0, 0
    while count > touch.sum():
        n = int(touch.sum())
        fluid |= touch
        count -= n
        done += n
        layers += 1
        touch = _touching(fluid)
        if not touch.any():
            break
    if 0 < count <= touch.sum():
        rate = np.float32(count / float(touch.sum()))
        libc = ctypes.CDLL("libc.so.6")
        libc.srand(1)
        for _ in range(fluid.size):
            libc.rand()
        for z in range(fluid.shape[2]):
            for y in range(fluid.shape[1]):
                for x in range(fluid.shape[0]):
                    if touch[x, y, z] and np.float32(libc.rand()) / np.float32(2147483647) < rate:
                        fluid[x, y, z] = True
                        done += 1
    return fluid, done, layers


def _read(tmp_path, tag, extra):
    pos = tmp_path / "spheres.txt"
    pos.write_text("".join(f"{v}\n" for sph in EROSION_SPHERES for v in sph))
    p = tmp_path / f"erosion_{tag}.txt"
    p.write_text(EROSION_INPUT.format(folder=tmp_path, positions=pos, extra=extra))
    return H.read_data(p)


@pytest.mark.parametrize("layers_wanted", [0, 1, 2])
def test_erosion_against_brute_force(tmp_path, layers_wanted):
    """0: the target lies within the first layer of touching solids (the
    reference's own path: one random pass whose draws follow its rand()
    stream); 1, 2: the target lies beyond one or two whole layers, where the
    count is recounted after each layer (the reference's size_t counter is
    not; see erode() in the harness)."""
    _, fluid0, _, e0 = _read(tmp_path, "none", "")
    assert e0 == 0
    fluid0 = fluid0.astype(bool)
    solids = int((~fluid0).sum())
    pd = np.float32(solids) / np.float32(fluid0.size)
    # layer sizes of successive whole-layer erosion
    sizes, f = [], fluid0.copy()
    for _ in range(3):
        t = _touching(f)
        sizes.append(int(t.sum()))
        f |= t
    # an erosion count halfway into the wanted layer
    want = sum(sizes[:layers_wanted]) + sizes[layers_wanted] // 2
    target = float(pd * np.float32(1 - want / solids))
    d, fluid, rho, eroded = _read(tmp_path, f"t{layers_wanted}", f"target_density {target!r}")
    assert d["erosion"] == 1
    expect, done, layers = _erode_replay(fluid0, d["target_density"])
    assert layers == layers_wanted
    assert np.array_equal(fluid.astype(bool), expect)
    assert eroded == done == int((fluid.astype(bool) & ~fluid0).sum())
    # eroded nodes are near the requested count (random pass: binomial)
    sd = math.sqrt(sizes[layers_wanted] * 0.25)
    assert abs(done - want) <= 5 * sd + 2
    assert np.all(fluid.astype(bool) >= fluid0)  # erosion only turns solids fluid
    # the density initialisation is untouched by the erosion
    assert np.array_equal(rho, H.rho_init(fluid.shape, 0.7, 0.0285714))


def test_random_sequential_addition_spheres_do_not_overlap():
    """The 64^3 workload's geometry: 40 spheres, radii in [3, 6], placed by
    seeded random sequential addition, pairwise periodic centre distance at
    least the sum of the radii; the same seed gives the same spheres."""
    from tests.helpers import config1_case, random_spheres
    shape = (64, 64, 64)
    sph = random_spheres(shape, 40, 3.0, 6.0, 1, nonoverlap=True)
    assert sph.shape == (40, 4)
    assert (sph[:, 3] >= 3.0).all() and (sph[:, 3] <= 6.0).all()
    L = np.array(shape, float)
    for i in range(40):
        for j in range(i):
            d = np.abs(sph[i, :3] - sph[j, :3])
            d = np.minimum(d, L - d)
            assert np.sqrt((d * d).sum()) >= sph[i, 3] + sph[j, 3]
    assert np.array_equal(sph, random_spheres(shape, 40, 3.0, 6.0, 1, nonoverlap=True))
    _, fluid, rho = config1_case()
    # 40 disjoint balls of radius 3..6 occupy 3.4%..28% of the box
    assert 0.70 < fluid.mean() < 0.97 and rho.dtype == np.float32

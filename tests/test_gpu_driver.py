"""The reference program end to end (src/main.cc:4-17: read_data, init,
run_gpu) through the host harness on the HIP library: statistics computed
on the device (mplbm_calc_cluster, the default) and on the host after a
field read (the reference's way, input token host_statistics 1) must write
the same Stats.txt rows and VTK files."""
import os

import numpy as np
import pytest

from multiphase_lbm_amd import harness

pytestmark = pytest.mark.gpu

INPUT = """result_folder {folder}
statPeriod 40
vtkPeriod 80
lx 36
ly 30
lz 28
rho0 0.7
IniPerturbRate 0.0285714
rhos 0.7
G -5.0
tau 1.0
rhol_spinodal 1.285931
rhog_spinodal 0.323507
resize
place_sphere 9 8 7 5
place_sphere 27 20 21 6
place_sphere 30 5 12 4
place_sphere 3 25 25 4.5
step_max {step_max}
kernel {kernel}
host_statistics {host}
{extra}
"""


def _run(tmp_path, kernel, host, step_max=241, extra="", tag=""):
    folder = tmp_path / f"{kernel}_{host}{tag}"
    folder.mkdir()
    inp = tmp_path / f"in_{kernel}_{host}{tag}.txt"
    inp.write_text(INPUT.format(folder=folder, kernel=kernel, host=host, step_max=step_max, extra=extra))
    print(f"case driver kernel={kernel} host_statistics={host} {extra!r}", flush=True)
    assert harness.load().mplbm_host_run(str(inp).encode()) == 0, harness.load().mplbm_host_last_error()
    return folder


def _rows(folder):
    lines = (folder / "Stats.txt").read_text().splitlines()
    return [ln.split("\t") for ln in lines[1:] if ln.strip()]


@pytest.mark.parametrize("kernel", ["compact", "fused"])
def test_device_and_host_statistics_agree(tmp_path, kernel):
    dev = _run(tmp_path, kernel, 0)
    host = _run(tmp_path, kernel, 1)
    rd, rh = _rows(dev), _rows(host)
    assert len(rd) == len(rh) == 7  # steps 0, 40, ..., 240
    for a, b in zip(rd, rh):
        assert a == b, (a, b)
    vtks = sorted(f for f in os.listdir(dev) if f.endswith(".vtk"))
    assert vtks == sorted(f for f in os.listdir(host) if f.endswith(".vtk")) and len(vtks) == 4
    for f in vtks:
        assert (dev / f).read_bytes() == (host / f).read_bytes(), f


def _vtk_arrays(path, shape):
    """Point arrays of an ASCII VTK written by write_vtk, as [x][y][z]."""
    lines = path.read_text().splitlines()
    n = shape[0] * shape[1] * shape[2]
    out = {}
    for k, ln in enumerate(lines):
        if ln.startswith("SCALARS "):
            name = ln.split()[1]
            vals = np.array([float(v) for v in lines[k + 2:k + 2 + n]])
            out[name] = vals.reshape(shape[::-1]).transpose(2, 1, 0)
    return out


def test_checkpoint_resume_and_binary_vtr(tmp_path):
    # uninterrupted run with VTR files next to the VTK files
    full = _run(tmp_path, "compact", 0, extra="vtrPeriod 80", tag="_full")
    # the same run saved at step 120, then resumed from the checkpoint
    first = _run(tmp_path, "compact", 0, step_max=121, extra="checkpointPeriod 120", tag="_a")
    ck = first / "checkpoint.mplbm"
    assert ck.exists()
    resumed = _run(tmp_path, "compact", 0, extra=f"read_checkpoint {ck}", tag="_b")
    rows_full = {r[0]: r for r in _rows(full)}
    rows_res = _rows(resumed)
    assert [r[0] for r in rows_res] == ["120", "160", "200", "240"]
    for r in rows_res:
        assert r == rows_full[r[0]], (r, rows_full[r[0]])
    # VTR holds the VTK arrays (the VTK rounds floats to 5 digits)
    shape = (36, 30, 28)
    for k in range(4):
        v = harness.read_vtr(full / f"result{k:06d}.vtr")
        a = _vtk_arrays(full / f"result{k:06d}.vtk", shape)
        for name in ("Zones", "LiqClusters", "GasClusters", "Solids"):
            assert np.array_equal(v[name], a[name].astype(np.int64)), name
        for name in ("Density", "Pressure"):
            assert np.allclose(v[name], a[name], rtol=6e-5, atol=1e-30), name


HYSTERESIS = """result_folder {folder}
statPeriod 25
lx 32
ly 32
lz 32
rho0 0.7
IniPerturbRate 0.0285714
rhos 0.7
G -5.0
tau 1.0
rhol_spinodal 1.285931
rhog_spinodal 0.323507
resize
place_sphere 16 16 16 6
rho_inject_period 5
rho_increment 0.004
inject_type 0
inject_position 1
max_saturation 0.5
min_saturation 0.3
step_max {step_max}
kernel compact
{extra}
"""


def _run_hyst(tmp_path, tag, step_max=3001, extra="saturation_cycle 1"):
    folder = tmp_path / f"hyst{tag}"
    folder.mkdir()
    inp = tmp_path / f"hyst{tag}.txt"
    inp.write_text(HYSTERESIS.format(folder=folder, step_max=step_max, extra=extra))
    print(f"case driver hysteresis {tag} {extra!r}", flush=True)
    assert harness.load().mplbm_host_run(str(inp).encode()) == 0, harness.load().mplbm_host_last_error()
    return folder


def _sat(folder):
    return np.array([float(r[8]) for r in _rows(folder)])


def test_hysteresis_cycles_injection_and_drainage(tmp_path):
    """max_saturation / min_saturation with saturation_cycle 1 (the
    reference parses both and never acts on them): the driver injects until
    the saturation reaches the maximum, drains to the minimum, then injects
    again.  A checkpoint taken while draining resumes draining, so the
    resumed run writes the uninterrupted run's Stats.txt rows.  Without
    saturation_cycle the reference's input keeps injecting."""
    full = _run_hyst(tmp_path, "_full")
    sat = _sat(full)
    steps = [int(r[0]) for r in _rows(full)]
    assert np.isfinite(sat).all()
    i = int(np.argmax(sat >= 0.5))
    assert sat[i] >= 0.5, "never reached max_saturation"
    j = i + int(np.argmax(sat[i:] <= 0.3))
    assert sat[j] <= 0.3, "never drained to min_saturation"
    assert sat[j:].max() > sat[j] + 0.05, "injection did not resume"
    # checkpoint two statistics steps into the drainage, then resume
    assert j >= i + 3
    ck = steps[i + 2]
    first = _run_hyst(tmp_path, "_a", step_max=ck + 1, extra=f"saturation_cycle 1\ncheckpointPeriod {ck}")
    assert (first / "checkpoint.mplbm").exists()
    assert (first / "checkpoint.mplbm.driver").read_text().split() == ["reverse", "-1"]
    resumed = _run_hyst(tmp_path, "_b", extra=f"saturation_cycle 1\nread_checkpoint {first / 'checkpoint.mplbm'}")
    rows_full = {r[0]: r for r in _rows(full)}
    rows_res = _rows(resumed)
    assert rows_res[0][0] == str(ck) and len(rows_res) == len(steps) - (i + 2)
    for r in rows_res:
        assert r == rows_full[r[0]], (r, rows_full[r[0]])
    # the reference's behaviour: the same input without saturation_cycle
    plain = _run_hyst(tmp_path, "_plain", step_max=steps[i + 4] + 1, extra="")
    sp = _sat(plain)
    assert np.array_equal(sp[:i + 1], sat[:i + 1])
    assert sp[-1] > sp[i] and sp[-1] > sat[i + 4], "drained without saturation_cycle"

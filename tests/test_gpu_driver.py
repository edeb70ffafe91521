"""The reference program end to end (src/main.cc:4-17: read_data, init,
run_gpu) through the host harness on the HIP library: statistics computed
on the device (mplbm_calc_cluster, the default) and on the host after a
field read (the reference's way, input token host_statistics 1) must write
the same Stats.txt rows and VTK files."""
import os

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
step_max 241
kernel {kernel}
host_statistics {host}
"""


def _run(tmp_path, kernel, host):
    folder = tmp_path / f"{kernel}_{host}"
    folder.mkdir()
    inp = tmp_path / f"in_{kernel}_{host}.txt"
    inp.write_text(INPUT.format(folder=folder, kernel=kernel, host=host))
    print(f"case driver kernel={kernel} host_statistics={host}", flush=True)
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

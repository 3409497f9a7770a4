"""GPU tests of the raw-plane command line (autorift.jl_amd/bin/autorift, the reference's
app/src/AutoRIFTApp.jl contract) and of the map conventions end to end."""
import os
import subprocess

import numpy as np
import pytest

from _util import shifted_pair

pytestmark = pytest.mark.gpu

autorift = pytest.importorskip("autorift")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "autorift.jl_amd", "bin", "autorift")


@pytest.mark.parametrize("chip,radius", [(None, None), (16, 10)])
def test_app_raw_planes_equal_the_api(tmp_path, chip, radius):
    ref, sec = shifted_pair(600, (3.0, -2.0), seed=21, dtype=np.float32)
    ref, sec = ref * 255, sec * 255
    ny, nx = ref.shape
    fr, fs, fo = tmp_path / "ref.f32", tmp_path / "sec.f32", tmp_path / "out.f32"
    np.asfortranarray(ref).T.tofile(fr)      # column-major bytes: the transpose in C order
    np.asfortranarray(sec).T.tofile(fs)
    args = [APP, str(fr), str(fs), str(ny), str(nx), str(fo)]
    kw = {}
    if chip is not None:
        args += [str(chip), str(radius)]
        kw = dict(chip_size=chip, search_radius=radius)
    print("app", args[3:], flush=True)
    subprocess.run(args, check=True, timeout=120)
    exp = autorift.autorift(ref, sec, **kw)
    nr, nc = exp.dx.shape
    planes = np.fromfile(fo, dtype=np.float32)
    assert planes.size == 3 * nr * nc
    got = [planes[k * nr * nc:(k + 1) * nr * nc].reshape((nr, nc), order="F") for k in range(3)]
    for g, e in zip(got, (exp.dx, exp.dy, exp.correlation)):
        np.testing.assert_array_equal(g.view(np.uint32), np.asarray(e, np.float32).view(np.uint32))
    assert np.nanmedian(got[0]) == -3.0 and np.nanmedian(got[1]) == 2.0


def test_app_argument_errors(tmp_path):
    r = subprocess.run([APP], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "autorift <reference>" in r.stdout
    r = subprocess.run([APP, "a", "b", "10", "x", "o"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "positive integers" in r.stdout
    r = subprocess.run([APP, "a", "b", "10", "10", "o", "30"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "multiple of 4" in r.stdout


def test_geo_north_up_velocity():
    """Features moved 6 px east and 4 px south on a north-up raster: vx = 6, vy = -4, and the
    south-up storage of the same scene gives the same field (test/extensions.jl:48-109)."""
    ref, sec = shifted_pair(400, (6.0, 4.0), seed=5, dtype=np.float32)
    ref, sec = ref * 255, sec * 255
    y = 3_000_000.0 - 30.0 * np.arange(400)
    x = 500_000.0 + 30.0 * np.arange(400)
    print("geo north-up", flush=True)
    north = autorift.autorift_geo(ref, sec, y, x, grid_spacing=16)
    assert np.nanmedian(north["vx"]) == 6.0 and np.nanmedian(north["vy"]) == -4.0
    south = autorift.autorift_geo(ref[::-1].copy(), sec[::-1].copy(), y[::-1].copy(), x,
                                  grid_spacing=16)
    assert np.nanmedian(south["vx"]) == 6.0 and np.nanmedian(south["vy"]) == -4.0
    assert np.all(np.diff(north["y"]) < 0) and np.all(np.diff(south["y"]) > 0)
    st = autorift.autorift_geo(ref, sec, y, x, grid_spacing=16, dt=np.timedelta64(16, "D"))
    yrs = 16 / 365.25
    assert np.nanmedian(st["vx"]) == pytest.approx(6 * 30 / yrs, rel=1e-5)
    assert np.nanmedian(st["vy"]) == pytest.approx(-4 * 30 / yrs, rel=1e-5)

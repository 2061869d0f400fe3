"""C-ABI checks that need no GPU: the library loads, exports every entry
point the header declares, validates arguments before touching a device,
and its host-side drift/kick tables agree with the closed forms."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

from cuco_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cuco_engine.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cuco_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = E.lib()
    names = declared_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", E.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (cuco_[a-z0-9_]+)", out))
    assert set(names) <= exported
    assert set(E.EXPORTED) == set(names)


def test_library_has_gfx950_code_object():
    out = subprocess.run(["readelf", "-S", E.LIB_PATH], capture_output=True,
                         text=True, check=True)
    assert ".hip_fatbin" in out.stdout
    blob = open(E.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_status_strings_and_null_handle():
    lib = E.lib()
    assert lib.cuco_status_string(0) == b"ok"
    assert lib.cuco_status_string(1) == b"invalid argument"
    rc = lib.cuco_pp_force(None, 0)
    assert rc == 1
    assert b"null engine" in lib.cuco_last_error()


@pytest.mark.parametrize("bad", [
    dict(pm_grid=10),                       # below 16
    dict(pm_grid=14 * 16),                  # prime factor 7
    dict(pm_grid=2048),                     # beyond 1024
    dict(box_size=-1.0),
    dict(soft_h=0.0),
    dict(rcut=600.0),                       # cell range would wrap onto itself
])
def test_create_rejects_bad_parameters_without_device(bad):
    kw = dict(pm_grid=64, box_size=1000.0, G=1.0, asmth=19.53, rcut=87.9,
              soft_h=1.0)
    kw.update(bad)
    with pytest.raises(E.CucoError) as ei:
        E.Engine(max_particles=100, **kw)
    assert ei.value.status in (1, 5)
    good = dict(pm_grid=64, box_size=1000.0, G=1.0, asmth=19.53, rcut=87.9,
                soft_h=1.0)
    with pytest.raises(E.CucoError) as ei:
        E.Ensemble([good, kw], [100, 100])   # one bad box fails the ensemble
    assert ei.value.status in (1, 5)


@pytest.mark.parametrize("counts", [[], [100, 0], [100, -5]])
def test_ensemble_create_rejects_bad_counts_without_device(counts):
    good = dict(pm_grid=64, box_size=1000.0, G=1.0, asmth=19.53, rcut=87.9,
                soft_h=1.0)
    with pytest.raises((E.CucoError, ValueError)):
        E.Ensemble([good] * len(counts), counts)


def test_drift_tables_einstein_de_sitter():
    a0, a1, H0 = 0.02, 0.05, 0.1
    d, k = E.init_drift_table(a0, a1, 1.0, 0.0, H0)
    a = np.exp(np.log(a0) + (np.log(a1) - np.log(a0)) / 1000 * (np.arange(1000) + 1))
    np.testing.assert_allclose(d, 2 / H0 * (a0 ** -0.5 - a ** -0.5), rtol=1e-11)
    np.testing.assert_allclose(k, 2 / H0 * (a ** 0.5 - a0 ** 0.5), rtol=1e-11)


def test_drift_tables_lcdm_match_scipy_and_oracle(orc, golden_dir):
    g = np.load(f"{golden_dir}/drift_kick.npz")
    a0, a1 = float(g["a0"]), float(g["a1"])
    args = (a0, a1, float(g["omega0"]), float(g["omega_lambda"]), float(g["hubble"]))
    d, k = E.init_drift_table(*args)
    od, ok = orc.drift_tables(*args)
    np.testing.assert_allclose(d, od, rtol=1e-11)
    np.testing.assert_allclose(k, ok, rtol=1e-11)
    for av, dr, kr in zip(g["avals"], g["drift"], g["kick"]):
        assert E.table_factor(d, a0, a1, math.log(a0), math.log(av)) == pytest.approx(dr, rel=1e-9)
        assert E.table_factor(k, a0, a1, math.log(a0), math.log(av)) == pytest.approx(kr, rel=1e-9)


def test_kdk_factors_sum_to_whole_interval():
    a0, a1 = 0.02, 0.05
    d, k = E.init_drift_table(a0, a1, 0.3, 0.7, 0.1)
    f = E.kdk_factors(d, k, a0, a1, a0, a1, 10)
    assert f.shape == (10, 3)
    la0, la1 = math.log(a0), math.log(a1)
    assert f[:, 1].sum() == pytest.approx(E.table_factor(d, a0, a1, la0, la1), rel=1e-12)
    assert (f[:, 0] + f[:, 2]).sum() == pytest.approx(E.table_factor(k, a0, a1, la0, la1), rel=1e-12)
    assert np.all(f > 0)


def test_particle_record_layout():
    assert C.sizeof(E.Particle) == 60          # allvars.h:35-45
    assert E.PARTICLE_DTYPE.itemsize == 60
    assert E.Particle.GravPM.offset == 40

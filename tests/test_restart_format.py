"""CPU tests of the restart file layout (palinstrophy_amd.restart): the
reference's variable/dimension/attribute names (turbo_main.py:1283-1338),
round trips, the FP64 extras, and the velocity written into ``uc``."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from palinstrophy_amd import restart as R

ATTRS = {"format_version": 1, "Nbase": 16, "Re": 1.0e4, "K0": 10.0, "start_spectrum": "KM3", "visc": 1.0e-4,
         "cflnum": 0.75, "seed_init": 4895, "t": 0.125, "dt": 3.94060425090338607e-03, "cn": 0.97,
         "cnm1": 0.96, "it": 500}


def rand_c(shape, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def test_layout_names_and_roundtrip(tmp_path):
    p = str(tmp_path / R.RESTART_FILE)
    om, fn = rand_c((16, 8), 1), rand_c((16, 8), 2)
    uc = R.velocity_from_omega(om)
    R.write_restart_file(p, ATTRS, {"uc": uc, "om2": om, "fnm1": fn})
    with netcdf_file(p, "r", mmap=False) as ds:
        assert set(ds.dimensions) == {"uc_z", "uc_k", "uc_component", "om2_z", "om2_x_half", "fnm1_z",
                                      "fnm1_x_half"}
        assert ds.dimensions["uc_k"] == 3 * 16 // 4 + 1 and ds.dimensions["uc_component"] == 3
        for n in ("uc", "om2", "fnm1"):
            assert ds.variables[f"{n}_real"].typecode() == "f"
            assert ds.variables[f"{n}_imag"].dimensions == ds.variables[f"{n}_real"].dimensions
        assert ds.variables["om2_real"].dimensions == ("om2_z", "om2_x_half")
    attrs, arr = R.read_restart_file(p)
    for k, v in ATTRS.items():
        assert type(attrs[k]) is type(v) and attrs[k] == v, k  # doubles round-trip exactly
    with netcdf_file(p, "r", mmap=False) as ds:
        assert np.asarray(ds.dt).dtype == np.float64 and np.asarray(ds.it).dtype == np.int32
    assert arr["om2"].dtype == np.complex64
    assert np.array_equal(arr["om2"], om.astype(np.complex64))
    assert np.array_equal(arr["fnm1"], fn.astype(np.complex64))
    assert np.array_equal(arr["uc"], uc)


def test_fp64_extras_restore_full_precision(tmp_path):
    p = str(tmp_path / R.RESTART_FILE)
    om, fn = rand_c((16, 8), 3), rand_c((16, 8), 4)
    R.write_restart_file(p, ATTRS, {"uc": R.velocity_from_omega(om), "om2": om, "fnm1": fn}, f8=("om2", "fnm1"))
    _, arr = R.read_restart_file(p)
    assert arr["om2"].dtype == np.complex128 and np.array_equal(arr["om2"], om)
    assert np.array_equal(arr["fnm1"], fn)
    with netcdf_file(p, "r", mmap=False) as ds:  # the float32 pairs a reference reader uses are still there
        assert np.array_equal(ds.variables["om2_real"][:], om.real.astype(np.float32))


def test_velocity_written_into_uc_is_the_curl_inverse_of_omega():
    N, H = 16, 8
    om = rand_c((N, H), 5)
    om[0, 0] = 0
    uc = R.velocity_from_omega(om).astype(np.complex128)
    u, w = uc[:, :H, 0], uc[:, :H, 1]
    kx = np.arange(H)[None, :]
    kz = np.array([z if z < N // 2 else z - N for z in range(N)])[:, None]
    assert np.max(np.abs(kx * u + kz * w)) < 1e-5  # divergence free
    curl = 1j * (kz * u - kx * w)  # omega = i(kz u - kx w)
    mask = (kz != 0) | (kx != 0)
    assert np.max(np.abs(curl - om)[mask]) < 1e-5 * np.max(np.abs(om))
    assert np.all(uc[:, H:, :] == 0) and np.all(uc[:, :, 2] == 0)


def test_hdf5_restart_is_refused_clearly(tmp_path):
    p = tmp_path / R.RESTART_FILE
    p.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(RuntimeError, match="netCDF4"):
        R.read_restart_file(str(p))


def test_missing_attribute_is_rejected(tmp_path):
    a = dict(ATTRS)
    del a["cnm1"]
    with pytest.raises(ValueError):
        R.write_restart_file(str(tmp_path / "x.nc"), a, {"om2": rand_c((16, 8), 6)})

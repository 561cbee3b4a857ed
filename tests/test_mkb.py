"""Solver control files (mkb.dat) as the reference's lsr_mkb_core_init reads
them (lsd_mkb/lsd_mkb_core/lss_mkb_core.c:148-705).  The fixtures are the
reference's own ctest control files (tests/golden/mkb/); the expected values
are what that routine stores for each of them: ILU codes split into precon and
k (:469-474), a negative block type selecting the storage (:464-468), the
multi-level solver clamped to one level (:427-440), overrides > 0 (:483-487)
and the defaults for an empty name (:237-266)."""
import os

import pytest

import modfem_amd as mf

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mkb")


def rd(name, *a):
    return mf.mkb_read(os.path.join(G, name), *a)


def test_jacobi_control_file():
    d = rd("laplace_std_in_cube.dat")
    assert (d["solver"], d["nr_level"], d["krylbas"]) == (1, 1, 50)
    assert (d["precon"], d["nr_prec"], d["block_type"], d["storage_type"], d["ilu_k"]) == (1, 1, 1, 1, 0)
    assert (d["max_iter"], d["conv_type"], d["conv_meas"]) == (500, 0, 1e-9)
    assert (d["monitor"], d["nr_pre_smooth"], d["nr_post_smooth"]) == (2, 1, 0)


def test_gs_and_crs_storage():
    d = rd("laplace_std.dat")
    assert (d["precon"], d["nr_prec"], d["block_type"], d["storage_type"]) == (2, 2, 1, 1)
    d = rd("laplace_std_in_box.dat")
    assert (d["krylbas"], d["precon"], d["nr_prec"]) == (20, 2, 1)
    assert (d["block_type"], d["storage_type"]) == (-1, 2)  # "-2": CRS storage
    assert (d["max_iter"], d["conv_type"], d["conv_meas"]) == (1000, 0, 1e-12)


def test_multilevel_file_clamped_and_short_output_line():
    d = rd("laplace_dg.dat")
    assert (d["solver"], d["nr_level"], d["krylbas"]) == (1, 1, 50)  # MULTI_GMRES -> GMRES
    assert (d["precon"], d["nr_prec"], d["block_type"]) == (2, 2, 2)
    d = rd("modfem_test_scripts.dat")
    assert (d["max_iter"], d["conv_type"], d["conv_meas"], d["monitor"]) == (50, 1, 1e-6, 2)


def test_overrides_and_defaults():
    d = rd("laplace_std_in_cube.dat", 77, 1, 1e-5, 3)
    assert (d["max_iter"], d["conv_type"], d["conv_meas"], d["monitor"]) == (77, 1, 1e-5, 3)
    d = rd("laplace_std_in_cube.dat", 0, 0, 0.0, 0)  # only values > 0 override
    assert (d["max_iter"], d["conv_type"], d["conv_meas"], d["monitor"]) == (500, 0, 1e-9, 2)
    d = mf.mkb_read("")
    assert (d["solver"], d["krylbas"], d["precon"], d["nr_prec"]) == (1, 50, 2, 1)
    assert (d["max_iter"], d["conv_type"], d["conv_meas"], d["storage_type"]) == (100, 0, 1e-6, -1)
    d = mf.mkb_read("", 300, 1, 1e-10, -1)
    assert (d["max_iter"], d["conv_type"], d["conv_meas"]) == (300, 1, 1e-10)


def test_ilu_codes_and_errors(tmp_path):
    f = tmp_path / "ilu.dat"
    f.write_text("FINE_LEVEL x\n1 1 30 c\n50 1 -2 c\n200 1 1e-10 c\n0 0 0\n")
    d = mf.mkb_read(str(f))
    assert (d["precon"], d["ilu_k"], d["storage_type"]) == (5, 0, 2)
    f.write_text("AMG\n3\n")
    with pytest.raises(RuntimeError):
        mf.mkb_read(str(f))
    with pytest.raises(RuntimeError):
        mf.mkb_read(str(tmp_path / "absent.dat"))

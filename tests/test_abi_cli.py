"""The C ABI library loads and exports every entry point include/fdt_xfibres.h
declares (no compute calls), and the xfibres_gpu driver's argument handling
matches the reference driver (xfibres_gpu.cc:27-99, xfibresoptions.cc:26-55)
on the paths that end before any GPU work."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fdt_xfibres.h")
CLI = os.path.join(ROOT, "fdt_amd", "bin", "xfibres_gpu")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*[\w\s\*]+?\b(fdt_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("fdt_create", "fdt_destroy", "fdt_set_acquisition", "fdt_load_data", "fdt_fit", "fdt_mcmc",
                 "fdt_finish", "fdt_get_outputs", "fdt_run", "fdt_lm_stage"):
        assert must in names


def test_library_exports_every_declared_symbol(fdt):
    lib = ctypes.CDLL(fdt.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(fdt.EXPORTS) == sorted(declared())


def test_library_exports_the_image_io(fdt):
    txt = open(os.path.join(ROOT, "include", "fdt_io.h")).read()
    names = sorted(set(re.findall(r"^\s*[\w\s\*]+?\b(fdt_\w+)\s*\(", txt, flags=re.M)))
    assert len(names) == 7, names
    lib = ctypes.CDLL(fdt.LIB_PATH)
    assert not [n for n in names if not hasattr(lib, n)]


def test_default_options_match_xfibres(fdt):
    o = fdt.default_opts()
    assert (o.nfibres, o.model, o.njumps, o.nburn, o.sampleevery, o.updateproposalevery, o.seed) == \
        (1, 1, 1250, 1000, 25, 40, 8665904)
    assert o.fudge == 1.0 and abs(o.R_prior_mean - 0.13) < 1e-7 and abs(o.R_prior_std - 0.03) < 1e-7
    assert o.R_prior_fudge == 0.0 and not (o.no_ard or o.all_ard or o.f0 or o.rician or o.nonlin)


def run_cli(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.fixture
def subj(tmp_path):
    (tmp_path / "S").mkdir()
    (tmp_path / "bvals").write_text("0 1000 1000 1000 1000 1000 1000 1000\n")
    (tmp_path / "bvecs").write_text("1 1 0 0 1 1 0 .5\n0 0 1 0 1 0 1 .5\n0 0 0 1 0 1 1 .5\n")
    return tmp_path


def test_cli_needs_four_positionals(subj):
    r = run_cli(["--data=x"], subj)
    assert r.returncode == 1
    assert "last 4 arguments" in r.stderr


def test_cli_missing_compulsory_exits_2(subj):
    r = run_cli(["--data=d", "-b", "bvals", "S", "0", "1", "10"], subj)  # no --mask, --bvecs
    assert r.returncode == 2


def test_cli_help_exits_2(subj):
    r = run_cli(["-h", "S", "0", "1", "10"], subj)
    assert r.returncode == 2


def test_cli_subject_not_a_directory(subj):
    r = run_cli(["--data=d", "--mask=m", "-b", "bvals", "-r", "bvecs", "--logdir=L", "--forcedir",
                 "nope", "0", "1", "10"], subj)
    assert r.returncode == 1
    assert "is not a directory" in r.stderr


def test_cli_bad_numbers(subj):
    base = ["--data=d", "--mask=m", "-b", "bvals", "-r", "bvecs", "--logdir=L", "--forcedir", "S"]
    r = run_cli(base + ["x", "1", "10"], subj)
    assert r.returncode == 1 and "NumThisPart" in r.stderr
    r = run_cli(base + ["0", "1", "0"], subj)
    assert r.returncode == 1 and "greater than 0" in r.stderr


def test_cli_bvals_bvecs_mismatch(subj):
    (subj / "bvals2").write_text("0 1000 1000\n")
    r = run_cli(["--data=d", "--mask=m", "-b", "bvals2", "-r", "bvecs", "--logdir=L", "--forcedir",
                 "S", "0", "1", "10"], subj)
    assert r.returncode == 1 and "must be the same" in r.stderr


def test_cli_unknown_option(subj):
    r = run_cli(["--bogus=1", "--data=d", "--mask=m", "-b", "bvals", "-r", "bvecs", "S", "0", "1", "10"], subj)
    assert r.returncode == 2 and "unknown option" in r.stderr


def test_cli_logdir_plus_suffix(subj):
    # without --forcedir an existing log directory name gets '+' appended (FSL Log::makeDir)
    (subj / "L").mkdir()
    r = run_cli(["--data=d", "--mask=m", "-b", "bvals", "-r", "bvecs", "--logdir=L", "nope", "0", "1", "10"], subj)
    assert "Log directory is: L+" in r.stdout
    assert (subj / "L+").is_dir()


def test_cli_gradnonlin_file_checked_before_gpu_work(subj):
    # a grad_dev part with fewer than 9 x nvox values is refused before the engine starts
    import numpy as np
    np.ones((8, 10)).tofile(subj / "data_0")
    np.zeros((9, 3)).tofile(subj / "grad_short")
    base = ["--data=data_0", "--mask=m", "-b", "bvals", "-r", "bvecs", "--logdir=L", "--forcedir"]
    r = run_cli(base + ["--gradnonlin=grad_short", "S", "0", "1", "10"], subj)
    assert r.returncode == 1 and "gradnonlin file grad_short holds fewer than 9 x 10" in r.stderr
    r = run_cli(base + ["--gradnonlin=missing", "S", "0", "1", "10"], subj)
    assert r.returncode == 1 and "cannot open gradnonlin file missing" in r.stderr


def test_cli_cpuinit_needs_nospat_without_a_grid(subj):
    # the part files carry no voxel grid: neighbour initialisation is refused before the engine starts
    import numpy as np
    np.ones((8, 10)).tofile(subj / "data_0")
    base = ["--data=data_0", "--mask=m", "-b", "bvals", "-r", "bvecs", "--logdir=L", "--forcedir"]
    r = run_cli(base + ["--cpuinit", "S", "0", "1", "10"], subj)
    assert r.returncode == 1 and "--cpuinit needs --nospat here" in r.stderr

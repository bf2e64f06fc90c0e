"""The pipeline's host-only parsers under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only): the NIfTI-1 reader/writer and mask
maps (fdt_nifti.cpp through include/fdt_io.h) over well-formed images of
every voxel type and byte order and over corrupt ones (the header fixtures of
tests/test_io_pipeline.py, truncated and garbage files, truncated gzip), and
the xfibres option and acquisition parsers (fdt_cli.h) over the drivers'
argument errors.  Every input must end in "OK" or "ERR" with no sanitizer
report."""
import gzip
import os
import struct
import subprocess

import numpy as np
import pytest

import nifti_np as NI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN = os.path.join(ROOT, "tests", "asan")


@pytest.fixture(scope="module")
def tools(tmp_path_factory):
    out = tmp_path_factory.mktemp("asan_build")
    subprocess.run(["make", "-s", "-C", ASAN, f"OUT={out}"], check=True, timeout=300)
    return str(out / "io_check"), str(out / "cli_check")


def run(cmd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=99",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=98")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    return r.stdout


def test_nifti_reader_writer_clean_under_sanitizers(tools, tmp_path):
    io_check, _ = tools
    good, bad = [], []
    for dt, big in ((np.float32, False), (np.int16, True), (np.uint8, False), (np.float64, True), (np.int32, False)):
        for ext in (".nii", ".nii.gz"):
            p = tmp_path / f"img_{np.dtype(dt).name}_{int(big)}{ext}"
            NI.write(p, (np.arange(60) % 7).astype(dt).reshape(5, 4, 3), big_endian=big)
            good.append(p)
    p4 = tmp_path / "vol4d.nii.gz"
    NI.write(p4, np.random.default_rng(0).standard_normal((4, 3, 2, 5)).astype(np.float32))
    good.append(p4)
    NI.write(tmp_path / "ok.nii", np.arange(24, dtype=np.float32).reshape(4, 3, 2))
    raw = bytearray((tmp_path / "ok.nii").read_bytes())

    def variant(name, edit):
        b = bytearray(raw)
        edit(b)
        (tmp_path / name).write_bytes(bytes(b))
        return tmp_path / name

    bad.append(variant("dims.nii", lambda b: struct.pack_into("<8h", b, 40, 7, *([32767] * 7))))
    bad.append(variant("negdim.nii", lambda b: struct.pack_into("<8h", b, 40, 3, -4, 3, 2, 1, 1, 1, 1)))
    bad.append(variant("rank9.nii", lambda b: struct.pack_into("<h", b, 40, 9)))
    for name, off in (("nan.nii", float("nan")), ("huge.nii", 1e30), ("big.nii", 3e9), ("neg.nii", -5.0)):
        bad.append(variant(name, lambda b, off=off: struct.pack_into("<f", b, 108, off)))
    bad.append(variant("dtype.nii", lambda b: struct.pack_into("<h", b, 70, 1234)))
    bad.append(variant("bitpix.nii", lambda b: struct.pack_into("<h", b, 72, 0)))
    bad.append(variant("sizeof.nii", lambda b: struct.pack_into("<i", b, 0, 12345)))
    bad.append(variant("ext.nii", lambda b: b.__setitem__(slice(348, 352), bytes([1, 0, 0, 0]))))
    (tmp_path / "short.nii").write_bytes(bytes(raw[:-10]))
    (tmp_path / "header_only.nii").write_bytes(bytes(raw[:200]))
    (tmp_path / "zeros.nii").write_bytes(b"\0" * 400)
    (tmp_path / "garbage.nii").write_bytes(np.random.default_rng(1).integers(0, 256, 4096, dtype=np.uint8).tobytes())
    gz = gzip.compress(bytes(raw))
    (tmp_path / "cut.nii.gz").write_bytes(gz[: len(gz) // 2])
    (tmp_path / "notgz.nii.gz").write_bytes(b"this is not gzip" * 10)
    bad += [tmp_path / n for n in ("short.nii", "header_only.nii", "zeros.nii", "garbage.nii", "cut.nii.gz",
                                   "notgz.nii.gz", "missing.nii")]
    out = run([io_check, str(tmp_path / "scratch.nii")] + [str(p) for p in good + bad])
    lines = out.strip().splitlines()
    assert len(lines) == len(good) + len(bad)
    for p, line in zip(good, lines[:len(good)]):
        assert line.startswith("OK"), (p, line)
    for line in lines[len(good):]:
        assert line.startswith(("OK", "ERR")), line
    # the refused headers of the C ABI tests are refused here too
    refused = {p.name for p, line in zip(bad, lines[len(good):]) if line.startswith("ERR")}
    assert {"dims.nii", "nan.nii", "huge.nii", "big.nii", "short.nii", "zeros.nii", "missing.nii"} <= refused


def test_cli_parsers_clean_under_sanitizers(tools, tmp_path):
    _, cli_check = tools
    (tmp_path / "bvals").write_text("0 1000 1000 1000 1000 1000 1000 1000\n")
    (tmp_path / "bvecs").write_text("1 1 0 0 1 1 0 .5\n0 0 1 0 1 0 1 .5\n0 0 0 1 0 1 1 .5\n")
    (tmp_path / "bvals_col").write_text("\n".join(["0"] + ["1000"] * 7) + "\n")
    (tmp_path / "bvecs_rows").write_text("\n".join("1 0 0" for _ in range(8)) + "\n")
    (tmp_path / "bvals_short").write_text("0 1000 1000\n")
    (tmp_path / "empty").write_text("")
    (tmp_path / "junk").write_text("x y z\n\x00\x01\n" + "9" * 5000 + "\n")
    b, c = str(tmp_path / "bvals"), str(tmp_path / "bvecs")
    cases = [
        ["--nf=3", "--model=2", "--bi=1000", "--nj=1250", "--se=25", "-b", b, "-r", c],
        ["--nf", "3", "--fudge=1", "--cnonlinear", "--rician", "--f0", "--ardf0", f"--bvals={b}", f"--bvecs={c}"],
        ["-b", str(tmp_path / "bvals_col"), "-r", str(tmp_path / "bvecs_rows")],
        ["-b", str(tmp_path / "bvals_short"), "-r", c],
        ["-b", str(tmp_path / "empty"), "-r", c],
        ["-b", str(tmp_path / "junk"), "-r", str(tmp_path / "junk")],
        ["-b", str(tmp_path / "missing"), "-r", c],
        ["--nf"], ["--nf=x3"], ["--bi=99999999999999999999"], ["--seed=-1"], ["--fudge=abc"], ["--unknown"],
        ["positional"], [""], ["-"], ["--"], ["--="], ["--model=", "--nj"], ["--logdir"], ["--Rmean=1e400"],
    ]
    for args in cases:
        out = run([cli_check] + args)
        assert out.startswith(("OK", "ERR")), (args, out)

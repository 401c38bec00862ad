"""The euler3d command line (main.cpp:17-101 of the reference) and the C++ host
classes behind it: exit codes, and on the GPU, every file the reference wrote
when it stopped (Convergence.dat, AerodynamicCoefficients.dat, the volume and
surface Tecplot files of each block) byte for byte, and a run resumed from the
restart files against the uninterrupted reference run.

The reference prints its files through default-formatted std::ostreams
(PostProcessing/src/OutputTecplot.cpp:32-473): 6 significant digits, the same
text as C printf's %g, so the convergence rows are also compared as text."""
import hashlib
import math
import os
import subprocess

import pytest

import euler3d_amd as E
from helpers import Golden, write_case_mesh, write_input

CLI = os.path.join(os.path.dirname(os.path.abspath(E.__file__)), "euler3d")


def run_cli(*args, cwd=None, timeout=300):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} is not built")
    return subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


def test_cli_requires_input():
    r = run_cli()
    assert r.returncode == 255 and "ERROR: No input file" in r.stdout


def test_cli_missing_input_is_io_error(tmp_path):
    r = run_cli(str(tmp_path / "absent"))
    assert r.returncode == 42, r.stdout + r.stderr


def test_cli_missing_topology_is_io_error(tmp_path):
    g = Golden("bump1_roe1")
    inp = tmp_path / "input"
    write_input(inp, str(tmp_path / "absent.top"), g.meta)
    r = run_cli(str(inp))
    assert r.returncode == 42, r.stdout + r.stderr


def test_cli_unknown_memory_order_is_rejected(tmp_path):
    r = run_cli(str(tmp_path / "input"), "--order", "hilbert")
    assert r.returncode == 2 and "unknown memory order" in r.stderr


def test_cli_unknown_flux_scheme_is_rejected(tmp_path):
    g = Golden("bump1_roe1")
    inp = tmp_path / "input"
    write_input(inp, str(tmp_path / "case.top"), g.meta)
    text = inp.read_text().replace("\nroe 1\n", "\nhllc 1\n")
    inp.write_text(text)
    r = run_cli(str(inp))
    assert r.returncode == 1 and "unknown flux scheme" in r.stderr


def expected_rows(g):
    """Convergence.dat rows the reference printed for a golden case."""
    mon = g.monitor()
    rms0 = mon[0, 1:6]
    rows = []
    for r in mon:
        drops = [math.log10(r[1 + k]) - math.log10(rms0[k]) for k in range(5)]
        # the fixture stores the reference's own ro/u/p drops; v/w follow the same formula
        assert drops[0] == r[11] and drops[1] == r[12] and drops[4] == r[13]
        vals = [r[6], r[7], r[8], r[9], r[10]] + drops
        rows.append(" ".join([str(int(r[0]))] + ["%g" % v for v in vals]))
    return rows


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["bump8_roe2_gg_vk", "naca_hex_roe2_gg_vk", "bump1_ausm2_lsq_bj", "bump8_ausm1_irs"])
@pytest.mark.parametrize("mode", ["fused", "granular", "fused_morton", "granular_rcm"])
def test_cli_convergence_history(tmp_path, case, mode):
    g = Golden(case)
    top = write_case_mesh(g, str(tmp_path))
    inp = tmp_path / "input"
    write_input(inp, top, g.meta)
    out = tmp_path / "bin"
    args = [str(inp), "--quiet", "--out", str(out)] + (["--granular"] if mode.startswith("granular") else [])
    if "_" in mode:  # another device memory order: the same files
        args += ["--order", mode.split("_")[1]]
    print(f"cli {case} {mode}", flush=True)
    r = run_cli(*args, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    lines = (out / "Convergence.dat").read_text().splitlines()
    assert lines[0].startswith("Iteration Cl Cd Cmx Cmy Cmz Density_Convergence")
    want = expected_rows(g)
    assert lines[1:] == want
    aero = (out / "AerodynamicCoefficients.dat").read_text().splitlines()
    last = g.monitor()[-1]
    assert aero[0] == "Angle of attack Cl Cd Cmx Cmy Cmz"
    assert aero[1] == " ".join("%g" % v for v in [g.meta["aoa"]] + list(last[6:11]))
    assert_reference_outputs(g, out, g.meta["outputs"])


def assert_reference_outputs(g, out, names):
    """Each named output file equals the reference's, byte for byte."""
    assert names
    for fn in names:
        data = (out / fn).read_bytes()
        ref_len = int(g.z["outlen:" + fn][0])
        assert len(data) == ref_len and hashlib.sha256(data).digest() == bytes(g.z["out:" + fn]), \
            f"{fn}: {len(data)} bytes vs the reference's {ref_len}"


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "granular"])
def test_cli_restart_resumes_bitwise(tmp_path, mode):
    """Stop after 6 of 12 iterations, resume from the restart files: the
    solution files equal those of the reference's uninterrupted 12-iteration
    run byte for byte, and the convergence rows continue at iteration 6."""
    g = Golden("bump8_roe2_gg_vk")
    top = write_case_mesh(g, str(tmp_path))
    first, second = tmp_path / "first", tmp_path / "second"
    extra = ["--granular"] if mode == "granular" else []
    write_input(tmp_path / "input6", top, g.meta, max_iter=6)
    print(f"cli restart {mode}: first half", flush=True)
    r = run_cli(str(tmp_path / "input6"), "--quiet", "--out", str(first), *extra, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert all((first / f"RestartFile{b:04d}.dat").exists() for b in range(8))
    write_input(tmp_path / "input12", top, g.meta)
    print(f"cli restart {mode}: second half", flush=True)
    r = run_cli(str(tmp_path / "input12"), "--quiet", "--out", str(second), "--restart", str(first), *extra,
                cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    rows = (second / "Convergence.dat").read_text().splitlines()
    assert rows == expected_rows(g)[6:]
    assert_reference_outputs(g, second, [f for f in g.meta["outputs"] if f != "Convergence.dat"])


@pytest.mark.gpu
def test_cli_restart_rejects_foreign_file(tmp_path):
    """A restart file that does not belong to the mesh is an I/O error (exit 42)."""
    g = Golden("bump1_roe1")
    top = write_case_mesh(g, str(tmp_path))
    write_input(tmp_path / "input", top, g.meta)
    bad = tmp_path / "bad"
    bad.mkdir()
    (bad / "RestartFile0000.dat").write_bytes(b"not a restart file")
    r = run_cli(str(tmp_path / "input"), "--quiet", "--restart", str(bad), cwd=str(tmp_path))
    assert r.returncode == 42 and "not a restart file" in r.stderr, r.stdout + r.stderr


@pytest.mark.gpu
def test_cli_stop_file(tmp_path):
    """A STOP file in the working directory ends the run after one iteration
    (PostProcessing/src/PostProcessing.cpp:72-80) and is removed."""
    g = Golden("bump1_roe1")
    top = write_case_mesh(g, str(tmp_path))
    inp = tmp_path / "input"
    write_input(inp, top, g.meta, max_iter=1000)
    (tmp_path / "STOP").write_text("")
    r = run_cli(str(inp), "--quiet", "--out", str(tmp_path / "bin"), cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    rows = (tmp_path / "bin" / "Convergence.dat").read_text().splitlines()[1:]
    assert len(rows) == 1 and rows[0] == expected_rows(g)[0]
    assert not (tmp_path / "STOP").exists()

"""The C++ host driver sweep_hip.x: the reference's executable shape with
SweepType HIP, SweeperAbstract::sweep driven by a host source iteration."""
import os
import subprocess

import numpy as np
import pytest

from tycho2_amd import build, psi_io

from conftest import golden

GOLD_DECK = """# input-gold.deck of the reference with SweepType HIP
snOrder         8
iterMax         100
errMax          1e-10
maxCellsPerStep 100
intraAngleP     3
interAngleP     1
nGroups         2
sigmaT1         10
sigmaS1         5
sigmaT2         10
sigmaS2         5
OutputFile      true
OutputFilename  {out}
SourceIteration true
MPIType         TychoTwoSided
SweepType       HIP
GaussElim       NoPivot
{extra}
"""


def run(deck_text, tmp_path, mesh=None):
    deck = tmp_path / "input.deck"
    deck.write_text(deck_text)
    return subprocess.run([build.DRIVER, mesh or golden("cube-208.smesh"),
                           str(deck)], capture_output=True, text=True,
                          timeout=300)


def test_driver_usage_and_errors(tmp_path):
    r = subprocess.run([build.DRIVER], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage" in r.stdout
    out = str(tmp_path / "o.psi")
    r = run(GOLD_DECK.format(out=out, extra="").replace(
        "SweepType       HIP", "SweepType       PBJ"), tmp_path)
    assert r.returncode == 1 and "Sweep type not recognized" in r.stderr
    r = run(GOLD_DECK.format(out=out, extra="").replace("snOrder         8",
                                                        ""), tmp_path)
    assert r.returncode == 1 and "snOrder" in r.stderr
    r = run(GOLD_DECK.format(out=out, extra="SweepEngine Fast"), tmp_path)
    assert r.returncode == 1 and "SweepEngine" in r.stderr
    r = run(GOLD_DECK.format(out=out, extra="SweepSchedule Heap"), tmp_path)
    assert r.returncode == 1 and "SweepSchedule" in r.stderr
    r = run(GOLD_DECK.format(out=out, extra="").replace(
        "GaussElim       NoPivot", "GaussElim       Fastest"), tmp_path)
    assert r.returncode == 1 and "GaussElim" in r.stderr
    r = run(GOLD_DECK.format(out=out, extra="Precision FP32"), tmp_path)
    assert r.returncode == 1 and "DeviceLoop true" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("extra,gauss", [
    ("DeviceLoop false", "NoPivot"), ("DeviceLoop true", "NoPivot"),
    ("DeviceLoop true\nSweepEngine Queue", "NoPivot"),
    ("DeviceLoop true", "Original"), ("DeviceLoop true", "CramerGlu"),
    ("DeviceLoop false", "CramerIntel"),
    ("DeviceLoop false\nSweepSchedule Priorities", "NoPivot"),
    ("DeviceLoop true\nSweepEngine Queue\nSweepSchedule Priorities",
     "NoPivot"),
    ("DeviceLoop false\nSWEEPTYPE OriginalTycho1", "NoPivot"),
    ("DeviceLoop true\nSWEEPTYPE OriginalTycho2", "NoPivot"),
    ("DeviceLoop true\nSWEEPTYPE TraverseGraph", "NoPivot")])
def test_driver_reproduces_gold(tmp_path, gold, extra, gauss):
    print(f"driver gold {extra} GaussElim {gauss}", flush=True)
    out = str(tmp_path / "out.psi")
    sweep_type = "HIP"
    if "SWEEPTYPE" in extra:   # the reference's other single-rank sweepers
        extra, sweep_type = extra.split("\nSWEEPTYPE ")
    deck = GOLD_DECK.format(out=out, extra=extra).replace(
        "GaussElim       NoPivot", f"GaussElim       {gauss}").replace(
        "SweepType       HIP", f"SweepType       {sweep_type}")
    r = run(deck, tmp_path)
    assert r.returncode == 0, r.stderr
    psi = psi_io.read_psi(out)
    assert psi.shape == (208, 128, 4, 2)
    got = psi[..., 0][:, gold["psi_angles"], :]
    l1, linf = psi_io.rel_diff(got, gold["psi_g0"])
    assert l1 < 1e-10 and linf < 1e-10
    assert "iterations: 34" in r.stdout
    # the reference executable's end-of-run lines (src/Main.cc:340-345)
    lines = dict(line.split(": ", 1) for line in r.stdout.splitlines()
                 if line.startswith(("L2 Relative Error", "Diff between")))
    from tycho2_amd import mesh as meshmod, problems
    sm = meshmod.read_smesh(golden("cube-208.smesh"))
    assert float(lines["L2 Relative Error"]) == pytest.approx(
        problems.hat_l2_error(sm, psi), rel=1e-6)
    assert float(lines["Diff between groups"]) == \
        psi_io.diff_between_groups(psi) == 0.0


@pytest.mark.gpu
def test_driver_reads_partitioned_mesh(tmp_path, gold):
    """The gold deck on cube-208 written as a 3-partition .pmesh (the
    reference executable's mesh input) reproduces gold.psi."""
    from tycho2_amd import mesh as meshmod
    print("driver gold from .pmesh", flush=True)
    sm = meshmod.read_smesh(golden("cube-208.smesh"))
    pm = str(tmp_path / "cube-208.pmesh")
    meshmod.write_pmesh(sm, pm, np.arange(sm.n_cells) % 3, 3)
    out = str(tmp_path / "out.psi")
    r = run(GOLD_DECK.format(out=out, extra="DeviceLoop true"), tmp_path,
            mesh=pm)
    assert r.returncode == 0, r.stderr
    psi = psi_io.read_psi(out)
    got = psi[..., 0][:, gold["psi_angles"], :]
    l1, linf = psi_io.rel_diff(got, gold["psi_g0"])
    assert l1 < 1e-10 and linf < 1e-10


def test_driver_rejects_bad_mesh(tmp_path):
    bad = tmp_path / "bad.pmesh"
    bad.write_bytes(b"Tycho 2 Parallel Mesh".ljust(32, b"\0") + b"\1" * 8)
    out = str(tmp_path / "o.psi")
    r = run(GOLD_DECK.format(out=out, extra=""), tmp_path, mesh=str(bad))
    assert r.returncode == 1 and "parallel mesh" in r.stderr



@pytest.mark.gpu
@pytest.mark.parametrize("extra", ["", "SweepEngine Queue"])
def test_driver_fp32_refined_reproduces_gold_psi(tmp_path, gold, extra):
    """Precision FP32 with DeviceLoop true runs the refined solve (FP32
    sweeps and FP32 psi inside FP64 defect correction); the psi it writes
    is psi32 plus its correction, and meets the reference's regression bar
    on psi itself: 1e-10 in diff.py's L1 and Linf metrics against gold.psi
    (test/diff.py:42-53, test/regression.py:4).  An uncorrected FP32 psi
    is ~1e-7 away."""
    print(f"driver gold FP32 refined {extra!r}", flush=True)
    out = str(tmp_path / "out.psi")
    deck = GOLD_DECK.format(out=out, extra="DeviceLoop true\nPrecision FP32\n"
                            + extra)
    r = run(deck, tmp_path)
    assert r.returncode == 0, r.stderr
    assert "refined solve:" in r.stdout
    psi = psi_io.read_psi(out)
    got = psi[..., 0][:, gold["psi_angles"], :]
    l1, linf = psi_io.rel_diff(got, gold["psi_g0"])
    print(f"  psi vs gold: L1 {l1:.2e}, Linf {linf:.2e}")
    assert l1 < 1e-10 and linf < 1e-10
    # whole-file norms of gold.psi (every angle and group)
    assert abs(np.abs(psi).sum() - gold["psi_l1"]) <= 1e-10 * gold["psi_l1"]
    assert abs(np.abs(psi).max() - gold["psi_linf"]) <= \
        1e-10 * gold["psi_linf"]


@pytest.mark.gpu
def test_driver_reports_iteration_limit(tmp_path):
    """iterMax below what the solve needs: the driver keeps the last
    iterate, as the reference does, and warns with the engine's
    T2_ERR_NOCONV message."""
    print("driver iteration limit", flush=True)
    out = str(tmp_path / "out.psi")
    deck = GOLD_DECK.format(out=out, extra="DeviceLoop true").replace(
        "iterMax         100", "iterMax         3")
    r = run(deck, tmp_path)
    assert r.returncode == 0, r.stderr
    assert "warning" in r.stderr and "source iteration" in r.stderr
    assert "iterations: 3" in r.stdout

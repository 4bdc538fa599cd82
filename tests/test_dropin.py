"""The source-level drop-in (CPU): the reference's MC/src/main.cpp and
MC/src/render.cpp, compiled unchanged against the reference's own headers,
link against libmmc.so in place of MC/src/sim.cpp (oracle/_ref/mc_dropin,
built by oracle/Makefile's dropin target), because the facade's meta::Simulator
and every struct of its API keep the reference's layout
(tests/abi/layout_check.cpp, static assertions on sizes, alignments and
offsets, plus the default values).  Needs the reference tree (this container);
the GPU run of the drop-in is tests/test_gpu_cli.py."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_MC = Path("/root/reference/projects/metamaterial-compression/industry-grade")
DROPIN = ROOT / "oracle" / "_ref" / "mc_dropin"
LIB = ROOT / "falling-dice_amd" / "libmmc.so"

needs_ref = pytest.mark.skipif(not (REF_MC / "include" / "sim.hpp").exists(), reason="reference tree absent")


@needs_ref
def test_facade_layout_equals_reference(tmp_path):
    exe = tmp_path / "layout_check"
    subprocess.run(["g++", "-std=c++20", "-Wno-invalid-offsetof", f"-I{REF_MC / 'include'}",
                    f"-I{ROOT / 'falling-dice_amd' / 'csrc'}", "-o", str(exe),
                    str(ROOT / "tests" / "abi" / "layout_check.cpp")], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True)
    print(p.stdout)
    assert p.returncode == 0 and "sizeof(Simulator) 600" in p.stdout


def _symbols(path, flag):
    """Demangled dynamic symbol names (nm lines: [address] type name)."""
    out = subprocess.run(["nm", "-D", "-C", flag, str(path)], capture_output=True, text=True, check=True).stdout
    names = set()
    for line in out.splitlines():
        m = re.match(r"^\s*(?:[0-9a-fA-F]+\s+)?[A-Za-z]\s+(.*)$", line)
        if m:
            names.add(m.group(1).strip())
    return names


@needs_ref
def test_reference_program_links_against_the_library():
    """Every meta:: symbol the reference's main/render objects need from
    sim.cpp is defined by libmmc.so."""
    if not DROPIN.exists():
        pytest.fail(f"{DROPIN} missing (make -C oracle dropin)")
    need = {s for s in _symbols(DROPIN, "--undefined-only") if s.startswith("meta::")}
    have = _symbols(LIB, "--defined-only")
    assert need, "the drop-in needs the solver's meta:: symbols"
    missing = sorted(need - have)
    assert not missing, missing
    for fn in ("meta::Simulator::Simulator(meta::SimulationConfig)", "meta::Simulator::run()",
               "meta::Simulator::rigidRotationTest(double*) const", "meta::Simulator::gradientCheck(double*) const",
               "meta::Simulator::exportSTL(std::filesystem::__cxx11::path const&, int) const"):
        assert fn in need, fn


@needs_ref
def test_reference_program_help_runs():
    p = subprocess.run([str(DROPIN), "--help"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "--mode sim|render|hero|stl|validate|convergence|all" in p.stdout

"""The C++ drop-in host (include/cfpq/*.h + cfpq-on-gpgpu_amd/host/*.cpp) on
the CPU: its grammar/graph parsers, worklist driver and print_results must
reproduce the reference driver's output file (oracle/_ref/cfpq_ref runs the
reference's own cfpq-cpp sources) or, without it, the oracle's fingerprints."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_results(path):
    out = {}
    for line in open(path):
        name, h, c = line.split()
        out[name] = (int(h), int(c))
    return out


@pytest.mark.parametrize("which", ["dyck", "sg", "random_cnf", "epsilon"])
def test_host_matches_reference_output(host_cpu_bin, oracle, gen, which):
    q = {"dyck": lambda: gen.dyck_er(300, 900, seed=41),
         "sg": lambda: gen.sg_forest(1500, seed=42, types_per_vertex=3),
         "random_cnf": lambda: gen.dense_stress(200, seed=43, n_nt=9, n_rules=16, mean_deg=2.0),
         "epsilon": lambda: gen.dyck_er(150, 400, seed=44)}[which]()
    if which == "epsilon":
        q.grammar = "S\nS A S1\nS1 S B\nS S S\nA a\nB b\n"
    with tempfile.TemporaryDirectory() as d:
        gp, ep, o1, o2 = (os.path.join(d, x) for x in ("g.txt", "e.txt", "host.txt", "ref.txt"))
        open(gp, "w").write(q.grammar)
        open(ep, "w").write(q.graph_text())
        subprocess.run([host_cpu_bin, gp, ep, o1], check=True, capture_output=True)
        got = read_results(o1)
        if os.path.exists(oracle.REF_BIN):
            ref, _ = oracle.run_reference_driver(gp, ep, o2)
            assert got == ref
            # same line order as the reference writer (std::map name order)
            assert [l.split()[0] for l in open(o1)] == [l.split()[0] for l in open(o2)]
        q.n = int(max(q.src.max(), q.dst.max())) + 1
        g, mats, _ = oracle.solve(q, "seminaive")
        assert got == oracle.fingerprints(g, mats, q.n)


def test_cli_fails_loudly_without_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    exe = os.path.join(ROOT, "cfpq-on-gpgpu_amd", "cfpq_hip")
    if not os.path.exists(exe):
        pytest.skip("cfpq_hip not built")
    (tmp_path / "g.txt").write_text("S a\n")
    (tmp_path / "e.txt").write_text("0 a 1\n")
    p = subprocess.run([exe, str(tmp_path / "g.txt"), str(tmp_path / "e.txt"), str(tmp_path / "o.txt")],
                       capture_output=True, text=True)
    assert p.returncode != 0 and "cfpq_create" in p.stderr


REFSRC = "/root/reference/cfpq-cpp"


def test_dropin_builds_against_reference_headers(tmp_path):
    """INTEGRATION's reference-side backend main (host/reference_main.cpp)
    compiles with the reference's unmodified cfpq-cpp headers first on the
    include path, together with the reference's Grammar.cpp and Graph.cpp,
    and links against libcfpq_hip.so.  Without a GPU the binary must fail
    loudly at context creation."""
    if not os.path.exists(os.path.join(REFSRC, "Grammar.cpp")):
        pytest.skip("reference sources absent")
    pkg = os.path.join(ROOT, "cfpq-on-gpgpu_amd")
    if not os.path.exists(os.path.join(pkg, "libcfpq_hip.so")):
        pytest.skip("libcfpq_hip.so not built")
    exe = str(tmp_path / "hip_main")
    p = subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", REFSRC, "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "include", "cfpq"), "-o", exe,
                        os.path.join(pkg, "host", "reference_main.cpp"), os.path.join(REFSRC, "Grammar.cpp"),
                        os.path.join(REFSRC, "Graph.cpp"), os.path.join(pkg, "host", "HipMatrix.cpp"),
                        "-L", pkg, "-lcfpq_hip", "-Wl,-rpath," + pkg], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    import torch
    if torch.cuda.is_available():
        return
    (tmp_path / "g.txt").write_text("S a\n")
    (tmp_path / "e.txt").write_text("0 a 1\n")
    r = subprocess.run([exe, str(tmp_path / "g.txt"), str(tmp_path / "e.txt"), str(tmp_path / "o.txt")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "cfpq_create" in r.stderr


def test_repo_and_reference_interfaces_do_not_mix(tmp_path):
    """Both Matrix.h headers in one translation unit is an error (distinct
    include guards, two definitions of Matrix), not a silent loss of the
    repository's hooks."""
    if not os.path.exists(os.path.join(REFSRC, "Matrix.h")):
        pytest.skip("reference sources absent")
    src = tmp_path / "mix.cpp"
    src.write_text(f'#include "{REFSRC}/Matrix.h"\n#include "{ROOT}/include/cfpq/Matrix.h"\nint main() {{ return 0; }}\n')
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", str(src)], capture_output=True, text=True)
    assert p.returncode != 0 and "redefinition" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["dyck", "sg", "epsilon"])
def test_dropin_reference_driver_on_gpu(oracle, gen, which):
    """The reference's unmodified driver (cfpq-cpp Grammar/Graph parsers,
    worklist and print_results) with HipMatrix plugged in, built by
    oracle/Makefile as _ref/cfpq_ref_hip: its output file equals the oracle's
    fingerprints."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cfpq_ref_hip")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/cfpq_ref_hip not built (reference sources absent at build time)")
    q = {"dyck": lambda: gen.dyck_er(300, 900, seed=51),
         "sg": lambda: gen.sg_forest(1500, seed=52, types_per_vertex=3),
         "epsilon": lambda: gen.dyck_er(150, 400, seed=54)}[which]()
    if which == "epsilon":
        q.grammar = "S\nS A S1\nS1 S B\nS S S\nA a\nB b\n"
    with tempfile.TemporaryDirectory() as d:
        gp, ep, o1 = (os.path.join(d, x) for x in ("g.txt", "e.txt", "out.txt"))
        open(gp, "w").write(q.grammar)
        open(ep, "w").write(q.graph_text())
        p = subprocess.run([exe, gp, ep, o1], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        total_ms, algo_ms = (int(x) for x in p.stdout.split())
        assert total_ms >= algo_ms >= 0
        got = read_results(o1)
    q.n = int(max(q.src.max(), q.dst.max())) + 1
    g, mats, _ = oracle.solve(q, "seminaive")
    assert got == oracle.fingerprints(g, mats, q.n)

"""Test-set harness with the reference's protocol: the counterpart of
test_utils/build_testset.py and testing_system.py, driving this engine's
command-line runner (cfpq_hip, the drop-in for the reference backends' mains)
and any other runner that speaks the same protocol.

    python -m cfpq-on-gpgpu_amd.harness build-testset DATA [-o tests.csv]
    python -m cfpq-on-gpgpu_amd.harness run tests.csv [--runner EXE[=NAME[=ARG,...]]]...
           [--check TEST]... [--repeat 30] [--strategy circle|simple]
           [--timeout 1800] [--threshold 600] [--axiom s] [-o result.csv]

build-testset  walks DATA for folders holding both Grammars/ and Matrices/
               and writes one test per (grammar, matrix) pair, named
               "<grammar stem>:<matrix stem>", with the rule count and the
               vertex / edge counts (build_testset.py:7-42).
run            cross-validates the runners on the --check tests (every runner's
               "name hash count" file must be identical, testing_system.py:78-106),
               then times every (runner, test) --repeat times in the chosen
               order (run_strategy.py:25-42) and writes result.csv with the
               reference's columns (testing_system.py:17-39): per runner the
               mean and std of the total and algorithm times in seconds and
               the axiom nonterminal's hash and count.

A runner is an executable called as ``EXE grammar graph out [ARG...]`` that
prints "<total_ms> <algorithm_ms>" and writes "name hash count" lines; a
non-empty stderr or a non-zero exit is a failure (test_utils/runners.py:16-31).
Without --runner the harness runs cfpq_hip three times: as the device
fixpoint ("c++_cfpq_hip"), as the reference's worklist over cfpq_add_mul
("c++_cfpq_hip_worklist") and as the sparse pair-list fixpoint
("c++_cfpq_hip_pairs").
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "cfpq_hip")


# ---------------------------------------------------------------------------
# test sets
# ---------------------------------------------------------------------------

def _stem(name: str) -> str:
    return name[:-4] if len(name) > 4 else name


def build_testset(data_folder: str, out: str = "tests.csv") -> list[list]:
    """One test per (grammar, matrix) of every folder that holds both
    Grammars/ and Matrices/ (test_utils/build_testset.py:7-42)."""
    tests, cache = [], {}
    for root, dirnames, _ in sorted(os.walk(data_folder)):
        if not ("Grammars" in dirnames and "Matrices" in dirnames):
            continue
        for grammar in sorted(os.listdir(os.path.join(root, "Grammars"))):
            for matrix in sorted(os.listdir(os.path.join(root, "Matrices"))):
                pg = os.path.join(root, "Grammars", grammar)
                pm = os.path.join(root, "Matrices", matrix)
                if pm not in cache:
                    n_vert = n_edge = 0
                    with open(pm) as f:
                        for line in f:
                            if not line.strip():
                                continue
                            u, _, v = line.split()
                            n_edge += 1
                            n_vert = max(n_vert, int(u), int(v))
                    cache[pm] = (n_vert + 1 if n_edge else 0, n_edge)
                if pg not in cache:
                    with open(pg) as f:
                        cache[pg] = sum(1 for _ in f)
                tests.append([f"{_stem(grammar)}:{_stem(matrix)}", pg, cache[pg], pm, *cache[pm]])
    with open(out, "w") as f:
        f.write("Name,Grammar,Number of rules,Matrix, Number of vertices, Number of edges\n")
        for t in tests:
            f.write(",".join(map(str, t)) + "\n")
    return tests


def read_tests(path: str) -> dict[str, tuple[str, str]]:
    """name -> (grammar path, matrix path), columns 0, 1 and 3 (testing_system.py:163-169)."""
    tests = {}
    with open(path) as f:
        f.readline()
        for line in f:
            if not line.strip():
                continue
            v = line.rstrip("\n").split(",")
            tests[v[0]] = (v[1], v[3])
    return tests


# ---------------------------------------------------------------------------
# runners
# ---------------------------------------------------------------------------

class RunnerError(RuntimeError):
    pass


@dataclass
class Runner:
    path: str
    name: str
    args: list = field(default_factory=list)
    timeout: float = 1800.0

    def run(self, grammar: str, graph: str, out: str) -> tuple[float, float]:
        """(total seconds, algorithm seconds)."""
        try:
            p = subprocess.run([self.path, grammar, graph, out, *self.args], capture_output=True,
                               timeout=self.timeout)
        except subprocess.TimeoutExpired:
            raise RunnerError("timeout")
        if p.returncode != 0 or p.stderr:
            raise RunnerError(f"exit {p.returncode}: {p.stderr.decode(errors='replace').strip()}")
        fields = p.stdout.decode().split()
        if len(fields) < 2:
            raise RunnerError(f"unexpected output {p.stdout!r}")
        return int(fields[0]) / 1000.0, int(fields[1]) / 1000.0


def parse_runner(spec: str, timeout: float) -> Runner:
    """EXE[=NAME[=ARG,ARG...]]"""
    parts = spec.split("=", 2)
    path = parts[0]
    name = parts[1] if len(parts) > 1 and parts[1] else "c++_" + os.path.basename(path)
    args = [a for a in parts[2].split(",") if a] if len(parts) > 2 else []
    return Runner(path, name, args, timeout)


def default_runners(timeout: float) -> list[Runner]:
    if not os.path.exists(CLI):
        raise FileNotFoundError(f"{CLI} is missing: build it with `make -C {HERE}`")
    return [Runner(CLI, "c++_cfpq_hip", [], timeout), Runner(CLI, "c++_cfpq_hip_worklist", ["--worklist"], timeout),
            Runner(CLI, "c++_cfpq_hip_pairs", ["--pairs"], timeout)]


def read_answer(path: str) -> dict[str, tuple[str, str]]:
    out = {}
    with open(path) as f:
        for line in f:
            if line.strip():
                name, h, c = line.split()
                out[name] = (h, c)
    return out


def axiom_of(answer: dict[str, tuple[str, str]], axiom: str) -> tuple[str, str]:
    """The axiom's (hash, count); like the reference's reader, the last line
    when the axiom is absent (testing_system.py:141-146)."""
    if axiom in answer:
        return answer[axiom]
    if not answer:
        return ("-", "-")
    return answer[list(answer)[-1]]


# ---------------------------------------------------------------------------
# run order, statistics
# ---------------------------------------------------------------------------

SAVE = "save"


def strategy(runners, tests, method: str, repeat: int):
    """(runner, test) order (run_strategy.py:25-42); SAVE marks where the
    circle order rewrites result.csv."""
    if method == "simple":
        return [(r, t) for r in runners for t in tests for _ in range(repeat)]
    if method == "circle":
        return [(r, t) for _ in range(repeat) for t in tests for r in list(runners) + [SAVE]]
    raise ValueError(f"unknown strategy {method!r}")


def write_results(path: str, runners, results) -> None:
    """result.csv with the reference's columns (testing_system.py:17-39)."""
    with open(path, "w") as f:
        header = "Test name"
        for r in runners:
            header += (f",{r.name}_time_all_mean,{r.name}_time_all_std"
                       f",{r.name}_time_mult_mean,{r.name}_time_mult_std"
                       f",{r.name}_axiom_hash,{r.name}_axiom_count")
        f.write(header + "\n")
        for test, per in results.items():
            line = test
            for r in runners:
                m = per[r.name]
                if m:
                    t_all = np.array([x[0] for x in m])
                    t_alg = np.array([x[1] for x in m])
                    line += f",{t_all.mean()},{t_all.std()},{t_alg.mean()},{t_alg.std()},{m[0][2]},{m[0][3]}"
                else:
                    line += ",-,-,-,-,-,-"
            f.write(line + "\n")


class Mismatch(RuntimeError):
    pass


def cross_validate(runners, tests, check, folder: str = "cross_val", log=print) -> dict:
    """Every runner's answer on each check test must equal the first runner's
    (testing_system.py:78-106).  Raises Mismatch; runners that fail are
    logged and skipped, as in the reference."""
    os.makedirs(folder, exist_ok=True)
    answers = {}
    for test in check:
        ref = None
        ref_name = None
        for r in runners:
            out = os.path.join(folder, f"{test}_{r.name}.txt")
            try:
                r.run(*tests[test], out)
            except RunnerError as e:
                log(f"{r.name} on {test}: cannot validate ({e})")
                continue
            ans = read_answer(out)
            answers[(test, r.name)] = ans
            if ref is None:
                ref, ref_name = ans, r.name
            elif ans != ref:
                raise Mismatch(f"{r.name} and {ref_name} have different answers on test {test}")
    return answers


def run_tests(runners, tests, *, check=(), repeat: int = 30, method: str = "circle", threshold: float = 600.0,
              stop_repeat: int = 1, axiom: str = "s", out: str = "result.csv", answer: str = "answer.txt",
              cross_val: str = "cross_val", log=print) -> dict:
    if check:
        cross_validate(runners, tests, check, cross_val, log)
    results = {t: {r.name: [] for r in runners} for t in tests}
    write_results(out, runners, results)
    failed = set()
    for r, t in strategy(runners, list(tests), method, repeat):
        if r == SAVE:
            write_results(out, runners, results)
            continue
        if (r.name, t) in failed:
            continue
        m = results[t][r.name]
        # a test that already took longer than the threshold is not repeated
        if len(m) >= stop_repeat and np.mean([x[0] for x in m]) > threshold:
            continue
        try:
            t_all, t_alg = r.run(*tests[t], answer)
            h, c = axiom_of(read_answer(answer), axiom)
            m.append((t_all, t_alg, h, c))
            log(f"{r.name} on {t}: {t_all:.3f} s (algorithm {t_alg:.3f} s)")
        except RunnerError as e:
            log(f"{r.name} on {t}: failed ({e})")
            failed.add((r.name, t))
    write_results(out, runners, results)
    return results


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="harness", description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build-testset")
    b.add_argument("data")
    b.add_argument("-o", "--out", default="tests.csv")
    r = sub.add_parser("run")
    r.add_argument("tests")
    r.add_argument("--runner", action="append", default=[], help="EXE[=NAME[=ARG,...]]")
    r.add_argument("--check", action="append", default=[], help="cross-validate the runners on this test")
    r.add_argument("--repeat", type=int, default=30)
    r.add_argument("--strategy", default="circle", choices=["circle", "simple"])
    r.add_argument("--timeout", type=float, default=1800.0)
    r.add_argument("--threshold", type=float, default=600.0)
    r.add_argument("--axiom", default="s")
    r.add_argument("-o", "--out", default="result.csv")
    a = ap.parse_args(argv)
    say = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "build-testset":
        tests = build_testset(a.data, a.out)
        say(f"{len(tests)} tests written to {a.out}")
        return 0
    tests = read_tests(a.tests)
    runners = [parse_runner(s, a.timeout) for s in a.runner] or default_runners(a.timeout)
    unknown = [c for c in a.check if c not in tests]
    if unknown:
        say(f"unknown check tests: {', '.join(unknown)}")
        return 2
    try:
        run_tests(runners, tests, check=a.check, repeat=a.repeat, method=a.strategy, threshold=a.threshold,
                  axiom=a.axiom, out=a.out, log=say)
    except Mismatch as e:
        say(str(e))
        return 1
    say(f"results written to {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

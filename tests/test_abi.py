"""The C ABI library loads and exports everything include/cfpq_hip.h declares
(CPU only: no compute call is made)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cfpq_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfpq_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert sorted(pkg.EXPORTS) == names


def test_struct_layouts_match_header(pkg):
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "cfpq_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(cfpq_opts), sizeof(cfpq_rule), sizeof(cfpq_stats),
         sizeof(cfpq_query), offsetof(cfpq_stats, fixpoint_ms), offsetof(cfpq_opts, schedule),
         sizeof(cfpq_pairs), offsetof(cfpq_pairs, cap));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [ctypes.sizeof(pkg.Opts), ctypes.sizeof(pkg.Rule), ctypes.sizeof(pkg.Stats), ctypes.sizeof(pkg.Query),
            pkg.Stats.fixpoint_ms.offset, pkg.Opts.schedule.offset, ctypes.sizeof(pkg.Pairs), pkg.Pairs.cap.offset]
    assert got == want


def test_create_without_gpu_fails_cleanly(pkg):
    import torch  # noqa: F401  (only to learn whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = pkg.lib().cfpq_create(0, None, ctypes.byref(h))
    assert rc == pkg.E_DEVICE and not h.value

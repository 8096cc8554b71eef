"""The C-ABI library loads and exports every entry point include/mhpc.h declares (CPU:
no compute calls), and the Python binding's struct layouts match the header."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (REPO / "include" / "mhpc.h").read_text()
    decl = re.compile(r"^(?:const\s+)?(?:int|void|int64_t|char)\s*\**\s*(mhpc_[a-z0-9_]+)\s*\(", re.M)
    return sorted(set(decl.findall(text)))


def test_library_exports_every_declared_symbol():
    from miluphpc_amd import lib

    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib.LIB, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (mhpc_[a-z0-9_]+)", out))
    assert set(syms) <= exported


def test_default_params_and_status_strings():
    from miluphpc_amd import lib

    p = lib.default_params()
    assert p.theta == 0.5 and p.max_nb == 180 and p.n_target == 64
    assert (p.band_lo, p.band_hi, p.max_counts) == (50, 80, 7)
    assert lib.LIB.mhpc_status_string(0) == b"ok"
    assert b"tree depth" in lib.LIB.mhpc_status_string(lib.E_DEPTH)


def test_struct_layout_matches_header():
    """Compile a tiny C program against include/mhpc.h and compare sizes / offsets."""
    from miluphpc_amd import lib

    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "mhpc.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(mhpc_params), offsetof(mhpc_params, cs_before_p),
         offsetof(mhpc_params, sml_max), offsetof(mhpc_params, smoothing_kernel),
         sizeof(mhpc_rhs_stats), offsetof(mhpc_rhs_stats, num_cells), offsetof(mhpc_rhs_stats, h_passes));
  return 0;
}
'''
    tmp = Path("/tmp/mhpc_layout_check")
    tmp.mkdir(exist_ok=True)
    (tmp / "t.c").write_text(src)
    subprocess.run(["gcc", "-I", str(REPO / "include"), str(tmp / "t.c"), "-o", str(tmp / "t")], check=True)
    vals = list(map(int, subprocess.run([str(tmp / "t")], capture_output=True, text=True).stdout.split()))
    assert vals == [C.sizeof(lib.Params), lib.Params.cs_before_p.offset, lib.Params.sml_max.offset,
                    lib.Params.smoothing_kernel.offset, C.sizeof(lib.RhsStats),
                    lib.RhsStats.num_cells.offset, lib.RhsStats.h_passes.offset]


def test_no_gpu_call_without_device_fails_cleanly():
    """Without a GPU, creating a context reports an error status instead of crashing."""
    import torch

    if torch.cuda.is_available():
        return
    from miluphpc_amd import lib

    h = C.c_void_p()
    st = lib.LIB.mhpc_create(C.byref(lib.default_params()), 16, 0, C.byref(h))
    assert st != lib.OK


def test_ics_shapes_and_invariants():
    from miluphpc_amd import ics

    s = ics.plummer(5000, seed=1)
    r = np.sqrt(s["x"] ** 2 + s["y"] ** 2 + s["z"] ** 2)
    assert r.max() < 10.5
    assert abs(s["mass"].sum() - 1.0) < 1e-12
    assert np.abs((s["mass"][:, None] * np.stack([s["vx"], s["vy"], s["vz"]], 1)).sum(0)).max() < 1e-14
    # virial ratio of the Plummer model: 2T / |W| ~ 1 (half-mass radius 1.305 a)
    assert 1.0 < np.median(r) < 1.6
    L = ics.sedov_lattice(8)
    assert len(L["x"]) == 512 and abs((L["mass"] * L["e"]).sum() - 1.0 - 1e-6) < 1e-9
    b = ics.boss_bodenheimer(2000)
    assert abs(b["mass"].sum() - 1.0) < 1e-12
    assert np.all(b["x"] ** 2 + b["y"] ** 2 + b["z"] ** 2 <= 1.0)

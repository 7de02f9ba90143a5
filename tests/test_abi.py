"""The C ABI library: it loads and exports every entry point include/bmg.h
declares; option defaults equal the reference dictionary defaults; host-only
entry points behave without a GPU.  No compute calls here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bmg.h")
UNET_HEADER = os.path.join(ROOT, "include", "bmg_unet.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|long|const char\*)\s+(bmg_\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("bmg_create", "bmg_destroy", "bmg_solve", "bmg_solve_device",
                 "bmg_set_operator_stencil", "bmg_set_operator_laplacian",
                 "bmg_get_operator_stencil", "bmg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from briscola_amd import bmg
    lib = bmg.load()
    out = subprocess.run(["nm", "-D", "--defined-only", bmg.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (bmg_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(lib, n)
    # the Python binding names exactly the declared set
    assert sorted(bmg.EXPORTS) == declared_functions()


def test_library_exports_learned_preconditioner_symbols():
    from briscola_amd import bmg, unet
    out = subprocess.run(["nm", "-D", "--defined-only", bmg.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (bmg_\w+)", out))
    names = declared_functions(UNET_HEADER)
    assert "bmg_unet_train_step" in names and "bmg_set_learned" in names
    assert not [n for n in names if n not in exported]
    assert sorted(unet.EXPORTS) == names
    unet._lib()   # every signature binds


def test_library_is_gfx950_code_object(tmp_path):
    """The device code bundled in libbmg.so targets gfx950 only."""
    from briscola_amd import bmg
    llvm = "/opt/rocm/llvm/bin"
    if not os.path.exists(os.path.join(llvm, "clang-offload-bundler")):
        pytest.skip("ROCm LLVM tools not installed")
    fb = tmp_path / "fatbin.bin"
    subprocess.run([os.path.join(llvm, "llvm-objcopy"), f"--dump-section=.hip_fatbin={fb}",
                    bmg.LIB_PATH, str(tmp_path / "stripped.so")], check=True,
                   capture_output=True)
    out = subprocess.run([os.path.join(llvm, "clang-offload-bundler"), "--list",
                          "--type=o", f"--input={fb}"], capture_output=True,
                         text=True, check=True).stdout
    targets = [t for t in out.split() if t.startswith("hip")]
    assert targets == ["hipv4-amdgcn-amd-amdhsa--gfx950"]


def test_default_opts_are_reference_defaults():
    """MG.C:264-304 and solver.C:119-122: nSweepsPre 0, nSweepsPost 2,
    cycleType F, coarseMode direct, coarseLevel 0, smoother rbgs,
    minIter 0, maxIter 999."""
    from briscola_amd import bmg
    o = bmg.default_opts()
    assert (o.n_sweeps_pre, o.n_sweeps_post) == (0, 2)
    assert o.cycle == bmg.FCYCLE
    assert o.coarse_mode == bmg.COARSE_DIRECT
    assert o.coarse_level == 0
    assert o.smoother == bmg.RBGS
    assert (o.min_iter, o.max_iter) == (0, 999)
    assert o.precision == bmg.FP64


def test_struct_layouts_match_header(tmp_path):
    """ctypes structures agree with the C compiler's layout of bmg.h."""
    from briscola_amd import bmg
    src = tmp_path / "layout.c"
    src.write_text("""#include <stdio.h>
#include <stddef.h>
#include "bmg.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(bmg_grid), offsetof(bmg_grid, bc),
         offsetof(bmg_grid, bc_value), sizeof(bmg_opts), offsetof(bmg_opts, precision),
         sizeof(bmg_stats), offsetof(bmg_stats, solve_ms),
         offsetof(bmg_stats, learned_fallback),
         offsetof(bmg_stats, learned_guess_residual));
  return 0;
}""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    c = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                        check=True).stdout.split()]
    py = [ctypes.sizeof(bmg.Grid), bmg.Grid.bc.offset, bmg.Grid.bc_value.offset,
          ctypes.sizeof(bmg.Opts), bmg.Opts.precision.offset,
          ctypes.sizeof(bmg.Stats), bmg.Stats.solve_ms.offset,
          bmg.Stats.learned_fallback.offset, bmg.Stats.learned_guess_residual.offset]
    assert py == c


def test_last_error_is_a_string():
    from briscola_amd import bmg
    assert isinstance(bmg.load().bmg_last_error(), bytes)


def test_null_arguments_rejected_without_gpu():
    """Argument checks come before any device call."""
    from briscola_amd import bmg
    lib = bmg.load()
    assert lib.bmg_create(None, None, None) == -1
    assert lib.bmg_destroy(None) == 0
    assert lib.bmg_num_levels(None) == -1

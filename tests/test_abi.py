"""The drop-in boundary: libpm.so loads and exports every entry point that
include/photon_mapper.h declares (the reference's three launchers,
photonMappingKernel.cu:1523,1549,1569, plus the pm_* API), and argument
validation works without a GPU."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

import pmamd

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "photon_mapper.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b((?:pm|launch)_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_reference_launchers():
    names = declared_functions()
    for n in ("launch_init_random_numbers_kernel", "launch_emit_photons_kernel", "launch_photon_mapping_kernel"):
        assert n in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pmamd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert sorted(pmamd.exported_symbols()) == declared_functions()


def test_presets_and_validation_without_gpu():
    lib = pmamd.load()
    for p in (1, 2, 3, 5):
        c = pmamd.preset(p)
        assert c.width > 0 and c.k_surface in (32, 64)
    bad = pmamd.PmConfig()
    assert lib.pm_config_preset(99, ctypes.byref(bad)) == pmamd.PM_ERR_INVALID
    h = ctypes.c_void_p()
    c = pmamd.preset(1)
    c.k_surface = 65  # above the 64-lane gather limit
    assert lib.pm_create(ctypes.byref(c), ctypes.byref(h)) == pmamd.PM_ERR_INVALID
    c = pmamd.preset(1)
    c.leaf_size = 0
    assert lib.pm_create(ctypes.byref(c), ctypes.byref(h)) == pmamd.PM_ERR_INVALID
    c = pmamd.preset(1)
    c.spheres[0].material = 7
    assert lib.pm_create(ctypes.byref(c), ctypes.byref(h)) == pmamd.PM_ERR_INVALID
    assert lib.pm_create(None, ctypes.byref(h)) == pmamd.PM_ERR_INVALID
    assert lib.pm_emit(None) == pmamd.PM_ERR_INVALID
    assert lib.pm_status_string(pmamd.PM_ERR_STATE) == b"call order violated"


def test_options_and_accessors_reject_bad_arguments_without_gpu():
    lib = pmamd.load()
    # no context: every option call is refused before any GPU work
    for opt in (pmamd.PM_OPT_GATHER_MODE, pmamd.PM_OPT_VOXEL_GRID, pmamd.PM_OPT_TOP_LEVELS,
                pmamd.PM_OPT_CACHE_NEAR_GROUPS, pmamd.PM_OPT_CACHE_EMA, pmamd.PM_OPT_CACHE_SEEDS):
        assert lib.pm_set_option(None, opt, 0) == pmamd.PM_ERR_INVALID
    t = (ctypes.c_uint64 * 8)()
    assert lib.pm_get_voxel_table(None, 0, t, 8) == pmamd.PM_ERR_INVALID


def test_struct_layout_matches_c():
    # sizeof checks: the ctypes mirror must match the C layout of the header.
    assert ctypes.sizeof(pmamd.PmSphere) == 36
    assert ctypes.sizeof(pmamd.PmStats) == 24 + 12 * 8 + 4 * 8 + 16


def test_header_compiles_as_c_with_the_mirror_layout(tmp_path):
    """The header is plain C99 (what a cgo or C caller includes), and every
    struct's size and every field's offset equal the ctypes mirror's."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    structs = {"pm_sphere": pmamd.PmSphere, "pm_config": pmamd.PmConfig, "pm_stats": pmamd.PmStats,
               "pm_cache_stats": pmamd.PmCacheStats, "pm_batch_tuning": pmamd.PmBatchTuning,
               "pm_batch_stats": pmamd.PmBatchStats}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "photon_mapper.h"', "int main(void) {"]
    want = []
    for cname, py in structs.items():
        src.append('  printf("%%zu\\n", sizeof(%s));' % cname)
        want.append(ctypes.sizeof(py))
        for f in py._fields_:
            src.append('  printf("%%zu\\n", offsetof(%s, %s));' % (cname, f[0]))
            want.append(getattr(py, f[0]).offset)
    src.append("  return 0;\n}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(REPO, "include")
    subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-I", inc, "-o", str(exe), str(c)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == want

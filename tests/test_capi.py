"""CPU checks of the C-ABI boundary: the HIP library loads and exports every entry point that
include/gato_mi355x.h declares, with the ctypes signatures used by gato_amd._lib; argument
validation that needs no device. No compute calls (no GPU here)."""
import ctypes as C
import re

import pytest

from conftest import ROOT


def header_symbols():
    src = (ROOT / "include" / "gato_mi355x.h").read_text()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gato_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from gato_amd import _lib

    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes table covers the header exactly
    assert sorted(_lib.SIGNATURES) == syms


def test_struct_layouts_match_header():
    from gato_amd._lib import GatoParams, GatoStats

    # 15 reference scalars + 4 int32 extensions, 4 bytes each
    assert C.sizeof(GatoParams) == 19 * 4
    assert C.sizeof(GatoStats) == 8 + 4 + 4 + 9 * 8


def test_struct_offsets_match_the_c_compiler(tmp_path):
    # the ctypes mirrors against what a C compiler makes of the header: size and every field offset
    import shutil
    import subprocess

    from gato_amd._lib import GatoParams, GatoStats

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "gato_mi355x.h"', "int main(void) {"]
    for cname, ct in (("gato_params", GatoParams), ("gato_stats", GatoStats)):
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                         text=True).stdout.splitlines())
    for cname, ct in (("gato_params", GatoParams), ("gato_stats", GatoStats)):
        assert int(got[f"{cname} size"]) == C.sizeof(ct)
        for f, _ in ct._fields_:
            assert int(got[f"{cname} {f}"]) == getattr(ct, f).offset, (cname, f)


def test_argument_validation_without_device():
    from gato_amd import _lib

    lib = _lib.load()
    nq, nx = C.c_int32(), C.c_int32()
    assert lib.gato_plant_dims(0, C.byref(nq), C.byref(nx)) == 0 and (nq.value, nx.value) == (7, 14)
    assert lib.gato_plant_dims(1, C.byref(nq), C.byref(nx)) == 0 and (nq.value, nx.value) == (6, 12)
    assert lib.gato_plant_dims(5, C.byref(nq), C.byref(nx)) == 1001
    assert b"unknown plant" in lib.gato_last_error()
    h = C.c_void_p()
    p = _lib.GatoParams()
    p.max_sqp_iters = 5
    # bad knot count / batch / plant are rejected before any device call
    assert lib.gato_create(C.byref(h), 0, 1, 4, C.byref(p)) == 1001
    assert lib.gato_create(C.byref(h), 0, 32, 0, C.byref(p)) == 1001
    assert lib.gato_create(C.byref(h), 9, 32, 4, C.byref(p)) == 1001
    assert lib.gato_solve(None, None, None, None, 0.01, None, None) == 1001
    assert lib.gato_destroy(None) == 0
    # backward, launch shape, tuner and mixed batching validate before touching a device
    assert lib.gato_backward(None, None, None, None, 0.01, None, None, None, None, None, None, None) == 1001
    assert lib.gato_set_backward_params(None, 10, 1e-6, 1) == 1001
    assert lib.gato_set_launch_shape(None, 8) == 1001
    assert lib.gato_autotune(None, None, None, None, 0.01, None, None) == 1001
    hm = C.c_void_p()
    one = (C.c_int32 * 1)(0)
    assert lib.gato_mixed_create(C.byref(hm), 0, one, one, one, C.byref(p)) == 1001
    assert lib.gato_mixed_solve(None, None, None, None, 0.01, None, None) == 1001
    assert lib.gato_mixed_destroy(None) == 0
    # MPC helpers: path generation and the wrench estimator check their arguments first
    off = (C.c_double * 3)(0.0, 0.5, 0.6)
    assert lib.gato_figure8_path(None, 600, 5, 0.4, 0.4, off, 0.7, None) == 1001
    assert lib.gato_figure8_path(C.c_void_p(16), 0, 5, 0.4, 0.4, off, 0.7, None) == 1001
    he = C.c_void_p()
    assert lib.gato_force_est_create(C.byref(he), 3, 5.0, 2.0, 20.0, 0.5, 0) == 1001  # needs batch > 3
    assert lib.gato_force_est_create(C.byref(he), 16, 5.0, 30.0, 20.0, 0.5, 0) == 1001  # min > max radius
    assert lib.gato_force_est_create(C.byref(he), 16, 5.0, 2.0, 20.0, 1.5, 0) == 1001  # smoothing outside [0, 1]
    assert lib.gato_force_est_generate(None, None, None, None, None) == 1001
    assert lib.gato_force_est_update(None, None, None, None, 0.5, 0.8, None, None) == 1001
    assert lib.gato_force_est_destroy(None) == 1001
    # the allocation-free solve refuses bad arguments and a workspace beyond its static buffer before any device call
    d = C.c_void_p(256)  # never dereferenced: every case below is refused first
    p.max_sqp_iters = 10
    assert lib.gato_solve_noalloc(0, 32, 4, C.byref(p), None, d, d, 0.01, None, None) == 1001
    assert lib.gato_solve_noalloc(7, 32, 4, C.byref(p), d, d, d, 0.01, None, None) == 1001
    assert lib.gato_solve_noalloc(0, 129, 4, C.byref(p), d, d, d, 0.01, None, None) == 1001
    assert lib.gato_solve_noalloc(0, 32, 4096, C.byref(p), d, d, d, 0.01, None, None) == 1001
    msg = lib.gato_last_error().decode()
    assert "static" in msg and str(32 << 20) in msg


def test_drop_in_module_surface():
    import importlib

    mod = importlib.import_module("gato_amd.bsqp.bsqpN32_iiwa14")
    assert mod.KNOT_POINTS == 32
    for B in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024):
        cls = getattr(mod, f"BSQP_{B}_float")
        for m in ("solve", "reset_dual", "set_f_ext_batch", "set_rho_penalty_batch", "set_drho_batch",
                  "set_mu_batch", "set_pcg_tol_batch", "sim_forward", "reset_rho", "set_rho_adaptation"):
            assert callable(getattr(cls, m))
    assert importlib.import_module("gato_amd.bsqp.bsqpN128_indy7").KNOT_POINTS == 128

"""C-ABI boundary checks that need no GPU: the library loads, exports every entry point declared
in include/vib.h, and the host-only helpers (whitening, preintegration) match the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vib.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vib_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(abi):
    lib = abi.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (vib_[a-z_0-9]+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    assert lib.vib_abi_version() == 3


def test_create_without_device_fails_cleanly(abi):
    import ctypes as C
    lib = abi.load_library()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("device present")
    except Exception:
        pass
    h = C.c_void_p()
    assert lib.vib_create(C.byref(h), 0) == 13  # VIB_ERR_NO_DEVICE


def test_footprint_is_positive(abi):
    from vibpy import synth
    w = synth.generate(K=5, L=100, seed=3)
    s = w.cstruct()
    import ctypes as C
    n = abi.load_library().vib_window_footprint(C.byref(s), 0)
    assert n > 8 * (w.F * 10 + w.L * 7)


def test_device_entry_points_reject_bad_arguments_before_any_device_call(abi):
    """The batch preintegration, whitening, gate and refit entry points check their arguments on
    the host first: a null context or a malformed request is VIB_ERR_INVALID_ARGUMENT, with no
    HIP call made (so this runs without a GPU)."""
    import ctypes as C
    lib = abi.load_library()
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    p = abi.imu_params()
    smp = np.zeros((4, 7))
    i1 = np.zeros(1, np.int32)
    d9 = np.zeros(9)
    out = np.zeros(abi.IMU_STRIDE)
    st = np.zeros(1, np.int32)
    f = lambda a: a.ctypes.data_as(dp)
    g = lambda a: a.ctypes.data_as(ip)
    assert lib.vib_preintegrate_batch(None, 1, f(smp), 4, g(i1), g(i1), f(d9), f(d9), f(d9), C.byref(p), f(d9), f(d9),
                                      f(out), g(st), None) == abi.VIB_ERR_INVALID_ARGUMENT
    px = np.zeros(2)
    intr = np.array([400.0, 400.0, 320.0, 240.0, 640.0, 480.0])
    b3 = np.zeros(3)
    assert lib.vib_whiten_observations(None, 1, f(px), None, f(intr), 1, None, 1.0, f(b3), f(b3), g(st),
                                       None) == abi.VIB_ERR_INVALID_ARGUMENT
    rep = abi.VibGateReport()
    assert lib.vib_nis_gate(None, 0, 9.21, 0, C.byref(rep), None) == abi.VIB_ERR_INVALID_ARGUMENT
    rr = abi.VibRefitReport()
    ro = abi.VibRefitOptions(9.21, 3, 0)
    opts = abi.options()
    assert lib.vib_solve_refit(None, 0, C.byref(opts), C.byref(ro), C.byref(rr)) == abi.VIB_ERR_INVALID_ARGUMENT


def test_result_structs_match_the_header(abi, tmp_path):
    """The ctypes mirrors of the gate, refit and solve records have the sizes and field offsets
    that the C compiler gives the structs of include/vib.h."""
    import ctypes as C
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "vib.h"\nint main(void) {\n'
                   'printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(vib_residual_stats), sizeof(vib_gate_report),\n'
                   'sizeof(vib_refit_options), sizeof(vib_refit_report), sizeof(vib_solve_report),\n'
                   'offsetof(vib_refit_report, pre), offsetof(vib_gate_report, admitted));\nreturn 0;\n}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(abi.VibResidualStats), C.sizeof(abi.VibGateReport), C.sizeof(abi.VibRefitOptions),
            C.sizeof(abi.VibRefitReport), C.sizeof(abi.VibSolveReport), abi.VibRefitReport.pre.offset,
            abi.VibGateReport.admitted.offset]
    assert got == want


def test_window_and_keyframe_structs_match_the_header(abi, tmp_path):
    """The ctypes mirrors of vib_window and vib_keyframe have the C compiler's sizes and offsets."""
    import ctypes as C
    src = tmp_path / "win.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "vib.h"\nint main(void) {\n'
                   'printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(vib_window), offsetof(vib_window, prior_rows),\n'
                   'offsetof(vib_window, lm_id), sizeof(vib_keyframe), offsetof(vib_keyframe, num_landmarks),\n'
                   'offsetof(vib_keyframe, obs_sqrt_info));\nreturn 0;\n}\n')
    exe = tmp_path / "win"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(abi.VibWindow), abi.VibWindow.prior_rows.offset, abi.VibWindow.lm_id.offset,
            C.sizeof(abi.VibKeyframe), abi.VibKeyframe.num_landmarks.offset, abi.VibKeyframe.obs_sqrt_info.offset]
    assert got == want

"""The C ABI: every entry point include/*.h declares is exported by
libcitcom_hip.so, and the library refuses to run without a GPU (no CPU
fallback)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header="citcom_hip.h", prefix="chip_"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_A-Z0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    from citcom_amd import hip
    assert os.path.exists(hip.LIB_PATH), "build the library first"
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (\w+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    assert len(declared()) >= 20
    learn = declared("citcom_learn.h", "cl_")
    assert len(learn) >= 8 and not [s for s in learn if s not in exported]


def test_no_cpu_fallback_without_device():
    import torch  # noqa: F401  (same process state as the product)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from citcom_amd import deck
    from citcom_amd.hip import ChipError, Citcom
    p = deck.problem_from_deck({k: str(v) for k, v in deck.workload("tiny").items()})
    with pytest.raises(ChipError, match="no HIP device"):
        Citcom(p)


def test_null_context_refused():
    """Every entry point that takes a context refuses a null one with an
    error code and a message (no device work is started)."""
    import ctypes as C
    from citcom_amd import hip
    L = hip.lib()
    calls = {
        "chip_solve_constrained_flow_iterative": (None, None),
        "chip_PG_timestep": (None, None),
        "chip_general_stokes_solver": (None, None),
        "chip_construct_stiffness_B_matrix": (None,),
        "chip_get_system_viscosity": (None,),
        "chip_set_option": (None, b"graphs", 1),
        "chip_heat_flux": (None, None),
    }
    for name, args in calls.items():
        assert getattr(L, name)(*args) == -1, name
        assert b"null context" in L.chip_last_error(), name
    buf = (C.c_double * 4)()
    assert L.chip_upload(None, b"U", 0, buf, 4) == -1
    assert L.chip_download(None, b"U", 0, buf, 4) == -1

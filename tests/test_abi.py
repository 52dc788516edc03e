"""The C-ABI library loads and exports every function include/*.h declares (CPU only)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(kt[g]?_[a-z0-9_]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "kt_build_b2b" in names and "kt_solve_both" in names and "ktg_generate" in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    from ktplace_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_python_binding_covers_context_api():
    from ktplace_amd import simpl
    declared = {n for n in declared_functions() if n.startswith("kt_")}
    missing = declared - set(simpl.EXPORTS)
    assert not missing, sorted(missing)


def test_no_gpu_fails_loudly():
    """Without a GPU the context cannot be created and says why (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from ktplace_amd.simpl import Context, KtError
    with pytest.raises(KtError):
        Context(0)

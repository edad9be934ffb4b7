"""The C ABI library loads on a machine without a GPU, exports every
function include/latttmat.h declares, and refuses to run without a device
(no CPU fallback)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "latttmat.h")).read()
    return sorted(set(re.findall(r"\b(tm_[a-z_]+)\s*\(", src)))


def test_header_matches_binding():
    import latttmat_amd._abi as abi
    assert sorted(abi.EXPORTS) == declared()


def test_library_exports_all():
    import latttmat_amd as lt
    assert sorted(lt.exported_symbols()) == declared()


def test_no_cpu_fallback():
    import latttmat_amd as lt
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(lt.TMError) as e:
        lt.TransferMatrix("A", 2, 1, 8)
    assert e.value.status == 3  # TM_EHIP
    assert "no HIP device" in str(e.value)


def test_version():
    import latttmat_amd as lt
    assert b"gfx950" in lt.lib().tm_version()

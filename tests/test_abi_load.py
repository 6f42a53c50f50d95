"""The C-ABI library loads, exports every symbol include/pstm_abi.h declares,
reports its ABI version and validates parameters on the host (no GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from seisproc_amd.abi_types import (ABI_VERSION, FLAG_AA, PSTM_ERR_INVALID, PSTM_ERR_UNSUPPORTED,
                                    PSTM_OK, PstmParams)
from seisproc_amd.processors import kernel_backend as kb

HEADER = Path(__file__).resolve().parent.parent / "include" / "pstm_abi.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pstm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = kb.load_library()
    names = _declared()
    assert len(names) >= 9
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(kb.SYMBOLS)
    assert lib.pstm_abi_version() == ABI_VERSION


def test_struct_size_matches_header():
    # offsetof-style check compiled from the header itself
    import subprocess, tempfile, os
    src = '#include <stdio.h>\n#include "%s"\nint main(){printf("%%zu", sizeof(pstm_params_t));}' % HEADER
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-o", os.path.join(d, "s"), c], check=True)
        out = subprocess.run([os.path.join(d, "s")], capture_output=True, text=True, check=True).stdout
    assert int(out) == ctypes.sizeof(PstmParams)


def _p(**kw):
    d = dict(n_traces=10, n_samples=100, n_il=4, n_xl=4, n_tau=100)
    d.update(kw)
    return PstmParams.from_dict(d)


@pytest.mark.parametrize("kw,status,msg", [
    (dict(), PSTM_OK, ""),
    (dict(dt=0.0), PSTM_ERR_INVALID, "dt"),
    (dict(max_angle_deg=95.0), PSTM_ERR_INVALID, "max_angle"),
    (dict(xl_dx=10.0, xl_dy=10.0), PSTM_ERR_INVALID, "orthogonal"),
    # any record length (windows longer than the LDS share are staged in segments)
    (dict(n_samples=10600), PSTM_OK, ""),
    (dict(n_samples=120000), PSTM_OK, ""),
    (dict(n_samples=(1 << 24) + 1), PSTM_ERR_UNSUPPORTED, "n_samples"),
    (dict(n_samples=1), PSTM_ERR_INVALID, "n_samples"),
    (dict(flags=FLAG_AA, aa_max=0.5), PSTM_ERR_INVALID, "aa_max"),
    (dict(flags=FLAG_AA, n_samples=50000), PSTM_OK, ""),
    (dict(min_offset_m=500.0, max_offset_m=400.0), PSTM_ERR_INVALID, "min_offset_m"),
    (dict(flags=1 << 12), PSTM_ERR_INVALID, "flag"),
    (dict(tau0=-0.1), PSTM_ERR_INVALID, "tau0"),
])
def test_validate_on_host(kw, status, msg):
    lib = kb.load_library()
    p = _p(**kw)
    rc = lib.pstm_validate(ctypes.byref(p))
    assert rc == status
    if msg:
        assert msg in kb.last_error(lib)


def test_missing_library_is_an_error(tmp_path):
    with pytest.raises(kb.BackendUnavailable):
        kb.load_library(tmp_path / "nope.so")


def test_only_hip_backend():
    assert kb.get_effective_backend() is kb.KernelBackend.HIP
    with pytest.raises(ValueError):
        kb.get_effective_backend("metal_cpp")

"""The C-ABI library loads and exports every symbol include/*.h declares; the
host-only entry points agree with the oracle.  No device calls."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(mcrg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared_functions()
    for required in ("mcrg_ctx_create", "mcrg_chain_create", "mcrg_chain_sweep",
                     "mcrg_chain_get_records", "mcrg_thresholds"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import mcrg_amd
    out = subprocess.run(["nm", "-D", "--defined-only", mcrg_amd.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = _declared_functions() - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    for name in _declared_functions():
        assert hasattr(mcrg_amd.lib(), name)


def test_abi_version():
    import mcrg_amd
    assert mcrg_amd.abi_version() == 3


@pytest.mark.parametrize("J", [0.0, 0.1, 0.42, 0.4406868, 0.4688157, 0.489652, 1.0, 9.0])
def test_thresholds_match_oracle(orc, J):
    import mcrg_amd
    assert list(mcrg_amd.thresholds(J)) == list(orc.thresholds(J))


def test_thresholds_reject_antiferromagnetic():
    import mcrg_amd
    with pytest.raises(mcrg_amd.MCRGError) as e:
        mcrg_amd.thresholds(-0.44)
    assert e.value.status == mcrg_amd.ERR_INVALID_ARG
    assert "coupling" in str(e.value)


@pytest.mark.parametrize("nl", [0, 1, 5, 10])
def test_record_size_matches_oracle(orc, nl):
    import mcrg_amd
    assert mcrg_amd.record_size(nl) == orc.record_size(nl) == 22 * nl + 16

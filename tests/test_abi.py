"""The C-ABI library: it loads, exports every function include/qarp_hip.h
declares, and the Python struct images match the C layout.  No GPU calls."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from openqarp_amd import _abi, _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "qarp_hip.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^int\s+(qhip_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_python_export_list():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_loads_and_exports_every_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.library_path())], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (qhip_\w+)", out))
    assert set(declared_functions()) <= exported


def _c_offsets():
    """offsetof/sizeof of the ABI structs as a C compiler lays them out."""
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "qarp_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(qhip_op), offsetof(qhip_op, gate),
    offsetof(qhip_op, n_qubits), offsetof(qhip_op, qubit_offset), offsetof(qhip_op, n_params),
    offsetof(qhip_op, flags), offsetof(qhip_op, params), offsetof(qhip_op, unitary_offset),
    offsetof(qhip_op, dep_offset), offsetof(qhip_op, n_deps), offsetof(qhip_op, cbit));
  printf("%zu %zu %zu %zu %zu\n", sizeof(qhip_pauli_term), offsetof(qhip_pauli_term, offset),
    offsetof(qhip_pauli_term, n_letters), offsetof(qhip_pauli_term, coeff_re),
    offsetof(qhip_pauli_term, coeff_im));
  printf("%zu %zu %zu %zu %zu\n", sizeof(qhip_stream), sizeof(qhip_observable), sizeof(qhip_config),
    sizeof(qhip_apply_stats), sizeof(qhip_batch_info));
  return 0;
}
'''
    return src


def test_struct_layouts_match_the_header(tmp_path):
    c = tmp_path / "layout.c"
    c.write_text(_c_offsets())
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(c), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    op = list(map(int, lines[0].split()))
    d = _abi.OP_DTYPE
    assert op[0] == d.itemsize
    names = ["gate", "n_qubits", "qubit_offset", "n_params", "flags", "params", "unitary_offset",
             "dep_offset", "n_deps", "cbit"]
    assert op[1:] == [d.fields[n][1] for n in names]
    term = list(map(int, lines[1].split()))
    t = _abi.TERM_DTYPE
    assert term[0] == t.itemsize
    assert term[1:] == [t.fields[n][1] for n in ["offset", "n_letters", "coeff_re", "coeff_im"]]
    sizes = list(map(int, lines[2].split()))
    assert sizes == [C.sizeof(_abi.QhipStream), C.sizeof(_abi.QhipObservable), C.sizeof(_abi.QhipConfig),
                     C.sizeof(_abi.QhipApplyStats), C.sizeof(_abi.QhipBatchInfo)]


def test_missing_library_raises_backend_unavailable(monkeypatch, tmp_path):
    from openqarp_amd.errors import BackendUnavailable
    monkeypatch.setenv("QHIP_LIBRARY", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(BackendUnavailable):
        _lib.load()
    monkeypatch.setattr(_lib, "_lib", None)

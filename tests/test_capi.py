"""CPU checks of the C ABI library: it loads, exports every function include/xnb_amd.h declares,
and its host-only generator has the stated properties.  No GPU compute call is made."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import exanbody_amd
from exanbody_amd import _lib, generate_fcc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "xnb_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xnb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers the header exactly
    assert sorted(_lib.exported_symbols()) == names


def test_package_refuses_to_run_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(ImportError):
        _lib.load()


def test_error_strings_and_no_device_path():
    L = exanbody_amd.load()
    assert L.xnb_error_string(0) == b"ok"
    assert L.xnb_error_string(_lib.XNB_E_OVERFLOW) == b"neighbour list overflow"
    if exanbody_amd.device_count() == 0:
        d = _lib.Domain()
        for a in range(3):
            d.bounds_max[a] = 9.0
            d.grid_dims[a] = 3
            d.periodic[a] = 1
        d.cell_size = 3.0
        h = C.c_void_p()
        assert L.xnb_create(C.byref(h), C.byref(d), 0) == _lib.XNB_E_HIP
        assert not h.value


def test_domain_validation_before_device():
    L = exanbody_amd.load()
    d = _lib.Domain()
    for a in range(3):
        d.bounds_max[a] = 10.0
        d.grid_dims[a] = 3
    d.cell_size = 3.0  # 3 * 3 != 10: check_domain fails before any device call
    h = C.c_void_p()
    assert L.xnb_create(C.byref(h), C.byref(d), 0) == _lib.XNB_E_DOMAIN


def test_generator_properties():
    r, v, ids, L = generate_fcc(6, 0.8442, jitter=0.01, temperature=0.72, seed=5)
    n = len(r)
    assert n == 4 * 6 ** 3
    assert L == pytest.approx(6 * (4 / 0.8442) ** (1 / 3), rel=1e-15)
    assert np.all(r >= 0) and np.all(r < L)
    a = L / 6
    base = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    sites = ((g[:, None, :] + base[None]) * a).reshape(-1, 3)
    d = r - sites
    d -= L * np.round(d / L)
    assert np.abs(d).max() <= 0.01 + 1e-12
    assert np.abs(v.sum(0)).max() < 1e-12
    assert (v * v).sum() / (3 * n - 3) == pytest.approx(0.72, rel=1e-13)
    assert np.array_equal(ids, np.arange(n))
    r2, v2, _, _ = generate_fcc(6, 0.8442, jitter=0.01, temperature=0.72, seed=5)
    assert np.array_equal(r, r2) and np.array_equal(v, v2)
    r3, _, _, _ = generate_fcc(6, 0.8442, jitter=0.01, temperature=0.72, seed=6)
    assert not np.array_equal(r, r3)


def test_generator_partial_fill_density():
    r, v, ids, L = generate_fcc(5, 0.8, n_atoms=300, jitter=0.0, temperature=1.0, seed=1)
    assert len(r) == 300 and L ** 3 * 0.8 == pytest.approx(300, rel=1e-12)
    assert len({tuple(x) for x in np.round(r, 9)}) == 300


# C struct of include/xnb_amd.h -> its ctypes mirror in the package
def _ctypes_mirrors():
    from exanbody_amd import serve
    return {"xnb_domain": _lib.Domain, "xnb_lj_params": _lib.LJParams,
            "xnb_chunk_neighbors_config": _lib.ChunkNeighborsConfig, "xnb_run_params": _lib.RunParams,
            "xnb_job": serve.Job, "xnb_job_result": serve.JobResult, "xnb_serve_config": serve.ServeConfig,
            "xnb_serve_stats": serve.ServeStats}


def test_header_compiles_as_c_and_ctypes_layouts_match(tmp_path):
    """include/xnb_amd.h is a C header (gcc -std=c99 -pedantic -Werror); every struct the
    bindings mirror has the C compiler's size and field offsets in its ctypes mirror."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    mirrors = _ctypes_mirrors()
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "xnb_amd.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe), str(src)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), f"{cname}: C size {got[(cname, 'size')]} != ctypes {C.sizeof(py)}"
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, f"{cname}.{f[0]}: offset differs"
    # every field of the C struct is mirrored (no C field left out of the ctypes table)
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "xnb_amd.h")).read(), flags=re.S)
    for cname, py in mirrors.items():
        body = re.search(r"typedef struct \{([^{}]*)\}\s*" + cname + ";", text).group(1)
        names = []
        for decl in body.split(";"):
            for part in decl.split(","):
                m = re.search(r"([A-Za-z_]\w*)\s*(?:\[[^\]]*\])?\s*$", part.strip())
                if m:
                    names.append(m.group(1))
        assert names == [f[0] for f in py._fields_], f"{cname}: {names}"


def test_job_bytes_estimate():
    """xnb_job_bytes_for grows with the atoms, with the neighbours per atom and with the ghost
    share of a small periodic box; xnb_job_bytes(n) is the value at 0.4 n ghosts, 110
    neighbours and n cells."""
    from exanbody_amd.serve import job_bytes, make_job
    L = _lib.load()
    a = make_job(4000, 0.6, 1.0, 10, _lib.PRECISION_FP64, seed=1)
    b = make_job(4000, 1.2, 1.0, 10, _lib.PRECISION_FP64, seed=1)
    c = make_job(8000, 1.2, 1.0, 10, _lib.PRECISION_FP64, seed=1)
    assert job_bytes(a) < job_bytes(b) < job_bytes(c)
    dom = b["domain"]
    D = np.asarray(dom["grid_dims"], float)
    vol = np.prod(np.asarray(dom["bounds_max"]) - np.asarray(dom["bounds_min"]))
    nb = 4000 / vol * 4.0 / 3.0 * np.pi * 2.8 ** 3
    ghost = np.prod((D + 2) / D)
    want = int(4000 * ghost * 500 + 4000 * 4 * nb + np.prod(D + 2) * 64) + (16 << 20)
    assert abs(job_bytes(b) - want) <= 2  # last-bit rounding
    # a non-periodic box has no ghost slots
    d = dict(b, domain=dict(dom, periodic=(0, 0, 0)))
    assert job_bytes(d) < job_bytes(b)
    assert L.xnb_job_bytes(4000) == int(1.4 * 4000 * 500 + 4000 * 4 * 110.0 + 4000 * 64) + (16 << 20)

"""The C ABI without a GPU: the library loads, exports every function that
include/dte.h declares, validates arguments before touching a device, and
reports a missing device with its own status code.  Also checks that the
product never links or imports the oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from datamininglibs_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_function():
    L = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in capi.declared_symbols():
        assert n in exported, n  # unmangled: extern "C"


def test_product_does_not_link_or_import_the_oracle():
    deps = subprocess.run(["ldd", capi.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in deps
    for dirpath, _, files in os.walk(os.path.join(ROOT, "datamininglibs_amd")):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cc", ".cpp")):
                txt = open(os.path.join(dirpath, fn), errors="replace").read()
                assert "pyoracle" not in txt and "liboracle" not in txt and "dte_oracle" not in txt, fn


def test_params_default_matches_reference_defaults():
    for alg in (capi.RANDOM_FOREST, capi.EXTRA_TREES):
        p = capi.default_params(alg)
        # ensembles_interface.cc:34-56: 100 trees, depth 1000, min node 5, k from round(ln F)+1
        assert (p.algorithm, p.n_trees, p.max_depth, p.min_node_size, p.n_features_per_node) == (alg, 100, 1000, 5, 0)
        assert p.task == capi.CLASSIFICATION and p.build_mode == capi.BUILD_BATCHED


def _fit_rc(params, X=None, y=None, n=8, f=2):
    L = capi.lib()
    X = np.zeros((f, n), np.float32) if X is None else X
    y = np.zeros(n, np.float32) if y is None else y
    h = ctypes.c_void_p()
    rc = L.dte_fit(X.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), n, f,
                   ctypes.byref(params) if params is not None else None, ctypes.byref(h))
    return rc


def _last_error():
    buf = ctypes.create_string_buffer(512)
    capi.lib().dte_last_error(buf, 512)
    return buf.value.decode()


@pytest.mark.parametrize("field,value", [("algorithm", 7), ("task", 3), ("n_trees", 0), ("max_depth", -1),
                                         ("min_node_size", -2), ("n_classes", 0), ("n_classes", 33),
                                         ("build_mode", 9), ("n_features_per_node", -1), ("tree_offset", -1),
                                         ("cut_team", 32), ("cut_team", 513), ("cut_block", 4096),
                                         ("bucket_target", 16), ("bucket_target", 1024), ("cut_ert", 8)])
def test_invalid_params_rejected_before_device(field, value):
    p = capi.default_params(capi.RANDOM_FOREST)
    setattr(p, field, value)
    assert _fit_rc(p) == capi.ERR_INVALID_ARGUMENT
    assert _last_error()


def test_null_and_empty_arguments():
    assert _fit_rc(None) == capi.ERR_INVALID_ARGUMENT
    p = capi.default_params(capi.EXTRA_TREES)
    L = capi.lib()
    h = ctypes.c_void_p()
    assert L.dte_fit(None, None, 8, 2, ctypes.byref(p), ctypes.byref(h)) == capi.ERR_INVALID_ARGUMENT
    assert _fit_rc(p, n=0) == capi.ERR_INVALID_ARGUMENT
    assert L.dte_forest_free(None) == capi.OK
    assert L.dte_params_default(5, ctypes.byref(p)) == capi.ERR_INVALID_ARGUMENT
    assert L.dte_last_error(None, 0) in (capi.OK, capi.ERR_INVALID_ARGUMENT)


@pytest.mark.skipif(capi.device_count() > 0, reason="a device is present")
def test_no_device_is_reported():
    p = capi.default_params(capi.EXTRA_TREES)
    assert _fit_rc(p) == capi.ERR_NO_DEVICE
    assert "device" in _last_error()
    s = ctypes.c_void_p()
    assert capi.lib().dte_server_create(0, 1 << 30, 65536, ctypes.byref(s)) == capi.ERR_NO_DEVICE
    assert capi.lib().dte_server_create(0, 1 << 30, 0, ctypes.byref(s)) == capi.ERR_INVALID_ARGUMENT


def test_error_state_is_per_call():
    p = capi.default_params(capi.RANDOM_FOREST)
    p.n_trees = 0
    assert _fit_rc(p) == capi.ERR_INVALID_ARGUMENT
    n = ctypes.c_int32(-1)
    assert capi.lib().dte_device_count(ctypes.byref(n)) == capi.OK
    assert n.value >= 0


def test_cut_block_below_cut_team_rejected():
    p = capi.default_params(capi.RANDOM_FOREST, cut_team=512, cut_block=256)
    assert _fit_rc(p) == capi.ERR_INVALID_ARGUMENT
    assert "cut_block" in _last_error()


def test_autotune_validates_before_device():
    L = capi.lib()
    p = capi.default_params(capi.RANDOM_FOREST)
    out = capi.Params()
    X = np.zeros((2, 8), np.float32)
    y = np.zeros(8, np.float32)
    assert L.dte_autotune(X.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), 8, 2,
                          ctypes.byref(p), -1, ctypes.byref(out), None) == capi.ERR_INVALID_ARGUMENT
    assert L.dte_autotune(None, None, 8, 2, ctypes.byref(p), 0, ctypes.byref(out), None) == capi.ERR_INVALID_ARGUMENT


def test_forest_load_rejects_bad_files_before_device(tmp_path):
    L = capi.lib()
    h = ctypes.c_void_p()
    missing = tmp_path / "missing.dtef"
    assert L.dte_forest_load(str(missing).encode(), 0, ctypes.byref(h)) == capi.ERR_IO
    junk = tmp_path / "junk.dtef"
    junk.write_bytes(b"not a forest file at all, just some bytes" * 4)
    assert L.dte_forest_load(str(junk).encode(), 0, ctypes.byref(h)) == capi.ERR_INVALID_ARGUMENT
    assert "magic" in _last_error()
    bad = tmp_path / "bad.dtef"
    bad.write_bytes(b"DTEFRST1" + bytes(96))
    assert L.dte_forest_load(str(bad).encode(), 0, ctypes.byref(h)) == capi.ERR_INVALID_ARGUMENT
    assert "checksum" in _last_error()
    assert L.dte_forest_load(None, 0, ctypes.byref(h)) == capi.ERR_INVALID_ARGUMENT


def _fnv1a64(b):
    h = 0xcbf29ce484222325
    for x in b:
        h = ((h ^ x) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def _forest_file(path, *, left=(1, -1, -1), feature=(0, -1, -1), depth=(0, 1, 1), n_features=2, classes=2):
    """One classification tree: a root split into two leaves (format DTEFRST1, version 1)."""
    n = 3
    # magic, version, n_trees, n_classes, n_features, task, algorithm, fixed_shift, has_stats,
    # tree_offset, n_nodes, fit_rows, seed
    hdr = b"DTEFRST1" + np.array([1], np.uint32).tobytes() + \
        np.array([1, classes, n_features, 0, 0, 0, 0, 0], np.int32).tobytes() + \
        np.array([n, 8], np.int64).tobytes() + np.array([7], np.uint64).tobytes()
    body = (np.array([0, n], np.int64).tobytes() + np.array(feature, np.int32).tobytes() +
            np.array([0.5, -1, -1], np.float32).tobytes() + np.array(left, np.int32).tobytes() +
            np.array(depth, np.int32).tobytes() + np.zeros(n, np.float32).tobytes() +
            np.array([3, 2, 3, 0, 0, 2], np.uint32).tobytes())
    data = hdr + body
    path.write_bytes(data + np.array([_fnv1a64(data)], np.uint64).tobytes())


def test_forest_file_structure_checked_before_device(tmp_path):
    L = capi.lib()
    h = ctypes.c_void_p()
    good = tmp_path / "good.dtef"
    _forest_file(good)
    # a well-formed file passes every check and stops at the missing device
    assert L.dte_forest_load(str(good).encode(), 0, ctypes.byref(h)) == capi.ERR_NO_DEVICE
    for name, kw, msg in [("link", dict(left=(2, -1, -1)), "child link"),
                          ("feat", dict(feature=(5, -1, -1)), "feature"),
                          ("depth", dict(depth=(0, 2, 1)), "depth")]:
        p = tmp_path / f"{name}.dtef"
        _forest_file(p, **kw)
        assert L.dte_forest_load(str(p).encode(), 0, ctypes.byref(h)) == capi.ERR_INVALID_ARGUMENT, name
        assert msg in _last_error(), (name, _last_error())
    trunc = tmp_path / "trunc.dtef"
    data = good.read_bytes()[:-20]
    trunc.write_bytes(data + np.array([_fnv1a64(data)], np.uint64).tobytes())
    assert L.dte_forest_load(str(trunc).encode(), 0, ctypes.byref(h)) == capi.ERR_INVALID_ARGUMENT


def test_ctypes_structures_match_the_c_header(tmp_path):
    """Sizes and field offsets of the ctypes mirrors equal the C compiler's."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"dte_params": capi.Params, "dte_node": capi.Node, "dte_forest_info": capi.ForestInfo,
               "dte_server_stats": capi.ServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dte.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-I", inc, "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert got[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])


def test_wrappers_check_the_feature_dimension():
    """Forest.predict and Server.submit refuse arrays whose feature extent
    differs from the forest's before any library call (dte_predict reads
    n_rows * n_features floats, dte_server_submit the same)."""
    from types import SimpleNamespace
    fake = SimpleNamespace(n_features=5, task=capi.CLASSIFICATION, n_classes=3)
    with pytest.raises(ValueError):
        capi.Forest.predict(fake, np.zeros((4, 10), np.float32))
    with pytest.raises(ValueError):
        capi.Forest.predict(fake, np.zeros(10, np.float32))
    srv = SimpleNamespace(_meta={0: (5, 3)}, _h=None)
    with pytest.raises(ValueError):
        capi.Server.submit(srv, 0, np.zeros((10, 4), np.float32))
    with pytest.raises(ValueError):
        capi.Server.submit(srv, 0, np.zeros(5, np.float32))

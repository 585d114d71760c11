"""The C ABI of libsph_hip.so on the CPU: the library loads, exports every entry
point include/sph_hip.h declares, validates arguments before touching a device,
and its host-side scene generators match independent restatements."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, lattice_numpy

HEADER = os.path.join(ROOT, "include", "sph_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_phase_api():
    names = declared_symbols()
    for must in ("sph_create", "sph_destroy", "sph_upload", "sph_append", "sph_set_gravity", "sph_phase",
                 "sph_step", "sph_download", "sph_debug", "sph_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import sphhip
    lib = sphhip.load()
    out = subprocess.run(["nm", "-D", "--defined-only", sphhip.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (sph_[a-z_0-9]+)$", out.stdout, flags=re.M))
    missing = [n for n in declared_symbols() if n not in exported]
    assert not missing, missing
    for n in declared_symbols():
        getattr(lib, n)
    # the Python binding covers every entry point
    assert set(sphhip._lib.SIGNATURES) == set(declared_symbols())


def test_library_does_not_link_the_oracle():
    import sphhip
    out = subprocess.run(["ldd", sphhip.LIB_PATH], capture_output=True, text=True, check=True)
    assert "oracle" not in out.stdout
    syms = subprocess.run(["nm", "-D", sphhip.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in syms


def test_argument_validation_before_device(hip_host):
    lib, _lib = hip_host
    p = _lib.sph_params()
    ctx = C.c_void_p()
    lib.sph_default_params(C.byref(p), 0.4, 0)
    assert lib.sph_create(C.byref(p), C.byref(ctx)) == _lib.SPH_ERR_INVALID
    assert b"capacity" in lib.sph_last_error()
    lib.sph_default_params(C.byref(p), -1.0, 10)
    assert lib.sph_create(C.byref(p), C.byref(ctx)) == _lib.SPH_ERR_INVALID
    assert lib.sph_create(None, C.byref(ctx)) == _lib.SPH_ERR_INVALID
    assert lib.sph_upload(None, None, None, None, 0) == _lib.SPH_ERR_INVALID
    assert lib.sph_step(None, 1, None) == _lib.SPH_ERR_INVALID
    lib.sph_destroy(None)


def test_noalloc_limits_checked_before_device(hip_host):
    """sph_step_noalloc rejects null arrays, sizes above its stated maximum and
    grids above its cell maximum before any device call; the header, the binding
    and the library agree on the limits."""
    lib, _lib = hip_host
    text = open(HEADER).read()
    assert int(re.search(r"#define SPH_NOALLOC_MAX_PARTICLES (\d+)u", text).group(1)) == _lib.NOALLOC_MAX_PARTICLES
    assert int(re.search(r"#define SPH_NOALLOC_MAX_CELLS (\d+)u", text).group(1)) == _lib.NOALLOC_MAX_CELLS
    p = _lib.sph_params()
    lib.sph_default_params(C.byref(p), 0.4, 8)
    st = _lib.sph_device_state()
    assert lib.sph_step_noalloc(C.byref(p), C.byref(st), 8, 1, None) == _lib.SPH_ERR_INVALID
    for name, _ in _lib.sph_device_state._fields_:   # never dereferenced on these paths
        setattr(st, name, 0x1000)
    assert lib.sph_step_noalloc(None, C.byref(st), 8, 1, None) == _lib.SPH_ERR_INVALID
    assert lib.sph_step_noalloc(C.byref(p), C.byref(st), _lib.NOALLOC_MAX_PARTICLES + 1, 1, None) == \
        _lib.SPH_ERR_CAPACITY
    assert b"65536" in lib.sph_last_error()
    lib.sph_default_params(C.byref(p), 3.0, 8)   # 66^3 = 287,496 cells
    assert lib.sph_step_noalloc(C.byref(p), C.byref(st), 8, 1, None) == _lib.SPH_ERR_CAPACITY
    assert b"262144" in lib.sph_last_error()


def test_variant_table_matches_header_and_binding(hip_host):
    """SPH_NUM_VARIANTS in the header, the binding's NUM_VARIANTS and the library's
    table agree: every index below it has a shape, the next one is rejected, and
    the default is a word-queue shape."""
    lib, _lib = hip_host
    n = int(re.search(r"#define SPH_NUM_VARIANTS (\d+)", open(HEADER).read()).group(1))
    assert _lib.NUM_VARIANTS == n
    vals = [C.c_int() for _ in range(5)]
    for v in range(n):
        assert lib.sph_variant_shape(v, *[C.byref(x) for x in vals]) == _lib.SPH_OK, v
        threads, unroll, pairs, stall, queue = (x.value for x in vals)
        assert threads in (64, 128, 256) and unroll in (2, 4, 8) and pairs >= 16 and queue in (0, 1)
    assert lib.sph_variant_shape(n, *[C.byref(x) for x in vals]) == _lib.SPH_ERR_INVALID
    assert lib.sph_variant_shape(_lib.DEFAULT_VARIANT, *[C.byref(x) for x in vals]) == _lib.SPH_OK
    assert vals[4].value == 1


def test_default_params_are_the_reference_constants(hip_host):
    lib, _lib = hip_host
    p = _lib.sph_params()
    lib.sph_default_params(C.byref(p), 0.4, 1620)
    assert list(p.box) == [np.float32(0.4)] * 3
    assert np.float32(p.gravity[1]) == np.float32(-9.80665) and p.gravity[0] == 0 and p.gravity[2] == 0
    assert np.float32(p.dt) == np.float32(0.01)
    assert p.capacity == 1620


def test_no_device_is_reported(hip_host):
    lib, _lib = hip_host
    buf = C.create_string_buffer(128)
    if lib.sph_device_name(buf, 128) == _lib.SPH_OK:
        pytest.skip("a HIP device is visible")
    p = _lib.sph_params()
    lib.sph_default_params(C.byref(p), 0.4, 100)
    ctx = C.c_void_p()
    assert lib.sph_create(C.byref(p), C.byref(ctx)) == _lib.SPH_ERR_NO_DEVICE
    assert not ctx.value


def test_dam_break_generator_matches_oracle(orc):
    import sphhip
    for box in (0.1, 0.4, 0.9):
        pos = sphhip.scenes.dam_break(box)
        o = orc.OracleSim(box)
        o.setup_scene()
        assert np.array_equal(pos.view(np.uint32), o.state()["pos"].view(np.uint32))


def test_lattice_generator_matches_independent_restatement():
    import sphhip
    for dims, box, seed in (((16, 16, 16), 0.4, 0), ((7, 3, 5), (0.3, 0.2, 0.5), 9), ((32, 16, 32), 0.8, 1)):
        a = sphhip.scenes.lattice(*dims, box, jitter_h=0.01, seed=seed)
        b = lattice_numpy(*dims, box, 0.01, seed)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), dims
    p = sphhip.scenes.lattice(16, 16, 16, 0.4, jitter_h=0.01, seed=0)
    spacing = np.float32(0.0457) / 2
    assert np.all(np.abs(p[1:, 2] - p[:-1, 2])[:15] < spacing * 1.1)


def test_workload_1m_shape():
    import sphhip
    pos, w = sphhip.scenes.workload("dam_break_1m")
    assert pos.shape == (1048576, 3) and w["box"] == 3.2
    assert sphhip.scenes.grid_res(3.2) == 71


@pytest.fixture(scope="module")
def hip_host():
    import sphhip
    return sphhip.load(), sphhip._lib

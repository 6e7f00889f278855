"""C ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, and its host-only entry points (parameter
derivation, pa.dat I/O, option defaults, error reporting) behave like the
reference's; the *.dat layout round-trips through read_data_2D's reading."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import effplast2d_amd as ep
from effplast2d_amd import _lib
from effplast2d_amd.datfile import dat_name, read_data_2D, write_dat
from oracle import ora

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ep2d_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in sorted(names):
        assert hasattr(L, n), n
    # the Python binding types every declared entry point
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_version_and_device_count():
    assert b"gfx950" in _lib.lib().ep2d_version()
    assert ep.device_count() >= 0


def test_derive_params_matches_golden_and_oracle():
    gold = json.load(open(os.path.join(GOLDEN, "pa_golden.json")))
    for name, g in gold.items():
        i = g["inputs"]
        pa, nx = ep.derive_params(Lx=i["Lx"], Ly=i["Ly"], nGrid=i["nGrid"], nPores=i["nPores"],
                                  porosity=i["porosity"], Y=i["Y"], K0=i["K0"], G0=i["G0"],
                                  rho0=i["rho0"], CFL=i["CFL"])
        assert nx == g["nX"]
        assert list(pa) == [float.fromhex(h) for h in g["pa_hex"]], name


def test_pa_file_round_trip(tmp_path):
    src = os.path.join(GOLDEN, "pa_rve1024_plastic.dat")
    pa = ep.read_pa(src)
    out = tmp_path / "pa.dat"
    ep.write_pa(out, pa)
    assert open(out, "rb").read() == open(src, "rb").read()


def test_read_pa_errors():
    with pytest.raises(ep.Ep2dError) as e:
        ep.read_pa("/nonexistent/data/pa.dat")
    assert e.value.status == -3 and "Cannot open file" in str(e.value)


def test_derive_params_rejects_bad_inputs():
    with pytest.raises(ep.Ep2dError) as e:
        ep.derive_params(nGrid=0)
    assert e.value.status == -1
    with pytest.raises(ep.Ep2dError):
        ep.derive_params(porosity=1.5)


def test_default_options_are_the_reference_macros():
    o = _lib.Options()
    _lib.lib().ep2d_default_options(C.byref(o))
    assert (o.nTasks, o.nIter, o.eIter, o.checkInterval, o.device, o.nTimeSteps) == \
        (3, 100000, 1e-10, 10000, 0, 1)


def test_create_reports_errors_without_exiting():
    """Without a GPU (or with bad arguments) creation returns an error code;
    the reference would exit(-1) from gpuAssert."""
    pa, nx = ep.derive_params(nGrid=1)
    h = C.c_void_p()
    L = _lib.lib()
    o = _lib.Options()
    L.ep2d_default_options(C.byref(o))
    rc = L.ep2d_create(pa.ctypes.data_as(C.POINTER(C.c_double)), nx, nx + 1, C.byref(o), C.byref(h))
    assert rc == -1 and b"square" in L.ep2d_last_error()
    o.nIter = 0
    rc = L.ep2d_create(pa.ctypes.data_as(C.POINTER(C.c_double)), nx, nx, C.byref(o), C.byref(h))
    assert rc == -1
    if ep.device_count() == 0:
        L.ep2d_default_options(C.byref(o))
        rc = L.ep2d_create(pa.ctypes.data_as(C.POINTER(C.c_double)), nx, nx, C.byref(o),
                           C.byref(h))
        assert rc in (-1, -2) and not h.value
    assert L.ep2d_iterate(None, 1) == -1
    L.ep2d_destroy(None)


def test_dat_file_layout_round_trip(tmp_path, monkeypatch):
    """SaveMatrix writes A[j*m + i] raw; read_data_2D.m reshapes to (m, n)
    column-major, so MATLAB's field(i+1, j+1) is A[j*m + i]."""
    m, n = 5, 3
    A = np.arange(m * n, dtype=np.float64) * 1.5 - 2.0
    monkeypatch.chdir(tmp_path)
    write_dat(dat_name("Pc", 128), A)
    F = read_data_2D("Pc", 128, m, n)
    assert F.shape == (m, n)
    for i in range(m):
        for j in range(n):
            assert F[i, j] == A[j * m + i]
    assert os.path.getsize(dat_name("Pc", 128)) == 8 * m * n
    Z = read_data_2D("missing", 128, m, n)
    assert not Z.any()


def test_python_mirror_validates_arguments():
    with pytest.raises(ValueError):
        ep.EffPlast2D(np.zeros(5), 64)


def test_oracle_is_not_imported_by_the_package():
    import sys
    pkg_dir = os.path.join(ROOT, "effplast2d_amd")
    for fn in os.listdir(pkg_dir):
        if fn.endswith(".py"):
            assert "oracle" not in open(os.path.join(pkg_dir, fn)).read(), fn
    for fn in os.listdir(os.path.join(pkg_dir, "csrc")):
        if fn.endswith((".cpp", ".hip", ".h", "Makefile")):
            assert "ep2d_oracle" not in open(os.path.join(pkg_dir, "csrc", fn)).read(), fn
    assert ora.LIB_PATH not in [getattr(m, "__file__", "") for m in sys.modules.values()]


def test_batch_options_and_errors():
    L = _lib.lib()
    o = _lib.BatchOptions()
    L.ep2d_batch_default_options(C.byref(o))
    assert o.mode == _lib.MODES["split"] and o.useGraphs == 1 and o.checkInterval > 0
    h = C.c_void_p()
    o.mode = _lib.MODES["plain"]
    assert L.ep2d_batch_create(C.byref(o), C.byref(h)) == -1 and not h.value
    assert b"mode" in L.ep2d_last_error()
    L.ep2d_batch_default_options(C.byref(o))
    o.tileRows = 7
    assert L.ep2d_batch_create(C.byref(o), C.byref(h)) == -1
    L.ep2d_batch_default_options(C.byref(o))
    o.quad2Rows = -1
    assert L.ep2d_batch_create(C.byref(o), C.byref(h)) == -1
    assert b"quad2Rows" in L.ep2d_last_error()
    L.ep2d_batch_default_options(C.byref(o))
    o.quadRows, o.triple2Rows = 64, 64
    assert L.ep2d_batch_create(C.byref(o), C.byref(h)) == -1
    assert b"at most one" in L.ep2d_last_error()
    if ep.device_count() == 0:
        L.ep2d_batch_default_options(C.byref(o))
        assert L.ep2d_batch_create(C.byref(o), C.byref(h)) in (-1, -2) and not h.value
    assert L.ep2d_batch_step(None, None, None) == -1
    assert L.ep2d_batch_submit(None, None, None) == -1
    L.ep2d_batch_destroy(None)


def test_batch_request_bytes():
    """The admission estimate covers the four displacement and four velocity
    buffers of the fused schemes (ping-pong U, double-buffered V), the mask
    byte, five cell and two vertex stress fields per RVE; FP32 velocities
    make split smaller by 4 (nX+1) nX x 4 bytes; bad arguments are errors."""
    pa = ora.derive_params(20.0, 20.0, 1024, 1, 0.01, 2.5e-3)
    f = ep.request_bytes("fp64", pa, 1024)
    s = ep.request_bytes("split", pa, 1024)
    pts, nC, nV = 1025 * 1024, 1024 * 1024, 1023 * 1023
    assert f - s == 4 * pts * 4
    floor = 4 * pts * 8 + 4 * pts * 8 + nC * (1 + 5 * 8) + 2 * nV * 8
    assert floor <= f <= floor + (1 << 20)
    assert ep.request_bytes("split", pa, 2048) > 3.9 * s
    L = _lib.lib()
    o = _lib.BatchOptions()
    L.ep2d_batch_default_options(C.byref(o))
    n = C.c_size_t()
    from effplast2d_amd.batch import make_request
    rq = make_request(pa, 4)
    assert L.ep2d_batch_request_bytes(C.byref(o), C.byref(rq), C.byref(n)) == -1
    o.mode = _lib.MODES["plain"]
    rq = make_request(pa, 64)
    assert L.ep2d_batch_request_bytes(C.byref(o), C.byref(rq), C.byref(n)) == -1
    assert L.ep2d_batch_request_bytes(None, None, None) == -1


def test_cli_usage_error():
    """Missing arguments: the usage message and exit status 1 (main.cu:6-10,
    39-41), before any device work."""
    import subprocess
    cli = os.path.join(ROOT, "effplast2d_amd", "ep2d_cli64")
    r = subprocess.run([cli, "1", "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "ERROR:  missing arguments" in r.stderr and "<time steps>" in r.stderr


def test_struct_layouts_match_the_header():
    """Every ctypes mirror of an include/ep2d.h struct has the C size."""
    from effplast2d_amd import _lib as L
    lib = L.lib()
    pairs = {"ep2d_physics": L.Physics, "ep2d_options": L.Options, "ep2d_moduli": L.Moduli,
             "ep2d_record": L.Record, "ep2d_result": L.Result, "ep2d_an_static": L.AnStatic,
             "ep2d_batch_options": L.BatchOptions, "ep2d_request": L.Request,
             "ep2d_batch_result": L.BatchResult, "ep2d_batch_stats": L.BatchStats}
    for name, cls in pairs.items():
        assert lib.ep2d_struct_size(name.encode()) == C.sizeof(cls), name
    assert lib.ep2d_struct_size(b"no_such_struct") == 0


@pytest.mark.parametrize("nX,kw,msg", [
    (23170, {}, "grid too large"),
    (16383, {"marchRows": 16, "itersPerLaunch": 3}, "under 2 GiB"),
    (64, {"marchRows": 16, "itersPerLaunch": 5}, "itersPerLaunch must be"),
    (64, {"itersPerLaunch": 3}, "marchRows > 0"),
    (64, {"marchRows": 16, "itersPerLaunch": 4, "stageWaves": 3}, "stageWaves must be"),
    (64, {"marchRows": 16, "itersPerLaunch": 4, "laneColumns": 3}, "laneColumns must be"),
])
def test_shape_and_size_limits_fail_before_the_device(nX, kw, msg):
    """Shape options and grid-size limits are checked before any device call:
    fields under 4 GiB (32-bit offsets of the fused kernels), multi-iteration
    launches only with fields under 2 GiB, 1 to 4 iterations per launch, and
    multi-iteration launches only with a segment height."""
    pa, _ = ep.derive_params(Lx=20.0, Ly=20.0, nX=nX, nPores=1, porosity=0.01)
    with pytest.raises(ep.Ep2dError, match=msg) as ei:
        ep.EffPlast2D(pa, nX, mode="split", **kw)
    assert "EP2D_ERR_ARG" in str(ei.value)


def _chain_list(sizes, ns, w, rows):
    """Work list of the chain kernels as the engine builds it: per RVE,
    bands of `rows` rows, strips of 60w+4-4ns output columns starting at
    column -2ns."""
    cols = 60 * w + 4 - 4 * ns
    segs = []
    for rve, (nx, ny) in enumerate(sizes):
        for j0 in range(0, ny, rows):
            for k in range(-(-nx // cols)):
                segs.append((rve, cols * k - 2 * ns, j0, min(j0 + rows, ny)))
    return segs


def _shares(n):
    q, r = n // 8, n % 8
    return [(x * q + min(x, r), q + (1 if x < r else 0)) for x in range(8)]


def _covered(segs):
    return {(a, s, j) for a, s, z, e in segs for j in range(z, e)}


@pytest.mark.parametrize("resident", [10 ** 6, 768])
@pytest.mark.parametrize("rows", [8, 96])
def test_chain_launch_order_covers_the_same_work(rows, resident):
    """The launch order of a chain work list (ep2d_diag_chain_order, the
    engine's own function, host only): with every XCD's share of the list
    taken as the kernel takes it (contiguous, n/8 entries and one more for
    the first n%8), each share covers exactly the rows of exactly the strips
    its share of the input list covered; every piece lies inside one input
    segment; column-edge segments come first in each share; a list of at
    most `resident` entries keeps its length, a longer one grows by the
    pieces cut from the end of each share (one-row segments listed again
    instead of cut)."""
    ns, w = 4, 2
    uw = 60 * w + 4
    sizes = [(161, 161)] * 5 + [(300, 257), (1024, 1024), (96, 130)]
    segs = _chain_list(sizes, ns, w, rows)
    n = len(segs)
    flat = np.array(segs, dtype=np.int32).ravel()
    dims = np.array(sizes, dtype=np.int32).ravel()
    m = C.c_int()
    L = _lib.lib()
    assert L.ep2d_diag_chain_order(flat.ctypes.data, n, w, dims.ctypes.data, len(sizes),
                                   resident, None, 0, C.byref(m)) == 0
    out = np.zeros(4 * m.value, dtype=np.int32)
    got = C.c_int()
    assert L.ep2d_diag_chain_order(flat.ctypes.data, n, w, dims.ctypes.data, len(sizes),
                                   resident, out.ctypes.data, m.value, C.byref(got)) == 0
    assert got.value == m.value
    res = [tuple(int(v) for v in out[4 * k:4 * k + 4]) for k in range(m.value)]
    if n <= resident:
        assert m.value == n
    else:
        assert m.value > n and (m.value - n) % 8 == 0
    by_key = {}
    for a, s, z, e in segs:
        by_key.setdefault((a, s), []).append((z, e))
    for a, s, z, e in res:
        assert z < e
        assert any(z0 <= z and e <= e0 for z0, e0 in by_key[(a, s)]), (a, s, z, e)
    for (lo, cnt), (lo2, cnt2) in zip(_shares(n), _shares(m.value)):
        src, dst = segs[lo:lo + cnt], res[lo2:lo2 + cnt2]
        assert _covered(dst) == _covered(src)
        edge = [not (s >= 1 and s + uw - 1 <= sizes[a][0] - 2) for a, s, z, e in dst]
        assert edge == sorted(edge, reverse=True)


def test_chain_launch_order_rejects_bad_arguments():
    L = _lib.lib()
    segs = np.array([0, -8, 0, 8, 3, -8, 0, 8], dtype=np.int32)
    dims = np.array([64, 64], dtype=np.int32)
    m = C.c_int()
    assert L.ep2d_diag_chain_order(segs.ctypes.data, 2, 2, dims.ctypes.data, 1, 0, None, 0,
                                   C.byref(m)) != 0          # RVE 3 of 1
    assert L.ep2d_diag_chain_order(segs.ctypes.data, 1, 3, dims.ctypes.data, 1, 0, None, 0,
                                   C.byref(m)) != 0          # three waves per stage

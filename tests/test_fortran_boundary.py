"""The Fortran side of the drop-in boundary: the nuppn_hip bind(c) module
(nuppn_modular2-ininet3_amd/fortran/nuppn_hip.F90), compiled with flang, against the C ABI.

CPU: every bind(c) type has the C struct's size and field offsets (a Fortran and a C program
print both); every entry point the header declares has a bind(c) interface of the same name
(c_fortran_dgssv_ excepted: the reference's BD step calls it with an implicit interface,
solver/source/bader_deuflhard.F90:438-464), and a program taking the address of every interface
links against libnuppn_hip.so.

GPU: the ppn-style driver (fortran/ppn_driver.F90, built by csrc/build.sh) integrates a
40-cycle one-zone trajectory through nuppn_integrate_zone and gets the Python binding's results
bit for bit, hence the oracle's decisions within 1e-10; its batch mode gets
nuppn_integrate_batch_mixed's bits.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nuppn_modular2-ininet3_amd")
LIB = os.path.join(PKG, "lib")
FC = os.environ.get("FC", "/opt/rocm/llvm/bin/flang")
DRIVER = os.path.join(LIB, "nuppn_ppn_driver")

need_fc = pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler (flang)")


@need_fc
def test_bind_c_types_match_c_structs(tmp_path):
    """c_sizeof and the field offsets of nuppn_network, nuppn_knobs, nuppn_zone_stats and
    nuppn_batch_opts, nuppn_mppnp_opts, nuppn_rate_physics, nuppn_tune_result equal the C sizeof /
    offsetof; all 32 interfaces link."""
    if not os.path.exists(os.path.join(LIB, "libnuppn_hip.so")):
        pytest.skip("libnuppn_hip.so not built")
    d = str(tmp_path)
    subprocess.run([FC, "-J", d, os.path.join(PKG, "fortran", "nuppn_hip.F90"),
                    os.path.join(ROOT, "tests", "fortran", "layout_f.F90"), "-o",
                    os.path.join(d, "layout_f"), "-L" + LIB, "-lnuppn_hip",
                    "-Wl,-rpath," + LIB], check=True, cwd=d)
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "fortran", "layout_c.c"), "-o",
                    os.path.join(d, "layout_c")], check=True)
    f = subprocess.run([os.path.join(d, "layout_f")], check=True, capture_output=True,
                       text=True).stdout.split("\n")
    c = subprocess.run([os.path.join(d, "layout_c")], check=True, capture_output=True,
                       text=True).stdout.split("\n")
    f_types = {ln.split()[0]: ln.split()[1:] for ln in f if ln.strip()}
    c_types = {ln.split()[0]: ln.split()[1:] for ln in c if ln.strip()}
    print(f_types)
    assert f_types.pop("bound_procedures") == ["32"]
    assert f_types == c_types
    assert set(c_types) == {"nuppn_network", "nuppn_knobs", "nuppn_zone_stats",
                            "nuppn_batch_opts", "nuppn_mppnp_opts", "nuppn_rate_physics",
                            "nuppn_tune_result"}


def test_module_binds_every_declared_entry_point():
    """Each function the header declares has a bind(c) interface with that C name."""
    hdr = open(os.path.join(ROOT, "include", "nuppn_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(\w+)\s*\(", hdr, flags=re.M))
    declared -= {"if", "while", "for", "sizeof"}
    src = open(os.path.join(PKG, "fortran", "nuppn_hip.F90")).read()
    bound = set(re.findall(r'bind\(c,\s*name="(\w+)"\)', src))
    print(sorted(declared))
    assert "nuppn_integrate_zone" in declared and "slu_solve" in declared
    assert declared - bound == {"c_fortran_dgssv_"}
    assert bound <= declared


def _write_net(path, net):
    with open(path, "wb") as f:
        np.array([net.nsp, net.nre, 0], np.int32).tofile(f)
        np.ascontiguousarray(net.k, np.int32).tofile(f)
        np.ascontiguousarray(net.rtype, np.int32).tofile(f)
        np.ascontiguousarray(net.coef, np.float64).tofile(f)
        np.ascontiguousarray(net.aw, np.float64).tofile(f)
        np.ascontiguousarray(net.zw, np.float64).tofile(f)


def _driver():
    if not os.path.exists(DRIVER):
        pytest.skip("the Fortran driver is not built (no flang at build time)")
    return DRIVER


@pytest.mark.gpu
def test_fortran_ppn_driver_trajectory(pkg, oracle, tmp_path):
    """40 ppn cycles of the 271-species network from Fortran: the Python binding's bits, the
    oracle's step decisions, <= 1e-10 on every X > 1e-20."""
    drv = _driver()
    w = pkg.network.config(0)
    ncyc = 40
    _write_net(tmp_path / "net.bin", w.net)
    with open(tmp_path / "traj.bin", "wb") as f:
        np.array([ncyc], np.int32).tofile(f)
        w.x0[0].astype(np.float64).tofile(f)
        for s in range(ncyc):
            np.array([w.t9[s, 0], w.t9[s, 1], w.rho[s, 0], w.rho[s, 1], w.dt[s]]).tofile(f)
    r = subprocess.run([drv, "zone", str(tmp_path / "net.bin"), str(tmp_path / "traj.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = open(tmp_path / "out.bin", "rb").read()
    rec = np.dtype([("rc", "<i4"), ("ierr", "<i4"), ("nsubt", "<i4"), ("titers", "<i4"),
                    ("n_newton", "<i8")])
    cyc = np.frombuffer(raw[:ncyc * rec.itemsize], rec)
    tail = np.frombuffer(raw[ncyc * rec.itemsize:], np.float64)
    y_f, ye_f = tail[:w.net.nsp], tail[w.net.nsp]
    # the same cycles through the Python binding, and through the oracle
    ctx = pkg.hip.Context(w.net)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    y_p = w.x0[0].copy()
    y_o = w.x0[0].copy()
    try:
        for s in range(ncyc):
            args = (w.t9[s, 0], w.t9[s, 1], w.rho[s, 0], w.rho[s, 1], w.dt[s])
            ie_p, y_p, ye_p, st_p = ctx.integrate_network(y_p, *args)
            ie_o, y_o, ye_o, st_o = o.integrate_zone(y_o, *args)
            assert cyc["rc"][s] == 0 and cyc["ierr"][s] == ie_p == ie_o == 0
            assert (cyc["nsubt"][s], cyc["n_newton"][s]) == (st_p.nsubt, st_p.n_newton) == \
                (st_o.nsubt, st_o.n_newton)
    finally:
        ctx.close()
    np.testing.assert_array_equal(y_f, y_p)
    assert ye_f == ye_p
    m = y_o > 1e-20
    assert np.max(np.abs(y_f[m] - y_o[m]) / y_o[m]) <= 1e-10


@pytest.mark.gpu
def test_fortran_driver_batch_mixed(pkg, tmp_path):
    """mppnp's batch call from Fortran (column-major yps(nsp, nz)): the Python binding's bits."""
    drv = _driver()
    w = pkg.network.config(1, nz=8)
    sel = np.array([0, 2, 5, 6])
    nz = len(sel)
    _write_net(tmp_path / "net.bin", w.net)
    args = [a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]
    with open(tmp_path / "zones.bin", "wb") as f:
        np.array([nz], np.int32).tofile(f)
        np.ascontiguousarray(w.x0[sel], np.float64).tofile(f)   # yps(nsp, nz), zone-major
        for a in args:
            np.ascontiguousarray(a, np.float64).tofile(f)
    r = subprocess.run([drv, "batch", str(tmp_path / "net.bin"), str(tmp_path / "zones.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = open(tmp_path / "out.bin", "rb").read()
    o = 0
    rc = np.frombuffer(raw, np.int32, 1, o)[0]
    o += 4
    ierr = np.frombuffer(raw, np.int32, nz, o)
    o += 4 * nz
    nsubt = np.frombuffer(raw, np.int32, nz, o)
    o += 4 * nz
    nnew = np.frombuffer(raw, np.int64, nz, o)
    o += 8 * nz
    y = np.frombuffer(raw, np.float64, nz * w.net.nsp, o).reshape(nz, w.net.nsp)
    o += 8 * nz * w.net.nsp
    ye = np.frombuffer(raw, np.float64, nz, o)
    ctx = pkg.hip.Context(w.net)
    try:
        ie_p, y_p, ye_p, st_p = ctx.integrate_batch_mixed(w.x0[sel], *args)
    finally:
        ctx.close()
    assert rc == 0
    np.testing.assert_array_equal(ierr, ie_p)
    np.testing.assert_array_equal(nsubt, st_p["nsubt"])
    np.testing.assert_array_equal(nnew, st_p["n_newton"])
    np.testing.assert_array_equal(y, y_p)
    np.testing.assert_array_equal(ye, ye_p)

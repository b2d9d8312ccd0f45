"""Output handlers and the HDF5 file layer on the CPU (no propagation).

Mirrors test/test_output.jl's "HDF5" and "Memory" testsets (synthetic saves
and statistics fed through the per-step output call), the continuation-cache
bookkeeping of HDF5Output (src/Output.jl:146-302, 462-487) and the file
layout HDF5.jl gives a Julia array (dimensions reversed, complex as {r, i}),
checked with the HDF5 library's own h5dump.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from luna_jl_amd import output as O

pytestmark = pytest.mark.skipif(not os.path.isfile(O.H5_LIB_PATH), reason="libluna_h5.so not built (no HDF5)")

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


class StubStats:
    """A statsfun whose records give one scalar and one 2-vector per step,
    like test_output.jl's statsfun(y, t, dt) = Dict("stat" => stat,
    "stat2d" => [stat, stat])."""

    def __init__(self, stat):
        self.stat = stat

    def to_dict(self, records):
        n = len(records)
        return {"stat": np.full(n, self.stat), "stat2d": np.full((n, 2), self.stat)}


def _feed(o, y0, t):
    rec = np.zeros((1, 15))
    for ii, ti in enumerate(t):
        o.on_step(ii, y0[..., None], np.array([ti]), ti, 0.0, y0, rec)


def test_hdf5_output(tmp_path):
    fpath = str(tmp_path / "sub" / "test.h5")
    rng = np.random.default_rng(1)
    shape, n = (1024, 4, 2), 11
    stat = rng.normal()
    t = O.save_grid(0, 10, n)
    omega = rng.normal(size=1024)
    o = O.HDF5Output(fpath, 0, 10, n, StubStats(stat), yname="y", tname="t")
    extra = {"ω": omega, "git_commit": O.git_commit()}
    o(extra)
    o({"meta1": 100, "meta2": "src"}, meta=True)
    y0 = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    _feed(o, y0, t)
    with pytest.warns(UserWarning):
        o(extra, force=True)
    with O.H5File(fpath) as f:
        assert np.array_equal(f.read("t"), t)
        yr = f.read("y")
        assert yr.shape == shape + (n,)
        assert all(np.array_equal(yr[..., i], y0) for i in range(n))
        assert np.array_equal(f.read("ω"), omega)
        assert f.read("git_commit") == O.git_commit() == f.read("meta/git_commit")
        assert np.all(f.read("stats/stat") == stat)
        s2 = f.read("stats/stat2d")
        assert s2.shape == (2, n) and np.all(s2 == stat)
        assert f.read("meta/sourcecode") == O.sourcecode()
        assert f.read("meta/meta1") == 100 and f.read("meta/meta2") == "src"
        assert f.read("meta/cache/saved") == n and f.read("meta/cache/t") == t[-1]
    assert np.all(o["stats", "stat"] == stat)
    assert np.array_equal(o["y"], yr)
    assert o["meta"]["meta1"] == 100
    assert np.array_equal(o["y", slice(None), 0, 1, 3], y0[:, 0, 1])


def test_hdf5_layout_matches_hdf5jl(tmp_path):
    """HDF5.jl stores a Julia array of size (a, b, c) as an HDF5 dataset of
    dims (c, b, a) and ComplexF64 as the compound {r, i}; Output.jl makes the
    solution chunked by one z-slice and unlimited along z."""
    if not os.path.isfile(H5DUMP):
        pytest.skip("h5dump not available")
    fpath = str(tmp_path / "layout.h5")
    o = O.HDF5Output(fpath, 0, 1, 5)
    y0 = np.arange(12, dtype=float).reshape(3, 4) + 0j
    o.on_step(0, y0[..., None], np.array([0.0]), 0.1, 0.01, y0, np.zeros((0, 15)))
    out = subprocess.run([H5DUMP, "-H", "-p", "-d", "/Eω", fpath], capture_output=True, text=True).stdout
    assert "( 5, 4, 3 ) / ( H5S_UNLIMITED, 4, 3 )" in out
    assert 'H5T_IEEE_F64LE "r"' in out and 'H5T_IEEE_F64LE "i"' in out
    assert "CHUNKED ( 1, 4, 3 )" in out
    out = subprocess.run([H5DUMP, "-d", "/Eω", "-s", "0,2,1", "-c", "1,1,1", fpath], capture_output=True,
                         text=True).stdout
    # y0[1, 2] = 6 (Julia's y0[2, 3]) sits at file index [z][j][i] = [0][2][1]
    assert "(0,2,1):{6,0}" in "".join(out.split())
    z = subprocess.run([H5DUMP, "-H", "-d", "/meta/cache/t", fpath], capture_output=True, text=True).stdout
    assert "SCALAR" in z


def test_memory_output():
    rng = np.random.default_rng(2)
    shape, n = (1024, 4, 2), 11
    stat = rng.normal()
    t = O.save_grid(0, 10, n)
    o = O.MemoryOutput(0, 10, n, StubStats(stat), yname="y", tname="t")
    extra = {"ω": rng.normal(size=1024), "git_commit": O.git_commit()}
    o(extra)
    o({"meta1": 100, "meta2": "src"}, meta=True)
    y0 = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    _feed(o, y0, t)
    assert np.array_equal(o.data["t"], t)
    assert all(np.array_equal(o.data["y"][..., i], y0) for i in range(n))
    with pytest.raises(KeyError):
        o(extra)
    with pytest.warns(UserWarning):
        o(extra, force=True)
    assert np.all(o.data["stats"]["stat"] == stat)
    assert o.data["stats"]["stat2d"].shape == (2, n)
    assert np.all(o["stats"]["stat2d"] == stat)
    assert o.data["meta"]["git_commit"] == O.git_commit()
    assert o.data["meta"]["sourcecode"] == O.sourcecode()
    assert o.data["meta"]["meta1"] == 100 and o.data["meta"]["meta2"] == "src"
    assert o["y", 0, 1, 1, 3] == o.data["y"][0, 1, 1, 3]


def test_cache_bookkeeping(tmp_path):
    fpath = str(tmp_path / "cache.h5")
    rng = np.random.default_rng(3)
    y = rng.normal(size=64) + 1j * rng.normal(size=64)
    o = O.HDF5Output(fpath, 0, 1.0, 11, StubStats(0.5))
    # before any save the cache has no t: check_cache hands back its arguments
    assert o.check_cache(y, 0.0, 1e-4)[1:] == (0.0, 1e-4)
    zs = O.save_grid(0, 1.0, 11)
    for k in range(4):
        o.on_step(k, (y * (k + 1))[:, None], zs[k:k + 1], zs[k] + 0.01, 0.02 * (k + 1), y * (k + 1),
                  np.zeros((1, 15)))
    y2, t2, dt2 = o.check_cache(y, 0.0, 1e-4)
    assert t2 == zs[3] + 0.01 and dt2 == 0.08 and np.array_equal(y2, 4 * y)
    # a cache behind the requested start is ignored (src/Output.jl:478-480)
    assert o.check_cache(y, 0.9, 1e-4)[1:] == (0.9, 1e-4)
    # reopening the file continues its saves and keeps its hash
    o2 = O.HDF5Output(fpath, 0, 1.0, 11, StubStats(0.5))
    assert o2.saved == 4 and o2.cachehash == o.cachehash
    o2.on_step(4, (5 * y)[:, None], zs[4:5], zs[4] + 0.01, 0.1, 5 * y, np.zeros((1, 15)))
    assert np.array_equal(o2["Eω"][:, 4], 5 * y) and np.array_equal(o2["Eω"][:, 0], y)
    assert len(o2["stats"]["stat"]) == 5 and o2["meta"]["cache"]["saved"] == 5
    # a different propagation (other statistics) may not continue this cache
    o3 = O.HDF5Output(fpath, 0, 1.0, 11, None)
    with pytest.raises(RuntimeError, match="hash"):
        o3.on_step(5, (6 * y)[:, None], zs[5:6], zs[5] + 0.01, 0.1, 6 * y, np.zeros((0, 15)))
    # a file without a cache group refuses a cached output; read-only refuses writes
    nc = str(tmp_path / "nocache.h5")
    O.HDF5Output(nc, 0, 1.0, 11, cache=False)
    with pytest.raises(RuntimeError, match="no cache"):
        O.HDF5Output(nc, 0, 1.0, 11)
    ro = O.HDF5Output.open(fpath)
    with pytest.raises(RuntimeError, match="read-only"):
        ro("x", 1.0)
    with pytest.raises(FileNotFoundError):
        O.HDF5Output.open(str(tmp_path / "missing.h5"))


def test_h5_round_trip(tmp_path):
    fpath = str(tmp_path / "types.h5")
    rng = np.random.default_rng(4)
    vals = {
        "f": 1.25, "i": -7, "u": np.uint64(2 ** 63 + 5), "c": 1 - 2j, "b": np.array([True, False, True]),
        "s": "ωλ text", "sl": ["a", "bb", ""], "a1": rng.normal(size=5),
        "a3": np.asfortranarray(rng.normal(size=(3, 4, 2))), "c2": rng.normal(size=(4, 3)) + 1j,
        "ints": np.arange(6, dtype=np.int64).reshape(2, 3), "g": {"x": 2.0, "h": {"y": "deep"}},
    }
    with O.H5File(fpath, "w") as f:
        for k, v in vals.items():
            f.write(k, v)
    with O.H5File(fpath, "r") as f:
        assert f.read("f") == 1.25 and f.read("i") == -7 and f.read("u") == np.uint64(2 ** 63 + 5)
        assert f.read("c") == 1 - 2j
        assert np.array_equal(f.read("b").astype(bool), vals["b"])
        assert f.read("s") == "ωλ text" and f.read("sl") == ["a", "bb", ""]
        for k in ("a1", "a3", "c2", "ints"):
            assert np.array_equal(f.read(k), vals[k]), k
        assert f.read("g") == {"x": 2.0, "h": {"y": "deep"}}
        assert sorted(f.keys()) == sorted(vals)
    with O.H5File(fpath, "r+") as f:
        f.write("a1", np.zeros(5))        # same shape: in place
        f.write("a3", np.ones(2))         # other shape: replaced
        f.delete("g/h")
        assert np.array_equal(f.read("a1"), np.zeros(5)) and np.array_equal(f.read("a3"), np.ones(2))
        assert "g/h" not in f and "g/x" in f and "nope/x" not in f


def test_scan_points_and_runscan(tmp_path):
    from luna_jl_amd import scans
    scan = scans.Scan("pscan", energy=[1.0, 2.0, 3.0], pressure=[10.0, 20.0])
    assert len(scan) == 6 and scan.size() == (3, 2)
    # Iterators.product: the first variable varies fastest
    assert scan.points() == [(1.0, 10.0), (2.0, 10.0), (3.0, 10.0), (1.0, 20.0), (2.0, 20.0), (3.0, 20.0)]
    assert scan.getvalue("pressure", 4) == 20.0 and scan.getvalue("energy", 5) == 2.0
    assert scans.makefilename(scan, 5) == "pscan_00005.h5"
    assert scans.chunks(list(range(1, 11)), 3) == [[1, 4, 7, 10], [2, 5, 8], [3, 6, 9]]
    seen = []

    def f(scanidx, e, p):
        seen.append(scanidx)
        if scanidx == 3:
            raise ValueError("bad point")
        return e * p

    with pytest.warns(RuntimeWarning, match="Error at scanidx 3"):
        out = scans.runscan(f, scan)
    assert seen == [1, 2, 3, 4, 5, 6] and out == [10.0, 20.0, 20.0, 40.0, 60.0]
    scan.exec = scans.RangeExec(range(2, 5))
    assert [i for i, _ in scans._share(scan)] == [1, 2, 3]  # the reference numbers the sub-range from 1
    assert [a for _, a in scans._share(scan)] == scan.points()[1:4]
    scan.exec = scans.BatchExec(3, 2)
    assert [i for i, _ in scans._share(scan)] == [2, 5]
    scan.addvariable("extra", [0.5])
    assert len(scan) == 6 and scan.values(6) == (3.0, 20.0, 0.5)
    # savescan writes the scan's description into the file's meta group
    o = O.ScanHDF5Output(scan, 4, 0, 1.0, 3, fdir=str(tmp_path / "scan"))
    assert os.path.basename(o.fpath) == "pscan_00004.h5"
    with O.H5File(o.fpath) as f:
        assert f.read("meta/scanidx") == 4
        assert f.read("meta/scanvars/pressure") == 20.0 and f.read("meta/scanvars/energy") == 1.0
        assert np.array_equal(f.read("meta/scanarrays/energy"), [1.0, 2.0, 3.0])
        assert f.read("meta/scanorder") == ["energy", "pressure", "extra"]
        assert np.array_equal(f.read("meta/scanshape"), [3, 2, 1])

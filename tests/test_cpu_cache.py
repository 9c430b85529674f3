"""The cache-file reading the `vfRay <volume>` command shares with the
reference (main.cc:242-342): .con built from a .incid file by the reference's
buildCon (vfRayPreComp.cc:424-501) over the lists the file holds, .incid
reading (:153-212), and the truncated .off (main.cc:257-262).  The fixtures
were written by the compiled reference (tests/golden/make_golden.py)."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import vfray_amd as V
from vfray_amd import workloads as WL
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "vfray_amd", "vfRay")


@pytest.fixture(scope="module")
def small():
    return dict(np.load(os.path.join(GOLD, "ref_small.npz")))


def test_con_from_canonical_incid_is_the_reference_con(small):
    """The .incid the reference writes for the 8^3 mesh, read back and run
    through buildCon, gives the reference's .con body and external-face count
    (and so the oracle's sort-and-match connectivity)."""
    x, t, _, _ = WL.plain_cpu(8)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "s8.incid")
        open(p, "wb").write(bytes(small["incid"]))
        off, ids = V.read_incid(p, x.shape[0])
    assert off[-1] == ids.size and np.all(np.diff(off) > 0)
    con, n_ext = V.con_from_incid(t, x.shape[0], off, ids)
    txt = bytes(small["con"]).decode().split()
    assert n_ext == int(txt[0]) and int(txt[1]) == t.shape[0]
    assert np.array_equal(con, np.array([int(v) for v in txt[2:]], np.uint32).reshape(-1, 4))
    assert np.array_equal(con, O.connectivity(t, x.shape[0])[0])


def test_con_from_out_of_order_incid_matches_reference():
    """A non-manifold mesh (a face shared by three tets) with a .incid whose
    tet lists are in descending order for three vertices: the reference's
    buildCon gives up on a list at the first id above its candidate, so it
    misses neighbours; the host buildCon reproduces the .con the reference
    wrote, entry for entry, and its count of external faces."""
    nm = dict(np.load(os.path.join(GOLD, "ref_cache.npz")))
    x, t = nm["nm_xyzs"], nm["nm_tets"]
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "nm.incid")
        open(p, "wb").write(bytes(nm["nm_incid"]))
        off, ids = V.read_incid(p, x.shape[0])
    con, n_ext = V.con_from_incid(t, x.shape[0], off, ids)
    assert np.array_equal(con, nm["nm_con"]) and n_ext == int(nm["nm_n_ext"][0])
    # with the lists in ascending order the non-manifold face goes to the
    # smallest other tet, as the oracle's (brute-force pinned) connectivity
    order = [np.sort(ids[off[i]:off[i + 1]]) for i in range(x.shape[0])]
    con2, _ = V.con_from_incid(t, x.shape[0], off, np.concatenate(order).astype(np.uint32))
    assert np.array_equal(con2, O.connectivity(t, x.shape[0])[0])
    assert not np.array_equal(con2, con)


def test_read_incid_rejects_bad_files(small):
    """readIncid's failures: a vertex count other than the mesh's, and a file
    cut short (both make the reference rebuild the .incid)."""
    x, _, _, _ = WL.plain_cpu(8)
    b = bytes(small["incid"])
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "a.incid")
        open(p, "wb").write(b)
        with pytest.raises(V.VfrError, match="vertices"):
            V.read_incid(p, x.shape[0] + 1)
        open(p, "wb").write(b[: len(b) // 2])
        with pytest.raises(V.VfrError):
            V.read_incid(p, x.shape[0])
        with pytest.raises(V.VfrError):
            V.read_incid(os.path.join(td, "missing.incid"), x.shape[0])


def test_cli_truncated_off(small):
    """An .off cut in the middle of its tetrahedra: the reference's message
    sequence and exit status 1, no cache file touched."""
    import sys
    sys.path.insert(0, GOLD)
    from cache_states import make_state, cache_lines
    ref = json.load(open(os.path.join(GOLD, "ref_cache.json")))["off_truncated"]
    x, t, _, _ = WL.plain_cpu(8)
    with tempfile.TemporaryDirectory() as td:
        V.write_off(os.path.join(td, "s8.off"), x, t)
        d = os.path.join(td, "tet_offs")
        make_state("off_truncated", d, small, open(os.path.join(td, "s8.off"), "rb").read())
        before = {e: open(os.path.join(d, "s8." + e), "rb").read() for e in ("con", "tf", "extf", "lmt", "incid")}
        r = subprocess.run([CLI, "--dir", d, "s8"], cwd=td, capture_output=True, text=True, timeout=120)
        assert r.returncode == ref["rc"] == 1
        assert cache_lines(r.stdout) == ref["lines"]
        for e, b in before.items():
            assert open(os.path.join(d, "s8." + e), "rb").read() == b

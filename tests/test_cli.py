"""The `vfRay <volume>` command (vfray_amd/vfRay): the reference's drop-in
entry point, main.cc:242-380."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import vfray_amd as V
from vfray_amd import workloads as WL
from helpers import legacy_tf_text
import hashlib
import json

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "vfray_amd", "vfRay")
GOLD = os.path.join(ROOT, "tests", "golden", "ref_small.npz")


def _run(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def test_cli_usage_error():
    """No volume name: usage text and exit status 1 (main.cc:245-248)."""
    with tempfile.TemporaryDirectory() as td:
        r = _run([], td)
        assert r.returncode == 1 and "use:" in r.stdout


def test_cli_missing_off():
    """Unreadable .off: message and exit status 1 (main.cc:257-260)."""
    with tempfile.TemporaryDirectory() as td:
        r = _run(["--dir", td, "nothing"], td)
        assert r.returncode == 1 and "Cannot read" in r.stdout


@pytest.mark.gpu
def test_cli_first_run_writes_reference_caches():
    """First run on the 8^3 mesh with no cache files: .incid, .con, .tf,
    .extf and .lmt come out byte-identical to the files the reference CLI
    wrote for the same .off, and the 512^2 image matches the reference's
    (within 1 per channel where the reference's psi table departs from the
    closed form).  A second run reads the caches and leaves them unchanged."""
    gold = dict(np.load(GOLD))
    with tempfile.TemporaryDirectory() as td:
        d = os.path.join(td, "tet_offs")
        os.makedirs(d)
        x, t, _, _ = WL.plain_cpu(8)
        V.write_off(os.path.join(d, "s8.off"), x, t)
        print("CLI first run s8", flush=True)
        r = _run(["--dir", d, "s8"], td)
        assert r.returncode == 0, r.stdout + r.stderr
        for ext, key in (("incid", "incid"), ("con", "con"), ("tf", "tf_default"), ("extf", "extf"),
                         ("lmt", "lmt")):
            got = np.frombuffer(open(os.path.join(d, "s8." + ext), "rb").read(), np.uint8)
            assert np.array_equal(got, gold[key]), ext
        ppm = np.frombuffer(open(os.path.join(td, "s8_gpu_img.ppm"), "rb").read(), np.uint8)
        ref = gold["ppm512_fast"]
        assert ppm.size == ref.size and np.array_equal(ppm[:15], ref[:15])
        # the golden image comes from the reference's -ffast-math build: a
        # handful of pixels (7 of 262,144) take another face or column break
        diff = np.abs(ppm[15:].astype(int) - ref[15:].astype(int)).reshape(-1, 3)
        assert (diff > 1).any(-1).mean() < 1e-4 and (diff > 0).any(-1).mean() < 0.01
        before = {e: open(os.path.join(d, "s8." + e), "rb").read() for e in ("con", "tf", "extf", "lmt")}
        print("CLI second run s8", flush=True)
        r2 = _run(["--dir", d, "s8"], td)
        assert r2.returncode == 0, r2.stdout + r2.stderr
        assert "not found" not in r2.stdout
        for e, b in before.items():
            assert open(os.path.join(d, "s8." + e), "rb").read() == b, e


@pytest.mark.gpu
def test_cli_image_matches_strict_reference():
    """With the seeded TF in s8.tf and a 128x128 image, the command's output
    equals the reference's strict-IEEE image except where the reference's psi
    table departs from the closed form (within 1)."""
    gold = dict(np.load(GOLD))
    with tempfile.TemporaryDirectory() as td:
        d = os.path.join(td, "tet_offs")
        os.makedirs(d)
        x, t, tf, _ = WL.plain_cpu(8)
        V.write_off(os.path.join(d, "s8.off"), x, t)
        V.write_tf(os.path.join(d, "s8.tf"), tf)
        print("CLI s8 128x128", flush=True)
        r = _run(["--dir", d, "--size", "128", "128", "s8"], td)
        assert r.returncode == 0, r.stdout + r.stderr
        ppm = np.frombuffer(open(os.path.join(td, "s8_gpu_img.ppm"), "rb").read(), np.uint8)
        assert bytes(ppm[:15]) == b"P6\n128 128\n255\n"
        img = ppm[15:].reshape(128, 128, 3).astype(int)
        diff = np.abs(img - gold["rgb128_strict"].astype(int))
        assert diff.max() <= 1 and (diff > 0).any(-1).mean() < 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("composite", [("linear", 1), ("linear-opacity", 2)])
def test_cli_linear_compositing(composite):
    """--composite linear|linear-opacity (the reference's USEPARTIALPREINT 0
    build switch): the image equals the oracle's linear render byte for
    byte."""
    import oracle as O
    name, comp = composite
    with tempfile.TemporaryDirectory() as td:
        d = os.path.join(td, "tet_offs")
        os.makedirs(d)
        x, t, tf, _ = WL.plain_cpu(8)
        V.write_off(os.path.join(d, "s8.off"), x, t)
        V.write_tf(os.path.join(d, "s8.tf"), tf)
        print(f"CLI s8 128x128 --composite {name}", flush=True)
        r = _run(["--dir", d, "--size", "128", "128", "--composite", name, "s8"], td)
        assert r.returncode == 0, r.stdout + r.stderr
        ppm = np.frombuffer(open(os.path.join(td, "s8_gpu_img.ppm"), "rb").read(), np.uint8)
        xn = O.normalize(x)
        con, _ = O.connectivity(t, x.shape[0])
        tfr = V.read_tf(os.path.join(d, "s8.tf"))
        orc = O.render_ref(xn, t, con, O.extf(con), tfr, V.psi_generate(), O.lmt(xn, t), 128, 128,
                           composite=comp)
        assert np.array_equal(ppm[15:].reshape(128, 128, 3), orc["rgb8"])
        assert orc["rgb8"].max() > 0


@pytest.mark.gpu
def test_cli_rotate():
    """--rotate 30 20 twice (two presses of the viewer's space key): the
    image has the compiled reference's owner faces (via the oracle) and is
    within 1 of its strict image where the psi tables differ, equal to the
    oracle's."""
    import oracle as O
    rot = dict(np.load(os.path.join(ROOT, "tests", "golden", "ref_rot.npz")))
    with tempfile.TemporaryDirectory() as td:
        d = os.path.join(td, "tet_offs")
        os.makedirs(d)
        x, t, tf, _ = WL.plain_cpu(8)
        V.write_off(os.path.join(d, "s8.off"), x, t)
        V.write_tf(os.path.join(d, "s8.tf"), tf)
        print("CLI s8 128x128 --rotate 30 20 x2", flush=True)
        args = ["--dir", d, "--size", "128", "128"]
        for a, b in rot["angles"]:
            args += ["--rotate", f"{a:g}", f"{b:g}"]
        r = _run(args + ["s8"], td)
        assert r.returncode == 0, r.stdout + r.stderr
        img = np.frombuffer(open(os.path.join(td, "s8_gpu_img.ppm"), "rb").read(), np.uint8)[15:]
        img = img.reshape(128, 128, 3)
        diff = np.abs(img.astype(int) - rot["rgb128_strict"].astype(int))
        assert diff.max() <= 1 and (diff > 0).any(-1).mean() < 0.01
        xn = O.normalize(x)
        con, _ = O.connectivity(t, x.shape[0])
        orc = O.render_ref(rot["verts"], t, con, O.extf(con), tf, V.psi_generate(), O.lmt(xn, t), 128, 128)
        assert np.array_equal(img, orc["rgb8"])


CLEANTF = os.path.join(ROOT, "vfray_amd", "cleanTF")


def test_cleantf_matches_reference():
    """cleanTF (cleanTF.cc:14-78): the legacy file's 256 entries come out as
    the .tf text byte for byte as the reference's converter writes them (sha256
    recorded from it), read back by vfr_read_tf."""
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "cleantf.json")))
    with tempfile.TemporaryDirectory() as td:
        src, dst = os.path.join(td, "old.off.tf"), os.path.join(td, "new.tf")
        open(src, "w").write(legacy_tf_text())
        r = subprocess.run([CLEANTF, src, dst], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        out = open(dst, "rb").read()
        assert hashlib.sha256(out).hexdigest() == gold["sha256"]
        ref = os.path.join(ROOT, "oracle", "_ref", "cleanTF")
        if os.path.exists(ref):
            dref = os.path.join(td, "ref.tf")
            subprocess.run([ref, src, dref], capture_output=True, timeout=60, check=True)
            assert open(dref, "rb").read() == out


def test_cleantf_errors():
    """Usage error: exit 1; unreadable input: a message, exit 0 and no output,
    as the reference's converter."""
    with tempfile.TemporaryDirectory() as td:
        r = subprocess.run([CLEANTF], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "use:" in r.stderr
        dst = os.path.join(td, "new.tf")
        r = subprocess.run([CLEANTF, os.path.join(td, "missing.tf"), dst], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 0 and not os.path.exists(dst)
        # a header line of 200 characters stops the reader; 199 do not
        for n, ok in ((199, True), (200, False)):
            lines = legacy_tf_text().split("\n")
            lines[2] = "x" * n
            src = os.path.join(td, f"h{n}.tf")
            open(src, "w").write("\n".join(lines))
            out = os.path.join(td, f"h{n}.out")
            r = subprocess.run([CLEANTF, src, out], capture_output=True, timeout=60)
            assert r.returncode == 0 and os.path.exists(out) == ok
            ref = os.path.join(ROOT, "oracle", "_ref", "cleanTF")
            if os.path.exists(ref):
                rout = os.path.join(td, f"h{n}.ref")
                subprocess.run([ref, src, rout], capture_output=True, timeout=60)
                assert os.path.exists(rout) == ok
                if ok:
                    assert open(rout, "rb").read() == open(out, "rb").read()


CACHE_STATES = ("con_count_incid", "con_count_no_incid", "incid_count", "tf_255", "extf_count", "extf_stale")


@pytest.mark.gpu
@pytest.mark.parametrize("state", CACHE_STATES)
def test_cli_cache_states_match_reference(state):
    """Partial and damaged cache files (tests/golden/cache_states.py): started
    from the same state as the reference's CLI, the command prints the same
    cache messages in the same order and leaves every cache file byte for byte
    as the reference does: .con rebuilt from a readable .incid (count header
    off by one), .incid rebuilt when unreadable, the default TF written over a
    255-entry .tf, .extf rebuilt when its count is off, and a reversed .extf
    with the right count kept after a .con rebuild (its bytes are what tell
    a kept file from a rebuilt one).  For the default-TF and reversed-.extf
    states the 128x128 image matches the reference's strict-IEEE image within
    1 where the psi tables differ."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from cache_states import make_state, cache_lines
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_cache.json")))[state]
    imgs = dict(np.load(os.path.join(ROOT, "tests", "golden", "ref_cache.npz")))
    gold = dict(np.load(GOLD))
    x, t, _, _ = WL.plain_cpu(8)
    with tempfile.TemporaryDirectory() as td:
        V.write_off(os.path.join(td, "s8.off"), x, t)
        off = open(os.path.join(td, "s8.off"), "rb").read()
        d = os.path.join(td, "tet_offs")
        make_state(state, d, gold, off)
        print(f"CLI cache state {state}", flush=True)
        r = _run(["--dir", d, "s8"], td)
        assert r.returncode == ref["rc"] == 0, r.stdout + r.stderr
        assert cache_lines(r.stdout) == ref["lines"]
        for ext, h in ref["files"].items():
            assert hashlib.sha256(open(os.path.join(d, "s8." + ext), "rb").read()).hexdigest() == h, ext
        if state + "_rgb128" in imgs:
            d2 = os.path.join(td, "again")
            make_state(state, d2, gold, off)
            r = _run(["--dir", d2, "--size", "128", "128", "s8"], td)
            assert r.returncode == 0, r.stdout + r.stderr
            img = np.frombuffer(open(os.path.join(td, "s8_gpu_img.ppm"), "rb").read(), np.uint8)[15:]
            diff = np.abs(img.reshape(128, 128, 3).astype(int) - imgs[state + "_rgb128"].astype(int))
            assert diff.max() <= 1 and (diff > 0).any(-1).mean() < 0.01

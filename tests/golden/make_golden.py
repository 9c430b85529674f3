"""Writes the golden fixtures under tests/golden/ from the reference CPU path
compiled by oracle/Makefile (oracle/_ref) and from the reference's psi table.
Run in a container that has the reference tree; the fixtures are data only.

  ref_small.npz   8^3 grid (3,072 tets) seeded like the 512^2 workload:
                  the reference's .con/.extf/.lmt/.tf/.incid bytes, per-pixel
                  owner faces and uchar image at 128x128 (strict-IEEE build),
                  and the image of the -ffast-math build.
  ref_c1.json     the 16^3 / 512x512 workload: sha256 of the reference's
                  owner-face array and of its images (strict and fast builds),
                  numVisFaces, counts.
  psi_ref.json    where the reference's 512x512 psi table departs from the
                  closed form by more than 1e-6 (index pairs and values), the
                  largest departure elsewhere, sha256 of the table.
  cleantf.json    sha256 of what the reference's cleanTF writes for the
                  seeded legacy TF file of tests/helpers.legacy_tf_text().
  ref_rot.npz     8^3 mesh with the seeded TF after two presses of the CUDA
                  viewer's space key at angles (30, 20) (rotateVolume +
                  normalizeVertices, limits kept): the vertices, per-pixel
                  owner faces and the 128x128 strict-IEEE image.
  ref_cache.json  for each partial or damaged cache state of cache_states.py:
  ref_cache.npz   the reference CLI's exit status, its cache-related messages
                  and the sha256 of every cache file after the run (the
                  -ffast-math build), and for the states that change the image the
                  128x128 strict-IEEE image and owner faces; plus a
                  non-manifold mesh (a face shared by three tets) with a
                  .incid whose lists are out of order, and the .con the
                  reference builds from it.

  python3 tests/golden/make_golden.py [cleantf|rotated|caches]   (all, or one)
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
import vfray_amd as V  # noqa: E402
from vfray_amd import workloads as WL  # noqa: E402
from helpers import legacy_tf_text  # noqa: E402

REF = "/root/reference"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def read_psi_header(path: str) -> np.ndarray:
    s = open(path).read()
    body = s[s.index("{{"):]
    vals = re.findall(r"[-+]?\d*\.\d*(?:[eE][-+]?\d+)?|\d+", body)
    a = np.array([float(v) for v in vals], np.float32)
    assert a.size == 512 * 512
    return a.reshape(512, 512)


def dataset(td: str, name: str, n: int):
    x, t, tf, _ = WL.plain_cpu(n)
    os.makedirs(os.path.join(td, "tet_offs"), exist_ok=True)
    V.write_off(os.path.join(td, "tet_offs", name + ".off"), x, t)
    return x, t, tf


def cleantf() -> None:
    exe = os.path.join(ROOT, "oracle", "_ref", "cleanTF")
    assert os.path.exists(exe), "build oracle/_ref first"
    with tempfile.TemporaryDirectory() as td:
        src, dst = os.path.join(td, "old.off.tf"), os.path.join(td, "new.tf")
        open(src, "w").write(legacy_tf_text())
        import subprocess
        subprocess.run([exe, src, dst], check=True, capture_output=True)
        b = open(dst, "rb").read()
    json.dump({"sha256": hashlib.sha256(b).hexdigest(), "bytes": len(b)},
              open(os.path.join(HERE, "cleantf.json"), "w"), indent=1)


ROT_ANGLES = [(30.0, 20.0), (30.0, 20.0)]


def rotated() -> None:
    assert O.ref_available("strict"), "build oracle/_ref first"
    mv = np.stack([V.gl_rotation(a, b) for a, b in ROT_ANGLES])
    with tempfile.TemporaryDirectory() as td:
        x, t, tf = dataset(td, "s8", 8)
        V.write_tf(os.path.join(td, "tet_offs", "s8.tf"), tf)
        h = O.run_ref_harness("strict", os.path.join(td, "tet_offs", "s8"), 128, 128, rotations=mv)
    np.savez_compressed(os.path.join(HERE, "ref_rot.npz"), angles=np.array(ROT_ANGLES, np.float32), mv=mv,
                        verts=h["verts"], owner128=h["entry_ext"], rgb128_strict=h["rgb8"],
                        n_vis=np.array([h["info"]["numVisFaces"]], np.uint32))


IMAGE_STATES = ("tf_255", "extf_stale")


def nonmanifold_mesh():
    """Five tets: three share the face (0, 1, 2) (a non-manifold mesh), one
    more on each side elsewhere; returns (xyzs, tets, incid text with some
    tet lists in descending order)."""
    x = np.array([[0, 0, 0, 0.1], [1, 0, 0, 0.3], [0, 1, 0, 0.5], [0.3, 0.3, 1, 0.7], [0.3, 0.3, -1, 0.9],
                  [0.2, 0.2, 0.6, 0.2], [1, 1, 0.5, 0.4], [-0.5, 0.2, 0.5, 0.6]], np.float32)
    t = np.array([[0, 1, 2, 3], [0, 2, 1, 4], [1, 0, 2, 5], [1, 2, 3, 6], [2, 0, 3, 7]], np.uint32)
    nv = x.shape[0]
    tl = [[j for j in range(t.shape[0]) if i in t[j]] for i in range(nv)]
    vl = [sorted({int(v) for j in tl[i] for v in t[j] if v != i}) for i in range(nv)]
    lines = [str(nv)]
    for i in range(nv):
        ts = tl[i][::-1] if i in (0, 2, 3) else tl[i]
        lines.append(" ".join([str(len(ts))] + [str(j) for j in ts]))
        lines.append(" ".join([str(len(vl[i]))] + [str(v) for v in vl[i]]))
    return x, t, "\n".join(lines) + "\n"


def caches() -> None:
    import subprocess
    import shutil
    from cache_states import STATES, CACHES, make_state, cache_lines
    assert O.ref_available("strict"), "build oracle/_ref first"
    canonical = dict(np.load(os.path.join(HERE, "ref_small.npz")))
    # the reference's own build flags (-O3 -ffast-math): with strict IEEE its
    # 512x512 walk meets a ray that circles an edge on this mesh and never ends
    exe = os.path.join(ROOT, "oracle", "_ref", "fast", "vfRayCPU")
    out, arrays = {}, {}
    with tempfile.TemporaryDirectory() as td:
        x, t, _, _ = WL.plain_cpu(8)
        V.write_off(os.path.join(td, "s8.off"), x, t)
        off = open(os.path.join(td, "s8.off"), "rb").read()
        for st in STATES:
            wd = os.path.join(td, st)
            make_state(st, os.path.join(wd, "tet_offs"), canonical, off)
            if st in IMAGE_STATES:
                shutil.copytree(os.path.join(wd, "tet_offs"), os.path.join(wd, "h"))
                h = O.run_ref_harness("strict", os.path.join(wd, "h", "s8"), 128, 128)
                arrays[st + "_rgb128"] = h["rgb8"]
                arrays[st + "_owner128"] = h["entry_ext"]
            os.makedirs(os.path.join(wd, "run"))
            r = subprocess.run([exe, "s8"], cwd=os.path.join(wd, "run"), capture_output=True, text=True,
                               timeout=600)
            files = {}
            for ext in CACHES:
                fp = os.path.join(wd, "tet_offs", "s8." + ext)
                if os.path.exists(fp):
                    files[ext] = hashlib.sha256(open(fp, "rb").read()).hexdigest()
            out[st] = {"rc": r.returncode, "lines": cache_lines(r.stdout), "files": files}
        # buildCon over out-of-order lists of a non-manifold mesh
        xm, tm, incid = nonmanifold_mesh()
        wd = os.path.join(td, "nm")
        os.makedirs(os.path.join(wd, "tet_offs"))
        os.makedirs(os.path.join(wd, "run"))
        V.write_off(os.path.join(wd, "tet_offs", "nm.off"), xm, tm)
        open(os.path.join(wd, "tet_offs", "nm.incid"), "w").write(incid)
        subprocess.run([exe, "nm"], cwd=os.path.join(wd, "run"), capture_output=True, text=True, timeout=600)
        con_txt = open(os.path.join(wd, "tet_offs", "nm.con")).read().split()
        arrays["nm_xyzs"] = xm
        arrays["nm_tets"] = tm
        arrays["nm_incid"] = np.frombuffer(incid.encode(), np.uint8)
        arrays["nm_con"] = np.array([int(v) for v in con_txt[2:]], np.uint32).reshape(-1, 4)
        arrays["nm_n_ext"] = np.array([int(con_txt[0])], np.uint32)
    json.dump(out, open(os.path.join(HERE, "ref_cache.json"), "w"), indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_cache.npz"), **arrays)


def main() -> None:
    if sys.argv[1:] == ["cleantf"]:
        cleantf()
        return
    if sys.argv[1:] == ["rotated"]:
        rotated()
        return
    if sys.argv[1:] == ["caches"]:
        caches()
        return
    cleantf()
    rotated()
    assert O.ref_available("strict") and O.ref_available("fast"), "build oracle/_ref first"
    out = {}
    with tempfile.TemporaryDirectory() as td:
        # 8^3: reference CLI (default TF, caches written), then harness on a fresh copy
        x, t, tf = dataset(td, "s8", 8)
        run = O.run_ref_cli("fast", td, "s8")
        files = {}
        for ext in ("con", "extf", "lmt", "tf", "incid"):
            files[ext] = np.frombuffer(open(os.path.join(td, "tet_offs", "s8." + ext), "rb").read(), np.uint8)
        ppm = open(os.path.join(run, "s8_cpu_img.ppm"), "rb").read()
        td2 = os.path.join(td, "fresh")
        dataset(td2, "s8", 8)
        V.write_tf(os.path.join(td2, "tet_offs", "s8.tf"), tf)
        h = O.run_ref_harness("strict", os.path.join(td2, "tet_offs", "s8"), 128, 128)
        td3 = os.path.join(td, "fresh2")
        dataset(td3, "s8", 8)
        V.write_tf(os.path.join(td3, "tet_offs", "s8.tf"), tf)
        hf = O.run_ref_harness("fast", os.path.join(td3, "tet_offs", "s8"), 128, 128)
        np.savez_compressed(os.path.join(HERE, "ref_small.npz"),
                            con=files["con"], extf=files["extf"], lmt=files["lmt"], tf_default=files["tf"],
                            incid=files["incid"], ppm512_fast=np.frombuffer(ppm, np.uint8),
                            owner128=h["entry_ext"], rgb128_strict=h["rgb8"], rgb128_fast=hf["rgb8"],
                            verts_normalized=h["verts"])
        # 16^3 at 512^2 with the seeded TF, fresh directories (in-memory limits)
        td4 = os.path.join(td, "c1s")
        dataset(td4, "c1", 16)
        V.write_tf(os.path.join(td4, "tet_offs", "c1.tf"), WL.plain_cpu()[2])
        hs = O.run_ref_harness("strict", os.path.join(td4, "tet_offs", "c1"), 512, 512)
        td5 = os.path.join(td, "c1f")
        dataset(td5, "c1", 16)
        V.write_tf(os.path.join(td5, "tet_offs", "c1.tf"), WL.plain_cpu()[2])
        hfc = O.run_ref_harness("fast", os.path.join(td5, "tet_offs", "c1"), 512, 512)
        out["c1"] = {"owner_sha256": sha(hs["entry_ext"]), "rgb_strict_sha256": sha(hs["rgb8"]),
                     "rgb_fast_sha256": sha(hfc["rgb8"]), "info": hs["info"],
                     "owner_fast_equals_strict": bool(np.array_equal(hs["entry_ext"], hfc["entry_ext"])),
                     "fast_vs_strict_pixels_differing": int((hs["rgb8"] != hfc["rgb8"]).any(-1).sum()),
                     "covered_pixels": int((hs["entry_ext"] != 0xFFFFFFFF).sum())}
    json.dump(out["c1"], open(os.path.join(HERE, "ref_c1.json"), "w"), indent=1)
    # psi table
    pr = read_psi_header(os.path.join(REF, "psiGammaTable512.h"))
    pc = V.psi_generate()
    d = np.abs(pr.astype(np.float64) - pc)
    idx = np.argwhere(d > 1e-6)
    json.dump({"sha256": sha(pr), "threshold": 1e-6, "n_deviations": int(len(idx)),
               "deviations": [[int(i), int(j), float(pr[i, j])] for i, j in idx],
               "max_other_abs_diff": float(d[d <= 1e-6].max())},
              open(os.path.join(HERE, "psi_ref.json"), "w"))
    caches()
    print("golden fixtures written")


if __name__ == "__main__":
    main()

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

  python3 tests/golden/make_golden.py [cleantf|rotated]   (all, or one)
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


def main() -> None:
    if sys.argv[1:] == ["cleantf"]:
        cleantf()
        return
    if sys.argv[1:] == ["rotated"]:
        rotated()
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
    print("golden fixtures written")


if __name__ == "__main__":
    main()

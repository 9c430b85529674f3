"""Partial and damaged cache-file states of the 8^3 dataset "s8", shared by
the fixture writer (make_golden.py, which records what the reference's own
CLI does from each state) and the CLI tests (which start the command from the
same state and compare).

A state is built in <dir>/tet_offs from the canonical cache files the
reference writes on a first run (ref_small.npz: con, extf, lmt, tf_default,
incid) and the seeded .off:

  con_count_incid     .con's tet count off by one; .incid present
  con_count_no_incid  the same without .incid
  incid_count         no .con; .incid's vertex count off by one
  tf_255              .tf announcing 255 entries (its 256 lines kept)
  extf_count          .extf's face count off by one
  extf_stale          no .con, .incid present, .extf with the right count and
                      its faces in reverse order (the reference keeps it after
                      rebuilding .con)
  off_truncated       the .off cut in the middle of its tetrahedra
"""
from __future__ import annotations

import os
import re

import numpy as np

STATES = ("con_count_incid", "con_count_no_incid", "incid_count", "tf_255", "extf_count", "extf_stale",
          "off_truncated")
CACHES = ("con", "incid", "tf", "extf", "lmt")


def make_state(state: str, tet_offs: str, canonical: dict, off_bytes: bytes) -> None:
    """Writes s8.off and the cache files of `state` into tet_offs."""
    os.makedirs(tet_offs, exist_ok=True)
    files = {ext: bytes(np.asarray(canonical[key], np.uint8))
             for ext, key in (("con", "con"), ("extf", "extf"), ("lmt", "lmt"), ("tf", "tf_default"),
                              ("incid", "incid"))}
    off = off_bytes

    def bump_line(b: bytes, line: int, delta: int) -> bytes:
        lines = b.split(b"\n")
        toks = lines[line].split(b" ")
        toks[0] = str(int(toks[0]) + delta).encode()
        lines[line] = b" ".join(toks)
        return b"\n".join(lines)

    if state == "con_count_incid":
        files["con"] = bump_line(files["con"], 1, 1)
    elif state == "con_count_no_incid":
        files["con"] = bump_line(files["con"], 1, 1)
        del files["incid"]
    elif state == "incid_count":
        del files["con"]
        files["incid"] = bump_line(files["incid"], 0, 1)
    elif state == "tf_255":
        files["tf"] = bump_line(files["tf"], 0, -1)
    elif state == "extf_count":
        files["extf"] = bump_line(files["extf"], 0, 1)
    elif state == "extf_stale":
        del files["con"]
        lines = files["extf"].rstrip(b"\n").split(b"\n")
        files["extf"] = b"\n".join([lines[0]] + lines[1:][::-1]) + b"\n"
    elif state == "off_truncated":
        lines = off.split(b"\n")
        off = b"\n".join(lines[: len(lines) - len(lines) // 8])
    else:
        raise ValueError(state)
    open(os.path.join(tet_offs, "s8.off"), "wb").write(off)
    for ext, b in files.items():
        open(os.path.join(tet_offs, "s8." + ext), "wb").write(b)


_KEEP = re.compile(r"^(Reading |Writing |Deleting |Cannot read | Computing | Building )|file not found!")


def cache_lines(stdout: str) -> list[str]:
    """The cache-related lines of a run's output, file paths reduced to their
    base names (the reference reads ../tet_offs/, the tests another dir)."""
    out = []
    for ln in stdout.splitlines():
        if _KEEP.search(ln):
            out.append(re.sub(r"\S*/(s8\.\w+)", r"\1", ln).rstrip())
    return out

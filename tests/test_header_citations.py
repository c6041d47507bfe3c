"""Every C-ABI declaration says which reference call it replaces (a src/
file:line or a pEVSL F90 name) or that it has no reference counterpart."""
import re
from pathlib import Path

HEADER = Path(__file__).resolve().parents[1] / "include" / "nmgpu.h"


def _declarations():
    text = HEADER.read_text()
    out = []
    blanket = False  # a section comment covering every entry below it
    comments = []
    for tok in re.finditer(r"/\*.*?\*/|(?:int|const char\s*\*)\s+(nmg_\w+)\s*\([^;]*\);", text, re.S):
        s = tok.group(0)
        if s.startswith("/*"):
            if "none of the entries" in s and "reference counterpart" in s:
                blanket = True
            comments.append(s)
            continue
        com = re.sub(r"\s*\n\s*\*?\s*", " ", comments[-1]) if comments else ""
        out.append((tok.group(1), com, blanket))
        comments = []
    return out


def test_every_declaration_cites_the_reference_or_says_it_has_none():
    decls = _declarations()
    assert len(decls) > 60
    cite = re.compile(r"src/\w+\.f90|_F90|_f90|no reference\s+counterpart|not called by the reference|"
                      r"pEVSL's|pEVSL (?:draws|spslicer)|chebiter_set_mode|kernel_probe|as nmg_set_tiling")
    missing = [name for name, com, blanket in decls if not blanket and not cite.search(com)]
    # entries declared in a group after one shared comment
    grouped = {"nmg_version", "nmg_device_info", "nmg_kernel_probe"}
    assert [m for m in missing if m not in grouped] == []

"""CPU tests of the build-time autotuner's plumbing (pyspectrum_amd/tune.py,
pyspectrum_amd/build.py): every candidate definition is a real compile-time
knob of the sources, and build() turns tuning.json into -D flags."""
import json
import os
import re

from pyspectrum_amd import build as B
from pyspectrum_amd import tune as T


def test_candidates_name_real_knobs():
    src = ""
    for f in os.listdir(B.CSRC):
        if f.endswith((".hip", ".h")):
            with open(os.path.join(B.CSRC, f)) as fh:
                src += fh.read()
    # knobs with a default (#ifndef X / #define X v) and optional overrides (#ifdef X)
    knobs = set(re.findall(r"#ifndef (PSX_[A-Z0-9_]+)\n#define \1 ", src))
    knobs |= set(re.findall(r"#ifdef (PSX_[A-Z0-9_]+)\n", src))
    assert "default" in T.CANDIDATES and T.CANDIDATES["default"] == {}
    for name, d in T.CANDIDATES.items():
        assert set(d) <= knobs, (name, set(d) - knobs)


def test_tuned_defines_from_json(tmp_path, monkeypatch):
    p = tmp_path / "tuning.json"
    p.write_text(json.dumps({"defines": {"PSX_CT_KC_BF": 128, "PSX_SHELL_X_PAIRS": 4}}))
    monkeypatch.setattr(B, "TUNING", str(p))
    assert B.tuned_defines() == ["-DPSX_CT_KC_BF=128", "-DPSX_SHELL_X_PAIRS=4"]
    monkeypatch.setattr(B, "TUNING", str(tmp_path / "absent.json"))
    assert B.tuned_defines() == []


def test_committed_tuning_is_consistent():
    if not os.path.isfile(B.TUNING):
        return
    with open(B.TUNING) as f:
        doc = json.load(f)
    assert doc["defines"] == T.CANDIDATES[doc["candidate"]]

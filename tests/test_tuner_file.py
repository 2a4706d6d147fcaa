"""The tuner file: decisions load into a context, merge, save and load back;
malformed files are refused; COF_TUNER_FILE picks the file a context opens
with; the build's own tuned_gfx950.json parses and names only variants the
launchers know."""
import json
import os
import subprocess
import sys

import pytest

import casoff

SIG_A = "fused|1c|NNNNNNNNNNNNNNNNNNNNNGG|7|4"
SIG_B = "bulge|1c|2f|1,0,0,0,0,|2070,2071,|6|6"


def test_load_merge_save_roundtrip(tmp_path):
    c = casoff.Context(-1)
    before = c.tuner_info()
    f1, f2, f3 = (str(tmp_path / n) for n in ("a.json", "b.json", "c.json"))
    with open(f1, "w") as f:
        json.dump({SIG_A: 4}, f)
    with open(f2, "w") as f:
        f.write('{\n  "%s": 2 ,"%s":3}\n' % (SIG_B, SIG_A))
    c.tuner_load(f1)
    assert c.tuner_info()[SIG_A] == 4
    c.tuner_load(f2)  # merges; later files win
    got = c.tuner_info()
    assert got[SIG_A] == 3 and got[SIG_B] == 2 and set(before) <= set(got)
    c.tuner_save(f3)
    with open(f3) as f:
        assert json.load(f) == got
    c.close()


@pytest.mark.parametrize("text", ['{"a": 1', '["a", 1]', '{"a": -1}', '{"a": "1"}', '{"a\\\\b": 1}', '{"a": 1} x', ""])
def test_malformed_files_are_refused(tmp_path, text):
    p = str(tmp_path / "bad.json")
    with open(p, "w") as f:
        f.write(text)
    c = casoff.Context(-1)
    with pytest.raises(casoff.CofError, match="malformed tuner file"):
        c.tuner_load(p)
    with pytest.raises(casoff.CofError, match="cannot read"):
        c.tuner_load(str(tmp_path / "missing.json"))
    c.close()


def test_env_selects_the_file(tmp_path):
    good, bad = str(tmp_path / "g.json"), str(tmp_path / "b.json")
    with open(good, "w") as f:
        json.dump({SIG_A: 1}, f)
    with open(bad, "w") as f:
        f.write("{")
    code = ("import casoff, sys\n"
            "try:\n c = casoff.Context(-1)\nexcept casoff.CofError as e:\n print('ERR', e); sys.exit(0)\n"
            "print(sorted(c.tuner_info().items()))\n")
    env = dict(os.environ, PYTHONPATH=os.path.dirname(casoff.__file__))
    for path, want in ((good, "[('%s', 1)]" % SIG_A), (bad, "ERR"), ("none", "[]")):
        out = subprocess.run([sys.executable, "-c", code], env=dict(env, COF_TUNER_FILE=path), capture_output=True,
                             text=True, timeout=60).stdout
        assert out.startswith(want), (path, out)


def test_build_tuner_file_parses():
    p = os.path.join(os.path.dirname(casoff.LIB_PATH), "tuned_gfx950.json")
    if not os.path.exists(p):
        pytest.skip("no tuner file built")
    with open(p) as f:
        d = json.load(f)
    assert d and all(isinstance(v, int) for v in d.values())
    for sig, v in d.items():
        assert sig.split("|")[0] in ("fused", "bulge"), sig
        assert 0 <= v < (5 if sig.startswith("fused") else 4), (sig, v)
    c = casoff.Context(-1)
    assert c.tuner_info() == d  # a context opens with it
    c.close()

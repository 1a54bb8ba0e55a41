"""The tile-shape table of the 2D kernels (include/fgtpw.h,
fgtpw_tuning_table): defaults, the table beside the library and the
FGTPW_TUNE override, with invalid entries ignored.  Host logic only; each
case runs in a fresh process because the table is read once per process."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _table(env_tune=None):
    env = dict(os.environ)
    env.pop("FGTPW_TUNE", None)
    if env_tune is not None:
        env["FGTPW_TUNE"] = env_tune
    code = "import json, fgt2d_amd; print(json.dumps(fgt2d_amd.tuning_table()))"
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True,
                       text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_override_and_validation():
    base = _table()
    t = _table("sort=2,segf1=128,sege2=64,spw1=4,spw2=99,sege1=512,segf9=128,bogus=1")
    assert t["sort"] == 2 and t["segf1"] == 128 and t["sege2"] == 64 and t["spw1"] == 4
    # out of range or unknown: the previous value stays
    assert t["spw2"] == base["spw2"] and t["sege1"] == base["sege1"]
    assert t["source"].endswith("FGTPW_TUNE")


def test_committed_table_parses():
    path = os.path.join(ROOT, "fgt2d_amd", "tuning.txt")
    t = _table()
    if os.path.exists(path):
        assert t["source"] == path
        for line in open(path):
            line = line.split("#", 1)[0].strip()
            if "=" in line:
                k, v = line.split("=")
                assert t[k.strip()] == int(v), k
    else:
        assert t["source"] == "defaults"

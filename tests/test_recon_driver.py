"""Host logic of the multichannel driver (cudasirecon_amd/recon.py, the
counterpart of the reference's recon.py:21-153): channel extraction for each
interleave, merging, OTF/config lookup, and the command line handed to the
executable, with a stand-in executable so no GPU is needed."""
import json
import os
import sys

import numpy as np
import pytest

from cudasirecon_amd import mrc, recon

NX, NY, NZ, NT = 6, 5, 3, 2
WAVES = [528, 608]


def _value(w, t, z):
    return 1000 * w + 100 * t + z


def _multiwave(path, interleaved, waves=WAVES, nt=NT):
    """A float MRC whose every section is filled with a code of its (wave, time, z)."""
    h = mrc.make_header(NX, NY, NZ * len(waves) * nt, 2, num_waves=len(waves), num_times=nt,
                        wave=waves[0], interleaved=interleaved, labels=["multi-wave test"])
    for i, w in enumerate(waves[1:], start=2):
        h["iwav%d" % i] = w
    data = np.zeros((h["nz"], NY, NX), np.float32)
    for w in range(len(waves)):
        for t in range(nt):
            for z in range(NZ):
                data[mrc.section_index(h, z, w, t, NZ)] = _value(w, t, z)
    mrc.write(path, data, h)
    return h


def _otfs_and_configs(d, waves=WAVES):
    for w in waves:
        (d / ("otf%d.otf" % w)).write_bytes(b"")
        (d / ("config%d.txt" % w)).write_text("wiener=0.00%d\n" % (w % 7 + 1))


@pytest.mark.parametrize("interleaved", [0, 1, 2])
def test_split_follows_the_interleave(tmp_path, interleaved):
    src = str(tmp_path / "cells.dv")
    _multiwave(src, interleaved)
    _otfs_and_configs(tmp_path)
    with recon.file_splitter(src, str(tmp_path), str(tmp_path)) as parts:
        assert [p[1] for p in parts] == WAVES
        for c, (f, w, otf, cfg) in enumerate(parts):
            assert f == str(tmp_path / ("cells-%d.dv" % w))
            assert otf.endswith("otf%d.otf" % w) and cfg.endswith("config%d.txt" % w)
            h, d = mrc.read(f)
            assert (h["num_waves"], h["num_times"], h["iwav1"], h["interleaved"]) == (1, NT, w, 0)
            assert h["iwav2"] == 0 and h["nz"] == NZ * NT
            assert h["labels"][0].startswith("multi-wave test")
            want = [_value(c, t, z) for t in range(NT) for z in range(NZ)]
            assert d[:, 0, 0].tolist() == want
            assert np.all(d == d[:, :1, :1])
        files = [p[0] for p in parts]
    # the extracted channels are removed again
    assert not any(os.path.exists(f) for f in files)
    assert os.path.exists(src)


def test_single_wave_file_is_not_split(tmp_path):
    src = str(tmp_path / "one.dv")
    _multiwave(src, 0, waves=[528])
    _otfs_and_configs(tmp_path, [528])
    with recon.file_splitter(src, str(tmp_path), str(tmp_path)) as parts:
        assert parts == [(src, 528, str(tmp_path / "otf528.otf"), str(tmp_path / "config528.txt"))]
    assert os.path.exists(src)


def test_channel_without_otf_is_skipped(tmp_path, capsys):
    src = str(tmp_path / "cells.dv")
    _multiwave(src, 1)
    _otfs_and_configs(tmp_path, [608])
    with recon.file_splitter(src, str(tmp_path), str(tmp_path)) as parts:
        assert [p[1] for p in parts] == [608]
    assert "Could not find otf file: %s" % (tmp_path / "otf528.otf") in capsys.readouterr().out
    (tmp_path / "otf528.otf").write_bytes(b"")
    assert recon.get_otf_and_config(528, str(tmp_path), str(tmp_path / "nowhere")) is None
    assert "Could not find config file" in capsys.readouterr().out


def test_merge_writes_wave_fastest(tmp_path):
    files = []
    for c, w in enumerate(WAVES):
        h = mrc.make_header(NX, NY, NZ * NT, 2, num_times=NT, wave=w, xlen=0.04, zlen=0.125)
        d = np.stack([np.full((NY, NX), _value(c, t, z), np.float32) for t in range(NT) for z in range(NZ)])
        f = str(tmp_path / ("p%d.dv" % w))
        mrc.write(f, d, h)
        files.append(f)
    out = recon.merge_files(files, str(tmp_path / "merged.dv"), delete=True)
    h, d = mrc.read(out)
    assert (h["num_waves"], h["num_times"], h["interleaved"], h["nz"]) == (2, NT, 1, 2 * NT * NZ)
    assert [h["iwav%d" % i] for i in range(1, 6)] == WAVES + [0, 0, 0]
    assert h["xlen"] == pytest.approx(0.04)
    for c in range(2):
        for t in range(NT):
            for z in range(NZ):
                assert d[mrc.section_index(h, z, c, t), 0, 0] == _value(c, t, z)
    assert not any(os.path.exists(f) for f in files)


def test_merge_refuses_mismatched_channels(tmp_path):
    a, b = str(tmp_path / "a.dv"), str(tmp_path / "b.dv")
    mrc.write(a, np.zeros((2, 4, 4), np.float32), mrc.make_header(4, 4, 2, 2))
    mrc.write(b, np.zeros((3, 4, 4), np.float32), mrc.make_header(4, 4, 3, 2))
    with pytest.raises(ValueError, match="differ in shape"):
        recon.merge_files([a, b], str(tmp_path / "m.dv"))


STAND_IN = r'''
import json, sys
sys.path.insert(0, %(root)r)
from cudasirecon_amd import mrc
a = sys.argv[1:]
if "608" in a[0] and %(fail608)r:
    sys.exit(3)
h, d = mrc.read(a[0])
mrc.write(a[1], 2 * d.astype("float32"), dict(h, mode=2))
with open(a[1] + ".argv.json", "w") as f:
    json.dump(a, f)
'''


def _stand_in(tmp_path, fail608=False):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    s = tmp_path / "stand_in.py"
    s.write_text(STAND_IN % {"root": root, "fail608": fail608})
    return [sys.executable, str(s)]


def test_reconstruct_runs_each_channel_and_merges(tmp_path):
    src = str(tmp_path / "cells.dv")
    h_in = _multiwave(src, 2)
    _otfs_and_configs(tmp_path)
    out = recon.reconstruct(src, otf_dir=str(tmp_path), config_dir=str(tmp_path), app=_stand_in(tmp_path),
                            wiener=0.002, **{"--ls": 0.2035})
    assert out == str(tmp_path / "cells_PROC.dv")
    h, d = mrc.read(out)
    assert (h["num_waves"], h["interleaved"], h["iwav1"], h["iwav2"]) == (2, 1, 528, 608)
    for c in range(2):
        for t in range(NT):
            for z in range(NZ):
                assert d[mrc.section_index(h, z, c, t), 0, 0] == 2 * _value(c, t, z)
    # the per-channel command lines: input, output, otf, -c config, then the extra pairs
    for w in WAVES:
        argv = json.load(open(str(tmp_path / ("cells-%d_PROC.dv.argv.json" % w))))
        assert argv[0].endswith("cells-%d.dv" % w) and argv[1].endswith("cells-%d_PROC.dv" % w)
        assert argv[2:] == [str(tmp_path / ("otf%d.otf" % w)), "-c", str(tmp_path / ("config%d.txt" % w)),
                            "--wiener", "0.002", "--ls", "0.2035"]
    # split inputs and per-channel outputs are gone; the input is untouched
    left = sorted(p.name for p in tmp_path.iterdir() if p.suffix == ".dv")
    assert left == ["cells.dv", "cells_PROC.dv"]
    assert mrc.read(src)[0]["num_waves"] == h_in["num_waves"]


def test_failed_channel_is_reported_and_skipped(tmp_path, capsys):
    src = str(tmp_path / "cells.dv")
    _multiwave(src, 0)
    _otfs_and_configs(tmp_path)
    out = recon.reconstruct(src, otf_dir=str(tmp_path), config_dir=str(tmp_path),
                            app=_stand_in(tmp_path, fail608=True))
    assert "Cannot reconstruct file %s" % (tmp_path / "cells-608.dv") in capsys.readouterr().out
    h, d = mrc.read(out)
    # one channel left: renamed to the output, not merged
    assert (h["num_waves"], h["iwav1"]) == (1, 528)
    assert d[:, 0, 0].tolist() == [2 * _value(0, t, z) for t in range(NT) for z in range(NZ)]


def test_no_channel_raises(tmp_path):
    src = str(tmp_path / "cells.dv")
    _multiwave(src, 0)
    with pytest.raises(RuntimeError, match="no channel"):
        recon.reconstruct(src, otf_dir=str(tmp_path), config_dir=str(tmp_path), app=_stand_in(tmp_path))


def test_command_line(tmp_path, monkeypatch):
    src = str(tmp_path / "cells.dv")
    _multiwave(src, 1)
    _otfs_and_configs(tmp_path)
    seen = {}
    monkeypatch.setattr(recon, "reconstruct", lambda f, **kw: seen.update(f=f, **kw) or "x")
    assert recon.main([src, "--otf-dir", str(tmp_path), "--wiener", "0.003", "--config-dir", "/c"]) == 0
    assert seen == {"f": os.path.abspath(src), "otf_dir": str(tmp_path), "config_dir": "/c", "app": None,
                    "wiener": "0.003"}
    with pytest.raises(SystemExit):
        recon.main([src, "--wiener"])
    with pytest.raises(SystemExit):
        recon.main([str(tmp_path / "missing.dv")])

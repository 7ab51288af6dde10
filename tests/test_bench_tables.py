"""The bench's per-stage tables against the engine and the stage geometry.

`bench.py` prices every stage the engine times (`Engine::stage_begin`) with
an algorithmic byte count, and names the kernels each stage launches.  These
checks run on the CPU only: every stage name the reconstruction path times
has a table entry, and the byte counts of the plain 512x512x48 stack (3
directions x 5 phases, 2x lateral zoom) match their closed forms.
"""
import glob
import os
import re
import sys
import types

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

# stages outside the plain reconstruction path (the deskew resampling runs
# only with a non-zero deskew angle)
OFF_PATH = {"resample"}


def engine_stage_names():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "cudasirecon_amd", "csrc", "*.hip")):
        with open(path) as f:
            names |= set(re.findall(r'stage_begin\("([a-z0-9_]+)"', f.read()))
    return names


def plain_stack():
    return types.SimpleNamespace(nx=512, ny=512, nz=48, ndirs=3, nphases=5, zoomfact=2.0, z_zoom=1,
                                 wave=515.0, nimm=1.515, na=1.4, dz=0.125)


def test_every_timed_stage_has_a_kernel_entry():
    names = engine_stage_names()
    assert {"filter", "asm_x", "asm_y", "ov_cols", "ov_rows"} <= names
    missing = sorted(names - OFF_PATH - set(bench.KERNEL_OF_STAGE))
    assert not missing, missing


def test_byte_and_flop_tables_cover_the_same_stages():
    p = plain_stack()
    L = bench.geometry(p)
    sb = bench.stage_bytes(p, L)
    sf = bench.stage_fft_flops(p, L)
    assert set(sb) <= set(bench.KERNEL_OF_STAGE)
    assert set(sf) <= set(sb)
    assert all(v > 0 for v in sb.values())


def test_plain_stack_bytes_closed_forms():
    p = plain_stack()
    L = bench.geometry(p)
    sb = bench.stage_bytes(p, L)
    H = 257 * 512 * 48  # half-spectrum elements of one band
    # every band of the stack in one filter launch: order 0 read and written,
    # each side-band pair (two bands) read and written
    assert sb["filter"] == 3 * (2 + 4 * 2) * H * 8 == 1515847680
    # the x pass reads the 7 y-transformed bands once and writes the float
    # volume once
    assert sb["asm_x"] == 7 * 48 * 1024 * 512 * 8 + 48 * 1024 * 1024 * 4 == 1610612736
    # the two-pass overlap construction saves exactly one write and one read
    # of both overlaps' stored planes against the three-pass form
    nxy, nzs = 512 * 512, L["nzs"]
    three = sb["ov_x"] + sb["ov_y"] + sb["ov_zsum"]
    two = sb["ov_cols"] + sb["ov_rows"]
    assert three - two == p.ndirs * 2 * 2 * nzs * nxy * 8


def test_profiled_traffic_covers_the_default_path():
    # the FP64 fit (the default) builds its overlaps in two passes; the
    # three-pass stages run only with the FP32 fit and the dumps
    three_pass = {"ov_x", "ov_y", "ov_zsum"}
    p = plain_stack()
    sb = bench.stage_bytes(p, bench.geometry(p))
    for stage in sorted(set(sb) - three_pass):
        t = bench.profiled_traffic(stage)
        assert t is not None and t > 0, stage

"""Multichannel driver: reconstruct every emission wavelength of an MRC/DV file.

The cudasirecon executable reconstructs the first wave of its input only, so
a multi-wave acquisition is split into one single-wave file per channel, each
channel is reconstructed with the OTF and config file named for its
wavelength (``otf<wave>.otf``, ``config<wave>.txt``), and the reconstructions
are merged into one multi-wave file.  This is the counterpart of the
reference's recon.py:21-153 (get_otf_and_config, merge_files,
write_single_channel, file_splitter, reconstruct_single, reconstruct and its
command line), with these deliberate differences:

- channels are extracted by the header's interleave (ZTW, WZT or ZWT; every
  time point of the wave, time-major), where recon.py:86-88 takes the
  third-to-last array axis, which is the wave axis for WZT files only;
- the merged file says what it holds: W fastest (WZT, interleave code 1),
  the layout of the channels stacked per section, with NumWaves and the
  wavelength list of the reconstructed channels (recon.py:33-46 stacks the
  same way but keeps the first channel's interleave code);
- extra ``--key value`` pairs on the command line are passed on with one
  pair of dashes (recon.py:150-152 keeps the key's own dashes and adds two);
- a channel whose executable run fails (non-zero status or no output file)
  is reported and skipped, as recon.py:132-136 does for an exception, and a
  file with no reconstructable channel raises instead of indexing an empty
  list.

Usage: ``python -m cudasirecon_amd.recon infile [--otf-dir D] [--config-dir D]
[--app PATH] [--key value ...]``.
"""
import argparse
import contextlib
import os
import subprocess
import sys

import numpy as np

from . import mrc

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# the reconstruction executable (built by cudasirecon_amd/csrc/Makefile)
APP = os.path.join(PKG_DIR, "bin", "cudasirecon")


def get_otf_and_config(wave, otf_dir=None, config_dir=None):
    """(otf path, config path) for an emission wavelength in nm, or None with a
    message when either file is missing (recon.py:21-31).  Directories default
    to the current working directory."""
    otf = os.path.join(otf_dir or os.getcwd(), "otf%d.otf" % wave)
    config = os.path.join(config_dir or os.getcwd(), "config%d.txt" % wave)
    if not os.path.exists(otf):
        print("Could not find otf file: %s" % otf)
        return None
    if not os.path.exists(config):
        print("Could not find config file: %s" % config)
        return None
    return otf, config


def _waves(h):
    return [h["iwav%d" % (i + 1)] for i in range(5)]


def _single_wave_header(h, nz, wave):
    out = dict(h)
    out.update(nz=nz, mz=nz, num_waves=1, interleaved=0, iwav1=wave, iwav2=0, iwav3=0, iwav4=0, iwav5=0)
    return out


def channel_sections(h, w):
    """Section indices of wave w, every time point, in (time, z) order."""
    nw, nt = h["num_waves"], h["num_times"]
    nzz = h["nz"] // (nw * nt)
    return [mrc.section_index(h, z, w, t, nzz) for t in range(nt) for z in range(nzz)]


def write_single_channel(array, outpath, hdr, wave):
    """A single-wave file of `array` ([sections][ny][nx], z fastest then time)
    with the other header fields of `hdr` (recon.py:51-61)."""
    print("Extracting channel %d..." % wave)
    mrc.write(outpath, array, _single_wave_header(hdr, array.shape[0], wave))


@contextlib.contextmanager
def file_splitter(fname, otf_dir=None, config_dir=None):
    """Yields [(file, wave, otf, config)] for the channels of `fname` that have
    an OTF and a config; multi-wave files are split into `<stem>-<wave><ext>`
    files, removed again on exit (recon.py:64-101)."""
    h, data = mrc.read(fname)
    stem, ext = os.path.splitext(fname)
    splitfiles = []
    try:
        if h["num_waves"] == 1:
            wave = h["iwav1"]
            found = get_otf_and_config(wave, otf_dir, config_dir)
            if found is not None:
                splitfiles = [(fname, wave) + found]
        else:
            for c in range(h["num_waves"]):
                wave = _waves(h)[c]
                found = get_otf_and_config(wave, otf_dir, config_dir)
                if found is None:
                    continue
                out = "%s-%d%s" % (stem, wave, ext)
                write_single_channel(data[channel_sections(h, c)], out, h, wave)
                splitfiles.append((out, wave) + found)
        yield splitfiles
    finally:
        if h["num_waves"] > 1:
            print("\nCleaning up extracted channels...")
            for f in splitfiles:
                if os.path.exists(f[0]):
                    os.remove(f[0])


def reconstruct_single(file, otf, config=None, outfile=None, app=None, **kwargs):
    """Run the executable on one single-wave file (recon.py:104-118): the output
    defaults to `<stem>_PROC<ext>`; keyword arguments become `--key value`."""
    if outfile is None:
        stem, ext = os.path.splitext(file)
        outfile = "%s_PROC%s" % (stem, ext)
    app = app or APP
    cmd = (list(app) if isinstance(app, (list, tuple)) else [app]) + [file, outfile, otf]
    if config:
        cmd += ["-c", config]
    for key, value in kwargs.items():
        cmd += ["--%s" % key.lstrip("-"), str(value)]
    print("Running command: %s" % " ".join(cmd))
    r = subprocess.run(cmd)
    if r.returncode != 0:
        raise RuntimeError("%s exited with status %d" % (cmd[0], r.returncode))
    if not os.path.exists(outfile):
        raise RuntimeError("%s wrote no output file %s" % (cmd[0], outfile))
    return outfile


def merge_files(file_list, outfile, delete=False):
    """One multi-wave file from single-wave reconstructions (recon.py:33-48):
    sections W fastest (interleave code 1), NumWaves and wavelengths set, the
    other header fields from the first file."""
    if not 1 <= len(file_list) <= 5:
        raise ValueError("merge_files: 1 to 5 channels, got %d" % len(file_list))
    read = [mrc.read(f) for f in file_list]
    h0 = read[0][0]
    shapes = {d.shape for _, d in read}
    if len(shapes) != 1:
        raise ValueError("merge_files: channels differ in shape: %s" % sorted(shapes))
    merged = np.stack([d for _, d in read], axis=1)  # [section][wave][ny][nx]
    merged = merged.reshape(-1, merged.shape[2], merged.shape[3])
    h = dict(h0)
    h.update(nz=merged.shape[0], mz=merged.shape[0], num_waves=len(file_list), interleaved=1)
    waves = [hh["iwav1"] for hh, _ in read] + [0] * (5 - len(read))
    for i in range(5):
        h["iwav%d" % (i + 1)] = waves[i]
    mrc.write(outfile, merged, h)
    if delete:
        for f in file_list:
            try:
                os.remove(f)
            except OSError:
                pass
    return outfile


def reconstruct(inFile, outfile=None, otf_dir=None, config_dir=None, app=None, **kwargs):
    """Split, reconstruct every channel with its own OTF and config, merge
    (recon.py:121-143).  Returns the output path (`<stem>_PROC<ext>` by default)."""
    if outfile is None:
        stem, ext = os.path.splitext(inFile)
        outfile = "%s_PROC%s" % (stem, ext)
    with file_splitter(inFile, otf_dir, config_dir) as splitfiles:
        processed = []
        for file, wave, otf, config in splitfiles:
            try:
                processed.append(reconstruct_single(file, otf, config, app=app, **kwargs))
            except Exception as e:  # noqa: BLE001 - report and go on with the other channels
                print("Cannot reconstruct file %s due to error %s" % (file, e))
        if not processed:
            raise RuntimeError("no channel of %s could be reconstructed" % inFile)
        if len(processed) > 1:
            print("\nMerging multi-channel reconstructions...")
            return merge_files(processed, outfile, delete=True)
        if processed[0] != outfile:
            os.replace(processed[0], outfile)
        return outfile


def main(argv=None):
    ap = argparse.ArgumentParser(description="Reconstruct every channel of a multi-wave MRC/DV file.")
    ap.add_argument("infile")
    ap.add_argument("--otf-dir", default=None, help="where otf<wave>.otf files are (default: cwd)")
    ap.add_argument("--config-dir", default=None, help="where config<wave>.txt files are (default: cwd)")
    ap.add_argument("--app", default=None, help="reconstruction executable (default: the package's cudasirecon)")
    args, extras = ap.parse_known_args(argv)
    if len(extras) % 2:
        ap.error("extra options come in '--key value' pairs: %s" % " ".join(extras))
    if not os.path.exists(args.infile):
        ap.error("no such file: %s" % args.infile)
    kw = {k.lstrip("-"): v for k, v in zip(extras[::2], extras[1::2])}
    print(reconstruct(os.path.abspath(args.infile), otf_dir=args.otf_dir, config_dir=args.config_dir,
                      app=args.app, **kw))
    return 0


if __name__ == "__main__":
    sys.exit(main())

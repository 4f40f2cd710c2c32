"""Drop-in command-line tools (cucm_amd/bin/cucm_compress, cucm_decompress):
the reference's options, defaults, size suffixes and error exits
(cucm_compress_v2.cu:1252-1418, cucm_decompress_v2.cu:1250-1313).  Option
handling is checked on the CPU (the tools report their parsed settings before
touching the GPU); compression itself runs in the GPU tests."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cucm_amd", "bin")
COMP = os.path.join(BIN, "cucm_compress")
DECOMP = os.path.join(BIN, "cucm_decompress")

pytestmark = pytest.mark.skipif(not os.path.exists(COMP), reason="tools not built")


def _run(args):
    return subprocess.run(args, capture_output=True, text=True, timeout=60)


def test_no_arguments_prints_help_and_succeeds():
    r = _run([COMP])
    assert r.returncode == 0 and "-s, --chunksize" in r.stdout and "-p, --prelearn" in r.stdout


def test_unknown_option_fails_with_help():
    r = _run([COMP, "-x"])
    assert r.returncode == 1 and "Unknown option: -x" in r.stderr and "Usage:" in r.stdout


@pytest.mark.parametrize("flag,value,msg", [("-c", "0", "Invalid chunknum: 0"),
                                            ("-s", "abc", "Invalid chunk size: abc"),
                                            ("-u", "-4", "Invalid unit size: -4"),
                                            ("-b", "0", "Invalid batch number: 0")])
def test_invalid_values_are_rejected(flag, value, msg):
    r = _run([COMP, flag, value])
    assert r.returncode == 1 and msg in r.stderr


def test_missing_input_file():
    r = _run([COMP, "-f", "/nonexistent/input"])
    assert r.returncode == 1 and "Failed to open input file: /nonexistent/input" in r.stderr


@pytest.mark.parametrize("arg,bytes_", [("64K", 65536), ("2m", 2 << 20), ("3 G", 3 << 30),
                                        ("4096", 4096), ("7Q", 7), ("12k9", 12288)])
def test_size_suffixes(tmp_path, arg, bytes_):
    src = tmp_path / "in.bin"
    src.write_bytes(b"x" * 1000)
    # an output path that cannot be opened stops the tool right after it has
    # reported its parsed settings, before any GPU work
    r = _run([COMP, "-f", str(src), "-s", arg, "-o", str(tmp_path / "no" / "out")])
    assert r.returncode == 1 and "Failed to open output file" in r.stderr
    assert ("chunksize: %d," % bytes_) in r.stdout
    assert ("chunknum: %d," % ((1000 + bytes_ - 1) // bytes_)) in r.stdout


def test_defaults_match_the_reference(tmp_path):
    src = tmp_path / "in.bin"
    src.write_bytes(b"y" * 3000)
    r = _run([COMP, "-f", str(src), "-o", str(tmp_path / "no" / "out")])
    line = [l for l in r.stdout.splitlines() if l.startswith("Compressing file:")][0]
    assert "chunksize: 1048576, chunknum: 1, unit: 128, prelearn: 0, batch: 100" in line


def test_decompressor_usage_errors(tmp_path):
    r = _run([DECOMP])
    assert r.returncode == 1 and "Usage:" in r.stdout
    r = _run([DECOMP, "-i", "/nonexistent", "-o", str(tmp_path / "o")])
    assert r.returncode == 1 and "cannot open input file" in r.stderr
    bad = tmp_path / "bad"
    bad.write_bytes(b"\x00" * 21 + b"\x07")
    bad.write_bytes(b"\x00" * 20 + b"\x07")
    r = _run([DECOMP, "-i", str(bad), "-o", str(tmp_path / "o")])
    assert r.returncode == 1 and "not a container" in r.stderr

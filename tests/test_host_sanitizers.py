"""The library's host C++ under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU): the generator and flattening (csrc/netgen.cc), the Bookshelf reader
and writer (csrc/bookshelf.cc, which parses untrusted files) on well-formed
and malformed designs, and the legalizer's host geometry (csrc/lal_geom.cc),
compiled with g++ -fsanitize=address,undefined into tests/host_sanitize_check.cc.
Any report fails the test."""
import shutil
import subprocess
from pathlib import Path

import pytest

from test_bookshelf import NETS, NODES, PL, SCL, WTS, write_design

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "ktplace_amd" / "csrc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("san") / "host_sanitize_check"
    r = subprocess.run([gxx, "-O1", "-g", "-std=c++20", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-fno-omit-frame-pointer", f"-I{ROOT / 'include'}", f"-I{CSRC}",
                        str(ROOT / "tests" / "host_sanitize_check.cc"), str(CSRC / "netgen.cc"),
                        str(CSRC / "bookshelf.cc"), str(CSRC / "lal_geom.cc"), "-lz", "-o", str(exe)],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        pytest.fail(r.stderr[-4000:])
    return exe


def run(exe, *args):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300,
                       env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert r.returncode == 0 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    return r.stdout


def test_generator_and_flattening(checker):
    assert run(checker, "gen").count("gen seed") == 6


def test_lal_geometry(checker):
    assert "geom ok" in run(checker, "geom")


MALFORMED_NETS = [
    NETS,
    NETS.replace("NetDegree : 3 alpha", "NetDegree : nan alpha"),
    NETS.replace("NetDegree : 3 alpha", "NetDegree : 1e300 alpha"),
    NETS.replace("NetDegree : 3 alpha", "NetDegree : -4 alpha"),
    NETS.replace("NetDegree : 3 alpha", "NetDegree :"),
    NETS.replace("a I : 0.5 0.25", "a I : x y"),
    NETS.replace("a I : 0.5 0.25", "a I :"),
    "UCLA nets 1.0\nNetDegree : 2\n",
    "",
]


@pytest.mark.parametrize("k", range(len(MALFORMED_NETS)))
def test_bookshelf_reader_on_malformed_nets(checker, tmp_path, k):
    d = write_design(tmp_path, nodes=NODES, nets=MALFORMED_NETS[k], pl=PL, scl=SCL, wts=WTS)
    out = run(checker, "read", d, tmp_path / "placed.pl")
    assert "read" in out or "error" in out


@pytest.mark.parametrize("nodes,pl,scl", [
    (NODES.replace("a   2  1", "a   nan  inf"), PL, SCL),
    (NODES.replace("NumTerminals : 1", "NumTerminals : 99"), PL, SCL),
    (NODES, PL.replace("a 1 2 : N", "a : N"), SCL),
    (NODES, PL, SCL.replace("Coordinate : 0", "Coordinate : nan")),
    (NODES, PL, SCL.replace("SubrowOrigin : 1 Numsites : 4", "SubrowOrigin : 1e308 Numsites : 1e308")),
    (NODES, PL, "UCLA scl 1.0\nCoreRow Horizontal\n"),
    ("", "", ""),
])
def test_bookshelf_reader_on_malformed_nodes_pl_scl(checker, tmp_path, nodes, pl, scl):
    d = write_design(tmp_path, nodes=nodes, nets=NETS, pl=pl, scl=scl, wts=WTS)
    out = run(checker, "read", d, tmp_path / "placed.pl")
    assert "read" in out or "error" in out


def test_bookshelf_reader_on_ibm01(checker, tmp_path):
    src = ROOT / "tests" / "golden" / "bookshelf" / "ibm01"
    d = tmp_path / "ibm01"
    shutil.copytree(src, d)
    out = run(checker, "read", d, tmp_path / "placed.pl")
    assert "read 12752 vertices" in out


def test_bookshelf_reader_on_a_truncated_gzip(checker, tmp_path):
    d = write_design(tmp_path, gz=("nets",), nodes=NODES, nets=NETS, pl=PL, scl=SCL, wts=WTS)
    path = d / "tiny.nets.gz"
    path.write_bytes(path.read_bytes()[:20])
    out = run(checker, "read", d, tmp_path / "placed.pl")
    assert "read" in out or "error" in out

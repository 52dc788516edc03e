"""Bookshelf reader and placement writer (host C++ in libktplace_amd), CPU only.

Small designs written here check each rule the reader follows
(src/adaptor/bookshelfToKTAdaptor.cc, src/datamodel/kt_die.cc:45-103); the
ICCAD04 ibm01 benchmark (tests/golden/bookshelf/ibm01, as the reference's
benchmark/ directory holds it) checks sizes and, through an independent
parse of its files here, the pin offsets and nets: the HPWL of its .pl
placement computed from the raw files equals the oracle's HPWL of the design
the reader built.
"""
import gzip
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from ktplace_amd import design as dz

IBM01 = Path(__file__).parent / "golden" / "bookshelf" / "ibm01"

NODES = """UCLA nodes 1.0
# comment line
NumNodes : 5
NumTerminals : 1
   a   2  1
   b   4  1
   m   3  2
   io  1  1 terminal
   lone 1 1
"""

NETS = """UCLA nets 1.0
NumNets : 4
NumPins : 10
NetDegree : 3 alpha
   a I : 0.5 0.25
   b O : -1 0
   io I
NetDegree : 2
   b I : 0 0
   m O : 1 -1
NetDegree : 2 alpha
   lone I : 0 0
   a O : 0 0
NetDegree : 2 beta
   a I : 0 0
   ghost O : 0 0
   m I : 0 0
"""

PL = """UCLA pl 1.0

a 1 2 : N
b 5 0 : N
m 10 3 : N /FIXED
io 0 9 : N
"""

SCL = """UCLA scl 1.0
NumRows : 2
CoreRow Horizontal
  Coordinate : 0
  Height : 1
  Sitewidth : 1
  Sitespacing : 2
  SubrowOrigin : 1 Numsites : 4
  SubrowOrigin : 12 NumSites : 3
End
CoreRow Horizontal
  Coordinate : 1
  Height : 1
  Sitewidth : 1
  SubrowOrigin : 0 NumSites : 0
End
CoreRow Horizontal
  Coordinate : 2
  Height : 1
  Sitewidth : 1
  Sitespacing : 1
  SubrowOrigin : -3 NumSites : 5
End
"""

WTS = """UCLA wts 1.0
alpha 2.5
beta 0.5
"""


def write_design(root: Path, name="tiny", gz=(), **files):
    d = root / name
    d.mkdir(parents=True, exist_ok=True)
    for ext, text in files.items():
        path = d / f"{name}.{ext}"
        if ext in gz:
            with gzip.open(str(path) + ".gz", "wt") as f:
                f.write(text)
        else:
            path.write_text(text)
    return d


def pins_of(d, net):
    """(vertex, off_x, off_y) of a flat net."""
    f = d.flat
    s, e = int(f.net_start[net]), int(f.net_start[net + 1])
    return [(int(f.pin_cell[k]), float(f.pin_off_x[k]), float(f.pin_off_y[k])) for k in range(s, e)]


@pytest.fixture
def tiny(tmp_path):
    return dz.read_bookshelf(write_design(tmp_path, nodes=NODES, nets=NETS, pl=PL, scl=SCL, wts=WTS))


def test_cells_terminals_and_fixed(tiny):
    d = tiny
    assert d.meta["names"] == ["a", "b", "m", "io", "lone"]
    assert list(d.raw.is_fixed) == [0, 0, 1, 1, 0]            # m by /FIXED, io as a terminal
    assert d.n == 3 and list(d.flat.mov_vertex) == [0, 1, 4]
    assert list(d.raw.pos_x) == [1, 5, 10, 0, 0] and list(d.raw.pos_y) == [2, 0, 3, 9, 0]
    assert list(d.flat.width) == [2, 4, 1] and list(d.flat.height) == [1, 1, 1]
    # obstacles: fixed cells with an area (m and the 1 x 1 terminal)
    assert sorted(map(tuple, d.flat.obstacles.tolist())) == [(0, 9, 1, 10), (10, 3, 13, 5)]


def test_nets_offsets_weights_and_merges(tiny):
    d = tiny
    # alpha (twice: the second block adds its pins), the nameless net, beta
    assert d.flat.num_nets == 3
    assert list(d.raw.net_weight) == [2.5, 1.0, 0.5]
    # pins keep offsets from the lower-left: (dx + w/2, dy + h/2); the flat
    # net is stable-sorted by vertex keeping the first pin of a vertex
    assert pins_of(d, 0) == [(0, 1.5, 0.75), (1, 1.0, 0.5), (3, 0.5, 0.5), (4, 0.5, 0.5)]
    assert pins_of(d, 1) == [(1, 2.0, 0.5), (2, 2.5, 0.0)]
    # beta: NetDegree 2 keeps two pin lines, and the pin of the unknown cell is dropped
    assert pins_of(d, 2) == [(0, 1.0, 0.5)]


def test_rows_subrows_and_die(tiny):
    d = tiny
    r = d.raw
    # the row with only an empty subrow is dropped; pitch = sitespacing (2), else sitewidth
    assert list(r.row_y) == [0, 0, 2]
    assert list(zip(r.row_x0, r.row_x1)) == [(1, 9), (12, 18), (-3, 2)]
    assert list(r.row_pitch) == [2, 2, 1]
    # die: fixed cells (each at least 1 x 1) extended by the rows' box
    assert list(d.die) == [-3.0, 0.0, 18.0, 10.0]


def test_die_from_rows_without_fixed_cells(tmp_path):
    nodes = "UCLA nodes 1.0\nNumNodes : 2\nNumTerminals : 0\n a 1 1\n b 1 1\n"
    nets = "UCLA nets 1.0\nNetDegree : 2 n\n a I\n b O\n"
    scl = ("UCLA scl 1.0\nCoreRow Horizontal\n Coordinate : 4\n Height : 2\n Sitewidth : 1\n"
           " SubrowOrigin : 3 NumSites : 7\nEnd\n")
    d = dz.read_bookshelf(write_design(tmp_path, nodes=nodes, nets=nets, scl=scl))
    assert list(d.die) == [3.0, 4.0, 10.0, 6.0]
    assert d.n == 2 and list(d.raw.pos_x) == [0, 0]  # no .pl: positions stay at 0


def test_gzip_and_plain_read_alike(tmp_path):
    plain = dz.read_bookshelf(write_design(tmp_path / "p", nodes=NODES, nets=NETS, pl=PL, scl=SCL, wts=WTS))
    gz = dz.read_bookshelf(write_design(tmp_path / "g", nodes=NODES, nets=NETS, pl=PL, scl=SCL, wts=WTS,
                                        gz=("nodes", "nets", "pl", "scl", "wts")))
    for a, b in ((plain.flat.pin_off_x, gz.flat.pin_off_x), (plain.raw.row_x1, gz.raw.row_x1),
                 (plain.flat.net_weight, gz.flat.net_weight), (plain.die, gz.die)):
        assert np.array_equal(a, b)


def test_missing_files_fail_cleanly(tmp_path):
    with pytest.raises(RuntimeError, match="nodes"):
        dz.read_bookshelf(tmp_path / "nothing")
    with pytest.raises(RuntimeError, match="nets"):
        dz.read_bookshelf(write_design(tmp_path, nodes=NODES))


def test_write_pl_round_trip(tiny, tmp_path):
    d = tiny
    x = np.array([1.25, 7.5, 3.0 + 1.0 / 3.0])
    y = np.array([0.0, 2.0, 1e-7])
    out = tmp_path / "placed.pl"
    dz.write_pl(d, out, x, y)
    lines = out.read_text().splitlines()
    assert lines == ["a\t1.25\t0\t: N", "b\t7.5\t2\t: N", "m\t10\t3\t: N /FIXED", "io\t0\t9\t: N /FIXED",
                     "lone\t3.333333333\t1e-07\t: N"]
    # read back as the design's .pl: the movable cells land where written
    again = write_design(tmp_path, nodes=NODES, nets=NETS, scl=SCL, wts=WTS, pl=out.read_text())
    back = dz.read_bookshelf(again)
    assert np.allclose(back.raw.pos_x[[0, 1, 4]], x, rtol=1e-9) and np.allclose(back.raw.pos_y[[0, 1, 4]], y)
    assert list(back.raw.is_fixed) == [0, 0, 1, 1, 0]
    with pytest.raises(ValueError):
        dz.write_pl(d, out, x[:2], y[:2])


@pytest.fixture(scope="module")
def ibm01():
    return dz.read_bookshelf(IBM01)


def test_ibm01_sizes(ibm01):
    d = ibm01
    assert d.flat.num_vertices == 12752 and d.n == 12752 - 246
    assert d.flat.num_nets == 14111 and d.flat.num_pins == 50566
    assert len(d.raw.row_y) == 144 and set(d.raw.row_height) == {16.0}
    assert int((d.flat.height > 16).sum()) == 246   # movable cells taller than a row
    assert len(d.flat.obstacles) == 246


def parse_raw(root: Path, name: str):
    """Independent restatement: cells (w, h), lower-left positions, and each
    net's pin positions as centre + offset, from the files alone."""
    def lines(ext):
        with gzip.open(root / f"{name}.{ext}.gz", "rt") as f:
            return [l.split("#")[0].split() for l in f]
    size, pos = {}, {}
    body = False
    for t in lines("nodes"):
        if body and len(t) >= 3:
            size[t[0]] = (float(t[1]), float(t[2]))
        if t and t[0] == "NumTerminals":
            body = True
    for t in lines("pl"):
        if len(t) >= 3 and t[0] in size:
            pos[t[0]] = (float(t[1]), float(t[2]))
    nets, cur, left = [], None, 0
    for t in lines("nets"):
        if t and t[0] == "NetDegree":
            cur = []
            nets.append(cur)
            left = int(t[2] if t[1] == ":" else t[1])
        elif cur is not None and t and left > 0:
            left -= 1
            dx, dy = (float(t[3]), float(t[4])) if len(t) >= 5 and t[2] == ":" else (0.0, 0.0)
            (x, y), (w, h) = pos.get(t[0], (0.0, 0.0)), size[t[0]]
            cur.append((t[0], x + w / 2 + dx, y + h / 2 + dy))
    return size, pos, nets


def test_ibm01_hpwl_of_its_placement_matches_an_independent_parse(ibm01):
    d = ibm01
    size, pos, nets = parse_raw(IBM01, "ibm01")
    hpwl = 0.0
    for net in nets:
        first = {}
        for c, x, y in net:
            first.setdefault(c, (x, y))
        if len(first) < 2:
            continue
        xs = [p[0] for p in first.values()]
        ys = [p[1] for p in first.values()]
        hpwl += (max(xs) - min(xs)) + (max(ys) - min(ys))
    names = d.meta["names"]
    mov = d.flat.mov_vertex
    x = np.array([pos[names[v]][0] for v in mov])
    y = np.array([pos[names[v]][1] for v in mov])
    o = ol.Oracle.of(d)
    assert o.hpwl(x, y) == pytest.approx(hpwl, rel=1e-12)


ADAPTEC1 = Path(__file__).parent / "golden" / "bookshelf" / "adaptec1"


def test_adaptec1_loads_as_the_reference_reports_it():
    """The ISPD 2005 adaptec1 files as the reference's benchmark directory
    holds them: the reader's counts are the reference's own load line for this
    design (.github/workflows/ci.yml:150, "Loaded: 211447 cells (543
    terminals), 221142 nets, 944053 pins, 890 rows"); pins counted as the
    files list them, before each net keeps one pin per cell.  The HPWL of the
    design's own placement equals an independent parse of its files."""
    d = dz.read_bookshelf(ADAPTEC1)
    assert d.flat.num_vertices == 211447 and d.flat.num_vertices - d.n == 543
    assert d.flat.num_nets == 221142 and len(d.raw.pin_vertex) == 944053
    assert len(d.raw.row_y) == 890
    size, pos, nets = parse_raw(ADAPTEC1, "adaptec1")
    hpwl = 0.0
    for net in nets:
        first = {}
        for c, x, y in net:
            first.setdefault(c, (x, y))
        if len(first) < 2:
            continue
        xs = [q[0] for q in first.values()]
        ys = [q[1] for q in first.values()]
        hpwl += (max(xs) - min(xs)) + (max(ys) - min(ys))
    names = d.meta["names"]
    mov = d.flat.mov_vertex
    x = np.array([pos[names[v]][0] for v in mov])
    y = np.array([pos[names[v]][1] for v in mov])
    assert ol.Oracle.of(d).hpwl(x, y) == pytest.approx(hpwl, rel=1e-12)


def test_malformed_net_degrees_are_skipped(tmp_path):
    """NetDegree values that are not finite non-negative numbers (nan, inf,
    -2, 1e300, 2^64) are malformed headers: the reader skips them (their pin
    lines join the previous net, as any line after a header does) instead of
    converting them to an integer; a degree beyond the file is capped."""
    nets = """UCLA nets 1.0
NumNets : 6
NumPins : 8
NetDegree : 2 good
   a I : 0 0
   b O : 0 0
NetDegree : nan n1
NetDegree : inf n2
NetDegree : -2 n3
NetDegree : 1e300 huge
   a I : 0 0
   m O : 0 0
NetDegree : 18446744073709551616 big
   b I : 0 0
   m O : 0 0
"""
    d = dz.read_bookshelf(write_design(tmp_path, nodes=NODES, nets=nets, pl=PL, scl=SCL, wts=WTS))
    names = d.meta["names"]
    got = [sorted(names[v] for v, _, _ in pins_of(d, k)) for k in range(d.flat.num_nets)]
    assert got == [["a", "b"], ["a", "m"], ["b", "m"]]

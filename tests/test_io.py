"""Input entry points (trn_medit_*, trn_trnm_*): the text parse against the
reference's own Mesh::load (oracle/_ref/ref_harness `load`, the unmodified
src/io.cpp) and the Python restatement in meshgen, the binary format round
trip, and the errors a malformed file gives.  Host only: no GPU needed,
except the device read at the end."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import trinity_amd as T
from conftest import load_golden
from trinity_amd import meshgen as mg


def _grid4():
    g = load_golden("grid4_shock4")
    return mg.Mesh(xy=g["xy_in"], tri=g["tri"], tags=g["tags"], solut=g["solut"])


def _hub():
    m = mg.workload("hub40:fronts")
    return mg.renumber(m, np.random.default_rng(3).permutation(m.nv))


def _ref_load(tmp_path, mp, bp, nv, nt):
    out = tmp_path / "ref"
    out.mkdir()
    subprocess.run([oracle.REF_HARNESS, "load", mp, bp, str(out)], check=True, timeout=120,
                   stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="2"))

    def rd(name, dt, shape):
        return np.fromfile(str(out / name), dtype=dt).reshape(shape)
    return (rd("xy.bin", "<f8", (nv, 2)), rd("tri.bin", "<i4", (nt, 3)),
            rd("tags.bin", "u1", (nv,)), rd("solut.bin", "<f8", (nv,)))


@pytest.mark.parametrize("make", [_grid4, _hub])
def test_medit_parse_matches_reference_load(tmp_path, make):
    m = make()
    mp, bp = str(tmp_path / "m.mesh"), str(tmp_path / "m.bb")
    mg.write_medit(mp, bp, m)
    xy, tri, tags, solut = T.read_medit(mp, bp)
    # the Python restatement and the generator's own arrays
    py = mg.read_medit(mp, bp)
    assert np.array_equal(xy, py.xy) and np.array_equal(tri, py.tri)
    assert np.array_equal(tags, py.tags) and np.array_equal(solut, py.solut)
    assert np.array_equal(xy.view(np.uint64), m.xy.view(np.uint64))   # %.17g round trip
    assert np.array_equal(tri, m.tri) and np.array_equal(tags, m.tags)
    if os.path.exists(oracle.REF_HARNESS):
        rxy, rtri, rtags, rsol = _ref_load(tmp_path, mp, bp, m.nv, m.nt)
        assert np.array_equal(xy.view(np.uint64), rxy.view(np.uint64))
        assert np.array_equal(tri, rtri) and np.array_equal(tags, rtags)
        assert np.array_equal(solut.view(np.uint64), rsol.view(np.uint64))


def test_medit_parse_follows_io_find(tmp_path):
    """Comments, blank lines, leading blanks and keywords later in the file
    than a section of the same name: io::find's rules (src/io.cpp:25-48)."""
    m = mg.workload("uniform6:tanh")
    mp, bp = str(tmp_path / "m.mesh"), str(tmp_path / "m.bb")
    mg.write_medit(mp, bp, m)
    lines = open(mp).read().splitlines()
    out = []
    for k, l in enumerate(lines):
        if l == "Vertices":
            out += ["# Vertices are below", "% Triangles too", "   Vertices"]
            continue
        out.append(l)
        if k % 7 == 3 and lines[k - 1] not in ("Vertices", "Edges", "Triangles", "Corners"):
            out.append("")          # blank lines between records
    open(mp, "w").write("\n".join(out) + "\n")
    xy, tri, tags, solut = T.read_medit(mp, bp)
    assert np.array_equal(xy, m.xy) and np.array_equal(tri, m.tri)
    assert np.array_equal(tags, m.tags) and np.array_equal(solut, m.solut)
    if os.path.exists(oracle.REF_HARNESS):
        rxy, rtri, rtags, rsol = _ref_load(tmp_path, mp, bp, m.nv, m.nt)
        assert np.array_equal(xy, rxy) and np.array_equal(tri, rtri)
        assert np.array_equal(tags, rtags) and np.array_equal(solut, rsol)


def _expect_error(mp, bp):
    with pytest.raises(T.TrinityError) as e:
        T.read_medit(mp, bp)
    assert e.value.code == T.TRN_ERR_ARG
    return str(e.value)


def test_medit_parse_rejects_what_it_cannot_read_exactly(tmp_path):
    m = mg.workload("uniform4:tanh")
    mp, bp = str(tmp_path / "m.mesh"), str(tmp_path / "m.bb")
    mg.write_medit(mp, bp, m)
    good = open(mp).read()
    # two vertex records on one line
    lines = good.splitlines()
    k = lines.index("Vertices") + 2
    lines[k] = lines[k] + " " + lines[k + 1]
    del lines[k + 1]
    open(mp, "w").write("\n".join(lines) + "\n")
    assert "one per line" in _expect_error(mp, bp)
    # a triangle corner out of range
    lines = good.splitlines()
    k = lines.index("Triangles") + 2
    lines[k] = "1 2 %d 1" % (m.nv + 1)
    open(mp, "w").write("\n".join(lines) + "\n")
    assert "out of range" in _expect_error(mp, bp)
    # a truncated field file
    open(mp, "w").write(good)
    open(bp, "w").write("".join(open(bp).read().splitlines(True)[:5]))
    assert "file ends" in _expect_error(mp, bp)
    with pytest.raises(T.TrinityError):
        T.read_medit(str(tmp_path / "missing.mesh"))


def test_trnm_round_trip_on_the_host(tmp_path):
    m = _hub()
    a, b = str(tmp_path / "a.trnm"), str(tmp_path / "b.trnm")
    mg.write_trnm(a, m)                       # the generator's writer
    T.write_trnm(b, m.xy, m.tri, m.tags, m.solut)
    assert open(a, "rb").read() == open(b, "rb").read()
    assert T.trnm_info(a) == (m.nv, m.nt)
    xy, tri, tags, solut = T.read_trnm(a)
    assert np.array_equal(xy, m.xy) and np.array_equal(tri, m.tri)
    assert np.array_equal(tags, m.tags) and np.array_equal(solut, m.solut)
    _, _, _, none = T.read_trnm(a, with_solut=False)
    assert none is None
    open(b, "wb").write(open(a, "rb").read()[:-9])
    with pytest.raises(T.TrinityError):
        T.read_trnm(b)


@pytest.mark.gpu
def test_trnm_read_into_device_arrays(tmp_path, ctx):
    """Through the pinned staging buffers, several 64 MiB stages long."""
    m = mg.workload("jitter1600:fronts")
    p = str(tmp_path / "m.trnm")
    mg.write_trnm(p, m)
    xy, tri, tags, solut = T.read_trnm(p, ctx)
    assert np.array_equal(xy.cpu().numpy(), m.xy) and np.array_equal(tri.cpu().numpy(), m.tri)
    assert np.array_equal(tags.cpu().numpy(), m.tags)
    assert np.array_equal(solut.cpu().numpy(), m.solut)

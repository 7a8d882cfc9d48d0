"""Particle lists (#LSX / #LST, libpyEM/EMAN2.py:2471-2640,
libEM/io/lstfastio.cpp) and the caller loop over a list."""
import json

import numpy as np
import pytest

from eman2_amd import lst, mrc
from eman2_amd.make3d import reconstruct_lst


def test_transform_json_is_get_matrix_string_6():
    """EMAN2jsondb.py:192-196: {"__class__": "Transform", "matrix":
    get_matrix_string(6)}, twelve %1.6g values row by row
    (transform.cpp:169-178)"""
    m = np.array([[1, 0, 0, 2.5], [0, 0.5, -0.8660254, 0], [0, 0.8660254, 0.5, -1.25]])
    j = lst.transform_to_json(m)
    assert j == {"__class__": "Transform", "matrix": "[1,0,0,2.5,0,0.5,-0.866025,0,0,0.866025,0.5,-1.25]"}


def test_reads_a_hand_written_lsx(tmp_path):
    """a list laid out as LSXFile writes it: #LSX, a comment, "# L", records
    of exactly L bytes with the JSON overrides after a tab"""
    L = 140
    recs = ["0\tstack.mrcs",
            '3\tstack.mrcs\t{"score":-0.5,"xform.projection":{"__class__":"Transform",'
            '"matrix":"[0,1,0,0,-1,0,0,0,0,0,1,4]"}}']
    text = "#LSX\n# any comment\n# %d\n" % L + "".join(r.ljust(L - 1) + "\n" for r in recs)
    p = tmp_path / "a.lst"
    p.write_text(text)
    got = lst.read(str(p))
    assert got[0] == (0, "stack.mrcs", {})
    n, f, d = got[1]
    assert (n, f, d["score"]) == (3, "stack.mrcs", -0.5)
    np.testing.assert_array_equal(d["xform.projection"], [[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 1, 4]])


def test_write_read_round_trip_fixed_records(tmp_path):
    rng = np.random.default_rng(3)
    mats = rng.uniform(-1, 1, (5, 3, 4)).round(4)
    recs = [(i * 2, "p.mrcs", {"xform.projection": mats[i], "class": i % 2}) for i in range(5)] + [(11, "q.mrcs", None)]
    p = tmp_path / "b.lst"
    lst.write(str(p), recs)
    raw = p.read_text().split("\n")
    assert raw[0] == "#LSX" and raw[1].startswith("#")
    L = int(raw[2][1:])
    body = raw[3:-1]
    assert len(body) == 6 and all(len(r) + 1 == L for r in body)
    longest = max(len(r.rstrip()) for r in body)
    assert L == longest + 1 + 4                       # LSXFile.rewrite's sizing
    got = lst.read(str(p))
    for (n, f, d), i in zip(got[:5], range(5)):
        assert (n, f, d["class"]) == (2 * i, "p.mrcs", i % 2)
        np.testing.assert_allclose(d["xform.projection"], mats[i], atol=5e-6)
    assert got[5] == (11, "q.mrcs", {})
    # the stored dict is sorted compact JSON, as json.dumps(sort_keys, separators) writes it
    assert json.loads(body[0].split("\t")[2])["class"] == 0


def test_legacy_lst_and_bad_files(tmp_path):
    p = tmp_path / "c.lst"
    p.write_text("#LST\n4\tstack.mrcs\n# comment\n7\tother.mrcs\n")
    assert lst.read(str(p)) == [(4, "stack.mrcs", {}), (7, "other.mrcs", {})]
    p.write_text("#LSX\r\n# c\r\n# 20\r\n")
    with pytest.raises(ValueError, match="r"):
        lst.read(str(p))
    # a 12-byte tail in a 20-byte layout is no record (lstfastio.cpp:104-105 counts whole records)
    p.write_text("#LSX\n# c\n# 20\n" + "2\tstack.mrcs".ljust(19) + "\n0\tstack.mrcs\n")
    assert lst.read(str(p)) == [(2, "stack.mrcs", {})]
    p.write_text("P6\n")
    with pytest.raises(ValueError, match="LSX"):
        lst.read(str(p))


class _Recorder:
    """the reconstructor surface reconstruct() drives, recording its inputs"""

    def setup(self):
        self.calls = []

    def insert_batch(self, imgs, xforms, weights, defocus=None, ctf=None):
        self.calls.append((np.array(imgs), np.array(xforms), np.array(weights)))

    def finish(self, doift):
        return np.zeros((4, 4, 4), np.float32)


def test_reconstruct_from_list(tmp_path):
    """reconstruct_lst reads every listed particle from its MRC stack, in list
    order, with the record's orientation, and writes the map with the
    stack's pixel size"""
    rng = np.random.default_rng(5)
    stack = rng.standard_normal((6, 8, 8)).astype(np.float32)
    mrc.write(str(tmp_path / "s.mrcs"), stack, apix=1.7)
    order = [5, 0, 3]
    mats = [np.eye(3, 4) * (k + 1) for k in range(3)]
    lst.write(str(tmp_path / "sel.lst"), [(n, "s.mrcs", {"xform.projection": m}) for n, m in zip(order, mats)])
    r = _Recorder()
    out, info = reconstruct_lst(r, str(tmp_path / "sel.lst"), out_path=str(tmp_path / "map.mrc"), niter=1)
    imgs, xf, w = r.calls[0]
    np.testing.assert_array_equal(imgs, stack[order])
    np.testing.assert_allclose(xf, np.stack(mats), atol=1e-6)
    assert w.tolist() == [1, 1, 1] and info["ptcl_repr"] == 3
    assert abs(mrc.read_header(str(tmp_path / "map.mrc"))["apix"] - 1.7) < 1e-6
    lst.write(str(tmp_path / "bad.lst"), [(9, "s.mrcs", {"xform.projection": mats[0]})])
    with pytest.raises(IndexError):
        lst.load_particles(str(tmp_path / "bad.lst"))


def test_image_paths_relative_to_working_directory(tmp_path, monkeypatch):
    """a project's sets/x.lst names particles/y.mrcs relative to the project
    root (the working directory), as the reference opens it; a name that
    only resolves beside the list still loads"""
    rng = np.random.default_rng(6)
    (tmp_path / "particles").mkdir()
    (tmp_path / "sets").mkdir()
    stack = rng.standard_normal((4, 8, 8)).astype(np.float32)
    mrc.write(str(tmp_path / "particles" / "y.mrcs"), stack, apix=2.1)
    other = rng.standard_normal((4, 8, 8)).astype(np.float32)
    mrc.write(str(tmp_path / "sets" / "y.mrcs"), other, apix=2.1)   # a decoy beside the list
    m = np.eye(3, 4)
    lst.write(str(tmp_path / "sets" / "x.lst"), [(2, "particles/y.mrcs", {"xform.projection": m}),
                                                  (1, "particles/y.mrcs", {"xform.projection": m})])
    monkeypatch.chdir(tmp_path)
    imgs, _, _ = lst.load_particles("sets/x.lst")
    np.testing.assert_array_equal(imgs, stack[[2, 1]])
    lst.write(str(tmp_path / "sets" / "z.lst"), [(3, "y.mrcs", {"xform.projection": m})])
    monkeypatch.chdir(tmp_path / "particles")       # y.mrcs here is the working directory's
    imgs, _, _ = lst.load_particles(str(tmp_path / "sets" / "z.lst"))
    np.testing.assert_array_equal(imgs, stack[[3]])
    monkeypatch.chdir(tmp_path)                     # only beside the list
    imgs, _, _ = lst.load_particles("sets/z.lst")
    np.testing.assert_array_equal(imgs, other[[3]])

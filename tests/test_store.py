"""Saving a built H-matrix and replaying it without the host build
(bigwham_dev_amd.store; the reference's Hmat::writeToFile / readFromFile,
src/hmat/hmatrix/hmat.cc:749-935, write the same content to HDF5)."""
import json

import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HostHmat, store
from bigwham_dev_amd.meshes import rect_grid, segment_line


@pytest.fixture(scope="module")
def built():
    coor, conn = rect_grid(24)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    assert h.flat().n_lr > 0 and h.flat().n_fr > 0
    return h


def test_round_trip_is_bitwise(built, tmp_path):
    p = store.save(built, tmp_path / "h")
    f = store.load(p)
    a, b = built.flat(), f.flat()
    assert (a.dof, a.n_row_pts, a.n_col_pts, a.n_fr, a.n_lr) == \
           (b.dof, b.n_row_pts, b.n_col_pts, b.n_fr, b.n_lr)
    assert f.nb_entries() == O.nb_entries(a) == built.nb_entries()
    assert np.array_equal(O.dense(a), O.dense(b))
    x = np.random.default_rng(0).uniform(-1, 1, built.size)
    assert np.array_equal(O.matvec_original(a, x, n_threads=1), O.matvec_original(b, x, n_threads=1))
    assert np.array_equal(O.diagonal(a, original=True), O.diagonal(b, original=True))
    meta = json.loads((p / "meta.json").read_text())
    assert meta["m"] == meta["n"] == built.size and meta["dof_dimension"] == 3
    # saving the loaded file again gives identical arrays
    q = store.save(f, tmp_path / "h2")
    for name in ("frb_data.npy", "lrb_A.npy", "lrb_B.npy", "pattern_LRB.npy", "perm_cols.npy"):
        assert np.array_equal(np.load(p / name), np.load(q / name))


def test_low_rank_free_matrix(tmp_path):
    """A matrix small enough to have only full-rank blocks stores empty LR arrays."""
    coor, conn = segment_line(20)
    h = HostHmat.from_mesh("2DS0-H", coor, conn, [1.0, 0.25])
    h.build(32, 3.0, 1e-6)
    assert h.flat().n_lr == 0
    f = store.load(store.save(h, tmp_path / "s"))
    x = np.arange(h.size, dtype=float)
    assert np.array_equal(O.matvec(h.flat(), x, n_threads=1), O.matvec(f.flat(), x, n_threads=1))


def test_rejects_damaged_files(built, tmp_path):
    p = store.save(built, tmp_path / "h")
    with pytest.raises(ValueError, match="not a saved"):
        store.load(tmp_path / "missing")
    meta = json.loads((p / "meta.json").read_text())
    (p / "meta.json").write_text(json.dumps({**meta, "version": 99}))
    with pytest.raises(ValueError, match="unsupported"):
        store.load(p)
    (p / "meta.json").write_text(json.dumps(meta))
    offs = np.load(p / "lrb_A_offsets.npy")
    offs[1] += 1
    np.save(p / "lrb_A_offsets.npy", offs)
    with pytest.raises(ValueError, match="LRB A"):
        store.load(p)


def test_pattern_only_is_refused(tmp_path):
    coor, conn = rect_grid(8)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build_pattern(64, 3.0)
    with pytest.raises(ValueError, match="build it first"):
        store.save(h, tmp_path / "p")


@pytest.mark.parametrize("bad", ["float32", "big-endian", "2-D", "int perm", "float offsets"])
def test_rejects_wrong_dtypes(built, tmp_path, bad):
    """Payloads the native code would read 8 bytes at a time must be 1-D
    little-endian float64; permutations and offsets must be integers.  A file
    with the right element counts but another type is refused, not read."""
    p = store.save(built, tmp_path / "h")
    if bad == "float32":
        np.save(p / "lrb_B.npy", np.load(p / "lrb_B.npy").astype(np.float32))
    elif bad == "big-endian":
        np.save(p / "frb_data.npy", np.load(p / "frb_data.npy").astype(">f8"))
    elif bad == "2-D":
        a = np.load(p / "lrb_A.npy")
        np.save(p / "lrb_A.npy", a.reshape(1, -1))
    elif bad == "int perm":
        np.save(p / "perm_cols.npy", np.load(p / "perm_cols.npy").astype(np.float64))
    else:
        np.save(p / "frb_offsets.npy", np.load(p / "frb_offsets.npy").astype(np.float64))
    with pytest.raises(ValueError):
        store.load(p)

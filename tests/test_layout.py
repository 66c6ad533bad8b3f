"""CPU checks of the device slab layout: bounds of every work item and a
numpy replay of the kernels against the oracle (no GPU needed)."""
import numpy as np
import pytest

import oracle as O
from bigwham_dev_amd import HostHmat
from layout_emu import LayoutDump
from bigwham_dev_amd.meshes import rect_grid, segment_line


def _cases():
    h = HostHmat.gaussian(700, 10.0)
    h.build(32, 3.0, 1e-6)
    yield "gaussian", h
    coor, conn = segment_line(900)
    h = HostHmat.from_mesh("2DS0-H", coor, conn, [1.0, 0.25])
    h.build(32, 3.0, 1e-6)
    yield "2DS0", h
    coor, conn = rect_grid(16)
    h = HostHmat.from_mesh("3DR0-H", coor, conn, [1.0, 0.25])
    h.build(64, 3.0, 1e-6)
    yield "3DR0", h


CASES = dict(_cases())


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("wi_bytes", [0, 2048])
def test_layout_replay(name, precision, adjoint, wi_bytes):
    h = CASES[name]
    f = h.flat()
    lay = LayoutDump(f, precision, adjoint, wi_bytes)
    lay.check_bounds()
    x = np.random.default_rng(3).uniform(-1, 1, h.size)
    y = lay.replay(x)
    y_ref = O.rmatvec(f, x) if adjoint else O.matvec(f, x)
    tol = 1e-13 if precision == "fp64" else 1e-6
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < tol
    perm = h.permutation()
    y = lay.replay(x, perm, perm)
    y_ref = O.rmatvec_original(f, x) if adjoint else O.matvec_original(f, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < tol


def test_layout_padding_and_split():
    """Work items of one slab are split at wi_bytes and reduced in order; the
    stored entries exceed the algorithmic ones only by column padding."""
    h = CASES["3DR0"]
    f = h.flat()
    lay = LayoutDump(f, "fp64", False, 2048)
    assert len(lay.passes[1]["reds"]) > 0
    stored = sum(P["data"].size for P in lay.passes)
    E = O.nb_entries(f)
    assert E <= stored <= 1.2 * E

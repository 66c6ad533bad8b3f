"""CPU replay of the slab layout (test helper).

Exports the layout through bw_layout_* (host-only C-ABI of libbwhip.so),
checks every index each kernel would touch against its buffer, and replays
k_gather / k_slab / k_reduce in numpy, so the layout is verified on the CPU
before any kernel runs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from bigwham_dev_amd import _native as N


class LayoutDump:
    def __init__(self, flat, precision="fp64", adjoint=False, wi_bytes=0):
        lib = N.hip()
        out = C.c_void_p()
        prec = N.BW_PREC_FP64 if precision == "fp64" else N.BW_PREC_FP32
        if isinstance(flat, (list, tuple)):  # block-diagonal composite of several meshes
            flats = list(flat)
            arr = (C.POINTER(N.FlatHmat) * len(flats))(*[C.pointer(f) for f in flats])
            rc = lib.bw_layout_build_multi(arr, len(flats), prec, int(adjoint), int(wi_bytes),
                                           C.byref(out))
            n_row = sum(f.n_row_pts for f in flats)
            n_col = sum(f.n_col_pts for f in flats)
            dof = flats[0].dof
        else:
            rc = lib.bw_layout_build(C.byref(flat), prec, int(adjoint), int(wi_bytes),
                                     C.byref(out))
            n_row, n_col, dof = flat.n_row_pts, flat.n_col_pts, flat.dof
        if rc != 0:
            raise N.BigWhamError(rc, lib.bw_layout_last_error().decode())
        info = np.zeros(16, dtype=np.int64)
        lib.bw_layout_info(out, info.ctypes.data_as(C.POINTER(C.c_int64)))
        self.xt_t, self.xt_size, self.V = int(info[0]), int(info[1]), int(info[14])
        self.passes = []
        for p in range(2):
            nwi, ngidx, nred, ndata, npart, nout = (int(v) for v in info[2 + 6 * p: 8 + 6 * p])
            wis = np.zeros((max(nwi, 1), 8), dtype=np.int64)
            gidx = np.zeros(max(ngidx, 1), dtype=np.int32)
            reds = np.zeros((max(nred, 1), 4), dtype=np.int64)
            data = np.zeros(max(ndata, 1))
            lib.bw_layout_export(out, p + 1, wis.ctypes.data_as(C.POINTER(C.c_int64)),
                                 gidx.ctypes.data_as(C.POINTER(C.c_int32)),
                                 reds.ctypes.data_as(C.POINTER(C.c_int64)),
                                 data.ctypes.data_as(C.POINTER(C.c_double)))
            self.passes.append(dict(wis=wis[:nwi], gidx=gidx[:ngidx], reds=reds[:nred],
                                    data=data[:ndata], npart=npart, nout=nout))
        self.xcd = []
        for p in (1, 2):
            xs = np.zeros(9, dtype=np.int32)
            nb = C.c_int32()
            lib.bw_layout_export_xcd(out, p, xs.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nb))
            self.xcd.append((xs.astype(np.int64), nb.value))
        self.rowmap = np.zeros(max(int(info[15]), 1), dtype=np.int32)
        lib.bw_layout_export_rowmap(out, self.rowmap.ctypes.data_as(C.POINTER(C.c_int32)))
        self.rowmap = self.rowmap[:int(info[15])].astype(np.int64)
        lib.bw_layout_free(out)
        self.dof = dof
        self.n_in = dof * (n_row if adjoint else n_col)
        self.n_out = dof * (n_col if adjoint else n_row)

    def check_bounds(self):
        """Every address k_slab / k_reduce would form lies inside its buffer."""
        V = self.V
        for p, P in enumerate(self.passes):
            for (doff, gbeg, ng, R, orow, poff, _, _) in P["wis"]:
                assert 1 <= R <= 64
                assert doff >= 0 and doff + ng * R * V <= P["data"].size
                assert gbeg >= 0 and gbeg + ng <= P["gidx"].size
                if ng:
                    z = P["gidx"][gbeg:gbeg + ng].astype(np.int64)
                    assert z.min() >= 0 and z.max() + V <= self.xt_size
                if poff >= 0:
                    assert poff + R <= P["npart"]
                elif p == 0:
                    self._check_t_rows(orow, R)
                else:
                    assert orow >= 0 and orow + R <= self.n_out
            for (orow, R, poff, npart) in P["reds"]:
                assert poff + npart * R <= P["npart"]
                if p == 0:
                    self._check_t_rows(orow, R)
                else:
                    assert 0 <= orow and orow + R <= self.n_out
        # pass 1 writes every t entry exactly once
        t = self.rowmap
        assert t.size == np.unique(t).size
        assert t.size == 0 or (t.min() >= self.xt_t and t.max() < self.xt_size - 64)

    def check_xcd_cover(self):
        """The launch covers every work item exactly once: XCD x's waves
        w = xcd[x] + 4k + wave, k < blocks / 8, stop at xcd[x + 1]."""
        for p, P in enumerate(self.passes):
            xs, nb = self.xcd[p]
            nwi = len(P["wis"])
            assert xs[0] == 0 and xs[-1] == nwi and np.all(np.diff(xs) >= 0)
            assert nb % 8 == 0 and 4 * (nb // 8) >= np.diff(xs).max(initial=0)
            seen = np.zeros(nwi, dtype=int)
            for b in range(nb):
                x, k = b % 8, b // 8
                for wave in range(4):
                    w = xs[x] + 4 * k + wave
                    if w < xs[x + 1]:
                        seen[w] += 1
            assert np.all(seen == 1)

    def _check_t_rows(self, orow, R):
        """Pass-1 rows are row-map indices; their t entries lie in xt's t part."""
        assert 0 <= orow and orow + R <= self.rowmap.size
        t = self.rowmap[orow:orow + R]
        assert t.min() >= self.xt_t and t.max() < self.xt_size - 64

    def replay(self, x: np.ndarray, perm_in=None, perm_out=None, dtype=np.float64):
        """y = H x through the layout, with the kernels' arithmetic order."""
        V, d = self.V, self.dof
        xt = np.zeros(self.xt_size, dtype=dtype)
        if perm_in is not None:
            idx = (d * perm_in[:, None] + np.arange(d)).ravel()
            xt[:self.n_in] = x[idx]
        else:
            xt[:self.n_in] = x
        y = np.zeros(self.n_out)
        for p, P in enumerate(self.passes):
            part = np.zeros(P["npart"], dtype=dtype)
            for (doff, gbeg, ng, R, orow, poff, _, _) in P["wis"]:
                blk = P["data"][doff:doff + ng * R * V].reshape(ng, R, V).astype(dtype)
                z = P["gidx"][gbeg:gbeg + ng].astype(np.int64)
                zz = xt[z[:, None] + np.arange(V)[None, :]]  # (ng, V)
                acc = np.einsum("grv,gv->r", blk, zz)
                if poff >= 0:
                    part[poff:poff + R] = acc
                elif p == 0:
                    xt[self.rowmap[orow:orow + R]] = acc
                else:
                    y[orow:orow + R] = acc
            for (orow, R, poff, npart) in P["reds"]:
                s = part[poff:poff + npart * R].reshape(npart, R).sum(0)
                if p == 0:
                    xt[self.rowmap[orow:orow + R]] = s
                else:
                    y[orow:orow + R] = s
        if perm_out is not None:
            idx = (d * perm_out[:, None] + np.arange(d)).ravel()
            out = np.empty_like(y)
            out[idx] = y
            return out
        return y

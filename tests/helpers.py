"""Shared helpers of the test-suite."""
import numpy as np


def rel_l2(a, b):
    cplx = np.iscomplexobj(a) or np.iscomplexobj(b)
    a = np.asarray(a, dtype=np.complex128 if cplx else np.float64)
    b = np.asarray(b, dtype=np.complex128 if cplx else np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def random_field(mesh, which, seed, scale=1e-6):
    """Random complex point field in the canonical layout, masked like the
    reference masks displacements (imaginary alpha=0, axis, Nyquist)."""
    import axisem3d_amd._abi as abi
    rng = np.random.default_rng(seed)
    n = mesh.field_size(which)
    x = rng.standard_normal(n) * scale
    nr = mesh.array("point_nr")
    fl = mesh.array("point_flags")
    solid = which < 4
    ncomp = 3 if solid else 1
    pos = 0
    for p in range(mesh.d.npoint):
        if not (fl[p] & (abi.PT_SOLID if solid else abi.PT_FLUID)):
            continue
        nu = nr[p] // 2
        blk = x[pos:pos + (nu + 1) * ncomp * 2].reshape(nu + 1, ncomp, 2)
        blk[0, :, 1] = 0.
        if fl[p] & abi.PT_AXIAL:
            if solid:
                blk[0, 0:2, :] = 0.
                if nu >= 1:
                    blk[1, 2, :] = 0.
                    # a field that survives the axis rotation: u_phi = i u_s
                    blk[1, 1, 0] = -blk[1, 0, 1]
                    blk[1, 1, 1] = blk[1, 0, 0]
                blk[2:, :, :] = 0.
            else:
                blk[1:, :, :] = 0.
        if nr[p] % 2 == 0:
            blk[nu] = 0.
        pos += (nu + 1) * ncomp * 2
    return x


class Lifted3D:
    """The same mesh with every solid (and with fluid=True every fluid)
    element's 1-D material block repeated on its Nr azimuthal samples and
    flagged 3-D (a laterally constant 3-D model)."""

    def __init__(self, mesh, fluid=False):
        import ctypes as C
        import axisem3d_amd._abi as abi
        self.mesh = mesh
        ne = mesh.d.nelem
        flags = mesh.array("elem_flags").copy()
        off = mesh.array("mat_off")
        pool = mesh.array("mat_pool")
        enr = mesh.array("point_nr")[mesh.array("elem_points").reshape(ne, 25)].max(axis=1)
        blocks, new_off = [], [0]
        for e in range(ne):
            blk = pool[off[e]:off[e + 1]]
            if fluid or not flags[e] & abi.EL_FLUID:
                npar = blk.size // 25
                blk = np.repeat(blk.reshape(npar, 1, 25), enr[e], axis=1).ravel()
                flags[e] |= abi.EL_3D
            blocks.append(blk)
            new_off.append(new_off[-1] + blk.size)
        self._flags = flags.astype(np.int32)
        self._off = np.array(new_off, dtype=np.int64)
        self._pool = np.concatenate(blocks).astype(np.float32)
        # attenuation blocks: dkappa [1][np] then dmu [1][np] -> [nr][np] each
        aoff = mesh.array("att_off")
        apool = mesh.array("att_pool")
        ablocks, anew = [], [0]
        for e in range(ne):
            blk = apool[aoff[e]:aoff[e + 1]]
            if blk.size and (flags[e] & abi.EL_3D):
                h = blk.size // 2
                blk = np.concatenate([np.tile(blk[:h], enr[e]), np.tile(blk[h:], enr[e])])
            ablocks.append(blk)
            anew.append(anew[-1] + blk.size)
        self._aoff = np.array(anew, dtype=np.int64)
        self._apool = (np.concatenate(ablocks) if anew[-1] else np.zeros(1)).astype(np.float32)
        d = abi.MeshDesc.from_buffer_copy(mesh.d)
        d.elem_flags = self._flags.ctypes.data_as(abi.i32p)
        d.mat_off = self._off.ctypes.data_as(abi.i64p)
        d.mat_pool = self._pool.ctypes.data_as(abi.f32p)
        d.att_off = self._aoff.ctypes.data_as(abi.i64p)
        d.att_pool = self._apool.ctypes.data_as(abi.f32p)
        self._d = d
        self.desc = C.pointer(d)

    @property
    def d(self):
        return self._d

    def field_size(self, which):
        return self.mesh.field_size(which)


class LiftedMass3D:
    """The same mesh with every point's Mass1D written as Mass3D (the inverse
    mass repeated on the point's Nr slices) and every SFCoupling1D as
    SFCoupling3D (its row repeated on the slices): the same operator."""

    def __init__(self, mesh):
        import ctypes as C
        import axisem3d_amd._abi as abi
        self.mesh = mesh
        nr = mesh.array("point_nr")
        self._keep = []
        d = abi.MeshDesc.from_buffer_copy(mesh.d)
        for side in ("solid", "fluid"):
            off = mesh.array("mass_%s_off" % side)
            inv = mesh.array("inv_mass_%s" % side)
            blocks, new = [], [0]
            for p in range(mesh.d.npoint):
                blk = inv[off[p]:off[p + 1]]
                if blk.size == 1:
                    blk = np.repeat(blk, nr[p])
                blocks.append(blk)
                new.append(new[-1] + blk.size)
            o = np.array(new, dtype=np.int64)
            v = np.concatenate(blocks).astype(np.float32)
            self._keep += [o, v]
            setattr(d, "mass_%s_off" % side, o.ctypes.data_as(abi.i64p))
            setattr(d, "inv_mass_%s" % side, v.ctypes.data_as(abi.f32p))
        soff, sn, sp = mesh.array("sf_off"), mesh.array("sf_normal").reshape(-1, 6), mesh.array("sf_point")
        rows, new = [], [0]
        for q in range(mesh.d.nsf):
            blk = sn[soff[q]:soff[q + 1]]
            if blk.shape[0] == 1:
                blk = np.repeat(blk, nr[sp[q]], axis=0)
            rows.append(blk)
            new.append(new[-1] + blk.shape[0])
        o = np.array(new, dtype=np.int64)
        v = np.concatenate(rows).astype(np.float32).ravel() if rows else np.zeros(6, np.float32)
        self._keep += [o, v]
        d.sf_off = o.ctypes.data_as(abi.i64p)
        d.sf_normal = v.ctypes.data_as(abi.f32p)
        self._d = d
        self.desc = C.pointer(d)

    @property
    def d(self):
        return self._d

    def field_size(self, which):
        return self.mesh.field_size(which)


class OceanRows:
    """The same mesh with the ocean rows replaced: ocean_off [npoint+1] and rows
    [n][4] in the descriptor's layout (MassOcean1D / MassOcean3D)."""

    def __init__(self, mesh, off, rows):
        import ctypes as C
        import axisem3d_amd._abi as abi
        self.mesh = mesh
        d = abi.MeshDesc.from_buffer_copy(mesh.d)
        self._off = np.ascontiguousarray(off, dtype=np.int64)
        self._rows = np.ascontiguousarray(np.asarray(rows, dtype=np.float32).reshape(-1), dtype=np.float32)
        if self._rows.size == 0:
            self._rows = np.zeros(4, np.float32)
        d.ocean_off = self._off.ctypes.data_as(abi.i64p)
        d.ocean = self._rows.ctypes.data_as(abi.f32p)
        self._d = d
        self.desc = C.pointer(d)

    @property
    def d(self):
        return self._d

    def field_size(self, which):
        return self.mesh.field_size(which)

    def array(self, name):
        if name == "ocean_off":
            return self._off
        if name == "ocean":
            return self._rows[:4 * int(self._off[-1])]
        return self.mesh.array(name)


def lift_ocean3d(mesh):
    """MassOcean1D rows {1/(m+mo), 1/m, sin t, cos t} of a mesh rewritten as
    MassOcean3D rows {1/m, (sin t, 0, cos t) sqrt(mo / (m (m + mo)))} on each of
    the point's Nr slices (points of Nr 1 keep their one row): the same
    operator, through the FFT path."""
    off, rows = mesh.array("ocean_off"), mesh.array("ocean").reshape(-1, 4).astype(np.float64)
    nr = mesh.array("point_nr")
    new_off, new_rows = [0], []
    for p in range(mesh.d.npoint):
        blk = rows[off[p]:off[p + 1]]
        if blk.shape[0] == 1 and nr[p] > 1:
            iz, ir, st, ct = blk[0]
            m, mo = 1. / ir, 1. / iz - 1. / ir
            sc = np.sqrt(mo / (m * (m + mo)))
            blk = np.repeat([[ir, st * sc, 0., ct * sc]], nr[p], axis=0)
        new_rows.append(blk)
        new_off.append(new_off[-1] + blk.shape[0])
    return OceanRows(mesh, new_off, np.concatenate(new_rows) if new_rows else np.zeros((0, 4)))


class WithPRT:
    """The same mesh with a particle relabelling (PRT_1D) on the elements of
    `elems`: X [len(elems)][4][25] (X0..X3 per GLL point).  Those elements get
    the TIso flag, as a relabelled element always rotates (SolidElement.cpp:21)."""

    def __init__(self, mesh, elems, X):
        import ctypes as C
        import axisem3d_amd._abi as abi
        self.mesh = mesh
        d = abi.MeshDesc.from_buffer_copy(mesh.d)
        ne = mesh.d.nelem
        # one block per element: [4][25] (PRT_1D) or [rows][4][25] (PRT_3D)
        blocks = [np.asarray(x, dtype=np.float32).ravel() for x in X]
        off = np.zeros(ne + 1, dtype=np.int64)
        cnt = np.zeros(ne, dtype=np.int64)
        for e, b in zip(elems, blocks):
            cnt[e] = b.size
        off[1:] = np.cumsum(cnt)
        pool = np.zeros(max(int(off[-1]), 1), dtype=np.float32)
        for e, b in zip(elems, blocks):
            pool[off[e]:off[e] + b.size] = b
        flags = mesh.array("elem_flags").copy()
        flags[np.asarray(elems, dtype=np.int64)] |= abi.EL_TISO
        self._keep = [off, pool, flags]
        d.prt_off = off.ctypes.data_as(abi.i64p)
        d.prt_pool = pool.ctypes.data_as(abi.f32p)
        d.elem_flags = flags.ctypes.data_as(abi.i32p)
        self._d = d
        self.desc = C.pointer(d)

    @property
    def d(self):
        return self._d

    def field_size(self, which):
        return self.mesh.field_size(which)

    def array(self, name):
        if name == "elem_flags":
            return self._keep[2]
        if name == "prt_off":
            return self._keep[0]
        if name == "prt_pool":
            return self._keep[1]
        return self.mesh.array(name)


def random_prt(n, seed, amp=0.1, rows=1):
    """X0..X3 of n elements: X0, X3 near 1, X1, X2 small (a smooth undulation's
    inverse Jacobian factors), [n][4][25], or [n][rows][4][25] per slice."""
    rng = np.random.default_rng(seed)
    X = np.empty((n, rows, 4, 25), dtype=np.float32)
    X[:, :, 0] = 1. + amp * rng.uniform(-1, 1, (n, rows, 25))
    X[:, :, 1] = amp * rng.uniform(-1, 1, (n, rows, 25))
    X[:, :, 2] = amp * rng.uniform(-1, 1, (n, rows, 25))
    X[:, :, 3] = 1. + amp * rng.uniform(-1, 1, (n, rows, 25))
    return X[:, 0] if rows == 1 else X

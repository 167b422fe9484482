"""Device-resident driven and cavity problems on structured boxes.

The driven cube: an L-metre cube of n^3 boxes (6 Kuhn tetrahedra each), eps_r
constant per box and drawn with ``numpy.random.default_rng(seed)`` from a
list of values (mu_r = 1), first-order ABC on the whole boundary, and a dipole
source.  Everything (mesh, material lookup, numbering, pattern, assembly, RHS)
is built by libsucem_hip without a host copy of the mesh.
"""
from __future__ import division

import ctypes

import numpy as np

from .. import _hip
from ..Consts import c0, Z0
from ..Mesh import FunctionSpace, Mesh
from ..SystemMatrices import AssembledSystem
from ..runtime import get_context


def box_regions(n, values, seed):
    """Region id per cell: one draw per box from ``values`` (cells of a box are consecutive)."""
    rng = np.random.default_rng(seed)
    nb = int(np.prod(n))
    box = rng.integers(0, len(values), size=nb)
    return np.repeat(box, 6).astype(np.int64)


class DeviceProblem(object):
    """Mesh + space + assembled system living on the device."""

    def __init__(self, n, dims, order, which, eps_values=(1.0,), seed=0, shift=(0.0, 0.0, 0.0),
                 ctx=None, region=None):
        self.ctx = ctx or get_context()
        self.n = (int(n),) * 3 if np.isscalar(n) else tuple(int(v) for v in n)
        self.dims = np.asarray(dims, np.float64) * np.ones(3)
        self.order = order
        self.mesh = Mesh.box_on_device(self.n, scale=self.dims, shift=shift, ctx=self.ctx,
                                       export=False)
        mh = self.mesh._device[1]
        eps_values = np.asarray(eps_values, np.float64)
        if region is not None:
            self.region = np.ascontiguousarray(region, np.int64)
        else:
            self.region = box_regions(self.n, eps_values, seed) if len(eps_values) > 1 else None
        ids = np.arange(len(eps_values), dtype=np.int64)
        mu = np.ones(len(eps_values))
        _hip.call("sucem_mesh_set_materials", mh.h, _hip.ptr(self.region), len(ids), _hip.ptr(ids),
                  _hip.ptr(np.ascontiguousarray(eps_values)), _hip.ptr(mu))
        self.space = FunctionSpace(self.mesh, "Nedelec 1st kind H(curl)", order, ctx=self.ctx)
        self.system = AssembledSystem(self.space, which)
        info = self.space.info()
        self.ndofs, self.nnz, self.nnz0 = int(info[0]), int(info[1]), int(info[6])

    def rhs_dipoles(self, positions, moments, k0, rhs_index=None, nrhs=1):
        """b[r] = -j k0 Z0 sum N(x_s).p_s on the device (complex128, nrhs x N)."""
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
        val = np.ascontiguousarray(moments, np.complex128).reshape(-1, 3)
        idx = None if rhs_index is None else np.ascontiguousarray(rhs_index, np.int64)
        b = self.ctx.tensor(nrhs * self.ndofs)
        _hip.call("sucem_rhs_point_sources", self.ctx.handle, self.space.handle.h, len(pos),
                  _hip.ptr(pos), _hip.ptr(val), _hip.ptr(idx), nrhs, 0.0, -k0 * Z0, _hip.ptr(b))
        return b

    def _check_vec(self, t, dtype, nvec=1):
        """Device vectors handed to the library must have the dtype and length it will use."""
        import torch
        if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
            raise ValueError("expected a contiguous device tensor")
        if t.dtype != dtype or t.numel() < nvec * self.ndofs:
            raise ValueError("expected %s with at least %d elements, got %s with %d"
                             % (dtype, nvec * self.ndofs, t.dtype, t.numel()))

    def solve(self, k0, b, x, precision=64, tol=1e-8, max_iter=50000, method=_hip.METHOD_COCG,
              check_every=16, ru_delta=0.01, hist=None, precond=_hip.PRECOND_AUXILIARY,
              precond_shift=-1.0):
        import torch
        self._check_vec(b, torch.complex128)
        self._check_vec(x, torch.complex128)
        if hist is not None and (hist.dtype != np.float64 or not hist.flags["C_CONTIGUOUS"]):
            raise ValueError("hist must be a contiguous float64 array")
        o = _hip.SolveOpts()
        _hip.load().sucem_solve_opts_default(ctypes.byref(o))
        o.method, o.precision, o.tol, o.max_iter = method, precision, tol, max_iter
        o.check_every, o.ru_delta, o.precond = check_every, ru_delta, precond
        o.precond_shift = float(precond_shift)
        st = _hip.SolveStats()
        _hip.call("sucem_solve_driven", self.ctx.handle, self.system.handle.h, float(k0),
                  _hip.ptr(b), _hip.ptr(x), ctypes.byref(o), ctypes.byref(st), _hip.ptr(hist),
                  0 if hist is None else len(hist))
        return st.as_dict()

    def solve_coef(self, cs, ct, c0, b, x, precision=64, tol=1e-8, max_iter=50000,
                   method=_hip.METHOD_COCG, check_every=16, ru_delta=0.01,
                   precond=_hip.PRECOND_AUXILIARY, eliminate_pec=False, paired=False):
        """Solve (cs S + ct T + i c0 S0) x = b (e.g. the shifted cavity operator with the
        PEC DOFs eliminated).  ``paired``: b and x hold two real systems in their real and
        imaginary parts, solved as two independent real recursions (c0 = 0)."""
        import torch
        self._check_vec(b, torch.complex128)
        self._check_vec(x, torch.complex128)
        o = _hip.SolveOpts()
        _hip.load().sucem_solve_opts_default(ctypes.byref(o))
        o.method, o.precision, o.tol, o.max_iter = method, precision, tol, max_iter
        o.check_every, o.ru_delta, o.precond = check_every, ru_delta, precond
        o.eliminate_pec = 1 if eliminate_pec else 0
        o.paired = 1 if paired else 0
        st = _hip.SolveStats()
        _hip.call("sucem_solve_system", self.ctx.handle, self.system.handle.h, float(cs),
                  float(ct), float(c0), _hip.ptr(b), _hip.ptr(x), ctypes.byref(o),
                  ctypes.byref(st), None, 0)
        return st.as_dict()

    def solve_batch(self, k0s, B, X, tol=1e-8, max_iter=50000, check_every=16,
                    precond=_hip.PRECOND_AUXILIARY, precision=64):
        """Solve len(k0s) systems: B, X are (nb, N) complex128 device tensors.  FP64 with
        the gradient preconditioner runs them interleaved on the assembled matrix; other
        options one after another with the system's operator."""
        import torch
        nb = len(k0s)
        self._check_vec(B, torch.complex128, nb)
        self._check_vec(X, torch.complex128, nb)
        o = _hip.SolveOpts()
        _hip.load().sucem_solve_opts_default(ctypes.byref(o))
        o.method, o.precision, o.tol, o.max_iter = _hip.METHOD_COCG, precision, tol, max_iter
        o.check_every, o.precond = check_every, precond
        st = (_hip.SolveStats * nb)()
        k0 = np.ascontiguousarray(k0s, np.float64)
        _hip.call("sucem_solve_driven_batch", self.ctx.handle, self.system.handle.h, nb,
                  _hip.ptr(k0), _hip.ptr(B), _hip.ptr(X), ctypes.byref(o), st)
        return [s.as_dict() for s in st]

    def solve_sweep(self, k0s, B, X, precision=32, tol=1e-8, max_iter=50000, check_every=16,
                    width=0, use_graph=True, ru_delta=0.01, precond_shift=-1.0):
        """Solve len(k0s) driven systems together (``sucem_solve_driven_sweep``): ``width``
        interleaved COCG recursions at a time sharing each operator pass (0: the width is
        autotuned on the first systems), the iteration replayed from a hipGraph.  B, X:
        (nb, N) complex128 device tensors."""
        import torch
        nb = len(k0s)
        self._check_vec(B, torch.complex128, nb)
        self._check_vec(X, torch.complex128, nb)
        o = _hip.SolveOpts()
        _hip.load().sucem_solve_opts_default(ctypes.byref(o))
        o.method, o.precision, o.tol, o.max_iter = _hip.METHOD_COCG, precision, tol, max_iter
        o.check_every, o.precond, o.ru_delta = check_every, _hip.PRECOND_AUXILIARY, ru_delta
        o.precond_shift = float(precond_shift)
        st = (_hip.SolveStats * nb)()
        k0 = np.ascontiguousarray(k0s, np.float64)
        _hip.call("sucem_solve_driven_sweep", self.ctx.handle, self.system.handle.h, nb,
                  _hip.ptr(k0), _hip.ptr(B), _hip.ptr(X), ctypes.byref(o), int(width),
                  int(bool(use_graph)), st)
        return [s.as_dict() for s in st]

    def eigs(self, sigma, nev, ncv=0, tol=1e-10, inner_tol=1e-11, vectors=False,
             inner_precision=64, block=0):
        """Lowest nev k^2 > sigma of the PEC cavity: (k2, relative residuals, stats[, modes]).
        ``block``: Lanczos block size 1 or 2 (paired inner solves); 0 picks 2 where the
        paired solves run."""
        o = _hip.EigOpts()
        _hip.load().sucem_eig_opts_default(ctypes.byref(o))
        o.ncv, o.tol, o.inner_tol = int(ncv), tol, inner_tol
        o.inner_precision = int(inner_precision)
        o.block = int(block)
        k2 = np.zeros(nev)
        resid = np.zeros(nev)
        vecs = self.ctx.tensor(nev * self.ndofs) if vectors else None
        st = _hip.EigStats()
        _hip.call("sucem_eigs", self.ctx.handle, self.system.handle.h, float(sigma), int(nev),
                  ctypes.byref(o), _hip.ptr(k2), _hip.ptr(resid), _hip.ptr(vecs), ctypes.byref(st))
        if vectors:
            return k2, resid, st.as_dict(), vecs.reshape(nev, self.ndofs)
        return k2, resid, st.as_dict()

    def tune(self, precision, reps=10):
        """Autotune the fused operator for this system (assembled launch shapes and the
        matrix-free operator); returns ({matrix_free, lanes, loads, blocks}, ms per apply)."""
        cfg = np.zeros(4, np.int32)
        ms = ctypes.c_double()
        _hip.call("sucem_system_tune", self.ctx.handle, self.system.handle.h, int(precision),
                  int(reps), _hip.ptr(cfg), ctypes.byref(ms))
        return dict(matrix_free=int(cfg[0]), lanes=int(cfg[1]), loads=int(cfg[2]),
                    blocks=int(cfg[3])), ms.value

    def set_spmv(self, precision, matrix_free=0, lanes=16, loads=4, blocks=2048):
        """Select the operator explicitly (see ``tune``): matrix_free 0 (assembled CSR: lanes
        per row, loads per lane), 1 (per-cell element products: lanes per gather row 0, 1, 4;
        loads 1 = whole element, 2 = half the rows at a time) or 2 (patches: loads = cells
        per patch, 128, 192 or 256); blocks = grid of the row / shared-DOF kernel."""
        cfg = np.array([matrix_free, lanes, loads, blocks], np.int32)
        _hip.call("sucem_system_set_spmv", self.ctx.handle, self.system.handle.h, int(precision),
                  _hip.ptr(cfg))

    def apply(self, k0, x, y, precision=32, nvec=1):
        import torch
        dt = torch.complex64 if precision == 32 else torch.complex128
        self._check_vec(x, dt, nvec)
        self._check_vec(y, dt, nvec)
        _hip.call("sucem_system_apply", self.ctx.handle, self.system.handle.h, float(k0), precision,
                  0, nvec, _hip.ptr(x), _hip.ptr(y))

    def patch_info(self, cb):
        """{patches, patch_dofs, shared_dofs, partial_sums, lds, face_rows} of the patch
        operator with patches of ``cb`` cells."""
        info = np.zeros(6, np.int64)
        _hip.call("sucem_system_patch_info", self.ctx.handle, self.system.handle.h, int(cb),
                  _hip.ptr(info))
        return dict(zip(("patches", "patch_dofs", "shared_dofs", "partial_sums", "lds",
                         "face_rows"), (int(v) for v in info)))

    def operator_bytes(self, precision, matrix_free, epilogue=False, cb=256, deferred=False):
        """Algorithmic HBM bytes of one application of the fused operator (one vector);
        ``epilogue``: plus the read of w for the fused dot product sum_r w_r y_r;
        ``matrix_free`` 2 (patches of ``cb`` cells): per cell 12 coefficients and nloc
        (patch-DOF index, product place) pairs, per patch DOF its id, segment end,
        destination and x value, per shared DOF its id and segment pointer, the partial
        sums written and read, per boundary face 6 coefficients, its cell and local face
        and its products' places, y once.  ``deferred`` (patch form inside COCG): the shared
        DOFs' partial sums are written but summed by the x/r update, so the bytes are the
        patch kernel's: no shared-DOF segments or partial reads, y written for the DOFs of
        one patch only, and (FP32) w read at every finished row and every partial sum;
        FP64 forms the epilogue from the element quadratic forms and reads no w."""
        if deferred and matrix_free == 2:
            s = 4 if precision == 32 else 8
            sx = 2 * s
            nc = int(np.prod(self.n)) * 6
            nbf = int(self.space.info()[4])
            nloc, nfd = (6, 3) if self.order == 1 else (20, 8)
            pi = self.patch_info(cb)
            own = self.ndofs - pi["shared_dofs"]
            return (nc * (12 * s + nloc * 4) + pi["patch_dofs"] * (4 + 2 + 4 + sx)
                    + pi["partial_sums"] * sx + nbf * (6 * s + 4 + nfd * 2)
                    + 2 * (pi["patches"] + 1) * 4 + own * sx
                    + ((own + pi["partial_sums"]) * sx if epilogue and precision == 32
                       else 0))
        if epilogue:
            return (self.operator_bytes(precision, matrix_free, cb=cb)
                    + self.ndofs * (8 if precision == 32 else 16))
        if not matrix_free:
            return self.spmv_bytes(precision)
        if matrix_free == 2:
            s = 4 if precision == 32 else 8
            sx = 2 * s
            nc = int(np.prod(self.n)) * 6
            nbf = int(self.space.info()[4])
            nloc, nfd = (6, 3) if self.order == 1 else (20, 8)
            pi = self.patch_info(cb)
            return (nc * (12 * s + nloc * 4) + pi["patch_dofs"] * (4 + 2 + 4 + sx)
                    + pi["shared_dofs"] * 8 + 2 * pi["partial_sums"] * sx
                    + nbf * (6 * s + 4 + nfd * 2) + 2 * (pi["patches"] + 1) * 4
                    + self.ndofs * sx)
        s = 4 if precision == 32 else 8            # real value
        sx = 2 * s                                  # complex vector element
        n = self.ndofs
        nc = int(np.prod(self.n)) * 6
        nbf = int(self.space.info()[4])
        nloc, nfd = (6, 3) if self.order == 1 else (20, 8)
        cells = nc * (12 * s + nloc * 4 + nloc * 4)   # coefficients, DOF ids, product positions
        faces = nbf * (6 * s + 8 + 2 * nfd * 4)       # coefficients, cell + local face, DOFs, positions
        products = 2 * (nc * nloc + nbf * nfd) * sx   # products written, then read
        return cells + faces + products + (n + 1) * 4 + 2 * n * sx   # segments; x once, y once

    def assembly_bytes(self):
        """Algorithmic HBM bytes of the S + T assembly gather (``k_gather_volume``): the
        segment pointers (nnz + 1 int32), every contribution id of the (row, col)-sorted
        gather map once (nc * nloc^2 uint32), the 12 FP64 geometric coefficients of each cell
        once, and the FP64 and FP32 (S, T) pairs written per nonzero (16 + 8 bytes)."""
        nc = int(np.prod(self.n)) * 6
        nloc = 6 if self.order == 1 else 20
        return 4 * (self.nnz + 1) + 4 * nc * nloc * nloc + 96 * nc + 24 * self.nnz

    def spmv_bytes(self, precision):
        """Algorithmic HBM bytes of one assembled fused SpMV (one vector)."""
        sv = 8 if precision == 32 else 16          # (S, T) pair
        s0v = 4 if precision == 32 else 8
        sx = 8 if precision == 32 else 16          # complex vector element
        n, nnz, nnz0 = self.ndofs, self.nnz, self.nnz0
        return (nnz * (4 + sv) + (n + 1) * 4 + nnz0 * (4 + s0v) + (n + 1) * 4 + 2 * n * sx)


def dielectric_block(n, seed, frac=(0.1, 0.3)):
    """A box-aligned block of whole boxes occupying a volume fraction in ``frac`` of an
    (nx, ny, nz) box grid, drawn with ``numpy.random.default_rng(seed)``: per axis an
    extent of round(a n_i) boxes (at least 1), a ~ U[0.3, 0.9], redrawn until the fraction
    fits, at a uniform offset.  Returns (region per cell: 1 inside, 0 outside; the block's
    (lo, hi) box indices; its volume fraction).  Cells of box (ix, iy, iz) are cells
    6 b .. 6 b + 5 with b = ix + nx (iy + ny iz)."""
    n = np.asarray(n, np.int64)
    rng = np.random.default_rng(seed)
    for _ in range(10000):
        ext = np.maximum(1, np.rint(rng.uniform(0.3, 0.9, size=3) * n)).astype(np.int64)
        f = float(np.prod(ext / n))
        if frac[0] <= f <= frac[1]:
            break
    else:
        raise ValueError("no block with a volume fraction in %r on a %r grid" % (frac, tuple(n)))
    lo = np.array([rng.integers(0, n[i] - ext[i] + 1) for i in range(3)], np.int64)
    hi = lo + ext
    ix, iy, iz = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    inside = ((ix >= lo[0]) & (ix < hi[0]) & (iy >= lo[1]) & (iy < hi[1])
              & (iz >= lo[2]) & (iz < hi[2]))
    box = inside.transpose(2, 1, 0).reshape(-1)           # box index ix + nx (iy + ny iz)
    return np.repeat(box.astype(np.int64), 6), (lo, hi), f


def dielectric_cavity(n=(120, 60, 90), dims=(1.0, 0.5, 0.75), order=2, eps_block=6.0, seed=0,
                      ctx=None):
    """The large eigenproblem: a PEC box cavity of n boxes with a seeded box-aligned
    dielectric block (eps_r = eps_block, 10-30 % of the volume), S and T assembled."""
    region, block, f = dielectric_block(n, seed)
    prob = DeviceProblem(n, dims, order, _hip.MAT_S | _hip.MAT_T, (1.0, float(eps_block)),
                         seed, ctx=ctx, region=region)
    prob.block, prob.block_fraction = block, f
    return prob


def sweep_sources(nfreq=64, ndip=8, L=1.0, f_lo=0.3e9, f_hi=1.5e9, seed=0):
    """The frequency sweep's inputs: nfreq frequencies uniform in [f_lo, f_hi] and, per
    frequency, ndip unit dipoles with directions uniform on the sphere and positions uniform
    in the inner 60 % of the cube (drawn with ``numpy.random.default_rng(seed)``).  Returns
    (freqs (nfreq,), positions (nfreq, ndip, 3), moments (nfreq, ndip, 3))."""
    rng = np.random.default_rng(seed)
    freqs = np.linspace(f_lo, f_hi, nfreq)
    pos = rng.uniform(0.2 * L, 0.8 * L, size=(nfreq, ndip, 3))
    d = rng.standard_normal(size=(nfreq, ndip, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    return freqs, pos, d


def sweep_rhs(prob, freqs, pos, moments):
    """Right-hand sides of the sweep on the device: (nfreq * ndip, N) complex128, system
    f * ndip + s = dipole s at frequency f, and the k0 of each system."""
    nfreq, ndip = pos.shape[:2]
    k0s = np.repeat(2 * np.pi * np.asarray(freqs) / c0, ndip)
    blocks = []
    for f in range(nfreq):
        blocks.append(prob.rhs_dipoles(pos[f], moments[f].astype(np.complex128), k0s[f * ndip],
                                       rhs_index=np.arange(ndip), nrhs=ndip))
    import torch
    return torch.cat(blocks).contiguous(), k0s


def driven_cube(n=40, L=1.0, order=2, freq=1e9, eps_values=(1.0, 2.2, 4.4), seed=0, ctx=None):
    """The driven-cube workload: z dipole of moment lambda/1000 at the cube centre."""
    prob = DeviceProblem(n, L, order, _hip.MAT_S | _hip.MAT_T | _hip.MAT_S0, eps_values, seed,
                         ctx=ctx)
    lam = c0 / freq
    prob.k0 = 2 * np.pi * freq / c0
    prob.source_pos = np.array([[L / 2, L / 2, L / 2]])
    prob.source_moment = np.array([[0, 0, lam / 1000]], np.complex128)
    return prob

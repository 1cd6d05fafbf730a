"""Python binding of the C ABI with the reference's class and method names.

Mirrors ``VarLayerSampler`` (src/core/VarLayerSampler.h:8-37), ``RCWAIntegrator``
(src/core/RCWAIntegrator.h:8-40), ``RedhefferIntegrator``
(src/core/RedhefferIntegrator.h:10-90) and ``DifferentialIntegrator``
(src/core/DifferentialIntegrator.h:8-40).  Matrices returned to Python are numpy
complex128 arrays; the computation itself runs in ``libvarrcwa.so`` on the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, List, Optional, Sequence

import numpy as np

from ._lib import (RULE_FN, SAMPLE_DEVICE_FN, SAMPLE_HOST_FN, Context, VrcError, VrcLayer, VrcRegion,
                   VrcStackInstance, c_double_p, cplx, from_dev_colmajor, load, ptr, to_dev_colmajor)

__all__ = ["Context", "Layer", "LayerSampler", "VarLayerSampler", "StackLayerSampler", "RCWAIntegrator", "RedhefferIntegrator",
           "DifferentialIntegrator", "batch_stack_rcwa", "StackTape", "assemble_toeplitz", "assemble_pq", "eig_pq",
           "layer_smatrix", "star_product", "project_left", "redheffer_segment", "assemble_toeplitz_vjp",
           "assemble_pq_vjp", "eig_pq_vjp", "layer_smatrix_vjp", "star_product_vjp", "project_left_vjp",
           "VrcError"]

_DEFAULT_CTX: Optional[Context] = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


class Layer:
    """Layer(x, y, eps) (src/rcwa/Layer.h:25-27); eps has shape (len(y)-1, len(x)-1)."""

    def __init__(self, x, y, eps):
        self.x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        self.y = np.ascontiguousarray(np.asarray(y, dtype=np.float64))
        e = np.asarray(eps, dtype=np.complex128)
        if e.shape != (self.y.size - 1, self.x.size - 1):
            raise ValueError(f"eps shape {e.shape} does not match coordinates "
                             f"({self.y.size - 1}, {self.x.size - 1})")
        self.eps = np.asfortranarray(e)
        self._eps_flat = np.ascontiguousarray(self.eps.reshape(-1, order="F")).view(np.float64)
        self.struct = VrcLayer(self.x.size, self.x.ctypes.data_as(c_double_p), self.y.size,
                               self.y.ctypes.data_as(c_double_p), self._eps_flat.ctypes.data_as(c_double_p))


def _as_layer(l) -> Layer:
    return l if isinstance(l, Layer) else Layer(*l)


def _layer_array(layers: Sequence[Layer]):
    arr = (VrcLayer * max(1, len(layers)))()
    for i, l in enumerate(layers):
        arr[i] = l.struct
    return arr


def _host_mat(buf: np.ndarray, n: int) -> np.ndarray:
    """Column-major host buffer of n*n complex -> numpy matrix."""
    return buf.reshape(n, n).T.copy()


class _Sampler:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = C.c_void_p()
        self._keep: List = []

    def nDim(self) -> int:
        return int(self.ctx.lib.vrc_sampler_ndim(self.h))

    def sample(self, z: float):
        """LayerSampler::sample -> (P, Q) as numpy arrays."""
        N = self.nDim()
        P = np.empty(N * N, np.complex128)
        Q = np.empty(N * N, np.complex128)
        self.ctx.check(self.ctx.lib.vrc_sampler_sample(self.h, float(z), P.ctypes.data_as(c_double_p),
                                                       Q.ctypes.data_as(c_double_p)), "sample")
        return _host_mat(P, N), _host_mat(Q, N)

    def sampleOnGPU(self, z: float, P, Q):
        """LayerSampler::sampleOnGPU into device tensors (column-major N x N)."""
        self.ctx.check(self.ctx.lib.vrc_sampler_sample_gpu(self.h, float(z), ptr(P), ptr(Q)), "sampleOnGPU")

    def _K(self):
        n = self.nDim() // 2
        Kx = np.empty(n * n, np.complex128)
        Ky = np.empty(n * n, np.complex128)
        self.ctx.check(self.ctx.lib.vrc_sampler_k(self.h, Kx.ctypes.data_as(c_double_p),
                                                  Ky.ctypes.data_as(c_double_p)), "K")
        return _host_mat(Kx, n), _host_mat(Ky, n)

    def Kx(self):
        return self._K()[0]

    def Ky(self):
        return self._K()[1]

    def close(self):
        if self.h:
            self.ctx.lib.vrc_sampler_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VarLayerSampler(_Sampler):
    """VarLayerSampler(lambda, nx, ny, x, y, eps, enablePML=True)."""

    def __init__(self, lam, nx, ny, x, y, eps, enablePML=True, ctx: Optional[Context] = None):
        super().__init__(ctx or default_context())
        self.lam, self.nx, self.ny = float(lam), int(nx), int(ny)
        self.base = Layer(x, y, eps)
        self.ctx.check(self.ctx.lib.vrc_var_sampler_create(self.ctx.h, self.lam, self.nx, self.ny,
                                                           C.byref(self.base.struct), int(bool(enablePML)),
                                                           C.byref(self.h)), "VarLayerSampler")

    def _rule(self, axis: int, layout: int, rule: Callable[[float], float]):
        cb = RULE_FN(lambda z, user: float(rule(z)))
        self._keep.append(cb)
        self.ctx.check(self.ctx.lib.vrc_sampler_add_rule(self.h, axis, int(layout), cb, None), "addRule")

    def addXRule(self, layout: int, rule: Callable[[float], float]):
        self._rule(0, layout, rule)

    def addYRule(self, layout: int, rule: Callable[[float], float]):
        self._rule(1, layout, rule)


class LayerSampler(_Sampler):
    """Base class of a caller's own sampler -- the reference's abstract LayerSampler
    (src/core/LayerSampler.h:5-77). Subclass it, override ``sample(z) -> (P, Q)`` (numpy
    N x N complex) and ``assignSimulationRegion()`` (set Lx_, Ly_, b0_ .. r1_), optionally
    ``sampleOnGPU(z, P, Q)`` (fill the device tensors, column-major N x N, on the context's
    stream), and call ``postInitialization()`` at the end of ``__init__``. The integrators
    then call back into the subclass for every P and Q."""

    def __init__(self, lam, nx, ny, enablePML=True, ctx: Optional[Context] = None):
        super().__init__(ctx or default_context())
        self.lam, self.nx, self.ny, self.enablePML = float(lam), int(nx), int(ny), bool(enablePML)
        self.Lx_ = self.Ly_ = self.b0_ = self.b1_ = self.u0_ = self.u1_ = 0.0
        self.l0_ = self.l1_ = self.r0_ = self.r1_ = 0.0

    def assignSimulationRegion(self):
        raise NotImplementedError("a LayerSampler subclass assigns its simulation region")

    def lambda_(self):
        return self.lam

    def k0(self):
        return 2 * 3.1415926535897932 / self.lam

    def Lx(self):
        return self.Lx_

    def Ly(self):
        return self.Ly_

    def postInitialization(self):
        self.assignSimulationRegion()
        r = VrcRegion(self.Lx_, self.Ly_, self.b0_, self.b1_, self.u0_, self.u1_, self.l0_, self.l1_, self.r0_,
                      self.r1_)
        N = 2 * self.nx * self.ny

        def host(z, Pp, Qp, user):
            try:
                P, Q = self.sample(z)
                for src, dst in ((P, Pp), (Q, Qp)):
                    a = np.asarray(src, dtype=np.complex128)
                    if a.shape != (N, N):
                        return 3
                    out = np.ctypeslib.as_array(C.cast(dst, C.POINTER(C.c_double)), shape=(2 * N * N,))
                    out[:] = np.asfortranarray(a).reshape(-1, order="F").view(np.float64)
                return 0
            except Exception:
                return 2

        def device(z, Pd, Qd, stream, user):
            try:
                return 1 if self.sampleOnGPU(z, Pd, Qd) is NotImplemented else 0
            except Exception:
                return 2

        self._cbs = (SAMPLE_HOST_FN(host), SAMPLE_DEVICE_FN(device))
        self.ctx.check(self.ctx.lib.vrc_callback_sampler_create(
            self.ctx.h, self.lam, self.nx, self.ny, C.byref(r), int(self.enablePML), self._cbs[0], self._cbs[1],
            None, C.byref(self.h)), "postInitialization")

    def sample(self, z: float):
        raise NotImplementedError("a LayerSampler subclass computes P and Q in sample(z)")

    def sampleOnGPU(self, z: float, P, Q):
        """Default: not provided (the integrators call sample() and copy). An override
        receives raw device pointers (ints) of the N x N column-major complex buffers."""
        return NotImplemented


class StackLayerSampler(_Sampler):
    """Piecewise-constant stack of Layer slabs between half-space media."""

    def __init__(self, lam, nx, ny, layers, z_edges, bottom, top, enablePML=False, ctx: Optional[Context] = None):
        super().__init__(ctx or default_context())
        self.layers = [_as_layer(l) for l in layers]
        self.bottom, self.top = _as_layer(bottom), _as_layer(top)
        self.z_edges = np.ascontiguousarray(np.asarray(z_edges, np.float64))
        arr = _layer_array(self.layers)
        self._keep.append(arr)
        self.ctx.check(self.ctx.lib.vrc_stack_sampler_create(
            self.ctx.h, float(lam), int(nx), int(ny), len(self.layers), arr,
            self.z_edges.ctypes.data_as(c_double_p), C.byref(self.bottom.struct), C.byref(self.top.struct),
            int(bool(enablePML)), C.byref(self.h)), "StackLayerSampler")


class GDSIISampler(_Sampler):
    """GDSIISampler (src/core/GDSIISampler.cpp:17-146): x-z cross-sections of the first cell's
    polygons in a GDSII file (micrometers), PML always on. Defaults follow
    src/core/GDSIISampler.h:27-35 (permittivities: SiO2 1.445^2 background, Si 3.48^2 device)."""

    def __init__(self, lam, nx, ny, filename, shift=(0.0, 0.0, 0.0), lPML=0.5, rPML=0.5, lspace=1.5, rspace=1.5,
                 bPML=0.5, uPML=0.5, bspace=1.08, uspace=1.0, permBackground=1.445 * 1.445,
                 permDevice=3.48 * 3.48, thickness=0.22, ctx: Optional[Context] = None):
        super().__init__(ctx or default_context())
        sh = np.ascontiguousarray(shift, dtype=np.float64)
        ps = np.ascontiguousarray([lPML, rPML, lspace, rspace, bPML, uPML, bspace, uspace], dtype=np.float64)
        eb = np.array([complex(permBackground).real, complex(permBackground).imag])
        ed = np.array([complex(permDevice).real, complex(permDevice).imag])
        self.ctx.check(self.ctx.lib.vrc_gds_sampler_create(
            self.ctx.h, float(lam), int(nx), int(ny), str(filename).encode(), sh.ctypes.data, ps.ctypes.data,
            eb.ctypes.data, ed.ctypes.data, float(thickness), C.byref(self.h)), "GDSIISampler")

    def _bounds(self):
        b = np.zeros(4)
        self.ctx.check(self.ctx.lib.vrc_gds_sampler_bounds(self.h, b.ctypes.data), "bounds")
        return b

    def minX(self):
        return float(self._bounds()[0])

    def maxX(self):
        return float(self._bounds()[1])

    def minZ(self):
        return float(self._bounds()[2])

    def maxZ(self):
        return float(self._bounds()[3])

    def sampleLayout(self, z: float):
        """(x, y, eps) of the cross-section at z; eps[i, j] for y strip i and x cell j."""
        nxc, nyc = C.c_int(0), C.c_int(0)
        lib = self.ctx.lib
        self.ctx.check(lib.vrc_sampler_layout(self.h, float(z), 0, None, C.byref(nxc), 0, None, C.byref(nyc), None),
                       "sampleLayout")
        x = np.zeros(nxc.value)
        y = np.zeros(nyc.value)
        e = np.zeros((nyc.value - 1) * (nxc.value - 1), dtype=np.complex128)
        self.ctx.check(lib.vrc_sampler_layout(self.h, float(z), nxc.value, x.ctypes.data, C.byref(nxc), nyc.value,
                                              y.ctypes.data, C.byref(nyc), e.ctypes.data), "sampleLayout")
        return x, y, e.reshape((nyc.value - 1, nxc.value - 1), order="F")


class RCWASolver:
    """The legacy stacked-layer solver RCWASolver (src/rcwa/RCWASolver.h:60-80): addLayer,
    solve(lambda, layerStack, thickness, translation=None), then Tdd/Rud and the
    transmittance/reflectance helpers (src/rcwa/RCWASolver.cpp:582-670). nX, nY are
    harmonic orders (nx = 2 nX + 1); PML always on."""

    def __init__(self, nX: int, nY: int, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.h = C.c_void_p()
        self._layers = []
        self.ctx.check(self.ctx.lib.vrc_rcwa_solver_create(self.ctx.h, int(nX), int(nY), C.byref(self.h)),
                       "RCWASolver")

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self.ctx.lib.vrc_rcwa_solver_destroy(self.h)
            self.h = C.c_void_p()

    def addLayer(self, x, y=None, eps=None):
        L = _as_layer(x if y is None else (x, y, eps))
        self._layers.append(L)
        self.ctx.check(self.ctx.lib.vrc_rcwa_solver_add_layer(self.h, C.byref(L.struct)), "addLayer")

    def nLayerTypes(self):
        return len(self._layers)

    def solve(self, lam: float, layerStack, thickness, translation=None):
        st = np.ascontiguousarray(layerStack, dtype=np.int32)
        th = np.ascontiguousarray(thickness, dtype=np.float64)
        tr = None if translation is None else np.ascontiguousarray(translation, dtype=np.float64)
        if tr is not None and tr.size != st.size:
            raise ValueError("translation needs one value per stack entry")
        self.ctx.check(self.ctx.lib.vrc_rcwa_solver_solve(self.h, float(lam), st.size, st.ctypes.data, th.ctypes.data,
                                                          None if tr is None else tr.ctypes.data), "solve")

    def _dims(self):
        N, t = C.c_int(), C.c_int()
        self.ctx.lib.vrc_rcwa_solver_dims(self.h, C.byref(N), C.byref(t))
        return N.value

    def _get(self, which, k=0):
        N = self._dims()
        cnt = N if which == 4 else N * N
        buf = np.empty(cnt, np.complex128)
        self.ctx.check(self.ctx.lib.vrc_rcwa_solver_get(self.h, which, int(k), buf.ctypes.data), "get")
        return buf if which == 4 else buf.reshape((N, N), order="F")

    def Tdd(self):
        return self._get(0)

    def Rud(self):
        return self._get(1)

    def eigvecE(self, layerType):
        return self._get(2, layerType)

    def eigvecH(self, layerType):
        return self._get(3, layerType)

    def gamma(self, layerType):
        return self._get(4, layerType)

    def evaluateIntermediateField(self, input_vec, nstack: int):
        """RCWASolver::evaluateIntermediateField (src/rcwa/RCWASolver.cpp:1373-1621) for the
        stack of the last solve (nstack entries): input_vec (N,) or (N, k) incident amplitudes
        of the first stack entry, or an int (inputMode: that unit vector). Returns (c_m, c_p),
        lists of nstack - 2 arrays shaped like input_vec."""
        N = self._dims()
        if isinstance(input_vec, (int, np.integer)):
            v = np.zeros(N, np.complex128)
            v[int(input_vec)] = 1.0
        else:
            v = np.asarray(input_vec, np.complex128)
        one = v.ndim == 1
        V = np.asfortranarray(v.reshape(N, -1))
        k = V.shape[1]
        nin = int(nstack) - 2
        cm = np.empty((max(nin, 1), N, k), np.complex128, order="C")
        cp = np.empty_like(cm)
        bufm = np.empty(max(nin, 1) * N * k, np.complex128)
        bufp = np.empty_like(bufm)
        dp = C.POINTER(C.c_double)
        self.ctx.check(self.ctx.lib.vrc_rcwa_solver_intermediate_field(
            self.h, k, V.ctypes.data_as(dp), bufm.ctypes.data_as(dp), bufp.ctypes.data_as(dp)),
            "evaluateIntermediateField")
        for j in range(nin):
            cm[j] = bufm[j * N * k:(j + 1) * N * k].reshape((N, k), order="F")
            cp[j] = bufp[j * N * k:(j + 1) * N * k].reshape((N, k), order="F")
        out_m = [cm[j][:, 0] if one else cm[j].copy() for j in range(nin)]
        out_p = [cp[j][:, 0] if one else cp[j].copy() for j in range(nin)]
        return out_m, out_p

    def tdd(self, row, col):
        return complex(self.Tdd()[row, col])

    def rud(self, row, col):
        return complex(self.Rud()[row, col])

    def transmittance(self, inputMode, outputMode=None):
        """|Tdd(out, in)|^2, or the column's power sum without outputMode (:582-594, :632-650)."""
        col = self.Tdd()[:, inputMode]
        return float(np.sum(np.abs(col) ** 2)) if outputMode is None else float(abs(col[outputMode]) ** 2)

    def reflectance(self, inputMode, outputMode=None):
        col = self.Rud()[:, inputMode]
        return float(np.sum(np.abs(col) ** 2)) if outputMode is None else float(abs(col[outputMode]) ** 2)


def gds_polygons(filename, unit: float = 1e-6, cell: int = 0):
    """The polygons of one cell of a GDSII file as the library reads them (gdstk::read_gds,
    src/gdstk/library.cpp:160-420): a list of [(x, y), ...] in units of `unit` meters. Host only."""
    lib = load()
    n = C.c_int(0)
    fn = str(filename).encode()
    rc = lib.vrc_gds_polygons(fn, float(unit), int(cell), 0, C.byref(n), None, 0, None)
    if rc != 0:
        raise VrcError(f"gds_polygons failed ({rc}) for {filename}")
    counts = np.zeros(max(n.value, 1), dtype=np.int32)
    lib.vrc_gds_polygons(fn, float(unit), int(cell), n.value, C.byref(n), counts.ctypes.data, 0, None)
    total = int(counts[:n.value].sum())
    xy = np.zeros(2 * max(total, 1))
    if lib.vrc_gds_polygons(fn, float(unit), int(cell), n.value, C.byref(n), counts.ctypes.data, total,
                            xy.ctypes.data) != 0:
        raise VrcError(f"gds_polygons failed for {filename}")
    out, o = [], 0
    for k in range(n.value):
        out.append([(float(xy[2 * (o + i)]), float(xy[2 * (o + i) + 1])) for i in range(int(counts[k]))])
        o += int(counts[k])
    return out


_GETTERS = ["Tdd", "Tuu", "Rud", "Rdu", "W0", "V0", "gam0", "W1", "V1", "gam1"]


class _Integrator:
    KIND = 0

    def __init__(self, z0: float, z1: float, sampler: _Sampler):
        self.ctx = sampler.ctx
        self.sampler = sampler
        self.h = C.c_void_p()
        self.ctx.check(self.ctx.lib.vrc_integrator_create(self.ctx.h, self.KIND, float(z0), float(z1), sampler.h,
                                                          C.byref(self.h)), "integrator")

    def _get(self, which: int):
        N = self.sampler.nDim()
        n = N if which in (6, 9) else N * N
        buf = np.empty(n, np.complex128)
        self.ctx.check(self.ctx.lib.vrc_integrator_get(self.h, which, buf.ctypes.data_as(c_double_p)), "get")
        return buf if which in (6, 9) else _host_mat(buf, N)

    def counts(self):
        a, b = C.c_int(), C.c_int()
        self.ctx.lib.vrc_integrator_counts(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def N(self):
        return self.counts()[0]

    def nEigVal(self):
        return self.counts()[1]

    def close(self):
        if self.h:
            self.ctx.lib.vrc_integrator_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


for _i, _name in enumerate(_GETTERS):
    setattr(_Integrator, _name, (lambda i: lambda self: self._get(i))(_i))


class RCWAIntegrator(_Integrator):
    """RCWAIntegrator(z0, z1, sampler).compute(N)."""
    KIND = 0

    def compute(self, N: int):
        self.ctx.check(self.ctx.lib.vrc_rcwa_compute(self.h, int(N)), "RCWAIntegrator.compute")

    def compute_batched(self, N: int):
        self.ctx.check(self.ctx.lib.vrc_rcwa_compute_batched(self.h, int(N)), "RCWAIntegrator.compute_batched")

    def compute_edges(self, z_edges):
        ze = np.ascontiguousarray(np.asarray(z_edges, np.float64))
        self.ctx.check(self.ctx.lib.vrc_rcwa_compute_edges(self.h, ze.size, ze.ctypes.data_as(c_double_p)),
                       "RCWAIntegrator.compute_edges")


class RedhefferIntegrator(_Integrator):
    """RedhefferIntegrator(z0, z1, sampler).compute(p, maxN=10, pe=1e-2)."""
    KIND = 1

    def compute(self, p: int, maxN: int = 10, pe: float = 1e-2):
        self.ctx.check(self.ctx.lib.vrc_redheffer_compute(self.h, int(p), int(maxN), float(pe)),
                       "RedhefferIntegrator.compute")

    @staticmethod
    def compute_many(integrators, p: int, maxN: int = 10, pe: float = 1e-2):
        """compute(p, maxN, pe) of several integrators (one context, equal N) in one call;
        each new reference eigensystem level is solved in one batched launch."""
        its = list(integrators)
        if not its:
            return
        arr = (C.c_void_p * len(its))(*[it.h for it in its])
        ctx = its[0].ctx
        ctx.check(ctx.lib.vrc_redheffer_compute_many(len(its), arr, int(p), int(maxN), float(pe)),
                  "RedhefferIntegrator.compute_many")

    def tree(self):
        n = self.ctx.lib.vrc_integrator_tree(self.h, 0, None)
        buf = np.empty(2 * max(n, 1), np.float64)
        self.ctx.lib.vrc_integrator_tree(self.h, n, buf.ctypes.data_as(c_double_p))
        return [(buf[2 * i], buf[2 * i + 1]) for i in range(n)]


class DifferentialIntegrator(_Integrator):
    """DifferentialIntegrator(z0, z1, sampler).compute(N)."""
    KIND = 2

    def compute(self, N: int):
        self.ctx.check(self.ctx.lib.vrc_differential_compute(self.h, int(N)), "DifferentialIntegrator.compute")


# ------------------------------------------------------------------ batched stacks
def _stack_array(instances):
    keep = []
    arr = (VrcStackInstance * max(1, len(instances)))()
    for i, inst in enumerate(instances):
        ls = [_as_layer(l) for l in inst["layers"]]
        la = _layer_array(ls)
        th = np.ascontiguousarray(np.asarray(inst["thickness"], np.float64))
        bot, top = _as_layer(inst["bottom"]), _as_layer(inst["top"])
        keep += [ls, la, th, bot, top]
        arr[i] = VrcStackInstance(float(inst["lam"]), len(ls), la, th.ctypes.data_as(c_double_p),
                                  C.pointer(bot.struct), C.pointer(top.struct))
    return arr, keep


class StackTape:
    """Reverse-pass state of one ``batch_stack_rcwa(..., tape=True)`` call."""

    def __init__(self, ctx: Context, h: C.c_void_p):
        self.ctx, self.h = ctx, h

    def cells(self) -> int:
        """Length of the permittivity gradient (all slab cells of all instances)."""
        return int(self.ctx.lib.vrc_tape_cells(self.h))

    def vjp(self, S_bar, chunk: int = 0):
        """Cotangent of every slab eps cell (device complex128 vector, instance then
        slab order, each slab's cells column-major) for the cotangent S_bar of S."""
        torch = self.ctx.torch
        S_bar = S_bar.contiguous()
        out = torch.empty(self.cells(), dtype=torch.complex128, device=S_bar.device)
        self.ctx.check(self.ctx.lib.vrc_batch_stack_vjp(self.h, ptr(S_bar), ptr(out), int(chunk)), "batch_stack_vjp")
        return out

    def bytes(self) -> int:
        """Device memory the tape holds."""
        return int(self.ctx.lib.vrc_tape_bytes(self.h))

    def vjp_shape(self, S_bar, shape, chunk: int = 0, want_eps: bool = False):
        """Gradient with respect to shape parameters of the slabs' x coordinates.

        shape: one entry per slab (instance-major, then layer): None, or (var_index,
        coeff) with coeff of shape (len(var_index), npara) (coordinate var_index[k] moves
        by coeff[k, p] per unit of parameter p). Returns (per-slab list of numpy arrays of
        npara gradients (empty for None), eps cotangents or None)."""
        torch = self.ctx.torch
        S_bar = S_bar.contiguous()
        npara, nvar, vidx, coef = [], [], [], []
        for e in shape:
            if e is None:
                npara.append(0)
                nvar.append(0)
                continue
            vi, cf = e
            cf = np.asarray(cf, dtype=np.float64).reshape(len(vi), -1)
            npara.append(cf.shape[1])
            nvar.append(len(vi))
            vidx.extend(int(v) for v in vi)
            coef.extend(cf.reshape(-1).tolist())
        np_a = np.ascontiguousarray(npara, dtype=np.int32)
        nv_a = np.ascontiguousarray(nvar, dtype=np.int32)
        vi_a = np.ascontiguousarray(vidx or [0], dtype=np.int32)
        cf_a = np.ascontiguousarray(coef or [0.0], dtype=np.float64)
        out = np.zeros(max(1, int(np_a.sum())), dtype=np.float64)
        eps = torch.empty(self.cells(), dtype=torch.complex128, device=S_bar.device) if want_eps else None
        ip = C.POINTER(C.c_int)
        dp = C.POINTER(C.c_double)
        self.ctx.check(self.ctx.lib.vrc_batch_stack_vjp_shape(
            self.h, ptr(S_bar), np_a.ctypes.data_as(ip), nv_a.ctypes.data_as(ip), vi_a.ctypes.data_as(ip),
            cf_a.ctypes.data_as(dp), out.ctypes.data_as(dp), ptr(eps) if want_eps else None, int(chunk)),
            "batch_stack_vjp_shape")
        per, o = [], 0
        for k in npara:
            per.append(out[o:o + k].copy())
            o += k
        return per, eps

    def close(self):
        if self.h:
            self.ctx.lib.vrc_tape_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_stack_rcwa(ctx: Context, nx: int, ny: int, instances, chunk: int = 0, want_bases: bool = False,
                     tape: bool = False, return_info: bool = False, recompute: bool = False, layer_block: int = 0):
    """instances: list of dicts with keys lam, layers (Layer list), thickness, bottom, top.

    Returns a dict of device tensors: S (ninst, 4, N, N) as column-major storage
    (use ``from_dev_colmajor``), W0, V0, gam0, W1, V1, gam1 when requested, and a
    StackTape under "tape" when tape=True (recompute=True: the tape keeps no slab
    eigen-systems, the reverse pass solves them again; layer_block: layers per block of
    the reverse pass, 0 = automatic). A numerical failure of any instance (an
    unconverged eigensolve, NaN or infinity in its S-matrix) raises VrcError naming the
    instance, unless return_info=True: then the per-instance codes (0 ok, 1 eigensolver,
    4 non-finite) are returned under "info" (numpy int array) and nothing is raised for them.
    """
    torch = ctx.torch
    n = nx * ny
    N = 2 * n
    arr, keep = _stack_array(instances)
    dev = f"cuda:{ctx.device}"
    S = torch.empty((len(instances), 4, N, N), dtype=torch.complex128, device=dev)
    out = {"S": S}
    ptrs = [None] * 6
    if want_bases:
        for k, name in enumerate(["W0", "V0", "gam0", "W1", "V1", "gam1"]):
            shape = (len(instances), N) if name.startswith("gam") else (len(instances), N, N)
            out[name] = torch.empty(shape, dtype=torch.complex128, device=dev)
            ptrs[k] = ptr(out[name])
    info = np.zeros(max(1, len(instances)), dtype=np.int32)
    info_p = info.ctypes.data_as(C.POINTER(C.c_int))
    if tape:
        h = C.c_void_p()
        rc = ctx.lib.vrc_batch_stack_forward_ex(ctx.h, nx, ny, len(instances), arr, int(chunk), ptr(S), *ptrs,
                                                int(bool(recompute)), int(layer_block), C.byref(h), info_p)
        if h:
            out["tape"] = StackTape(ctx, h)
        what = "batch_stack_forward"
    else:
        rc = ctx.lib.vrc_batch_stack_rcwa(ctx.h, nx, ny, len(instances), arr, int(chunk), ptr(S), *ptrs, info_p)
        what = "batch_stack_rcwa"
    if not (return_info and rc > 0):
        ctx.check(rc, what)
    if return_info:
        out["info"] = info[:len(instances)].copy()
    del keep
    return out


class BatchPlan:
    """A recorded batch (vrc_batch_plan_*): the work of batch_stack_rcwa(ctx, nx, ny,
    instances, chunk, want_bases) recorded once as a HIP graph; launch() replays it into
    the plan's output tensors (returned as the same dict batch_stack_rcwa gives)."""

    def __init__(self, ctx: Context, nx: int, ny: int, instances, chunk: int = 0, want_bases: bool = False):
        torch = ctx.torch
        self.ctx = ctx
        self.ninst = len(instances)
        N = 2 * nx * ny
        dev = f"cuda:{ctx.device}"
        self.out = {"S": torch.empty((self.ninst, 4, N, N), dtype=torch.complex128, device=dev)}
        ptrs = [None] * 6
        if want_bases:
            for k, name in enumerate(["W0", "V0", "gam0", "W1", "V1", "gam1"]):
                shape = (self.ninst, N) if name.startswith("gam") else (self.ninst, N, N)
                self.out[name] = torch.empty(shape, dtype=torch.complex128, device=dev)
                ptrs[k] = ptr(self.out[name])
        arr, keep = _stack_array(instances)
        h = C.c_void_p()
        torch.cuda.synchronize()
        ctx.check(ctx.lib.vrc_batch_plan_create(ctx.h, nx, ny, self.ninst, arr, int(chunk), ptr(self.out["S"]), *ptrs,
                                                C.byref(h)), "batch_plan_create")
        del keep
        self.h = h

    def launch(self, return_info: bool = False):
        info = np.zeros(max(1, self.ninst), dtype=np.int32)
        rc = self.ctx.lib.vrc_batch_plan_launch(self.h, info.ctypes.data_as(C.POINTER(C.c_int)))
        if not (return_info and rc > 0):
            self.ctx.check(rc, "batch_plan_launch")
        out = dict(self.out)
        if return_info:
            out["info"] = info[:self.ninst].copy()
        return out

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.vrc_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_stack_rcwa_mixed(ctx: Context, harmonics, instances, chunk: int = 0, want_bases: bool = False,
                           return_info: bool = False):
    """batch_stack_rcwa with a harmonic count per instance: harmonics[i] = (nx, ny).

    Returns a dict with "S": a list of per-instance device tensors (4, N_i, N_i)
    (column-major storage, views into one packed buffer), the end bases as lists when
    requested, and "info" when return_info=True (see batch_stack_rcwa)."""
    torch = ctx.torch
    B = len(instances)
    if len(harmonics) != B:
        raise ValueError("one (nx, ny) pair per instance expected")
    nxa = np.ascontiguousarray([int(h[0]) for h in harmonics] or [1], dtype=np.int32)
    nya = np.ascontiguousarray([int(h[1]) for h in harmonics] or [1], dtype=np.int32)
    Ns = [2 * int(a) * int(b) for a, b in zip(nxa[:B], nya[:B])]
    dev = f"cuda:{ctx.device}"
    Sflat = torch.empty(max(1, sum(4 * N * N for N in Ns)), dtype=torch.complex128, device=dev)
    mats = {k: torch.empty(max(1, sum(N * N for N in Ns)), dtype=torch.complex128, device=dev)
            for k in (["W0", "V0", "W1", "V1"] if want_bases else [])}
    vecs = {k: torch.empty(max(1, sum(Ns)), dtype=torch.complex128, device=dev)
            for k in (["gam0", "gam1"] if want_bases else [])}
    ptrs = [ptr(mats["W0"]), ptr(mats["V0"]), ptr(vecs["gam0"]), ptr(mats["W1"]), ptr(mats["V1"]),
            ptr(vecs["gam1"])] if want_bases else [None] * 6
    arr, keep = _stack_array(instances)
    info = np.zeros(max(1, B), dtype=np.int32)
    rc = ctx.lib.vrc_batch_stack_rcwa_mixed(ctx.h, B, nxa.ctypes.data_as(C.POINTER(C.c_int)),
                                            nya.ctypes.data_as(C.POINTER(C.c_int)), arr, int(chunk), ptr(Sflat),
                                            *ptrs, info.ctypes.data_as(C.POINTER(C.c_int)))
    del keep
    if not (return_info and rc > 0):
        ctx.check(rc, "batch_stack_rcwa_mixed")

    def split(flat, per):
        out, o = [], 0
        for N in Ns:
            out.append(flat[o:o + per(N)])
            o += per(N)
        return out

    out = {"S": [t.view(4, N, N) for t, N in zip(split(Sflat, lambda N: 4 * N * N), Ns)]}
    for k, t in mats.items():
        out[k] = [u.view(N, N) for u, N in zip(split(t, lambda N: N * N), Ns)]
    for k, t in vecs.items():
        out[k] = split(t, lambda N: N)
    if return_info:
        out["info"] = info[:B].copy()
    return out


def stack_efficiencies(ctx: Context, nx: int, ny: int, instances, solved, e_inc):
    """Plane-wave diffraction efficiencies (RCWASolver::scatterPlaneWave,
    src/rcwa/RCWASolver.cpp:1636-1733) of stacks solved by ``batch_stack_rcwa`` with
    ``want_bases=True``. e_inc: device complex128 (ninst, 2 nx ny), the incident field
    harmonics [Ex; Ey] in the bottom medium, travelling up. Returns (REF, TRN), device
    float64 (ninst, nx ny): efficiency per order of the reflected and transmitted field."""
    torch = ctx.torch
    n = nx * ny
    B = len(instances)
    if "W0" not in solved or "W1" not in solved:
        raise ValueError("stack_efficiencies needs the end bases (batch_stack_rcwa(..., want_bases=True))")
    if tuple(e_inc.shape) != (B, 2 * n) or e_inc.dtype != torch.complex128 or not e_inc.is_contiguous():
        raise ValueError("e_inc must be a contiguous complex128 tensor of shape (ninst, 2 nx ny)")
    for name in ("S", "W0", "W1"):
        if solved[name].shape[0] != B:
            raise ValueError(f"{name} holds {solved[name].shape[0]} instances, expected {B}")
    arr, keep = _stack_array(instances)
    dev = f"cuda:{ctx.device}"
    REF = torch.empty((B, n), dtype=torch.float64, device=dev)
    TRN = torch.empty((B, n), dtype=torch.float64, device=dev)
    ctx.check(ctx.lib.vrc_stack_efficiencies(ctx.h, nx, ny, B, arr, ptr(solved["S"]), ptr(solved["W0"]),
                                             ptr(solved["W1"]), ptr(e_inc), ptr(REF), ptr(TRN)),
              "stack_efficiencies")
    del keep
    return REF, TRN


def modal_power(ctx: Context, nx: int, ny: int, W, V, Lx: float, Ly: float, nres_x: int, nres_y: int,
                x0: float, x1: float, y0: float, y1: float):
    """Power of modes through a window, evaluateEnergy (src/tests/test_cpu.cpp:335-412).

    W, V: device complex128 tensors (nmodes, 2 nx ny); row b holds mode b's field harmonics
    [Ex; Ey] and [Hx; Hy] (a column of the integrator's W/V; the column-major device storage
    of batch_stack_rcwa's W0[k] is exactly this layout). Returns device float64 (nmodes,)."""
    torch = ctx.torch
    n2 = 2 * nx * ny
    for name, t in (("W", W), ("V", V)):
        if t.dim() != 2 or t.shape[1] != n2 or t.dtype != torch.complex128 or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous complex128 tensor of shape (nmodes, 2 nx ny)")
    if W.shape[0] != V.shape[0]:
        raise ValueError("W and V must hold the same number of modes")
    energy = torch.empty((W.shape[0],), dtype=torch.float64, device=f"cuda:{ctx.device}")
    ctx.check(ctx.lib.vrc_modal_power(ctx.h, nx, ny, W.shape[0], ptr(W), n2, ptr(V), n2, float(Lx), float(Ly),
                                      int(nres_x), int(nres_y), float(x0), float(x1), float(y0), float(y1),
                                      ptr(energy)), "modal_power")
    return energy


def find_guided_modes(ctx: Optional[Context], gamma, k0: float, imag_coeff: float, re_min: float, re_max: float):
    """findGuidedModes (src/tests/test_cpu.cpp:311-333): [(index, gamma / k0^2)] of the modes with
    re_min <= Re <= re_max and Im < imag_coeff, by decreasing real part. Host only: ctx may be
    None (no device is touched)."""
    g = np.ascontiguousarray(np.asarray(gamma, dtype=np.complex128))
    n = g.shape[0]
    idx = np.zeros(max(n, 1), dtype=np.int32)
    neff = np.zeros(max(n, 1), dtype=np.complex128)
    cnt = C.c_int(0)
    lib = ctx.lib if ctx is not None else load()
    rc = lib.vrc_find_guided_modes(ctx.h if ctx is not None else None, n, g.ctypes.data, float(k0),
                                   float(imag_coeff), float(re_min), float(re_max), idx.ctypes.data,
                                   neff.ctypes.data, C.byref(cnt))
    if ctx is not None:
        ctx.check(rc, "find_guided_modes")
    elif rc != 0:
        raise VrcError(f"find_guided_modes failed ({rc}): bad arguments")
    return [(int(idx[k]), complex(neff[k])) for k in range(cnt.value)]


# ------------------------------------------------------------------ stage entry points
def assemble_toeplitz(ctx: Context, nx: int, ny: int, layers):
    torch = ctx.torch
    ls = [_as_layer(l) for l in layers]
    n = nx * ny
    fe = torch.empty((3, len(ls), n, n), dtype=torch.complex128, device=f"cuda:{ctx.device}")
    arr = _layer_array(ls)
    ctx.check(ctx.lib.vrc_assemble_toeplitz(ctx.h, nx, ny, len(ls), arr, ptr(fe[0]), ptr(fe[1]), ptr(fe[2]),
                                            n * n), "assemble_toeplitz")
    return fe


def assemble_pq(ctx: Context, nx: int, ny: int, layers, lam: float, enable_pml: bool, region):
    torch = ctx.torch
    ls = [_as_layer(l) for l in layers]
    reg = _as_layer(region)
    N = 2 * nx * ny
    P = torch.empty((len(ls), N, N), dtype=torch.complex128, device=f"cuda:{ctx.device}")
    Q = torch.empty_like(P)
    arr = _layer_array(ls)
    ctx.check(ctx.lib.vrc_assemble_pq(ctx.h, nx, ny, len(ls), arr, float(lam), int(bool(enable_pml)),
                                      C.byref(reg.struct), ptr(P), ptr(Q), N * N), "assemble_pq")
    return P, Q


def wave_eqn_derivatives(ctx: Context, nx: int, ny: int, layer, lam: float, enable_pml: bool, region,
                         para_x: dict, npara: int = 0):
    """Shape and wavelength derivatives of P and Q (Layer::waveEqnCoeff's dP, dQ, dPl, dQl,
    src/rcwa/Layer.cpp:36-162). para_x: Parameterization {x-coordinate index: coefficients}
    (src/rcwa/Parameterization.h:5-25; missing coefficients are 0). Returns device tensors in
    column-major storage (read with from_dev_colmajor): dP, dQ (npara, N, N), dfe11, dfe22,
    dfe33 (npara, n, n), dPl, dQl (N, N)."""
    torch = ctx.torch
    npara = max(int(npara), max((len(v) for v in para_x.values()), default=0))
    n, N = nx * ny, 2 * nx * ny
    keys = sorted(int(k) for k in para_x)
    idx = np.ascontiguousarray(keys, dtype=np.int32)
    coef = np.zeros((len(keys), max(npara, 1)), dtype=np.float64)
    for r, k in enumerate(keys):
        v = np.asarray(para_x[k], dtype=np.float64)
        coef[r, :v.size] = v
    coef = np.ascontiguousarray(coef[:, :npara])
    dev = f"cuda:{ctx.device}"
    out = {name: torch.empty((npara, N, N), dtype=torch.complex128, device=dev) for name in ("dP", "dQ")}
    for name in ("dfe11", "dfe22", "dfe33"):
        out[name] = torch.empty((npara, n, n), dtype=torch.complex128, device=dev)
    for name in ("dPl", "dQl"):
        out[name] = torch.empty((N, N), dtype=torch.complex128, device=dev)
    L, R = _as_layer(layer), _as_layer(region)
    ctx.check(ctx.lib.vrc_wave_eqn_derivatives(ctx.h, nx, ny, C.byref(L.struct), float(lam), int(bool(enable_pml)),
                                               C.byref(R.struct), len(keys), idx.ctypes.data, coef.ctypes.data,
                                               npara, ptr(out["dP"]), ptr(out["dQ"]), ptr(out["dfe11"]),
                                               ptr(out["dfe22"]), ptr(out["dfe33"]), ptr(out["dPl"]),
                                               ptr(out["dQl"])), "wave_eqn_derivatives")
    return out


def eig_pq(ctx: Context, P, Q):
    torch = ctx.torch
    B, N, _ = P.shape
    W = torch.empty_like(P)
    V = torch.empty_like(P)
    g = torch.empty((B, N), dtype=torch.complex128, device=P.device)
    info = torch.zeros(B, dtype=torch.int32, device=P.device)
    ctx.check(ctx.lib.vrc_eig_pq(ctx.h, N, B, ptr(P), ptr(Q), N * N, ptr(W), ptr(V), ptr(g), ptr(info)), "eig_pq")
    return W, V, g, info


def layer_smatrix(ctx: Context, lam, dz, W, V, gam, W0, V0, shared_ref: bool = True):
    torch = ctx.torch
    B, N, _ = W.shape
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float64), (B,)))
    dz = np.ascontiguousarray(np.broadcast_to(np.asarray(dz, np.float64), (B,)))
    Tuu = torch.empty_like(W)
    Rud = torch.empty_like(W)
    ctx.check(ctx.lib.vrc_layer_smatrix(ctx.h, N, B, lam.ctypes.data_as(c_double_p), dz.ctypes.data_as(c_double_p),
                                        ptr(W), ptr(V), ptr(gam), ptr(W0), ptr(V0), 0 if shared_ref else N * N,
                                        ptr(Tuu), ptr(Rud)), "layer_smatrix")
    return Tuu, Rud


def star_product(ctx: Context, S, s):
    """S, s: tensors (4, B, N, N) [dd, uu, ud, du] (S updated in place)."""
    _, B, N, _ = S.shape
    ctx.check(ctx.lib.vrc_star_product(ctx.h, N, B, ptr(S[0]), ptr(S[1]), ptr(S[2]), ptr(S[3]), ptr(s[0]), ptr(s[1]),
                                       ptr(s[2]), ptr(s[3]), N * N), "star_product")


def project_left(ctx: Context, S, W1, V1, W0, V0):
    _, B, N, _ = S.shape
    ctx.check(ctx.lib.vrc_project_left(ctx.h, N, B, ptr(S[0]), ptr(S[1]), ptr(S[2]), ptr(S[3]), ptr(W1), ptr(V1),
                                       ptr(W0), ptr(V0), N * N), "project_left")


def redheffer_segment(ctx: Context, lam, z0, z1, Pr, Qr, W, V, gam, P3, Q3):
    torch = ctx.torch
    N = W.shape[-1]
    outs = [torch.empty((N, N), dtype=torch.complex128, device=W.device) for _ in range(4)]
    err = C.c_double()
    ctx.check(ctx.lib.vrc_redheffer_segment(ctx.h, N, float(lam), float(z0), float(z1), ptr(Pr), ptr(Qr), ptr(W),
                                            ptr(V), ptr(gam), ptr(P3[0]), ptr(P3[1]), ptr(P3[2]), ptr(Q3[0]),
                                            ptr(Q3[1]), ptr(Q3[2]), *[ptr(o) for o in outs], C.byref(err)),
              "redheffer_segment")
    return outs, err.value


# ------------------------------------------------------------------ stage adjoints
def _cells(layers) -> int:
    return sum(l.eps.size for l in layers)


def assemble_toeplitz_vjp(ctx: Context, nx: int, ny: int, layers, fe_bar):
    """fe_bar: (3, L, n, n) cotangents of (fe11, fe22, fe33) in the storage layout of
    assemble_toeplitz -> eps cotangents (all layers' cells, column-major, concatenated)."""
    torch = ctx.torch
    ls = [_as_layer(l) for l in layers]
    n = nx * ny
    fe_bar = fe_bar.contiguous()
    out = torch.empty(_cells(ls), dtype=torch.complex128, device=fe_bar.device)
    ctx.check(ctx.lib.vrc_assemble_toeplitz_vjp(ctx.h, nx, ny, len(ls), _layer_array(ls), ptr(fe_bar[0]),
                                                ptr(fe_bar[1]), ptr(fe_bar[2]), n * n, ptr(out)),
              "assemble_toeplitz_vjp")
    return out


def assemble_pq_vjp(ctx: Context, nx: int, ny: int, layers, lam: float, region, P_bar, Q_bar):
    torch = ctx.torch
    ls = [_as_layer(l) for l in layers]
    reg = _as_layer(region)
    N = 2 * nx * ny
    P_bar, Q_bar = P_bar.contiguous(), Q_bar.contiguous()
    out = torch.empty(_cells(ls), dtype=torch.complex128, device=P_bar.device)
    ctx.check(ctx.lib.vrc_assemble_pq_vjp(ctx.h, nx, ny, len(ls), _layer_array(ls), float(lam), C.byref(reg.struct),
                                          ptr(P_bar), ptr(Q_bar), N * N, ptr(out)), "assemble_pq_vjp")
    return out


def eig_pq_vjp(ctx: Context, P, Q, W, V, gam, W_bar, V_bar, gam_bar):
    torch = ctx.torch
    B, N, _ = P.shape
    Wb, gb = W_bar.clone().contiguous(), gam_bar.clone().contiguous()
    Pb, Qb = torch.empty_like(P), torch.empty_like(Q)
    ctx.check(ctx.lib.vrc_eig_pq_vjp(ctx.h, N, B, ptr(P), ptr(Q), N * N, ptr(W), ptr(V), ptr(gam), ptr(Wb),
                                     ptr(V_bar.contiguous()), ptr(gb), ptr(Pb), ptr(Qb)), "eig_pq_vjp")
    return Pb, Qb


def layer_smatrix_vjp(ctx: Context, lam, dz, W, V, gam, W0, V0, Tuu_bar, Rud_bar, shared_ref: bool = True):
    torch = ctx.torch
    B, N, _ = W.shape
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float64), (B,)))
    dz = np.ascontiguousarray(np.broadcast_to(np.asarray(dz, np.float64), (B,)))
    Wb, Vb = torch.empty_like(W), torch.empty_like(V)
    gb = torch.empty_like(gam)
    ctx.check(ctx.lib.vrc_layer_smatrix_vjp(ctx.h, N, B, lam.ctypes.data_as(c_double_p),
                                            dz.ctypes.data_as(c_double_p), ptr(W), ptr(V), ptr(gam), ptr(W0),
                                            ptr(V0), 0 if shared_ref else N * N, ptr(Tuu_bar.contiguous()),
                                            ptr(Rud_bar.contiguous()), ptr(Wb), ptr(Vb), ptr(gb)),
              "layer_smatrix_vjp")
    return Wb, Vb, gb


def star_product_vjp(ctx: Context, S0, s, S_bar):
    """S0, s, S_bar: (4, B, N, N) [dd, uu, ud, du]. Returns (cotangent of S0, of s)."""
    _, B, N, _ = S0.shape
    Sb = S_bar.clone().contiguous()
    sb = ctx.torch.empty_like(S0)
    ctx.check(ctx.lib.vrc_star_product_vjp(ctx.h, N, B, *[ptr(S0[k]) for k in range(4)], *[ptr(s[k]) for k in range(4)],
                                           N * N, *[ptr(Sb[k]) for k in range(4)], *[ptr(sb[k]) for k in range(4)]),
              "star_product_vjp")
    return Sb, sb


def project_left_vjp(ctx: Context, S, W1, V1, W0, V0, S_bar):
    _, B, N, _ = S.shape
    Sb = S_bar.clone().contiguous()
    ctx.check(ctx.lib.vrc_project_left_vjp(ctx.h, N, B, *[ptr(S[k]) for k in range(4)], ptr(W1), ptr(V1), ptr(W0),
                                           ptr(V0), N * N, *[ptr(Sb[k]) for k in range(4)]), "project_left_vjp")
    return Sb

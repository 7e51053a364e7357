"""rooftop_mom_amd: the +rooftopMoM object interface (Mesh, Problem, Operator, Preconditioner, Excitation, Solver,
Result) over librooftop.so, the gfx950 HIP library of the iterative-solve path.

This package directory is named `rooftop-mom_amd`; it is imported under the module name `rooftop_mom_amd` (see
load() in this repository's __graft_entry__.py). Every numerical step of the path -- generators, spectra, products,
preconditioners, Krylov loops -- runs in librooftop.so on the GPU; without the library every class raises.

Names, argument meanings, defaults and error identifiers follow the reference MATLAB package rooftopMoM 3.5.0:
  Mesh           src/+rooftopMoM/Mesh.m:1-63
  Problem        src/+rooftopMoM/Problem.m:25-206
  Operator       src/+rooftopMoM/Operator.m:7-54 and Product.m:25-57
  Preconditioner src/+rooftopMoM/Preconditioner.m:7-70 (kinds here: 'nearField', 'jacobi')
  Excitation     src/+rooftopMoM/Excitation.m:29-117
  Solver         src/+rooftopMoM/Solver.m:17-169
  Result         src/+rooftopMoM/Result.m:14-94
Additions: 'precision' ('c128', 'c64', 'mixed') on products, and 'precision'/'refine' on the solver (complex64
inner loop inside an FP64 refinement loop on the true residual).
"""

import ctypes as C
import math

import numpy as np

from . import _abi
from ._abi import RooftopError, check, library

Z0 = 376.730313668          # src/+rooftopMoM/+assembler/Constants.m:11
SPEED_OF_LIGHT = 299792458.0

_PRECISIONS = {'c128': _abi.RT_C128, 'c64': _abi.RT_C64, 'mixed': _abi.RT_MIXED}
_PRODUCTS = {'fft': _abi.RT_PRODUCT_FFT, 'splitFFT': _abi.RT_PRODUCT_SPLIT_FFT,
             'leanSplitFFT': _abi.RT_PRODUCT_LEAN_SPLIT_FFT}
_MODIFIERS = {None: _abi.RT_MODIFIER_NONE, 'real': _abi.RT_MODIFIER_REAL, 'imag': _abi.RT_MODIFIER_IMAG}
_PRECONDITIONERS = {'jacobi': _abi.RT_PRECOND_JACOBI, 'nearField': _abi.RT_PRECOND_NEARFIELD}


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """One GPU, one HIP stream and the memory ceiling of the reference's Resources/Budget (Resources.m:9-47)."""

    _default = None

    def __init__(self, device=0, memory=0):
        lib = library()
        h = C.c_void_p()
        check(lib.rt_ctx_create(int(device), int(memory), C.byref(h)))
        self.handle = h
        self._objects = []

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context()
        return cls._default

    def memory(self):
        """(held, ceiling) device bytes (Resources.held/ceiling, Resources.m:55-71)."""
        held, ceiling = C.c_uint64(), C.c_uint64()
        check(library().rt_ctx_memory(self.handle, C.byref(held), C.byref(ceiling)))
        return held.value, ceiling.value

    def peakMemory(self, reset=False):
        """The most device bytes held at once since the context was created or the last reset."""
        peak = C.c_uint64()
        check(library().rt_ctx_memory_peak(self.handle, int(bool(reset)), C.byref(peak)))
        return peak.value

    def synchronize(self):
        check(library().rt_ctx_synchronize(self.handle))

    def stream(self):
        s = C.c_void_p()
        check(library().rt_ctx_stream(self.handle, C.byref(s)))
        return s.value


class Mesh:
    """Uniform grid of countX x countY faces of sizeX x sizeY m (Mesh.m:10-30)."""

    def __init__(self, countX, countY, sizeX, sizeY):
        if int(countX) != countX or int(countY) != countY or countX < 1 or countY < 1:
            raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBePositive', 'face counts are positive')
        if not (sizeX > 0 and sizeY > 0 and math.isfinite(sizeX) and math.isfinite(sizeY)):
            raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBePositive', 'face sizes are positive')
        if countX == 1 and countY == 1:
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:mesh:noRooftops',
                               'A single face has no interior edge, so it carries no rooftop.')
        self.countX, self.countY = int(countX), int(countY)
        self.sizeX, self.sizeY = float(sizeX), float(sizeY)

    def count(self):
        return self.countX, self.countY

    def faceSize(self):
        return self.sizeX, self.sizeY

    def radius(self):
        return math.hypot(self.countX * self.sizeX, self.countY * self.sizeY) / 2


class Problem:
    """The EFIE of a Mesh at one wavenumber (Problem.m:25-92); generators assembled on the GPU on first use.

    mask (countX x countY booleans, optional) keeps the rooftops whose two faces are metal.
    """

    def __init__(self, mesh, wavenumber=None, electricalSize=None, quadratureOrder=(16, 8), mask=None,
                 context=None):
        if (wavenumber is None) == (electricalSize is None):
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:problem:invalidWavenumber',
                               "A problem takes either the wavenumber or the 'electricalSize', not both or neither.")
        if wavenumber is None:
            wavenumber = electricalSize / mesh.radius()
        orders = np.atleast_1d(quadratureOrder)
        if orders.size not in (1, 2):
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:problem:invalidQuadratureOrder',
                               'The quadrature order is one order or a pair [near, far].')
        self.mesh = mesh
        self.context = context or Context.default()
        self.orders = (int(orders[0]), int(orders[-1]))
        self._wavenumber = float(wavenumber)
        self.mask = None
        mptr = None
        if mask is not None:
            self.mask = np.asfortranarray(np.asarray(mask, dtype=np.uint8).reshape(mesh.countX, mesh.countY))
            mptr = C.c_void_p(self.mask.ctypes.data)
        h = C.c_void_p()
        check(library().rt_problem_create(self.context.handle, mesh.countX, mesh.countY, mesh.sizeX, mesh.sizeY,
                                          self._wavenumber, self.orders[0], self.orders[1], mptr, C.byref(h)))
        self.handle = h

    def __del__(self):
        if getattr(self, 'handle', None) is not None and _abi._lib is not None:
            _abi._lib.rt_problem_destroy(self.handle)
            self.handle = None

    def wavenumber(self):
        return self._wavenumber

    def electricalSize(self):
        return self._wavenumber * self.mesh.radius()

    def rooftopCount(self):
        n = C.c_int64()
        check(library().rt_problem_count(self.handle, C.byref(n)))
        return n.value

    def countDir(self, direction):
        cx, cy = self.mesh.count()
        return (cx - 1) * cy if direction == 'x' else cx * (cy - 1)

    def generators(self):
        """{'magneticXX', 'magneticYY', 'electric'} as complex128 arrays (None for a missing direction)."""
        cx, cy = self.mesh.count()
        out = {'magneticXX': np.zeros((cx - 1, cy), np.complex128, order='F') if cx > 1 else None,
               'magneticYY': np.zeros((cx, cy - 1), np.complex128, order='F') if cy > 1 else None,
               'electric': np.zeros((cx, cy), np.complex128, order='F')}
        check(library().rt_problem_generators(self.handle, _ptr(out['magneticXX']), _ptr(out['magneticYY']),
                                              _ptr(out['electric'])))
        return out

    def setGenerators(self, magneticXX, magneticYY, electric):
        """Uses given generators (as the reference tests build Generators from stored data)."""
        arrs = [None if g is None else np.asfortranarray(g, dtype=np.complex128)
                for g in (magneticXX, magneticYY, electric)]
        check(library().rt_problem_set_generators(self.handle, *[_ptr(a) for a in arrs]))


class Product:
    """Base of the products (Product.m): apply(x) for N x 1 (or N x R) vectors, memory(), workingMemory()."""

    handle = None

    def __init__(self, problem, handle, precision, count=None, context=None):
        self._problem = problem
        self.handle = handle
        self.precision = precision
        self._count = problem.rooftopCount() if count is None else int(count)
        self.context = problem.context if context is None else context

    def __del__(self):
        if getattr(self, 'handle', None) is not None and _abi._lib is not None:
            _abi._lib.rt_product_destroy(self.handle)
            self.handle = None

    def problem(self):
        return self._problem

    def apply(self, x, precision=None):
        """Product.apply (Product.m:25-35); raises rooftopMoM:product:sizeMismatch on a wrong size."""
        n = self._count
        a = np.asarray(x)
        if not np.issubdtype(a.dtype, np.number) or a.ndim not in (1, 2) or a.shape[0] != n:
            raise RooftopError(_abi.RT_ESIZE, 'rooftopMoM:product:sizeMismatch',
                               'The input must be a numeric %dx1 vector.' % n)
        a = np.asfortranarray(a, dtype=np.complex128)
        nrhs = 1 if a.ndim == 1 else a.shape[1]
        out = np.empty_like(a)
        prec = _PRECISIONS[precision] if precision else (_abi.RT_C128 if self.precision == _abi.RT_MIXED
                                                         else self.precision)
        status = library().rt_product_apply(self.handle, _ptr(a), _ptr(out), n, nrhs, prec)
        if isinstance(self, CallbackOperator):
            self.raise_pending()
        check(status)
        return out

    def apply_device(self, x_ptr, y_ptr, nrhs=1, precision=None):
        """Enqueues the product on device buffers (complex128) on the context's stream."""
        prec = _PRECISIONS[precision] if precision else (_abi.RT_C128 if self.precision == _abi.RT_MIXED
                                                         else self.precision)
        check(library().rt_product_apply_device(self.handle, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                                self._count, nrhs, prec))

    def memory(self):
        s = C.c_uint64()
        check(library().rt_product_memory(self.handle, C.byref(s), None))
        return s.value

    def workingMemory(self):
        w = C.c_uint64()
        check(library().rt_product_memory(self.handle, None, C.byref(w)))
        return w.value

    def spectra(self):
        """Applicator.spectra() (Applicator.m:20-22): {'period': (Mx, My), 'magneticXX', 'magneticYY', 'electric'}, each
        plane the DFT of the generator's even embedding scaled by 1/(Mx My), on the folded frequencies
        [0, Mx/2] x [0, My/2] (Hx x Hy, complex128)."""
        dims = (C.c_int32 * 4)()
        check(library().rt_product_spectra(self.handle, dims, None))
        hx, hy = dims[2], dims[3]
        planes = np.zeros(3 * hx * hy, np.complex128)
        check(library().rt_product_spectra(self.handle, dims, _ptr(planes)))
        planes = planes.reshape(3, hy, hx).transpose(0, 2, 1)
        return {'period': (dims[0], dims[1]), 'magneticXX': planes[0], 'magneticYY': planes[1],
                'electric': planes[2]}


class Operator(Product):
    """rooftopMoM.Operator(problem, 'product', ..., 'modifier', ...) (Operator.m:7-54)."""

    def __init__(self, problem, product='splitFFT', modifier=None, precision='c128'):
        if product not in _PRODUCTS:
            raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBeMember',
                               "'product' must be 'fft', 'splitFFT' or 'leanSplitFFT'.")
        h = C.c_void_p()
        check(library().rt_product_create(problem.handle, _PRODUCTS[product], _PRECISIONS[precision],
                                          _MODIFIERS[modifier], C.byref(h)))
        super().__init__(problem, h, _PRECISIONS[precision])

    @staticmethod
    def estimateMemory(countX, countY, precision='c128'):
        """(stored, working) device bytes of the product of a countX x countY grid before it exists (the static
        Applicator.estimateWorkingMemory, Applicator.m:44-46)."""
        stored, working = C.c_uint64(), C.c_uint64()
        check(library().rt_product_estimate_memory(int(countX), int(countY), _PRECISIONS[precision],
                                                   C.byref(stored), C.byref(working)))
        return stored.value, working.value

    def tune(self, nrhs=1, precision=None):
        """Times the launch shapes of the matvec passes for nrhs right-hand sides and keeps the fastest (the FFT
        tile and radix schedule); returns ({'bits', 'group', 'tile', 'teams', 'chunks'}, milliseconds of one
        application)."""
        prec = _PRECISIONS[precision] if precision else (_abi.RT_C128 if self.precision == _abi.RT_MIXED
                                                         else self.precision)
        shape = (C.c_int32 * 5)()
        ms = C.c_double()
        check(library().rt_product_tune(self.handle, nrhs, prec, shape, C.byref(ms)))
        return {'bits': shape[0], 'group': shape[1], 'tile': shape[2], 'teams': shape[3], 'chunks': shape[4]}, ms.value

    def setShape(self, shape, nrhs=1, precision=None):
        """Uses shape {'bits', 'group', 'tile', 'teams', 'chunks'} for nrhs right-hand sides (a saved tune() result;
        teams and chunks default to 1; a group of -2 or -4 runs that many pass A/C columns in turn)."""
        prec = _PRECISIONS[precision] if precision else (_abi.RT_C128 if self.precision == _abi.RT_MIXED
                                                         else self.precision)
        arr = (C.c_int32 * 5)(shape['bits'], shape['group'], shape['tile'], shape.get('teams', 1),
                              shape.get('chunks', 1))
        check(library().rt_product_set_shape(self.handle, nrhs, prec, arr))

    def profile(self, nrhs=1, reps=20, precision=None):
        """Kernel milliseconds and algorithmic bytes of the three passes in one application, and the wall
        milliseconds of one application (HIP events): (pass_ms, pass_bytes, wall_ms)."""
        prec = _PRECISIONS[precision] if precision else (_abi.RT_C128 if self.precision == _abi.RT_MIXED
                                                         else self.precision)
        ms = (C.c_double * 4)()
        by = (C.c_double * 3)()
        check(library().rt_product_profile(self.handle, nrhs, reps, prec, ms, by))
        return list(ms)[:3], list(by), ms[3]


class DenseOperator(Product):
    """A matrix operator on the GPU (the matrix form of convertToHandle, +solver/+utils/convertToHandle.m:10-16):
    the N x N complex matrix is copied to the device and applied by a batched matrix-vector product."""

    def __init__(self, matrix, precision='c128', context=None):
        a = np.asarray(matrix)
        if a.ndim != 2 or a.shape[0] != a.shape[1] or not np.issubdtype(a.dtype, np.number):
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:solver:invalidOperator',
                               'A matrix operator is a square numeric matrix.')
        a = np.asfortranarray(a, dtype=np.complex128)
        context = context or Context.default()
        h = C.c_void_p()
        check(library().rt_product_create_dense(context.handle, a.shape[0], _ptr(a), _PRECISIONS[precision],
                                                C.byref(h)))
        super().__init__(None, h, _PRECISIONS[precision], count=a.shape[0], context=context)


class CallbackOperator(Product):
    """A function handle or an object with apply as an operator of the GPU solver (convertToHandle.m:10-16): the
    Krylov loop calls it on the host with every product, one N x 1 vector at a time."""

    def __init__(self, function, count, context=None):
        fn = function.apply if hasattr(function, 'apply') and not isinstance(function, Product) else function
        if not callable(fn):
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:solver:invalidOperator',
                               'An operator is a matrix, a function handle or an object with apply.')
        self._function = fn
        self.error = None
        count = int(count)

        def trampoline(user, inp, out, n, nrhs):
            try:
                size = int(n) * int(nrhs)
                x = np.ctypeslib.as_array(C.cast(inp, C.POINTER(C.c_double)), (2 * size,)).view(np.complex128)
                y = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_double)), (2 * size,)).view(np.complex128)
                x = x.reshape((int(n), int(nrhs)), order='F')
                y = y.reshape((int(n), int(nrhs)), order='F')
                for r in range(int(nrhs)):
                    v = np.asarray(fn(x[:, r].copy()), dtype=np.complex128).reshape(-1)
                    if v.size != n:
                        raise RooftopError(_abi.RT_ESIZE, 'rooftopMoM:solver:sizeMismatch',
                                           'The operator returned %d values for %d unknowns.' % (v.size, n))
                    y[:, r] = v
            except BaseException as e:   # reported by the call that applied the operator
                if self.error is None:
                    self.error = e

        self._trampoline = _abi.APPLY_FN(trampoline)   # kept alive with the product
        context = context or Context.default()
        h = C.c_void_p()
        check(library().rt_product_create_callback(context.handle, count, self._trampoline, None, C.byref(h)))
        super().__init__(None, h, _abi.RT_MIXED, count=count, context=context)

    def raise_pending(self):
        if self.error is not None:
            e, self.error = self.error, None
            raise e


def as_product(operator, count, context=None):
    """A Product for what the reference's solver accepts as an operator or preconditioner (convertToHandle.m:10-16):
    a product as is, a matrix (DenseOperator), a function handle or an object with apply (CallbackOperator)."""
    if operator is None or isinstance(operator, Product):
        return operator
    if isinstance(operator, np.ndarray) or (hasattr(operator, 'shape') and hasattr(operator, 'dtype')):
        return DenseOperator(operator, context=context)
    return CallbackOperator(operator, count, context=context)


_KERNELS = {'strang': 0, 'tChan': 1, 'rChan': 2, 'huckle': 3, 'superoptimal': 4}
_BUILDERS = {'partwise': 0, 'blockwise': 1, 'hybrid': 2}


class Preconditioner(Product):
    """A preconditioner product: 'nearField' (3x3-cell neighbourhoods), 'jacobi' (diag(Z)) or 'circulant', the
    reference's rooftopMoM.Preconditioner(problem, 'kernel', k, 'builder', b, 'bandwidth', [bx by], 'modifier', m)
    (Preconditioner.m:7-70): the inverse of a block circulant approximation on the face grid. Passing kernel= (or
    builder=, bandwidth=) selects 'circulant', whose defaults are the reference's ('rChan', 'partwise').

    Errors rooftopMoM:preconditioner:singular like Preconditioner.m:60-67.
    """

    def __init__(self, problem, kind=None, precision='c128', kernel=None, builder=None, bandwidth=None,
                 modifier=None):
        if kind is None:
            kind = 'circulant' if (kernel or builder or bandwidth is not None) else 'nearField'
        if kind == 'circulant':
            kernel, builder = kernel or 'rChan', builder or 'partwise'
            if kernel not in _KERNELS:
                raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBeMember',
                                   "'kernel' must be 'strang', 'tChan', 'rChan', 'huckle' or 'superoptimal'.")
            if builder not in _BUILDERS:
                raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBeMember',
                                   "'builder' must be 'partwise', 'blockwise' or 'hybrid'.")
            bw = (0, 0)
            if bandwidth is not None:
                bw = tuple(np.atleast_1d(bandwidth))
                if len(bw) != 2:
                    raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:preconditioner:invalidBandwidth',
                                       'The bandwidth is a pair [bandwidthX, bandwidthY].')
                if any(int(v) != v or v <= 0 for v in bw):
                    raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:preconditioner:invalidBandwidth',
                                       'The bandwidth is a pair of positive integers.')
            h = C.c_void_p()
            check(library().rt_precond_create_circulant(problem.handle, _KERNELS[kernel], _BUILDERS[builder],
                                                        int(bw[0]), int(bw[1]), _MODIFIERS[modifier],
                                                        _PRECISIONS[precision], C.byref(h)))
            self.kernel, self.builder = kernel, builder
        else:
            if kind not in _PRECONDITIONERS:
                raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBeMember',
                                   "'kind' must be 'nearField', 'jacobi' or 'circulant'.")
            h = C.c_void_p()
            check(library().rt_precond_create(problem.handle, _PRECONDITIONERS[kind], _PRECISIONS[precision],
                                              C.byref(h)))
        self.kind = kind
        super().__init__(problem, h, _PRECISIONS[precision])


# ---------------------------------------------------------------------------------------------------------------
# Excitations (O(N) host formulas of the boundary inputs; the solve itself runs on the GPU).
# ---------------------------------------------------------------------------------------------------------------
def _gauss_legendre01(order):
    x, w = np.polynomial.legendre.leggauss(order)
    return (x + 1) / 2, w / 2


class Excitation:
    """Right-hand side of a Problem (Excitation.m): planeWave or deltaGap."""

    def __init__(self, problem, kind, vector, parameters):
        self._problem, self._type, self._vector, self._parameters = problem, kind, vector, parameters

    def problem(self):
        return self._problem

    def type(self):
        return self._type

    def vector(self):
        return self._vector

    def parameters(self):
        return self._parameters

    @staticmethod
    def planeWave(problem, direction=(0, 0, -1), polarization=(1, 0, 0), intensity=1.0):
        """E = intensity * polarization * exp(-1i k direction . r) tested with every rooftop, computed in O(N) by
        translation phases (generatePlaneWave.m:18-42) with the problem's far quadrature order."""
        d = np.asarray(direction, float)
        p = np.asarray(polarization, float)
        for v, name in ((d, 'direction'), (p, 'polarization')):
            if not np.any(v):
                raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:excitation:zeroVector', 'The %s must not be zero.' % name)
        d, p = d / np.linalg.norm(d), p / np.linalg.norm(p)
        if abs(d @ p) > 10 * np.finfo(float).eps:
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:excitation:notPerpendicular',
                               'The polarization must be perpendicular to the direction of propagation.')
        vector = plane_wave_vector(problem, d, p, float(intensity))
        return Excitation(problem, 'planeWave', vector,
                          {'direction': d, 'polarization': p, 'intensity': float(intensity)})

    @staticmethod
    def deltaGap(problem, port=None, direction='x', voltage=1.0):
        """voltage x edge length on the port rooftops (generateDeltaGap.m:13-14); the default port is the middle
        line of edges (generateMiddlePort.m:14-28). port holds 0-based rooftop indices."""
        cx, cy = problem.mesh.count()
        n = problem.rooftopCount()
        if port is None:
            if problem.countDir(direction) == 0:
                raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:excitation:noRooftops',
                                   'The mesh has no %s-directed rooftops.' % direction)
            port = middle_port(problem, direction)
        port = np.asarray(port, dtype=np.int64)
        if np.any(port >= n) or np.any(port < 0) or np.unique(port).size < port.size:
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:excitation:invalidPort',
                               'The port must hold distinct rooftop indices up to %d.' % n)
        vector = np.zeros(n, np.complex128)
        vector[port] = voltage * edge_lengths(problem, port)
        return Excitation(problem, 'deltaGap', vector, {'port': port, 'voltage': voltage})


class DeviceBatch:
    """n x count complex128 in device memory (a column per right-hand side), owned through a Context."""

    def __init__(self, context, n, count):
        self.context, self.n, self.count = context, int(n), int(count)
        p = C.c_void_p()
        check(library().rt_device_alloc(context.handle, 16 * self.n * self.count, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, context, array):
        a = np.asfortranarray(array, dtype=np.complex128)
        if a.ndim == 1:
            a = a[:, None]
        out = cls(context, a.shape[0], a.shape[1])
        check(library().rt_copy_to_device(context.handle, C.c_void_p(out.ptr), a.ctypes.data_as(C.c_void_p),
                                          a.nbytes))
        return out

    def to_host(self):
        a = np.empty((self.n, self.count), np.complex128, order='F')
        check(library().rt_copy_to_host(self.context.handle, a.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                        a.nbytes))
        return a

    def free(self):
        if getattr(self, 'ptr', None) and _abi._lib is not None:
            _abi._lib.rt_device_free(self.context.handle, C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        self.free()


def _rows3(v, count, name):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (count, 3)))
    return a


def plane_waves(problem, directions, polarizations, intensities=None, out=None):
    """The plane waves of count incidences (directions, polarizations: count x 3) as a DeviceBatch, computed on the
    device (rt_plane_waves_device; generatePlaneWave.m:18-42 per column)."""
    d = np.atleast_2d(np.asarray(directions, dtype=np.float64))
    count = d.shape[0]
    d = _rows3(d, count, 'direction')
    p = _rows3(polarizations, count, 'polarization')
    e = None if intensities is None else np.ascontiguousarray(np.broadcast_to(np.asarray(intensities, float), count))
    out = out or DeviceBatch(problem.context, problem.rooftopCount(), count)
    dp = C.POINTER(C.c_double)
    check(library().rt_plane_waves_device(problem.handle, count, d.ctypes.data_as(dp), p.ctypes.data_as(dp),
                                          None if e is None else e.ctypes.data_as(dp), C.c_void_p(out.ptr)))
    return out


def radar_cross_sections(problem, b, x, intensities=None):
    """sigma per column of the DeviceBatches b (plane waves) and x (their solutions), on the device
    (rt_radar_cross_sections_device; computeRadarCrossSection.m:18-19)."""
    count = b.count
    e = None if intensities is None else np.ascontiguousarray(np.broadcast_to(np.asarray(intensities, float), count))
    sigma = np.zeros(count)
    dp = C.POINTER(C.c_double)
    check(library().rt_radar_cross_sections_device(problem.handle, count, C.c_void_p(b.ptr), C.c_void_p(x.ptr),
                                                   None if e is None else e.ctypes.data_as(dp),
                                                   sigma.ctypes.data_as(dp)))
    return sigma


def input_impedances(problem, port, x, voltages=None):
    """Z_in per column of the DeviceBatch x for one port (0-based rooftop indices), on the device
    (rt_input_impedances_device; computeInputImpedance.m:17)."""
    port = np.ascontiguousarray(port, dtype=np.int64)
    count = x.count
    v = None if voltages is None else np.ascontiguousarray(np.broadcast_to(np.asarray(voltages, float), count))
    z = np.zeros(2 * count)
    dp = C.POINTER(C.c_double)
    check(library().rt_input_impedances_device(problem.handle, count, port.ctypes.data_as(C.POINTER(C.c_int64)),
                                               port.size, None if v is None else v.ctypes.data_as(dp),
                                               C.c_void_p(x.ptr), z.ctypes.data_as(dp)))
    return z[0::2] + 1j * z[1::2]


def edge_lengths(problem, idx):
    idx = np.asarray(idx)
    out = np.full(idx.shape, problem.mesh.sizeX)
    out[idx < problem.countDir('x')] = problem.mesh.sizeY
    return out


def middle_port(problem, direction='x'):
    cx, cy = problem.mesh.count()
    if direction == 'x':
        rows = cx - 1
        return (cx // 2 - 1) + rows * np.arange(cy)
    rows = cx
    return problem.countDir('x') + np.arange(rows) + rows * (cy // 2 - 1)


def plane_wave_vector(problem, d, p, intensity):
    a, b = problem.mesh.faceSize()
    cx, cy = problem.mesh.count()
    k = problem.wavenumber()
    t, w = _gauss_legendre01(problem.orders[1])
    pu, pv = np.meshgrid(t, t, indexing='ij')   # u fastest after the column-major ravel
    wu, wv = np.meshgrid(w, w, indexing='ij')
    pu, pv, w2 = pu.ravel(order='F'), pv.ravel(order='F'), (wu * wv).ravel(order='F')
    steps = -1j * k * d[:2] * np.array([a, b])
    phases = np.exp(steps[0] * pu + steps[1] * pv)
    parts = []
    for n, (rows, cols, coord, second) in enumerate(((cx - 1, cy, pu, (1, 0)), (cx, cy - 1, pv, (0, 1)))):
        if rows * cols == 0:
            continue
        amp = coord + np.exp(steps[0] * second[0] + steps[1] * second[1]) * (1 - coord)
        ref = intensity * p[n] * a * b * (phases @ (w2 * amp))
        trans = np.exp(steps[0] * np.arange(rows))[:, None] * np.exp(steps[1] * np.arange(cols))[None, :]
        parts.append(ref * trans.ravel(order='F'))
    vec = np.concatenate(parts)
    if problem.mask is not None:
        vec = vec * rooftop_existence(problem)
    return vec


def rooftop_existence(problem):
    """Per rooftop: 1 if both its faces are metal (all ones without a mask)."""
    cx, cy = problem.mesh.count()
    if problem.mask is None:
        return np.ones(problem.rooftopCount())
    m = problem.mask.astype(bool)
    ex = (m[:-1, :] & m[1:, :]).ravel(order='F') if cx > 1 else np.zeros(0, bool)
    ey = (m[:, :-1] & m[:, 1:]).ravel(order='F') if cy > 1 else np.zeros(0, bool)
    return np.concatenate([ex, ey]).astype(float)


class Result:
    """Outcome of Solver.solve (Result.m:14-94)."""

    def __init__(self, solution, converged, iterationCount, residualNorms, residualNorm, time, excitation,
                 outerCount=0):
        self._solution, self._converged, self._iterations = solution, converged, iterationCount
        self._norms, self._residual, self._time = residualNorms, residualNorm, time
        self._excitation, self.outerCount = excitation, outerCount

    def solution(self):
        return self._solution

    def converged(self):
        return self._converged

    def iterationCount(self):
        return self._iterations

    def residualNorms(self):
        return self._norms

    def residualNorm(self):
        return self._residual

    def time(self):
        return self._time

    def excitation(self):
        return self._excitation

    def _params(self, kind):
        if self._excitation is None or self._excitation.type() != kind:
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:result:wrongExcitation',
                               'This quantity needs a %s excitation.' % kind)
        return self._excitation.parameters()

    def inputImpedance(self):
        """V / (edge lengths . I(port)) (computeInputImpedance.m:17)."""
        p = self._params('deltaGap')
        prob = self._excitation.problem()
        return p['voltage'] / (edge_lengths(prob, p['port']) @ self._solution[p['port']])

    def radarCrossSection(self):
        """(k Z0)^2 |b.' I|^2 / (4 pi E^4) (computeRadarCrossSection.m:18-19)."""
        p = self._params('planeWave')
        k = self._excitation.problem().wavenumber()
        return (k * Z0) ** 2 * abs(self._excitation.vector() @ self._solution) ** 2 / (4 * np.pi * p['intensity'] ** 4)

    def normalizedRadarCrossSection(self):
        return self.radarCrossSection() / (2 * np.pi / self._excitation.problem().wavenumber()) ** 2


class Solver:
    """rooftopMoM.Solver('method', ..., 'tolerance', ..., 'maxIterationCount', ..., 'restart', ...) (Solver.m:17-38).

    precision 'c64' runs the products of the Krylov loop in complex64 inside an FP64 refinement loop on the true
    residual (refine is then implied); refine=True alone wraps the FP64 loop the same way. basis 'c128' keeps the
    GMRES basis of a complex64 inner loop in complex128 (by default it is stored in complex64).
    """

    def __init__(self, method='cg', tolerance=1e-6, maxIterationCount=1000, restart=50, precision='c128',
                 refine=False, innerTolerance=0.0, maxOuter=20, basis=None):
        if method not in ('cg', 'gmres', 'blockcg'):
            raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBeMember',
                               "'method' must be 'cg', 'gmres' or 'blockcg'.")
        if not (0 < tolerance < 1):
            raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBeLessThan', "'tolerance' must be below 1.")
        if not (restart == math.inf or (restart > 0 and restart == int(restart))):
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:solver:invalidRestart',
                               'The restart is a positive integer or Inf.')
        if precision not in ('c128', 'c64'):
            raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBeMember', "'precision' is 'c128' or 'c64'.")
        self.method, self.tolerance, self.maxIterationCount = method, float(tolerance), int(maxIterationCount)
        self.restart = restart
        self.precision = precision
        self.refine = bool(refine) or precision == 'c64'
        if basis not in (None, 'c64', 'c128'):
            raise RooftopError(_abi.RT_EINVAL, 'MATLAB:validators:mustBeMember', "'basis' is 'c64' or 'c128'.")
        self.basis = basis
        self.innerTolerance, self.maxOuter = float(innerTolerance), int(maxOuter)

    def options(self):
        o = _abi.rt_solve_opts()
        o.method = {'cg': _abi.RT_METHOD_CG, 'gmres': _abi.RT_METHOD_GMRES,
                    'blockcg': _abi.RT_METHOD_BLOCK_CG}[self.method]
        o.tolerance = self.tolerance
        o.max_iterations = self.maxIterationCount
        o.restart = 0 if self.restart == math.inf else int(self.restart)
        o.precision = _PRECISIONS[self.precision]
        if self.precision == 'c64' and self.basis == 'c128':
            o.precision = _abi.RT_C64_FP64_BASIS
        o.refine = int(self.refine)
        o.inner_tolerance = self.innerTolerance
        o.max_outer = self.maxOuter
        return o

    def solve(self, operator, rightHandSide, preconditioner=None, initialGuess=None):
        """Solves operator * x = rightHandSide on the GPU (Solver.m:42-92).

        A matrix right-hand side (N x R) solves its R columns together, in lockstep, and returns one Result per column,
        each equal to the Result of solving that column alone.
        """
        excitation = rightHandSide if isinstance(rightHandSide, Excitation) else None
        b = excitation.vector() if excitation else rightHandSide
        b = np.asarray(b)
        if not np.issubdtype(b.dtype, np.number) or not np.all(np.isfinite(b)):
            raise RooftopError(_abi.RT_EINVAL, 'rooftopMoM:solver:invalidRightHandSide',
                               'The right-hand side is a rooftopMoM.Excitation or a numeric vector of finite values.')
        b = np.asfortranarray(b, dtype=np.complex128)
        # Matrices, function handles and objects with apply become device products (Solver.m:146-154).
        context = next((q.context for q in (operator, preconditioner) if isinstance(q, Product)), None)
        operator = as_product(operator, b.shape[0], context)
        preconditioner = as_product(preconditioner, b.shape[0], operator.context)
        n = operator._count
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        if b.shape[0] != n:
            raise RooftopError(_abi.RT_ESIZE, 'rooftopMoM:solver:sizeMismatch',
                               'The operator does not match the %d unknowns of the right-hand side.' % b.shape[0])
        if preconditioner is not None and preconditioner._count != n:
            raise RooftopError(_abi.RT_ESIZE, 'rooftopMoM:solver:sizeMismatch',
                               'The preconditioner does not match the %d unknowns of the right-hand side.' % n)
        x0 = None
        if initialGuess is not None:
            x0 = np.asfortranarray(initialGuess, dtype=np.complex128)
            if x0.size != b.size:
                raise RooftopError(_abi.RT_ESIZE, 'rooftopMoM:solver:sizeMismatch',
                                   'The initial guess must have %d elements.' % n)
        x = np.empty_like(b)
        cap = self.maxIterationCount * 2 + 64
        norms = np.zeros((nrhs, cap))
        outs = (_abi.rt_output * nrhs)()
        for r in range(nrhs):
            outs[r].residual_norms_capacity = cap
            outs[r].residual_norms = norms[r].ctypes.data_as(C.POINTER(C.c_double))
        o = self.options()
        status = library().rt_solve(operator.context.handle, operator.handle,
                                    preconditioner.handle if preconditioner is not None else None, C.byref(o),
                                    _ptr(b), n, nrhs, _ptr(x0), _ptr(x), outs)
        for q in (operator, preconditioner):
            if isinstance(q, CallbackOperator):
                q.raise_pending()
        check(status)
        results = [Result(x if b.ndim == 1 else x[:, r].copy(), bool(out.converged), out.iteration_count,
                          norms[r, :min(cap, out.residual_norms_count)].copy(), out.residual_norm, out.time_s,
                          excitation, out.outer_count) for r, out in enumerate(outs)]
        return results[0] if b.ndim == 1 else results

    def solveDevice(self, operator, b, preconditioner=None, x=None):
        """Solves the columns of the DeviceBatch b together with inputs and outputs in HBM (rt_solve_device); returns
        (x as a DeviceBatch, [Result without a host solution] per column)."""
        n = operator._count
        if b.n != n:
            raise RooftopError(_abi.RT_ESIZE, 'rooftopMoM:solver:sizeMismatch',
                               'The operator does not match the %d unknowns of the right-hand side.' % b.n)
        x = x or DeviceBatch(b.context, n, b.count)
        cap = self.maxIterationCount * 2 + 64
        norms = np.zeros((b.count, cap))
        outs = (_abi.rt_output * b.count)()
        for r in range(b.count):
            outs[r].residual_norms_capacity = cap
            outs[r].residual_norms = norms[r].ctypes.data_as(C.POINTER(C.c_double))
        o = self.options()
        check(library().rt_solve_device(b.context.handle, operator.handle,
                                        preconditioner.handle if preconditioner is not None else None, C.byref(o),
                                        C.c_void_p(b.ptr), n, b.count, None, C.c_void_p(x.ptr), outs))
        return x, [Result(None, bool(out.converged), out.iteration_count,
                          norms[r, :min(cap, out.residual_norms_count)].copy(), out.residual_norm, out.time_s, None,
                          out.outer_count) for r, out in enumerate(outs)]


# ---------------------------------------------------------------------------------------------------------------
# Serving: the library's request loop (rt_server_*), and the connected-blob masks of the served-stream workload.
# ---------------------------------------------------------------------------------------------------------------
def blob_mask(countX, countY, seed, fill):
    """A 4-connected blob of metal faces grown by random frontier growth from a uniformly drawn seed face until
    round(fill * countX * countY) faces are metal (countX x countY booleans; rt_blob_mask)."""
    m = np.zeros(countX * countY, np.uint8)
    check(library().rt_blob_mask(int(countX), int(countY), int(seed), float(fill), m.ctypes.data_as(C.c_void_p)))
    return m.reshape((countX, countY), order='F').astype(bool)


class Server:
    """Requests of different grids, frequencies and masks admitted first come, first served by an HBM budget and
    solved concurrently by `slots` workers, each on its own HIP stream (rt_server_*). The result of a request is
    the one Solver.solve gives on the same inputs. batch > 1: queued requests of the same embedding size (and solver
    options) are solved together, up to `batch` in one batched product and Krylov loop (rt_server_set_batching)."""

    def __init__(self, slots=4, budget=0, device=0, batch=1):
        h = C.c_void_p()
        check(library().rt_server_create(int(device), int(budget), int(slots), C.byref(h)))
        self.handle = h
        if batch != 1:
            check(library().rt_server_set_batching(self.handle, int(batch)))

    @staticmethod
    def _request(countX, countY, sizeX, sizeY, wavenumber, rightHandSide, mask, solver, preconditioner,
                 quadratureOrder):
        solver = solver or Solver('cg', tolerance=1e-6)
        b = np.ascontiguousarray(rightHandSide, dtype=np.complex128)
        req = _abi.rt_request()
        req.count_x, req.count_y = int(countX), int(countY)
        req.size_x, req.size_y, req.wavenumber = float(sizeX), float(sizeY), float(wavenumber)
        req.near_order, req.far_order = int(quadratureOrder[0]), int(quadratureOrder[-1])
        m = None
        if mask is not None:
            mask = np.asarray(mask, dtype=np.uint8)
            if mask.size != int(countX) * int(countY):
                raise RooftopError(_abi.RT_ESIZE, 'rooftopMoM:solver:sizeMismatch',
                                   'The mask must hold countX x countY = %d faces.' % (int(countX) * int(countY)))
            m = np.asfortranarray(mask.reshape(int(countX), int(countY))).ravel(order='F')
            req.mask = m.ctypes.data
        req.rhs = b.ctypes.data
        req.n = b.size
        req.preconditioner = _PRECONDITIONERS[preconditioner] if preconditioner else 0
        req.opts = solver.options()
        return req, (b, m)   # the arrays must outlive the call that reads the request

    @staticmethod
    def requestBytes(countX, countY, solver=None, preconditioner='nearField'):
        """Device bytes a request of a countX x countY grid reserves while it runs (its admission size)."""
        req, keep = Server._request(countX, countY, 1.0, 1.0, 1.0, np.zeros(1), None, solver, preconditioner,
                                    (16, 8))
        b = C.c_uint64()
        check(library().rt_request_bytes(C.byref(req), C.byref(b)))
        return b.value

    def submit(self, countX, countY, sizeX, sizeY, wavenumber, rightHandSide, mask=None, solver=None,
               preconditioner='nearField', quadratureOrder=(16, 8)):
        """Queues one solve and returns its ticket at once (the inputs are copied)."""
        req, keep = self._request(countX, countY, sizeX, sizeY, wavenumber, rightHandSide, mask, solver,
                                  preconditioner, quadratureOrder)
        ticket = C.c_int64()
        check(library().rt_server_submit(self.handle, C.byref(req), C.byref(ticket)))
        return ticket.value

    def wait(self, ticket, n):
        """(solution, Result, {'queue', 'service', 'latency'} seconds) of a ticket; blocks until it has run."""
        x = np.empty(n, np.complex128)
        cap = 8192
        norms = np.zeros(cap)
        out = _abi.rt_output()
        out.residual_norms_capacity = cap
        out.residual_norms = norms.ctypes.data_as(C.POINTER(C.c_double))
        times = (C.c_double * 3)()
        check(library().rt_server_wait(self.handle, int(ticket), _ptr(x), n, C.byref(out), times))
        res = Result(x, bool(out.converged), out.iteration_count, norms[:min(cap, out.residual_norms_count)].copy(),
                     out.residual_norm, out.time_s, None, out.outer_count)
        return x, res, {'queue': times[0], 'service': times[1], 'latency': times[2]}

    def stats(self):
        counts = (C.c_int64 * 5)()
        by = (C.c_uint64 * 2)()
        check(library().rt_server_stats(self.handle, counts, by))
        groups, grouped = C.c_int64(), C.c_int64()
        check(library().rt_server_batch_stats(self.handle, C.byref(groups), C.byref(grouped)))
        return {'submitted': counts[0], 'completed': counts[1], 'queued': counts[2], 'running': counts[3],
                'peak_running': counts[4], 'reserved_bytes': by[0], 'budget_bytes': by[1],
                'batched_solves': groups.value, 'batched_requests': grouped.value}

    def close(self):
        if getattr(self, 'handle', None) is not None and _abi._lib is not None:
            _abi._lib.rt_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

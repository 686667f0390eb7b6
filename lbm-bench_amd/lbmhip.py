"""ctypes binding of the GPU library (liblbmhip-dp.so / liblbmhip-sp.so).

Mirrors the structs of include/lbm.h (LBM-Bench's src/lbm.h:54-103 with
the appended fields) and wraps the kernel registry the way LBM-Bench's
driver uses it (src/main.c:93-152): geometryCreate -> GKernels[k].Init ->
SetNode / device init -> Kernel -> GetNode / statistics -> Deinit.

There is no fallback: if the library is not built, loading raises.
Logical PDF arrays are float64 numpy arrays (nX, nY, nZ, 19).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

KERNELS = ("pull-soa", "pull-aos", "push-soa", "push-aos",
           "blk-push-soa", "blk-pull-soa", "aa-soa")
STORAGE = {"native": 0, "fp64": 0, "fp32": 1, "fp16": 2, "bf16": 3}
INIT = {"zero": 0, "rest": 1, "random": 2}

c_int, c_double, c_char_p, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_void_p
I3 = c_int * 3


def library_path(precision="dp"):
    """the in-tree build; LBMHIP_LIBRARY_DIR selects another build directory"""
    return os.path.join(os.environ.get("LBMHIP_LIBRARY_DIR", HERE), f"liblbmhip-{precision}.so")


def driver_path(precision="dp"):
    return os.path.join(HERE, f"lbmbench-HIP-{precision}")


def _structs(pdf):
    class LatticeDesc(ctypes.Structure):
        _fields_ = [("Dims", I3), ("Lattice", ctypes.POINTER(c_int)), ("nCells", c_int),
                    ("nFluid", c_int), ("nObst", c_int), ("nInlet", c_int), ("nOutlet", c_int),
                    ("PeriodicX", c_int), ("PeriodicY", c_int), ("PeriodicZ", c_int),
                    ("Name", c_char_p)]

    class CaseData(ctypes.Structure):
        _fields_ = [("Omega", pdf), ("RhoIn", pdf), ("RhoOut", pdf), ("XForce", pdf),
                    ("MaxIterations", c_int), ("StatisticsModulus", c_int),
                    ("Extensions", c_int), ("MagicParam", c_double), ("BodyForce", c_int), ("Storage", c_int),
                    ("StorageScale", c_double), ("RhoRef", c_double),
                    ("PressureBoundaries", c_int)]

    class KernelData(ctypes.Structure):
        pass

    KDP = ctypes.POINTER(KernelData)
    LDP = ctypes.POINTER(LatticeDesc)
    CDP = ctypes.POINTER(CaseData)
    node_fn = ctypes.CFUNCTYPE(None, KDP, c_int, c_int, c_int, ctypes.POINTER(pdf))
    bcget_fn = ctypes.CFUNCTYPE(None, KDP, c_int, c_int, c_int, c_int, ctypes.POINTER(pdf))
    bcset_fn = ctypes.CFUNCTYPE(None, KDP, c_int, c_int, c_int, c_int, pdf)
    kern_fn = ctypes.CFUNCTYPE(None, LDP, KDP, CDP)
    KernelData._fields_ = [
        ("Pdfs", ctypes.POINTER(pdf) * 2), ("PdfsActive", ctypes.POINTER(pdf)),
        ("Dims", I3), ("GlobalDims", I3), ("Offsets", I3),
        ("BounceBackPdfsSrc", ctypes.POINTER(c_int)), ("BounceBackPdfsDst", ctypes.POINTER(c_int)),
        ("nBounceBackPdfs", c_int), ("Blk", I3), ("Iteration", c_int),
        ("LoopBalance", c_double), ("Duration", c_double),
        ("GetNode", node_fn), ("SetNode", node_fn),
        ("BoundaryConditionsGetPdf", bcget_fn), ("BoundaryConditionsSetPdf", bcset_fn),
        ("Kernel", kern_fn), ("Dev", c_void_p), ("Name", c_char_p)]
    init_fn = ctypes.CFUNCTYPE(None, LDP, ctypes.POINTER(KDP), CDP)
    deinit_fn = ctypes.CFUNCTYPE(None, LDP, ctypes.POINTER(KDP))

    class KernelFunctions(ctypes.Structure):
        _fields_ = [("Name", c_char_p), ("Init", init_fn), ("Deinit", deinit_fn)]

    return LatticeDesc, CaseData, KernelData, KernelFunctions


class Library:
    """One loaded build of the library ("dp": PdfType double, "sp": float)."""

    def __init__(self, precision="dp"):
        path = library_path(precision)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run make -C lbm-bench_amd")
        self.precision = precision
        self.pdf = ctypes.c_float if precision == "sp" else ctypes.c_double
        self.lib = ctypes.CDLL(path)
        (self.LatticeDesc, self.CaseData, self.KernelData,
         self.KernelFunctions) = _structs(self.pdf)
        L = self.lib
        LDP = ctypes.POINTER(self.LatticeDesc)
        KDP = ctypes.POINTER(self.KernelData)
        CDP = ctypes.POINTER(self.CaseData)
        dp = ctypes.POINTER(c_double)
        L.geometryCreate.argtypes = [c_char_p, I3, I3, LDP]
        L.geometrySetSeed.argtypes = [ctypes.c_uint64]
        L.geometryDestroy.argtypes = [LDP]
        L.caseDataDefaults.argtypes = [CDP]
        L.lbmGetNodes.argtypes = [LDP, KDP, dp]
        L.lbmSetNodes.argtypes = [LDP, KDP, dp]
        L.lbmInitPdfs.argtypes = [LDP, KDP, c_int, ctypes.c_uint64, c_double, c_double]
        L.lbmMassMomentum.argtypes = [LDP, KDP, dp]
        L.lbmCompareMomentum.argtypes = [LDP, KDP, KDP, dp]
        L.lbmCompareMomentumNoAlloc.argtypes = [LDP, KDP, KDP, dp]
        L.lbmCompareMomentumNoAlloc.restype = c_int
        L.lbmRunStepsTimed.argtypes = [LDP, KDP, CDP, c_int, dp, ctypes.POINTER(c_int)]
        L.lbmRunStepsTimed.restype = c_double
        L.lbmRunSteps.argtypes = [LDP, KDP, CDP, c_int]
        L.lbmRunSteps.restype = c_double
        L.lbmDeviceBytes.argtypes = [KDP]
        L.lbmDeviceBytes.restype = ctypes.c_size_t
        L.lbmPlacement.argtypes = [KDP, ctypes.POINTER(c_int), ctypes.POINTER(c_double)]
        L.lbmPlacement.restype = None
        L.lbmCopyBandwidth.argtypes = [ctypes.c_size_t, c_int]
        L.lbmCopyBandwidth.restype = c_double
        L.lbmCopyBandwidthShape.argtypes = [ctypes.c_size_t, c_int, c_int, c_int, c_int]
        L.lbmCopyBandwidthShape.restype = c_double
        L.lbmPatternBandwidth.argtypes = [c_int, c_int, c_int, ctypes.c_size_t, c_int]
        L.lbmPatternBandwidth.restype = c_double
        L.lbmDeviceSynchronize.restype = c_int
        L.lbmSelectDevice.argtypes = [c_int]
        L.lbmSelectDevice.restype = c_int
        L.lbmDeviceInfo.argtypes = [c_char_p, c_int, ctypes.POINTER(c_int),
                                    ctypes.POINTER(ctypes.c_size_t)]
        L.kernelDensity.argtypes = [KDP, LDP]
        L.kernelDensity.restype = self.pdf
        L.kernelVerify.argtypes = [LDP, KDP, CDP, ctypes.POINTER(self.pdf)]
        L.kernelSetInitialDensity.argtypes = [LDP, KDP, CDP]
        L.kernelSetInitialVelocity.argtypes = [LDP, KDP, CDP]
        L.kernelAddBodyForce.argtypes = [KDP, LDP, CDP]
        L.kernelComputeBoundaryConditions.argtypes = [KDP, LDP, CDP]
        L.lbmPlaneVelocity.argtypes = [LDP, KDP, c_int, dp]
        L.lbmPlaneVelocity.restype = None
        L.lbmCellMoments.argtypes = [LDP, KDP, dp]
        L.lbmCellMoments.restype = None
        L.lbmSaveState.argtypes = [LDP, KDP, CDP, c_char_p]
        L.lbmSaveState.restype = c_int
        L.lbmLoadState.argtypes = [LDP, KDP, c_char_p]
        L.lbmLoadState.restype = c_int
        L.kernelComputeBoundaryConditions.restype = None
        L.lbmAdjointCreate.argtypes = [LDP, CDP, c_int]
        L.lbmAdjointCreate.restype = c_void_p
        L.lbmAdjointDestroy.argtypes = [c_void_p]
        L.lbmAdjointSetField.argtypes = [c_void_p, dp]
        L.lbmAdjointSetState.argtypes = [c_void_p, dp]
        L.lbmAdjointInitState.argtypes = [c_void_p, c_int, ctypes.c_uint64, c_double, c_double]
        L.lbmAdjointForward.argtypes = [c_void_p, c_int, dp]
        L.lbmAdjointForward.restype = c_double
        L.lbmAdjointGetState.argtypes = [c_void_p, dp]
        L.lbmAdjointGradient.argtypes = [c_void_p, c_int, ctypes.POINTER(c_int), c_int, dp, dp]
        L.lbmAdjointGradient.restype = c_double
        L.lbmAdjointStepTimes.argtypes = [c_void_p, c_int, dp]
        L.lbmAdjointDeviceBytes.argtypes = [c_void_p]
        L.lbmAdjointDeviceBytes.restype = ctypes.c_size_t
        L.lbmPorousField.argtypes = [I3, ctypes.c_uint64, c_double, c_double, c_double, dp]
        self.n_kernels = c_int.in_dll(L, "G_N_KERNELS").value
        self.registry = (self.KernelFunctions * self.n_kernels).in_dll(L, "GKernels")

    def kernel_names(self):
        return [self.registry[i].Name.decode() for i in range(self.n_kernels)]

    def device_info(self):
        name = ctypes.create_string_buffer(256)
        cus = c_int(0)
        mem = ctypes.c_size_t(0)
        rc = self.lib.lbmDeviceInfo(name, 256, ctypes.byref(cus), ctypes.byref(mem))
        if rc != 0:
            raise RuntimeError("no HIP device")
        return name.value.decode(), cus.value, mem.value

    def copy_bandwidth(self, nbytes, reps):
        return self.lib.lbmCopyBandwidth(nbytes, reps)

    def copy_bandwidth_shape(self, nbytes, reps, vectors, non_temporal, persistent_groups=0):
        """GB/s of a copy of the given shape (lbmCopyBandwidthShape)"""
        gbs = self.lib.lbmCopyBandwidthShape(nbytes, reps, vectors, int(bool(non_temporal)),
                                             persistent_groups)
        if gbs < 0:
            raise ValueError("copy_bandwidth_shape: unsupported shape")
        return gbs

    def pattern_bandwidth(self, planes, element_bytes, in_place, cells, reps):
        """GB/s of the step's loads and stores without arithmetic"""
        gbs = self.lib.lbmPatternBandwidth(planes, element_bytes, int(bool(in_place)), cells, reps)
        if gbs < 0:
            raise ValueError("pattern_bandwidth: unsupported planes / element size")
        return gbs

    def synchronize(self):
        if self.lib.lbmDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")


class Lattice:
    """A lattice bound to one registered kernel (Init ... Deinit)."""

    def __init__(self, lib, kernel, geometry, dims, periodic=(0, 0, 0), omega=1.0,
                 xforce=0.0, body_force=False, magic=1.0 / 12.0, storage="native",
                 scale=0.0, rho_ref=1.0, seed=0, pressure=None):
        """pressure=(rho_in, rho_out): the Zou-He inlet / outlet after every
        step (CaseDataType.PressureBoundaries) on the cells flagged inlet /
        outlet (set_flags)"""
        self.lib = lib
        L = lib.lib
        self.ld = lib.LatticeDesc()
        L.geometrySetSeed(seed)
        per = I3(*periodic)
        L.geometryCreate(geometry.encode(), I3(*dims), per, ctypes.byref(self.ld))
        self.dims = tuple(dims)
        self.periodic = (self.ld.PeriodicX, self.ld.PeriodicY, self.ld.PeriodicZ)
        self.cd = lib.CaseData()
        self.cd.MaxIterations = 0
        self.cd.RhoIn = 1.0
        self.cd.RhoOut = 1.0
        self.cd.StatisticsModulus = 100
        L.caseDataDefaults(ctypes.byref(self.cd))
        self.cd.Omega = omega
        self.cd.XForce = xforce
        self.cd.BodyForce = int(body_force)
        self.cd.MagicParam = magic
        self.cd.Storage = STORAGE[storage]
        self.cd.StorageScale = scale
        self.cd.RhoRef = rho_ref
        if pressure is not None:
            self.cd.RhoIn, self.cd.RhoOut = pressure
            self.cd.PressureBoundaries = 1
        names = lib.kernel_names()
        if kernel not in names:
            L.geometryDestroy(ctypes.byref(self.ld))
            raise ValueError(f"unknown kernel {kernel}")
        self.kf = lib.registry[names.index(kernel)]
        self.kd = ctypes.POINTER(lib.KernelData)()
        self.kf.Init(ctypes.byref(self.ld), ctypes.byref(self.kd), ctypes.byref(self.cd))
        self.kernel = kernel

    # -- geometry -------------------------------------------------------
    def lattice_flags(self):
        n = self.dims[0] * self.dims[1] * self.dims[2]
        a = np.ctypeslib.as_array(self.ld.Lattice, shape=(n,))
        return a.reshape(self.dims[2], self.dims[1], self.dims[0]).copy()

    def fluid_mask(self):
        """bool (nX, nY, nZ): True on non-obstacle cells"""
        return self.lattice_flags().transpose(2, 1, 0) != 0

    def set_flags(self, flags):
        """relabel non-obstacle cells (fluid / inlet / outlet, LatticeDescType
        flags, (nZ, nY, nX)); which cells are obstacles is fixed at Init"""
        flags = np.ascontiguousarray(flags, dtype=np.int32)
        if ((flags == 0) != (self.lattice_flags() == 0)).any():
            raise ValueError("set_flags: the obstacle cells cannot change after Init")
        n = self.dims[0] * self.dims[1] * self.dims[2]
        np.ctypeslib.as_array(self.ld.Lattice, shape=(n,))[:] = flags.reshape(-1)

    def cell_moments(self):
        """float64 (nX, nY, nZ, 4): rho, jx, jy, jz per cell (0 on obstacles)"""
        out = np.empty(tuple(self.dims) + (4,), dtype=np.float64)
        self.lib.lib.lbmCellMoments(ctypes.byref(self.ld), self.kd,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def plane_velocity(self, x):
        """float64 (nY, nZ): x momentum of the logical PDFs of plane x
        (lbmPlaneVelocity; 0 on obstacles)"""
        out = np.empty((self.dims[1], self.dims[2]), dtype=np.float64)
        self.lib.lib.lbmPlaneVelocity(ctypes.byref(self.ld), self.kd,
                                      x, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def save_state(self, path):
        """lbmSaveState: flags and logical PDFs to a state file"""
        if self.lib.lib.lbmSaveState(ctypes.byref(self.ld), self.kd, ctypes.byref(self.cd),
                                     os.fsencode(path)) != 0:
            raise OSError(f"lbmSaveState failed for {path}")

    def load_state(self, path):
        """lbmLoadState: PDFs from a state file of the same lattice"""
        if self.lib.lib.lbmLoadState(ctypes.byref(self.ld), self.kd, os.fsencode(path)) != 0:
            raise OSError(f"lbmLoadState failed for {path}")

    def boundary_conditions(self, rho_in, rho_out):
        """kernelComputeBoundaryConditions: Zou-He pressure inlet at x = 0 and
        outlet at x = nX-1 on the cells flagged inlet / outlet"""
        self.cd.RhoIn, self.cd.RhoOut = rho_in, rho_out
        self.lib.lib.kernelComputeBoundaryConditions(self.kd, ctypes.byref(self.ld),
                                                     ctypes.byref(self.cd))

    @property
    def n_fluid(self):
        return self.ld.nFluid

    @property
    def loop_balance(self):
        return self.kd.contents.LoopBalance

    @property
    def iteration(self):
        return self.kd.contents.Iteration

    # -- state ----------------------------------------------------------
    def set_nodes(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        assert arr.shape == self.dims + (19,)
        self.lib.lib.lbmSetNodes(ctypes.byref(self.ld), self.kd,
                                 arr.ctypes.data_as(ctypes.POINTER(c_double)))

    def get_nodes(self):
        out = np.empty(self.dims + (19,), dtype=np.float64)
        self.lib.lib.lbmGetNodes(ctypes.byref(self.ld), self.kd,
                                 out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out

    def get_node(self, x, y, z):
        buf = (self.lib.pdf * 19)()
        self.kd.contents.GetNode(self.kd, x, y, z, buf)
        return np.array(buf[:], dtype=np.float64)

    def set_node(self, x, y, z, values):
        buf = (self.lib.pdf * 19)(*[float(v) for v in values])
        self.kd.contents.SetNode(self.kd, x, y, z, buf)

    def init(self, mode="random", seed=0, rho=1.0, amp=1e-3):
        self.lib.lib.lbmInitPdfs(ctypes.byref(self.ld), self.kd, INIT[mode], seed, rho, amp)

    def init_reference(self):
        """LBM-Bench's kernelSetInitialDensity + kernelSetInitialVelocity"""
        L = self.lib.lib
        L.kernelSetInitialDensity(ctypes.byref(self.ld), self.kd, ctypes.byref(self.cd))
        L.kernelSetInitialVelocity(ctypes.byref(self.ld), self.kd, ctypes.byref(self.cd))

    # -- stepping -------------------------------------------------------
    def run(self, steps):
        """runs `steps` steps through kd->Kernel; returns hipEvent seconds"""
        self.cd.MaxIterations = steps
        self.kd.contents.Kernel(ctypes.byref(self.ld), self.kd, ctypes.byref(self.cd))
        return self.kd.contents.Duration

    def run_timed(self, steps):
        """runs `steps` steps, each launched on its own with a hipEvent after it
        -> (run seconds, {"step": (seconds, launches), "odd": (seconds, launches)})
        where "step" are the step launches (aa-soa: even steps) and "odd"
        aa-soa's odd steps"""
        t = (c_double * 2)()
        n = (c_int * 2)()
        total = self.lib.lib.lbmRunStepsTimed(ctypes.byref(self.ld), self.kd, ctypes.byref(self.cd),
                                              steps, t, n)
        return total, {"step": (t[0], n[0]), "odd": (t[1], n[1])}

    def add_body_force(self):
        self.lib.lib.kernelAddBodyForce(self.kd, ctypes.byref(self.ld), ctypes.byref(self.cd))

    # -- checks ---------------------------------------------------------
    def mass_momentum(self):
        out = np.zeros(4)
        self.lib.lib.lbmMassMomentum(ctypes.byref(self.ld), self.kd,
                                     out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out

    def density(self):
        return float(self.lib.lib.kernelDensity(self.kd, ctypes.byref(self.ld)))

    def verify(self):
        norm = self.lib.pdf(0.0)
        self.lib.lib.kernelVerify(ctypes.byref(self.ld), self.kd, ctypes.byref(self.cd),
                                  ctypes.byref(norm))
        return float(norm.value)

    def compare_momentum(self, other):
        out = np.zeros(4)
        self.lib.lib.lbmCompareMomentum(ctypes.byref(self.ld), self.kd, other.kd,
                                        out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out

    def compare_momentum_noalloc(self, other):
        """lbmCompareMomentumNoAlloc: compare_momentum without device
        allocations; raises ValueError when the library refuses the call"""
        out = np.zeros(4)
        rc = self.lib.lib.lbmCompareMomentumNoAlloc(ctypes.byref(self.ld), self.kd, other.kd,
                                                    out.ctypes.data_as(ctypes.POINTER(c_double)))
        if rc != 0:
            raise ValueError("lbmCompareMomentumNoAlloc refused the lattices")
        return out

    def device_bytes(self):
        return self.lib.lib.lbmDeviceBytes(self.kd)

    def placement(self):
        """(candidate allocations probed, GB/s of the one kept; 1, 0.0 without a probe)"""
        n, g = c_int(0), c_double(0.0)
        self.lib.lib.lbmPlacement(self.kd, ctypes.byref(n), ctypes.byref(g))
        return n.value, g.value

    def close(self):
        if self.kd:
            self.kf.Deinit(ctypes.byref(self.ld), ctypes.byref(self.kd))
            self.kd = ctypes.POINTER(self.lib.KernelData)()
        if self.ld.Lattice:
            self.lib.lib.geometryDestroy(ctypes.byref(self.ld))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def porous_field(lib, dims, seed, sigma, a=4.0, b=-2.0):
    """solid fraction s in (0, 1), float64 (nX, nY, nZ) (lbmPorousField)"""
    out = np.zeros(tuple(dims))
    lib.lib.lbmPorousField(I3(*dims), seed, sigma, a, b,
                           out.ctypes.data_as(ctypes.POINTER(c_double)))
    return out


class Adjoint:
    """The differentiation form (lbmAdjoint*): Brinkman TRT forward from a
    fixed initial state, J = sum over x = outlet_x of 1/2 rho |j|^2, and dJ/ds
    by the hand-written adjoint with multi-level checkpointing."""

    def __init__(self, lib, geometry, dims, periodic=(0, 0, 0), omega=1.0, xforce=0.0,
                 magic=1.0 / 12.0, storage="native", outlet_x=None, scale=0.0, seed=0):
        self.lib = lib
        L = lib.lib
        self.ld = lib.LatticeDesc()
        L.geometrySetSeed(seed)
        L.geometryCreate(geometry.encode(), I3(*dims), I3(*periodic), ctypes.byref(self.ld))
        self.dims = tuple(dims)
        self.cd = lib.CaseData()
        L.caseDataDefaults(ctypes.byref(self.cd))
        self.cd.Omega = omega
        self.cd.XForce = xforce
        self.cd.MagicParam = magic
        self.cd.Storage = STORAGE[storage]
        self.cd.StorageScale = scale
        self.cd.RhoRef = 1.0
        self.outlet_x = dims[0] - 1 if outlet_x is None else outlet_x
        self.h = L.lbmAdjointCreate(ctypes.byref(self.ld), ctypes.byref(self.cd), self.outlet_x)

    def fluid_mask(self):
        n = self.dims[0] * self.dims[1] * self.dims[2]
        a = np.ctypeslib.as_array(self.ld.Lattice, shape=(n,))
        return a.reshape(self.dims[2], self.dims[1], self.dims[0]).transpose(2, 1, 0) != 0

    def lattice_flags(self):
        n = self.dims[0] * self.dims[1] * self.dims[2]
        return np.ctypeslib.as_array(self.ld.Lattice, shape=(n,)).reshape(
            self.dims[2], self.dims[1], self.dims[0]).copy()

    def set_field(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        assert s.shape == self.dims
        self.lib.lib.lbmAdjointSetField(self.h, s.ctypes.data_as(ctypes.POINTER(c_double)))

    def set_state(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        self.lib.lib.lbmAdjointSetState(self.h, f.ctypes.data_as(ctypes.POINTER(c_double)))

    def init(self, mode="random", seed=0, rho=1.0, amp=1e-3):
        self.lib.lib.lbmAdjointInitState(self.h, INIT[mode], seed, rho, amp)

    def forward(self, steps):
        """-> (J, seconds)"""
        t = c_double(0.0)
        J = self.lib.lib.lbmAdjointForward(self.h, steps, ctypes.byref(t))
        return J, t.value

    def state(self):
        out = np.empty(self.dims + (19,))
        self.lib.lib.lbmAdjointGetState(self.h, out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out

    def gradient(self, steps, intervals=None):
        """-> (J, dJ/ds (nX, nY, nZ), stats {forward_s, recompute_s, adjoint_s,
        device_bytes}); intervals default to sqrt(T) checkpoints"""
        if intervals is None:
            intervals = [max(1, int(round(steps ** 0.5)))]
        iv = (c_int * max(1, len(intervals)))(*intervals)
        grad = np.zeros(self.dims)
        stats = np.zeros(4)
        J = self.lib.lib.lbmAdjointGradient(self.h, steps, iv, len(intervals),
                                            grad.ctypes.data_as(ctypes.POINTER(c_double)),
                                            stats.ctypes.data_as(ctypes.POINTER(c_double)))
        return J, grad, dict(zip(("forward_s", "recompute_s", "adjoint_s", "device_bytes"), stats))

    def step_times(self, reps=20):
        out = np.zeros(2)
        self.lib.lib.lbmAdjointStepTimes(self.h, reps, out.ctypes.data_as(ctypes.POINTER(c_double)))
        return {"forward_s": out[0], "adjoint_s": out[1]}

    def close(self):
        if self.h:
            self.lib.lib.lbmAdjointDestroy(self.h)
            self.h = None
        if self.ld.Lattice:
            self.lib.lib.geometryDestroy(ctypes.byref(self.ld))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


STATE_HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"),
                         ("dims", "<i4", 3), ("periodic", "<i4", 3), ("q", "<i4"),
                         ("iteration", "<i4"), ("omega", "<f8"), ("xforce", "<f8"),
                         ("magic_param", "<f8"), ("flag_bytes", "<u8"), ("pdf_bytes", "<u8"),
                         ("checksum", "<u8"), ("kernel", "S32")])


def read_state(path, verify=True):
    """A state file of lbmSaveState (format in include/lbm_hip.h) ->
    (header dict, flags int8 (nZ, nY, nX), logical PDFs float64
    (nX, nY, nZ, 19), memory-mapped); verify recomputes the checksum"""
    h = np.fromfile(path, dtype=STATE_HEADER, count=1)[0]
    if h["magic"] != b"LBMSTAT1" or h["version"] != 1 or h["header_bytes"] != 128:
        raise ValueError(f"{path}: not a version-1 LBM state file")
    nx, ny, nz = (int(v) for v in h["dims"])
    ncells = nx * ny * nz
    flags = np.memmap(path, dtype=np.int8, mode="r", offset=128, shape=(ncells,))
    pdfs = np.memmap(path, dtype="<f8", mode="r", offset=128 + int(h["flag_bytes"]),
                     shape=(nx, ny, nz, int(h["q"])))
    if verify:
        words = np.memmap(path, dtype="<u8", mode="r", offset=128,
                          shape=((int(h["flag_bytes"]) + int(h["pdf_bytes"])) // 8,))
        if _fnv64_words(words) != int(h["checksum"]):
            raise ValueError(f"{path}: checksum mismatch")
    header = {k: (h[k].tolist() if k not in ("magic", "kernel") else h[k].decode())
              for k in STATE_HEADER.names}
    return header, flags.reshape(nz, ny, nx), pdfs


def _fnv64_words(words, chunk=1 << 22):
    """FNV-1a over 64-bit words (offset basis 0xcbf29ce484222325, prime
    0x100000001b3), the state file checksum"""
    h = 0xCBF29CE484222325
    prime, mask = 0x100000001B3, (1 << 64) - 1
    for start in range(0, len(words), chunk):
        for w in words[start:start + chunk].tolist():
            h = ((h ^ w) * prime) & mask
    return h


"""ctypes bindings for the oracle (liboracle.so) and for the compiled
reference (_ref/libref.so, _ref/libref_sp.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker of the GPU path.

Logical PDF arrays are float64 numpy arrays of shape (nX, nY, nZ, 19);
lattice flag arrays are int32 arrays of shape (nZ, nY, nX) (x fastest in
memory, LatticeDescType.Lattice order of LBM-Bench src/lbm.h:108-111).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

KERNELS = ("pull-soa", "pull-aos", "push-soa", "push-aos",
           "blk-push-soa", "blk-pull-soa", "aa-soa")

_I3 = ctypes.c_int * 3
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a, typ):
    return a.ctypes.data_as(typ)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
    return ctypes.CDLL(path)


class Oracle:
    """The CPU restatement (oracle/lbm_oracle.c)."""

    def __init__(self, path=None):
        self.lib = _load(path or os.path.join(HERE, "liboracle.so"))
        L = self.lib
        L.oracle_geometry.argtypes = [ctypes.c_char_p, _I3, _I3, ctypes.c_uint64, _ip, _ip]
        L.oracle_geometry.restype = ctypes.c_int
        L.oracle_run.argtypes = [ctypes.c_char_p, _I3, _ip, _I3, _dp, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, _dp, _ip]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_run_bc.argtypes = [ctypes.c_char_p, _I3, _ip, _I3, _dp, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _ip]
        L.oracle_run_bc.restype = ctypes.c_int
        L.oracle_init_random.argtypes = [_I3, _ip, ctypes.c_uint64, ctypes.c_double,
                                         ctypes.c_double, _dp]
        L.oracle_init_random.restype = None
        L.oracle_density.argtypes = [_I3, _ip, _dp]
        L.oracle_density.restype = ctypes.c_double
        L.oracle_verify.argtypes = [_I3, _ip, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.oracle_verify.restype = ctypes.c_double
        L.oracle_moments.argtypes = [_I3, _ip, _dp, _dp]
        L.oracle_moments.restype = None
        L.oracle_hash.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_hash.restype = ctypes.c_uint64
        L.oracle_brinkman.argtypes = [_I3, _ip, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.oracle_brinkman.restype = ctypes.c_double
        L.oracle_zou_he.argtypes = [_I3, _ip, _dp, ctypes.c_double, ctypes.c_double]
        L.oracle_zou_he.restype = None
        L.oracle_porous_field.argtypes = [_I3, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, _dp]
        L.oracle_porous_field.restype = None

    def geometry(self, gtype, dims, periodic=(0, 0, 0), seed=0):
        """-> (lattice[nZ,nY,nX] int32, periodic tuple, (nFluid, nObst))"""
        nx, ny, nz = dims
        lat = np.empty((nz, ny, nx), dtype=np.int32)
        per = _I3(*periodic)
        counts = (ctypes.c_int * 2)()
        rc = self.lib.oracle_geometry(gtype.encode(), _I3(*dims), per, seed,
                                      _ptr(lat, _ip), counts)
        if rc != 0:
            raise ValueError(f"unknown geometry {gtype}")
        return lat, tuple(per), (counts[0], counts[1])

    def run(self, kernel, dims, lattice, periodic, init, iterations, omega,
            magic=1.0 / 12.0, xforce=0.0, bodyforce=False):
        """-> (logical pdfs after the run, steps actually run)"""
        out = np.zeros(tuple(dims) + (19,), dtype=np.float64)
        lat = np.ascontiguousarray(lattice, dtype=np.int32)
        ip = None
        if init is not None:
            init = np.ascontiguousarray(init, dtype=np.float64)
            ip = _ptr(init, _dp)
        done = ctypes.c_int(0)
        rc = self.lib.oracle_run(kernel.encode(), _I3(*dims), _ptr(lat, _ip),
                                 _I3(*periodic), ip, iterations, omega, magic,
                                 xforce, int(bodyforce), _ptr(out, _dp),
                                 ctypes.byref(done))
        if rc != 0:
            raise ValueError(f"unknown kernel {kernel}")
        return out, done.value

    def run_bc(self, kernel, dims, lattice, periodic, init, iterations, omega, magic,
               xforce, bodyforce, rho_in, rho_out):
        """run() with the Zou-He inlet / outlet (flags 2 / 4 of lattice) after
        every step -> (logical pdfs, steps actually run)"""
        out = np.zeros(tuple(dims) + (19,), dtype=np.float64)
        lat = np.ascontiguousarray(lattice, dtype=np.int32)
        init = np.ascontiguousarray(init, dtype=np.float64)
        done = ctypes.c_int(0)
        rc = self.lib.oracle_run_bc(kernel.encode(), _I3(*dims), _ptr(lat, _ip), _I3(*periodic),
                                    _ptr(init, _dp), iterations, omega, magic, xforce,
                                    int(bodyforce), rho_in, rho_out, _ptr(out, _dp),
                                    ctypes.byref(done))
        if rc != 0:
            raise ValueError(f"unknown kernel {kernel}")
        return out, done.value

    def init_random(self, dims, lattice, seed, amp, rho=1.0):
        out = np.zeros(tuple(dims) + (19,), dtype=np.float64)
        lat = np.ascontiguousarray(lattice, dtype=np.int32)
        self.lib.oracle_init_random(_I3(*dims), _ptr(lat, _ip), seed, amp, rho,
                                    _ptr(out, _dp))
        return out

    def density(self, dims, lattice, pdfs):
        lat = np.ascontiguousarray(lattice, dtype=np.int32)
        pdfs = np.ascontiguousarray(pdfs, dtype=np.float64)
        return self.lib.oracle_density(_I3(*dims), _ptr(lat, _ip), _ptr(pdfs, _dp))

    def verify(self, dims, lattice, pdfs, omega, xforce, force_correction=True):
        lat = np.ascontiguousarray(lattice, dtype=np.int32)
        pdfs = np.ascontiguousarray(pdfs, dtype=np.float64)
        return self.lib.oracle_verify(_I3(*dims), _ptr(lat, _ip), _ptr(pdfs, _dp),
                                      omega, xforce, int(force_correction))

    def moments(self, dims, lattice, pdfs):
        lat = np.ascontiguousarray(lattice, dtype=np.int32)
        pdfs = np.ascontiguousarray(pdfs, dtype=np.float64)
        out = np.zeros(4)
        self.lib.oracle_moments(_I3(*dims), _ptr(lat, _ip), _ptr(pdfs, _dp), _ptr(out, _dp))
        return out

    def hash(self, seed, a, b):
        return self.lib.oracle_hash(seed, a, b)

    def brinkman(self, dims, lattice, s, init, steps, omega, magic, xforce, outlet_x,
                 gradient=False):
        """-> (J, final state, dJ/ds or None) of the Brinkman TRT model"""
        lat = np.ascontiguousarray(lattice, dtype=np.int32)
        s = np.ascontiguousarray(s, dtype=np.float64)
        init = np.ascontiguousarray(init, dtype=np.float64)
        final = np.zeros(tuple(dims) + (19,))
        grad = np.zeros(tuple(dims)) if gradient else None
        J = self.lib.oracle_brinkman(_I3(*dims), _ptr(lat, _ip), _ptr(s, _dp), _ptr(init, _dp),
                                     steps, omega, magic, xforce, outlet_x, _ptr(final, _dp),
                                     _ptr(grad, _dp) if gradient else None)
        return J, final, grad

    def zou_he(self, dims, lattice, pdfs, rho_in, rho_out):
        """Zou-He inlet (x = 0, flag 2) / outlet (x = nX-1, flag 4) on a copy
        of the logical PDFs (src/lbm.c:676-785)"""
        out = np.ascontiguousarray(pdfs, dtype=np.float64).copy()
        lat = np.ascontiguousarray(lattice, dtype=np.int32)
        self.lib.oracle_zou_he(_I3(*dims), _ptr(lat, _ip), _ptr(out, _dp), rho_in, rho_out)
        return out

    def porous_field(self, dims, seed, sigma, a=4.0, b=-2.0):
        out = np.zeros(tuple(dims))
        self.lib.oracle_porous_field(_I3(*dims), seed, sigma, a, b, _ptr(out, _dp))
        return out


class Reference:
    """LBM-Bench's own sources compiled into _ref/ (ref_harness.c); omp
    selects the OpenMP build (the reference's parallel push-soa and
    pull-soa loops)."""

    def __init__(self, sp=False, omp=False):
        name = "libref_sp.so" if sp else ("libref_omp.so" if omp else "libref.so")
        self.lib = _load(os.path.join(REF_DIR, name))
        L = self.lib
        L.ref_geometry.argtypes = [ctypes.c_char_p, _I3, _I3, _ip, _ip]
        L.ref_geometry.restype = ctypes.c_int
        L.ref_run.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _I3, _I3, _dp, ctypes.c_int,
                              ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.ref_run.restype = ctypes.c_int
        L.ref_zou_he.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _I3, _I3, _dp, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, _dp]
        L.ref_zou_he.restype = ctypes.c_int
        L.ref_run_bc.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _I3, _I3, _dp, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, _dp]
        L.ref_run_bc.restype = ctypes.c_int
        L.ref_verification_norm.argtypes = [ctypes.c_char_p]
        L.ref_verification_norm.restype = ctypes.c_double

    @staticmethod
    def available(sp=False, omp=False):
        name = "libref_sp.so" if sp else ("libref_omp.so" if omp else "libref.so")
        return os.path.exists(os.path.join(REF_DIR, name))

    def geometry(self, gtype, dims, periodic=(0, 0, 0)):
        nx, ny, nz = dims
        lat = np.empty((nz, ny, nx), dtype=np.int32)
        per = _I3(*periodic)
        counts = (ctypes.c_int * 2)()
        self.lib.ref_geometry(gtype.encode(), _I3(*dims), per, _ptr(lat, _ip), counts)
        return lat, tuple(per), (counts[0], counts[1])

    def run(self, kernel, geometry, dims, periodic, init, iterations, omega, xforce,
            ref_init=False):
        """-> (logical pdfs, (density, verify norm, duration))"""
        out = np.zeros(tuple(dims) + (19,), dtype=np.float64)
        stats = np.zeros(3)
        ip = None
        if init is not None:
            init = np.ascontiguousarray(init, dtype=np.float64)
            ip = _ptr(init, _dp)
        rc = self.lib.ref_run(kernel.encode(), geometry.encode(), _I3(*dims), _I3(*periodic),
                              ip, int(ref_init), iterations, omega, xforce,
                              _ptr(out, _dp), _ptr(stats, _dp))
        if rc != 0:
            raise ValueError(f"unknown kernel {kernel}")
        return out, tuple(stats)

    def zou_he(self, kernel, geometry, dims, periodic, init, iterations, omega, xforce,
               rho_in, rho_out):
        """a run, then the x faces marked inlet/outlet and the reference's
        kernelComputeBoundaryConditions once -> logical pdfs"""
        out = np.zeros(tuple(dims) + (19,), dtype=np.float64)
        init = np.ascontiguousarray(init, dtype=np.float64)
        rc = self.lib.ref_zou_he(kernel.encode(), geometry.encode(), _I3(*dims), _I3(*periodic),
                                 _ptr(init, _dp), iterations, omega, xforce, rho_in, rho_out,
                                 _ptr(out, _dp))
        if rc != 0:
            raise ValueError(f"unknown kernel {kernel}")
        return out

    def run_bc(self, kernel, geometry, dims, periodic, init, iterations, omega, xforce,
               rho_in, rho_out):
        """x faces marked inlet/outlet, then per step the kernel's one-step run
        and the reference's kernelComputeBoundaryConditions -> logical pdfs
        (two-lattice kernels only)"""
        out = np.zeros(tuple(dims) + (19,), dtype=np.float64)
        init = np.ascontiguousarray(init, dtype=np.float64)
        rc = self.lib.ref_run_bc(kernel.encode(), geometry.encode(), _I3(*dims), _I3(*periodic),
                                 _ptr(init, _dp), iterations, omega, xforce, rho_in, rho_out,
                                 _ptr(out, _dp))
        if rc != 0:
            raise ValueError(f"ref_run_bc: kernel {kernel} (rc {rc})")
        return out

    def verification_norm(self, kernel):
        return self.lib.ref_verification_norm(kernel.encode())

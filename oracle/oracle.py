"""ctypes view of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only as the checker.  The HIP engine never calls it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dptr = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
iptr = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class ScfOptions(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("energy_tol", C.c_double), ("residual_tol", C.c_double),
                ("density_tol", C.c_double), ("conv_algo", C.c_int), ("spherical", C.c_int),
                ("direct", C.c_int), ("verbose", C.c_int)]


class ScfResult(C.Structure):
    _fields_ = [("total_energy", C.c_double), ("electronic_energy", C.c_double),
                ("nuclear_repulsion", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("n_occ", C.c_int), ("fock_builds", C.c_int), ("seconds_eri", C.c_double),
                ("seconds_fock", C.c_double), ("seconds_total", C.c_double)]


class DfhfResult(C.Structure):
    _fields_ = [("total_energy", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("n_aux", C.c_int), ("residual_norm", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.hfo_last_error.restype = C.c_char_p
        L.hfo_boys.argtypes = [C.c_int, C.c_double, dptr]
        L.hfo_basis_build.argtypes = [C.c_char_p, C.c_char_p, dptr, C.c_int, C.POINTER(C.c_void_p)]
        L.hfo_basis_free.argtypes = [C.c_void_p]
        L.hfo_basis_nfunc.argtypes = [C.c_void_p]
        L.hfo_basis_nprim_total.argtypes = [C.c_void_p]
        L.hfo_basis_export.argtypes = [C.c_void_p, iptr, dptr, iptr, iptr, dptr, dptr]
        L.hfo_basis_from_arrays.argtypes = [C.c_int, iptr, dptr, iptr, iptr, dptr, dptr, C.POINTER(C.c_void_p)]
        L.hfo_nshells.argtypes = [C.c_void_p, C.c_void_p]
        L.hfo_shell_quartet.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, dptr]
        L.hfo_contracted_eri.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.hfo_contracted_eri.restype = C.c_double
        L.hfo_eri_packed.argtypes = [C.c_void_p, dptr, C.c_size_t]
        L.hfo_schwarz.argtypes = [C.c_void_p, dptr]
        L.hfo_contract_jk.argtypes = [dptr, dptr, C.c_int, dptr, dptr]
        L.hfo_direct_jk.argtypes = [C.c_void_p, dptr, dptr, dptr, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.hfo_one_electron.argtypes = [C.c_void_p, dptr, C.c_int, dptr, dptr, dptr]
        L.hfo_eigh.argtypes = [dptr, C.c_int, dptr, dptr]
        L.hfo_spherical_transform.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        L.hfo_scf_rhf.argtypes = [C.c_void_p, dptr, C.c_int, C.c_int, C.POINTER(ScfOptions),
                                  C.POINTER(ScfResult), C.c_void_p, C.c_void_p, C.c_void_p]
        L.hfo_sad_density.argtypes = [C.c_void_p, dptr, C.c_int, C.c_int, dptr, dptr]
        L.hfo_atomic_hf.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_void_p]
        L.hfo_df_metric.argtypes = [C.c_void_p, dptr]
        L.hfo_df_three_center.argtypes = [C.c_void_p, C.c_void_p, dptr]
        L.hfo_df_build.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_int),
                                   C.POINTER(C.POINTER(C.c_double))]
        L.hfo_free.argtypes = [C.c_void_p]
        L.hfo_df_jk.argtypes = [dptr, C.c_int, C.c_int, dptr, C.c_int, dptr, dptr]
        L.hfo_df_rhf.argtypes = [C.c_void_p, C.c_void_p, dptr, C.c_int, C.c_int, C.c_int, C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.POINTER(DfhfResult), C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
        L.hfo_rhf_gradient.argtypes = [C.c_void_p, dptr, C.c_int, dptr, dptr, dptr]
        L.hfo_cholesky_eri.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_int),
                                       C.POINTER(C.c_double), C.c_void_p, C.c_int,
                                       C.POINTER(C.POINTER(C.c_double))]
        L.hfo_cd_rhf.argtypes = [C.c_void_p, dptr, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                 C.POINTER(DfhfResult), C.c_void_p]
        L.hfo_md_flops.argtypes = [C.c_int] * 4
        L.hfo_md_flops.restype = C.c_double
        L.hfo_fock_flops.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.hfo_fock_flops_sample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _LIB = L
    return _LIB


def _check(status):
    if status != 0:
        raise RuntimeError("oracle: %s" % lib().hfo_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def boys(m, t):
    out = np.zeros(m + 1)
    _check(lib().hfo_boys(m, t, out))
    return out


class Basis:
    """A Cartesian basis (one row per ContractedGaussian, src/basis.rs:89-98)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.n = lib().hfo_basis_nfunc(self.h)

    @classmethod
    def build(cls, g94_text, symbols, xyz_bohr):
        xyz = np.ascontiguousarray(np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3))
        h = C.c_void_p()
        _check(lib().hfo_basis_build(g94_text.encode(), " ".join(symbols).encode(), xyz.ravel().copy(),
                                     len(symbols), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, arrays):
        h = C.c_void_p()
        a = arrays
        _check(lib().hfo_basis_from_arrays(len(a["atom_index"]), a["atom_index"], a["center"].ravel().copy(),
                                           a["lxyz"].ravel().copy(), a["prim_offset"], a["exponent"],
                                           a["coefficient"], C.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.hfo_basis_free(self.h)
            self.h = None

    def arrays(self):
        n = self.n
        npt = lib().hfo_basis_nprim_total(self.h)
        out = dict(atom_index=np.zeros(n, np.int32), center=np.zeros(3 * n), lxyz=np.zeros(3 * n, np.int32),
                   prim_offset=np.zeros(n + 1, np.int32), exponent=np.zeros(npt), coefficient=np.zeros(npt))
        _check(lib().hfo_basis_export(self.h, out["atom_index"], out["center"], out["lxyz"], out["prim_offset"],
                                      out["exponent"], out["coefficient"]))
        out["center"] = out["center"].reshape(n, 3)
        out["lxyz"] = out["lxyz"].reshape(n, 3)
        return out

    def shells(self):
        ns = lib().hfo_nshells(self.h, None)
        ss = np.zeros(ns + 1, np.int32)
        lib().hfo_nshells(self.h, ss.ctypes.data_as(C.c_void_p))
        return ss

    def shell_quartet(self, a, b, c, d):
        ss = self.shells()
        dims = [ss[s + 1] - ss[s] for s in (a, b, c, d)]
        out = np.zeros(int(np.prod(dims)))
        _check(lib().hfo_shell_quartet(self.h, a, b, c, d, out))
        return out.reshape(dims)

    def contracted_eri(self, i, j, k, l):
        return lib().hfo_contracted_eri(self.h, i, j, k, l)

    def eri_packed(self):
        np_ = self.n * (self.n + 1) // 2
        out = np.zeros(np_ * (np_ + 1) // 2)
        _check(lib().hfo_eri_packed(self.h, out, out.size))
        return out

    def schwarz(self):
        ns = len(self.shells()) - 1
        q = np.zeros(ns * (ns + 1) // 2)
        _check(lib().hfo_schwarz(self.h, q))
        return q

    def direct_jk(self, p, stride=1, offset=0):
        p = np.ascontiguousarray(p, dtype=np.float64)
        j = np.zeros((self.n, self.n))
        k = np.zeros((self.n, self.n))
        sec = C.c_double()
        _check(lib().hfo_direct_jk(self.h, p, j, k, stride, offset, C.byref(sec)))
        return j, k, sec.value

    def one_electron(self, zxyz):
        zxyz = np.ascontiguousarray(zxyz, dtype=np.float64)
        s, t, v = (np.zeros((self.n, self.n)) for _ in range(3))
        _check(lib().hfo_one_electron(self.h, zxyz.ravel(), len(zxyz), s, t, v))
        return s, t, v

    def spherical_transform(self):
        nc = C.c_int()
        _check(lib().hfo_spherical_transform(self.h, None, C.byref(nc)))
        t = np.zeros((self.n, nc.value))
        _check(lib().hfo_spherical_transform(self.h, t.ctypes.data_as(C.c_void_p), C.byref(nc)))
        return t

    def fock_flops(self, stride=1, offset=0):
        f, q, pq = C.c_double(), C.c_double(), C.c_double()
        _check(lib().hfo_fock_flops_sample(self.h, stride, offset, C.byref(f), C.byref(q), C.byref(pq)))
        return f.value, q.value, pq.value


def contract_jk(p, eri, n):
    j = np.zeros((n, n))
    k = np.zeros((n, n))
    _check(lib().hfo_contract_jk(np.ascontiguousarray(p), np.ascontiguousarray(eri), n, j, k))
    return j, k


def eigh(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n))
    _check(lib().hfo_eigh(a, n, w, v))
    return w, v


CONV = {"adiissoscf": 0, "default": 0, "cdiis": 1, "diis": 1, "adiis": 2, "adiiscdiis": 3}


def scf_rhf(basis, zxyz, charge=0, max_iter=100, tol=1e-10, residual=None, density=None, conv="cdiis",
            spherical=False, direct=False, verbose=False):
    """RHF with the derived tolerances of run_scf_in_space (src/scf.rs:384-400)."""
    import math
    if residual is None:
        residual = max(math.sqrt(tol) * 1e-2, 1e-11)
    if density is None:
        density = min(max(math.sqrt(tol) * 0.1, 1e-12), 1e-10)
    opt = ScfOptions(max_iter, tol, residual, density, CONV[conv.replace("-", "").replace("_", "").lower()],
                     int(spherical), int(direct), int(verbose))
    res = ScfResult()
    n = basis.n
    p = np.zeros((n, n))
    f = np.zeros((n, n))
    eps = np.zeros(n)
    zxyz = np.ascontiguousarray(zxyz, dtype=np.float64)
    _check(lib().hfo_scf_rhf(basis.h, zxyz.ravel(), len(zxyz), charge, C.byref(opt), C.byref(res), _ptr(p),
                             _ptr(f), _ptr(eps)))
    return res, p, f, eps


def sad_density(basis, zxyz, s, charge=0):
    zxyz = np.ascontiguousarray(zxyz, dtype=np.float64)
    p = np.zeros((basis.n, basis.n))
    _check(lib().hfo_sad_density(basis.h, zxyz.ravel(), len(zxyz), charge, np.ascontiguousarray(s), p))
    return p


def atomic_hf(basis, z):
    e = C.c_double()
    d = np.zeros((basis.n, basis.n))
    _check(lib().hfo_atomic_hf(basis.h, z, C.byref(e), _ptr(d)))
    return e.value, d


def df_metric(aux):
    m = np.zeros((aux.n, aux.n))
    _check(lib().hfo_df_metric(aux.h, m))
    return m


def df_three_center(basis, aux):
    t = np.zeros((aux.n, basis.n * (basis.n + 1) // 2))
    _check(lib().hfo_df_three_center(basis.h, aux.h, t))
    return t


def df_build(basis, aux, cutoff=1e-10):
    rank = C.c_int()
    ptr = C.POINTER(C.c_double)()
    _check(lib().hfo_df_build(basis.h, aux.h, cutoff, C.byref(rank), C.byref(ptr)))
    np_ = basis.n * (basis.n + 1) // 2
    b = np.ctypeslib.as_array(ptr, shape=(rank.value * np_,)).copy().reshape(rank.value, np_)
    lib().hfo_free(ptr)
    return b


def df_jk(bt, n, cocc):
    cocc = np.ascontiguousarray(cocc, dtype=np.float64)
    j = np.zeros((n, n))
    k = np.zeros((n, n))
    _check(lib().hfo_df_jk(np.ascontiguousarray(bt), bt.shape[0], n, cocc, cocc.shape[1], j, k))
    return j, k


def df_rhf(basis, aux, zxyz, charge=0, max_iter=100, e_tol=1e-12, r_tol=1e-10, d_tol=1e-10, cutoff=1e-10):
    zxyz = np.ascontiguousarray(zxyz, dtype=np.float64)
    res = DfhfResult()
    n = basis.n
    p, f, j, k = (np.zeros((n, n)) for _ in range(4))
    _check(lib().hfo_df_rhf(basis.h, aux.h, zxyz.ravel(), len(zxyz), charge, max_iter, e_tol, r_tol, d_tol, cutoff,
                            C.byref(res), _ptr(p), _ptr(f), _ptr(j), _ptr(k)))
    return res, p, f, j, k


def rhf_gradient(basis, zxyz, p, w):
    """Nuclear gradient (natom x 3) for density p and energy-weighted density w
    (compute_gradient_from_density, src/gradient.rs:287-512)."""
    zxyz = np.ascontiguousarray(zxyz, dtype=np.float64)
    g = np.zeros((len(zxyz), 3))
    _check(lib().hfo_rhf_gradient(basis.h, zxyz.ravel(), len(zxyz), np.ascontiguousarray(p, dtype=np.float64),
                                  np.ascontiguousarray(w, dtype=np.float64), g))
    return g


CD_BUDGET = 512 * 1024 * 1024  # bytes of stored factors (src/hf_variants/cholesky_df.rs:22)


def cholesky_eri(basis, threshold, budget=CD_BUDGET):
    """Pivoted Cholesky factors of the AO ERIs: (factors rank x np packed,
    max residual diagonal, pivot list)."""
    np_ = basis.n * (basis.n + 1) // 2
    rank = C.c_int()
    mr = C.c_double()
    piv = np.zeros(np_, np.int32)
    ptr = C.POINTER(C.c_double)()
    _check(lib().hfo_cholesky_eri(basis.h, threshold, float(budget), C.byref(rank), C.byref(mr), _ptr(piv), np_,
                                  C.byref(ptr)))
    b = np.ctypeslib.as_array(ptr, shape=(rank.value * np_,)).copy().reshape(rank.value, np_)
    lib().hfo_free(ptr)
    return b, mr.value, [int(x) for x in piv[:rank.value]]


def cd_rhf(basis, zxyz, delta, charge=0, max_iter=300, tol=1e-12):
    """CD-HF energy (run_cd_hf): (result, density)."""
    zxyz = np.ascontiguousarray(zxyz, dtype=np.float64)
    res = DfhfResult()
    p = np.zeros((basis.n, basis.n))
    _check(lib().hfo_cd_rhf(basis.h, zxyz.ravel(), len(zxyz), charge, delta, max_iter, tol, C.byref(res), _ptr(p)))
    return res, p


def md_flops(la, lb, lc, ld):
    return lib().hfo_md_flops(la, lb, lc, ld)

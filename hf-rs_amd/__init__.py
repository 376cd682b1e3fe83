"""hf_rs_amd -- Python view of the MI355X Fock-build engine (libhfg.so).

Mirrors the hf-rs calls that reach the Fock-build path: build_basis
(src/basis.rs:402-452), compute_shell_quartet (src/shell_integrals.rs:138),
compute_eri_tensor (src/scf.rs:275), contract_jk / build_fock_matrix
(src/scf.rs:337-369) and the RHF driver run_scf_in_space_with_convergence
(src/scf.rs:410-627), through the C ABI of include/hfg.h.  There is no CPU
fallback: importing works without a GPU, but every compute call needs the
compiled library and a visible device, and raises otherwise.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhfg.so")
_LIB = None

HFG_OK, HFG_INVALID_INPUT, HFG_INTEGRAL, HFG_DEVICE, HFG_NOT_CONVERGED, HFG_UNSUPPORTED = range(6)
HFG_SERVE_KEEP_DENSITY, HFG_SERVE_GRADIENT = 1, 2  # hfg_server_submit flags


class HfError(RuntimeError):
    """Mirror of hf-rs HfError (src/error.rs:5-15); .code is the C status."""

    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


class BasisView(C.Structure):
    _fields_ = [("n_functions", C.c_int32), ("atom_index", C.c_void_p), ("center", C.c_void_p),
                ("lxyz", C.c_void_p), ("prim_offset", C.c_void_p), ("exponent", C.c_void_p),
                ("coefficient", C.c_void_p)]


class ScfOptions(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("energy_tol", C.c_double), ("residual_tol", C.c_double),
                ("density_tol", C.c_double), ("conv_algo", C.c_int32), ("spherical", C.c_int32),
                ("verbose", C.c_int32), ("incremental_fock", C.c_int32)]


class ScfResult(C.Structure):
    _fields_ = [("total_energy", C.c_double), ("electronic_energy", C.c_double),
                ("nuclear_repulsion", C.c_double), ("iterations", C.c_int32), ("converged", C.c_int32),
                ("n_occ", C.c_int32), ("fock_builds", C.c_int32), ("seconds_total", C.c_double),
                ("seconds_fock", C.c_double), ("seconds_setup", C.c_double), ("seconds_one_electron", C.c_double),
                ("seconds_guess", C.c_double), ("seconds_linalg", C.c_double)]


EXPORTED = ["hfg_last_error", "hfg_device_count", "hfg_ctx_create", "hfg_ctx_destroy", "hfg_ctx_info",
            "hfg_schwarz", "hfg_shell_quartet", "hfg_eri_packed", "hfg_contract_jk", "hfg_build_fock",
            "hfg_contract_jk_device", "hfg_scf_rhf", "hfg_basis_from_g94", "hfg_basis_free",
            "hfg_basis_get_view", "hfg_set_density", "hfg_jk_resident", "hfg_get_jk", "hfg_profile",
            "hfg_class_stats", "hfg_fock_flops", "hfg_md_flops", "hfg_persistent_blocks", "hfg_df_build",
            "hfg_df_destroy", "hfg_df_info", "hfg_df_jk", "hfg_df_rhf", "hfg_calibration_copy",
            "hfg_batch_create", "hfg_batch_destroy", "hfg_batch_info", "hfg_batch_set_density", "hfg_batch_jk",
            "hfg_batch_get_jk", "hfg_batch_jk_each", "hfg_df_jk_ex", "hfg_contract_jk_multi", "hfg_server_create",
            "hfg_server_destroy", "hfg_server_submit", "hfg_server_run", "hfg_server_result", "hfg_server_stats",
            "hfg_scf_rhf_guess", "hfg_eri_mo", "hfg_batch_profile", "hfg_batch_class_stats", "hfg_cd_build",
            "hfg_cd_rhf", "hfg_rhf_gradient", "hfg_graph_launches", "hfg_one_electron", "hfg_sad_density",
            "hfg_server_gradient"]


class ServerOptions(C.Structure):
    _fields_ = [("capacity", C.c_int32), ("hbm_budget_gb", C.c_double)]


class StepResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("iterations", C.c_int32), ("converged", C.c_int32),
                ("fock_builds", C.c_int32), ("n", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admitted_s", C.c_double), ("finished_s", C.c_double),
                ("steps", C.c_int32), ("converged", C.c_int32), ("status", C.c_int32), ("n", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("fused_builds", C.c_int64), ("iterations", C.c_int64), ("seconds", C.c_double),
                ("seconds_setup", C.c_double), ("seconds_sync", C.c_double), ("hbm_budget_bytes", C.c_double),
                ("jk_config", C.c_int32), ("jk_candidates", C.c_int32)]


class DfhfOptions(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("energy_tol", C.c_double), ("residual_tol", C.c_double),
                ("density_tol", C.c_double), ("metric_cutoff", C.c_double), ("k_precision", C.c_int32),
                ("switch_residual", C.c_double)]


class CdInfo(C.Structure):
    _fields_ = [("rank", C.c_int32), ("max_residual", C.c_double)]


class DfhfResult(C.Structure):
    _fields_ = [("total_energy", C.c_double), ("iterations", C.c_int32), ("converged", C.c_int32),
                ("n_aux", C.c_int32), ("fock_builds", C.c_int32), ("seconds_total", C.c_double),
                ("seconds_fock", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhfg.so is not built: run __graft_entry__.build() (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        L.hfg_last_error.restype = C.c_char_p
        vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
        L.hfg_ctx_create.argtypes = [C.POINTER(BasisView), C.POINTER(vp)]
        L.hfg_ctx_destroy.argtypes = [vp]
        L.hfg_ctx_info.argtypes = [vp] + [C.POINTER(C.c_int64)] * 4
        L.hfg_schwarz.argtypes = [vp, vp]
        L.hfg_shell_quartet.argtypes = [vp, i32, i32, i32, i32, vp]
        L.hfg_eri_packed.argtypes = [vp, vp, C.c_size_t]
        L.hfg_eri_mo.argtypes = [vp, vp, i32, vp, C.c_size_t]
        L.hfg_contract_jk.argtypes = [vp, vp, vp, vp]
        L.hfg_build_fock.argtypes = [vp, vp, vp, vp]
        L.hfg_contract_jk_multi.argtypes = [vp, vp, i32, vp, vp, vp]
        L.hfg_contract_jk_device.argtypes = [vp, vp, vp, vp]
        L.hfg_scf_rhf.argtypes = [C.POINTER(BasisView), vp, i32, i32, C.POINTER(ScfOptions), C.POINTER(ScfResult),
                                  vp, vp, vp]
        L.hfg_scf_rhf_guess.argtypes = [C.POINTER(BasisView), vp, i32, i32, C.POINTER(ScfOptions), vp,
                                        C.POINTER(ScfResult), vp, vp, vp]
        L.hfg_basis_from_g94.argtypes = [C.c_char_p, C.c_char_p, vp, i32, C.POINTER(vp)]
        L.hfg_basis_free.argtypes = [vp]
        L.hfg_basis_get_view.argtypes = [vp, C.POINTER(BasisView)]
        L.hfg_set_density.argtypes = [vp, vp]
        L.hfg_jk_resident.argtypes = [vp]
        L.hfg_get_jk.argtypes = [vp, vp, vp]
        L.hfg_profile.argtypes = [vp, i32]
        L.hfg_class_stats.argtypes = [vp, i32, vp, C.POINTER(dbl), C.POINTER(C.c_int64), C.POINTER(dbl),
                                      C.POINTER(C.c_int64)]
        L.hfg_fock_flops.argtypes = [vp, C.POINTER(dbl), C.POINTER(dbl)]
        L.hfg_md_flops.argtypes = [i32] * 4
        L.hfg_md_flops.restype = dbl
        L.hfg_calibration_copy.argtypes = [C.c_size_t, C.POINTER(dbl)]
        L.hfg_persistent_blocks.restype = i32
        L.hfg_df_build.argtypes = [C.POINTER(BasisView), C.POINTER(BasisView), dbl, C.POINTER(vp)]
        L.hfg_df_destroy.argtypes = [vp]
        L.hfg_df_info.argtypes = [vp] + [C.POINTER(C.c_int64)] * 3
        L.hfg_df_jk.argtypes = [vp, vp, i32, vp, vp]
        L.hfg_df_jk_ex.argtypes = [vp, vp, i32, vp, vp, i32, C.POINTER(dbl)]
        L.hfg_df_rhf.argtypes = [C.POINTER(BasisView), C.POINTER(BasisView), vp, i32, i32, C.POINTER(DfhfOptions),
                                 C.POINTER(DfhfResult), vp, vp, vp, vp]
        L.hfg_rhf_gradient.argtypes = [vp, vp, i32, vp, vp, vp]
        L.hfg_graph_launches.argtypes = [vp, C.POINTER(C.c_int64)]
        L.hfg_one_electron.argtypes = [vp, vp, i32, vp, vp, vp]
        L.hfg_sad_density.argtypes = [vp, vp, i32, i32, vp]
        L.hfg_cd_build.argtypes = [C.POINTER(BasisView), dbl, dbl, C.POINTER(vp), C.POINTER(CdInfo), vp, i32]
        L.hfg_cd_rhf.argtypes = [C.POINTER(BasisView), vp, i32, i32, dbl, dbl, C.POINTER(DfhfOptions),
                                 C.POINTER(DfhfResult), C.POINTER(CdInfo), vp, vp]
        L.hfg_batch_create.argtypes = [vp, i32, C.POINTER(vp)]
        L.hfg_batch_destroy.argtypes = [vp]
        L.hfg_batch_info.argtypes = [vp] + [C.POINTER(C.c_int64)] * 3 + [C.POINTER(dbl)]
        L.hfg_batch_set_density.argtypes = [vp, i32, vp]
        L.hfg_batch_jk.argtypes = [vp]
        L.hfg_batch_jk_each.argtypes = [vp]
        L.hfg_batch_get_jk.argtypes = [vp, i32, vp, vp]
        L.hfg_batch_profile.argtypes = [vp, i32]
        L.hfg_batch_class_stats.argtypes = [vp, i32, vp, C.POINTER(dbl), C.POINTER(C.c_int64), C.POINTER(dbl),
                                            C.POINTER(C.c_int64)]
        L.hfg_server_create.argtypes = [C.POINTER(ServerOptions), C.POINTER(vp)]
        L.hfg_server_destroy.argtypes = [vp]
        L.hfg_server_submit.argtypes = [vp, dbl, i32, vp, vp, i32, i32, C.POINTER(ScfOptions), i32,
                                        C.POINTER(C.c_int)]
        L.hfg_server_run.argtypes = [vp]
        L.hfg_server_result.argtypes = [vp, i32, C.POINTER(RequestResult), vp, vp]
        L.hfg_server_gradient.argtypes = [vp, i32, vp]
        L.hfg_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
        _LIB = L
    return _LIB


def _check(status):
    if status != HFG_OK:
        raise HfError(status, lib().hfg_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _square(a, n, name):
    """a as a C-contiguous float64 (n, n) array; ValueError otherwise (the C
    side reads exactly n*n doubles)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != (n, n):
        raise ValueError("%s must have shape (%d, %d), got %s" % (name, n, n, a.shape))
    return a


def _columns(a, n, name):
    """a as a C-contiguous float64 (n, k) array with k >= 1."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != n or a.shape[1] < 1:
        raise ValueError("%s must have shape (%d, k) with k >= 1, got %s" % (name, n, a.shape))
    return a


def _atoms(zxyz):
    """Nuclear charges and positions as a (natom, 4) float64 array."""
    z = np.ascontiguousarray(np.asarray(zxyz, dtype=np.float64))
    if z.ndim == 1:
        if z.size % 4:
            raise ValueError("zxyz must hold (Z, x, y, z) rows, got %d values" % z.size)
        z = z.reshape(-1, 4)
    if z.ndim != 2 or z.shape[1] != 4 or z.shape[0] < 1:
        raise ValueError("zxyz must have shape (natom, 4), got %s" % (z.shape,))
    return np.ascontiguousarray(z)


def md_flops(la, lb, lc, ld):
    return lib().hfg_md_flops(la, lb, lc, ld)


def calibration_copy(nbytes):
    """Time (ms) of one plain device-to-device copy of nbytes (counter calibration)."""
    ms = C.c_double()
    _check(lib().hfg_calibration_copy(int(nbytes), C.byref(ms)))
    return ms.value


def device_count():
    return lib().hfg_device_count()


class Basis:
    """Cartesian basis as flat arrays, one row per ContractedGaussian."""

    KEYS = ("atom_index", "center", "lxyz", "prim_offset", "exponent", "coefficient")

    def __init__(self, arrays):
        self.a = {
            "atom_index": np.ascontiguousarray(arrays["atom_index"], dtype=np.int32),
            "center": np.ascontiguousarray(np.asarray(arrays["center"], dtype=np.float64).reshape(-1, 3)),
            "lxyz": np.ascontiguousarray(np.asarray(arrays["lxyz"], dtype=np.int32).reshape(-1, 3)),
            "prim_offset": np.ascontiguousarray(arrays["prim_offset"], dtype=np.int32),
            "exponent": np.ascontiguousarray(arrays["exponent"], dtype=np.float64),
            "coefficient": np.ascontiguousarray(arrays["coefficient"], dtype=np.float64),
        }
        self.n = len(self.a["atom_index"])
        self.view = BasisView(self.n, *[_p(self.a[k]) for k in self.KEYS])

    @classmethod
    def from_g94(cls, text, symbols, xyz_bohr):
        xyz = np.ascontiguousarray(np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3))
        h = C.c_void_p()
        st = lib().hfg_basis_from_g94(text.encode(), " ".join(symbols).encode(), _p(xyz), len(symbols), C.byref(h))
        if st != HFG_OK:
            raise HfError(st, "basis construction failed")
        v = BasisView()
        _check(lib().hfg_basis_get_view(h, C.byref(v)))
        n = v.n_functions

        def arr(ptr, count, ct):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy()

        po = arr(v.prim_offset, n + 1, C.c_int32)
        out = dict(atom_index=arr(v.atom_index, n, C.c_int32), center=arr(v.center, 3 * n, C.c_double),
                   lxyz=arr(v.lxyz, 3 * n, C.c_int32), prim_offset=po,
                   exponent=arr(v.exponent, int(po[-1]), C.c_double),
                   coefficient=arr(v.coefficient, int(po[-1]), C.c_double))
        lib().hfg_basis_free(h)
        return cls(out)


class Batch:
    """Many molecules whose J/K builds run in fused launches (hfg_batch_*)."""

    def __init__(self, bases):
        self.bases = list(bases)
        arr = (BasisView * len(self.bases))(*[b.view for b in self.bases])
        self._arr = arr
        self.h = C.c_void_p()
        _check(lib().hfg_batch_create(C.cast(arr, C.c_void_p), len(self.bases), C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) and self.h.value and _LIB is not None:
            _LIB.hfg_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def info(self):
        v = [C.c_int64() for _ in range(3)]
        f = C.c_double()
        _check(lib().hfg_batch_info(self.h, *[C.byref(x) for x in v], C.byref(f)))
        return dict(count=v[0].value, n_max=v[1].value, quartets=v[2].value, flops=f.value)

    def set_density(self, mol, p):
        if not 0 <= mol < len(self.bases):
            raise ValueError("molecule index %d out of range" % mol)
        p = _square(p, self.bases[mol].n, "density")
        _check(lib().hfg_batch_set_density(self.h, mol, _p(p)))

    def jk(self):
        """J/K of every molecule's resident density; results stay in HBM."""
        _check(lib().hfg_batch_jk(self.h))

    def jk_each(self):
        """The same builds molecule after molecule (unbatched launches)."""
        _check(lib().hfg_batch_jk_each(self.h))

    def profile(self, on=True, exclusive=False):
        """Per-class timing of the following builds; exclusive: every class on
        one stream, so each time is the class's alone."""
        _check(lib().hfg_batch_profile(self.h, (2 if exclusive else 1) if on else 0))

    def class_stats(self):
        out = []
        for c in range(21):
            l4 = np.zeros(4, np.int32)
            fl, ms = C.c_double(), C.c_double()
            q, nl = C.c_int64(), C.c_int64()
            _check(lib().hfg_batch_class_stats(self.h, c, _p(l4), C.byref(fl), C.byref(q), C.byref(ms),
                                               C.byref(nl)))
            out.append(dict(cls=c, l=tuple(int(x) for x in l4), flops=fl.value, quartets=q.value, ms=ms.value,
                            launches=nl.value))
        return out

    def get_jk(self, mol):
        n = self.bases[mol].n
        j = np.zeros((n, n))
        k = np.zeros((n, n))
        _check(lib().hfg_batch_get_jk(self.h, mol, _p(j), _p(k)))
        return j, k


def scf_options(max_iter=100, tol=1e-10, residual=None, density=None, conv="adiissoscf", spherical=False,
                verbose=False, incremental=False):
    """hfg_scf_options with the derived tolerances of run_scf_in_space
    (src/scf.rs:384-400)."""
    if residual is None:
        residual = max(math.sqrt(tol) * 1e-2, 1e-11)
    if density is None:
        density = min(max(math.sqrt(tol) * 0.1, 1e-12), 1e-10)
    key = conv.replace("-", "").replace("_", "").lower()
    if key not in CONV:
        raise ValueError("unknown convergence algorithm %r" % conv)
    return ScfOptions(max_iter, tol, residual, density, CONV[key], int(spherical), int(verbose), int(incremental))


class Server:
    """Served / batched RHF (hfg_server_*): requests of geometry trajectories,
    admitted into `capacity` resident slots as they arrive and slots free up,
    every iteration of all slots with one fused Fock-build launch per class."""

    def __init__(self, capacity, hbm_budget_gb=0.0):
        self.h = C.c_void_p()
        self.opt = ServerOptions(int(capacity), float(hbm_budget_gb))
        _check(lib().hfg_server_create(C.byref(self.opt), C.byref(self.h)))
        self.requests = []

    def close(self):
        if getattr(self, "h", None) and self.h.value and _LIB is not None:
            _LIB.hfg_server_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def submit(self, bases, zxyzs, charge=0, arrival=0.0, keep_density=False, gradients=False, **scf_kw):
        """Queues one request: a trajectory of len(bases) geometries of one
        molecule (bases[k] built at zxyzs[k]); returns its id.  gradients:
        the analytic nuclear gradient of every step as well (result()'s
        "gradients", nsteps x natom x 3)."""
        bases = list(bases)
        atoms = [_atoms(z) for z in zxyzs]
        if not bases or len(bases) != len(atoms):
            raise ValueError("one basis per geometry step is required")
        natom = atoms[0].shape[0]
        if any(a.shape[0] != natom for a in atoms):
            raise ValueError("every step must have the same atoms")
        views = (BasisView * len(bases))(*[b.view for b in bases])
        zp = (C.c_void_p * len(atoms))(*[a.ctypes.data for a in atoms])
        opt = scf_options(**scf_kw)
        rid = C.c_int()
        _check(lib().hfg_server_submit(self.h, float(arrival), len(bases), C.cast(views, C.c_void_p),
                                       C.cast(zp, C.c_void_p), natom, int(charge), C.byref(opt),
                                       (HFG_SERVE_KEEP_DENSITY if keep_density else 0) |
                                       (HFG_SERVE_GRADIENT if gradients else 0), C.byref(rid)))
        self.requests.append((rid.value, [b.n for b in bases], bases, atoms, bool(gradients)))
        return rid.value

    def run(self):
        _check(lib().hfg_server_run(self.h))

    def result(self, rid, density=False):
        n_steps = len(self.requests[rid][1])
        r = RequestResult()
        steps = (StepResult * n_steps)()
        n_last = self.requests[rid][1][-1]
        p = np.zeros((n_last, n_last)) if density else None
        st = lib().hfg_server_result(self.h, rid, C.byref(r), C.cast(steps, C.c_void_p),
                                     _p(p) if p is not None else None)
        out = {k: getattr(r, k) for k, _ in RequestResult._fields_}
        out["steps_result"] = [{k: getattr(x, k) for k, _ in StepResult._fields_} for x in steps]
        out["ok"] = st == HFG_OK
        if st != HFG_OK:
            out["error"] = lib().hfg_last_error().decode()
        if density:
            out["density"] = p
        if self.requests[rid][4] and st == HFG_OK:
            g = np.zeros((n_steps, self.requests[rid][3][0].shape[0], 3))
            _check(lib().hfg_server_gradient(self.h, rid, _p(g)))
            out["gradients"] = g
        return out

    def stats(self):
        s = ServerStats()
        _check(lib().hfg_server_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}


class Context:
    """One molecule on one GPU: primitive pairs, Schwarz bounds, quartet lists."""

    def __init__(self, basis):
        self.basis = basis
        self.h = C.c_void_p()
        _check(lib().hfg_ctx_create(C.byref(basis.view), C.byref(self.h)))
        self.n = basis.n

    def close(self):
        if getattr(self, "h", None) and self.h.value and _LIB is not None:
            _LIB.hfg_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def info(self):
        v = [C.c_int64() for _ in range(4)]
        _check(lib().hfg_ctx_info(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("n_functions", "n_shells", "n_pairs", "n_quartets"), (x.value for x in v)))

    def schwarz(self):
        q = np.zeros(self.info()["n_pairs"])
        _check(lib().hfg_schwarz(self.h, _p(q)))
        return q

    def shell_quartet(self, a, b, c, d, dims):
        out = np.zeros(int(np.prod(dims)))
        _check(lib().hfg_shell_quartet(self.h, a, b, c, d, _p(out)))
        return out.reshape(dims)

    def eri_packed(self):
        np_ = self.n * (self.n + 1) // 2
        out = np.zeros(np_ * (np_ + 1) // 2)
        _check(lib().hfg_eri_packed(self.h, _p(out), out.size))
        return out

    def eri_mo(self, c):
        """(pq|rs) in the MO basis of the columns of c (n x n_mo), packed
        [rs][pq] with compound indices p >= q (transform_eri_ao_to_mo,
        src/mo_integrals.rs:156-265)."""
        c = _columns(c, self.n, "MO coefficients")
        nmo = c.shape[1]
        mp = nmo * (nmo + 1) // 2
        out = np.zeros(mp * mp)
        _check(lib().hfg_eri_mo(self.h, _p(c), nmo, _p(out), out.size))
        return out

    def contract_jk(self, p):
        p = _square(p, self.n, "density")
        j = np.zeros((self.n, self.n))
        k = np.zeros((self.n, self.n))
        _check(lib().hfg_contract_jk(self.h, _p(p), _p(j), _p(k)))
        return j, k

    def graph_launches(self):
        """Fock builds that replayed the captured HIP graph of their launches."""
        v = C.c_int64()
        _check(lib().hfg_graph_launches(self.h, C.byref(v)))
        return v.value

    def one_electron(self, zxyz):
        """S, T, V (n x n each) of the context's basis for the nuclei zxyz,
        computed on the device: compute_one_electron_integrals
        (src/scf.rs:246-271)."""
        zxyz = _atoms(zxyz)
        s, t, v = (np.zeros((self.n, self.n)) for _ in range(3))
        _check(lib().hfg_one_electron(self.h, _p(zxyz), len(zxyz), _p(s), _p(t), _p(v)))
        return s, t, v

    def sad_density(self, zxyz, charge=0):
        """The SAD initial density (compute_sad_density, src/sad.rs:217-259),
        superposed and scaled on the device."""
        zxyz = _atoms(zxyz)
        p = np.zeros((self.n, self.n))
        _check(lib().hfg_sad_density(self.h, _p(zxyz), len(zxyz), int(charge), _p(p)))
        return p

    def rhf_gradient(self, zxyz, p, w):
        """Nuclear gradient (natom x 3, Eh/bohr) for density p and energy-weighted
        density w: compute_gradient_from_density (src/gradient.rs:287-512)."""
        zxyz = _atoms(zxyz)
        p = _square(p, self.n, "density")
        w = _square(w, self.n, "energy-weighted density")
        g = np.zeros((len(zxyz), 3))
        _check(lib().hfg_rhf_gradient(self.h, _p(zxyz), len(zxyz), _p(p), _p(w), _p(g)))
        return g

    def contract_jk_multi(self, total, spins):
        """J of the total density and K of each spin density (contract_jk(total,
        spins), src/scf.rs:337-364).  A spin density that is the total object
        itself (RHF) shares its build."""
        tot = _square(total, self.n, "total density")
        arrs = [tot if s is total else _square(s, self.n, "spin density") for s in spins]
        j = np.zeros((self.n, self.n))
        ks = [np.zeros((self.n, self.n)) for _ in arrs]
        pin = (C.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
        kout = (C.c_void_p * max(1, len(ks)))(*[k.ctypes.data for k in ks])
        _check(lib().hfg_contract_jk_multi(self.h, _p(tot), len(arrs), C.cast(pin, C.c_void_p), _p(j),
                                           C.cast(kout, C.c_void_p)))
        return j, ks

    def build_fock(self, h, p):
        h = _square(h, self.n, "core Hamiltonian")
        p = _square(p, self.n, "density")
        f = np.zeros((self.n, self.n))
        _check(lib().hfg_build_fock(self.h, _p(h), _p(p), _p(f)))
        return f

    def set_density(self, p):
        p = _square(p, self.n, "density")
        _check(lib().hfg_set_density(self.h, _p(p)))

    def jk_resident(self):
        """J/K of the resident density; results stay in HBM."""
        _check(lib().hfg_jk_resident(self.h))

    def get_jk(self):
        j = np.zeros((self.n, self.n))
        k = np.zeros((self.n, self.n))
        _check(lib().hfg_get_jk(self.h, _p(j), _p(k)))
        return j, k

    def profile(self, on=True):
        _check(lib().hfg_profile(self.h, int(on)))

    def class_stats(self):
        out = []
        for c in range(21):
            l4 = np.zeros(4, np.int32)
            fl, ms = C.c_double(), C.c_double()
            q, nl = C.c_int64(), C.c_int64()
            _check(lib().hfg_class_stats(self.h, c, _p(l4), C.byref(fl), C.byref(q), C.byref(ms), C.byref(nl)))
            out.append(dict(cls=c, l=tuple(int(x) for x in l4), flops=fl.value, quartets=q.value, ms=ms.value,
                            launches=nl.value))
        return out

    def fock_flops(self):
        f, pq = C.c_double(), C.c_double()
        _check(lib().hfg_fock_flops(self.h, C.byref(f), C.byref(pq)))
        return f.value, pq.value

    def contract_jk_device(self, d_p, d_j, d_k):
        """J/K on HBM-resident buffers (device pointers as ints, e.g. torch data_ptr())."""
        _check(lib().hfg_contract_jk_device(self.h, C.c_void_p(d_p), C.c_void_p(d_j), C.c_void_p(d_k)))


class DfTensor:
    """Mirror of DfTensor (src/df.rs:80-174), resident in HBM.  aux None: the
    factors come from cholesky()."""

    def __init__(self, basis, aux, metric_cutoff=1e-10, _handle=None):
        self.basis, self.aux = basis, aux
        self.h = C.c_void_p()
        if _handle is not None:
            self.h = _handle
        else:
            _check(lib().hfg_df_build(C.byref(basis.view), C.byref(aux.view), metric_cutoff, C.byref(self.h)))
        v = [C.c_int64() for _ in range(3)]
        _check(lib().hfg_df_info(self.h, *[C.byref(x) for x in v]))
        self.n, self.n_aux, self.rank = (x.value for x in v)

    def close(self):
        if getattr(self, "h", None) and self.h.value and _LIB is not None:
            _LIB.hfg_df_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def coulomb_exchange_rhf(self, cocc):
        cocc = _columns(cocc, self.n, "occupied coefficients")
        j = np.zeros((self.n, self.n))
        k = np.zeros((self.n, self.n))
        _check(lib().hfg_df_jk(self.h, _p(cocc), cocc.shape[1], _p(j), _p(k)))
        return j, k

    PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2}

    def jk_precision(self, cocc, precision="fp64"):
        """J (FP64) and K with the exchange contraction in the given precision;
        also returns the device milliseconds of the contraction."""
        if precision not in self.PRECISION:
            raise ValueError("precision must be one of %s" % sorted(self.PRECISION))
        cocc = _columns(cocc, self.n, "occupied coefficients")
        j = np.zeros((self.n, self.n))
        k = np.zeros((self.n, self.n))
        ms = C.c_double()
        _check(lib().hfg_df_jk_ex(self.h, _p(cocc), cocc.shape[1], _p(j), _p(k), self.PRECISION[precision],
                                  C.byref(ms)))
        return j, k, ms.value


DF_PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2}


CD_HF_FACTOR_BUDGET = 512 * 1024 * 1024  # bytes (src/hf_variants/cholesky_df.rs:22)


def rhf_gradient(basis, zxyz, density, fock):
    """compute_rhf_gradient (src/gradient.rs:272-285) from a converged RHF
    density and Fock matrix: W = 2 C_occ eps C_occ^T = 1/2 P F P."""
    p = np.asarray(density, dtype=np.float64)
    w = 0.5 * p @ np.asarray(fock, dtype=np.float64) @ p
    ctx = Context(basis)
    try:
        return ctx.rhf_gradient(zxyz, p, w)
    finally:
        ctx.close()


def cholesky(basis, threshold, budget=CD_HF_FACTOR_BUDGET):
    """pivoted_cholesky_eri (src/cholesky.rs:44-192): (DfTensor of the
    Cholesky vectors, info dict with max_residual and pivots)."""
    h = C.c_void_p()
    info = CdInfo()
    np_ = basis.n * (basis.n + 1) // 2
    piv = np.zeros(np_, np.int32)
    _check(lib().hfg_cd_build(C.byref(basis.view), threshold, float(budget), C.byref(h), C.byref(info), _p(piv), np_))
    t = DfTensor(basis, None, _handle=h)
    return t, {"rank": info.rank, "max_residual": info.max_residual, "pivots": [int(x) for x in piv[:info.rank]]}


def cd_rhf(basis, zxyz, delta, charge=0, max_iter=300, tol=1e-12, budget=CD_HF_FACTOR_BUDGET,
           k_precision="fp64", switch_residual=1e-4):
    """CD-HF (run_cd_hf, src/hf_variants/cholesky_df.rs:80-95): the Cholesky
    factors to delta in the factorized RHF driver, with energy tolerance tol,
    residual tolerance max(sqrt(tol) 1e-2, 1e-11) and density tolerance
    sqrt(tol) 0.1 clamped to [1e-12, 1e-10]."""
    if k_precision not in DF_PRECISION:
        raise ValueError("k_precision must be one of %s" % sorted(DF_PRECISION))
    r_tol = max(math.sqrt(tol) * 1e-2, 1e-11)
    d_tol = min(max(math.sqrt(tol) * 0.1, 1e-12), 1e-10)
    opt = DfhfOptions(max_iter, tol, r_tol, d_tol, 1e-10, DF_PRECISION[k_precision], switch_residual)
    res = DfhfResult()
    info = CdInfo()
    zxyz = _atoms(zxyz)
    n = basis.n
    p, f = np.zeros((n, n)), np.zeros((n, n))
    _check(lib().hfg_cd_rhf(C.byref(basis.view), _p(zxyz), len(zxyz), charge, delta, float(budget), C.byref(opt),
                            C.byref(res), C.byref(info), _p(p), _p(f)))
    out = {k_: getattr(res, k_) for k_, _ in DfhfResult._fields_}
    out.update(density=p, fock=f, n_cholesky_vectors=info.rank, cholesky_max_error=info.max_residual)
    return out


def df_rhf(basis, aux, zxyz, charge=0, max_iter=100, e_tol=1e-12, r_tol=1e-10, d_tol=1e-10, metric_cutoff=1e-10,
           k_precision="fp64", switch_residual=1e-4):
    """DF-RHF (run_df_hf_with_aux, src/hf_variants/ri_df_hf.rs:122-263); defaults of DFHFOptions.

    k_precision "fp32" or "bf16x3" runs the exchange contraction of the early
    iterations in that precision until the projected residual is below
    switch_residual, then finishes with FP64 builds (no reference
    counterpart; "fp64" is the reference's scheme)."""
    if k_precision not in DF_PRECISION:
        raise ValueError("k_precision must be one of %s" % sorted(DF_PRECISION))
    opt = DfhfOptions(max_iter, e_tol, r_tol, d_tol, metric_cutoff, DF_PRECISION[k_precision], switch_residual)
    res = DfhfResult()
    zxyz = _atoms(zxyz)
    n = basis.n
    p, f, j, k = (np.zeros((n, n)) for _ in range(4))
    _check(lib().hfg_df_rhf(C.byref(basis.view), C.byref(aux.view), _p(zxyz), len(zxyz), charge, C.byref(opt),
                            C.byref(res), _p(p), _p(f), _p(j), _p(k)))
    out = {k_: getattr(res, k_) for k_, _ in DfhfResult._fields_}
    out.update(density=p, fock=f, j=j, k=k)
    return out


CONV = {"adiissoscf": 0, "default": 0, "cdiis": 1, "diis": 1, "pulay": 1, "adiis": 2, "adiiscdiis": 3}


def scf_rhf(basis, zxyz, charge=0, max_iter=100, tol=1e-10, residual=None, density=None, conv="adiissoscf",
            spherical=False, verbose=False, want_matrices=False, incremental=False, guess=None):
    """RHF with the derived tolerances of run_scf_in_space (src/scf.rs:384-400).
    incremental: J/K of the density change inside the iterations (full first
    build and final rebuild).  guess: initial density instead of SAD
    (run_scf_with_initial_density, src/scf.rs:630-639)."""
    opt = scf_options(max_iter, tol, residual, density, conv, spherical, verbose, incremental)
    res = ScfResult()
    zxyz = _atoms(zxyz)
    n = basis.n
    g = None if guess is None else _square(guess, n, "initial density")
    p = np.zeros((n, n)) if want_matrices else None
    f = np.zeros((n, n)) if want_matrices else None
    eps = np.zeros(n) if want_matrices else None
    outs = (_p(p) if p is not None else None, _p(f) if f is not None else None, _p(eps) if eps is not None else None)
    if g is None:
        st = lib().hfg_scf_rhf(C.byref(basis.view), _p(zxyz), len(zxyz), charge, C.byref(opt), C.byref(res), *outs)
    else:
        st = lib().hfg_scf_rhf_guess(C.byref(basis.view), _p(zxyz), len(zxyz), charge, C.byref(opt), _p(g),
                                     C.byref(res), *outs)
    if st not in (HFG_OK, HFG_NOT_CONVERGED):
        _check(st)
    out = {k: getattr(res, k) for k, _ in ScfResult._fields_}
    out["converged"] = bool(res.converged)
    if want_matrices:
        out.update(density=p, fock=f, orbital_energies=eps)
    return out


class Molecule:
    """Small mirror of hf_rs::Molecule (src/api.rs:437-638) for the RHF path."""

    def __init__(self, symbols, positions_bohr, basis="sto-3g", charge=0, spherical=False):
        from . import molecules as M
        self.mol = M.Molecule(symbols, positions_bohr, charge)
        self.basis = Basis.from_g94(M.basis_text(basis), symbols, positions_bohr)
        self.charge = charge
        self.spherical = spherical

    def scf(self, **kw):
        kw.setdefault("spherical", self.spherical)
        return scf_rhf(self.basis, self.mol.zxyz, charge=self.charge, **kw)


def run(method, symbols, positions_bohr, basis="sto-3g", charge=0, spherical=False, **kw):
    """Mirror of api::run (src/api.rs:1729-1738) for method "hf"/"rhf"."""
    if method.lower() not in ("hf", "rhf", "scf"):
        raise HfError(HFG_INVALID_INPUT, "only the RHF path is provided by this engine")
    r = Molecule(symbols, positions_bohr, basis, charge, spherical).scf(**kw)
    if not r["converged"]:
        raise HfError(HFG_NOT_CONVERGED, "SCF did not converge")
    return r

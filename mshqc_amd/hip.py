"""ctypes binding of libmshqc_hip.so for callers holding torch device tensors.

PyTorch-ROCm is used here only for device memory and streams: every call
passes ``tensor.data_ptr()`` through the C ABI declared in
``include/mshqc_hip.h`` and queues its kernels on torch's current stream.
There is no CPU fallback: if the library is missing, importing this module
raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmshqc_hip.so")

PREC_FP64 = 0

_c_int, _c_double, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)

# name -> argtypes (all return int status unless noted)
_SIGS = {
    "mshqc_hip_ctx_create": [_c_int, ctypes.POINTER(_vp)],
    "mshqc_hip_ctx_destroy": [_vp],
    "mshqc_hip_set_stream": [_vp, _vp],
    "mshqc_hip_last_ms": [_vp, ctypes.POINTER(ctypes.c_float)],
    "mshqc_hip_set_zblock": [_vp, ctypes.c_size_t],
    "mshqc_hip_profile": [_vp, _c_int],
    "mshqc_hip_profile_get": [_vp, _c_int, ctypes.POINTER(ctypes.c_long),
                              ctypes.POINTER(ctypes.c_double)],
    "mshqc_hip_alloc": [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)],
    "mshqc_hip_free": [_vp, _vp],
    "mshqc_hip_memcpy_h2d": [_vp, _vp, _vp, ctypes.c_size_t],
    "mshqc_hip_memcpy_d2h": [_vp, _vp, _vp, ctypes.c_size_t],
    "mshqc_hip_jk": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _c_int, _vp, _vp],
    "mshqc_hip_jk_factor": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _vp],
    "mshqc_hip_jk_rohf": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp],
    "mshqc_hip_half_transform": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp],
    "mshqc_hip_ump3": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                       _vp, _vp, _D],
    "mshqc_hip_rmp3": [_vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _D, _vp],
    "mshqc_hip_cholesky": [_vp, _vp, _c_int, ctypes.c_double, _c_int, _vp, ctypes.POINTER(_c_int),
                           ctypes.POINTER(_c_int), _D],
    "mshqc_hip_omp2_opdm": [_vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                            _D, _vp, _vp, _vp, _vp],
    "mshqc_hip_omp2_grad2p": [_vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp],
    "mshqc_hip_omp2_fock": [_vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _c_int] + [_vp] * 8,
    "mshqc_hip_omp2_grad_norm": [_vp, _c_int, _c_int, _vp, _vp, _D],
    "mshqc_hip_omp2_rotate": [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp],
    "mshqc_hip_transform": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp],
    "mshqc_hip_rmp2_energy": [_vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _D],
    "mshqc_hip_rmp2_grad": [_vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _D,
                            _vp, _vp, _vp],
    "mshqc_hip_rmp2_t2": [_vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp],
    "mshqc_hip_rmp2_energy_noalloc": [_vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp],
    "mshqc_hip_ump2_energy": [_vp, _vp, _vp] + [_c_int] * 5 + [_vp] * 4 + [_c_int, _D],
    "mshqc_hip_ump2_grad": [_vp, _vp, _vp] + [_c_int] * 5 + [_vp] * 4 + [_c_int, _D]
    + [_vp] * 6,
    "mshqc_hip_ump2_t2": [_vp, _vp, _vp] + [_c_int] * 5 + [_vp] * 7,
}

#: Default cap (doubles) of the per-block amplitude workspaces of a context.
ZBLOCK_DEFAULT = 1 << 30

#: Every symbol include/mshqc_hip.h declares.
EXPORTED = sorted(list(_SIGS) + ["mshqc_hip_last_error", "mshqc_hip_version"])


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libmshqc_hip.so and attach the C signatures (no GPU needed)."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m mshqc_amd.build` "
            "(the HIP path has no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _c_int
    lib.mshqc_hip_last_error.argtypes = [_vp]
    lib.mshqc_hip_last_error.restype = ctypes.c_char_p
    lib.mshqc_hip_version.argtypes = []
    lib.mshqc_hip_version.restype = ctypes.c_char_p
    return lib


_LIB = None
_LOCK = threading.Lock()


def lib() -> ctypes.CDLL:
    """The process-wide library: MSHQC_HIP_LIBRARY names another build of it
    (same ABI), else the in-tree libmshqc_hip.so."""
    global _LIB
    with _LOCK:
        if _LIB is None:
            _LIB = load_library(os.environ.get("MSHQC_HIP_LIBRARY", LIB_PATH))
    return _LIB


class Context:
    """One device context (stream-following, workspace cache) of the C ABI."""

    def __init__(self, device: int = 0):
        import torch

        self._torch = torch
        self.device = device
        h = _vp()
        rc = lib().mshqc_hip_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"mshqc_hip_ctx_create(device={device}) failed with code {rc}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mshqc_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- plumbing -----------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            msg = lib().mshqc_hip_last_error(self.h).decode()
            raise RuntimeError(f"{what} failed (code {rc}): {msg}")

    def _follow_stream(self):
        s = self._torch.cuda.current_stream(self.device)
        lib().mshqc_hip_set_stream(self.h, _vp(s.cuda_stream))

    def _ptr(self, t, shape=None, name="tensor"):
        torch = self._torch
        if t is None:
            return None
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError(f"{name}: need a contiguous float64 CUDA tensor")
        if shape is not None and tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name}: shape {tuple(t.shape)} != expected {tuple(shape)}")
        return _vp(t.data_ptr())

    def set_zblock(self, elems: int):
        """Cap (doubles) of the per-block amplitude workspaces; 0 restores
        the default (ZBLOCK_DEFAULT)."""
        self._check(lib().mshqc_hip_set_zblock(self.h, int(elems)), "mshqc_hip_set_zblock")

    def profile(self, enable: bool = True):
        """Clear and enable (or disable) the per-kernel-class event timers."""
        self._check(lib().mshqc_hip_profile(self.h, int(bool(enable))), "mshqc_hip_profile")

    def profile_get(self, kind: int):
        """(launches, total ms) of kernel class `kind` (0 pair, 1 GEMM, 2 J)."""
        n, ms = ctypes.c_long(), ctypes.c_double()
        self._check(lib().mshqc_hip_profile_get(self.h, kind, ctypes.byref(n), ctypes.byref(ms)),
                    "mshqc_hip_profile_get")
        return int(n.value), float(ms.value)

    def last_ms(self) -> float:
        ms = ctypes.c_float()
        lib().mshqc_hip_last_ms(self.h, ctypes.byref(ms))
        return float(ms.value)

    # -- Fock pieces --------------------------------------------------------
    def jk(self, L, P, screen=False):
        """J [n,n] and K (shape of P) from L [Q,n,n] and P [n,n] / [nspin,n,n]."""
        torch = self._torch
        if L.dim() != 3 or L.shape[1] != L.shape[2]:
            raise ValueError(f"L: need [Q, n, n], got {tuple(L.shape)}")
        Q, n, _ = L.shape
        if P.dim() == 2:
            nspin = 1
            pshape = (n, n)
        elif P.dim() == 3 and P.shape[0] in (1, 2):
            nspin = P.shape[0]
            pshape = (nspin, n, n)
        else:
            raise ValueError(f"P: need [n, n] or [nspin, n, n] with nspin 1 or 2, got {tuple(P.shape)}")
        J = torch.empty((n, n), dtype=torch.float64, device=L.device)
        K = torch.empty_like(P)
        self._follow_stream()
        rc = lib().mshqc_hip_jk(self.h, self._ptr(L, name="L"), Q, n, self._ptr(P, pshape, "P"),
                                nspin, int(bool(screen)), self._ptr(J), self._ptr(K))
        self._check(rc, "mshqc_hip_jk")
        return J, K

    def jk_factor(self, L, Y):
        torch = self._torch
        Q, n, _ = L.shape
        m = Y.shape[1]
        J = torch.empty((n, n), dtype=torch.float64, device=L.device)
        K = torch.empty((n, n), dtype=torch.float64, device=L.device)
        self._follow_stream()
        rc = lib().mshqc_hip_jk_factor(self.h, self._ptr(L, name="L"), Q, n,
                                       self._ptr(Y, (n, m), "Y"), m, self._ptr(J), self._ptr(K))
        self._check(rc, "mshqc_hip_jk_factor")
        return J, K

    def cholesky(self, eri, threshold, max_vectors=None):
        """Pivoted Cholesky of a device ERI matrix [n*n, n*n] (or [n,n,n,n]).
        Returns (L [Q, n*n] device tensor, converged, max_residual)."""
        torch = self._torch
        np2 = eri.shape[0] if eri.dim() == 2 else eri.shape[0] * eri.shape[1]
        n = int(round(np2 ** 0.5))
        if n * n != np2 or eri.numel() != np2 * np2:
            raise ValueError("eri: need [n*n, n*n] or [n, n, n, n]")
        cap = np2 if max_vectors is None else int(max_vectors)
        L = torch.empty((cap, np2), dtype=torch.float64, device=eri.device)
        nv, conv, res = _c_int(), _c_int(), ctypes.c_double()
        self._follow_stream()
        rc = lib().mshqc_hip_cholesky(self.h, self._ptr(eri, name="eri"), n, float(threshold), cap,
                                      self._ptr(L), ctypes.byref(nv), ctypes.byref(conv), ctypes.byref(res))
        self._check(rc, "mshqc_hip_cholesky")
        return L[: nv.value], bool(conv.value), res.value

    def jk_rohf(self, L, C, na, nb):
        """J[Pa + Pb], Ka, Kb for alpha = C[:, :na], beta = C[:, :nb]."""
        torch = self._torch
        Q, n, _ = L.shape
        ldc = C.shape[1]
        out = [torch.empty((n, n), dtype=torch.float64, device=L.device) for _ in range(3)]
        self._follow_stream()
        rc = lib().mshqc_hip_jk_rohf(self.h, self._ptr(L, name="L"), Q, n, self._ptr(C, (n, ldc), "C"),
                                     ldc, int(na), int(nb), *[self._ptr(t) for t in out])
        self._check(rc, "mshqc_hip_jk_rohf")
        return tuple(out)

    def half_transform(self, L, C):
        torch = self._torch
        Q, n, _ = L.shape
        m = C.shape[1]
        X = torch.empty((Q, n, m), dtype=torch.float64, device=L.device)
        self._follow_stream()
        rc = lib().mshqc_hip_half_transform(self.h, self._ptr(L, name="L"), Q, n,
                                            self._ptr(C, (n, m), "C"), m, self._ptr(X))
        self._check(rc, "mshqc_hip_half_transform")
        return X

    def transform(self, L, Cocc, Cvir, precision=PREC_FP64):
        """B [Q, o, v] = C_occ^T L^Q C_vir; precision k in [2, 4] runs it on split-BF16 GEMMs."""
        torch = self._torch
        Q, n, _ = L.shape
        o, v = Cocc.shape[1], Cvir.shape[1]
        B = torch.empty((Q, o, v), dtype=torch.float64, device=L.device)
        self._follow_stream()
        rc = lib().mshqc_hip_transform(self.h, self._ptr(L, name="L"), Q, n,
                                       self._ptr(Cocc, (n, o), "Cocc"), o,
                                       self._ptr(Cvir, (n, v), "Cvir"), v, int(precision), self._ptr(B))
        self._check(rc, "mshqc_hip_transform")
        return B

    # -- MP2 ----------------------------------------------------------------
    def rmp2_energy(self, B, eps_o, eps_v, precision=PREC_FP64) -> float:
        Q, o, v = B.shape
        E = ctypes.c_double()
        self._follow_stream()
        rc = lib().mshqc_hip_rmp2_energy(self.h, self._ptr(B, name="B"), Q, o, v,
                                         self._ptr(eps_o, (o,), "eps_o"),
                                         self._ptr(eps_v, (v,), "eps_v"), int(precision),
                                         ctypes.byref(E))
        self._check(rc, "mshqc_hip_rmp2_energy")
        return E.value

    def rmp2_energy_noalloc(self, B, eps_o, eps_v, E):
        """FP64 CD-RMP2 energy into the caller's device tensor E (one float64),
        allocating nothing; queued on torch's current stream, not synchronised."""
        Q, o, v = B.shape
        self._follow_stream()
        rc = lib().mshqc_hip_rmp2_energy_noalloc(self.h, self._ptr(B, name="B"), Q, o, v,
                                                 self._ptr(eps_o, (o,), "eps_o"),
                                                 self._ptr(eps_v, (v,), "eps_v"), self._ptr(E, (1,), "E"))
        self._check(rc, "mshqc_hip_rmp2_energy_noalloc")
        return E

    def rmp2_grad(self, B, eps_o, eps_v, precision=PREC_FP64, want=(True, True, True)):
        """Returns (E, dE/dB, dE/deps_o, dE/deps_v); entries not wanted are None."""
        torch = self._torch
        Q, o, v = B.shape
        dB = torch.empty_like(B) if want[0] else None
        deo = torch.empty((o,), dtype=torch.float64, device=B.device) if want[1] else None
        dev = torch.empty((v,), dtype=torch.float64, device=B.device) if want[2] else None
        E = ctypes.c_double()
        self._follow_stream()
        rc = lib().mshqc_hip_rmp2_grad(self.h, self._ptr(B, name="B"), Q, o, v,
                                       self._ptr(eps_o, (o,), "eps_o"),
                                       self._ptr(eps_v, (v,), "eps_v"), int(precision),
                                       ctypes.byref(E), self._ptr(dB), self._ptr(deo),
                                       self._ptr(dev))
        self._check(rc, "mshqc_hip_rmp2_grad")
        return E.value, dB, deo, dev

    def rmp2_t2(self, B, eps_o, eps_v):
        torch = self._torch
        Q, o, v = B.shape
        t2 = torch.empty((o, o, v, v), dtype=torch.float64, device=B.device)
        self._follow_stream()
        rc = lib().mshqc_hip_rmp2_t2(self.h, self._ptr(B, name="B"), Q, o, v,
                                     self._ptr(eps_o, (o,), "eps_o"),
                                     self._ptr(eps_v, (v,), "eps_v"), self._ptr(t2))
        self._check(rc, "mshqc_hip_rmp2_t2")
        return t2

    def ump2_energy(self, Ba, Bb, eoa, eva, eob, evb, precision=PREC_FP64):
        Q, oa, va = Ba.shape
        _, ob, vb = Bb.shape
        E = (ctypes.c_double * 3)()
        self._follow_stream()
        rc = lib().mshqc_hip_ump2_energy(
            self.h, self._ptr(Ba, name="Ba"), self._ptr(Bb, (Q, ob, vb), "Bb"), Q, oa, va, ob, vb,
            self._ptr(eoa, (oa,)), self._ptr(eva, (va,)), self._ptr(eob, (ob,)),
            self._ptr(evb, (vb,)), int(precision), E)
        self._check(rc, "mshqc_hip_ump2_energy")
        return tuple(E)

    def ump2_grad(self, Ba, Bb, eoa, eva, eob, evb, precision=PREC_FP64):
        torch = self._torch
        Q, oa, va = Ba.shape
        _, ob, vb = Bb.shape
        dBa, dBb = torch.empty_like(Ba), torch.empty_like(Bb)
        d = [torch.empty((k,), dtype=torch.float64, device=Ba.device) for k in (oa, va, ob, vb)]
        E = (ctypes.c_double * 3)()
        self._follow_stream()
        rc = lib().mshqc_hip_ump2_grad(
            self.h, self._ptr(Ba, name="Ba"), self._ptr(Bb, (Q, ob, vb), "Bb"), Q, oa, va, ob, vb,
            self._ptr(eoa, (oa,)), self._ptr(eva, (va,)), self._ptr(eob, (ob,)),
            self._ptr(evb, (vb,)), int(precision), E, self._ptr(dBa), self._ptr(dBb),
            *[self._ptr(x) for x in d])
        self._check(rc, "mshqc_hip_ump2_grad")
        return tuple(E), (dBa, dBb, *d)

    def ump2_t2(self, Ba, Bb, eoa, eva, eob, evb):
        """(t2_aa [oa,oa,va,va], t2_bb [ob,ob,vb,vb], t2_ab [oa,ob,va,vb]) = g / D."""
        torch = self._torch
        Q, oa, va = Ba.shape
        _, ob, vb = Bb.shape
        kw = dict(dtype=torch.float64, device=Ba.device)
        t = [torch.zeros(s, **kw) for s in ((oa, oa, va, va), (ob, ob, vb, vb), (oa, ob, va, vb))]
        self._follow_stream()
        rc = lib().mshqc_hip_ump2_t2(
            self.h, self._ptr(Ba, name="Ba"), self._ptr(Bb, (Q, ob, vb), "Bb"), Q, oa, va, ob, vb,
            self._ptr(eoa, (oa,)), self._ptr(eva, (va,)), self._ptr(eob, (ob,)), self._ptr(evb, (vb,)),
            *[self._ptr(x) for x in t])
        self._check(rc, "mshqc_hip_ump2_t2")
        return tuple(t)

    def omp2_opdm(self, Ba, Bb, eoa, eva, eob, evb):
        """(E_aa, E_bb, E_ab), (Goo_a, Gvv_a, Goo_b, Gvv_b) of one OMP2 iteration."""
        torch = self._torch
        Q, oa, va = Ba.shape
        _, ob, vb = Bb.shape
        G = [torch.empty((k, k), dtype=torch.float64, device=Ba.device) for k in (oa, va, ob, vb)]
        E = (ctypes.c_double * 3)()
        self._follow_stream()
        rc = lib().mshqc_hip_omp2_opdm(
            self.h, self._ptr(Ba, name="Ba"), self._ptr(Bb, (Q, ob, vb), "Bb"), Q, oa, va, ob, vb,
            self._ptr(eoa, (oa,)), self._ptr(eva, (va,)), self._ptr(eob, (ob,)), self._ptr(evb, (vb,)),
            E, *[self._ptr(g) for g in G])
        self._check(rc, "mshqc_hip_omp2_opdm")
        return tuple(E), tuple(G)

    def omp2_grad2p(self, G, Loo, Lvv):
        """Two-particle term w2 [v, o] of the OMP2 orbital gradient of one spin,
        from G = dE/dB [Q, o, v] and the same orbitals' Loo [Q, o, o], Lvv [Q, v, v]."""
        torch = self._torch
        Q, o, v = G.shape
        w2 = torch.zeros((v, o), dtype=torch.float64, device=G.device)
        self._follow_stream()
        rc = lib().mshqc_hip_omp2_grad2p(self.h, self._ptr(G, name="G"), self._ptr(Loo, (Q, o, o), "Loo"),
                                         self._ptr(Lvv, (Q, v, v), "Lvv"), Q, o, v, self._ptr(w2))
        self._check(rc, "mshqc_hip_omp2_grad2p")
        return w2

    def omp2_fock(self, L, H, Co_a, Cv_a, Co_b, Cv_b, P_a, P_b, Goo_a, Gvv_a, Goo_b, Gvv_b):
        """MO-basis generalised Fock (F_a, F_b) [n, n] of one OMP2 iteration."""
        torch = self._torch
        Q, n, _ = L.shape
        na, nb = Co_a.shape[1], Co_b.shape[1]
        F = [torch.empty((n, n), dtype=torch.float64, device=L.device) for _ in range(2)]
        self._follow_stream()
        p = self._ptr
        rc = lib().mshqc_hip_omp2_fock(self.h, p(L, name="L"), Q, n, p(H, (n, n), "H"), p(Co_a), p(Cv_a), na,
                                       p(Co_b), p(Cv_b), nb, p(P_a, (n, n)), p(P_b, (n, n)), p(Goo_a), p(Gvv_a),
                                       p(Goo_b), p(Gvv_b), p(F[0]), p(F[1]))
        self._check(rc, "mshqc_hip_omp2_fock")
        return tuple(F)

    def omp2_grad_norm(self, F, o, w2=None) -> float:
        n = F.shape[0]
        g2 = ctypes.c_double()
        self._follow_stream()
        rc = lib().mshqc_hip_omp2_grad_norm(self.h, n, int(o), self._ptr(F, (n, n), "F"),
                                            self._ptr(w2) if w2 is not None else None, ctypes.byref(g2))
        self._check(rc, "mshqc_hip_omp2_grad_norm")
        return g2.value

    def omp2_rotate(self, F, Co, Cv, P, w2=None):
        """One OMP2 orbital step in place on Co [n, o], Cv [n, v], P [n, n]."""
        n = F.shape[0]
        o = Co.shape[1]
        self._follow_stream()
        rc = lib().mshqc_hip_omp2_rotate(self.h, n, o, self._ptr(F, (n, n), "F"),
                                         self._ptr(w2) if w2 is not None else None, self._ptr(Co, (n, o), "Co"),
                                         self._ptr(Cv, (n, n - o), "Cv"), self._ptr(P, (n, n), "P"))
        self._check(rc, "mshqc_hip_omp2_rotate")

    def rmp3(self, B, Loo, Lvv, t1, eps_o, eps_v, want_t2=False):
        """(E3, t2_2 or None) of CD-RMP3 from B [Q,o,v], Loo [Q,o,o], Lvv [Q,v,v], t1 [o,o,v,v]."""
        torch = self._torch
        Q, o, v = B.shape
        t22 = torch.empty((o, o, v, v), dtype=torch.float64, device=B.device) if want_t2 else None
        E = ctypes.c_double()
        self._follow_stream()
        rc = lib().mshqc_hip_rmp3(self.h, self._ptr(B, name="B"), self._ptr(Loo, (Q, o, o), "Loo"),
                                  self._ptr(Lvv, (Q, v, v), "Lvv"), Q, o, v, self._ptr(t1, (o, o, v, v), "t1"),
                                  self._ptr(eps_o, (o,)), self._ptr(eps_v, (v,)), ctypes.byref(E),
                                  self._ptr(t22) if want_t2 else None)
        self._check(rc, "mshqc_hip_rmp3")
        return E.value, t22

    def ump3(self, Ba, Bb, Looa, Loob, Lvva, Lvvb, eoa, eva, eob, evb):
        """(E_ss_aa, E_ss_bb, E_os) of CD-UMP3 from per-spin MO blocks."""
        Q, oa, va = Ba.shape
        _, ob, vb = Bb.shape
        E = (ctypes.c_double * 3)()
        self._follow_stream()
        rc = lib().mshqc_hip_ump3(
            self.h, self._ptr(Ba, name="Ba"), self._ptr(Bb, (Q, ob, vb), "Bb"), self._ptr(Looa, (Q, oa, oa), "Looa"),
            self._ptr(Loob, (Q, ob, ob), "Loob"), self._ptr(Lvva, (Q, va, va), "Lvva"),
            self._ptr(Lvvb, (Q, vb, vb), "Lvvb"), Q, oa, va, ob, vb, self._ptr(eoa, (oa,)), self._ptr(eva, (va,)),
            self._ptr(eob, (ob,)), self._ptr(evb, (vb,)), E)
        self._check(rc, "mshqc_hip_ump3")
        return tuple(E)


_CTX = {}


def context(device: int = 0) -> Context:
    """Process-wide context per device."""
    if device not in _CTX:
        _CTX[device] = Context(device)
    return _CTX[device]

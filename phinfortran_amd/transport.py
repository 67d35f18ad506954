"""Host-side mirror of phinfortran's per-(omega,k) transport interface on the
MI355X engine.

The functions keep the reference's names, argument meaning and error
behaviour (a nonzero status raises, where the reference would `stop`):

    Engine.surface_green_tdb   negf/surface.f90:385-605   (tb/surface.f90:129-343)
    Engine.get_G_full_tdb      negf/lesser.f90:475-809
    Engine.G_block             negf/device.f90:289-560    (tb/device.f90:291-575)
    Engine.solve_tdb           the driver's ie loop, negf/phinfortran.f90:1196-1251,
                               batched over energies and k-points on the GPU
    Engine.solve_buttiker      the same loop with Buttiker probes: get_G_full_tdb +
                               get_G_lesser_tdb (negf/lesser.f90:475-809, 1199-1503)

Matrices are numpy / torch complex128 in the natural (row-major) orientation;
the conversion to the column-major C ABI happens here.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _native as N

# default lifetime prefactor B of the probes (negf/config.f90:97)
B_PROBE = 5.0e-20 * (2.0 * 3.1415926 * 1.0e12)


def _f(a):
    """complex128, Fortran-ordered (the C ABI layout of one matrix)."""
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _stack_f(blocks):
    """(..., m, n) array -> contiguous buffer of column-major matrices."""
    a = np.asarray(blocks, dtype=np.complex128)
    return np.ascontiguousarray(np.swapaxes(a, -1, -2))


def params(kind: str = "phonon", **kw) -> N.NtParams:
    """Scalar parameters of the reference's module globals.

    kind='phonon': negf conventions, eta(E) = E*eta + eta0, device eta x10.
    kind='tb':     tb conventions, constant eta in leads and device."""
    p = N.NtParams()
    N.lib().nt_params_default(C.byref(p))
    if kind == "tb":
        p.eta_mode = N.NT_ETA_ABSOLUTE
        p.device_eta_factor = 1.0
        p.eta0 = 0.0
        p.dos_mode = N.NT_DOS_TB
    elif kind == "phonon":
        p.dos_mode = N.NT_DOS_NEGF
    else:
        raise ValueError(kind)
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = {"fp64": N.NT_PREC_FP64, "mixed": N.NT_PREC_MIXED}[v]
        setattr(p, k, v)
    return p


class Engine:
    """One GPU context (HIP stream + device workspace)."""

    def __init__(self, device: int = 0, batch_limit: Optional[int] = None, use_torch_stream: bool = True):
        try:
            import torch
            # torch sets up the HIP runtime before the library's first HIP call
            # (the other order leaves torch without devices in this process)
            torch.cuda.is_available()
        except ImportError:
            pass
        self._L = N.lib()
        ctx = C.c_void_p()
        st = self._L.nt_create(C.byref(ctx), int(device))
        if st != 0:
            raise N.TransportError(st, f"nt_create failed: {N.status_string(st)}")
        self.ctx = ctx
        self.device = device
        if batch_limit is not None:
            self._check(self._L.nt_set_batch_limit(self.ctx, int(batch_limit)), "nt_set_batch_limit")
        self._torch_stream = use_torch_stream

    def close(self):
        if self.ctx:
            self._L.nt_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != 0:
            err = self._L.nt_last_error(self.ctx).decode() if self.ctx else ""
            raise N.TransportError(st, f"{what}: {N.status_string(st)} ({st}) {err}".strip())

    def sync(self):
        self._check(self._L.nt_sync(self.ctx), "nt_sync")

    # ------------------------------------------------------------ per point
    def surface_green_tdb(self, Haml, Hamr, Hamlv, Hamrv, E, p: N.NtParams, want_plus_minus=True):
        """surface_green_tdb: returns gl_adv, gr_ret, gl_plus, gr_minus, gml_adv,
        gmr_ret, sl_adv, sr_ret (numpy, row-major) and the iteration counts."""
        Haml, Hamr, Hamlv, Hamrv = map(_f, (Haml, Hamr, Hamlv, Hamrv))
        nl, nr = Haml.shape[0], Hamr.shape[0]
        out = {k: np.zeros((nl, nl), np.complex128, order="F") for k in ("gl_adv", "gl_plus", "gml_adv", "sl_adv")}
        out.update({k: np.zeros((nr, nr), np.complex128, order="F") for k in ("gr_ret", "gr_minus", "gmr_ret", "sr_ret")})
        iters = (C.c_int * 4)()
        st = self._L.nt_surface_green_tdb(
            self.ctx, C.byref(p), nl, nr, float(E), _p(Haml), _p(Hamr), _p(Hamlv), _p(Hamrv),
            _p(out["gl_adv"]), _p(out["gr_ret"]), _p(out["gl_plus"]) if want_plus_minus else None,
            _p(out["gr_minus"]) if want_plus_minus else None, _p(out["gml_adv"]), _p(out["gmr_ret"]),
            _p(out["sl_adv"]), _p(out["sr_ret"]), C.cast(iters, C.c_void_p))
        self._check(st, "surface_green_tdb")
        res = {k: np.ascontiguousarray(v) for k, v in out.items()}
        res["iters"] = list(iters)
        return res

    def get_G_full_tdb(self, gl, gr, gml, gmr, sl, sr, E, hamd, hamdv, hamld, hamrd, haml, hamr,
                       p: N.NtParams, first_point=False):
        """get_G_full_tdb with the retarded lead quantities of get_G_full_tdb_mpi
        (gl = gl_adv^+, gml = -gml_adv, sl = sl_adv^+).  hamd: (L, b, b)."""
        hamd = np.asarray(hamd, np.complex128)
        L, b = hamd.shape[0], hamd.shape[1]
        hd = _stack_f(hamd)
        hdv = _stack_f(hamdv) if L > 1 else np.zeros((1, b, b), np.complex128)
        gl, gr, gml, gmr, sl, sr, hamld, hamrd, haml, hamr = map(_f, (gl, gr, gml, gmr, sl, sr, hamld, hamrd, haml, hamr))
        nl, nr = haml.shape[0], hamr.shape[0]
        G_D_RL = np.zeros((nr, nl), np.complex128, order="F")
        G_D_LR = np.zeros((nl, nr), np.complex128, order="F")
        G_L = np.zeros((nl, nl), np.complex128, order="F")
        G_R = np.zeros((nr, nr), np.complex128, order="F")
        tl, tr = C.c_double(), C.c_double()
        st = self._L.nt_get_G_full_tdb(
            self.ctx, C.byref(p), int(bool(first_point)), nl, nr, b, L, float(E),
            _p(gl), _p(gr), _p(gml), _p(gmr), _p(sl), _p(sr), _p(hd), _p(hdv), _p(hamld), _p(hamrd),
            _p(haml), _p(hamr), C.cast(C.byref(tl), C.c_void_p), C.cast(C.byref(tr), C.c_void_p),
            _p(G_D_RL), _p(G_D_LR), _p(G_L), _p(G_R))
        self._check(st, "get_G_full_tdb")
        return dict(transl=tl.value, transr=tr.value, G_D_RL=np.ascontiguousarray(G_D_RL),
                    G_D_LR=np.ascontiguousarray(G_D_LR), G_L=np.ascontiguousarray(G_L),
                    G_R=np.ascontiguousarray(G_R))

    def G_block(self, gl, gr, gml, gmr, sl, sr, E, Hd: Sequence, Hdv: Sequence, Hld, Hrd, Haml, Hamr,
                p: N.NtParams, Hdv_back: Optional[Sequence] = None, Hrd_fwd=None):
        """G_block on variable layer sizes (layer_list = [h.shape[0] for h in Hd])."""
        L = len(Hd)
        Hd = [_f(h) for h in Hd]
        Hdv = [_f(h) for h in Hdv]
        layer_list = (C.c_int * L)(*[h.shape[0] for h in Hd])
        pHd = (C.c_void_p * L)(*[h.ctypes.data for h in Hd])
        pHdv = (C.c_void_p * max(L - 1, 1))(*[h.ctypes.data for h in Hdv])
        pback = None
        if Hdv_back is not None:
            Hdv_back = [_f(h) for h in Hdv_back]
            pback = (C.c_void_p * max(L - 1, 1))(*[h.ctypes.data for h in Hdv_back])
        gl, gr, gml, gmr, sl, sr, Hld, Hrd, Haml, Hamr = map(_f, (gl, gr, gml, gmr, sl, sr, Hld, Hrd, Haml, Hamr))
        Hrd_fwd = None if Hrd_fwd is None else _f(Hrd_fwd)
        nl, nr = Haml.shape[0], Hamr.shape[0]
        G_D_RL = np.zeros((nr, nl), np.complex128, order="F")
        G_D_LR = np.zeros((nl, nr), np.complex128, order="F")
        G_L = np.zeros((nl, nl), np.complex128, order="F")
        G_R = np.zeros((nr, nr), np.complex128, order="F")
        tl, tr = C.c_double(), C.c_double()
        st = self._L.nt_G_block(
            self.ctx, C.byref(p), nl, nr, L, C.cast(layer_list, C.c_void_p), float(E),
            _p(gl), _p(gr), _p(gml), _p(gmr), _p(sl), _p(sr), C.cast(pHd, C.c_void_p), C.cast(pHdv, C.c_void_p),
            None if pback is None else C.cast(pback, C.c_void_p), _p(Hld), _p(Hrd), _p(Hrd_fwd), _p(Haml), _p(Hamr),
            C.cast(C.byref(tl), C.c_void_p), C.cast(C.byref(tr), C.c_void_p),
            _p(G_D_RL), _p(G_D_LR), _p(G_L), _p(G_R))
        self._check(st, "G_block")
        return dict(transl=tl.value, transr=tr.value, G_D_RL=np.ascontiguousarray(G_D_RL),
                    G_D_LR=np.ascontiguousarray(G_D_LR), G_L=np.ascontiguousarray(G_L),
                    G_R=np.ascontiguousarray(G_R))

    def get_ldos(self, gl, gr, E, Hd: Sequence, Hdv: Sequence, Hld, Hrd, p: N.NtParams,
                 Hdv_back: Optional[Sequence] = None, Hrd_fwd=None, maxlen: Optional[int] = None):
        """get_ldos: (ldos (L, maxlen), ddos) of the device layers."""
        L = len(Hd)
        Hd = [_f(h) for h in Hd]
        Hdv = [_f(h) for h in Hdv]
        maxlen = maxlen or max(h.shape[0] for h in Hd)
        layer_list = (C.c_int * L)(*[h.shape[0] for h in Hd])
        pHd = (C.c_void_p * L)(*[h.ctypes.data for h in Hd])
        pHdv = (C.c_void_p * max(L - 1, 1))(*[h.ctypes.data for h in Hdv])
        pback = None
        if Hdv_back is not None:
            Hdv_back = [_f(h) for h in Hdv_back]
            pback = (C.c_void_p * max(L - 1, 1))(*[h.ctypes.data for h in Hdv_back])
        gl, gr, Hld, Hrd = map(_f, (gl, gr, Hld, Hrd))
        Hrd_fwd = None if Hrd_fwd is None else _f(Hrd_fwd)
        nl, nr = gl.shape[0], gr.shape[0]
        ldos = np.zeros((L, maxlen), order="F")
        ddos = C.c_double()
        st = self._L.nt_get_ldos(self.ctx, C.byref(p), nl, nr, L, C.cast(layer_list, C.c_void_p), float(E),
                                 _p(gl), _p(gr), C.cast(pHd, C.c_void_p), C.cast(pHdv, C.c_void_p),
                                 None if pback is None else C.cast(pback, C.c_void_p), _p(Hld), _p(Hrd),
                                 _p(Hrd_fwd), int(maxlen), ldos.ctypes.data_as(C.c_void_p),
                                 C.cast(C.byref(ddos), C.c_void_p))
        self._check(st, "get_ldos")
        return np.ascontiguousarray(ldos), ddos.value

    # -------------------------------------------------------------- batched
    def upload_blocks(self, blocks_per_k):
        """Stack tdb block sets of nk k-points into device tensors (ABI layout)."""
        import torch
        if not isinstance(blocks_per_k, (list, tuple)):
            blocks_per_k = [blocks_per_k]
        dev = torch.device("cuda", self.device)
        out = {}
        for name in ("Haml", "Hamlv", "Hamr", "Hamrv", "Hamd", "Hamdv", "Hamld", "Hamrd"):
            arr = np.stack([np.asarray(getattr(b, name) if not isinstance(b, dict) else b[name]) for b in blocks_per_k])
            if name == "Hamdv" and arr.shape[1] == 0:
                arr = np.zeros((arr.shape[0], 1) + arr.shape[2:], np.complex128)
            out[name] = torch.from_numpy(_stack_f(arr)).to(dev)
        b0 = blocks_per_k[0]
        get = (lambda k: b0[k]) if isinstance(b0, dict) else (lambda k: getattr(b0, k))
        out["nl"] = get("Haml").shape[0]
        out["nr"] = get("Hamr").shape[0]
        out["b"] = get("Hamd").shape[1]
        out["L"] = get("Hamd").shape[0]
        out["nk"] = len(blocks_per_k)
        return out

    def solve_tdb(self, dblocks, E, p: N.NtParams, kidx=None, first_point_mask=None, extras=(), stream_sync=True,
                  noalloc=False):
        """Batched (omega,k)-solve on device-resident blocks (see upload_blocks).
        E, kidx: torch tensors (float64 / int32) on the GPU or array-likes.
        noalloc: solve through nt_solve_tdb_batched_noalloc_dev (workspace in the
        library's static scratch; the outputs are still allocated here).
        Returns dict(T_LR, T_RL, status, iters[, extras...]) of torch tensors."""
        import torch
        dev = torch.device("cuda", self.device)
        E = torch.as_tensor(E, dtype=torch.float64, device=dev).contiguous()
        npts = E.numel()
        nl, nr, b, L, nk = dblocks["nl"], dblocks["nr"], dblocks["b"], dblocks["L"], dblocks["nk"]
        if kidx is not None:
            kidx = torch.as_tensor(kidx, dtype=torch.int32, device=dev).contiguous()
        if first_point_mask is not None:
            first_point_mask = torch.as_tensor(first_point_mask, dtype=torch.int32, device=dev).contiguous()
        T_LR = torch.empty(npts, dtype=torch.float64, device=dev)
        T_RL = torch.empty(npts, dtype=torch.float64, device=dev)
        status = torch.zeros(npts, dtype=torch.int32, device=dev)
        iters = torch.zeros(npts, dtype=torch.int32, device=dev)
        ex, res = self._extras(extras, npts, nl, nr, dev, L=L, b=b)
        if self._torch_stream:
            self._L.nt_set_stream(self.ctx, C.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        entry = self._L.nt_solve_tdb_batched_noalloc_dev if noalloc else self._L.nt_solve_tdb_batched_dev
        st = entry(
            self.ctx, C.byref(p), nl, nr, b, L, npts, ptr(E), ptr(kidx), nk,
            ptr(dblocks["Haml"]), ptr(dblocks["Hamlv"]), ptr(dblocks["Hamr"]), ptr(dblocks["Hamrv"]),
            ptr(dblocks["Hamd"]), ptr(dblocks["Hamdv"]), ptr(dblocks["Hamld"]), ptr(dblocks["Hamrd"]),
            ptr(first_point_mask), ptr(T_LR), ptr(T_RL), ptr(status), ptr(iters), C.byref(ex))
        self._check(st, "solve_tdb")
        if stream_sync:
            torch.cuda.synchronize(dev)
        out = dict(T_LR=T_LR, T_RL=T_RL, status=status, iters=iters)
        for name, t in res.items():
            out[name] = t if not t.is_complex() else t.transpose(-1, -2)
        return out

    @staticmethod
    def _extras(names, npts, nl, nr, dev, L=None, b=None):
        """device buffers of the requested nt_solve_extras fields (complex
        matrices stored column-major: tensor [p, col, row]; ldos [p, layer,
        orbital] and ddos [p] of get_ldos, which need the device's L and b)"""
        import torch
        shapes = dict(gl_adv=(nl, nl), gr_ret=(nr, nr), gl_plus=(nl, nl), gr_minus=(nr, nr), gml_adv=(nl, nl),
                      gmr_ret=(nr, nr), sl_adv=(nl, nl), sr_ret=(nr, nr), G_D_LR=(nr, nl), G_D_RL=(nl, nr),
                      G_L=(nl, nl), G_R=(nr, nr))
        ex = N.NtSolveExtras()
        res = {}
        for name in names:
            if name in ("sdos_l", "sdos_r", "ddos"):
                t = torch.zeros(npts, dtype=torch.float64, device=dev)
            elif name == "ldos":
                if L is None or b is None:
                    raise ValueError("ldos is an output of the batched solve (solve_tdb) only")
                t = torch.zeros((npts, L, b), dtype=torch.float64, device=dev)
            else:
                t = torch.zeros((npts,) + shapes[name], dtype=torch.complex128, device=dev)
            res[name] = t
            setattr(ex, name, t.data_ptr())
        return ex, res

    def solve_tdb_adjoint(self, dblocks, E, p: N.NtParams, weight, lo: int, hi: int, kidx=None,
                          first_point_mask=None, stream_sync=True):
        """Batched solve plus the gradient of W = sum_i weight[i] T_LR[i] with
        respect to the device blocks of layers lo..hi (0-based).

        Returns dict(T_LR, T_RL, status, iters, dHamd, dHamdv): dHamd is
        (nk, hi-lo+1, b, b) and dHamdv (nk, nc, b, b) for the couplings
        l -> l+1, l = lo..min(hi, L-2), both row-major views with torch's
        complex-gradient convention (dW/dRe + i dW/dIm)."""
        import torch
        dev = torch.device("cuda", self.device)
        E = torch.as_tensor(E, dtype=torch.float64, device=dev).contiguous()
        npts = E.numel()
        w = torch.as_tensor(weight, dtype=torch.float64, device=dev).contiguous()
        if w.numel() != npts:
            raise ValueError("weight must have one entry per point")
        nl, nr, b, L, nk = dblocks["nl"], dblocks["nr"], dblocks["b"], dblocks["L"], dblocks["nk"]
        if not (0 <= lo <= hi < L):
            raise ValueError(f"layer window [{lo}, {hi}] outside [0, {L})")
        nc = min(hi, L - 2) - lo + 1
        if kidx is not None:
            kidx = torch.as_tensor(kidx, dtype=torch.int32, device=dev).contiguous()
        if first_point_mask is not None:
            first_point_mask = torch.as_tensor(first_point_mask, dtype=torch.int32, device=dev).contiguous()
        T_LR = torch.empty(npts, dtype=torch.float64, device=dev)
        T_RL = torch.empty(npts, dtype=torch.float64, device=dev)
        status = torch.zeros(npts, dtype=torch.int32, device=dev)
        iters = torch.zeros(npts, dtype=torch.int32, device=dev)
        gd = torch.zeros((nk, hi - lo + 1, b, b), dtype=torch.complex128, device=dev)
        gv = torch.zeros((nk, max(nc, 0), b, b), dtype=torch.complex128, device=dev)
        if self._torch_stream:
            self._L.nt_set_stream(self.ctx, C.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        st = self._L.nt_solve_tdb_adjoint_batched_dev(
            self.ctx, C.byref(p), nl, nr, b, L, npts, ptr(E), ptr(kidx), nk,
            ptr(dblocks["Haml"]), ptr(dblocks["Hamlv"]), ptr(dblocks["Hamr"]), ptr(dblocks["Hamrv"]),
            ptr(dblocks["Hamd"]), ptr(dblocks["Hamdv"]), ptr(dblocks["Hamld"]), ptr(dblocks["Hamrd"]),
            ptr(first_point_mask), ptr(w), int(lo), int(hi), ptr(T_LR), ptr(T_RL), ptr(status), ptr(iters),
            ptr(gd), ptr(gv) if nc > 0 else None)
        self._check(st, "solve_tdb_adjoint")
        if stream_sync:
            torch.cuda.synchronize(dev)
        return dict(T_LR=T_LR, T_RL=T_RL, status=status, iters=iters,
                    dHamd=gd.transpose(-1, -2), dHamdv=gv.transpose(-1, -2))

    def solve_buttiker(self, dblocks, E, p: N.NtParams, temperature, t_left: float, t_right: float,
                       B_probe: float = B_PROBE, T_probe: float = 430.0, lifetime: int = 1, kidx=None,
                       first_point_mask=None, stream_sync=True, extras=()):
        """Batched solve with a Buttiker probe in every device layer
        (get_G_full_tdb + get_G_lesser_tdb, negf/lesser.f90:475-809, 1199-1503).

        temperature: the L probe temperatures.  Returns dict(T_LR, T_RL, status,
        iters, flux (npts, L), jacobian (npts, L, L) with [p, i, j] =
        d flux_i / d T_j, flux_left, flux_right (npts), a_sum, gn_sum (npts, L)),
        plus the requested extras as solve_tdb returns them (with the probes
        in the sweeps)."""
        import torch
        dev = torch.device("cuda", self.device)
        E = torch.as_tensor(E, dtype=torch.float64, device=dev).contiguous()
        npts = E.numel()
        nl, nr, b, L, nk = dblocks["nl"], dblocks["nr"], dblocks["b"], dblocks["L"], dblocks["nk"]
        temp = torch.as_tensor(temperature, dtype=torch.float64, device=dev).contiguous()
        if temp.numel() != L:
            raise ValueError(f"{temp.numel()} probe temperatures for {L} layers")
        if kidx is not None:
            kidx = torch.as_tensor(kidx, dtype=torch.int32, device=dev).contiguous()
        if first_point_mask is not None:
            first_point_mask = torch.as_tensor(first_point_mask, dtype=torch.int32, device=dev).contiguous()
        f64 = dict(dtype=torch.float64, device=dev)
        T_LR = torch.empty(npts, **f64)
        T_RL = torch.empty(npts, **f64)
        status = torch.zeros(npts, dtype=torch.int32, device=dev)
        iters = torch.zeros(npts, dtype=torch.int32, device=dev)
        res = dict(flux=torch.zeros(npts, L, **f64), jacobian=torch.zeros(npts, L, L, **f64),
                   flux_left=torch.zeros(npts, **f64), flux_right=torch.zeros(npts, **f64),
                   a_sum=torch.zeros(npts, L, **f64), gn_sum=torch.zeros(npts, L, **f64))
        po = N.NtProbeOut(**{k: res[k].data_ptr() for k in N.PROBE_OUT_FIELDS})
        ex, exres = self._extras(extras, npts, nl, nr, dev)
        pd = N.NtProbeDesc(temp.data_ptr(), float(t_left), float(t_right), float(B_probe), float(T_probe),
                           int(lifetime))
        if self._torch_stream:
            self._L.nt_set_stream(self.ctx, C.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        st = self._L.nt_buttiker_tdb_batched_dev(
            self.ctx, C.byref(p), C.byref(pd), nl, nr, b, L, npts, ptr(E), ptr(kidx), nk,
            ptr(dblocks["Haml"]), ptr(dblocks["Hamlv"]), ptr(dblocks["Hamr"]), ptr(dblocks["Hamrv"]),
            ptr(dblocks["Hamd"]), ptr(dblocks["Hamdv"]), ptr(dblocks["Hamld"]), ptr(dblocks["Hamrd"]),
            ptr(first_point_mask), ptr(T_LR), ptr(T_RL), ptr(status), ptr(iters), C.byref(ex), C.byref(po))
        self._check(st, "solve_buttiker")
        if stream_sync:
            torch.cuda.synchronize(dev)
        res["jacobian"] = res["jacobian"].transpose(-1, -2)     # column-major per point -> [p, i, j]
        for name, t in exres.items():
            res[name] = t if t.dim() == 1 else t.transpose(-1, -2)
        return dict(T_LR=T_LR, T_RL=T_RL, status=status, iters=iters, **res)

    def solve_tdb_multi(self, devices, p: N.NtParams, stream_sync=True):
        """Several devices (uneven b, L, leads) in one call.  devices: list of
        dicts with 'blocks' (from upload_blocks), 'E' and optional 'kidx',
        'first_point_mask'.  Returns a list of dict(T_LR, T_RL, status); with
        stream_sync=False they are ready on torch's current stream."""
        import torch
        dev = torch.device("cuda", self.device)
        descs = (N.NtDeviceDesc * max(1, len(devices)))()
        outs, keep = [], []
        ptr = lambda t: None if t is None else t.data_ptr()
        # devices whose leads and points are equal by value get the same
        # buffers, so the engine computes their lead Green's functions once
        shared = []

        def share(t):
            if t is None:
                return None
            for u in shared:
                if u.shape == t.shape and u.dtype == t.dtype and bool(torch.equal(u, t)):
                    return u
            shared.append(t)
            return t
        for i, d in enumerate(devices):
            blk = dict(d["blocks"])
            for name in ("Haml", "Hamlv", "Hamr", "Hamrv"):
                blk[name] = share(blk[name])
            E = share(torch.as_tensor(d["E"], dtype=torch.float64, device=dev).contiguous())
            kidx = d.get("kidx")
            kidx = None if kidx is None else share(torch.as_tensor(kidx, dtype=torch.int32, device=dev).contiguous())
            mask = d.get("first_point_mask")
            mask = None if mask is None else torch.as_tensor(mask, dtype=torch.int32, device=dev).contiguous()
            n = E.numel()
            o = dict(T_LR=torch.empty(n, dtype=torch.float64, device=dev),
                     T_RL=torch.empty(n, dtype=torch.float64, device=dev),
                     status=torch.zeros(n, dtype=torch.int32, device=dev))
            outs.append(o)
            keep += [E, kidx, mask]
            D = descs[i]
            D.nl, D.nr, D.b, D.L, D.nk, D.npts = blk["nl"], blk["nr"], blk["b"], blk["L"], blk["nk"], n
            for name in ("Haml", "Hamlv", "Hamr", "Hamrv", "Hamd", "Hamdv", "Hamld", "Hamrd"):
                setattr(D, name, ptr(blk[name]))
            D.E, D.kidx, D.first_point_mask = ptr(E), ptr(kidx), ptr(mask)
            D.T_LR, D.T_RL, D.status = ptr(o["T_LR"]), ptr(o["T_RL"]), ptr(o["status"])
        if self._torch_stream:
            self._L.nt_set_stream(self.ctx, C.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
        self._check(self._L.nt_solve_tdb_multi_dev(self.ctx, C.byref(p), len(devices), C.cast(descs, C.c_void_p)),
                    "solve_tdb_multi")
        if stream_sync:
            torch.cuda.synchronize(dev)
        return outs

    MODE_FIELDS = ("gl_adv", "gr_ret", "gl_plus", "gr_minus", "gml_adv", "gmr_ret", "G_D_RL", "G_D_LR", "G_L", "G_R")

    def solve_modes(self, dblocks, E, p: N.NtParams, a_z_L: float, a_z_R: float, kidx=None, first_point_mask=None,
                    tb: bool = False, want=("trl", "tlr", "rrr", "rll"), unfold=None, ksc=None, pc=None, probe=None,
                    extras=()):
        """Batched solve followed by the mode-resolved transmission (vel_mat_tdb).

        Returns the solve's dict plus k_z_* (npts, n), n_prop (npts, 4),
        t_mode_L (npts, nl), t_mode_R (npts, nr), sums (npts, 4: T_RL, T_LR,
        R_RR, R_LL over propagating modes) and the requested t/r matrices as
        row-major (npts, rows, cols) views.

        unfold (phinfortran_amd.unfold.UnfoldGeometry) with ksc (nk x 3, the
        k-points in crystal coordinates of its recivec) adds the Brillouin-zone
        unfolding: bz_* (npts, n) labels, k_xyz_* (npts, n, 3) unfolded k,
        split (npts, 8) in the order of unfold.SPLIT_NAMES, and spec_trl,
        spec_tlr, spec_rrr, spec_rll (npts, n_incident, 2): (total, specular)
        per incident mode.  pc (unfold.PrimitiveCell, with unfold; or a
        (left, right) pair, as the tight-binding run's find_k_pc_l / _r with
        unfold.TbPrimitiveCell) adds the primitive-cell k and group velocity
        of every mode: kp_* and v_* (npts, n, 3), ncand_* (npts, n)
        (find_k_pc, see unfold_pc).

        probe (dict of solve_buttiker's keyword arguments: temperature, t_left,
        t_right[, B_probe, T_probe, lifetime]) solves with Buttiker probes in
        the sweeps, as vel_mat_tdb sees them in the reference's Buttiker loop,
        and adds the probe outputs.  extras: further outputs of the batched
        solve (names as in solve_tdb: sdos_l, sdos_r, ldos, ...)."""
        import torch
        dev = torch.device("cuda", self.device)
        ex = tuple(self.MODE_FIELDS) + tuple(x for x in extras if x not in self.MODE_FIELDS)
        if probe is None:
            r = self.solve_tdb(dblocks, E, p, kidx=kidx, first_point_mask=first_point_mask, extras=ex)
        else:
            r = self.solve_buttiker(dblocks, E, p, kidx=kidx, first_point_mask=first_point_mask, extras=ex, **probe)
        E = torch.as_tensor(E, dtype=torch.float64, device=dev).contiguous()
        npts = E.numel()
        nl, nr, nk = dblocks["nl"], dblocks["nr"], dblocks["nk"]
        if kidx is not None:
            kidx = torch.as_tensor(kidx, dtype=torch.int32, device=dev).contiguous()
        if first_point_mask is not None:
            first_point_mask = torch.as_tensor(first_point_mask, dtype=torch.int32, device=dev).contiguous()
        shapes = dict(trl=(nl, nr), tlr=(nr, nl), rrr=(nr, nr), rll=(nl, nl))    # storage [p, col, row]
        out = N.NtModesOut()
        res = {}
        for name in want:
            t = torch.zeros((npts,) + shapes[name], dtype=torch.complex128, device=dev)
            res[name] = t
            setattr(out, name, t.data_ptr())
        for name, n in (("k_z_L_adv", nl), ("k_z_L_ret", nl), ("k_z_R_ret", nr), ("k_z_R_adv", nr),
                        ("t_mode_L", nl), ("t_mode_R", nr), ("sums", 4)):
            t = torch.zeros((npts, n), dtype=torch.float64, device=dev)
            res[name] = t
            setattr(out, name, t.data_ptr())
        res["n_prop"] = torch.zeros((npts, 4), dtype=torch.int32, device=dev)
        out.n_prop = res["n_prop"].data_ptr()
        status = torch.zeros(npts, dtype=torch.int32, device=dev)
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        # extras come back transposed (row-major views); the ABI wants the storage
        g = {k: r[k].transpose(-1, -2) for k in self.MODE_FIELDS}
        args = (self.ctx, C.byref(p), int(tb), nl, nr, npts, ptr(E), ptr(kidx), nk,
                ptr(dblocks["Haml"]), ptr(dblocks["Hamlv"]), ptr(dblocks["Hamr"]), ptr(dblocks["Hamrv"]),
                *[ptr(g[k]) for k in self.MODE_FIELDS], float(a_z_L), float(a_z_R), ptr(first_point_mask),
                C.byref(out))
        if unfold is None:
            st = self._L.nt_modes_tdb_batched_dev(*args, ptr(status))
        else:
            if ksc is None:
                raise ValueError("unfolding needs the k-points (ksc)")
            desc, keep = unfold.desc(torch, dev, ksc)
            if keep["ksc"].shape[0] != nk:
                raise ValueError(f"ksc has {keep['ksc'].shape[0]} k-points, the blocks {nk}")
            uo = N.NtUnfoldOut()
            nn = dict(L_adv=nl, L_ret=nl, R_ret=nr, R_adv=nr)
            for key, n in nn.items():
                res["bz_" + key] = torch.zeros((npts, n), dtype=torch.int32, device=dev)
                res["k_xyz_" + key] = torch.zeros((npts, 3, n), dtype=torch.float64, device=dev)
            for key, n in (("trl", nl), ("tlr", nr), ("rrr", nr), ("rll", nl)):
                res["spec_" + key] = torch.zeros((npts, 2, n), dtype=torch.float64, device=dev)
            res["split"] = torch.zeros((npts, 8), dtype=torch.float64, device=dev)
            for name in N.UNFOLD_OUT_FIELDS:
                setattr(uo, name, res[name].data_ptr())
            st = self._L.nt_modes_unfold_tdb_batched_dev(*args, C.byref(desc), C.byref(uo), ptr(status))
        self._check(st, "modes_tdb")
        torch.cuda.synchronize(dev)
        for name in want:
            res[name] = res[name].transpose(-1, -2)
        if unfold is not None:
            for name in [k for k in res if k.startswith("k_xyz_") or k.startswith("spec_")]:
                res[name] = res[name].transpose(-1, -2)
            if pc is not None:
                pcs = tuple(pc) if isinstance(pc, (tuple, list)) else (pc, pc)
                for key in ("L_adv", "L_ret", "R_ret", "R_adv"):
                    kx = res["k_xyz_" + key]
                    n = kx.shape[1]
                    Em = E.reshape(-1, 1).expand(npts, n)
                    u = self.unfold_pc(pcs[0] if key[0] == "L" else pcs[1], Em.contiguous(), kx.contiguous())
                    res["kp_" + key] = u["kp"].reshape(npts, n, 3)
                    res["v_" + key] = u["v"].reshape(npts, n, 3)
                    res["ncand_" + key] = u["ncand"].reshape(npts, n)
        res["modes_status"] = status
        r.update(res)
        return r

    def unfold_pc(self, pc, E, k_xyz):
        """find_k_pc (negf/dyn.f90:770-889) for modes of energies E (nmodes)
        and unfolded k k_xyz (nmodes x 3, k_z >= 99.9 for evanescent modes) on
        the GPU; pc: phinfortran_amd.unfold.PrimitiveCell, or
        unfold.TbPrimitiveCell for the tight-binding run's find_k_pc_l / _r
        (tb/hamiltonian.f90:666-916).  Returns dict kp, v (nmodes x 3) and
        ncand (nmodes) as torch tensors."""
        import torch
        dev = torch.device("cuda", self.device)
        E = torch.as_tensor(E, dtype=torch.float64, device=dev).contiguous().reshape(-1)
        k = torch.as_tensor(k_xyz, dtype=torch.float64, device=dev).contiguous().reshape(-1, 3)
        n = E.numel()
        if k.shape[0] != n:
            raise ValueError("E and k_xyz disagree on the number of modes")
        kp = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        v = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        nc = torch.zeros(n, dtype=torch.int32, device=dev)
        ptr = lambda t: C.c_void_p(t.data_ptr())
        ru = (C.c_double * 9)(*pc.reci_uc.reshape(-1))
        rp = (C.c_double * 9)(*pc.recivec_pc.reshape(-1))
        if getattr(pc, "tb", False):
            el_off, coef, dr = pc.device_table(torch, dev)
            st = self._L.nt_unfold_pc_tb_dev(self.ctx, n, ptr(E), ptr(k), pc.l, ru, rp, pc.n_pc, ptr(el_off),
                                             int(coef.numel()), ptr(coef), ptr(dr), ptr(kp), ptr(v), ptr(nc))
        else:
            seg_off, seg_rc, coef, dr = pc.device_table(torch, dev)
            nseg, nent = len(pc.table[0]) - 1, len(pc.table[2])
            st = self._L.nt_unfold_pc_dev(self.ctx, n, ptr(E), ptr(k), pc.l, ru, rp, pc.n_pc, nseg, ptr(seg_off),
                                          ptr(seg_rc), nent, ptr(coef), ptr(dr), ptr(kp), ptr(v), ptr(nc))
        self._check(st, "unfold_pc")
        return dict(kp=kp, v=v, ncand=nc)

    # ------------------------------------------------------------ tuning
    def autotune_inverse(self, n: int, batch: int, precision: str = "fp64"):
        """Time every inverse panel width for n x n batches; keeps the fastest.
        Returns (panel width, ms per batched launch)."""
        prec = {"fp64": N.NT_PREC_FP64, "mixed": N.NT_PREC_MIXED}[precision]
        nb, ms = C.c_int(), C.c_double()
        self._check(self._L.nt_autotune_inverse(self.ctx, int(n), int(batch), prec, C.byref(nb), C.byref(ms)),
                    "nt_autotune_inverse")
        return nb.value, ms.value

    def autotune_gemm(self, n: int, batch: int, precision: str = "fp64"):
        """Time the 64- and 32-wide GEMM output tiles on n x n x n batches;
        keeps the faster for GEMMs whose larger output dimension is n.
        Returns (tile, ms per batched launch)."""
        prec = {"fp64": N.NT_PREC_FP64, "mixed": N.NT_PREC_MIXED}[precision]
        tb, ms = C.c_int(), C.c_double()
        self._check(self._L.nt_autotune_gemm(self.ctx, int(n), int(batch), prec, C.byref(tb), C.byref(ms)),
                    "nt_autotune_gemm")
        return tb.value, ms.value

    def set_gemm_tile(self, n: int, tile: int, precision: str = "fp64"):
        """GEMM output tile (64, 32; 0 = default) for larger output dimension n"""
        prec = {"fp64": N.NT_PREC_FP64, "mixed": N.NT_PREC_MIXED}[precision]
        self._check(self._L.nt_set_gemm_tile(self.ctx, int(n), prec, int(tile)), "nt_set_gemm_tile")

    def set_gemm_3m(self, on: bool):
        """FP64 GEMMs with three real products per complex product."""
        self._check(self._L.nt_set_gemm_3m(self.ctx, int(bool(on))), "nt_set_gemm_3m")

    def set_sweeps(self, n: int):
        """batched solves: 1 = left-to-right sweep + corner correction for the
        right-to-left quantities (default), 2 = both sweeps as the reference"""
        self._check(self._L.nt_set_sweeps(self.ctx, int(n)), "nt_set_sweeps")

    def set_decimation_check(self, every: int):
        """decimation: read the unconverged count every `every` iterations
        (default 4); 0 = never (max_iter iterations, no host synchronisation
        inside the FP64 batched solve, which can then be captured in a graph)"""
        self._check(self._L.nt_set_decimation_check(self.ctx, int(every)), "nt_set_decimation_check")

    def set_multi_workers(self, n: int):
        """uneven-device calls: groups of devices solved by n worker contexts
        side by side (default 4; 1 = one after the other on this context)"""
        self._check(self._L.nt_set_multi_workers(self.ctx, int(n)), "nt_set_multi_workers")

    def set_adjoint_checkpoint(self, every: int):
        """adjoint storage: every layer (0), every n-th (n), about sqrt (-1)"""
        self._check(self._L.nt_set_adjoint_checkpoint(self.ctx, int(every)), "nt_set_adjoint_checkpoint")

    def adjoint_bytes_per_point(self, dblocks, lo: int, hi: int) -> int:
        out = C.c_longlong()
        self._check(self._L.nt_adjoint_bytes_per_point(self.ctx, dblocks["nl"], dblocks["nr"], dblocks["b"],
                                                       dblocks["L"], int(lo), int(hi), C.byref(out)),
                    "nt_adjoint_bytes_per_point")
        return int(out.value)

    def set_overlap(self, on: bool):
        """With both sweeps (set_sweeps(2)): run them on two streams (default)
        or on one."""
        self._check(self._L.nt_set_overlap(self.ctx, int(bool(on))), "nt_set_overlap")

    def set_point_split(self, n: int):
        """Batched FP64 solves of at least n * 256 points run as n parts side
        by side on worker contexts (default 2); 1: one stream."""
        self._check(self._L.nt_set_point_split(self.ctx, int(n)), "nt_set_point_split")

    def set_workspace_poison(self, on: bool):
        """Debugging aid: fill the workspace with NaN bit patterns at the start
        of every call, so that a read of workspace the call did not write
        shows as a NaN status or output."""
        self._check(self._L.nt_set_workspace_poison(self.ctx, int(bool(on))), "nt_set_workspace_poison")

    def set_inverse_panel(self, n: int, nb: int, precision: str = "fp64"):
        prec = {"fp64": N.NT_PREC_FP64, "mixed": N.NT_PREC_MIXED}[precision]
        self._check(self._L.nt_set_inverse_panel(self.ctx, int(n), prec, int(nb)), "nt_set_inverse_panel")

    # ---------------------------------------------------------- measurement
    def profile(self, on: bool):
        self._check(self._L.nt_profile_enable(self.ctx, int(on)), "nt_profile_enable")

    def profile_reset(self):
        self._check(self._L.nt_profile_reset(self.ctx), "nt_profile_reset")

    def profile_get(self, kind: int):
        n, ms, fl = C.c_longlong(), C.c_double(), C.c_double()
        self._check(self._L.nt_profile_get(self.ctx, int(kind), C.byref(n), C.byref(ms), C.byref(fl)), "nt_profile_get")
        return dict(launches=n.value, ms=ms.value, flops=fl.value)

    def mixed_counts(self):
        """(points, second refinement steps, FP64 recomputes) of mixed-precision
        solves since the last profile_reset."""
        out = (C.c_longlong * 3)()
        self._check(self._L.nt_mixed_counts(self.ctx, C.cast(out, C.c_void_p)), "mixed_counts")
        return tuple(int(v) for v in out)

    def last_mean_iters(self) -> float:
        return float(self._L.nt_last_mean_iters(self.ctx))

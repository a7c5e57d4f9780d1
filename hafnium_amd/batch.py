"""Batched DF-J/K of many instances of uneven size.

Each instance i has its fitted tensor ``B_i [naux_i, nao_i, nao_i]`` (tiled
packed, laid end to end in one batch tensor), its occupied orbitals
``C_i [nao_i, nocc_i]`` and density ``D_i``; one run computes every
``J_i = sum_P B_i[P] <B_i[P], D_i>`` and ``K_i = sum_P (B_i[P] C_i)(B_i[P] C_i)^T``
with one persistent kernel launch per phase (``hfjk_batch_run``), which a
hipGraph can capture (:meth:`BatchedCoulombExchange.capture`).

The single-instance counterpart is ``MI355XCoulombExchange``; the reference
builds one molecule at a time (hafnium/scf/coulomb_exchange.py:176-190).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from hafnium_amd import _lib
from hafnium_amd.errors import BackendError, InputError


def _torch():
    import torch

    return torch


class _Instance(ctypes.Structure):
    _fields_ = [("nao", ctypes.c_int), ("naux", ctypes.c_int), ("nocc", ctypes.c_int),
                ("b_offset", ctypes.c_longlong), ("c_offset", ctypes.c_longlong), ("d_offset", ctypes.c_longlong)]


@dataclass(frozen=True)
class InstanceShape:
    nao: int
    naux: int
    nocc: int


class BatchedCoulombExchange:
    """J and K of a batch of instances resident in HBM."""

    def __init__(self, shapes, *, device=None) -> None:
        from hafnium_amd.coulomb_exchange import _Context

        torch = _torch()
        if not torch.cuda.is_available():
            raise BackendError("no GPU is visible to PyTorch-ROCm")
        self.shapes = [InstanceShape(int(a), int(b), int(c)) for a, b, c in shapes]
        if not self.shapes:
            raise InputError("a batch holds at least one instance")
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self._ctx = _Context(self.device.index or 0)
        lib = self._ctx.lib
        b_off = c_off = d_off = 0
        self._inst = (_Instance * len(self.shapes))()
        self.b_offsets, self.c_offsets, self.d_offsets = [], [], []
        for i, sh in enumerate(self.shapes):
            if not (1 <= sh.nocc <= sh.nao) or sh.naux < 1:
                raise InputError(f"instance {i}: 1 <= nocc <= nao and naux >= 1 are required")
            p = int(lib.hfjk_packed_elems(sh.nao))
            self._inst[i] = _Instance(sh.nao, sh.naux, sh.nocc, b_off, c_off, d_off)
            self.b_offsets.append(b_off)
            self.c_offsets.append(c_off)
            self.d_offsets.append(d_off)
            b_off += sh.naux * p
            c_off += sh.nao * sh.nocc
            d_off += sh.nao * sh.nao
        self.tensor_elems, self.orbital_elems, self.matrix_elems = b_off, c_off, d_off
        self._b64 = torch.empty(b_off, dtype=torch.float64, device=self.device)
        self._copies = {}
        self._plans = {}
        self._layouts = {None: (self._inst, self.c_offsets, self.orbital_elems)}

    def layout(self, noccs=None):
        """``(instances, c_offsets, orbital_elems)`` for other occupied counts per instance
        (``None``: the counts the batch was made with). The tensor is shared."""
        key = None if noccs is None else tuple(int(x) for x in noccs)
        if key not in self._layouts:
            if len(key) != len(self.shapes):
                raise InputError("one occupied count per instance is required")
            inst = (_Instance * len(self.shapes))()
            offs, c_off = [], 0
            for i, (sh, no) in enumerate(zip(self.shapes, key)):
                if not 1 <= no <= sh.nao:
                    raise InputError(f"instance {i}: 1 <= nocc <= nao is required")
                inst[i] = _Instance(sh.nao, sh.naux, no, self.b_offsets[i], c_off, self.d_offsets[i])
                offs.append(c_off)
                c_off += sh.nao * no
            self._layouts[key] = (inst, offs, c_off)
        return self._layouts[key]

    # ------------------------------------------------------------- inputs

    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)

    def generate(self, seeds) -> None:
        """Fill every B_i from the package's seeded generator (``hafnium_amd.synthetic``)."""
        lib = self._ctx.lib
        for sh, off, seed in zip(self.shapes, self.b_offsets, seeds):
            ptr = ctypes.c_void_p(self._b64.data_ptr() + 8 * off)
            _lib.check(lib.hfjk_generate(self._ctx.handle, int(seed), 0, sh.naux, sh.nao, ptr, self._stream()))
        self._copies.clear()

    def set_tensors(self, tensors) -> None:
        """Pack full symmetric ``[naux, nao, nao]`` arrays into the batch tensor."""
        for i, b in enumerate(tensors):
            self.set_tensor(i, b)

    def set_tensor(self, i: int, b) -> None:
        """Pack instance ``i``'s full symmetric ``[naux, nao, nao]`` tensor."""
        torch = _torch()
        lib = self._ctx.lib
        sh, off = self.shapes[i], self.b_offsets[i]
        full = torch.as_tensor(np.asarray(b) if isinstance(b, np.ndarray) else b,
                               dtype=torch.float64).to(self.device).contiguous()
        if tuple(full.shape) != (sh.naux, sh.nao, sh.nao):
            raise InputError("a tensor does not match its instance shape")
        ptr = ctypes.c_void_p(self._b64.data_ptr() + 8 * off)
        _lib.check(lib.hfjk_pack(self._ctx.handle, ctypes.c_void_p(full.data_ptr()), sh.naux, sh.nao, ptr, None,
                                 self._stream()))
        self._copies.clear()

    def pack_orbitals(self, orbitals):
        """Concatenate ``[nao_i, nocc_i]`` matrices into the flat orbital buffer."""
        torch = _torch()
        out = torch.empty(self.orbital_elems, dtype=torch.float64, device=self.device)
        for sh, off, c in zip(self.shapes, self.c_offsets, orbitals):
            t = torch.as_tensor(np.asarray(c) if isinstance(c, np.ndarray) else c, dtype=torch.float64)
            if tuple(t.shape) != (sh.nao, sh.nocc):
                raise InputError("orbitals do not match their instance shape")
            out[off:off + sh.nao * sh.nocc] = t.reshape(-1).to(self.device)
        return out

    def pack_matrices(self, matrices):
        torch = _torch()
        out = torch.empty(self.matrix_elems, dtype=torch.float64, device=self.device)
        for sh, off, m in zip(self.shapes, self.d_offsets, matrices):
            t = torch.as_tensor(np.asarray(m) if isinstance(m, np.ndarray) else m, dtype=torch.float64)
            out[off:off + sh.nao * sh.nao] = t.reshape(-1).to(self.device)
        return out

    def matrix(self, flat, i):
        """Instance ``i``'s ``[nao, nao]`` view of a flat J/K/D buffer."""
        sh = self.shapes[i]
        off = self.d_offsets[i]
        return flat[off:off + sh.nao * sh.nao].view(sh.nao, sh.nao)

    # ------------------------------------------------------------- plans

    def _tensor(self, precision: int):
        torch = _torch()
        if precision == _lib.FP64:
            return self._b64
        if precision not in self._copies:
            lib = self._ctx.lib
            if precision == _lib.FP32:
                copy = torch.empty(self.tensor_elems, dtype=torch.float32, device=self.device)
            else:
                copy = torch.empty(2 * self.tensor_elems, dtype=torch.int16, device=self.device)
            _lib.check(lib.hfjk_split(self._ctx.handle, ctypes.c_void_p(self._b64.data_ptr()), self.tensor_elems,
                                      precision, ctypes.c_void_p(copy.data_ptr()), self._stream()))
            self._copies[precision] = copy
        return self._copies[precision]

    def release(self, precision: str) -> None:
        """Free the workspaces (every orbital layout) and derived tensor copy of ``precision``."""
        code = _lib.precision_code(precision)
        for key in [k for k in self._plans if k[0] == code]:
            plan = self._plans.pop(key)
            self._ctx.lib.hfjk_batch_destroy(plan[0])
        self._copies.pop(code, None)

    def plan(self, precision: str = "fp64", tile: int = 0, noccs=None):
        """Create (or reuse) the plan and workspace of ``precision`` with column tile ``tile``
        for the orbital layout of ``noccs`` (see :meth:`layout`)."""
        torch = _torch()
        code = _lib.precision_code(precision)
        inst, _, _ = self.layout(noccs)
        key = (code, None if noccs is None else tuple(int(x) for x in noccs))
        if key in self._plans and self._plans[key][2] == tile:
            return self._plans[key]
        if key in self._plans:
            self._ctx.lib.hfjk_batch_destroy(self._plans.pop(key)[0])
        lib = self._ctx.lib
        handle = ctypes.c_void_p()
        _lib.check(lib.hfjk_batch_create(self._ctx.handle, inst, len(self.shapes), code, int(tile),
                                         ctypes.byref(handle)))
        size = ctypes.c_size_t()
        _lib.check(lib.hfjk_batch_workspace_bytes(handle, ctypes.byref(size)))
        ws = torch.empty(int(size.value) + 256, dtype=torch.uint8, device=self.device)
        base = ws.data_ptr()
        aligned = (base + 255) & ~255
        _lib.check(lib.hfjk_batch_bind(handle, ctypes.c_void_p(aligned), int(size.value), self._stream()))
        self._plans[key] = (handle, ws, tile)
        return self._plans[key]

    # ------------------------------------------------------------- run

    def run(self, orbitals, densities=None, *, precision: str = "fp64", tile: int = 0, out=None, noccs=None):
        """One batched J/K build; returns flat ``(J, K)`` buffers (see :meth:`matrix`).
        ``orbitals`` follow the layout of ``noccs`` (default: the batch's own counts). Without
        ``densities`` each instance's density is ``C C^T`` of its orbitals and its Coulomb charges
        are formed from the half-transformed ``X = B C`` (the fitted tensor is read once less)."""
        torch = _torch()
        code = _lib.precision_code(precision)
        handle, _, _ = self.plan(precision, tile, noccs)
        tensor = self._tensor(code)
        if out is None:
            out = (torch.empty(self.matrix_elems, dtype=torch.float64, device=self.device),
                   torch.empty(self.matrix_elems, dtype=torch.float64, device=self.device))
        j, k = out
        _lib.check(self._ctx.lib.hfjk_batch_run(
            handle, ctypes.c_void_p(tensor.data_ptr()), ctypes.c_void_p(orbitals.data_ptr()),
            ctypes.c_void_p(densities.data_ptr()) if densities is not None else None,
            ctypes.c_void_p(j.data_ptr()), ctypes.c_void_p(k.data_ptr()),
            self._stream()))
        return j, k

    def coulomb_from_charges(self, charges, *, precision: str = "fp64", out=None):
        """``J_i = sum_P charges_i[P] B_i[P]`` (charges flat, instance after instance); flat J buffer."""
        torch = _torch()
        total = sum(sh.naux for sh in self.shapes)
        if charges.dtype != torch.float64 or charges.numel() != total or charges.device != self.device:
            raise InputError(f"charges are a float64 device vector of {total} values")
        code = _lib.precision_code(precision)
        handle, _, _ = self.plan(precision)
        tensor = self._tensor(code)
        if out is None:
            out = torch.empty(self.matrix_elems, dtype=torch.float64, device=self.device)
        _lib.check(self._ctx.lib.hfjk_batch_coulomb_charges(
            handle, ctypes.c_void_p(tensor.data_ptr()), ctypes.c_void_p(charges.contiguous().data_ptr()),
            ctypes.c_void_p(out.data_ptr()), self._stream()))
        return out

    def capture(self, orbitals, densities=None, *, precision: str = "fp64", tile: int = 0):
        """A hipGraph (``torch.cuda.CUDAGraph``) of one batched build and its output buffers.

        The plan and tensor copy are prepared before capture; the captured work
        is the counter reset and the persistent phase kernels only. The graph
        addresses the plan's workspace, the tensor copy and the input buffers
        by pointer, so it holds references to all of them: a later ``plan`` with
        another tile or ``release`` drops the batch's own references, and the
        buffers live on until the graph is freed.
        """
        torch = _torch()
        plan = self.plan(precision, tile)
        copy = self._tensor(_lib.precision_code(precision))
        out = (torch.empty(self.matrix_elems, dtype=torch.float64, device=self.device),
               torch.empty(self.matrix_elems, dtype=torch.float64, device=self.device))
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.run(orbitals, densities, precision=precision, tile=tile, out=out)
        graph.hfjk_buffers = (plan[1], copy, self._b64, orbitals, densities, out)
        return graph, out

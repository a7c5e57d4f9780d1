"""Batched restricted Hartree-Fock: many molecules converged together on one GPU.

Every iteration advances all molecules at once: one batched DF-J/K build
(``BatchedCoulombExchange``: one persistent launch per phase for every
molecule), strided-batched GEMMs and a batched eigensolver on matrices
zero-padded to the largest basis (for the Fock matrices the library's
one-launch solver of the occupied eigenpairs, ``hfjk_eigh_lowest_batched``;
rocSOLVER for the overlaps), and one host transfer of every
molecule's energy terms and newest DIIS overlaps (the DIIS overlap
matrices and coefficients stay on the device). Per molecule, the iteration
is the one of ``solve_scf`` (hafnium/scf/driver.py:154-213 of the
reference): F = h + J[2D] - K[D], the orbital gradient X^T (F D S - S D F) X,
DIIS over the last ``diis_size`` (at most 16) Fock matrices
(hafnium/scf/diis.py:27-106, solved by ``hfjk_diis_batched``), the
canonical orthogonaliser (orthogonalization.py:16-42) and aufbau filling.
Padded and discarded directions of the orthogonal basis get a large
diagonal shift, so they sort above every occupied level and never mix in.

Precision policy: all molecules share the arithmetic of the J/K build; it
switches to FP64 (and resets DIIS) once at least half of the unconverged
molecules have their gradient below ``options.switch_at(precision)`` or have stalled
(``mixed_phase_stalled``); switching early only moves molecules to FP64.
Convergence of a molecule (|dE| < energy_tolerance and gradient below
gradient_tolerance) is only declared on FP64 iterations; its energy and
iteration count are recorded at that iteration while the batch runs on.

Starts: from a density with a factor D = C C^T (the superposition of
atomic densities: blocks of atomic natural orbitals), or from a guess Fock
matrix (the learned guess), each evaluated with one batched build; the
learned start falls back to the atomic one per molecule when its energy is
higher.

No counterpart in the reference (its driver solves one molecule at a time).
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from hafnium_amd import _lib
from hafnium_amd.errors import BackendError, ConvergenceError, InputError
from hafnium_amd.molecule import BasisOnAtoms, Molecule
from hafnium_amd.scf import ScfOptions, mixed_phase_stalled

DISCARDED_SHIFT = 1.0e3  # Eh added to padded / discarded directions of the orthogonal basis
EIGH_LOWEST_MAX_N = 512  # largest matrix of hfjk_eigh_lowest_batched


def _torch():
    import torch

    return torch


def _p(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


@dataclass
class MoleculeInputs:
    """Host integrals of one molecule (closed shell)."""

    molecule: Molecule
    overlap: np.ndarray
    core: np.ndarray
    metric: np.ndarray
    three_center: np.ndarray
    n_alpha: int
    nuclear_repulsion: float
    aux_basis: str

    @property
    def nao(self) -> int:
        return self.overlap.shape[0]

    @property
    def naux(self) -> int:
        return self.metric.shape[0]


def molecule_inputs(molecule: Molecule, *, basis: str = "cc-pvdz", aux: str = "def2-universal-jkfit",
                    device_integrals: bool = True, device=None):
    """S, T + V (host) and the metric and three-centre integrals (device tensors computed on the
    GPU, or host arrays from the host engine with ``device_integrals=False``)."""
    from hafnium_amd import integrals

    if molecule.n_electrons % 2:
        raise InputError("the batched driver is restricted closed-shell")
    orb = BasisOnAtoms(basis, molecule)
    auxb = BasisOnAtoms(aux, molecule)
    s, t, v = integrals.one_electron(orb, molecule)
    if device_integrals:
        m = integrals.metric_device(auxb, device=device)
        tc = integrals.three_center_device(auxb, orb, device=device)
    else:
        m, tc = integrals.metric(auxb), integrals.three_center(auxb, orb)
    return MoleculeInputs(molecule, s, t + v, m, tc, molecule.n_electrons // 2, molecule.nuclear_repulsion(), aux)


def _host(a) -> np.ndarray:
    return a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)


@dataclass
class BatchResult:
    energy: float
    converged: bool
    iterations: int
    seconds: float
    history: list = field(default_factory=list)  # (energy, gradient norm, precision) per iteration


@dataclass
class _Start:
    density: object      # padded [N, nm, nm]
    fock: object         # padded [N, nm, nm]
    terms: np.ndarray    # [3, N]: tr(D h), tr(D J[D]), tr(D K[D])
    info: dict = field(default_factory=dict)


class BatchedRhf:
    """Host integrals of N molecules -> fitted tensors in one batch, padded FP64 state on the GPU."""

    def __init__(self, inputs, *, device=None, eigensolver: str = "lowest") -> None:
        from hafnium_amd.batch import BatchedCoulombExchange
        from hafnium_amd.coulomb_exchange import _Context
        from hafnium_amd.linalg import Dense

        torch = _torch()
        self.inputs = list(inputs)
        if not self.inputs:
            raise InputError("a batch holds at least one molecule")
        if eigensolver not in ("lowest", "jacobi", "divide"):
            raise InputError("eigensolver is 'lowest', 'jacobi' or 'divide'")
        # The overlaps need every eigenvector (rocSOLVER, divide and conquer or Jacobi); the Fock
        # solves need the occupied ones: "lowest" takes them from the library's one-launch solver
        # (matrices up to 512 functions; divide and conquer beyond).
        self.method = 1 if eigensolver == "jacobi" else 0
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.count = len(self.inputs)
        self.nm = max(x.nao for x in self.inputs)
        self.om = max(x.n_alpha for x in self.inputs)
        self.fock_lowest = eigensolver == "lowest" and self.nm <= EIGH_LOWEST_MAX_N
        self._lowest = None  # (work, info) of the lowest-eigenpair solver
        self.jk = BatchedCoulombExchange([(x.nao, x.naux, x.n_alpha) for x in self.inputs], device=self.device)
        self._ctx = self.jk._ctx
        self.lib = self._ctx.lib
        dense = Dense(self._ctx, self.device)
        self.dense = dense
        t0 = time.perf_counter()
        # Fit on the GPU, one molecule at a time: M = L L^T, B = L^-1 (P|uv); L is kept
        # (the learned guess needs L^-1 (P|Q) d).
        self.factors = []
        info = torch.zeros(len(self.inputs), dtype=torch.int32, device=self.device)
        for i, x in enumerate(self.inputs):
            m = torch.as_tensor(x.metric, dtype=torch.float64).to(self.device).clone().contiguous()
            _lib.check(self.lib.hfjk_dpotrf(self._ctx.handle, x.naux, _p(m), _p(info[i:i + 1]), self._stream()))
            b = torch.as_tensor(x.three_center, dtype=torch.float64).to(self.device).reshape(
                x.naux, -1).clone().contiguous()
            _lib.check(self.lib.hfjk_dtrsm_lower(self._ctx.handle, x.naux, x.nao * x.nao, _p(m), _p(b),
                                                 self._stream()))
            # (B + B^T) / 2: removes the triangular solve's rounding asymmetry (see density_fitting.df_tensor).
            _lib.check(self.lib.hfjk_symmetrize_batched(self._ctx.handle, x.nao, x.naux, _p(b), self._stream()))
            self.jk.set_tensor(i, b.view(x.naux, x.nao, x.nao))
            del b
            self.factors.append(m)
        torch.cuda.synchronize(self.device)
        bad = np.nonzero(info.cpu().numpy())[0]
        if len(bad):
            raise InputError(f"the metric of molecule(s) {bad.tolist()} is not positive definite")
        self.fit_seconds = time.perf_counter() - t0
        n, nm, cnt = self.nm, self.nm, self.count
        self._nn = nm * nm
        zeros = lambda *shape: torch.zeros(shape, dtype=torch.float64, device=self.device)  # noqa: E731
        s_pad, h_pad = zeros(cnt, nm, nm), zeros(cnt, nm, nm)
        for i, x in enumerate(self.inputs):
            s_pad[i, :x.nao, :x.nao] = torch.from_numpy(x.overlap)
            h_pad[i, :x.nao, :x.nao] = torch.from_numpy(x.core)
        self.s, self.h = s_pad, h_pad
        as_int = lambda v: torch.tensor(v, dtype=torch.int32, device=self.device)  # noqa: E731
        as_ll = lambda v: torch.tensor(v, dtype=torch.int64, device=self.device)  # noqa: E731
        self._nao = as_int([x.nao for x in self.inputs])
        self._nocc = as_int([x.n_alpha for x in self.inputs])
        self._moffs = as_ll(self.jk.d_offsets)
        self._coffs = as_ll(self.jk.c_offsets)
        self._max_mat = max(x.nao * x.nao for x in self.inputs)
        self._max_orb = max(x.nao * x.n_alpha for x in self.inputs)
        self.e_nuc = np.array([x.nuclear_repulsion for x in self.inputs])
        # Orthogonaliser from the batched eigendecomposition of the padded overlaps.
        vecs = s_pad.clone()
        vals = zeros(cnt, nm)
        # Failed batched diagonalisations (any molecule), accumulated on the device and read
        # with the per-iteration transfer.
        self._eig_failures = zeros(1)
        self._eig(vecs, vals)
        if int(self._eig_failures.item()):
            raise BackendError("the batched eigensolver failed on an overlap matrix")
        self.x = zeros(cnt, nm, nm)
        self.kept = torch.zeros((cnt, nm), dtype=torch.int32, device=self.device)
        _lib.check(self.lib.hfjk_orthogonalizer_batched(self._ctx.handle, nm, cnt, _p(vecs), _p(vals),
                                                        ScfOptions().overlap_threshold, _p(self.x), _p(self.kept),
                                                        self._stream()))
        kept = self.kept.sum(dim=1).cpu().numpy()
        for i, x in enumerate(self.inputs):
            if kept[i] < x.n_alpha:
                raise InputError(f"molecule {i}: {x.n_alpha} occupied orbitals do not fit {kept[i]} functions")
        # Work buffers.
        self._jc = torch.empty(self.jk.matrix_elems, dtype=torch.float64, device=self.device)
        self._kc = torch.empty_like(self._jc)
        self._cc = torch.empty(self.jk.orbital_elems, dtype=torch.float64, device=self.device)
        self._jpad, self._kpad = zeros(cnt, nm, nm), zeros(cnt, nm, nm)
        self._t1, self._t2 = zeros(cnt, nm, nm), zeros(cnt, nm, nm)

    # ------------------------------------------------------------ primitives
    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)

    def _eig(self, a, w):
        torch = _torch()
        cnt, nm = self.count, self.nm
        work = torch.empty(int(self.lib.hfjk_dsyev_batched_work(nm, cnt)), dtype=torch.float64, device=self.device)
        info = torch.zeros(cnt, dtype=torch.int32, device=self.device)
        _lib.check(self.lib.hfjk_dsyev_batched(self._ctx.handle, nm, cnt, _p(a), _p(w), _p(work), _p(info),
                                               self.method, self._stream()))
        self._eig_failures += (info != 0).any().to(torch.float64)

    def _eig_lowest(self, a, w):
        """The om lowest eigenpairs of every matrix of ``a`` (rows 0 .. om - 1, eigenvalues in ``w[:, :om]``)."""
        torch = _torch()
        cnt, nm, om = self.count, self.nm, self.om
        if self._lowest is None:
            work = torch.empty(int(self.lib.hfjk_eigh_lowest_batched_work(nm, cnt, om)), dtype=torch.float64,
                               device=self.device)
            self._lowest = (work, torch.zeros(cnt, dtype=torch.int32, device=self.device))
        work, info = self._lowest
        _lib.check(self.lib.hfjk_eigh_lowest_batched(self._ctx.handle, nm, cnt, om, _p(a), _p(w), _p(work), _p(info),
                                                     self._stream()))

    def _gemm(self, ta, tb, m, n, k, a, lda, sa, b, ldb, sb, c, ldc, sc, alpha=1.0, beta=0.0):
        _lib.check(self.lib.hfjk_dgemm_batched(self._ctx.handle, int(ta), int(tb), m, n, k, alpha, _p(a), lda, sa,
                                               _p(b), ldb, sb, beta, _p(c), ldc, sc, self.count, self._stream()))

    def _mm(self, a, b, c, ta=False, tb=False):
        nm, nn = self.nm, self._nn
        self._gemm(ta, tb, nm, nm, nm, a, nm, nn, b, nm, nn, c, nm, nn)

    def _copy(self, rows, cols, offs, max_elems, padded, ld, stride, compact, to_compact):
        _lib.check(self.lib.hfjk_copy_batched(self._ctx.handle, self.count, _p(rows), _p(cols), _p(offs), max_elems,
                                              _p(padded), ld, stride, _p(compact), int(to_compact), self._stream()))

    def _dots(self, x, y, out):
        _lib.check(self.lib.hfjk_ddot_batched(self._ctx.handle, self.count, self._nn, _p(x), self._nn, _p(y),
                                              self._nn, _p(out), self._stream()))

    def _combine(self, a, x, b, y, c, z, out):
        _lib.check(self.lib.hfjk_combine(self._ctx.handle, out.numel(), a, _p(x), b, _p(y) if y is not None else None,
                                         c, _p(z) if z is not None else None, _p(out), self._stream()))

    # ------------------------------------------------------------ SCF steps
    def fock(self, density, orbitals_compact, precision, out, terms, noccs=None):
        """F = h + 2 J[D] - K[D] into ``out`` for D = C C^T of ``orbitals_compact`` (the padded
        ``density`` is that product; J's charges come from the build's half-transform);
        terms[0:3] (device rows) = tr(D h), tr(D J), tr(D K)."""
        self.jk.run(orbitals_compact, None, precision=precision, out=(self._jc, self._kc), noccs=noccs)
        self._copy(self._nao, self._nao, self._moffs, self._max_mat, self._jpad, self.nm, self._nn, self._jc, False)
        self._copy(self._nao, self._nao, self._moffs, self._max_mat, self._kpad, self.nm, self._nn, self._kc, False)
        self._combine(1.0, self.h, 2.0, self._jpad, -1.0, self._kpad, out)
        self._dots(density, self.h, terms[0])
        self._dots(density, self._jpad, terms[1])
        self._dots(density, self._kpad, terms[2])

    def gradient(self, fock, density, out):
        """X^T (F D S - S D F) X into ``out``."""
        self._mm(fock, density, self._t1)
        self._mm(self._t1, self.s, self._t2)
        _lib.check(self.lib.hfjk_antisymmetrize_batched(self._ctx.handle, self.nm, self.count, _p(self._t2),
                                                        _p(self._t1), self._stream()))
        self._mm(self.x, self._t1, self._t2, ta=True)
        self._mm(self._t2, self.x, out)

    def solve(self, fock, mark=None):
        """Aufbau density and occupied orbitals of ``fock``: (D padded, C padded [N, nm, om], C compact)."""
        torch = _torch()
        mark = mark or (lambda stage: None)
        nm, om, nn = self.nm, self.om, self._nn
        self._mm(self.x, fock, self._t1, ta=True)
        fp = torch.empty_like(self._t1)
        self._mm(self._t1, self.x, fp)
        _lib.check(self.lib.hfjk_shift_discarded_batched(self._ctx.handle, nm, self.count, _p(self.kept),
                                                         DISCARDED_SHIFT, _p(fp), self._stream()))
        w = torch.empty((self.count, nm), dtype=torch.float64, device=self.device)
        mark("solve_transform")
        if self.fock_lowest:
            self._eig_lowest(fp, w)
        else:
            self._eig(fp, w)
        mark("solve_eigensolver")
        c_occ = torch.empty((self.count, nm, om), dtype=torch.float64, device=self.device)
        # C_occ = X V_occ^T: rows 0..om-1 of each eigenvector matrix are the lowest levels.
        self._gemm(False, True, nm, om, nm, self.x, nm, nn, fp, nm, nn, c_occ, om, nm * om)
        _lib.check(self.lib.hfjk_mask_columns_batched(self._ctx.handle, nm, om, self.count, _p(self._nocc),
                                                      _p(c_occ), self._stream()))
        density = torch.empty((self.count, nm, nm), dtype=torch.float64, device=self.device)
        self._gemm(False, True, nm, nm, om, c_occ, om, nm * om, c_occ, om, nm * om, density, nm, nn)
        compact = torch.empty(self.jk.orbital_elems, dtype=torch.float64, device=self.device)
        self._copy(self._nao, self._nocc, self._coffs, self._max_orb, c_occ, om, nm * om, compact, True)
        mark("solve_density")
        return density, c_occ, compact

    # ------------------------------------------------------------ starts
    def start_from_factors(self, densities, factors, precision: str = "fp64") -> _Start:
        """Start from host densities with factors D_i = C_i C_i^T ([nao_i, r_i])."""
        torch = _torch()
        cnt, nm = self.count, self.nm
        d_pad = torch.zeros((cnt, nm, nm), dtype=torch.float64, device=self.device)
        ranks = []
        for i, (d, c) in enumerate(zip(densities, factors)):
            n = self.inputs[i].nao
            if d.shape != (n, n) or c.shape[0] != n or c.shape[1] < 1:
                raise InputError(f"molecule {i}: density or factor has the wrong shape")
            d_pad[i, :n, :n] = torch.from_numpy(np.ascontiguousarray(d))
            ranks.append(int(c.shape[1]))
        _, offs, total = self.jk.layout(ranks)
        c_flat = np.empty(total)
        for i, c in enumerate(factors):
            c_flat[offs[i]:offs[i] + c.size] = np.ascontiguousarray(c).reshape(-1)
        c_dev = torch.from_numpy(c_flat).to(self.device)
        fock = torch.empty_like(d_pad)
        terms = torch.empty((3, cnt), dtype=torch.float64, device=self.device)
        self.fock(d_pad, c_dev, precision, fock, terms, noccs=ranks)
        return _Start(d_pad, fock, terms.cpu().numpy(), {"coulomb": self._jpad.clone()})

    def start_from_fock(self, guess_fock, precision: str = "fp64") -> _Start:
        """Start from the aufbau density of a guess Fock matrix (padded [N, nm, nm], device)."""
        torch = _torch()
        density, _, compact = self.solve(guess_fock)
        fock = torch.empty_like(density)
        terms = torch.empty((3, self.count), dtype=torch.float64, device=self.device)
        self.fock(density, compact, precision, fock, terms)
        return _Start(density, fock, terms.cpu().numpy())

    def energies(self, terms: np.ndarray) -> np.ndarray:
        return self.e_nuc + 2.0 * terms[0] + 2.0 * terms[1] - terms[2]

    def select(self, flags, first: _Start, second: _Start) -> _Start:
        """Per molecule, ``first`` where flags[i] else ``second``."""
        torch = _torch()
        f = torch.tensor(np.asarray(flags, dtype=np.int32), device=self.device)
        out_d = torch.empty_like(first.density)
        out_f = torch.empty_like(first.fock)
        for a, b, o in ((first.density, second.density, out_d), (first.fock, second.fock, out_f)):
            _lib.check(self.lib.hfjk_select_batched(self._ctx.handle, self.count, self._nn, _p(f), _p(a), _p(b),
                                                    _p(o), self._stream()))
        terms = np.where(np.asarray(flags, dtype=bool)[None, :], first.terms, second.terms)
        return _Start(out_d, out_f, terms)

    # ------------------------------------------------------------ the loop
    def run(self, start: _Start, options: ScfOptions | None = None, *, raise_on_failure: bool = True,
            profile: dict | None = None):
        """Converge every molecule from ``start``; returns one ``BatchResult`` per molecule.

        ``profile`` (a dict) receives wall seconds per stage, measured with a device
        synchronisation after each stage (which slows the run slightly)."""
        steps = self.iterate(start, options, profile=profile)
        while True:
            try:
                next(steps)
            except StopIteration as stop:
                return _finish(stop.value, raise_on_failure)

    def iterate(self, start: _Start, options: ScfOptions | None = None, *, profile: dict | None = None):
        """The SCF loop as a generator: it yields once per iteration, after the iteration's GPU work
        is enqueued on the current stream and before the host waits for its result, and returns one
        ``BatchResult`` per molecule (unconverged ones flagged). Resume it with that stream current
        (``run_pipelined`` interleaves several batches on their own streams this way)."""
        torch = _torch()
        options = ScfOptions() if options is None else options
        clock = [time.perf_counter()]

        def mark(stage):
            if profile is not None:
                torch.cuda.synchronize(self.device)
                now = time.perf_counter()
                profile[stage] = profile.get(stage, 0.0) + now - clock[0]
                clock[0] = now
        if options.level_shift or options.damping:
            raise InputError("the batched driver has no level shift or damping")
        if options.diis_size > 16:
            raise InputError("the batched driver keeps at most 16 DIIS matrices")
        t0 = time.perf_counter()
        cnt, nm, nn, m = self.count, self.nm, self._nn, options.diis_size
        fh = torch.zeros((m, cnt, nm, nm), dtype=torch.float64, device=self.device)
        gh = torch.zeros_like(fh)
        scal = torch.empty(3 * cnt + cnt * m + 1, dtype=torch.float64, device=self.device)
        terms_dev = scal[:3 * cnt].view(3, cnt)
        row_dev = scal[3 * cnt:3 * cnt + cnt * m].view(cnt, m)
        # 2 <G_a, G_b> per molecule, on the device, one row/column refreshed per iteration by the
        # DIIS kernel, which also solves for the coefficients there.
        ov_dev = torch.zeros((cnt, m, m), dtype=torch.float64, device=self.device)
        coef_dev = torch.empty((cnt, 1, m), dtype=torch.float64, device=self.device)
        precision = options.precision
        density = start.density
        orbitals = None
        results: list = [None] * cnt
        history = [[] for _ in range(cnt)]
        previous = np.full(cnt, math.inf)
        valid: list[int] = []
        slot = 0
        for index in range(options.max_iterations):
            fock = fh[slot]
            mark("other")
            if index == 0:
                fock.copy_(start.fock)
            else:
                self.fock(density, orbitals, precision, fock, terms_dev)
            mark("fock")
            self.gradient(fock, density, gh[slot])
            mark("gradient")
            # The newest error's overlaps with every slot, for every molecule, in one launch.
            _lib.check(self.lib.hfjk_history_dots_batched(self._ctx.handle, cnt, nn, m, _p(gh), slot, _p(row_dev),
                                                          self._stream()))
            scal[-1:] = self._eig_failures
            yield index
            host = scal.cpu().numpy()
            mark("overlaps_transfer")
            if host[-1] != 0.0:
                raise BackendError("the batched eigensolver failed on a Fock matrix")
            terms = start.terms if index == 0 else host[:3 * cnt].reshape(3, cnt)
            row = 2.0 * host[3 * cnt:3 * cnt + cnt * m].reshape(cnt, m)
            energy = self.energies(terms)
            gnorm = np.sqrt(np.maximum(row[:, slot], 0.0))
            change = energy - previous
            previous = energy
            now = time.perf_counter() - t0
            for i in range(cnt):
                history[i].append((float(energy[i]), float(gnorm[i]), precision))
                if results[i] is None and precision == "fp64" and abs(change[i]) < options.energy_tolerance \
                        and gnorm[i] < options.gradient_tolerance:
                    results[i] = BatchResult(float(energy[i]), True, index + 1, now, history[i])
            valid = (valid + [slot])[-m:]
            if all(r is not None for r in results):
                break
            if precision != "fp64":
                # Batch-wide switch once most unconverged molecules are below the threshold or have
                # stalled at the reduced-precision floor (FP64 iterations never hurt the others).
                pending = [i for i in range(cnt) if results[i] is None]
                mixed = [[g for (_, g, p) in history[i] if p != "fp64"] for i in pending]
                ready = sum(1 for g in mixed if g[-1] < options.switch_at(precision)
                            or mixed_phase_stalled(g, options.stall_window, options.stall_ratio))
                if 2 * ready >= len(mixed) or len(mixed[0]) >= options.max_iterations // 2:
                    precision = "fp64"
                    if options.reset_diis_on_switch:
                        valid = [slot]
            mark("host")
            # DIIS coefficients of every molecule over the valid slots on the device, then one
            # batched combination.
            slots = (ctypes.c_int * len(valid))(*valid)
            _lib.check(self.lib.hfjk_diis_batched(self._ctx.handle, cnt, m, _p(ov_dev), _p(row_dev), slot, slots,
                                                  len(valid), _p(coef_dev), self._stream()))
            extrapolated = torch.empty((cnt, nm, nm), dtype=torch.float64, device=self.device)
            _lib.check(self.lib.hfjk_diis_combine_batched(self._ctx.handle, cnt, nn, m, _p(fh), _p(coef_dev),
                                                          _p(extrapolated), self._stream()))
            mark("extrapolate")
            density, _, orbitals = self.solve(extrapolated, mark)
            slot = (slot + 1) % m
        total = time.perf_counter() - t0
        out = []
        for i in range(cnt):
            if results[i] is None:
                last = history[i][-1]
                results[i] = BatchResult(last[0], False, len(history[i]), total, history[i])
            out.append(results[i])
        return out


def _finish(out, raise_on_failure: bool):
    if raise_on_failure and not all(r.converged for r in out):
        bad = [i for i, r in enumerate(out) if not r.converged]
        raise ConvergenceError(f"molecules {bad} did not converge", out)
    return out


def run_pipelined(groups, options: ScfOptions | None = None, *, raise_on_failure: bool = True):
    """Converge several batches at once, each on its own stream, interleaving their iterations.

    ``groups`` is a list of ``(BatchedRhf, start)`` pairs; returns one result list per group. While
    the host waits for one batch's energies and DIIS overlaps, the other batches' enqueued work
    runs: the batched eigensolver of one (latency-bound kernels that leave most CUs idle) overlaps
    the J/K build of another (MFMA-bound). Each batch iterates exactly as in ``BatchedRhf.run``.
    """
    torch = _torch()
    if not groups:
        raise InputError("at least one batch")
    device = groups[0][0].device
    origin = torch.cuda.current_stream(device)
    streams = [torch.cuda.Stream(device=batch.device) for batch, _ in groups]
    for stream in streams:
        stream.wait_stream(origin)  # the starts were computed on the caller's stream
    steps = [batch.iterate(start, options) for batch, start in groups]
    outs: list = [None] * len(groups)
    live = list(range(len(groups)))
    while live:
        for g in list(live):
            with torch.cuda.stream(streams[g]):
                try:
                    next(steps[g])
                except StopIteration as stop:
                    outs[g] = stop.value
                    live.remove(g)
    for stream in streams:
        origin.wait_stream(stream)
    return [_finish(out, raise_on_failure) for out in outs]


def balanced_groups(sizes, count: int) -> list[list[int]]:
    """Indices of ``sizes`` dealt into ``count`` groups of about equal total cost (largest first,
    each to the group with the least nao^4 so far)."""
    if count < 1:
        raise InputError("at least one group")
    order = sorted(range(len(sizes)), key=lambda i: -sizes[i])
    groups: list[list[int]] = [[] for _ in range(min(count, len(sizes)))]
    cost = [0.0] * len(groups)
    for i in order:
        g = min(range(len(groups)), key=lambda k: cost[k])
        groups[g].append(i)
        cost[g] += float(sizes[i]) ** 4
    return [sorted(g) for g in groups]


# ---------------------------------------------------------------- guesses

def atomic_factors(molecule: Molecule, atom_densities: dict, tolerance: float = 1e-12):
    """Block-diagonal superposition of atomic densities and its factor (atomic natural orbitals
    scaled by the square roots of their occupations)."""
    blocks_d, blocks_c = [], []
    for z in molecule.charges:
        d = np.asarray(atom_densities[z])
        w, v = np.linalg.eigh(d)
        keep = w > tolerance * max(w.max(), 1e-300)
        blocks_d.append(d)
        blocks_c.append(v[:, keep] * np.sqrt(w[keep]))
    n = sum(b.shape[0] for b in blocks_d)
    r = sum(c.shape[1] for c in blocks_c)
    dens, fac = np.zeros((n, n)), np.zeros((n, r))
    at = col = 0
    for d, c in zip(blocks_d, blocks_c):
        dens[at:at + d.shape[0], at:at + d.shape[0]] = d
        fac[at:at + c.shape[0], col:col + c.shape[1]] = c
        at += d.shape[0]
        col += c.shape[1]
    return dens, fac


def atomic_start(batch: BatchedRhf, atom_densities: dict, precision: str = "fp64") -> _Start:
    pairs = [atomic_factors(x.molecule, atom_densities) for x in batch.inputs]
    return batch.start_from_factors([p[0] for p in pairs], [p[1] for p in pairs], precision)


def molecular_start(batch: BatchedRhf, water_orbitals: np.ndarray, precision: str = "fp64") -> _Start:
    """Superposition of water densities: each water's block is the isolated reference water's
    occupied orbitals rotated into its frame (``learned_guess.molecular_factors``)."""
    from hafnium_amd.learned_guess import molecular_factors

    pairs = [molecular_factors(x.molecule, water_orbitals) for x in batch.inputs]
    return batch.start_from_factors([p[0] for p in pairs], [p[1] for p in pairs], precision)


def learned_start(batch: BatchedRhf, model, sad: _Start, precision: str = "fp64") -> _Start:
    """F0 = h + J[d] - K[D_base] from predicted fitted coefficients d, D_base the density of the
    given start (``sad``: atomic or molecular superposition). Per molecule the start is the aufbau
    density of F0, with its Fock build, unless its energy exceeds the energy of D_base itself; there
    the given start is kept unchanged (the fallback)."""
    torch = _torch()
    t0 = time.perf_counter()
    coeffs = model.predict_many([x.molecule for x in batch.inputs])
    # Charges against B: L^-1 (P|Q) d, with the metric's Cholesky factor kept from the fit; all on the device.
    charges = []
    for x, f, d in zip(batch.inputs, batch.factors, coeffs):
        metric = torch.as_tensor(x.metric, dtype=torch.float64).to(batch.device).contiguous()
        v = batch.dense.gemm(metric, torch.from_numpy(d).to(batch.device).view(-1, 1).contiguous())
        _lib.check(batch.lib.hfjk_dtrsm_lower(batch._ctx.handle, x.naux, 1, _p(f), _p(v), batch._stream()))
        charges.append(v.view(-1))
    j_c = batch.jk.coulomb_from_charges(torch.cat(charges), precision=precision)
    j_pad = torch.zeros_like(batch.h)
    batch._copy(batch._nao, batch._nao, batch._moffs, batch._max_mat, j_pad, batch.nm, batch._nn, j_c, False)
    # F_sad = h + 2 J_sad - K_sad, so F0 = h + J[d] - K_sad = F_sad - 2 J_sad + J[d]: no second exchange build.
    guess = torch.empty_like(batch.h)
    batch._combine(1.0, sad.fock, -2.0, sad.info["coulomb"], 1.0, j_pad, guess)
    learned = batch.start_from_fock(guess, precision)
    e_learned = batch.energies(learned.terms)
    e_base = batch.energies(sad.terms)
    flags = e_learned <= e_base
    chosen = batch.select(flags, learned, sad)
    chosen.info = {"learned": int(flags.sum()), "fallback": int((~flags).sum()),
                   "energy_learned": e_learned.tolist(), "energy_base": e_base.tolist(),
                   "seconds": time.perf_counter() - t0}
    return chosen

"""Serving cluster-solve requests on one GPU with continuous batching.

A request is one cluster (geometry, materials, sources) at one or more
wavelengths; each (request, wavelength) is one system (I + T) p_inc = p_src
with one right-hand side per source, and the response carries, per
wavelength and source, the cross-sections and the per-particle forces.

The engine admits systems first-come first-served while their HBM footprint
(matrix or matrix-free operator workspace + Krylov state + solution) fits in
a free block of the engine's arena (one device allocation of the HBM
budget), assembles them (stored systems) and joins them to the running
batch.  A system runs matrix-free (the operator recomputed in every
product, O(n) memory) when `matrix_free` is True, or, with "auto" (the
default), when it is large enough for the matrix-free kernels to beat
reading a stored matrix (FP64: lmax <= 3 and n >= 6000; mixed: lmax <= 5
and n >= 6000) or its matrices would take more than a quarter of the
budget; stored and matrix-free systems share every batched launch; every engine step advances all active systems by
a few BiCGSTAB iterations in one batched launch sequence
(`mpa_systems_advance`), retires the systems that converged (observables,
memory freed) and admits waiting ones into the freed memory.  Requests that
arrive while others are on the GPU therefore start as soon as memory frees
up, and their iterations overlap with everyone else's.  Matrices, Krylov
vectors and per-RHS scalars stay on the device for the life of a system;
each system's arithmetic does not depend on what else is in the batch, so a
request's result is the same alone and in the mix.

With precision="mixed" every system is solved as FP32 + FP64 refinement
at the iteration level (`mpa_refine_systems_advance`): complex64 BiCGSTAB
corrections on a stored complex64 matrix or, for lmax <= 5, the FP16-split
matrix-free operator, FP64 residuals on the stored complex128 matrix or the
FP64 matrix-free operator; a system leaves the batch once every source
reached the tolerance in FP64.  A system with lmax > 5 that is to run
matrix-free (matrix_free=True, or "auto" with matrices above a quarter of
the budget) has no complex64 operator: it joins a second set, solved by
FP64 BiCGSTAB on the FP64 matrix-free operator (`mpa_systems_advance`),
which every engine step advances beside the refined set.

A system whose footprint alone exceeds the budget is rejected with status
"too_large" instead of blocking the queue.

A system retires when every source has stopped.  A source that stopped
without reaching the tolerance (maxiter, stalled or out of refinement steps,
non-finite) is not a result: a refined system is then re-solved by FP64
BiCGSTAB on its FP64 operator (the stored complex128 matrix or the FP64
matrix-free operator), the same safety net as `device.solve_refined`, and a
system that still does not converge, or an FP64 system that did not, is
recorded as failed.  Only converged systems count in `solves_done`.  A
request ends "done" when all its systems converged, "partial" when some
wavelengths were too large for the budget and the rest converged, "failed"
when any system failed, and "too_large" when none fit; only "done" requests
enter the latency figures.
"""

import collections
import ctypes
import time

import numpy as np
import torch

from . import _lib as L
from . import device as D


class _System:
    __slots__ = ("req", "index", "w", "n", "S", "k", "eps_b", "pos", "mie_scat", "mie_int", "p_src", "A", "x",
                 "state", "struct", "done", "iters", "footprint", "block", "matrix_free", "refine")


class Arena:
    """First-fit allocator over one device buffer: the engine's HBM budget.

    Matrices, Krylov state and solutions of admitted systems are carved out
    of a single allocation, so admission succeeds exactly when a contiguous
    block is free (no fragmentation inside the framework's caching
    allocator), and freed blocks coalesce with their neighbours."""

    ALIGN = 256

    def __init__(self, nbytes, device="cuda"):
        self.size = int(nbytes) // self.ALIGN * self.ALIGN
        self.buf = torch.empty(self.size, dtype=torch.uint8, device=device)
        self.free_list = [(0, self.size)]  # sorted (offset, size)
        self.live = {}

    def alloc(self, nbytes):
        need = (int(nbytes) + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        for k, (off, size) in enumerate(self.free_list):
            if size >= need:
                if size == need:
                    del self.free_list[k]
                else:
                    self.free_list[k] = (off + need, size - need)
                self.live[off] = need
                return off
        return None

    def free(self, off):
        size = self.live.pop(off)
        fl = self.free_list
        k = 0
        while k < len(fl) and fl[k][0] < off:
            k += 1
        fl.insert(k, (off, size))
        if k + 1 < len(fl) and fl[k][0] + fl[k][1] == fl[k + 1][0]:  # merge with the next block
            fl[k] = (fl[k][0], fl[k][1] + fl[k + 1][1])
            del fl[k + 1]
        if k > 0 and fl[k - 1][0] + fl[k - 1][1] == fl[k][0]:  # and with the previous one
            fl[k - 1] = (fl[k - 1][0], fl[k - 1][1] + fl[k][1])
            del fl[k]

    def view(self, off, nbytes, dtype, shape):
        return self.buf[off:off + nbytes].view(dtype).reshape(shape)

    def check_inside(self, block, end):
        """Raise if [block, end) runs past the block allocated at `block`."""
        if end > block + self.live[block]:
            raise RuntimeError(f"arena block at {block} overrun: {end - block} > {self.live[block]} bytes")

    @property
    def used(self):
        return sum(self.live.values())


class Request:
    def __init__(self, rid, workloads, arrival):
        self.id = rid
        self.workloads = list(workloads)
        self.arrival = arrival
        self.admitted = None
        self.finished = None
        self.status = "queued"
        self.results = [None] * len(self.workloads)
        self.pending = len(self.workloads)
        self.rejected = 0  # systems too large for the budget
        self.failed = 0    # systems that did not converge (after the FP64 safety net)

    def _finish(self, when):
        self.finished = when
        if self.rejected == len(self.workloads):
            self.status = "too_large"
        elif self.failed:
            self.status = "failed"
        elif self.rejected:
            self.status = "partial"
        else:
            self.status = "done"

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def _aligned(nbytes):
    return (int(nbytes) + Arena.ALIGN - 1) // Arena.ALIGN * Arena.ALIGN


def system_footprint(w, matrix_free=False, precision="fp64"):
    """Arena bytes of one system: matrices (c128, plus c64 for "mixed") or the
    matrix-free operator's workspace, solver state and solution."""
    n, S = w.n, len(w.sources)
    lib = L.load()
    if precision == "mixed":
        state = lib.mpa_refine_state_bytes(n, S)
        mats = _aligned(16 * n * n) + _aligned(8 * n * n)  # each matrix at an aligned offset
    else:
        state = lib.mpa_system_state_bytes(n, S, L.MPA_C128)
        mats = _aligned(16 * n * n)
    op = _aligned(D.ClusterOperator.workspace_bytes(w.N, w.lmax, S)) if matrix_free else mats
    return op + _aligned(state) + _aligned(16 * S * n)


class ServingEngine:
    def __init__(self, hbm_budget_bytes=200e9, iterations_per_step=2, maxiter=2000, clock=time.perf_counter,
                 matrix_free="auto", precision="fp64", inner_tol=1e-5, max_refine=8):
        D.require_gpu()
        if matrix_free not in (True, False, "auto"):
            raise ValueError("matrix_free must be True, False or 'auto'")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.matrix_free = matrix_free
        self.precision = precision
        self.inner_tol = float(inner_tol)
        self.max_refine = int(max_refine)
        self.arena = Arena(hbm_budget_bytes)
        self.budget = float(self.arena.size)
        self.k = int(iterations_per_step)
        self.maxiter = int(maxiter)
        self.clock = clock
        self.queue = collections.deque()  # systems waiting for memory, FIFO
        self.active = []
        self.finished = []
        self.rejected = []
        self.steps = 0
        self.solves_done = 0     # converged systems
        self.failed_systems = 0  # systems that stopped without converging
        self.fp64_fallbacks = 0  # refined systems finished by the FP64 safety net
        self._scratch = torch.empty(0, dtype=torch.uint8, device="cuda")
        self.lib = L.load()

    # ---- intake --------------------------------------------------------------------

    def _route(self, w):
        """(matrix_free, refined) for one system."""
        if self.precision == "mixed":
            if self.matrix_free is False:
                return False, True
            if w.lmax > 5:
                # no complex64 matrix-free operator above lmax 5: stored matrices, or FP64 matrix-free
                mf = self.matrix_free is True or 24.0 * w.n * w.n > 0.25 * self.budget
                return mf, not mf
            return self.matrix_free is True or w.n >= 6000 or 24.0 * w.n * w.n > 0.25 * self.budget, True
        if self.matrix_free != "auto":
            return bool(self.matrix_free), False
        # per-system matrix-free launches cost more than one batched product over many small matrices
        return (w.lmax <= 3 and w.n >= 6000) or 16.0 * w.n * w.n > 0.25 * self.budget, False

    def submit(self, request):
        for i, w in enumerate(request.workloads):
            s = _System()
            s.req, s.index, s.w = request, i, w
            s.matrix_free, s.refine = self._route(w)
            s.footprint = system_footprint(w, s.matrix_free, "mixed" if s.refine else "fp64")
            if s.footprint > self.budget:
                request.results[i] = {"status": "too_large"}
                request.rejected += 1
                request.pending -= 1
                continue
            self.queue.append(s)
        if request.pending == 0:
            request._finish(self.clock())
            self.rejected.append(request)

    def _inputs(self, systems):
        """Inputs of every system in `systems`: one upload of the per-particle
        data (positions, radius, k, eps, medium), the Mie coefficients of all
        their particles in one device call, and each system's plane-wave p_src
        on the device; each system gets views of the device buffers."""
        lmax = max(s.w.lmax for s in systems)
        rad = np.concatenate([np.broadcast_to(np.asarray(s.w.radius, dtype=float), (s.w.N,)) for s in systems])
        eps = np.concatenate([np.broadcast_to(np.asarray(s.w.eps, dtype=complex), (s.w.N,)) for s in systems])
        kk = np.concatenate([np.full(s.w.N, s.w.k) for s in systems])
        epsb = np.concatenate([np.full(s.w.N, complex(s.w.medium_eps)) for s in systems])
        cond = np.concatenate([np.zeros(s.w.N, dtype=bool) if s.w.conducting is None else
                               np.broadcast_to(np.asarray(s.w.conducting, dtype=bool), (s.w.N,)) for s in systems])
        ms_all, mi_all = D.mie_coefficients(rad, eps, np.ones(rad.shape[0]), kk, epsb, 1.0, lmax, cond)
        pos_d = D.to_device(np.concatenate([np.asarray(s.w.position, dtype=float).reshape(-1) for s in systems]),
                            torch.float64)
        o = 0
        for s in systems:
            w = s.w
            N = w.N
            s.pos = pos_d[3 * o:3 * (o + N)].view(N, 3)
            s.mie_scat = ms_all[o:o + N, :, :w.lmax].contiguous()
            s.mie_int = mi_all[o:o + N, :, :w.lmax].contiguous()
            s.p_src = D.plane_wave_structure(list(w.sources), s.pos, w.k, w.lmax)
            o += N

    def _prepare(self, s):
        w = s.w
        s.n, s.S = w.n, len(w.sources)
        s.k = w.k
        s.eps_b = float(np.real(w.medium_eps))
        off = s.block
        if s.refine:
            self._prepare_mixed(s, off)
            return
        nst = _aligned(self.lib.mpa_system_state_bytes(s.n, s.S, L.MPA_C128))
        if s.matrix_free:
            nw = D.ClusterOperator.workspace_bytes(w.N, w.lmax, s.S)
            ws = self.arena.view(off, nw, torch.uint8, (nw,))
            s.A = D.ClusterOperator(s.pos, s.mie_scat, s.k, max_rhs=s.S, workspace=ws)
            nA = _aligned(nw)
        else:
            nA = _aligned(16 * s.n * s.n)
            A = self.arena.view(off, 16 * s.n * s.n, D.C128, (s.n, s.n))
            s.A = D.assemble(s.pos, s.mie_scat, s.k, dtype=D.C128, add_identity=True, out=A)
        s.state = self.arena.view(off + nA, nst, torch.uint8, (nst,))
        s.x = self.arena.view(off + nA + nst, 16 * s.S * s.n, D.C128, (s.S, s.n))
        self.arena.check_inside(off, off + nA + nst + 16 * s.S * s.n)
        s.struct = D._sys_struct(s.A, s.p_src.reshape(s.S, s.n), s.x, s.state, w.tol, self.maxiter, False)
        s.done = False
        s.iters = 0

    def _prepare_mixed(self, s, off):
        w = s.w
        nst = _aligned(self.lib.mpa_refine_state_bytes(s.n, s.S))
        rs = L.MpaRefineSystem()
        # the stored matrices hold I + T; the matrix-free operator stands for T
        rs.n, rs.nrhs, rs.add_identity = s.n, s.S, 1 if s.matrix_free else 0
        if s.matrix_free:
            nw = D.ClusterOperator.workspace_bytes(w.N, w.lmax, s.S)
            ws = self.arena.view(off, nw, torch.uint8, (nw,))
            s.A = D.ClusterOperator(s.pos, s.mie_scat, s.k, max_rhs=s.S, workspace=ws)
            nA = _aligned(nw)
            rs.A32 = rs.A64 = None
            rs.op = ctypes.cast(ctypes.pointer(s.A.struct), ctypes.c_void_p)
        else:
            n64 = _aligned(16 * s.n * s.n)
            A64 = self.arena.view(off, 16 * s.n * s.n, D.C128, (s.n, s.n))
            A32 = self.arena.view(off + n64, 8 * s.n * s.n, D.C64, (s.n, s.n))
            D.assemble_dual(s.pos, s.mie_scat, s.k, add_identity=True, out64=A64, out32=A32)
            s.A = (A64, A32)
            nA = n64 + _aligned(8 * s.n * s.n)
            rs.A32, rs.A64 = A32.data_ptr(), A64.data_ptr()
        s.state = self.arena.view(off + nA, nst, torch.uint8, (nst,))
        s.x = self.arena.view(off + nA + nst, 16 * s.S * s.n, D.C128, (s.S, s.n))
        self.arena.check_inside(off, off + nA + nst + 16 * s.S * s.n)
        rs.b, rs.x, rs.state = s.p_src.data_ptr(), s.x.data_ptr(), s.state.data_ptr()
        rs.tol, rs.inner_tol, rs.maxiter, rs.max_refine = float(w.tol), self.inner_tol, self.maxiter, self.max_refine
        s.struct = rs
        s.done = False
        s.iters = 0

    @property
    def used(self):
        return self.arena.used

    def _admit(self):
        new = []
        while self.queue:
            off = self.arena.alloc(self.queue[0].footprint)
            if off is None:
                break  # FIFO: the head waits for memory, later systems wait behind it
            s = self.queue.popleft()
            s.block = off
            new.append(s)
        if new:
            self._inputs(new)
            for s in new:
                self._prepare(s)
                if s.req.admitted is None:
                    s.req.admitted = self.clock()
                if s.req.status == "queued":  # status only moves forward
                    s.req.status = "running"
            ref = [s for s in new if s.refine]
            f64 = [s for s in new if not s.refine]
            if ref:
                arr = (L.MpaRefineSystem * len(ref))(*[s.struct for s in ref])
                scratch = self._scratch_for(arr, len(ref), True)
                L.check(self.lib.mpa_refine_systems_start(len(ref), arr, D._ptr(scratch), D._stream()),
                        "refine_systems_start")
            if f64:
                arr = (L.MpaSystem * len(f64))(*[s.struct for s in f64])
                scratch = self._scratch_for(arr, len(f64), False)
                L.check(self.lib.mpa_systems_start(len(f64), arr, L.MPA_C128, D._ptr(scratch), D._stream()),
                        "systems_start")
            self.active.extend(new)
        return len(new)

    def _scratch_for(self, arr, count, refine):
        """One scratch buffer serves both sets: their launches are ordered on one stream."""
        if refine:
            need = self.lib.mpa_refine_scratch_bytes(count, arr)
        else:
            need = self.lib.mpa_batch_scratch_bytes(count, arr, L.MPA_C128)
        if self._scratch.numel() < need:
            self._scratch = torch.empty(int(need * 1.5) + 4096, dtype=torch.uint8, device="cuda")
        return self._scratch

    # ---- one engine step ---------------------------------------------------------------

    def step(self):
        """Admit what fits, advance every active system, retire the converged."""
        self._admit()
        if not self.active:
            return False
        keep = []
        for refine in (True, False):
            group = [s for s in self.active if s.refine == refine]
            if not group:
                continue
            count = len(group)
            total = sum(s.S for s in group)
            done = (ctypes.c_int * total)()
            its = (ctypes.c_int * total)()
            res = (ctypes.c_double * total)()
            stt = (ctypes.c_int * total)()
            if refine:
                arr = (L.MpaRefineSystem * count)(*[s.struct for s in group])
                outer = (ctypes.c_int * total)()
                scratch = self._scratch_for(arr, count, True)
                L.check(self.lib.mpa_refine_systems_advance(count, arr, D._ptr(scratch), self.k, done, outer, its,
                                                            res, stt, D._stream()), "refine_systems_advance")
            else:
                arr = (L.MpaSystem * count)(*[s.struct for s in group])
                scratch = self._scratch_for(arr, count, False)
                L.check(self.lib.mpa_systems_advance(count, arr, L.MPA_C128, D._ptr(scratch), self.k, done, its, res,
                                                     stt, D._stream()), "systems_advance")
            o = 0
            for s in group:
                if all(done[o + j] for j in range(s.S)):
                    self._retire(s, list(its[o:o + s.S]), list(res[o:o + s.S]), list(stt[o:o + s.S]))
                else:
                    keep.append(s)
                o += s.S
        self.steps += 1
        kept = {id(s) for s in keep}
        self.active = [s for s in self.active if id(s) in kept]
        return True

    def _fp64_safety_net(self, s, iters):
        """Re-solve a refined system whose refinement stopped short with FP64
        BiCGSTAB on its FP64 operator (stored I + T or the matrix-free T)."""
        b = s.p_src.reshape(s.S, s.n)
        if s.matrix_free:
            x, rep = D.bicgstab(s.A, b, tol=s.w.tol, maxiter=max(self.maxiter, 10 * s.n), add_identity=True,
                                check=False)
        else:
            x, rep = D.bicgstab(s.A[0], b, tol=s.w.tol, maxiter=max(self.maxiter, 10 * s.n), check=False)
        s.x.copy_(x.reshape(s.x.shape))
        self.fp64_fallbacks += 1
        return [a + b for a, b in zip(iters, rep.iterations)], list(rep.residual), list(rep.status)

    def _retire(self, s, iters, resid, status):
        fallback = False
        if any(v != L.MPA_OK for v in status) and s.refine:
            iters, resid, status = self._fp64_safety_net(s, iters)
            fallback = True
        ok = all(v == L.MPA_OK for v in status)
        p_inc = s.x.reshape(s.S, s.w.N, 2, s.w.rmax)
        obs = D.observables(p_inc, s.p_src, s.mie_scat, s.mie_int, s.k, s.eps_b, want_int=False,
                            want_torque=False)
        s.req.results[s.index] = {
            "wavelength": s.w.wavelength,
            "cross_sections": obs["C_total"].cpu().numpy(),  # [S, 3]
            "force": obs["F"].cpu().numpy(),                  # [S, N, 3]
            "iterations": iters, "residual": resid, "status": status, "matrix_free": s.matrix_free,
            "refined": s.refine, "fp64_fallback": fallback, "converged": ok, "done_at": self.clock(),
        }
        if ok:
            self.solves_done += 1
        else:
            self.failed_systems += 1
            s.req.failed += 1
        s.A = s.x = s.state = s.pos = s.mie_scat = s.mie_int = s.p_src = None
        self.arena.free(s.block)
        s.req.pending -= 1
        if s.req.pending == 0:
            s.req._finish(self.clock())
            self.finished.append(s.req)

    # ---- drivers ---------------------------------------------------------------------------

    def drain(self):
        while self.queue or self.active:
            if not self.step() and self.queue:
                raise RuntimeError("queue is stuck: the head system does not fit the budget")

    def run_stream(self, arrivals, max_wall=None, progress=None, progress_every=10.0):
        """Serve (arrival_offset_s, workloads) pairs in real time; returns the requests.

        progress(elapsed_s, engine) is called every `progress_every` seconds."""
        arrivals = list(arrivals)
        t0 = self.clock()
        next_report = progress_every
        reqs = []
        i = 0
        while i < len(arrivals) or self.queue or self.active:
            now = self.clock() - t0
            if progress is not None and now >= next_report:
                progress(now, self)
                next_report += progress_every
            while i < len(arrivals) and arrivals[i][0] <= now:
                r = Request(i, arrivals[i][1], t0 + arrivals[i][0])
                reqs.append(r)
                self.submit(r)
                i += 1
            if max_wall is not None and now > max_wall:
                break
            if not self.step() and i < len(arrivals):
                wait = arrivals[i][0] - (self.clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
        return reqs


def latency_summary(requests):
    lat = np.array([r.latency for r in requests if r.latency is not None and r.status == "done"])
    if lat.size == 0:
        return {"completed": 0}
    return {"completed": int(lat.size), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3)}

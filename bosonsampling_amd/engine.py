"""Device-resident MPO and the single-GPU simulation driver.

The state lives in HBM as torch tensors used purely as buffers:
  Gamma  [M, chi, chi] complex128, Lambda [M-1, chi] float64,
  charge [M+1, chi, 2] int32 (lexicographically sorted, padded with d),
with a host mirror of the charges for planning.  Every update runs through
libbsmpo (include/bsmpo.h): host C++ planning from the charges, then the
gfx950 kernels on the current torch stream.

Reference correspondence:
  DeviceMPO.update        RankWorker.RankProcess + FullWorker.Check
                          (src/beam_splitter_parallel/Rank.py:32-206, Full.py:113-137)
  simulate                FullWorker.LayerUpdate / Simulate (Full.py:149-202)
  observables             summary.Probability / EntanglementEntropy (src/summary.py:3-36)
  initialize              MPO.Initialize + Canonicalize (src/MPO.py:21-156)
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _capi
from ._capi import check, lib, ptr

C128, C64, BF16X3 = _capi.C128, _capi.C64, _capi.BF16X3


def _require_gpu():
    if not torch.cuda.is_available():
        raise _capi.BsmpoError("no GPU visible: the bsmpo engine runs on HIP devices only")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Workspace:
    """Grow-only device workspace and pinned plan staging buffer."""

    def __init__(self, device):
        self.device = device
        self.work = None
        self.plan = None

    def device_bytes(self, n):
        if self.work is None or self.work.numel() < n:
            self.work = torch.empty(max(n, 1 << 20) * 5 // 4, dtype=torch.uint8, device=self.device)
        return self.work

    def host_plan(self, n):
        if self.plan is None or self.plan.numel() < n:
            self.plan = torch.empty(max(n, 1 << 16) * 5 // 4, dtype=torch.uint8, pin_memory=True)
        return self.plan


class Plan:
    """Host plan of one update (opaque blob built by bsmpo_plan_build)."""

    def __init__(self, d, chi, CL, CC, CR, ws: Workspace = None, tau=None):
        CL = np.ascontiguousarray(CL, np.int32)
        CC = np.ascontiguousarray(CC, np.int32)
        CR = np.ascontiguousarray(CR, np.int32)
        t0, t1 = (-1, -1) if tau is None else (int(tau[0]), int(tau[1]))
        pb, wb = C.c_size_t(), C.c_size_t()
        check(lib().bsmpo_plan_query(d, chi, ptr(CL), ptr(CC), ptr(CR), t0, t1,
                                     C.byref(pb), C.byref(wb)), "plan_query")
        self.plan_bytes, self.work_bytes = pb.value, wb.value
        if ws is not None:
            self.buf = ws.host_plan(self.plan_bytes)
        else:
            self.buf = np.empty(self.plan_bytes, np.uint8)
        check(lib().bsmpo_plan_build(d, chi, ptr(CL), ptr(CC), ptr(CR), t0, t1, ptr(self.buf),
                                     self.plan_bytes), "plan_build")
        self.d, self.chi = d, chi

    @property
    def address(self):
        return ptr(self.buf)

    def info(self):
        inf = _capi.PlanInfo()
        check(lib().bsmpo_plan_info_get(C.c_void_p(self.address), C.byref(inf)), "plan_info")
        return inf

    def sectors(self):
        """[(tau0, tau1, L, R)], packed C offsets, rows pool, cols pool."""
        inf = self.info()
        out = np.zeros(4 * max(inf.n_sectors, 1), np.int32)
        coff = np.zeros(max(inf.n_sectors, 1), np.int64)
        check(lib().bsmpo_plan_sectors(C.c_void_p(self.address), ptr(out), ptr(coff), None, None),
              "plan_sectors")
        sec = out[:4 * inf.n_sectors].reshape(-1, 4)
        nr = int(sec[:, 2].sum())
        nc = int(sec[:, 3].sum())
        rows = np.zeros(max(nr, 1), np.int32)
        cols = np.zeros(max(nc, 1), np.int32)
        check(lib().bsmpo_plan_sectors(C.c_void_p(self.address), None, None, ptr(rows), ptr(cols)),
              "plan_sectors")
        return sec, coff[:inf.n_sectors], rows[:nr], cols[:nc]


def bs_tables(d: int, Bs: np.ndarray, device) -> torch.Tensor:
    """U3 tables [nb, d, d, d] complex128 on the device for 2x2 matrices Bs."""
    Bs = np.ascontiguousarray(np.asarray(Bs, np.complex128).reshape(-1, 2, 2))
    nb = Bs.shape[0]
    dB = torch.from_numpy(Bs.view(np.float64).copy()).to(device)
    U3 = torch.empty((nb, d, d, d), dtype=torch.complex128, device=device)
    check(lib().bsmpo_bs_table(d, nb, C.c_void_p(dB.data_ptr()), C.c_void_p(U3.data_ptr()),
                               _stream()), "bs_table")
    return U3


class DeviceMPO:
    """MPO state in HBM (see module docstring for the layout)."""

    def __init__(self, M: int, d: int, chi: int, device="cuda"):
        _require_gpu()
        self.M, self.d, self.chi = M, d, chi
        self.device = torch.device(device)
        self.Gamma = torch.zeros((M, chi, chi), dtype=torch.complex128, device=self.device)
        self.Lambda = torch.zeros((M - 1, chi), dtype=torch.float64, device=self.device)
        self.charge = torch.full((M + 1, chi, 2), d, dtype=torch.int32, device=self.device)
        self.edge = torch.ones(chi, dtype=torch.float64, device=self.device)
        self.charge_host = np.full((M + 1, chi, 2), d, np.int32)
        # kept count of the last update of each pair (mode, mode + 1): one slot per
        # mode, so the updates of one grouped run never write the same slot
        self.num_lambda = torch.zeros(max(M - 1, 1), dtype=torch.int32, device=self.device)
        self.ws = Workspace(self.device)

    # ------------------------------------------------------------ host I/O
    @classmethod
    def from_host(cls, Gamma, Lambda, charge, d, device="cuda"):
        M, chi = Gamma.shape[0], Gamma.shape[1]
        st = cls(M, d, chi, device)
        st.Gamma.copy_(torch.from_numpy(np.ascontiguousarray(Gamma, np.complex128)))
        st.Lambda.copy_(torch.from_numpy(np.ascontiguousarray(Lambda, np.float64)))
        st.charge.copy_(torch.from_numpy(np.ascontiguousarray(charge, np.int32)))
        st.charge_host[:] = charge
        return st

    def to_host(self):
        torch.cuda.synchronize(self.device)
        return (self.Gamma.cpu().numpy(), self.Lambda.cpu().numpy(), self.charge.cpu().numpy())

    def clone(self):
        st = DeviceMPO(self.M, self.d, self.chi, self.device)
        st.Gamma.copy_(self.Gamma)
        st.Lambda.copy_(self.Lambda)
        st.charge.copy_(self.charge)
        st.charge_host[:] = self.charge_host
        return st

    # -------------------------------------------------------------- update
    def plan(self, mode: int) -> Plan:
        ch = self.charge_host
        return Plan(self.d, self.chi, ch[mode], ch[mode + 1], ch[mode + 2], self.ws)

    def update(self, mode: int, U3: torch.Tensor, precision: int = C128, plan: Plan = None,
               sync_charges: bool = True):
        """Apply one beam splitter (U3 table on device) to sites (mode, mode+1)."""
        M = self.M
        if not 0 <= mode <= M - 2:
            raise ValueError(f"mode {mode} out of range for {M} sites")
        if U3.dtype != torch.complex128 or tuple(U3.shape) != (self.d,) * 3 or not U3.is_cuda:
            raise ValueError("U3 must be a device complex128 [d, d, d] table")
        plan = plan or self.plan(mode)
        work = self.ws.device_bytes(plan.work_bytes)
        LR = self.edge if mode == M - 2 else self.Lambda[mode + 1]
        check(lib().bsmpo_site_update(
            self.d, self.chi, C.c_void_p(plan.address),
            C.c_void_p(self.Gamma[mode].data_ptr()), C.c_void_p(self.Gamma[mode + 1].data_ptr()),
            C.c_void_p(LR.data_ptr()), C.c_void_p(U3.data_ptr()),
            C.c_void_p(self.charge[mode + 1].data_ptr()), C.c_void_p(self.Lambda[mode].data_ptr()),
            C.c_void_p(self.num_lambda[mode].data_ptr()), 1 if mode == M - 2 else 0, int(precision),
            C.c_void_p(work.data_ptr()), work.numel(), _stream()), "site_update")
        if sync_charges:
            self.charge_host[mode + 1] = self.charge[mode + 1].cpu().numpy()
        return plan

    # --------------------------------------------------------- observables
    def observables(self):
        """(trace, marginals[M, d]) over the ket = bra diagonal."""
        M, d, chi = self.M, self.d, self.chi
        R = torch.empty((M + 1, chi), dtype=torch.complex128, device=self.device)
        L = torch.empty((M + 1, chi), dtype=torch.complex128, device=self.device)
        marg = torch.empty((M, d), dtype=torch.float64, device=self.device)
        tr = torch.empty(1, dtype=torch.float64, device=self.device)
        check(lib().bsmpo_observables(M, d, chi, C.c_void_p(self.Gamma.data_ptr()),
                                      C.c_void_p(self.Lambda.data_ptr()),
                                      C.c_void_p(self.charge.data_ptr()), C.c_void_p(R.data_ptr()),
                                      C.c_void_p(L.data_ptr()), C.c_void_p(marg.data_ptr()),
                                      C.c_void_p(tr.data_ptr()), _stream()), "observables")
        return float(tr.item()), marg.cpu().numpy()

    def entanglement_entropy(self):
        """Entropy of every inner bond on the device (src/summary.py:3-8):
        S = -sum p log2 p, p = Lambda^2 / sum Lambda^2; returns M-1 floats."""
        S = torch.empty(self.M - 1, dtype=torch.float64, device=self.device)
        check(lib().bsmpo_entanglement_entropy(self.M, self.chi, C.c_void_p(self.Lambda.data_ptr()),
                                               C.c_void_p(S.data_ptr()), _stream()),
              "entanglement_entropy")
        return S.cpu().numpy()


# ---------------------------------------------------------- grouped updates

class GroupRunner:
    """Runs many independent updates with one launch per stage (libbsmpo's
    bsmpo_group_*): host plans, one pinned host image, one device workspace.

    The updates of one call must touch disjoint sites (e.g. all beam
    splitters of a brick-wall layer, across any number of DeviceMPO
    instances).  ``max_bytes`` bounds the device workspace of one group; a
    call with more work is split into several groups.  The groups of one call
    are independent, so with ``streams`` > 1 they go round the current stream
    and streams - 1 more (each with its own workspace): one group's
    contraction then fills the CUs that another's long eigensolver chains
    leave idle.  The call returns with the current stream joined to all.
    """

    def __init__(self, device="cuda", max_bytes: int = 24 << 30, use_graph: bool = False,
                 streams: int = 2, first_frac: float = 0.25):
        self.device = torch.device(device)
        self.max_bytes = int(max_bytes)
        self.first_frac = float(first_frac)  # workspace share of a call's first group
        self.use_graph = bool(use_graph)
        if streams not in (1, 2, 3, 4):
            raise ValueError("GroupRunner: streams must be in 1..4")
        self.streams = 1 if use_graph else int(streams)
        self.aux = []                  # launching streams besides the current one
        self.works = [None] * self.streams  # device workspace per launching stream
        self.hosts = [None, None]      # pinned host images, used alternately
        self.events = [None, None]     # consumption of each image by the device
        self.graphs = [None, None]     # graph that uploads each image (use_graph)
        self.side = None               # capture / replay stream (use_graph)
        self.turn = 0
        self.last_info = {}

    def _work(self, n, slot=0, stream=None):
        w = self.works[slot]
        if w is None or w.numel() < n:
            self.works[slot] = None
            # 25% headroom: a later, slightly larger group reuses the buffer
            # instead of allocating (and mapping) a new one mid-run
            w = torch.empty(n * 5 // 4, dtype=torch.uint8, device=self.device)
            if stream is not None:
                w.record_stream(stream)  # freed only after that stream's work on it
            self.works[slot] = w
        return w

    @property
    def work(self):
        return self.works[0]

    def _host(self, n):
        i = self.turn
        self.turn ^= 1
        if self.events[i] is not None:
            self.events[i].synchronize()   # the image's last upload has been consumed
        if self.graphs[i] is not None:
            check(lib().bsmpo_graph_destroy(self.graphs[i]), "graph_destroy")
            self.graphs[i] = None
        if self.hosts[i] is None or self.hosts[i].numel() < n:
            self.hosts[i] = None
            self.hosts[i] = torch.empty(n * 5 // 4, dtype=torch.uint8, pin_memory=True)
        return i, self.hosts[i]

    def run(self, updates, precision: int = C128):
        """updates: list of (DeviceMPO, mode, U3 device table).  Plans are
        built on the host group by group, so that planning the next group
        overlaps the device's work on the previous one."""
        if not updates:
            return
        main = torch.cuda.current_stream(self.device)
        while len(self.aux) < self.streams - 1:
            self.aux.append(torch.cuda.Stream(device=self.device))
        forked = [False] * len(self.aux)
        n_groups = 0

        def launch(grp, g):
            n = len(grp)
            arr = (C.c_void_p * n)(*[p.address for _, p in grp])
            hb, wb = C.c_size_t(), C.c_size_t()
            check(lib().bsmpo_group_query(n, arr, C.byref(hb), C.byref(wb)), "group_query")
            slot_s = g % self.streams
            side = slot_s > 0
            ts = self.aux[slot_s - 1] if side else main
            if side and not forked[slot_s - 1]:
                ts.wait_stream(main)  # the call's inputs, written on the current stream
                forked[slot_s - 1] = True
            stream = C.c_void_p(ts.cuda_stream)
            work = self._work(wb.value, slot_s, ts if side else None)
            slot, host = self._host(hb.value)
            io = (_capi.UpdateIO * n)()
            for j, (k, _) in enumerate(grp):
                st, mode, U3 = updates[k]
                LR = st.edge if mode == st.M - 2 else st.Lambda[mode + 1]
                io[j] = _capi.UpdateIO(st.Gamma[mode].data_ptr(), st.Gamma[mode + 1].data_ptr(),
                                       LR.data_ptr(), U3.data_ptr(),
                                       st.charge[mode + 1].data_ptr(), st.Lambda[mode].data_ptr(),
                                       st.num_lambda[mode].data_ptr(), 1 if mode == st.M - 2 else 0,
                                       0)
            check(lib().bsmpo_group_build(n, arr, io, C.c_void_p(work.data_ptr()),
                                          C.c_void_p(host.data_ptr()), hb.value), "group_build")
            if self.use_graph:
                # the group's upload + launch sequence as one HIP graph, replayed once
                if self.side is None:
                    self.side = torch.cuda.Stream(device=self.device)
                side = C.c_void_p(self.side.cuda_stream)
                self.side.wait_stream(torch.cuda.current_stream())
                ex = C.c_void_p()
                check(lib().bsmpo_group_graph_create(C.c_void_p(host.data_ptr()), int(precision),
                                                     C.c_void_p(work.data_ptr()), wb.value, side,
                                                     C.byref(ex)), "group_graph_create")
                self.graphs[slot] = ex
                check(lib().bsmpo_graph_launch(ex, side), "graph_launch")
                torch.cuda.current_stream().wait_stream(self.side)
            else:
                check(lib().bsmpo_group_run(C.c_void_p(host.data_ptr()), int(precision),
                                            C.c_void_p(work.data_ptr()), wb.value, stream),
                      "group_run")
            ev = torch.cuda.Event()
            ev.record(ts)
            self.events[slot] = ev

        cur, cur_bytes = [], 0
        for k, (st, mode, U3) in enumerate(updates):
            ch = st.charge_host
            p = Plan(st.d, st.chi, ch[mode], ch[mode + 1], ch[mode + 2])
            # the first group of a call is planned while the device idles (the
            # previous call ended on the charge read-back), so it is kept small
            cap = self.max_bytes * (self.first_frac if n_groups == 0 else 1.0)
            if cur and cur_bytes + p.work_bytes > cap:
                launch(cur, n_groups)  # asynchronous: the next plans are built meanwhile
                n_groups += 1
                cur, cur_bytes = [], 0
            cur.append((k, p))
            cur_bytes += p.work_bytes
        if cur:
            launch(cur, n_groups)
            n_groups += 1
        for a, f in zip(self.aux, forked):
            if f:
                main.wait_stream(a)
        self.last_info = {"groups": n_groups, "updates": len(updates)}
        # host mirror of the new centre charges: one device-to-host copy
        rows = [st.charge[mode + 1] for st, mode, _ in updates]
        new = torch.stack(rows).cpu().numpy()
        for (st, mode, _), ch in zip(updates, new):
            st.charge_host[mode + 1] = ch


class LayerGraphRunner:
    """Brick-wall layers of several instances, each layer planned on the
    device and replayed from captured HIP graphs (libbsmpo's bsmpo_layer_*).

    One layer program holds the two update sets of a brick-wall mesh (every
    even pair, every odd pair, across all instances) and is captured once;
    replay r applies layer r: set r % 2, BS tables from layer r's matrices
    (every layer's matrices are uploaded once), plans built on the device
    from the charges in HBM, nothing read back between layers.  Given the
    input modes, a replay skips the pairs beyond the light cone
    (Full.py:152-153), so any layers of a mesh can run this way.
    ``run(layers)`` takes consecutive layers and reads the program's state
    once, at its end.  A replay whose plan does not fit the capacities
    changes nothing and stops the later ones; ``run`` then re-creates the
    program with what the device reported it needed and runs the rest again
    (``fallbacks`` counts that).

    Capacities per group come from host plans of the first layer's updates
    (bsmpo_layer_needs) times ``headroom``.  Each update set is split into
    ``n_groups`` contiguous groups that run as ``n_streams`` chains.
    """

    def __init__(self, states, meshes, precision: int = C128, n_groups: int = 8,
                 n_streams: int = 2, headroom: float = 1.5, input_modes=None):
        self.states, self.meshes = list(states), list(meshes)
        # light cone (Full.py:152-153): the pair at `mode` first moves at layer
        # mode - (last occupied input); None: every pair from layer 0
        self.max_input = None if input_modes is None else max(int(m) for m in input_modes)
        st0 = self.states[0]
        self.M, self.d, self.chi, self.device = st0.M, st0.d, st0.chi, st0.device
        self.precision = int(precision)
        self.n_groups, self.n_streams, self.headroom = int(n_groups), int(n_streams), float(headroom)
        self.prog = None
        self.caps = None
        self.work = None
        self.stream = torch.cuda.Stream(device=self.device)
        self.fallbacks = 0
        self.updates = []
        for p in (0, 1):
            modes = list(range(p, self.M - 1, 2))
            self.updates.append([(k, st, i, mode) for k, st in enumerate(self.states)
                                 for i, mode in enumerate(modes)])
        mu = max(len(u) for u in self.updates)
        B = np.zeros((self.M, mu, 2, 2), np.complex128)
        for layer in range(self.M):
            for j, (k, _, i, _) in enumerate(self.updates[layer % 2]):
                B[layer, j] = self.meshes[k][layer, i]
        self.B = torch.from_numpy(B.view(np.float64).copy()).to(self.device)
        self.records = [self._records(p) for p in (0, 1)]

    # ----------------------------------------------------------- program
    def _records(self, p):
        recs = (_capi.LayerUpdate * len(self.updates[p]))()
        for j, (_, st, _, mode) in enumerate(self.updates[p]):
            edge = mode == st.M - 2
            LR = st.edge if edge else st.Lambda[mode + 1]
            first = 0 if self.max_input is None else max(0, mode - self.max_input)
            recs[j] = _capi.LayerUpdate(st.charge[mode].data_ptr(), st.charge[mode + 1].data_ptr(),
                                        st.charge[mode + 2].data_ptr(), st.Gamma[mode].data_ptr(),
                                        st.Gamma[mode + 1].data_ptr(), LR.data_ptr(),
                                        st.charge[mode + 1].data_ptr(), st.Lambda[mode].data_ptr(),
                                        st.num_lambda[mode].data_ptr(), 1 if edge else 0, first)
        return recs

    def _spec(self):
        spec = _capi.LayerSpec()
        spec.d, spec.chi, spec.precision = self.d, self.chi, self.precision
        spec.n_sets = 2
        for p in (0, 1):
            spec.n_upd[p] = len(self.updates[p])
            spec.upd[p] = C.cast(self.records[p], C.c_void_p)
        spec.n_groups = min(self.n_groups, *(len(u) for u in self.updates))
        spec.n_streams = self.n_streams
        spec.n_layers = self.M
        spec.d_B = self.B.data_ptr()
        spec.caps = self.caps
        return spec

    def needs(self, p):
        """What one group of update set p needs now (host plans of its
        updates from the host charge mirror): the maximum over its groups."""
        ups = self.updates[p]
        ng = min(self.n_groups, *(len(u) for u in self.updates))
        base, extra = divmod(len(ups), ng)
        worst, first = {}, 0
        for g in range(ng):
            n = base + (1 if g < extra else 0)
            plans = []
            for _, st, _, mode in ups[first:first + n]:
                ch = st.charge_host
                plans.append(Plan(st.d, st.chi, ch[mode], ch[mode + 1], ch[mode + 2]))
            arr = (C.c_void_p * n)(*[pl.address for pl in plans])
            c = _capi.LayerCaps()
            check(lib().bsmpo_layer_needs(n, arr, C.byref(c)), "layer_needs")
            for k, v in c.as_dict().items():
                worst[k] = max(worst.get(k, 0), v)
            first += n
        return worst

    def _set_caps(self, worst):
        c = _capi.LayerCaps()
        h = self.headroom
        c.var_bytes = int(worst["var_bytes"] * h) + (1 << 20)
        c.work_bytes = int(worst["work_bytes"] * h) + (1 << 20)
        for f in _capi.LayerCaps.INT_FIELDS:
            setattr(c, f, int(worst[f] * h) + 64)
        self.caps = c

    def _create(self):
        self.destroy()
        spec = self._spec()
        sz = C.c_size_t()
        check(lib().bsmpo_layer_query(C.byref(spec), C.byref(sz)), "layer_query")
        if self.work is None or self.work.numel() < sz.value:
            self.work = None
            self.work = torch.empty(sz.value, dtype=torch.uint8, device=self.device)
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        h = C.c_void_p()
        check(lib().bsmpo_layer_create(C.byref(spec), C.c_void_p(self.work.data_ptr()),
                                       self.work.numel(), C.c_void_p(self.stream.cuda_stream),
                                       C.byref(h)), "layer_create")
        self.prog = h

    def destroy(self):
        if self.prog is not None:
            check(lib().bsmpo_layer_destroy(self.prog), "layer_destroy")
            self.prog = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def state(self):
        """(bsmpo_layer_state, [what each group needed]) after a device sync."""
        st = _capi.LayerState()
        needs = (_capi.LayerCaps * max(1, self._spec().n_groups))()
        check(lib().bsmpo_layer_state_read(self.prog, C.byref(st), needs), "layer_state_read")
        return st, [n.as_dict() for n in needs]

    # ---------------------------------------------------------------- run
    def _sync_host_charges(self):
        for st in self.states:
            st.charge_host[:] = st.charge.cpu().numpy()

    def _active(self, layer):
        ups = self.updates[layer % 2]
        if self.max_input is None:
            return len(ups)
        return sum(1 for _, _, _, mode in ups if mode <= self.max_input + layer)

    def run(self, layers):
        """Applies the given consecutive layers (pairs beyond the light cone
        skipped when the runner knows the input modes, else every pair);
        returns the number of beam-splitter updates applied."""
        layers = list(layers)
        for a, b in zip(layers, layers[1:]):
            if b != a + 1:
                raise ValueError("layers must be consecutive")
        if not layers:
            return 0
        if layers[0] < 0 or layers[-1] >= self.M:
            raise ValueError("layer out of range")
        if self.caps is None:
            worst = self.needs(layers[0] % 2)
            if len(layers) > 1:
                w2 = self.needs(layers[1] % 2)
                worst = {k: max(v, w2[k]) for k, v in worst.items()}
            self._set_caps(worst)
        if self.prog is None:
            self._create()
        main = torch.cuda.current_stream(self.device)
        s = self.stream
        s.wait_stream(main)
        sp = C.c_void_p(s.cuda_stream)
        check(lib().bsmpo_layer_seek(self.prog, layers[0], sp), "layer_seek")
        for _ in layers:
            check(lib().bsmpo_layer_launch(self.prog, sp), "layer_launch")
        main.wait_stream(s)
        st, needs = self.state()  # the one synchronising read
        self._sync_host_charges()
        if st.status == 0:
            return sum(self._active(x) for x in layers)
        if st.status not in (1, 6):
            raise _capi.BsmpoError(f"layer program: {_capi.LAYER_STATUS.get(st.status, st.status)} "
                                   f"at layer {st.failed_at}")
        # the failed layer and every later one changed nothing: grow the
        # capacities to what the device found it needed, and run them again
        failed = st.failed_at
        self.fallbacks += 1
        grown = {}
        for f, _ in _capi.LayerCaps._fields_:
            if f != "status":
                need = max(nd[f] for nd in needs)
                grown[f] = max(need, int(getattr(self.caps, f) * 1.25 / self.headroom))
        self._set_caps(grown)
        self._create()
        done = sum(self._active(x) for x in layers if x < failed)
        return done + self.run([x for x in layers if x >= failed])


# Truncation rules per contraction precision: (tie band, drop rule), both
# relative.  Rounding of the contraction moves degenerate twin singular values
# apart by about its relative error, and leaves a floor of noise singular
# values below which nothing is signal: complex64 about 1e-7, split BF16x3
# about 1e-6 (hence its wider tie band and its drop rule at 1e-5 sigma_max).
PRECISION_RULES = {C128: (1e-6, 1e-6), C64: (1e-6, 1e-6), BF16X3: (1e-4, 1e-5)}


def set_precision_rules(precision: int):
    """Set the library's tie band and drop rule (process-wide, plans built
    afterwards) for a contraction precision; returns (tie_rel, drop_rel)."""
    tie, drop = PRECISION_RULES[int(precision)]
    check(lib().bsmpo_set_tie_rel(tie), "set_tie_rel")
    check(lib().bsmpo_set_drop_rel(drop), "set_drop_rel")
    return tie, drop


def simulate_batch(states, meshes, input_modes, layers=None, precision=C128, runner=None,
                   observe=False):
    """Layer-synchronous simulation of several independent instances: every
    non-skipped beam splitter of a layer, over all instances, in one grouped
    run (one launch per stage).  With a LayerGraphRunner for these states as
    `runner` (built with the same input modes and precision), consecutive
    layers are planned on the device and replayed as graphs instead."""
    if isinstance(runner, LayerGraphRunner):
        M = states[0].M
        todo = list(range(M) if layers is None else layers)
        probs = [[] for _ in states]
        n_updates = 0
        if observe:
            for layer in todo:
                n_updates += runner.run([layer])
                for k, st in enumerate(states):
                    probs[k].append(st.observables()[0])
        elif todo:
            n_updates = runner.run(todo)
        return n_updates, probs
    runner = runner or GroupRunner(states[0].device)
    M, d = states[0].M, states[0].d
    max_input = max(input_modes)
    probs = [[] for _ in states]
    n_updates = 0
    for layer in (range(M) if layers is None else layers):
        todo = layer_modes(M, layer, max_input)
        if not todo:
            continue
        ups = []
        for st, mesh in zip(states, meshes):
            U3 = bs_tables(d, np.stack([mesh[layer, i] for i, _ in todo]), st.device)
            ups.extend((st, mode, U3[k]) for k, (i, mode) in enumerate(todo))
        runner.run(ups, precision)
        n_updates += len(ups)
        if observe:
            for k, st in enumerate(states):
                probs[k].append(st.observables()[0])
    return n_updates, probs


# ------------------------------------------------------------ initial state

def autotune_eig_threshold(states, meshes, input_modes, layer, precision=C128, runner=None,
                           candidates=(0, 32, 64, 96), grouped=True, tie=0.01, repeats=2):
    """Pick the split between the two sector eigensolver routes (one-sided
    Jacobi in LDS for Gram size n <= threshold, the tridiagonal route above)
    by timing one layer per candidate on clones of `states`, and set the
    fastest with bsmpo_set_eig_threshold.  grouped: the layer runs as grouped
    launches (simulate_batch), else one update at a time (simulate).  Every
    candidate gives the same results to rounding; only the time differs.
    Each candidate is timed `repeats` times, the passes alternating in order
    (forward, backward, ...), and keeps its shortest time.  Candidates within
    `tie` (relative) of the fastest count as equal and the smallest of them is
    taken, so that timing noise does not flip the choice.
    Returns (best, {thr: seconds})."""
    import time
    runner = runner or GroupRunner(states[0].device)
    times = {}
    order = []
    for r in range(max(1, int(repeats))):
        order += list(candidates) if r % 2 == 0 else list(candidates)[::-1]
    for thr in order:
        check(lib().bsmpo_set_eig_threshold(int(thr)), "set_eig_threshold")
        clones = [st.clone() for st in states]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if grouped:
            simulate_batch(clones, meshes, input_modes, layers=[layer], precision=precision,
                           runner=runner)
        else:
            for st, mm in zip(clones, meshes):
                simulate(st, mm, input_modes, layers=[layer], precision=precision, observe=False)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        times[int(thr)] = min(el, times.get(int(thr), el))
        del clones
    fastest = min(times.values())
    best = min(t for t, sec in times.items() if sec <= (1.0 + tie) * fastest)
    check(lib().bsmpo_set_eig_threshold(best), "set_eig_threshold")
    return best, times


# Launch and tile geometry the layer-program autotuner searches, in the order
# it tunes them: the eigensolver route split (bsmpo_set_eig_threshold), the
# Householder block-shape split (bsmpo_set_tri_split) and the contraction's
# super-tile order (bsmpo_set_tile_order).  Every setting gives the same
# results to rounding.
LAYER_KNOBS = (("eig_threshold", (0, 32, 64, 96)),
               ("tri_split", (256, 384, 512)),
               ("tile_order", (4, 8, 16)))


def _knob(name, value=None):
    if value is None:
        return int(getattr(lib(), f"bsmpo_{name}")())
    check(getattr(lib(), f"bsmpo_set_{name}")(int(value)), f"set_{name}")
    return int(value)


def autotune_layer(states, meshes, layer, precision=C128, knobs=LAYER_KNOBS, n_groups=8,
                   n_streams=2, repeats=2, tie=0.01, headroom=1.3, log=None):
    """Tunes the launch and tile geometry of the layer-program path (device
    planning, graph replays) by timing one replay of `layer` per candidate on
    copies of `states` (restored before every trial), one knob at a time in
    `knobs` order, each knob's best kept for the next (coordinate descent).
    Each candidate is timed `repeats` times, the passes alternating in order,
    and keeps its shortest time; candidates within `tie` of the fastest count
    as equal, and among them the value in effect before tuning stays (else
    the fastest), so that timing noise does not move a knob.  The chosen
    values are set (process-wide) and returned with every timing:
    ({knob: value}, {knob: {value: seconds}})."""
    import time
    trial = [st.clone() for st in states]
    best, table = {}, {}
    for name, cands in knobs:
        incumbent = _knob(name)
        times = {}
        order = []
        for r in range(max(1, int(repeats))):
            order += list(cands) if r % 2 == 0 else list(cands)[::-1]
        for v in order:
            _knob(name, v)
            for t, st in zip(trial, states):
                t.Gamma.copy_(st.Gamma)
                t.Lambda.copy_(st.Lambda)
                t.charge.copy_(st.charge)
                t.charge_host[:] = st.charge_host
            runner = LayerGraphRunner(trial, meshes, precision, n_groups=n_groups,
                                      n_streams=n_streams, headroom=headroom)
            runner._set_caps(runner.needs(layer % 2))
            runner._create()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            runner.run([layer])
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            runner.destroy()
            del runner
            times[int(v)] = min(el, times.get(int(v), el))
        fastest = min(times.values())
        near = [v for v in cands if times[int(v)] <= (1.0 + tie) * fastest]
        pick = incumbent if incumbent in near else min(near, key=lambda v: times[int(v)])
        best[name] = _knob(name, pick)
        table[name] = times
        if log:
            log(f"[autotune] {name}: {pick} " + " ".join(f"{k}:{v:.3f}" for k, v in times.items()))
    del trial
    torch.cuda.empty_cache()
    return best, table


def input_density(d: int, eta: float) -> np.ndarray:
    """Single photon through a pure-loss channel of transmission eta:
    (1 - eta)|0><0| + eta|1><1| (src/MPO.py:41-44 with loss = 1 - eta)."""
    rho = np.zeros((d, d), np.float64)
    rho[0, 0] = 1.0 - eta
    rho[1, 1] = eta
    return rho


def product_mpo(M, d, chi, input_modes, eta, post_selected=None):
    """Product-state MPO (host arrays) before canonicalisation; generalises
    src/MPO.py:50-109 to any set of occupied modes and chi >= N + 1.
    Bond i carries the diagonal charges (c, c), c = 0..(photons in sites >= i);
    Gamma_i[(a,a), (b,b)] = rho_i[a-b, a-b]; Lambda = 1 on used slots."""
    inputs = sorted(set(int(m) for m in input_modes))
    rho_in = input_density(d, eta)
    rho_vac = np.zeros((d, d))
    rho_vac[0, 0] = 1.0
    n_right = [sum(1 for m in inputs if m >= i) for i in range(M + 1)]
    charge = np.full((M + 1, chi, 2), d, np.int32)
    slots = []
    for i in range(M + 1):
        if i == 0:
            cs = [post_selected] if post_selected is not None else list(range(min(n_right[0], d - 1) + 1))
        elif i == M:
            cs = [0]
        else:
            cs = list(range(min(n_right[i], d - 1) + 1))
        if len(cs) > chi:
            raise ValueError("chi too small for the product state")
        for k, c in enumerate(cs):
            charge[i, k] = (c, c)
        slots.append(cs)
    Gamma = np.zeros((M, chi, chi), np.complex128)
    for i in range(M):
        rho = rho_in if i in inputs else rho_vac
        for a_i, a in enumerate(slots[i]):
            for b_i, b in enumerate(slots[i + 1]):
                n = a - b
                if 0 <= n < d:
                    Gamma[i, a_i, b_i] = rho[n, n]
    Lambda = np.zeros((M - 1, chi), np.float64)
    for i in range(1, M):
        Lambda[i - 1, :len(slots[i])] = 1.0
    return Gamma, Lambda, charge


def canonicalisation_bs(d: int) -> np.ndarray:
    """Single-particle matrix of the reference canonicalisation sweep,
    Rand_BS_MPS(0, d, 0) (src/MPO.py:6-10,151): t = exp(2 pi i u), r = 0."""
    rs = np.random.RandomState(0)
    t = np.exp(1j * rs.rand() * 2 * np.pi)
    return np.array([[t, 0.0], [0.0, np.conj(t)]], dtype=np.complex128)


def initialize(M, d, chi, input_modes, eta, post_selected=None, device="cuda") -> DeviceMPO:
    """Initial MPO on the device: product state, then the right-to-left
    canonicalisation sweep of src/MPO.py:137-156, run by the GPU update."""
    G, L, ch = product_mpo(M, d, chi, input_modes, eta, post_selected)
    st = DeviceMPO.from_host(G, L, ch, d, device)
    U3 = bs_tables(d, canonicalisation_bs(d), st.device)[0]
    for mode in range(M - 2, -1, -1):
        st.update(mode, U3)
    return st


# ---------------------------------------------------------------- schedule

def layer_modes(M: int, layer: int, max_input: int):
    """(position, mode) of the non-skipped BSs of a brick-wall layer
    (Full.py:152-153; skip mode > last occupied input + layer)."""
    return [(i, mode) for i, mode in enumerate(range(layer % 2, M - 1, 2))
            if mode <= max_input + layer]


def simulate(st: DeviceMPO, mesh: np.ndarray, input_modes, layers=None, precision=C128,
             observe=True, callback=None):
    """Apply the mesh layer by layer (Full.py:169-202).  Returns per-layer
    trace prob[M+1] and EE[M-1, M+1] (NaN where not observed)."""
    M, d = st.M, st.d
    max_input = max(input_modes)
    prob = np.full(M + 1, np.nan)
    EE = np.full((M - 1, M + 1), np.nan)
    if observe:
        prob[0] = st.observables()[0]
        EE[:, 0] = st.entanglement_entropy()
    for layer in (range(M) if layers is None else layers):
        todo = layer_modes(M, layer, max_input)
        if not todo:
            continue
        U3 = bs_tables(d, np.stack([mesh[layer, i] for i, _ in todo]), st.device)
        for k, (i, mode) in enumerate(todo):
            st.update(mode, U3[k], precision)
            if callback is not None:
                callback(layer, i, mode, st)
        if observe:
            prob[layer + 1] = st.observables()[0]
            EE[:, layer + 1] = st.entanglement_entropy()
    return prob, EE

"""PM + learned correction: the fast self-gravity method for large N (3-D).

Force a = split mesh + near field (p3m.SplitForce) + the learned residual
(learned.LearnedCorrection) evaluated from the same mesh; leapfrog KDK in
FP64 state with libssim's kick/drift kernels.  The FP64 direct sum
(selfgrav.SelfGravSystem) stays available as the fallback and as the judge
of the force error (`force_errors`).

`make_training_set` builds the learned part's data from the method's own
run: it integrates with the split force and at every snapshot the FP64
all-pairs kernel computes the exact forces whose residual the network
learns (relative to the split force, see learned.py).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import _lib
from . import device as D
from . import pm as P
from .learned import LearnedCorrection
from .p3m import SplitForce
from .selfgrav import SelfGravSystem

DATA = Path(__file__).resolve().parent / "data"
DEFAULT_WEIGHTS = DATA / "learned_cube_250k.npz"

# split parameters of the 250k-mass configuration: 128^3 FP32 mesh with TSC
# assignment, split scale rs = 1 cell, near field to 5 rs, four empty cells
# around the bodies and the mesh placement kept while it fits (measured on
# the 1000-step run: median force error 1e-4, energy drift 1.2x the direct
# sum's; a mesh that follows the box every step, or CIC, drifts 2-3x more)
SPLIT = dict(mesh_n=128, rs_cells=1.0, rcut_rs=5.0, prec=P.F32, assign="tsc",
             margin=4.0, freeze=True)

# time step of the 250k-mass cube (G = 1, total mass 1 in the unit cube):
# 1000 steps end at t = 0.25, before the cold cube's free-fall collapse at
# t_ff = sqrt(3 pi / 32) = 0.54, where every method's cost and error are
# dominated by the singular core.
DT = 2.5e-4


def _torch():
    import torch
    return torch


class PMLearnedSystem:
    """n masses + t tracers, 3-D, G, Plummer eps2; KDK with the split force
    plus (optionally) the learned correction."""

    def __init__(self, dev, n: int, t: int, eps2: float, G: float = 1.0,
                 net: LearnedCorrection | None = None, **split):
        torch = _torch()
        self.dev, self.n, self.t = dev, int(n), int(t)
        self.eps2, self.G = float(eps2), float(G)
        cfg = dict(SPLIT)
        cfg.update(split)
        self.split = SplitForce(dev, cfg["mesh_n"], eps2, G, cfg["rs_cells"],
                                cfg["rcut_rs"], margin=cfg.get("margin", 2.0),
                                prec=cfg["prec"], freeze=cfg.get("freeze", False),
                                assign=cfg.get("assign", "cic"))
        self.net = net
        nb = self.n + self.t
        self.x = [torch.zeros(nb, dtype=torch.float64, device=dev) for _ in range(3)]
        self.v = [torch.zeros(nb, dtype=torch.float64, device=dev) for _ in range(3)]
        self.a = [torch.zeros(nb, dtype=torch.float64, device=dev) for _ in range(3)]
        self.m = torch.zeros(max(self.n, 1), dtype=torch.float64, device=dev)
        self.feat = None
        self.acc_valid = False
        D.on_shutdown(self)

    def close(self):
        """Release the captured step graph, the side streams, the pinned box
        buffer (which a kernel writes) and the workspaces, after the device
        has finished with them.  The system can run again afterwards (it
        re-creates them)."""
        torch = _torch()
        if torch.cuda.is_initialized():
            torch.cuda.synchronize(self.dev)
        for g in (getattr(self, "_graphs", None) or {}).values():
            g.reset()
        self._graphs, self._graph_key = {}, None
        self._box_dev = self._box_host = self._box_ws = self._kb_ws = None
        self._box_ev = None
        self.split._side = None

    def set_bodies(self, pos, vel, mass, tpos=None, tvel=None):
        for c in range(3):
            self.x[c][:self.n].copy_(D.to_dev(pos[c], self.dev))
            self.v[c][:self.n].copy_(D.to_dev(vel[c], self.dev))
            if self.t:
                self.x[c][self.n:].copy_(D.to_dev(tpos[c], self.dev))
                self.v[c][self.n:].copy_(D.to_dev(tvel[c], self.dev))
        self.m[:self.n].copy_(D.to_dev(mass, self.dev))
        self.mass_unit = float(np.mean(mass))
        self.split.set_mass_bound(self.m[:self.n])
        self.acc_valid = False

    def a_unit(self) -> float:
        return self.G * self.mass_unit / self.split.h ** 2

    def out_unit(self) -> float:
        """The network's output unit: out_scale (x |a| per body) in the
        relative mode, else G m_body / h^2."""
        return self.net.out_scale if self.net.relative else self.a_unit()

    def features(self):
        if self.feat is None:
            self.feat = _torch().empty((self.n + self.t, 32), dtype=_torch().int16,
                                       device=self.dev)
        self.net.mass_unit = self.mass_unit
        return self.net.features(self.x, self.split.rho, self.split.origin,
                                 self.split.h, self.split.prec, out=self.feat,
                                 n_mass=self.n)

    # fold the learned correction into the mesh gather (features and the
    # network run beside the convolution); False: a separate update after it
    fold = True

    def _correction(self):
        """Features and network outputs (no update) for the gather."""
        self._correction_features()
        return self._correction_outputs()

    def _correction_features(self):
        """The feature rows (read the deposited mass mesh)."""
        self.features()

    def _correction_outputs(self):
        """The network's outputs for the feature rows, as the gather's
        `corr` dict."""
        torch = _torch()
        if getattr(self, "_corr_out", None) is None:
            self._corr_out = torch.empty((self.n + self.t, 3), dtype=torch.float32,
                                         device=self.dev)
        self.net.predict(self.feat, out=self._corr_out)
        return dict(out=self._corr_out, scale=self.out_unit(),
                    relative=self.net.relative, vel=self.v, n_proj=self.n)

    def force(self, learned: bool = True, placed: bool = False, box=None):
        """a = split force (+ learned correction) at the current positions;
        `placed`: the mesh is placed for them already (run's steps).  `box`
        (_gather_box only): the gather also writes the box of the next
        drift (SplitForce.mesh_gather)."""
        fold = (learned and self.net is not None and self.fold and
                self.split.can_fold_correction())
        if box is not None and learned and self.net is not None and not fold:
            raise ValueError("the box comes from the gather only with the correction folded")
        # features behind the deposit, the network behind the near field
        corr = (self._correction_features, self._correction_outputs) if fold else None
        self.split.accelerations(self.x, self.n, self.m, self.a, placed=placed,
                                 correction=corr, box=box)
        if learned and self.net is not None and not fold:
            feat = self.features()
            # masses: no component along the velocity (no work done on
            # the self-gravitating system); tracers: the full correction
            self.net.predict(feat, add_to=self.a, scale=self.out_unit(),
                             relative=self.net.relative, vel=self.v,
                             n_proj=self.n)
        self.acc_valid = True
        return self.a

    # -- time stepping -------------------------------------------------------
    # A step is a fixed sequence of launches once the mesh is placed: the
    # box of the positions the step's drift will produce is computed on the
    # device at the end of the previous step (ssim_box_after_kdk, exact),
    # the host reads those six numbers, places the mesh (place_box), and
    # the step runs without another host round trip.  While the placement
    # is kept the step is replayed from a captured HIP graph, which removes
    # the ~100 kernel launches per step from the host's critical path; the
    # eager and the replayed steps run the same kernels with the same
    # arguments, so they give identical results.  The graph executor runs
    # the branches of a two-stream graph in capture order rather than side
    # by side, so the overlapped force is captured with the side stream's
    # binning first (split.side_first): the near field then runs beside the
    # deposit and the convolution.  Inside a run() call each step but the
    # last ends with the next step's opening kick and drift fused into its
    # closing kick (ssim_kick_drift_box), so only the first step launches
    # its own (ssim_kick_drift).  When the gather gives the final
    # accelerations (the correction folded into it, or none), it also
    # writes the next box (ssim_tsc_gather_box): the host waits for that
    # part of the step only, and places the mesh and queues the next step
    # while the step's kicks and drift (ssim_kicks_drift) still run.
    use_graph = True
    captures = 0          # step graphs captured so far (three per placement at most)

    def _box_buffers(self):
        torch = _torch()
        if getattr(self, "_box_dev", None) is None:
            nb = self.n + self.t
            self._box_dev = torch.empty(6, dtype=torch.float64, device=self.dev)
            self._box_host = torch.empty(6, dtype=torch.float64, pin_memory=True)
            lib = _lib.load()
            self._box_ws = D.workspace(lib.ssim_minmax_work_bytes(nb), self.dev)
            self._kb_ws = D.workspace(lib.ssim_kick_box_work_bytes(nb), self.dev)
            self._gb_ws = D.workspace(lib.ssim_tsc_gather_box_work_bytes(nb), self.dev)
            self._box_ev = torch.cuda.Event()

    def _next_box(self, half, full):
        ps = lambda lst: [_lib.ptr(c) for c in lst]  # noqa: E731
        _lib.call("ssim_box_after_kdk", 3, self.n + self.t, *ps(self.x), *ps(self.v),
                  *ps(self.a), half, full, _lib.ptr(self._box_dev),
                  _lib.ptr(self._box_ws), self._box_ws.numel(), D.stream(self.dev))

    def _read_box(self, in_host=False, recorded=False):
        """The box of the next drift: from the pinned host buffer (in_host:
        written there by the step) or copied from the device one.
        recorded: the box event was recorded after the launch that writes
        it (later work may still run)."""
        if not in_host:
            self._box_host.copy_(self._box_dev, non_blocking=True)
        if not (in_host and recorded):
            self._box_ev.record()
        self._box_ev.synchronize()
        b = self._box_host.numpy()
        return [(float(b[2 * c]), float(b[2 * c + 1])) for c in range(3)]

    def _opening(self, dt):
        """The opening half kick and drift of a step (ssim_kick_drift)."""
        ps = lambda lst: [_lib.ptr(c) for c in lst]  # noqa: E731
        _lib.call("ssim_kick_drift", 3, self.n + self.t, *ps(self.x), *ps(self.v),
                  *ps(self.a), 0.5 * dt, dt, D.stream(self.dev))

    def _step(self, dt, learned, snap_ptr=None, drift_next=False):
        """Launches of one KDK step after its opening kick and drift (the
        mesh already placed for the drifted positions), ending with the box
        of the next step's drift.  drift_next: another step follows, and its
        opening kick and drift are fused into this step's closing kick
        (ssim_kick_drift_box: the same operations in the same order)."""
        nb = self.n + self.t
        st = D.stream(self.dev)
        ps = lambda lst: [_lib.ptr(c) for c in lst]  # noqa: E731
        half = 0.5 * dt
        self.force(learned, placed=True)
        if snap_ptr is None:
            # closing kick (+ the next opening kick and drift) + next box in
            # one launch; the box goes straight into the pinned host buffer
            # (kernels can write it), so reading it needs no copy
            _lib.call("ssim_kick_drift_box" if drift_next else "ssim_kick_box", 3, nb,
                      *ps(self.x), *ps(self.v), *ps(self.a), half, dt,
                      _lib.ptr(self._box_host), _lib.ptr(self._kb_ws),
                      self._kb_ws.numel(), st)
            return
        _lib.call("ssim_kick_snap", 3, self.n, self.t, *ps(self.x), *ps(self.v),
                  *ps(self.a), half, snap_ptr, st)
        self._next_box(half, dt)

    def _gather_box(self, learned):
        """The gather can write the next box: TSC with interleaved FP32
        fields, and the correction (if any) folded into the gather."""
        return self.split.can_fold_correction() and (
            not learned or self.net is None or self.fold)

    def _force_box(self, dt, learned):
        """A step's force (its opening kick and drift done), the gather
        writing the box of the next drift into the pinned host buffer."""
        self.force(learned, placed=True,
                   box={"vel": self.v, "half": 0.5 * dt, "full": dt,
                        "out": _lib.ptr(self._box_host), "work": self._gb_ws})

    def _tail(self, dt, drift_next):
        """A step's closing half kick and, drift_next, the next step's
        opening half kick and drift (ssim_kicks_drift: the operations of
        ssim_kick then ssim_kick_drift)."""
        ps = lambda lst: [_lib.ptr(c) for c in lst]  # noqa: E731
        st = D.stream(self.dev)
        if drift_next:
            _lib.call("ssim_kicks_drift", 3, self.n + self.t, *ps(self.x), *ps(self.v),
                      *ps(self.a), 0.5 * dt, dt, st)
        else:
            _lib.call("ssim_kick", 3, self.n + self.t, *ps(self.v), *ps(self.a), 0.5 * dt,
                      st)

    def _graph_for(self, dt, learned, drift_next=False):
        """The captured step for the current placement (captured anew when
        the placement, dt or the learned flag changed); both forms of the
        step's end (drift_next) are kept, and the force with the gather's
        box (drift_next = "box", _force_box)."""
        torch = _torch()
        key = (self.split.placement(), float(dt), bool(learned), self.fold,
               self.split.overlap, self.split.near_hilo)
        if getattr(self, "_graph_key", None) != key:
            for g in (getattr(self, "_graphs", None) or {}).values():
                g.reset()
            self._graphs, self._graph_key = {}, key
        if drift_next not in self._graphs:
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream(device=self.dev)    # default priority
            side.wait_stream(torch.cuda.current_stream(self.dev))
            self.split.side_first = True
            try:
                with torch.cuda.graph(g, stream=side):
                    if drift_next == "box":
                        self._force_box(dt, learned)
                    else:
                        self._step(dt, learned, drift_next=drift_next)
            finally:
                self.split.side_first = False
            torch.cuda.current_stream(self.dev).wait_stream(side)
            self._graphs[drift_next] = g
            self.captures += 1
        return self._graphs[drift_next]

    def run(self, dt: float, steps: int, snap=None, learned: bool = True,
            graph: bool | None = None):
        """`steps` KDK steps; optional f32 snapshot tensor (steps, 6n+3t).
        graph: replay captured steps (default: use_graph, with a kept mesh
        placement and no snapshots); False launches every kernel."""
        self._box_buffers()
        if not self.acc_valid:
            self.force(learned)
        use = self.use_graph if graph is None else graph
        use = use and snap is None and self.split.freeze
        # the gather writes the next box, and the host does not wait for
        # the step's kicks and drift
        early = snap is None and self._gather_box(learned)
        self._next_box(0.5 * dt, dt)
        in_host = False
        for s in range(steps):
            self.split.place_box(self._read_box(in_host, recorded=early))
            in_host = snap is None    # the step writes it into the host buffer
            # the first step opens with its own kick and drift; each later
            # one's were fused into the step before (no snapshots)
            if s == 0 or snap is not None:
                self._opening(dt)
            drift_next = snap is None and s + 1 < steps
            if early:
                if use:
                    self._graph_for(dt, learned, "box").replay()
                else:
                    self._force_box(dt, learned)
                self._box_ev.record()
                self._tail(dt, drift_next)
            elif use:
                self._graph_for(dt, learned, drift_next).replay()
            else:
                self._step(dt, learned, None if snap is None else _lib.ptr(snap[s]),
                           drift_next)

    def host_state(self):
        return ([c.cpu().numpy() for c in self.x], [c.cpu().numpy() for c in self.v])


def direct_system(dev, n, t, eps2, G=1.0):
    return SelfGravSystem(dev, 3, n, t, _lib.PREC_F64, _lib.SCHEME_KDK, G, eps2)


def exact_accelerations(direct: SelfGravSystem, x, n, m):
    """FP64 all-pairs accelerations of the bodies x (masses first)."""
    for c in range(3):
        direct.x[c][:x[c].numel()].copy_(x[c])
    direct.m[:n].copy_(m[:n])
    return direct.accelerations()


def force_errors(a, ref):
    """Per-body relative error |a - ref| / |ref| (host arrays)."""
    a = np.array([c.cpu().numpy() if hasattr(c, "cpu") else c for c in a])
    r = np.array([c.cpu().numpy() if hasattr(c, "cpu") else c for c in ref])
    return np.linalg.norm(a - r, axis=0) / np.linalg.norm(r, axis=0)


def make_training_set(dev, pos, m, tpos, eps2, dt, n_snap=64, every=15,
                      G=1.0, net=None, log=None, **split):
    """Snapshots of the method's own run with exact FP64 targets.  Returns
    (features int16 [S*nb][32], targets f32 [S*nb][3]) on the device.  The
    bodies are loaded in a random order (the force does not depend on it),
    so every contiguous slice of a snapshot is a spatially mixed batch."""
    torch = _torch()
    n, t = len(m), len(tpos[0])
    nb = n + t
    rng = np.random.default_rng(0)
    pm_, pt_ = rng.permutation(n), rng.permutation(t)
    pos = [np.asarray(c)[pm_] for c in pos]
    m = np.asarray(m)[pm_]
    tpos = [np.asarray(c)[pt_] for c in tpos]
    net = net or LearnedCorrection(dev)
    sysm = PMLearnedSystem(dev, n, t, eps2, G, net=net, **split)
    sysm.set_bodies(pos, [np.zeros(n)] * 3, m, tpos, [np.zeros(t)] * 3)
    direct = direct_system(dev, n, t, eps2, G)
    feats = torch.empty((n_snap * nb, 32), dtype=torch.int16, device=dev)
    tgts = torch.empty((n_snap * nb, 3), dtype=torch.float32, device=dev)
    for s in range(n_snap):
        sysm.run(dt, every, learned=False)
        a_split = sysm.force(learned=False)
        a_ref = exact_accelerations(direct, sysm.x, n, sysm.m)
        sysm.features()
        feats[s * nb:(s + 1) * nb].copy_(sysm.feat)
        _lib.call("ssim_mlp_residual", nb, *[_lib.ptr(c) for c in a_ref],
                  *[_lib.ptr(c) for c in a_split], sysm.out_unit(),
                  int(net.relative), float(net.clip),
                  _lib.ptr(tgts[s * nb:(s + 1) * nb]), D.stream(dev))
        if log:
            log(f"snapshot {s + 1}/{n_snap}")
    return feats, tgts


def save_weights(net: LearnedCorrection, path=DEFAULT_WEIGHTS, **meta):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=net.host_params(), relative=np.int32(net.relative),
             out_scale=np.float64(net.out_scale), clip=np.float64(net.clip),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_weights(dev, path=DEFAULT_WEIGHTS) -> LearnedCorrection:
    with np.load(path, allow_pickle=False) as f:
        net = LearnedCorrection(dev, f["params"])
        if "relative" in f.files:
            net.relative = bool(f["relative"])
            net.out_scale = float(f["out_scale"])
            net.clip = float(f["clip"])
        else:                       # absolute units (G m_body / h^2)
            net.relative = False
    return net

"""One projection timestep: ``step_pure`` with the reference signature and a
preallocated ``CollocatedStepper`` for repeated stepping.

Mirrors solver/pure_functions/step.py:49-467 (collocated grid, von Karman
flow): [LES] -> rk_step_unified -> predictor BCs -> rhs = div*mask/dt ->
poisson_multigrid -> projection.  Every arithmetic stage is a HIP kernel.

The predictor BCs of step.py:182-191 (inlet + slip walls on u*, v*) are
already the last thing ``rk_step_unified`` applies (rk3_simple_new.py:
264-279 with slip walls), so re-imposing them is the identity and no kernel
is spent on it.  With ``use_les`` the reference computes nu_sgs and then
passes ``nu_sgs=None`` to the RK step (step.py:177-180), i.e. the eddy
viscosity never reaches the flow; ``les_coupling="discard"`` (default)
reproduces that and skips the dead computation, ``"apply"`` feeds it to the
RK step (the RK function's own contract, rk3_simple_new.py:304).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _arrays as A
from . import _lib
from .advection_schemes import rk_step_unified
from .operators import (constant_smagorinsky, dynamic_smagorinsky, pressure_rhs, project,
                        project_collocated, smagorinsky_total)
from .pressure_solvers import get_plan


@dataclass
class SimState:
    """Simulation state (solver/params.py:11-30)."""
    u: torch.Tensor
    v: torch.Tensor
    p: torch.Tensor
    u_prev: torch.Tensor
    v_prev: torch.Tensor
    c: torch.Tensor
    dt: float
    iteration: int
    grid_type: str = "collocated"
    integral: float = 0.0
    prev_error: float = 0.0


def initial_state(u, v, dt: float, device=None) -> SimState:
    ut = A.as_device(u, A.torch_dtype_of(u), device, "u")
    vt = A.as_device(v, ut.dtype, ut.device, "v")
    z = torch.zeros_like(ut)
    return SimState(u=ut, v=vt, p=z.clone(), u_prev=ut.clone(), v_prev=vt.clone(), c=z.clone(),
                    dt=float(dt), iteration=0)


def step_pure(state: SimState, mask, sdf, dx: float, dy: float, nu: float, U_inf: float,
              use_les: bool = False, smagorinsky_constant: float = 0.1,
              weno_epsilon: float = 1e-6, eps: float = 0.01, adaptive_dt: bool = False,
              dt_min: float = 5e-4, dt_max: float = 0.01, target_div: float = 1e-4,
              Kp: float = 0.5, Ki: float = 0.05, Kd: float = 0.1,
              flow_type: str = "von_karman", advection_scheme: str = "rk3",
              pressure_solver: str = "multigrid", les_model: str = "smagorinsky",
              limiter: str = "minmod", v_cycles: int = 1, fast_mode: bool = False,
              brinkman_eta: float = 0.005, grid_type: str = "collocated",
              verbose: bool = False, nn_pressure_model=None, *,
              les_coupling: str = "discard", mg_variant: str = "reference",
              tolerance: float = 1e-6, max_cycles: int = 40, warm_start=None) -> SimState:
    """One timestep; arguments as step_pure (step.py:49-66).

    Extensions (keyword-only): ``les_coupling`` ("discard" | "apply"),
    ``mg_variant`` ("reference": ``v_cycles`` reference V-cycles;
    "textbook": solve to ``||div|| <= tolerance``), ``warm_start``: callable
    ``(rhs, sdf) -> p0`` used as the multigrid initial guess.
    """
    if grid_type != "collocated":
        raise NotImplementedError("only grid_type='collocated' is implemented")
    if flow_type != "von_karman":
        raise NotImplementedError("only flow_type='von_karman' is implemented")
    if adaptive_dt:
        raise NotImplementedError("adaptive_dt: the PID dt controller is not implemented")
    if use_les and les_model not in ("smagorinsky", "dynamic_smagorinsky"):
        raise NotImplementedError(f"les_model={les_model!r}")
    u, v, dt = state.u, state.v, state.dt
    mt = A.as_device(mask, u.dtype, u.device, "mask")
    st = A.as_device(sdf, u.dtype, u.device, "sdf")
    nu_sgs = None
    if use_les and les_coupling == "apply":
        delta = (dx * dy) ** 0.5
        if les_model == "dynamic_smagorinsky":
            nu_sgs, _ = dynamic_smagorinsky(u, v, dx, dy, delta)
        else:
            nu_sgs = constant_smagorinsky(u, v, dx, dy, delta, smagorinsky_constant)
    us, vs = rk_step_unified(u, v, dt, nu, dx, dy, mt, sdf=st, U_inf=U_inf, nu_sgs=nu_sgs,
                             nu_hyper_ratio=0.0, slip_walls=True, fast_mode=fast_mode,
                             brinkman_eta=brinkman_eta, flow_type=flow_type)
    rhs = pressure_rhs(us, vs, mt, dt, dx, dy)
    nx, ny = u.shape
    if pressure_solver == "nn" and nn_pressure_model is not None:
        p = nn_pressure_model(rhs, mt)
    elif pressure_solver == "multigrid":
        p0 = warm_start(rhs, st) if warm_start is not None else None
        p = torch.empty_like(rhs)
        if mg_variant == "reference":
            plan = get_plan(nx, ny, dx, dy, u.dtype, "reference", 4, u.device)
            plan.solve(rhs, p, p0, v_cycles, tolerance)
        else:
            plan = get_plan(nx, ny, dx, dy, u.dtype, "textbook", 0, u.device)
            plan.solve(rhs, p, p0, max_cycles, tolerance)
    else:
        raise NotImplementedError(f"pressure_solver={pressure_solver!r}")
    un, vn = project(us, vs, p, mt, dt, dx, dy, U_inf)
    return SimState(u=un, v=vn, p=p, u_prev=u, v_prev=v, c=state.c, dt=dt,
                    iteration=state.iteration + 1, integral=state.integral,
                    prev_error=state.prev_error)


def inlet_velocity(U_inf: float, iteration: int, startup_ramp_steps: int) -> float:
    """Inlet speed with the cubic startup ramp (step_handlers.py:145-150);
    the reference GUI sets startup_ramp_steps = 0 (initializers.py:29)."""
    if iteration >= startup_ramp_steps:
        return float(U_inf)
    t = iteration / startup_ramp_steps
    return float(U_inf) * (3.0 * t * t - 2.0 * t ** 3)


def step_collocated(u, v, mask, dt: float, *, sdf, dx: float, dy: float, nu: float,
                    U_inf: float, use_les: bool = False, les_model: str = "smagorinsky",
                    smagorinsky_constant: float = 0.17, nu_hyper_ratio: float = 0.0,
                    slip_walls: bool = True, fast_mode: bool = False,
                    brinkman_eta: float = 0.01, v_cycles: int = 5, solver_dt: float = None,
                    iteration: int = 0, startup_ramp_steps: int = 0,
                    flow_type: str = "von_karman", pressure_solver: str = "multigrid",
                    nn_pressure_model=None, levels: int = 4):
    """One step of the GUI solver's ``_step_collocated``
    (solver/simulation/step_handlers.py:62-185, von Karman branch); returns
    ``(u, v, p)``.  The method's ``self.*`` attributes are keyword arguments
    here: ``nu``, ``U_inf`` (self.flow), ``use_les``, ``les_model``,
    ``smagorinsky_constant``, ``fast_mode``, ``brinkman_eta``, ``v_cycles``
    (self.sim_params; multigrid_v_cycles defaults to 5, params.py:274),
    ``nu_hyper_ratio`` (from ``get_re_parameters``), ``slip_walls`` (True,
    base_solver.py:211), ``solver_dt`` (self.dt, used by the correction;
    defaults to ``dt``), ``iteration`` / ``startup_ramp_steps`` (inlet ramp).

    Differences from ``step_pure`` that this composition keeps: with LES the
    RK step receives nu + nu_sgs as ``nu`` *and* nu_sgs (so the Laplacian
    sees nu + 2 nu_sgs and the biharmonic coefficient scales with
    nu + nu_sgs), no BC pass on u*, rhs = div/dt without the mask, no
    interface projection, no-slip walls in the correction.
    """
    if flow_type != "von_karman":
        raise NotImplementedError("only flow_type='von_karman' is implemented")
    if use_les and les_model not in ("smagorinsky", "dynamic_smagorinsky"):
        raise NotImplementedError(f"les_model={les_model!r}")
    A.require_gpu()
    want_np = A.is_numpy(u)
    dt_ = A.torch_dtype_of(u)
    ut = A.as_device(u, dt_, name="u")
    vt = A.as_device(v, dt_, ut.device, "v")
    mt = A.as_device(mask, dt_, ut.device, "mask")
    st = A.as_device(sdf, dt_, ut.device, "sdf")
    nu_arg, nu_sgs = nu, None
    if use_les:
        delta = (dx * dy) ** 0.5
        if les_model == "dynamic_smagorinsky":
            nu_sgs, nu_arg, _ = dynamic_smagorinsky(ut, vt, dx, dy, delta, nu=nu)
        else:
            nu_sgs, nu_arg = smagorinsky_total(ut, vt, nu, dx, dy, delta, smagorinsky_constant)
    us, vs = rk_step_unified(ut, vt, dt, nu_arg, dx, dy, mt, sdf=st, U_inf=U_inf, nu_sgs=nu_sgs,
                             nu_hyper_ratio=nu_hyper_ratio, slip_walls=slip_walls,
                             fast_mode=fast_mode, brinkman_eta=brinkman_eta, flow_type=flow_type)
    rhs = pressure_rhs(us, vs, None, dt, dx, dy)  # div / dt, no mask (step_handlers.py:92-98)
    nx, ny = ut.shape
    if pressure_solver == "nn" and nn_pressure_model is not None:
        p = nn_pressure_model(rhs, mt)
    elif pressure_solver == "multigrid":
        p = torch.empty_like(rhs)
        plan = get_plan(nx, ny, dx, dy, ut.dtype, "reference", levels, ut.device)
        plan.solve(rhs, p, None, v_cycles, 1e-6)
    else:
        raise NotImplementedError(f"pressure_solver={pressure_solver!r}")
    sdt = dt if solver_dt is None else solver_dt
    un, vn = project_collocated(us, vs, p, mt, sdt, dx, dy,
                                inlet_velocity(U_inf, iteration, startup_ramp_steps))
    return A.out_like(un, want_np), A.out_like(vn, want_np), A.out_like(p, want_np)


class CollocatedStepper:
    """Repeated steps on preallocated device buffers (no allocation per
    step): ping-pong velocity buffers, one multigrid plan.

    With a ``warm_start`` network and ``overlap_guess=True`` (textbook
    solver), the zero-guess solve is started at once on the current stream
    while the network's guess and its residual test run on a second stream
    (the test is the plan's: the guess is kept only if its residual is not
    above the zero guess's).  A kept guess is then solved from to the same
    tolerance and that solution is used; a rejected one costs nothing on the
    critical path.  ``guess_stats`` holds the last test.

    ``temporal_guess=True`` (textbook solver): a solve without another
    guess starts from the previous step's FP64 pressure, from the second
    step on from the linear extrapolation of the last two (same guess test).

    ``deferred_check=True`` (textbook solver, no ``overlap_guess``): a step
    queues its solve with the previous solve's iteration count, and the next
    step queues its predictor (LES, RK, rhs, guess) into a second buffer slot
    before it reads that solve's convergence state, so the host never drains
    the stream.  A solve that needed more iterations is completed then, its
    projection redone and the predictor queued again; results are those of
    the blocking step.  ``state``, ``load`` and ``finish()`` complete an open
    solve."""

    def __init__(self, nx, ny, dx, dy, mask, sdf, nu, U_inf, dt, *, dtype=torch.float32,
                 device="cuda", brinkman_eta=0.005, v_cycles=3, mg_variant="reference",
                 tolerance=1e-6, max_cycles=40, fast_mode=False, use_les=False,
                 les_coupling="discard", smagorinsky_constant=0.1, warm_start=None,
                 overlap_guess=False, deferred_check=False, temporal_guess=False):
        A.require_gpu()
        self.nx, self.ny, self.dx, self.dy = nx, ny, float(dx), float(dy)
        self.nu, self.U_inf, self.dt = float(nu), float(U_inf), float(dt)
        self.eta, self.v_cycles = float(brinkman_eta), int(v_cycles)
        self.mg_variant, self.tolerance, self.max_cycles = mg_variant, tolerance, max_cycles
        self.fast_mode = fast_mode
        self.warm_start = warm_start  # callable (rhs, sdf) -> p0, e.g. WarmStartNet
        self.les = use_les and les_coupling == "apply"
        self.cs = smagorinsky_constant
        dev = torch.device(device)
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.mask = A.as_device(mask, dtype, dev, "mask")
        self.sdf = A.as_device(sdf, dtype, dev, "sdf")
        e = lambda: torch.empty((nx, ny), dtype=dtype, device=dev)  # noqa: E731
        self.rhs, self.p = e(), e()
        self.nsgs = e() if self.les else None
        self.u = [e(), e()]
        self.v = [e(), e()]
        self.cur = 0
        variant = "reference" if mg_variant == "reference" else "textbook"
        deferred = bool(deferred_check) and variant == "textbook"
        self.temporal = bool(temporal_guess) and variant == "textbook"
        if deferred or self.temporal:
            # a deferred solve stays open on the plan from one step to the next,
            # and a temporal guess keeps this flow's previous pressures on the
            # plan, so such a stepper owns its plan (the shared plans serve
            # blocking callers, and a plan refuses a solve while one is open)
            from .pressure_solvers import MultigridPlan
            self.plan = MultigridPlan(nx, ny, dx, dy, dtype, variant, 0, dev)
            if self.temporal:
                self.plan.set_option(temporal=True)
        else:
            self.plan = get_plan(nx, ny, dx, dy, dtype, variant,
                                 4 if variant == "reference" else 0, dev)
        # the per-step launches go straight to the library with arguments
        # fixed here: the buffers are the stepper's own (shapes checked once),
        # so the Python cost per step is one call per kernel group
        self._lib = _lib.load()
        self._dc = A.dtype_code(self.mask)
        self._prm = _lib.RkParams(self.dt, self.nu, self.dx, self.dy, self.U_inf, self.eta, 0.0, 1,
                                  int(bool(fast_mode)))
        self._prm_ref = ctypes.byref(self._prm)
        self._cycles = self.v_cycles if variant == "reference" else int(max_cycles)
        # deferred convergence read (textbook): a step enqueues its solve, output
        # and projection without waiting; the next step (or finish / state)
        # reads the convergence state and redoes the projection if the solve
        # needed more iterations than the previous one
        self.deferred = deferred
        self._pending = False
        self._pend_stream = None  # torch stream the open solve was enqueued on
        self.topups = 0  # deferred solves that needed more iterations than enqueued
        self._own_plans = [self.plan] if (deferred or self.temporal) else []
        # predictor fields per slot: with the deferred read, a step's predictor
        # (LES, RK, rhs, guess) is queued before the previous solve's state is
        # read, into the other slot, so that a late top-up can still redo the
        # previous projection from its own predictor fields
        self._usv = [(e(), e())]
        self._usv.append((e(), e()) if self.deferred else self._usv[0])
        self.slot = 0
        # a WarmStartNet writes its guess into a stepper buffer (other
        # callables are called as warm_start(rhs, sdf) -> p0)
        self._net = warm_start if (hasattr(warm_start, "pscale") and hasattr(warm_start, "handle")
                                   and getattr(warm_start, "device", None) == dev) else None
        self.p0 = e() if warm_start is not None else None
        self.overlap = bool(overlap_guess) and warm_start is not None and variant == "textbook"
        self.guess_stats = None
        if self.overlap:
            from .pressure_solvers import MultigridPlan
            # a private plan for the guess test: stream-ordered (no host sync)
            self.guess_plan = MultigridPlan(nx, ny, dx, dy, dtype, "textbook", 0, dev)
            self.guess_plan.set_option(adaptive=False)
            self._own_plans.append(self.guess_plan)
            self.side = torch.cuda.Stream(device=dev)  # default priority
            self.p_guess = e()
            self.ev_rhs, self.ev_guess = torch.cuda.Event(), torch.cuda.Event()

    def load(self, u, v):
        self.finish()
        if self.temporal:
            self.plan.set_option(temporal=True)  # a new flow: no previous pressures
        self.u[self.cur].copy_(A.as_device(u, self.u[0].dtype, self.device, "u"))
        self.v[self.cur].copy_(A.as_device(v, self.u[0].dtype, self.device, "v"))

    @property
    def state(self):
        self.finish()
        return self.u[self.cur], self.v[self.cur]

    @property
    def us(self):
        """Predictor velocity u* of the last step."""
        return self._usv[1 - self.slot][0]

    @property
    def vs(self):
        return self._usv[1 - self.slot][1]

    def finish(self) -> bool:
        """Complete a deferred solve: read its convergence state and, if it
        needed more iterations, enqueue them and redo the projection (True).
        The top-up and the re-projection run on the stream the solve was
        enqueued on; if the current stream is another one, it waits for them."""
        if not self._pending:
            return False
        more = ctypes.c_int(0)
        with torch.cuda.device(self.device):
            _lib.check(self._lib.aj_poisson_mg_complete(self.plan.handle, ctypes.byref(more)),
                       "aj_poisson_mg_complete")
            self._pending = False
            if more.value:
                self.topups += 1
                ps = self._pend_stream
                us, vs = self._usv[1 - self.slot]
                _lib.check(self._lib.aj_project(
                    self._dc, us.data_ptr(), vs.data_ptr(), self.p.data_ptr(),
                    self.mask.data_ptr(), self.u[self.cur].data_ptr(), self.v[self.cur].data_ptr(),
                    self.nx, self.ny, self.dt, self.dx, self.dy, self.U_inf, ps.cuda_stream),
                    "aj_project")
                cur = torch.cuda.current_stream(self.device)
                if cur.cuda_stream != ps.cuda_stream:
                    cur.wait_stream(ps)
        return bool(more.value)

    def close(self):
        """Complete an open solve and release the plans this stepper owns
        (shared plans from ``get_plan`` stay); idempotent."""
        try:
            if getattr(self, "_pending", False):
                self.finish()
        finally:
            for pl in getattr(self, "_own_plans", []):
                pl.close()
            self._own_plans = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, p0=None):
        """One step (LES -> RK predictor -> rhs -> pressure -> projection) on
        the current stream; the same launches as ``step_collocated``'s
        functions, called without their per-call argument handling."""
        with torch.cuda.device(self.device):
            return self._step(p0)

    def _step(self, p0):
        if self.overlap or not self.deferred:
            self.finish()
            return self._correct(*self._predict(p0))
        # queue this step's predictor before reading the previous solve's
        # state, so that the GPU stays busy while the host waits; a top-up of
        # the previous solve re-projects its velocity, and the predictor is
        # then queued again from it
        pred = self._predict(p0)
        if self.finish():
            pred = self._predict(p0)
        return self._correct(*pred)

    def _predict(self, p0):
        """LES, RK predictor, rhs and the warm-start guess into this step's slot."""
        L, dc, s = self._lib, self._dc, A.stream_handle(self.device)
        nx, ny, dx, dy = self.nx, self.ny, self.dx, self.dy
        u, v = self.u[self.cur], self.v[self.cur]
        nsgs = None
        if self.les:
            nsgs = self.nsgs.data_ptr()
            _lib.check(L.aj_les_smagorinsky(dc, u.data_ptr(), v.data_ptr(), nsgs, nx, ny, dx, dy,
                                            (dx * dy) ** 0.5, self.cs, s), "aj_les_smagorinsky")
        us, vs = self._usv[self.slot]
        usp, vsp, mask = us.data_ptr(), vs.data_ptr(), self.mask.data_ptr()
        # RK predictor with rhs = div(u*, v*) * mask / dt in its epilogue
        _lib.check(L.aj_rk_step_rhs(dc, u.data_ptr(), v.data_ptr(), mask, A.ptr(self.sdf), nsgs,
                                    usp, vsp, self.rhs.data_ptr(), 1, nx, ny, self._prm_ref, s),
                   "aj_rk_step_rhs")
        if p0 is None and not self.overlap:
            if self._net is not None:
                net = self._net
                _lib.check(L.aj_nn_warm_start(net.handle, self.rhs.data_ptr(), A.ptr(self.sdf),
                                              self.p0.data_ptr(), nx, ny, net.r, net.sdf_ref,
                                              net.pscale, s), "aj_nn_warm_start")
                p0 = self.p0
            elif self.warm_start is not None:
                p0 = self.warm_start(self.rhs, self.sdf)
        return p0, (p0 is None and self.overlap)

    def _correct(self, p0, overlapped):
        """Pressure solve and projection of this step's slot into the next velocity."""
        L, dc, s = self._lib, self._dc, A.stream_handle(self.device)
        us, vs = self._usv[self.slot]
        un, vn = self.u[1 - self.cur], self.v[1 - self.cur]
        if overlapped:
            self._solve_overlapped()
        else:
            self._solve(p0, deferred=self.deferred)
        _lib.check(L.aj_project(dc, us.data_ptr(), vs.data_ptr(), self.p.data_ptr(),
                                self.mask.data_ptr(), un.data_ptr(), vn.data_ptr(), self.nx,
                                self.ny, self.dt, self.dx, self.dy, self.U_inf, s), "aj_project")
        self.cur = 1 - self.cur
        self.slot = 1 - self.slot
        return un, vn

    def _solve(self, p0, deferred=False):
        if p0 is not None and (p0.dtype != self.rhs.dtype or tuple(p0.shape) != (self.nx, self.ny)
                               or p0.device != self.device or not p0.is_contiguous()):
            raise ValueError("p0 must be a contiguous tensor of the stepper's shape, dtype and device")
        fn = self._lib.aj_poisson_mg_enqueue if deferred else self._lib.aj_poisson_mg
        stream = torch.cuda.current_stream(self.device)
        rc = fn(self.plan.handle, self.rhs.data_ptr(), A.ptr(p0), self.p.data_ptr(), self._cycles,
                float(self.tolerance), stream.cuda_stream)
        _lib.check(rc, "aj_poisson_mg")
        self._pending = deferred
        self._pend_stream = stream if deferred else None

    def _solve_overlapped(self):
        main = torch.cuda.current_stream(self.device)
        self.ev_rhs.record(main)
        with torch.cuda.stream(self.side):
            self.side.wait_event(self.ev_rhs)
            self.warm_start.warm_start(self.rhs, self.sdf, self.p0)
            # guess test only (no outer iteration): used_guess, guess / zero norms
            self.guess_plan.solve(self.rhs, self.p_guess, self.p0, 0, self.tolerance)
            self.ev_guess.record(self.side)
        self._solve(None)  # zero guess, concurrently with the network
        main.wait_event(self.ev_guess)
        self.guess_stats = self.guess_plan.stats()
        if self.guess_stats["used_guess"]:
            self._solve(self.p0)

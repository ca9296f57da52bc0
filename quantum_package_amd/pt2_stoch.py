"""Hybrid deterministic/stochastic EN-PT2 (Garniron et al., J. Chem. Phys.
147, 034101 (2017)), the reference's ``ZMQ_pt2``
(``plugins/Full_CI_ZMQ/pt2_stoch_routines.irp.f:6-588``).

Every determinant is a generator and a selector (the reference's
``pt2_stoch`` sets both thresholds to 1, ``pt2_stoch.irp.f:23-26``).  Per
state, with generators sorted by that state's weight c_I^2:

* the generators whose weight is at least half of a comb step of what is
  left form a deterministic set and are computed exactly;
* the rest is cut into 100 teeth of equal weight (``comb_step``); a comb is
  one uniform offset r in [0, comb_step), it picks in tooth j the generator
  whose cumulated-weight interval holds 1 - r - (teeth - j) comb_step;
* a comb gives, for every tooth j, the sample sum_{j' >= j} e_{I_j'} /
  w_{I_j'} comb_step of the PT2 of teeth j..teeth (``do_carlo``, :155-185);
* the estimate is the computed PT2 of the generators before the first tooth
  that is not complete, plus the first generator of that tooth's share of
  it, plus the comb mean for the remaining teeth; the error is the standard
  error of that mean (``pt2_collector``, :294-330);
* work is added in batches: the generators the new combs pick, and after each
  comb the next not-yet-computed generator in order, so that the exact
  prefix grows (``get_carlo_workbatch``, :448-472).  Batches continue until
  |error / estimate| < relative_error or |error| < absolute_error.

The exact PT2 of a set of generators is what the selection computes for
them (``select_connected``): here one ``qpx_select_generators`` call on the
GPU for the batch, which runs those generators' blocks only.

Differences from the reference, on purpose:

* ``pt2_find`` (:367-390) returns, in ranges over 64 generators, the index
  below the right one when the point falls strictly between two cumulated
  weights met by its bisection; here it is the first index whose cumulated
  weight reaches the point (a plain search), so that each generator is
  drawn with probability equal to its weight;
* when every generator has been computed, the result is their exact sum
  with a zero error (the reference's collector would fall back to the first
  tooth's comb mean);
* the random offsets come from numpy's PCG64 with a given seed, not from
  Fortran's ``random_number``, so individual samples are not reproduced.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

COMB_TEETH = 100


def pt2_find(v: float, w: np.ndarray, imin: int, imax: int) -> int:
    """First 1-based index i in [imin, imax - 1] with w(i) >= v, else imax
    (w nondecreasing)."""
    seg = w[imin - 1:imax - 1]
    return imin + int(np.searchsorted(seg, v, side="left"))


@dataclass
class Comb:
    """pt2_weight / pt2_cweight / comb_step / first_det_of_teeth /
    first_det_of_comb (pt2_stoch_routines.irp.f:517-575), 1-based indices."""
    weight: np.ndarray
    cweight: np.ndarray
    step: float
    first_det_of_comb: int
    first_det_of_teeth: np.ndarray   # (teeth + 2,), entries 1..teeth+1 used
    teeth: int = COMB_TEETH

    @property
    def n(self) -> int:
        return self.weight.size


def make_comb(c: np.ndarray, teeth: int = COMB_TEETH) -> Comb:
    """The comb of generators with coefficients c (one state, generator order)."""
    w = np.asarray(c, dtype=np.float64) ** 2
    n = w.size
    if n < 2:
        raise ValueError("the stochastic PT2 needs at least two generators")
    # suffix sums, accumulated backwards for precision
    cw = np.cumsum(w[::-1])[::-1].copy()
    total = cw[0]
    weight = w / total
    cw = cw / total
    cw[:-1] = 1.0 - cw[1:]
    cw[-1] = 1.0
    step = 1.0 / teeth
    first = 1
    norm_left = 1.0
    for i in range(n):
        if weight[i] / norm_left < 0.5 * step:
            first = i + 1
            break
        norm_left -= weight[i]
    first = max(2, first)
    step = (1.0 - cw[first - 2]) * step
    fdt = np.zeros(teeth + 2, dtype=np.int64)
    stato = 1.0 - step
    iloc = n
    for i in range(teeth, 0, -1):
        iloc = pt2_find(stato, cw, 1, iloc)
        fdt[i] = iloc
        stato -= step
    fdt[teeth + 1] = n + 1
    fdt[1] = first
    return Comb(weight=weight, cweight=cw, step=step, first_det_of_comb=first, first_det_of_teeth=fdt,
                teeth=teeth)


def get_comb(comb: Comb, r: float) -> np.ndarray:
    """Generators (1-based) of comb offset r, one per tooth (get_comb, :475-490)."""
    dets = np.zeros(comb.teeth + 1, dtype=np.int64)
    curs = 1.0 - r
    for j in range(comb.teeth, 0, -1):
        dets[j] = pt2_find(curs, comb.cweight, comb.first_det_of_teeth[j], comb.first_det_of_teeth[j + 1])
        curs -= comb.step
    return dets[1:]


def first_tooth(comb: Comb, computed: np.ndarray) -> int:
    """get_first_tooth (:396-421): the tooth holding the first generator at
    or after the comb that is not computed (teeth + 1 when all are)."""
    rest = np.nonzero(~computed[comb.first_det_of_comb - 1:])[0]
    if rest.size == 0:
        return comb.teeth + 1
    first_det = comb.first_det_of_comb + int(rest[0])
    t = 1
    for i in range(comb.teeth, 0, -1):
        if comb.first_det_of_teeth[i] < first_det:
            t = i
            break
    return t


@dataclass
class Pt2Stoch:
    """State of the estimator for one state."""
    comb: Comb
    e: np.ndarray                    # exact PT2 per generator (0 until computed)
    computed: np.ndarray             # bool per generator
    sumabove: np.ndarray = field(default=None)
    sum2above: np.ndarray = field(default=None)
    nabove: np.ndarray = field(default=None)

    def __post_init__(self):
        t = self.comb.teeth
        self.sumabove = np.zeros(t + 1)
        self.sum2above = np.zeros(t + 1)
        self.nabove = np.zeros(t + 1)

    def add_comb(self, dets: np.ndarray) -> None:
        """do_carlo for one comb whose generators are all computed."""
        inv = 1.0 / self.comb.weight
        val = 0.0
        for j in range(self.comb.teeth, 0, -1):
            d = dets[j - 1]
            val += self.e[d - 1] * inv[d - 1] * self.comb.step
            self.sumabove[j] += val
            self.sum2above[j] += val * val
            self.nabove[j] += 1.0

    def estimate(self) -> tuple[float, float]:
        """(PT2, standard error): pt2_collector's estimator (:309-322)."""
        c = self.comb
        if self.computed.all():
            return float(self.e.sum()), 0.0
        tooth = first_tooth(c, self.computed)
        f = c.first_det_of_teeth[tooth]
        e0 = float(self.e[: f - 1].sum())
        prop = (1.0 - (c.teeth - tooth + 1) * c.step) - c.cweight[f - 2]
        prop *= 1.0 / c.weight[f - 1]
        e0 += self.e[f - 1] * prop
        n = self.nabove[tooth]
        if n < 2:
            return e0 + (self.sumabove[tooth] / n if n else 0.0), float("inf")
        mean = self.sumabove[tooth] / n
        err = np.sqrt(1.0 / (n - 1) * abs(self.sum2above[tooth] / n - mean * mean))
        return e0 + mean, float(err)


def stochastic_pt2_state(coef_state: np.ndarray, compute, relative_error: float, absolute_error: float,
                         seed: int = 0, combs_per_batch: int = 64, min_combs: int = 5, log=None,
                         teeth: int = COMB_TEETH, max_seconds: float | None = None):
    """Stochastic PT2 of one state.  coef_state: generator coefficients in
    generator order (decreasing weight); compute(gens) returns the exact PT2
    of the 1-based generators gens (an array).  With max_seconds, batches
    stop once that much time has passed (the estimate and its error are
    those reached).  Returns (pt2, error, info)."""
    import time
    t0 = time.time()
    comb = make_comb(coef_state, teeth)
    n = comb.n
    S = Pt2Stoch(comb=comb, e=np.zeros(n), computed=np.zeros(n, dtype=bool))
    rng = np.random.default_rng(seed)

    def run(gens):
        gens = np.unique(np.asarray(gens, dtype=np.int64))
        gens = gens[~S.computed[gens - 1]]
        if gens.size:
            S.e[gens - 1] = compute(gens)
            S.computed[gens - 1] = True

    # the deterministic set
    run(np.arange(1, comb.first_det_of_comb))
    nxt = comb.first_det_of_comb
    n_combs = 0
    est, err = S.estimate()
    while True:
        # a batch: the combs' generators, and one more generator in order per comb
        rs = rng.random(combs_per_batch) * comb.step
        picks = [get_comb(comb, r) for r in rs]
        todo = []
        for dets in picks:
            todo.append(dets)
            while nxt <= n and S.computed[nxt - 1]:
                nxt += 1
            if nxt <= n:
                todo.append(np.array([nxt]))
                nxt += 1
        run(np.concatenate(todo))
        for dets in picks:
            S.add_comb(dets)
        n_combs += len(picks)
        est, err = S.estimate()
        if log:
            log(f"pt2_stoch: {n_combs} combs, {int(S.computed.sum())}/{n} generators, PT2 = {est:.10f} +/- {err:.3e}")
        if S.computed.all():
            break
        if n_combs >= min_combs and (abs(err) < absolute_error or abs(err / est) < relative_error):
            break
        if max_seconds is not None and n_combs >= min_combs and time.time() - t0 > max_seconds:
            break
    return est, err, {"combs": n_combs, "generators_computed": int(S.computed.sum()), "generators": n,
                      "deterministic": comb.first_det_of_comb - 1, "seconds": time.time() - t0}


def ZMQ_pt2(ctx, dets: np.ndarray, coef: np.ndarray, E0, relative_error: float = 1e-3,
            absolute_error: float = 1e-9, seed: int = 0, combs_per_batch: int = 64, log=None,
            max_seconds: float | None = None):
    """EN-PT2 of psi (dets (N, 2, N_int), coef (N, N_states)) estimated
    stochastically on the GPU context ctx (its Hamiltonian set), as the
    reference's ZMQ_pt2: per state, generators and selectors are all of psi
    sorted by that state's weight.  Returns (pt2[N_states], error[N_states],
    info per state)."""
    from .cipsi import sort_by_weight
    coef = np.asarray(coef, dtype=np.float64)
    if coef.ndim == 1:
        coef = coef[:, None]
    ns = coef.shape[1]
    E0 = np.atleast_1d(np.asarray(E0, dtype=np.float64))
    n_det = dets.shape[0]
    pt2 = np.zeros(ns)
    err = np.zeros(ns)
    infos = []
    if n_det < max(10, ns):
        # the reference runs the plain selection below 10 determinants (:33-36)
        ds, cs, _ = sort_by_weight(dets, coef)
        ctx.set_wavefunction(ds)
        pt2, _, _ = ctx.ZMQ_selection(np.asfortranarray(cs), n_det, n_det, E0, 1)
        return pt2, err, [{"exact": True}] * ns
    for ist in range(ns):
        w = np.zeros(ns)
        w[ist] = 1.0
        ds, cs, _ = sort_by_weight(dets, coef, state_weights=w)
        ctx.set_wavefunction(ds)
        cf = np.asfortranarray(cs)

        def compute(gens, cf=cf, ist=ist):
            return ctx.select_generators(cf, n_det, E0, gens)[:, ist]

        p, e, info = stochastic_pt2_state(cs[:, ist], compute, relative_error, absolute_error, seed=seed + ist,
                                          combs_per_batch=combs_per_batch, log=log, max_seconds=max_seconds)
        pt2[ist], err[ist] = p, e
        infos.append(info)
    return pt2, err, infos

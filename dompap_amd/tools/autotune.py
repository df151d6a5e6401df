"""GPU autotuner with the interface of dompap/tools/autotune.py:4-178.

What the reference tunes, and what it tunes here:

* neighbour-list and force method strings: every method string runs the
  same GPU kernels, so the simulation keeps its strings;
* the skin (autotune.py:62-90): the same search -- the skin grows from
  ``smallest_skin`` by ``step_skin`` as long as a timed run of ``steps``
  steps gets faster, and stops at the first slower run or at half the
  shortest box side; the fastest skin wins;
* no counterpart in the reference, for the GPU: the kernel shape (lanes per
  particle, fused or split step kernels), the cell width (cells of 1.0 or
  more times the list cutoff: larger cells make larger tiles) and the row
  capacity (``max_number_of_neighbors``: lowered, never raised, to 1.5 times
  the longest row seen while tuning, in multiples of 8).  They are timed on
  the system itself, so each system shape gets its own choice, and stored in
  the simulation (``Simulation.set_kernel_tuning``, ``set_neighbor_list``).
"""
from __future__ import annotations

from time import perf_counter

CELL_SCALES = (1.0, 1.05, 1.1, 1.15, 1.2, 1.3)


def tune_cell_scale(engine, steps=200, scales=CELL_SCALES, integrator="gjf"):
    """Cell width of an engine's system (cells at least ``scale`` list
    cutoffs wide; larger cells make fewer, larger tiles and longer candidate
    lists): ``steps`` timed steps at each width, the fastest kept.  Each
    width's new cells, list, plan and step graphs are made outside its timed
    run.  Returns (scale, {scale: seconds})."""
    times = {}
    for scale in scales:
        engine.set_cell_scale(scale)
        engine.step(max(steps // 10, 8), integrator)
        engine.prepare_step_graphs(integrator, timed=False)
        start = perf_counter()
        engine.step(steps, integrator)  # returns after the device finished
        times[scale] = perf_counter() - start
    best = min(times, key=times.get)
    engine.set_cell_scale(best)
    return best, times


def _seconds(sim, steps):
    start = perf_counter()
    sim.run(steps)  # returns after the device finished the steps
    return perf_counter() - start


def _longest_row(sim):
    _, _, longest = sim._device().stats()
    return int(longest)


def autotune(sim, steps=100, test_double_loop=True, smallest_skin=0.1, step_skin=0.1,
             verbose=False, plot=False, tune_capacity=True):
    """Tune skin, kernel shape, cell width and row capacity of ``sim``
    (dompap/tools/autotune.py:4-6); returns ``sim``."""
    probe = sim.copy()
    probe.step()  # compiles the kernels and builds the first list

    # -- skin: grow while the timed run gets faster ------------------------
    limit = min(sim.box_vectors) / 2
    tried = []  # (skin, seconds, rebuilds)
    candidate = smallest_skin
    previous = float("inf")
    longest = 0
    while candidate < limit:
        before = probe.number_of_neighbor_list_updates
        probe.set_neighbor_list(skin=candidate)
        probe.step()  # the new list and plan, outside the timed run
        seconds = _seconds(probe, steps)
        longest = max(longest, _longest_row(probe))
        tried.append((candidate, seconds, probe.number_of_neighbor_list_updates - before))
        candidate += step_skin
        if seconds > previous:
            break
        previous = seconds
    if verbose:
        print("skin   ms/step   rebuilds in the run")
        for sk, seconds, rebuilds in tried:
            print(f"{sk:5.2f}  {seconds / steps * 1e3:8.4f}  {rebuilds} in {steps} steps")
    skin, best_seconds, _ = min(tried, key=lambda t: t[1])
    sim.set_neighbor_list(skin=skin)

    # -- kernel shape and cell width at that skin ----------------------------
    probe = sim.copy()
    probe.step()
    timings = {}

    def timed(tpp, fused, scale):
        probe.set_kernel_tuning(tpp, 0, fused, scale)
        probe.step()  # re-plan (and, for a new cell width, rebuild) outside the timed run
        timings[(tpp, fused, scale)] = _seconds(probe, steps)
        if verbose:
            print(f"{tpp} lanes per particle, {'fused' if fused else 'split'} step, cells "
                  f"{scale:.2f} x cutoff: {timings[(tpp, fused, scale)] / steps * 1e3:.4f} ms per step")
        return _longest_row(probe)

    # shape first at the narrowest cells, then the cell width for that shape
    for tpp in (1, 2, 4, 8):
        for fused in (0, 1):
            longest = max(longest, timed(tpp, fused, CELL_SCALES[0]))
    tpp, fused, _ = min(timings, key=timings.get)
    for scale in CELL_SCALES[1:]:
        longest = max(longest, timed(tpp, fused, scale))
    tpp, fused, scale = min(timings, key=timings.get)
    sim.set_kernel_tuning(tpp, 0, fused, scale)

    # -- row capacity -----------------------------------------------------
    capacity = int(sim.max_number_of_neighbors)
    if tune_capacity and longest > 0:
        want = max(64, -(-int(1.5 * longest) // 8) * 8)
        if want < capacity:
            capacity = want
            sim.set_neighbor_list(max_number_of_neighbors=capacity)
    if verbose:
        print(f"chosen: skin {skin}, {tpp} lanes per particle, {'fused' if fused else 'split'} step, "
              f"cells {scale:.2f} x cutoff, max_number_of_neighbors {capacity} (longest row seen {longest})")
    if plot:
        import matplotlib.pyplot as plt
        plt.plot([t[0] for t in tried], [t[1] for t in tried], "o", label="runs")
        plt.plot(skin, best_seconds, "ro", label="fastest")
        plt.xlabel("skin")
        plt.ylabel("seconds")
        plt.legend()
        plt.show()
    return sim
